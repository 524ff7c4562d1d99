"""Helpers for the SDPA file-format tests: a dense (.dat) to sparse (.dat-s) conversion written from the
format descriptions of sdpa_io.cpp:410-751 (dense: every block in full, C first; sparse: "k l i j v" lines
with i <= j), and a reader of the solution printout of a result file (sdpa_io.cpp:1167-1219)."""
from __future__ import annotations

import re
from fractions import Fraction

_NUM = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")


def dense_to_sparse(src: str, dst: str) -> None:
    with open(src) as f:
        lines = f.readlines()
    k = 0
    while lines[k].startswith("*") or lines[k].startswith('"'):
        k += 1
    m = int(_NUM.findall(lines[k])[0])
    toks = _NUM.findall("".join(lines[k + 1:]))
    nb = int(toks[0])
    bs = [int(t) for t in toks[1:1 + nb]]
    b = toks[1 + nb:1 + nb + m]
    pos = 1 + nb + m
    out = [f"{m} = mDIM", f"{nb} = nBLOCK", " ".join(str(x) for x in bs) + " = bLOCKsTRUCT", " ".join(b)]
    for kk in range(m + 1):
        for l, s in enumerate(bs):
            if s >= 2:
                for i in range(s):
                    for j in range(s):
                        v = toks[pos]
                        pos += 1
                        if i <= j and Fraction(v) != 0:
                            out.append(f"{kk} {l + 1} {i + 1} {j + 1} {v}")
            else:
                for j in range(abs(s)):
                    v = toks[pos]
                    pos += 1
                    if Fraction(v) != 0:
                        out.append(f"{kk} {l + 1} {j + 1} {j + 1} {v}")
    with open(dst, "w") as f:
        f.write("\n".join(out) + "\n")


def solution(text: str) -> dict:
    """{'xVec': [...], 'xMat': [...], 'yMat': [...]} as Fractions in print order, and the layout of each
    part with every number replaced by '#'"""
    out = {}
    for name, nxt in (("xVec", "xMat"), ("xMat", "yMat"), ("yMat", None)):
        start = text.index(name + " = \n") + len(name) + 4
        end = text.index(nxt + " = \n") if nxt else None
        part = text[start:end]
        if nxt is None:  # the yMat block ends at its closing brace line
            part = part[:part.index("\n}\n") + 3]
        out[name] = [Fraction(v) for v in _NUM.findall(part)]
        out[name + "_layout"] = re.sub(r"[+-]?\d\.\d+e[+-]\d+", "#", part)
    return out
