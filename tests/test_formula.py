"""The formula table of Newton::computeFormula_SDP (sdpa_newton.cpp:694-819), pinned.

For every (constraint i, SDP block l) the reference picks F1, F2 or F3 from three cost estimates

    f1 = kappa n inz + n^3 + kappa up
    f2 = kappa n inz + kappa (n + 1) up
    f3 = kappa (2 kappa inz + 1) up / kappa

(inz = NonZeroEffect of A_i's block, up = inz + the NonZeroEffect of every partner j of the same block
with jnz < inz, or jnz == inz and j < i; sdpa_newton.cpp:728-753), restricted to F1 / F2 when the block
was switched to dense storage (20 % rule, sdpa_struct.cpp:491).  kappa is the double 1.2
(sdpa_main.cpp:27) converted to mpf, so the decisions are those of exact rational arithmetic with
kappa = Fraction(1.2): every difference of two estimates is a nonzero multiple of 2^-52 (or an
integer), far above the mpf rounding of the estimates.

This file restates the rule independently, in Python Fractions from the problem text, and checks
that the table the host driver uploads (sdpa_amd_formula) and the oracle's table (orc_formula) both
equal it on the reference's regression inputs, the seeded workloads and a sample of the batch
workload.  A second test shows that the inputs discriminate: each plausible slip in the rule (the
tie order, a strict / non-strict comparison, the kappa value, the f3 factor, the dense threshold)
changes the table of at least one input.
"""
from __future__ import annotations

import ctypes
import os
import re
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib
import sdpg

KAPPA = Fraction(1.2)  # the double 1.2 exactly: 5404319552844595 / 2^52

SDPLIB = ["example1", "control1", "control2", "truss1", "truss4", "hinf1", "hinf3", "theta1", "infp1", "infd1",
          "qap5"]

_NUM = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?")


def structure(path: str):
    """m, SDP block sizes and, per constraint, {block: (stored entries, NonZeroEffect)} as the
    reference's sparse reader counts them (sdpa_io.cpp:626-751, sdpa_struct.cpp:1259-1291)."""
    with open(path) as f:
        lines = f.readlines()
    k = 0
    while lines[k].startswith("*") or lines[k].startswith('"'):
        k += 1
    m = int(_NUM.findall(lines[k])[0])
    toks = _NUM.findall("".join(lines[k + 1:]))
    nblock = int(toks[0])
    bstruct = [int(t) for t in toks[1:1 + nblock]]
    pos = 1 + nblock + m  # skip b
    sdp_number, sizes = {}, []
    for l, s in enumerate(bstruct):
        if s >= 2:
            sdp_number[l] = len(sizes)
            sizes.append(s)
    cnt = [dict() for _ in range(m)]
    while pos + 5 <= len(toks):
        kk, ll, i, j = (int(float(t)) for t in toks[pos:pos + 4])
        pos += 5
        if kk == 0 or (ll - 1) not in sdp_number:
            continue
        b = sdp_number[ll - 1]
        stored, eff = cnt[kk - 1].get(b, (0, 0))
        cnt[kk - 1][b] = (stored + 1, eff + (1 if i == j else 2))
    return m, sizes, cnt


def table(path: str, *, tie_lower: bool = True, strict: bool = True, kappa: Fraction = KAPPA,
          f3_factor: int = 2, dense_ratio: float = 0.20) -> np.ndarray:
    """The formula table in exact arithmetic; the keyword arguments plant the slips of the
    discrimination test (the defaults are the reference's rule)."""
    m, sizes, cnt = structure(path)
    ns = len(sizes)
    out = np.zeros((m, ns), dtype=np.int32)
    for l, n in enumerate(sizes):
        cons = [i for i in range(m) if l in cnt[i]]  # SDP_constraint[l] (sdpa_dataset.cpp:254-325)
        eff, dense = {}, {}
        for i in cons:
            stored, e = cnt[i][l]
            dense[i] = not (stored < (n * n) * dense_ratio)  # changeToDense (sdpa_struct.cpp:491)
            eff[i] = n * n if dense[i] else e
        for i in cons:
            inz, up = eff[i], eff[i]
            for j in cons:
                jnz = eff[j]
                below = jnz < inz if strict else jnz <= inz
                if below or (jnz == inz and ((j < i) if tie_lower else (j > i))):
                    if not (not strict and jnz == inz and j == i):
                        up += jnz
            f1 = kappa * n * inz + n ** 3 + kappa * up
            f2 = kappa * n * inz + kappa * (n + 1) * up
            f3 = kappa * (f3_factor * kappa * inz + 1) * up / kappa
            if dense[i]:
                out[i, l] = 1 if f1 < f2 else 2
            elif f1 < f2 and f1 < f3:
                out[i, l] = 1
            elif f2 < f3:
                out[i, l] = 2
            else:
                out[i, l] = 3
    return out


def oracle_table(path: str) -> np.ndarray:
    lib = oracle_lib.lib()
    lib.orc_formula.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    m, ns = ctypes.c_int(), ctypes.c_int()
    cap = 1 << 22
    buf = (ctypes.c_int * cap)()
    rc = lib.orc_formula(path.encode(), 256, ctypes.byref(m), ctypes.byref(ns), buf, cap)
    assert rc == 0, rc
    return np.frombuffer(buf, dtype=np.int32, count=m.value * ns.value).reshape(m.value, ns.value).copy()


_INPUTS = {}


@pytest.fixture(scope="module")
def inputs(workload, sdplib_file):
    """name -> problem path: the reference's regression inputs, workloads 0-2, 12 batch problems."""
    if not _INPUTS:
        for name in SDPLIB:
            _INPUTS[name] = sdplib_file(name)
        for k in (0, 1, 2):
            _INPUTS[f"w{k}"] = workload(k)
        import gen_problems
        tmp = tempfile.mkdtemp(prefix="sdpg_formula_")
        for k, p in enumerate(gen_problems.generate_batch(tmp, 12)):
            _INPUTS[f"batch{k:03d}"] = p
    return _INPUTS


def test_formula_tables_match_exact_rule(inputs):
    seen = set()
    for name, path in inputs.items():
        ref = table(path)
        seen |= set(np.unique(ref).tolist())
        host = sdpg.formula(path)
        orc = oracle_table(path)
        assert host.shape == ref.shape, (name, host.shape, ref.shape)
        assert np.array_equal(host, ref), (name, np.argwhere(host != ref)[:5])
        assert np.array_equal(orc, ref), (name, np.argwhere(orc != ref)[:5])
    # the inputs exercise all three formulas
    assert {1, 2, 3} <= seen, seen


def test_inputs_discriminate_rule_slips(inputs):
    ref = {name: table(path) for name, path in inputs.items()}
    slips = {
        "tie order j > i": dict(tie_lower=False),
        "jnz <= inz": dict(strict=False),
        "kappa = 1": dict(kappa=Fraction(1)),
        "kappa = 6/5 exactly": dict(kappa=Fraction(6, 5)),
        "f3 without the factor 2": dict(f3_factor=1),
        "dense threshold 10 %": dict(dense_ratio=0.10),
    }
    for what, kw in slips.items():
        changed = [name for name, path in inputs.items() if not np.array_equal(table(path, **kw), ref[name])]
        if what == "kappa = 6/5 exactly":
            # 6/5 and the double 1.2 differ by 2^-52 / 5: no decision here is that close
            assert not changed, changed
        else:
            assert changed, f"no input distinguishes the slip '{what}'"
