"""Exact and approximate views of packed multiple-precision arrays for the parity tests.

A packed element (include/sdpg.h) is [e, sign, m_0 .. m_{L-1}] with value (-1)^sign M 2^(e - 32L),
M = sum m_k 2^(32k).  Every such value is a dyadic rational, so a whole array is exactly
(integer array) x 2^E for one E: products and sums of these integers are exact and fast (Python
ints in numpy object arrays).  The float64 view is for error bounds only.
"""
from __future__ import annotations

import re
from fractions import Fraction

import numpy as np

EZERO = -(1 << 29)


def _parts(arr, L):
    a = np.asarray(arr, dtype=np.uint32).reshape(-1, L + 2)
    e = a[:, 0].view(np.int32).astype(np.int64)
    neg = (a[:, 1] & 1).astype(bool)
    limbs = np.ascontiguousarray(a[:, 2:])
    return e, neg, limbs


def to_ints(arr, L):
    """(object array of Python ints v, E) with value_k = v[k] * 2^E exactly (E = 0 if all zero)."""
    e, neg, limbs = _parts(arr, L)
    n = len(e)
    M = [int.from_bytes(limbs[k].tobytes(), "little") for k in range(n)]
    nz = [k for k in range(n) if M[k]]
    if not nz:
        return np.array([0] * n, dtype=object), 0
    E = min(int(e[k]) - 32 * L for k in nz)
    out = np.empty(n, dtype=object)
    for k in range(n):
        v = M[k] << (int(e[k]) - 32 * L - E) if M[k] else 0
        out[k] = -v if neg[k] else v
    return out, E


def to_float(arr, L):
    """float64 approximation (top 64 bits of each mantissa); zero for zero."""
    e, neg, limbs = _parts(arr, L)
    top = limbs[:, L - 1].astype(np.float64) * 4294967296.0 + limbs[:, L - 2].astype(np.float64)
    v = np.ldexp(top, (e - 64).clip(-2000, 2000).astype(np.int32))
    v[limbs[:, L - 1] == 0] = 0.0
    return np.where(neg, -v, v)


def align(a, Ea, b, Eb):
    """Two exact integer arrays at exponents Ea, Eb brought to the common exponent min(Ea, Eb)."""
    E = min(Ea, Eb)
    return (a * (1 << (Ea - E)) if Ea > E else a), (b * (1 << (Eb - E)) if Eb > E else b), E


def diff_float(a, Ea, b, Eb):
    """float64 of a 2^Ea - b 2^Eb elementwise, computed exactly before rounding."""
    x, y, E = align(a, Ea, b, Eb)
    d = x - y
    return np.array([float(Fraction(int(v)) * Fraction(2) ** E) if v else 0.0 for v in d], dtype=np.float64)


def ldexp_obj(v, E):
    """float64 of each integer v 2^E (exact integers may exceed the float range before scaling)."""
    out = np.zeros(len(v), dtype=np.float64)
    for k, x in enumerate(v):
        x = int(x)
        if x:
            b = abs(x).bit_length()
            sh = max(0, b - 60)
            out[k] = np.ldexp(float(x >> sh if x > 0 else -((-x) >> sh)), E + sh)
    return out


_NUM = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdD][+-]?\d+)?")


def read_problem(path):
    """SDPA sparse file -> (m, sdp block sizes, n_lp, b as Fractions, C and A_1..A_m as lists of
    (sdp block, row, col, Fraction) with 0-based row <= col as read, C negated as the reference
    stores it (sdpa_io.cpp:661))."""
    with open(path) as f:
        lines = f.readlines()
    k = 0
    while lines[k].startswith("*") or lines[k].startswith('"'):
        k += 1
    m = int(_NUM.findall(lines[k])[0])
    toks = _NUM.findall("".join(lines[k + 1:]))
    nblock = int(toks[0])
    bstruct = [int(t) for t in toks[1:1 + nblock]]
    b = [Fraction(t) for t in toks[1 + nblock:1 + nblock + m]]
    pos = 1 + nblock + m
    sdp_number, sizes, n_lp = {}, [], 0
    for l, s in enumerate(bstruct):
        if s >= 2:
            sdp_number[l] = len(sizes)
            sizes.append(s)
        else:
            n_lp += abs(s)
    mats = [[] for _ in range(m + 1)]
    while pos + 5 <= len(toks):
        kk, ll, i, j = (int(float(t)) for t in toks[pos:pos + 4])
        v = Fraction(toks[pos + 4])
        pos += 5
        if (ll - 1) in sdp_number:
            mats[kk].append((sdp_number[ll - 1], i - 1, j - 1, -v if kk == 0 else v))
    return m, sizes, n_lp, b, mats[0], mats[1:]


def dense_blocks(entries, sizes):
    """list of (block, i, j, v) -> per block a symmetric n x n object array of Fractions."""
    out = [np.full((n, n), Fraction(0), dtype=object) for n in sizes]
    for blk, i, j, v in entries:
        out[blk][i, j] = v
        out[blk][j, i] = v
    return out


def block_views(flat, sizes):
    """dense linear space (blocks column-major back to back) -> list of n x n arrays (row i, col j)."""
    out, off = [], 0
    for n in sizes:
        out.append(np.asarray(flat[off:off + n * n]).reshape(n, n).T)
        off += n * n
    return out
