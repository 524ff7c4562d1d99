"""Seeded SDPA sparse-format (.dat-s) problem generator for the benchmark workloads.

Every number written is a dyadic rational (an integer times 2^-k) printed with its
exact finite decimal expansion, so any reader that converts decimal text at >= 64
bits (GMP's mpf_set_str in the reference, the C++ reader of this package) obtains
bit-identical inputs.  Feasible-by-construction problems follow the SDPA sign
convention: the file's F0 is read as C = -F0 (reference sdpa_io.cpp:661), so the
generator writes F0 = -C, c = b and F_i = A_i; then X0 is strictly feasible for the
internal primal (A_i . X = b_i) and (y0, Z0) for the internal dual
(sum y_i A_i + Z = C).

Workloads (random generator: numpy PCG64, seed 0x5D9A0000 + index):
  0: one dense block n=15, m=21, dense symmetric N(0,1) A_i
  1: blocks (50, 25), m=351, ~8% of positions per block, N(0,1), symmetrised
  2: max-cut relaxation n=m=1000, A_i = e_i e_i^T, b=1, C = -L/4 of G(1000, 0.01),
     edge weights uniform in {-1, +1}
  3: one block n=150, m=1326, rank-2 A_i = u v^T + v u^T, u, v ~ N(0, 1/n)
  4: 512 independent problems, 1-3 blocks of size U[8,40], m U[10,120], dense A_i
"""
from __future__ import annotations

import argparse
import hashlib
import os
import sys

import numpy as np

SEED_BASE = 0x5D9A0000
QBITS = 20  # values are multiples of 2^-QBITS


def _q(x: np.ndarray) -> np.ndarray:
    """Round float64 array to integers in units of 2^-QBITS (int64)."""
    return np.rint(np.asarray(x, dtype=np.float64) * (1 << QBITS)).astype(np.int64)


def dyadic_str(k: int, s: int) -> str:
    """Exact decimal text of k * 2^-s."""
    if k == 0:
        return "0"
    neg = k < 0
    k = -k if neg else k
    if s <= 0:
        txt = str(k << (-s))
    else:
        num = k * (5 ** s)  # k/2^s = k*5^s/10^s
        ip, fp = divmod(num, 10 ** s)
        frac = str(fp).rjust(s, "0").rstrip("0")
        txt = str(ip) + ("." + frac if frac else "")
    return ("-" if neg else "") + txt


def _orth(rng: np.random.Generator, n: int) -> np.ndarray:
    """Random orthogonal matrix by modified Gram-Schmidt on a Gaussian matrix."""
    g = rng.standard_normal((n, n))
    q = np.zeros((n, n))
    for j in range(n):
        v = g[:, j].copy()
        for i in range(j):
            v -= (q[:, i] @ v) * q[:, i]
        q[:, j] = v / np.sqrt(v @ v)
    return q


def _pd_block(rng: np.random.Generator, n: int) -> np.ndarray:
    """Q diag(U(0.5, 2)) Q^T quantised to the dyadic grid (symmetric int64)."""
    q = _orth(rng, n)
    d = rng.uniform(0.5, 2.0, n)
    x = (q * d) @ q.T
    x = 0.5 * (x + x.T)
    xi = _q(x)
    return np.triu(xi) + np.triu(xi, 1).T


class Problem:
    """Block-diagonal SDP held as exact integers: value = int * 2^-scale."""

    def __init__(self, blocks):
        self.blocks = list(blocks)
        self.m = 0
        self.A = []  # list over constraints of {block: sym int64 matrix (units 2^-QBITS)}
        self.b = []  # python ints, units 2^-2QBITS
        self.C = {}  # block -> python-int object matrix, units 2^-2QBITS

    def write(self, path: str, title: str) -> str:
        lines = [f'"{title}"', f"{self.m} = mDIM", f"{len(self.blocks)} = nBLOCK",
                 " ".join(str(b) for b in self.blocks) + " = bLOCKsTRUCT",
                 " ".join(dyadic_str(v, 2 * QBITS) for v in self.b)]
        # F0 = -C
        for l in range(len(self.blocks)):
            c = self.C.get(l)
            if c is None:
                continue
            n = self.blocks[l]
            for i in range(n):
                for j in range(i, n):
                    v = int(c[i][j])
                    if v != 0:
                        lines.append(f"0 {l + 1} {i + 1} {j + 1} {dyadic_str(-v, 2 * QBITS)}")
        for k, amap in enumerate(self.A):
            for l in sorted(amap):
                a = amap[l]
                if isinstance(a, tuple):  # (rows, cols, values) upper-triangle COO
                    ii, jj, vv = a
                else:
                    ii, jj = np.nonzero(np.triu(a))
                    vv = a[ii, jj]
                for i, j, v in zip(np.asarray(ii).tolist(), np.asarray(jj).tolist(), np.asarray(vv).tolist()):
                    lines.append(f"{k + 1} {l + 1} {i + 1} {j + 1} {dyadic_str(int(v), QBITS)}")
        text = "\n".join(lines) + "\n"
        with open(path, "w") as f:
            f.write(text)
        return hashlib.sha256(text.encode()).hexdigest()


def _finish_feasible(prob: Problem, rng: np.random.Generator, x0: dict, z0: dict) -> None:
    """b_i = <A_i, X0> exactly; C = sum y0_i A_i + Z0 exactly."""
    y0 = _q(rng.standard_normal(prob.m))
    prob.b = []
    for amap in prob.A:
        s = 0
        for l, a in amap.items():
            prod = a * x0[l]  # each term < 2^62 for the sizes generated here
            s += sum(int(v) for v in prod.sum(axis=1))
        prob.b.append(s)
    for l, n in enumerate(prob.blocks):
        acc = np.zeros((n, n), dtype=object)
        for k, amap in enumerate(prob.A):
            if l in amap and y0[k] != 0:
                acc = acc + amap[l].astype(object) * int(y0[k])
        acc = acc + z0[l].astype(object) * (1 << QBITS)
        prob.C[l] = acc


def _sym_dense(rng, n):
    g = _q(rng.standard_normal((n, n)))
    return np.triu(g) + np.triu(g, 1).T


def gen0(rng):
    p = Problem([15])
    p.m = 21
    p.A = [{0: _sym_dense(rng, 15)} for _ in range(p.m)]
    x0 = {0: _pd_block(rng, 15)}
    z0 = {0: _pd_block(rng, 15)}
    _finish_feasible(p, rng, x0, z0)
    return p


def gen1(rng):
    p = Problem([50, 25])
    p.m = 351
    p.A = []
    for _ in range(p.m):
        amap = {}
        for l, n in enumerate(p.blocks):
            npos = n * (n + 1) // 2
            mask = rng.random(npos) < 0.08
            if not mask.any():
                mask[rng.integers(npos)] = True
            iu, ju = np.triu_indices(n)
            a = np.zeros((n, n), dtype=np.int64)
            vals = _q(rng.standard_normal(int(mask.sum())))
            a[iu[mask], ju[mask]] = vals
            a = a + np.triu(a, 1).T
            amap[l] = a
        p.A.append(amap)
    x0 = {l: _pd_block(rng, n) for l, n in enumerate(p.blocks)}
    z0 = {l: _pd_block(rng, n) for l, n in enumerate(p.blocks)}
    _finish_feasible(p, rng, x0, z0)
    return p


def gen2(rng, n=1000, pedge=0.01):
    p = Problem([n])
    p.m = n
    p.A = [{0: ([i], [i], [1 << QBITS])} for i in range(n)]
    p.b = [1 << (2 * QBITS)] * n
    iu, ju = np.triu_indices(n, 1)
    mask = rng.random(iu.size) < pedge
    w = rng.choice(np.array([-1, 1]), size=int(mask.sum()))
    W = np.zeros((n, n), dtype=np.int64)
    W[iu[mask], ju[mask]] = w
    W = W + W.T
    L = np.diag(W.sum(axis=1)) - W
    # C = -L/4 in units of 2^-2QBITS
    p.C[0] = (-L).astype(object) * (1 << (2 * QBITS - 2))
    return p


def gen3(rng, n=150, m=1326):
    p = Problem([n])
    p.m = m
    p.A = []
    for _ in range(m):
        u = rng.standard_normal(n) / np.sqrt(n)
        v = rng.standard_normal(n) / np.sqrt(n)
        a = _q(np.outer(u, v) + np.outer(v, u))
        a = np.triu(a) + np.triu(a, 1).T
        p.A.append({0: a})
    x0 = {0: _pd_block(rng, n)}
    z0 = {0: _pd_block(rng, n)}
    _finish_feasible(p, rng, x0, z0)
    return p


def gen4_one(rng):
    nb = int(rng.integers(1, 4))
    blocks = [int(rng.integers(8, 41)) for _ in range(nb)]
    p = Problem(blocks)
    p.m = int(rng.integers(10, 121))
    p.A = [{l: _sym_dense(rng, n) for l, n in enumerate(blocks)} for _ in range(p.m)]
    x0 = {l: _pd_block(rng, n) for l, n in enumerate(blocks)}
    z0 = {l: _pd_block(rng, n) for l, n in enumerate(blocks)}
    _finish_feasible(p, rng, x0, z0)
    return p


GENERATORS = {0: gen0, 1: gen1, 2: gen2, 3: gen3}
TITLES = {
    0: "workload 0: dense n=15 m=21",
    1: "workload 1: blocks (50,25) m=351 8% sparse",
    2: "workload 2: max-cut n=m=1000 G(1000,0.01) weights +-1",
    3: "workload 3: n=150 m=1326 rank-2 constraints",
}


def generate(index: int, path: str) -> str:
    """Write workload `index` to `path`; returns the sha256 of the text."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + index))
    if index == 4:
        raise ValueError("workload 4 is a batch; use generate_batch")
    return GENERATORS[index](rng).write(path, TITLES[index])


def generate_batch(outdir: str, count: int = 512) -> list:
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 4))
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for k in range(count):
        p = gen4_one(rng)
        path = os.path.join(outdir, f"batch{k:03d}.dat-s")
        p.write(path, f"workload 4 problem {k}")
        paths.append(path)
    return paths


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("index", type=int)
    ap.add_argument("out")
    a = ap.parse_args(argv)
    if a.index == 4:
        generate_batch(a.out)
    else:
        print(generate(a.index, a.out))


if __name__ == "__main__":
    sys.exit(main())
