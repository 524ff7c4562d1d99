"""SDPA wire formats end to end (sdpa_io.cpp): the dense data reader, initial-point files and the solution
printout, checked against the reference's own example files and recorded result.

- example1.dat (dense) and example1.dat-s (sparse) are the same problem: identical runs, and both match
  example1.result (pdOPT, 39 iterations, objective -41.9);
- example2.dat (dense, an SDP block of 2, one of 3 and an LP block of 2) against its conversion to the
  sparse format: identical runs;
- example1.ini (dense) and example1.ini-s (sparse) are the same initial point (sdpa_io.cpp:150-274):
  identical runs from it, to the same optimum;
- the solution printout (xVec = -y, xMat = Z, yMat = X with the 40-digit formats of the parameter file)
  has the layout of example1.result and its values to 1e-25;
- restarting from the reference's printed solution (written as an initial-point file) ends at once in
  pdOPT: the printed point is optimal in the reading conventions (sign of y, which matrix is which).
"""
import json
import os
import tempfile
from decimal import Decimal, localcontext
from fractions import Fraction

import pytest

import formats
import sdpg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIB = os.path.join(GOLD, "sdplib")


def _run(data, **kw):
    log = tempfile.mktemp(suffix=".log")
    r = sdpg.run(data, log_path=log, **kw)
    with open(log) as f:
        return r, f.read()


def _same(r1, r2):
    assert r1.phase == r2.phase and r1.iterations == r2.iterations
    assert r1.objP70 == r2.objP70 and r1.objD70 == r2.objD70


@pytest.mark.gpu
def test_dense_and_sparse_data_agree():
    with open(os.path.join(GOLD, "sdplib_golden.json")) as f:
        gold = json.load(f)["example1"]
    print("case example1 dense / sparse", flush=True)
    rd, _ = _run(os.path.join(LIB, "example1.dat"))
    rs, _ = _run(os.path.join(LIB, "example1.dat-s"))
    _same(rd, rs)
    assert rd.phase.decode().strip() == gold["phase"]
    assert abs(rd.iterations - gold["iterations"]) <= 1
    assert abs(Decimal(rd.objP.decode()) - Decimal(gold["objP"])) < Decimal("1e-14")


@pytest.mark.gpu
def test_dense_data_with_lp_block():
    tmp = tempfile.mkdtemp()
    sparse = os.path.join(tmp, "example2.dat-s")
    formats.dense_to_sparse(os.path.join(LIB, "example2.dat"), sparse)
    print("case example2 dense / sparse", flush=True)
    rd, _ = _run(os.path.join(LIB, "example2.dat"))
    rs, _ = _run(sparse)
    print(f"  {rd.iterations} iterations, {rd.phase.decode().strip()}, objP {rd.objP.decode()}")
    _same(rd, rs)


@pytest.mark.gpu
def test_initial_point_files():
    data = os.path.join(LIB, "example1.dat-s")
    print("case example1 from example1.ini / .ini-s", flush=True)
    rd, ld = _run(data, init_path=os.path.join(LIB, "example1.ini"))
    rs, ls = _run(data, init_path=os.path.join(LIB, "example1.ini-s"))
    r0, _ = _run(data)
    print(f"  from the initial point: {rd.iterations} iterations, {rd.phase.decode().strip()} "
          f"(default start: {r0.iterations})")
    _same(rd, rs)
    assert rd.phase.decode().strip() == "pdOPT"
    assert abs(Decimal(rd.objP70.decode()) - Decimal(r0.objP70.decode())) < Decimal("1e-28")
    assert rd.iterations != r0.iterations  # the start matters


@pytest.mark.gpu
def test_solution_printout_matches_reference():
    with open(os.path.join(LIB, "example1.result")) as f:
        ref = formats.solution(f.read())
    print("case example1 solution printout", flush=True)
    _, log = _run(os.path.join(LIB, "example1.dat-s"))
    got = formats.solution(log)
    for name in ("xVec", "xMat", "yMat"):
        assert got[name + "_layout"] == ref[name + "_layout"], name
        assert len(got[name]) == len(ref[name])
        for a, b in zip(got[name], ref[name]):
            assert abs(a - b) <= Fraction(1, 10 ** 25) * (1 + abs(b)), (name, float(a), float(b))


def _dec(v: Fraction) -> str:
    with localcontext() as ctx:
        ctx.prec = 60
        return f"{Decimal(v.numerator) / Decimal(v.denominator):.45e}"


@pytest.mark.gpu
def test_restart_from_printed_solution():
    with open(os.path.join(LIB, "example1.result")) as f:
        ref = formats.solution(f.read())
    ini = os.path.join(tempfile.mkdtemp(), "restart.ini-s")
    n = 2
    lines = [" ".join(_dec(v) for v in ref["xVec"])]
    for target, vals in ((1, ref["xMat"]), (2, ref["yMat"])):  # 1: Z (printed as xMat), 2: X (yMat)
        for i in range(n):
            for j in range(i, n):
                lines.append(f"{target} 1 {i + 1} {j + 1} {_dec(vals[i * n + j])}")
    with open(ini, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("case example1 restarted from its printed solution", flush=True)
    r, _ = _run(os.path.join(LIB, "example1.dat-s"), init_path=ini)
    print(f"  {r.iterations} iterations, {r.phase.decode().strip()}, objP {r.objP.decode()}")
    assert r.phase.decode().strip() == "pdOPT"
    assert r.iterations <= 2
    assert abs(Decimal(r.objP.decode()) + Decimal("41.9")) < Decimal("1e-14")
