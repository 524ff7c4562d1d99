"""SDPA formats on the host (no device): the dense reader (sdpa_io.cpp:696-751) reads the same structure as
the sparse one, seen through the formula table the host driver computes from it (stored entries,
NonZeroEffect, block membership), and the test's dense-to-sparse conversion is checked against the
reference's own pair of files (example1.dat / example1.dat-s)."""
import os
import tempfile

import numpy as np

import formats
import sdpg
from exact import read_problem

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sdplib")


def test_conversion_matches_reference_pair():
    out = os.path.join(tempfile.mkdtemp(), "example1.dat-s")
    formats.dense_to_sparse(os.path.join(LIB, "example1.dat"), out)
    a = read_problem(out)
    b = read_problem(os.path.join(LIB, "example1.dat-s"))
    assert a[:3] == b[:3] and a[3] == b[3]
    assert sorted(a[4]) == sorted(b[4]) and [sorted(x) for x in a[5]] == [sorted(x) for x in b[5]]


def test_dense_reader_structure():
    tmp = tempfile.mkdtemp()
    for name in ("example1", "example2"):
        sparse = os.path.join(tmp, name + ".dat-s")
        formats.dense_to_sparse(os.path.join(LIB, name + ".dat"), sparse)
        fd = sdpg.formula(os.path.join(LIB, name + ".dat"))
        fs = sdpg.formula(sparse)
        assert fd.shape == fs.shape and np.array_equal(fd, fs), name
