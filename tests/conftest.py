import json
import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sdpa-gmp_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

PARAM_256 = os.path.join(PKG, "data", "param_256_gap1e-60.sdpa")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsdpg.so")


_WORKLOADS = {}


@pytest.fixture(scope="session")
def workload():
    """workload(k) -> path of the generated .dat-s of workload k (sha256-checked against inputs.json)."""
    import gen_problems

    with open(os.path.join(GOLDEN, "inputs.json")) as f:
        hashes = json.load(f)
    tmp = tempfile.mkdtemp(prefix="sdpg_w_")

    def get(k):
        if k not in _WORKLOADS:
            path = os.path.join(tmp, f"w{k}.dat-s")
            h = gen_problems.generate(k, path)
            key = f"w{k}"
            if key in hashes:
                assert h == hashes[key], f"generated workload {k} differs from the fixture inputs"
            _WORKLOADS[k] = path
        return _WORKLOADS[k]

    return get


@pytest.fixture(scope="session")
def sdplib_file():
    """sdplib_file(name) -> path of a reference regression input (decompressed)."""
    import lzma
    import shutil

    tmp = tempfile.mkdtemp(prefix="sdpg_lib_")

    def get(name):
        src = os.path.join(GOLDEN, "sdplib", name + ".dat-s")
        dst = os.path.join(tmp, name + ".dat-s")
        if not os.path.exists(dst):
            if os.path.exists(src):
                shutil.copy(src, dst)
            else:
                with lzma.open(src + ".xz", "rb") as fi, open(dst, "wb") as fo:
                    fo.write(fi.read())
        return dst

    return get
