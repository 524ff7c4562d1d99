"""The C-ABI boundary without a GPU: libsdpg.so loads and exports every function that
include/sdpg.h and include/sdpa_amd.h declare; the pure helpers answer; the Python binding's
export list and the headers agree.  No device call is made."""
import ctypes
import os
import re

import sdpg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("sdpg.h", "sdpa_amd.h")]


def declared():
    names = []
    for h in HEADERS:
        with open(h) as f:
            text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?(?:int|void|char|double)\s*\*?\s*((?:sdpg|sdpa_amd)_\w+)\s*\(", text,
                            flags=re.M)
    return sorted(set(names))


def test_headers_declare_the_expected_entry_points():
    names = declared()
    for must in ("sdpg_ctx_create", "sdpg_schur", "sdpg_potrf", "sdpg_potrs", "sdpg_gvec", "sdpg_dz", "sdpg_dx",
                 "sdpg_residuals", "sdpg_rgemm", "sdpg_factor_point", "sdpg_min_eigen", "sdpa_amd_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = sdpg.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_export_list_matches_headers():
    assert sorted(set(sdpg.EXPORTS)) == declared()


def test_limb_rule_matches_gmp_precision():
    lib = sdpg.load()
    # GMP mpf: prec = (max(53, p) + 127) / 64 64-bit limbs; the device format keeps as many bits
    for bits, limbs in ((53, 4), (128, 6), (200, 10), (256, 10), (320, 12), (512, 18), (1024, 34)):
        assert lib.sdpg_limbs(bits) == limbs == sdpg.limbs_for_bits(bits)


def test_last_error_is_a_string():
    lib = sdpg.load()
    assert isinstance(lib.sdpg_last_error(), bytes)


def test_packed_format_round_trip():
    from fractions import Fraction
    L = sdpg.limbs_for_bits(256)
    vals = [Fraction(0), Fraction(1), Fraction(-3, 7), Fraction(2) ** -300, Fraction(10) ** 40 + 1]
    back = sdpg.unpack(sdpg.pack(vals, L), L)
    for v, b in zip(vals, back):
        assert abs(v - b) <= abs(v) * Fraction(1, 2 ** (32 * L - 1))
