"""ctypes binding of the MI355X Newton engine (include/sdpg.h) and host driver
(include/sdpa_amd.h).

Mirrors the reference's Newton interface (sdpa_newton.h:49-168): a context is the
Newton object of one problem; gvec / schur / potrf / potrs / dz / dx are the steps of
Newton::Mehrotra (sdpa_newton.cpp:1375-1414).  Multiple-precision values cross the
boundary in the packed format of sdpg.h (L+2 uint32 words per element); helpers here
convert packed arrays to and from exact Python Fractions for tests.

The library must have been built (build.py); there is no fallback: every entry
point raises if the shared library or a HIP device is missing.
"""
from __future__ import annotations

import ctypes
import os
from fractions import Fraction

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")

SDPG_X, SDPG_Z, SDPG_INVZ, SDPG_Y, SDPG_RP, SDPG_RD, SDPG_GVEC = 1, 2, 3, 4, 5, 6, 7
SDPG_B, SDPG_L, SDPG_DY, SDPG_DZ, SDPG_DX, SDPG_RZINV = 8, 9, 10, 11, 12, 13
SDPG_INVCHOLX, SDPG_INVCHOLZ = 14, 15
SDPG_X0, SDPG_Z0 = 16, 17
(SDPG_IP_XZ, SDPG_IP_XDZ, SDPG_IP_DXZ, SDPG_IP_DXDZ, SDPG_IP_CX, SDPG_IP_CDX, SDPG_IP_BY, SDPG_IP_BDY,
 SDPG_IP_TRX, SDPG_IP_TRZ, SDPG_IP_X0Z, SDPG_IP_XZ0) = range(1, 13)
GEMM_ELEM, GEMM_PLAIN, GEMM_OZAKI = -1, 0, 1
SDPG_NOINFO, SDPG_PFEAS, SDPG_DFEAS, SDPG_PDFEAS = 0, 1, 2, 3
BATCH_NONE, BATCH_PREDICTOR, BATCH_CORRECTOR = 0, 1, 2
EZERO = -(1 << 29)

# every exported symbol of include/sdpg.h and include/sdpa_amd.h
EXPORTS = [
    "sdpg_limbs", "sdpg_ctx_create", "sdpg_ctx_destroy", "sdpg_set_block", "sdpg_set_lp", "sdpg_set_b",
    "sdpg_set_formula", "sdpg_commit", "sdpg_set_gemm", "sdpg_set_point", "sdpg_residuals", "sdpg_gvec",
    "sdpg_schur", "sdpg_potrf", "sdpg_potrs", "sdpg_dz", "sdpg_dx", "sdpg_get", "sdpg_put", "sdpg_timers",
    "sdpg_reset_timers", "sdpg_sync", "sdpg_last_error", "sdpg_rgemm", "sdpg_update", "sdpg_factor_point",
    "sdpg_inner", "sdpg_min_eigen", "sdpg_profile", "sdpg_rgemm_noalloc",
    "sdpg_tune_slices", "sdpg_valu_peak", "sdpg_batch_create", "sdpg_batch_destroy", "sdpg_batch_newton",
    "sdpg_batch_set_graph", "sdpg_batch_time", "sdpg_batch_gvec", "sdpg_batch_dz", "sdpg_batch_dx",
    "sdpg_batch_update", "sdpg_batch_factor_point", "sdpg_batch_inner", "sdpg_batch_residuals",
    "sdpg_batch_min_eigen", "sdpg_batch_wait_contexts", "sdpg_batch_timers",
    "sdpa_amd_solve", "sdpa_amd_bench", "sdpa_amd_ctx_from_file", "sdpa_amd_formula", "sdpa_amd_run",
    "sdpa_amd_batch_solve",
]


class SdpgError(RuntimeError):
    def __init__(self, msg, rc=None):
        super().__init__(msg)
        self.rc = rc


_lib = None


def load(path: str | None = None):
    """Load libsdpg.so (raises if it is not built); SDPG_LIB names another build of it."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("SDPG_LIB") or os.path.join(LIBDIR, "libsdpg.so")
    if not os.path.exists(path):
        raise SdpgError(f"{path} not built; run sdpa-gmp_amd/build.py")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u32p = P(ctypes.c_uint32)
    lib.sdpg_last_error.restype = ctypes.c_char_p
    lib.sdpg_limbs.argtypes = [ctypes.c_int]
    lib.sdpg_ctx_destroy.restype = None
    lib.sdpg_ctx_destroy.argtypes = [ctypes.c_void_p]
    for name in ("sdpg_schur", "sdpg_potrs", "sdpg_dx", "sdpg_sync", "sdpg_reset_timers", "sdpg_commit"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.sdpg_potrf.argtypes = [ctypes.c_void_p, P(ctypes.c_int)]
    lib.sdpg_dz.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.sdpg_gvec.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p, u32p, ctypes.c_int]
    lib.sdpg_set_point.argtypes = [ctypes.c_void_p, u32p, u32p, u32p, u32p]
    lib.sdpg_residuals.argtypes = [ctypes.c_void_p, u32p, u32p, u32p, u32p]
    lib.sdpg_get.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p]
    lib.sdpg_put.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p]
    lib.sdpg_set_gemm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.sdpg_timers.argtypes = [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int]
    lib.sdpg_update.argtypes = [ctypes.c_void_p, u32p, u32p]
    lib.sdpg_factor_point.argtypes = [ctypes.c_void_p, P(ctypes.c_int)]
    lib.sdpg_inner.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p]
    lib.sdpg_min_eigen.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p]
    lib.sdpg_rgemm.argtypes = [ctypes.c_int] * 7 + [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p] + \
        [ctypes.c_int] * 5 + [P(ctypes.c_double)]
    lib.sdpg_rgemm_noalloc.argtypes = lib.sdpg_rgemm.argtypes
    lib.sdpg_valu_peak.argtypes = [ctypes.c_int, P(ctypes.c_double)]
    lib.sdpg_tune_slices.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double),
                                     ctypes.c_int]
    lib.sdpa_amd_ctx_from_file.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           P(ctypes.c_void_p), P(ctypes.c_int), P(ctypes.c_longlong),
                                           P(ctypes.c_int)]
    lib.sdpa_amd_solve.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.sdpa_amd_bench.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                                   ctypes.c_char_p]
    lib.sdpg_profile.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double), ctypes.c_int]
    lib.sdpa_amd_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.sdpa_amd_batch_solve.argtypes = [P(ctypes.c_char_p), ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p, P(ctypes.c_double),
                                         P(ctypes.c_double)]
    lib.sdpg_batch_create.argtypes = [P(ctypes.c_void_p), ctypes.c_int, P(ctypes.c_void_p)]
    lib.sdpg_batch_destroy.restype = None
    lib.sdpg_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.sdpg_batch_newton.argtypes = [ctypes.c_void_p, P(ctypes.c_int), P(ctypes.c_int)]
    lib.sdpg_batch_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.sdpg_batch_time.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
    ip = P(ctypes.c_int)
    lib.sdpg_batch_gvec.argtypes = [ctypes.c_void_p, ip, u32p, u32p, ip]
    lib.sdpg_batch_dz.argtypes = [ctypes.c_void_p, ip, ip]
    lib.sdpg_batch_dx.argtypes = [ctypes.c_void_p, ip]
    lib.sdpg_batch_update.argtypes = [ctypes.c_void_p, ip, u32p, u32p]
    lib.sdpg_batch_factor_point.argtypes = [ctypes.c_void_p, ip, ip]
    lib.sdpg_batch_inner.argtypes = [ctypes.c_void_p, ip, u32p]
    lib.sdpg_batch_residuals.argtypes = [ctypes.c_void_p, ip, u32p]
    lib.sdpg_batch_min_eigen.argtypes = [ctypes.c_void_p, ip, ctypes.c_int, u32p]
    lib.sdpg_batch_wait_contexts.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.sdpg_batch_timers.argtypes = [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int]
    lib.sdpa_amd_formula.argtypes = [ctypes.c_char_p, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int), ctypes.c_int]
    _lib = lib
    return lib


def limbs_for_bits(bits: int) -> int:
    """32-bit limbs for a precision request: GMP's mpf prec = (max(53,p)+127)//64 64-bit limbs."""
    b = max(53, bits)
    return 2 * ((b + 127) // 64)


# ----------------------------------------------------------------- packed <-> exact

def frac_to_words(x: Fraction, L: int) -> list:
    """Round x to nearest (ties away) into packed words [e, s, m0..m_{L-1}]."""
    if x == 0:
        return [EZERO & 0xFFFFFFFF, 0] + [0] * L
    s = 1 if x < 0 else 0
    x = abs(x)
    num, den = x.numerator, x.denominator
    e = num.bit_length() - den.bit_length()
    if Fraction(1 << e if e >= 0 else 1, 1 if e >= 0 else 1 << -e) <= x:
        e += 1
    # now 2^(e-1) <= x < 2^e  (adjust once more if needed)
    while (Fraction(2) ** (e - 1)) > x:
        e -= 1
    while (Fraction(2) ** e) <= x:
        e += 1
    sh = 32 * L - e
    q = (num << sh) // den if sh >= 0 else num // (den << -sh)
    r2 = ((num << (sh + 1)) // den if sh + 1 >= 0 else num // (den << -(sh + 1))) & 1
    M = q + r2
    if M >> (32 * L):
        M >>= 1
        e += 1
    return [e & 0xFFFFFFFF, s] + [(M >> (32 * k)) & 0xFFFFFFFF for k in range(L)]


def words_to_frac(w, L: int) -> Fraction:
    M = 0
    for k in range(L - 1, -1, -1):
        M = (M << 32) | int(w[2 + k])
    if M == 0:
        return Fraction(0)
    e = int(np.int32(np.uint32(w[0])))
    sh = e - 32 * L
    v = Fraction(M << sh) if sh >= 0 else Fraction(M, 1 << -sh)
    return -v if int(w[1]) & 1 else v


def pack(values, L: int) -> np.ndarray:
    """Sequence of Fractions -> packed uint32 array of shape (n, L+2)."""
    out = np.zeros((len(values), L + 2), dtype=np.uint32)
    for i, v in enumerate(values):
        out[i] = frac_to_words(Fraction(v), L)
    return out


def unpack(arr: np.ndarray, L: int) -> list:
    arr = np.asarray(arr, dtype=np.uint32).reshape(-1, L + 2)
    return [words_to_frac(row, L) for row in arr]


def repack(arr: np.ndarray, L_from: int, L_to: int) -> np.ndarray:
    """Exact value in L_from limbs -> rounded into L_to limbs."""
    return pack(unpack(arr, L_from), L_to)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


# ----------------------------------------------------------------- Rgemm

def rgemm(prec_bits: int, A: np.ndarray, B: np.ndarray, C: np.ndarray | None = None, transa: bool = False,
          transb: bool = False, op: int = 0, gemm: int = GEMM_PLAIN, slices: int = 0, lower: bool = False,
          device: int = 0, noalloc: bool = False):
    """C op= op(A) op(B) on the device.  A, B, C: packed arrays of shape (rows, cols, L+2), column-major
    in the sense of Rgemm (element (i, j) = X[i, j]).  Returns (C, device milliseconds).  noalloc: through
    sdpg_rgemm_noalloc (static device scratch, no allocation)."""
    lib = load()
    L = limbs_for_bits(prec_bits)
    ra, ca = A.shape[:2]
    rb, cb = B.shape[:2]
    m, k = (ca, ra) if transa else (ra, ca)
    k2, n = (cb, rb) if transb else (rb, cb)
    if k != k2:
        raise ValueError("inner dimensions differ")
    colmaj = lambda X: np.ascontiguousarray(np.transpose(X, (1, 0, 2))).astype(np.uint32)  # noqa: E731
    a, b = colmaj(A), colmaj(B)
    if C is None:
        C = np.zeros((m, n, L + 2), dtype=np.uint32)
        C[:, :, 0] = EZERO & 0xFFFFFFFF
    c = colmaj(C)
    ms = ctypes.c_double(0)
    fn = lib.sdpg_rgemm_noalloc if noalloc else lib.sdpg_rgemm
    rc = fn(prec_bits, device, int(transa), int(transb), m, n, k, _ptr(a), max(1, ra), _ptr(b),
            max(1, rb), _ptr(c), max(1, m), op, gemm, slices, int(lower), ctypes.byref(ms))
    if rc:
        raise SdpgError(f"{fn.__name__}: {rc}: {lib.sdpg_last_error().decode()}", rc)
    return np.ascontiguousarray(np.transpose(c, (1, 0, 2))), ms.value


def valu_peak(device: int = 0) -> float:
    """sdpg_valu_peak: measured limb multiply-adds per second of the device"""
    lib = load()
    v = ctypes.c_double()
    rc = lib.sdpg_valu_peak(device, ctypes.byref(v))
    if rc:
        raise SdpgError(f"sdpg_valu_peak: {rc}: {lib.sdpg_last_error().decode()}", rc)
    return v.value


def tune_slices(prec_bits: int, target_bits: int = 0, device: int = 0):
    """sdpg_tune_slices: (chosen slice count, [(S, effective bits, ms of a 512^3 product, ms of a
    1024 x 1024 x 32 lower update) for every slice count built])"""
    lib = load()
    S = ctypes.c_int()
    tab = (ctypes.c_double * 64)()
    rc = lib.sdpg_tune_slices(prec_bits, device, target_bits, ctypes.byref(S), tab, 64)
    if rc:
        raise SdpgError(f"sdpg_tune_slices: {rc}: {lib.sdpg_last_error().decode()}", rc)
    rows = [tuple(tab[q:q + 4]) for q in range(0, 64, 4) if tab[q] > 0]
    return S.value, rows


# ----------------------------------------------------------------- context

class Context:
    """Newton engine of one problem (the Newton object of sdpa_main.cpp:214-225)."""

    def __init__(self, data_path: str, prec_bits: int = 256, device: int = 0, gemm: int = GEMM_PLAIN):
        self.lib = load()
        self.L = self.lib.sdpg_limbs(prec_bits)
        h = ctypes.c_void_p()
        m = ctypes.c_int()
        nd = ctypes.c_longlong()
        rc = self.lib.sdpa_amd_ctx_from_file(data_path.encode(), prec_bits, device, gemm, ctypes.byref(h),
                                             ctypes.byref(m), ctypes.byref(nd), None)
        if rc:
            raise SdpgError(f"sdpa_amd_ctx_from_file rc={rc}: {self.lib.sdpg_last_error().decode()}")
        self.h, self.m, self.nd = h, m.value, nd.value

    def close(self):
        if self.h:
            self.lib.sdpg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc < 0:
            raise SdpgError(f"{what} rc={rc}: {self.lib.sdpg_last_error().decode()}")
        return rc

    def factor_point(self) -> int:
        info = ctypes.c_int()
        self._chk(self.lib.sdpg_factor_point(self.h, ctypes.byref(info)), "sdpg_factor_point")
        return info.value

    def update(self, alphaP_words, alphaD_words):
        a = np.ascontiguousarray(alphaP_words, dtype=np.uint32)
        d = np.ascontiguousarray(alphaD_words, dtype=np.uint32)
        self._chk(self.lib.sdpg_update(self.h, _ptr(a), _ptr(d)), "sdpg_update")

    def inner(self, what: int) -> np.ndarray:
        out = np.zeros(self.L + 2, dtype=np.uint32)
        self._chk(self.lib.sdpg_inner(self.h, what, _ptr(out)), "sdpg_inner")
        return out

    def min_eigen(self, which: int) -> np.ndarray:
        out = np.zeros(self.L + 2, dtype=np.uint32)
        self._chk(self.lib.sdpg_min_eigen(self.h, which, _ptr(out)), "sdpg_min_eigen")
        return out

    def size_of(self, what: int) -> int:
        if what in (SDPG_Y, SDPG_RP, SDPG_GVEC, SDPG_DY):
            return self.m
        if what in (SDPG_B, SDPG_L):  # noqa: SIM102
            return self.m * self.m
        return self.nd

    def set_gemm(self, mode: int, slices: int = 0):
        self._chk(self.lib.sdpg_set_gemm(self.h, mode, slices), "sdpg_set_gemm")

    def set_point(self, X, Z, invZ, y):
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (X, Z, invZ, y)]
        self._chk(self.lib.sdpg_set_point(self.h, *[_ptr(a) for a in arrs]), "sdpg_set_point")

    def residuals(self):
        outs = [np.zeros(self.L + 2, dtype=np.uint32) for _ in range(4)]
        self._chk(self.lib.sdpg_residuals(self.h, *[_ptr(a) for a in outs]), "sdpg_residuals")
        return outs

    def gvec(self, corrector: bool, mu_words, beta_words, phase: int):
        mu = np.ascontiguousarray(mu_words, dtype=np.uint32)
        be = np.ascontiguousarray(beta_words, dtype=np.uint32)
        self._chk(self.lib.sdpg_gvec(self.h, 1 if corrector else 0, _ptr(mu), _ptr(be), phase), "sdpg_gvec")

    def schur(self):
        self._chk(self.lib.sdpg_schur(self.h), "sdpg_schur")

    def potrf(self) -> int:
        info = ctypes.c_int()
        self._chk(self.lib.sdpg_potrf(self.h, ctypes.byref(info)), "sdpg_potrf")
        return info.value

    def potrs(self):
        self._chk(self.lib.sdpg_potrs(self.h), "sdpg_potrs")

    def dz(self, phase: int):
        self._chk(self.lib.sdpg_dz(self.h, phase), "sdpg_dz")

    def dx(self):
        self._chk(self.lib.sdpg_dx(self.h), "sdpg_dx")

    def get(self, what: int) -> np.ndarray:
        out = np.zeros((self.size_of(what), self.L + 2), dtype=np.uint32)
        self._chk(self.lib.sdpg_get(self.h, what, _ptr(out)), "sdpg_get")
        return out

    def put(self, what: int, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        self._chk(self.lib.sdpg_put(self.h, what, _ptr(a)), "sdpg_put")

    def timers(self):
        t = (ctypes.c_double * 8)()
        self.lib.sdpg_timers(self.h, t, 8)
        return list(t)


class Batch:
    """Lockstep Newton steps of many contexts (sdpg_batch_*): one request per context."""

    def __init__(self, ctxs):
        self.lib = load()
        self.ctxs = list(ctxs)  # keeps the contexts alive while the batch holds them
        arr = (ctypes.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = ctypes.c_void_p()
        rc = self.lib.sdpg_batch_create(ctypes.byref(h), len(self.ctxs), arr)
        if rc:
            raise SdpgError(f"sdpg_batch_create rc={rc}: {self.lib.sdpg_last_error().decode()}", rc)
        self.h = h

    def close(self):
        if self.h:
            self.lib.sdpg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def newton(self, req) -> np.ndarray:
        r = np.ascontiguousarray(req, dtype=np.int32)
        assert r.shape == (len(self.ctxs),)
        info = np.zeros(len(self.ctxs), dtype=np.int32)
        rc = self.lib.sdpg_batch_newton(self.h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if rc:
            raise SdpgError(f"sdpg_batch_newton rc={rc}: {self.lib.sdpg_last_error().decode()}", rc)
        return info

    def set_graph(self, on: bool):
        self.lib.sdpg_batch_set_graph(self.h, 1 if on else 0)

    def _req(self, req):
        r = np.ascontiguousarray(req, dtype=np.int32)
        assert r.shape == (len(self.ctxs),)
        return r, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def _chk(self, rc, what):
        if rc:
            raise SdpgError(f"{what} rc={rc}: {self.lib.sdpg_last_error().decode()}", rc)

    def _scalars(self, vals):
        L = self.ctxs[0].L
        a = np.ascontiguousarray(np.asarray(vals, dtype=np.uint32).reshape(len(self.ctxs), L + 2))
        return a, _ptr(a)

    def gvec(self, req, mu, beta, phases):
        r, rp = self._req(req)
        m, mp_ = self._scalars(mu)
        b, bp = self._scalars(beta)
        ph, php = self._req(phases)
        self._chk(self.lib.sdpg_batch_gvec(self.h, rp, mp_, bp, php), "sdpg_batch_gvec")

    def dz(self, req, phases):
        r, rp = self._req(req)
        ph, php = self._req(phases)
        self._chk(self.lib.sdpg_batch_dz(self.h, rp, php), "sdpg_batch_dz")

    def dx(self, req):
        r, rp = self._req(req)
        self._chk(self.lib.sdpg_batch_dx(self.h, rp), "sdpg_batch_dx")

    def update(self, req, alphaP, alphaD):
        r, rp = self._req(req)
        a, ap = self._scalars(alphaP)
        d, dp = self._scalars(alphaD)
        self._chk(self.lib.sdpg_batch_update(self.h, rp, ap, dp), "sdpg_batch_update")

    def factor_point(self, req) -> np.ndarray:
        r, rp = self._req(req)
        info = np.zeros(len(self.ctxs), dtype=np.int32)
        self._chk(self.lib.sdpg_batch_factor_point(self.h, rp, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                  "sdpg_batch_factor_point")
        return info

    def inner(self, req) -> np.ndarray:
        """(n, 10, L+2): selectors SDPG_IP_XZ .. SDPG_IP_TRZ at index selector - 1"""
        r, rp = self._req(req)
        out = np.zeros((len(self.ctxs), 10, self.ctxs[0].L + 2), dtype=np.uint32)
        self._chk(self.lib.sdpg_batch_inner(self.h, rp, _ptr(out)), "sdpg_batch_inner")
        return out

    def residuals(self, req) -> np.ndarray:
        """(n, 4, L+2): max |Rp|, max |Rd|, C.X, b.y"""
        r, rp = self._req(req)
        out = np.zeros((len(self.ctxs), 4, self.ctxs[0].L + 2), dtype=np.uint32)
        self._chk(self.lib.sdpg_batch_residuals(self.h, rp, _ptr(out)), "sdpg_batch_residuals")
        return out

    def min_eigen(self, req, which: int) -> np.ndarray:
        r, rp = self._req(req)
        out = np.zeros((len(self.ctxs), self.ctxs[0].L + 2), dtype=np.uint32)
        self._chk(self.lib.sdpg_batch_min_eigen(self.h, rp, which, _ptr(out)), "sdpg_batch_min_eigen")
        return out

    def wait_contexts(self, on: bool):
        self.lib.sdpg_batch_wait_contexts(self.h, 1 if on else 0)

    def timers(self) -> list:
        t = (ctypes.c_double * 9)()
        self.lib.sdpg_batch_timers(self.h, t, 9)
        return list(t)

    def time_ms(self) -> float:
        t = ctypes.c_double()
        self.lib.sdpg_batch_time(self.h, ctypes.byref(t))
        return t.value


class Result(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int), ("phase", ctypes.c_char * 16),
                ("objP", ctypes.c_char * 96), ("objD", ctypes.c_char * 96), ("relgap", ctypes.c_char * 48),
                ("pfeas", ctypes.c_char * 48), ("dfeas", ctypes.c_char * 48), ("mu", ctypes.c_double),
                ("t_schur_ms", ctypes.c_double), ("t_chol_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double),
                ("t_gvec_ms", ctypes.c_double), ("t_dxdz_ms", ctypes.c_double), ("t_res_ms", ctypes.c_double),
                ("t_total_s", ctypes.c_double), ("limbs", ctypes.c_int), ("t_inv_ms", ctypes.c_double),
                ("t_eig_ms", ctypes.c_double), ("objP70", ctypes.c_char * 112), ("objD70", ctypes.c_char * 112),
                ("slices", ctypes.c_int)]


def solve(data_path: str, param_path: str | None = None, prec: int = 0, log_path: str | None = None,
          gemm: int = GEMM_PLAIN, slices: int = 0, device: int = 0) -> Result:
    """Whole SDPA solve (pinpal) with the Newton system on the GPU."""
    lib = load()
    r = Result()
    rc = lib.sdpa_amd_solve(data_path.encode(), param_path.encode() if param_path else None, prec,
                            log_path.encode() if log_path else None, gemm, slices, device, ctypes.byref(r))
    if rc:
        raise SdpgError(f"sdpa_amd_solve rc={rc}: {lib.sdpg_last_error().decode()}")
    return r


def batch_solve(data_paths, param_path: str | None = None, prec: int = 0, device: int = 0, graph: bool = True):
    """Lockstep solves of many problems with batched Newton steps (sdpa_amd_batch_solve): returns
    (list of Result, latencies in seconds, stats dict)."""
    lib = load()
    n = len(data_paths)
    paths = (ctypes.c_char_p * n)(*[p.encode() for p in data_paths])
    res = (Result * n)()
    lat = (ctypes.c_double * n)()
    st = (ctypes.c_double * 12)()
    rc = lib.sdpa_amd_batch_solve(paths, n, param_path.encode() if param_path else None, prec, device,
                                  1 if graph else 0, res, lat, st)
    if rc:
        raise SdpgError(f"sdpa_amd_batch_solve rc={rc}: {lib.sdpg_last_error().decode()}", rc)
    kinds = ("newton", "gvec", "dz", "dx", "update", "factor_point", "inner", "residuals", "min_eigen")
    stats = {"wall_s": st[0], "meetings": int(st[1]), "mean_problems_per_meeting": st[2],
             "device_ms": {k: st[3 + i] for i, k in enumerate(kinds)}}
    return list(res), list(lat), stats


def formula(data_path: str) -> np.ndarray:
    """The formula table (Newton::computeFormula_SDP) the host driver uploads, computed on the host
    without a device: array (m, n_sdp) of 1 / 2 / 3 (F1 / F2 / F3), 0 where A_i has no block."""
    lib = load()
    m, ns = ctypes.c_int(), ctypes.c_int()
    rc = lib.sdpa_amd_formula(data_path.encode(), ctypes.byref(m), ctypes.byref(ns), None, 0)
    if rc or m.value <= 0:
        raise SdpgError(f"sdpa_amd_formula rc={rc}")
    out = np.zeros(m.value * max(ns.value, 1), dtype=np.int32)
    rc = lib.sdpa_amd_formula(data_path.encode(), None, None, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              out.size)
    if rc:
        raise SdpgError(f"sdpa_amd_formula rc={rc}")
    return out.reshape(m.value, max(ns.value, 1))[:, :ns.value]


class Options(ctypes.Structure):
    _fields_ = [("data_path", ctypes.c_char_p), ("data_sparse", ctypes.c_int), ("init_path", ctypes.c_char_p),
                ("init_sparse", ctypes.c_int), ("param_path", ctypes.c_char_p), ("preset", ctypes.c_int),
                ("prec_override", ctypes.c_int), ("log_path", ctypes.c_char_p), ("gemm_mode", ctypes.c_int),
                ("slices", ctypes.c_int), ("device", ctypes.c_int)]


def run(data_path: str, *, data_sparse: int = -1, init_path: str | None = None, init_sparse: int = -1,
        param_path: str | None = None, preset: int = 0, prec: int = 0, log_path: str | None = None,
        gemm: int = GEMM_PLAIN, slices: int = 0, device: int = 0) -> Result:
    """sdpa_amd_run: a solve with the reference's command-line options (dense / sparse data and initial-point
    files, parameter presets)."""
    lib = load()
    enc = lambda x: x.encode() if x else None  # noqa: E731
    o = Options(enc(data_path), data_sparse, enc(init_path), init_sparse, enc(param_path), preset, prec,
                enc(log_path), gemm, slices, device)
    r = Result()
    rc = lib.sdpa_amd_run(ctypes.byref(o), ctypes.byref(r))
    if rc:
        raise SdpgError(f"sdpa_amd_run rc={rc}: {lib.sdpg_last_error().decode()}", rc)
    return r


def bench(data_path: str, param_path: str | None, prec: int, gemm: int, slices: int, warm_iters: int,
          warmup: int, steps: int, device: int = 0, dump_prefix: str | None = None):
    """Per-step device ms of [schur, potrf], wall ms, and the per-kernel figures of sdpg_profile
    (dict kind -> (launches, ms, work)) on the iterate after warm_iters iterations."""
    lib = load()
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    ks = (ctypes.c_double * 12)()
    rc = lib.sdpa_amd_bench(data_path.encode(), param_path.encode() if param_path else None, prec, gemm, slices,
                            device, warm_iters, warmup, steps, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ks,
                            dump_prefix.encode() if dump_prefix else None)
    if rc:
        raise SdpgError(f"sdpa_amd_bench rc={rc}: {lib.sdpg_last_error().decode()}")
    kinds = ("ozaki_mfma", "plain_limb", "chol_col", "ozaki_mfma_longk")
    kst = {k: (int(ks[3 * i]), ks[3 * i + 1], ks[3 * i + 2]) for i, k in enumerate(kinds)}
    return a.value, b.value, c.value, kst
