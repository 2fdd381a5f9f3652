"""The reference-named host primitives libpfsf exports beside the tick path
(pfsf_compute.h: pfsf_normalize_soa6, pfsf_downsample_2to1;
pfsf_diagnostics.h: pfsf_chebyshev_omega, pfsf_precompute_omega_table,
pfsf_estimate_spectral_radius): bit for bit against the reference's own
sources compiled into oracle/_ref (when the reference tree is present) and
against the oracle's restatement.  CPU only: these are host functions."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from pfsf_amd import capi

REF = O.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference tree)")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


RHOS = [0.0, 0.3, 0.5, 0.9, 0.95, 0.99, 0.999, 0.9999, 1.0, 1.2, -0.7, float("nan")]


def test_chebyshev_omega_matches_the_oracle():
    L = capi.load()
    for rho in RHOS:
        for it in range(-2, 80):
            a = np.float32(L.pfsf_chebyshev_omega(it, rho))
            b = np.float32(O.chebyshev_omega(it, rho))
            assert a.tobytes() == b.tobytes(), (it, rho)


@needs_ref
def test_chebyshev_schedule_bit_exact_against_reference():
    L = capi.load()
    for rho in RHOS + [float(x) for x in np.random.default_rng(0).uniform(0, 1, 50)]:
        for it in range(-2, 80):
            a = np.float32(L.pfsf_chebyshev_omega(it, rho))
            b = np.float32(REF.pfsf_chebyshev_omega(it, rho))
            assert a.tobytes() == b.tobytes(), (it, rho)
        for cap in (1, 2, 3, 64, 200):
            got = np.zeros(cap, np.float32)
            want = np.zeros(cap, np.float32)
            assert L.pfsf_precompute_omega_table(rho, _p(got), cap) == cap
            assert REF.pfsf_precompute_omega_table(rho, _p(want), cap) == cap
            assert got.tobytes() == want.tobytes(), (rho, cap)
    assert L.pfsf_precompute_omega_table(0.5, None, 4) == capi.INVALID_ARG
    assert L.pfsf_precompute_omega_table(0.5, _p(np.zeros(1, np.float32)), 0) == capi.INVALID_ARG
    for lmax in range(-1, 300, 7):
        for m in (0.95, 1.0, 0.5):
            a = np.float32(L.pfsf_estimate_spectral_radius(lmax, m))
            b = np.float32(REF.pfsf_estimate_spectral_radius(lmax, m))
            assert a.tobytes() == b.tobytes()


def _norm_case(rng, n, scale):
    return [rng.uniform(0, scale, k * n).astype(np.float32) for k in (1, 1, 1, 6)]


@pytest.mark.parametrize("scale", [0.5, 1.0, 37.0, 1e6])
def test_normalize_soa6(scale):
    L = capi.load()
    rng = np.random.default_rng(int(scale))
    n = 1000
    src, rc, rt, cond = _norm_case(rng, n, scale)
    cond[17] = np.nan  # ignored by the maximum, scaled like any other entry
    got = [a.copy() for a in (src, rc, rt, cond)]
    sm = C.c_float(0)
    L.pfsf_normalize_soa6(*[_p(a) for a in got], None, n, C.byref(sm))
    s2, r2, t2, c2, smax = O.normalize_soa6(src, rc, rt, cond)
    assert np.float32(sm.value) == np.float32(smax)
    for a, b in zip(got, (s2, r2, t2, c2)):
        assert a.tobytes() == b.tobytes()
    if REF is not None:
        want = [a.copy() for a in (src, rc, rt, cond)]
        sr = C.c_float(0)
        REF.pfsf_normalize_soa6(*[_p(a) for a in want], None, n, C.byref(sr))
        assert sr.value == sm.value
        for a, b in zip(got, want):
            assert a.tobytes() == b.tobytes()
    # empty or NULL input: nothing scaled, sigma max 1
    sm = C.c_float(0)
    L.pfsf_normalize_soa6(None, None, None, None, None, 5, C.byref(sm))
    assert sm.value == 1.0
    L.pfsf_normalize_soa6(*[_p(a) for a in got], None, 0, C.byref(sm))
    assert sm.value == 1.0


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 2, 2), (3, 5, 7), (8, 1, 9), (17, 16, 15)])
@pytest.mark.parametrize("with_type", [True, False])
def test_downsample_2to1(shape, with_type):
    L = capi.load()
    rng = np.random.default_rng(sum(shape))
    lz, ly, lx = shape
    fine = rng.standard_normal(shape).astype(np.float32)
    ft = rng.integers(0, 4, shape).astype(np.uint8) if with_type else None  # 3: unknown = air
    cs = ((lz + 1) // 2, (ly + 1) // 2, (lx + 1) // 2)
    got, gt = np.zeros(cs, np.float32), np.zeros(cs, np.uint8)
    L.pfsf_downsample_2to1(_p(fine), _p(ft) if with_type else None, lx, ly, lz, _p(got), _p(gt))
    if with_type:
        wc, wt = O.downsample_2to1(fine, ft, lx, ly, lz)
        assert got.tobytes() == wc.tobytes() and np.array_equal(gt, wt)
    if REF is not None:
        want, wt = np.zeros(cs, np.float32), np.zeros(cs, np.uint8)
        REF.pfsf_downsample_2to1(_p(fine), _p(ft) if with_type else None, lx, ly, lz,
                                 _p(want), _p(wt))
        assert got.tobytes() == want.tobytes() and np.array_equal(gt, wt)
    L.pfsf_downsample_2to1(None, None, lx, ly, lz, _p(got), None)  # NULL input: no-op
