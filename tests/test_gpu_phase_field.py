"""Phase-field damage evolution on the GPU (pfsf_hip_phase_field_evolve and
the engine tick) against the oracle's restatement of
SH/phase_field_evolve.comp.glsl:50-168: damage bit for bit over repeated
steps, and the CANTILEVER(1) records the tick reports once damage passes
0.95."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pfsf_amd import capi
from pfsf_amd import device as D
from pfsf_amd import workloads as W
from pfsf_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _t(a, cuda):
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


@pytest.mark.parametrize("seed", [0, 1])
def test_phase_field_steps_bit_exact(cuda, seed):
    import ctypes as C

    isl = W.small_box_union(seed, L=14)
    rng = np.random.default_rng(seed)
    sh = isl.vtype.shape
    cond = rng.uniform(0.1, 3.0, (6,) + sh).astype(np.float32)
    phi = rng.uniform(0, 5, sh).astype(np.float32)
    phi.ravel()[rng.integers(0, phi.size, 20)] = np.nan
    H = rng.uniform(-10, 3000, sh).astype(np.float32)
    hy = rng.uniform(-0.5, 1.5, sh).astype(np.float32)
    d = np.zeros(sh, np.float32)
    code = np.where(rng.random(sh) < 0.05, 2, 0).astype(np.uint8)
    L = capi.load()
    g = {k: _t(v, cuda) for k, v in (("vt", isl.vtype), ("cond", cond), ("phi", phi), ("H", H),
                                      ("hy", hy), ("code", code), ("d0", d))}
    g["d1"] = torch.empty_like(g["d0"])
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    for step in range(12):
        r = L.pfsf_hip_phase_field_evolve(
            capi.Dims(isl.lx, isl.ly, isl.lz), p(g["vt"]), p(g["cond"]), p(g["phi"]), p(g["H"]),
            p(g["hy"]) if step % 2 == 0 else None, p(g["code"]), p(g["d0"]), p(g["d1"]),
            float(O.PF_L0), float(O.PF_GC), float(O.PF_RELAX), 1,
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert r == capi.OK
        d, frac = O.phase_field(isl, cond, phi, H, d, code, hydration=hy if step % 2 == 0 else None)
        code = np.where(frac != 0, np.uint8(1), code)
        assert np.array_equal(g["d1"].cpu().numpy().view(np.uint32), d.view(np.uint32)), step
        assert np.array_equal(g["code"].cpu().numpy(), code), step
        g["d0"], g["d1"] = g["d1"], g["d0"]
    assert (code == 1).any()


def test_engine_ticks_report_cantilever_once_damage_passes_095(cuda):
    """Ticks of one island under a heavy load: each tick re-solves, updates
    H, steps the damage (voxels failed last tick frozen), scans and marks
    CANTILEVER where the scan finds nothing and d > 0.95; damage and records
    follow the oracle's replay of the same sequence on the tick's phi."""
    isl = W.small_box_union(5, L=12)
    isl.source[...] *= np.float32(3e4)
    isl.rcomp[...] = np.where(isl.vtype != 0, np.float32(1e12), np.float32(0))
    isl.rtens[...] = isl.rcomp
    cond = np.tile(isl.kappa.ravel(), 6).astype(np.float32)
    c6 = cond.reshape((6,) + isl.vtype.shape)
    H = np.zeros(isl.vtype.shape, np.float32)
    d = np.zeros(isl.vtype.shape, np.float32)
    code = np.zeros(isl.vtype.shape, np.uint8)
    seen = 0
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(1, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(1, isl, cond=cond, kappa=isl.kappa) == capi.OK
        for tick in range(14):
            e.lib.pfsf_mark_full_rebuild(e.h, 1)
            r, ev = e.tick_dbb([1], capacity=1 << 14)
            assert r == capi.OK
            phi = e.phi(1).reshape(isl.vtype.shape).copy()
            H = O.history_update(isl, phi, H)
            d, frac = O.phase_field(isl, c6, phi, H, d, code)
            scan = O.failure_scan(isl, phi)
            code = np.where((scan == 0) & (isl.vtype == 1) & (d > np.float32(0.95)), np.uint8(1),
                            scan)
            import ctypes as C
            got = np.zeros(isl.n, np.float32)
            cnt = C.c_int32()
            assert e.lib.pfsf_read_damage(e.h, 1, got.ctypes.data_as(C.c_void_p), isl.n,
                                          C.byref(cnt)) == capi.OK
            assert np.array_equal(got.view(np.uint32), d.ravel().view(np.uint32)), tick
            flat = np.flatnonzero(code.ravel())
            want = np.stack([flat % isl.lx, (flat // isl.lx) % isl.ly,
                             flat // (isl.lx * isl.ly), code.ravel()[flat]], 1).astype(np.int32)
            assert np.array_equal(ev, want), tick
            seen = max(seen, int((ev[:, 3] == 1).sum()) if len(ev) else 0)
    assert seen > 0
