"""The SoA-6 operator on the GPU: its kernels bit for bit against the oracle,
and the drop-in tick on face conductivities only (no per-voxel kappa), the
way the unchanged Java caller drives it (PFSFDataBuilder.java:77-133 data
prep, normalizeSoA6, pfsf_register_island_buffers, pfsf_tick_dbb).

Bars: couplings, operator, failure scan, history and RBGS bit-identical to
the oracle; the tick's phi within 1e-5 relative L2 of the oracle's FP64
direct solve of the same SoA-6 system (FP64 residual <= 1e-6 by the
oracle), its failure records and history bit-identical to the oracle's scan
and history of the returned phi."""
import ctypes as C

import numpy as np
import pytest
import torch

import domains
from oracle import oracle as O
from pfsf_amd import capi
from pfsf_amd import device as D
from pfsf_amd import workloads as W
from pfsf_amd.engine import Engine

pytestmark = pytest.mark.gpu

NAMES = domains.fixture_names()


def _t(a, cuda):
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def _aniso_case(seed, L=12):
    """A box-union island with anisotropic faces: vertical (y) faces ten
    times the horizontal ones, per-voxel scatter, some zero faces."""
    isl = W.small_box_union(seed, L=L)
    rng = np.random.default_rng(seed)
    base = rng.uniform(0.5, 2.0, isl.vtype.shape).astype(np.float32)
    cond = np.stack([base, base, 10 * base, 10 * base, base, base]).astype(np.float32)
    cond *= rng.uniform(0.8, 1.25, cond.shape).astype(np.float32)
    cond[:, isl.vtype == 0] = 0.0
    cond[rng.integers(0, 6), rng.random(isl.vtype.shape) < 0.05] = 0.0
    return isl, np.ascontiguousarray(cond, np.float32)


def _cases():
    out = [("aniso0", *_aniso_case(0)), ("aniso5", *_aniso_case(5, L=17))]
    for name in ("l_bracket_5x5x5", "arch_7x7x3", "simple_beam_7x3x10", "smoke_2x2x1"):
        out.append((name, *domains.soa6_island(name)))
    return out


@pytest.mark.parametrize("case", range(6))
def test_soa6_kernels_bit_exact(cuda, case):
    name, isl, cond = _cases()[case]
    vt, cd = _t(isl.vtype, cuda), _t(cond, cuda)
    cw = torch.empty((13,) + isl.vtype.shape, dtype=torch.float64, device=cuda)
    D.soa6_couplings(vt, cd, cw)
    want = O.soa6_couplings(isl, cond)
    assert np.array_equal(cw.cpu().numpy().view(np.uint64), want.view(np.uint64)), name
    rng = np.random.default_rng(case)
    x = rng.standard_normal(isl.vtype.shape)
    y = torch.empty_like(_t(x, cuda))
    D.soa6_apply_f64(vt, cw, _t(x, cuda), y)
    assert np.array_equal(y.cpu().numpy().view(np.uint64),
                          O.soa6_apply(isl, want, x).view(np.uint64)), name
    ref = O.direct_solve_soa6(isl, cond).astype(np.float32)
    rc, rt = _t(isl.rcomp, cuda), _t(isl.rtens, cuda)
    for scale in (1, 30, 300):
        p = (ref * np.float32(scale)).astype(np.float32)
        code = torch.empty_like(vt)
        D.soa6_failure_scan(vt, cw, _t(p, cuda), rc, rt, code)
        assert np.array_equal(code.cpu().numpy(), O.soa6_failure_scan(isl, want, p)), (name, scale)
        h0 = np.full(isl.vtype.shape, np.float32(1e-9))
        H = _t(h0, cuda)
        D.soa6_history_update(vt, cd, _t(p, cuda), H)
        assert np.array_equal(H.cpu().numpy().view(np.uint32),
                              O.soa6_history_update(isl, cond, p, h0).view(np.uint32)), name
    phi = torch.zeros(isl.vtype.shape, dtype=torch.float32, device=cuda)
    D.soa6_rbgs(vt, cw, _t(isl.source, cuda), phi, 40)
    assert np.array_equal(phi.cpu().numpy().view(np.uint32),
                          O.soa6_rbgs(isl, want, sweeps=40).view(np.uint32)), name


def test_chebyshev_jacobi_bit_exact(cuda):
    """Chebyshev-accelerated Jacobi (schedule pfsf_chebyshev_omega) against the
    oracle's Jacobi steps with the oracle's omega schedule."""
    for isl, steps in ((W.dense_block(0, 16), 25), (W.small_box_union(3, L=14), 40)):
        rho = np.float32(0.95 * np.cos(np.pi / max(isl.lx, isl.ly, isl.lz)))
        d = D.upload(isl, cuda)
        phi = torch.zeros_like(d["kappa"])
        tmp = torch.empty_like(phi)
        D.chebyshev_jacobi(d["vtype"], d["kappa"], d["source"], phi, tmp, steps, float(rho))
        want = np.zeros(isl.vtype.shape, np.float32)
        for k in range(steps):
            want = O.jacobi_step(isl, want, O.chebyshev_omega(k, float(rho)))
        assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32))
        res0 = O.residual(isl, np.zeros(isl.vtype.shape))
        assert O.residual(isl, phi.cpu().numpy().astype(np.float64)) < res0


def _fixture_arrays(name, cuda):
    """The fixture's material grid through the device data prep
    (pfsf_hip_build_island_data) and the library's host normalisation
    (pfsf_normalize_soa6), as the Java caller prepares its buffers."""
    mi, rows, anchor = domains.fixture_material_grid(name)
    table = np.zeros(len(rows), dtype=capi.MATERIAL_DTYPE)
    for k, (dens, rc, rt, _) in enumerate(rows):
        table[k]["density"], table[k]["rcomp"], table[k]["rtens"] = dens, rc, rt
    a = D.build_island_data(_t(mi, cuda), table, None, _t(anchor, cuda))
    a = {k: v.cpu().numpy() for k, v in a.items()}
    L = capi.load()
    sm = C.c_float(0)
    arr = [np.ascontiguousarray(a[k]) for k in ("source", "rcomp", "rtens", "cond")]
    L.pfsf_normalize_soa6(*[x.ctypes.data_as(C.c_void_p) for x in arr], None, a["type"].size,
                          C.byref(sm))
    src, rc, rt, cond = arr
    isl = domains.make(a["type"], 0.0, np.where(a["type"] == 1, src, np.float32(0)), rc, rt,
                       name=name)
    return isl, cond, a["max_phi"]


def _events(isl, code, origin):
    flat = np.flatnonzero(code.ravel())
    lx, ly = isl.lx, isl.ly
    return np.stack([flat % lx + origin[0], (flat // lx) % ly + origin[1],
                     flat // (lx * ly) + origin[2], code.ravel()[flat]], 1).astype(np.int32)


def test_fixtures_tick_with_soa6_faces_only(cuda):
    """All 20 fixtures, each through data prep -> normalise -> register
    (faces only) -> pfsf_tick_dbb as its own tick."""
    for name in NAMES:
        isl, cond, max_phi = _fixture_arrays(name, cuda)
        with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
            assert e.init() == capi.OK
            org = (3, -7, 11)
            assert e.add_island(1, isl.lx, isl.ly, isl.lz, origin=org) == capi.OK
            assert e.register(1, isl, cond=cond.ravel(), max_phi=max_phi.ravel()) == capi.OK
            r, ev = e.tick_dbb([1], capacity=1 << 14)
            assert r == capi.OK, name
            r, st = e.status(1)
            assert st.operator_mode == 1 and st.last_solve.converged == 1, name
            phi = e.phi(1).reshape(isl.vtype.shape)
            ref = O.direct_solve_soa6(isl, cond)
            nref = np.linalg.norm(ref)
            if nref == 0:
                assert np.all(phi == 0), name
            else:
                err = np.linalg.norm(phi - ref) / nref
                assert err <= 1e-5, (name, err)
                assert st.last_solve.rel_residual <= 1e-6, name
            cw = O.soa6_couplings(isl, cond)
            assert np.array_equal(ev, _events(isl, O.soa6_failure_scan(isl, cw, phi), org)), name
            r, H = e.read_history(1, isl.n)
            want = O.soa6_history_update(isl, cond, phi, np.zeros(isl.vtype.shape, np.float32))
            assert np.array_equal(H.view(np.uint32), want.ravel().view(np.uint32)), name


def test_fixtures_in_one_soa6_tick(cuda):
    """The 20 fixtures as islands of one tick on face conductivities."""
    cases = [(n, *_fixture_arrays(n, cuda)) for n in NAMES]
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, (name, isl, cond, mp) in enumerate(cases):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=(20 * k, 0, 0)) == capi.OK
            assert e.register(k + 1, isl, cond=cond.ravel(), max_phi=mp.ravel()) == capi.OK
        r, ev = e.tick_dbb(list(range(1, len(cases) + 1)), capacity=1 << 15)
        assert r == capi.OK
        want = []
        for k, (name, isl, cond, mp) in enumerate(cases):
            phi = e.phi(k + 1).reshape(isl.vtype.shape)
            ref = O.direct_solve_soa6(isl, cond)
            if np.linalg.norm(ref) > 0:
                assert np.linalg.norm(phi - ref) <= 1e-5 * np.linalg.norm(ref), name
            cw = O.soa6_couplings(isl, cond)
            want.append(_events(isl, O.soa6_failure_scan(isl, cw, phi), (20 * k, 0, 0)))
        assert np.array_equal(ev, np.concatenate(want))


@pytest.mark.parametrize("seed,L", [(0, 12), (5, 17), (9, 24)])
def test_anisotropic_island_tick(cuda, seed, L):
    isl, cond = _aniso_case(seed, L)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(4, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(4, isl, cond=cond.ravel()) == capi.OK
        assert e.tick_dbb([4])[0] == capi.OK
        phi = e.phi(4).reshape(isl.vtype.shape)
        r, st = e.status(4)
        ref = O.direct_solve_soa6(isl, cond)
        err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
        assert err <= 1e-5, err
        assert st.last_solve.converged == 1 and st.last_solve.rel_residual <= 1e-6
        # the per-voxel-maximum kappa the engine used before is a different
        # operator: its solution is far from this one
        proxy = domains.make(isl.vtype, cond.max(axis=0), isl.source)
        alt = O.direct_solve(proxy)
        assert np.linalg.norm(alt - ref) > 1e-2 * np.linalg.norm(ref)


def test_sparse_scatter_updates_then_tick(cuda):
    """pfsf_get_sparse_upload_buffer / pfsf_notify_sparse_updates: 48-byte
    records scattered on the device (SH/sparse_scatter.comp.glsl:21-67),
    out-of-range indices skipped; the tick then solves the scattered arrays."""
    isl, cond = _aniso_case(2, L=10)
    n = isl.n
    rng = np.random.default_rng(3)
    rec = np.zeros((n + 3, 12), np.uint32)
    order = rng.permutation(n)
    rec[:n, 0] = order.astype(np.uint32)
    rec[:n, 1] = isl.source.ravel()[order].view(np.uint32)
    rec[:n, 2] = isl.vtype.ravel()[order]
    rec[:n, 3] = np.float32(1e9).view(np.uint32)
    rec[:n, 4] = isl.rcomp.ravel()[order].view(np.uint32)
    rec[:n, 5] = isl.rtens.ravel()[order].view(np.uint32)
    rec[:n, 6:] = cond.reshape(6, -1)[:, order].T.view(np.uint32)
    rec[n:, 0] = np.array([-1, n, n + 7], np.int64).astype(np.uint32)  # skipped
    rec[n:, 1:] = np.float32(123.0).view(np.uint32)
    base = dict(source=np.zeros(n, np.float32), type=np.zeros(n, np.uint8),
                max_phi=np.zeros(n, np.float32), rcomp=np.zeros(n, np.float32),
                rtens=np.zeros(n, np.float32), cond=np.zeros(6 * n, np.float32))
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(9, isl.lx, isl.ly, isl.lz) == capi.OK
        addr, size = C.c_void_p(), C.c_int64()
        assert e.lib.pfsf_get_sparse_upload_buffer(e.h, 9, C.byref(addr), C.byref(size)) == 0
        assert size.value == 512 * 48
        buf = (C.c_uint32 * (512 * 12)).from_address(addr.value)
        done = 0
        while done < len(rec):  # at most 512 records per notification
            k = min(512, len(rec) - done)
            C.memmove(buf, rec[done:done + k].ctypes.data, 48 * k)
            assert e.lib.pfsf_notify_sparse_updates(e.h, 9, k) == capi.OK
            done += k
        assert e.tick_dbb([9])[0] == capi.OK
        phi = np.zeros(n, np.float32)
        cnt = C.c_int32()
        assert e.lib.pfsf_read_stress(e.h, 9, phi.ctypes.data_as(C.c_void_p), n,
                                      C.byref(cnt)) == capi.OK
    got = O.sparse_scatter(rec, base)
    sc = domains.make(got["type"].reshape(isl.vtype.shape), 0.0,
                      np.where(got["type"] == 1, got["source"], 0).reshape(isl.vtype.shape),
                      got["rcomp"].reshape(isl.vtype.shape), got["rtens"].reshape(isl.vtype.shape))
    gcond = got["cond"].reshape((6,) + isl.vtype.shape)
    assert np.array_equal(gcond, cond) and np.array_equal(sc.vtype, isl.vtype)
    ref = O.direct_solve_soa6(sc, gcond)
    assert np.linalg.norm(phi.reshape(ref.shape) - ref) <= 1e-5 * np.linalg.norm(ref)


def test_soa6_islands_in_one_batched_tick(cuda):
    """Several faces-only islands of one tick are solved together (the
    batched SoA-6 PCG: every vector pass and the proxies' V-cycles over all
    islands at once).  Each island's phi equals, bit for bit, the phi of the
    same island ticked alone (the single-island SoA-6 PCG), and is within
    1e-5 of the oracle's SOA6 direct solve.  A second tick with one island's
    load doubled and the others unchanged (warm starts) mixes islands that
    iterate with islands that are done at once."""
    cases = [_aniso_case(s, L) for s, L in [(0, 12), (5, 17), (9, 24), (3, 9)]]
    ids = [1, 2, 3, 4]

    def alone(isl, cond):
        with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
            assert e.init() == capi.OK
            assert e.add_island(1, isl.lx, isl.ly, isl.lz) == capi.OK
            assert e.register(1, isl, cond=cond.ravel()) == capi.OK
            assert e.tick_dbb([1])[0] == capi.OK
            return e.phi(1).copy()

    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, (isl, cond) in zip(ids, cases):
            assert e.add_island(k, isl.lx, isl.ly, isl.lz, origin=(40 * k, 0, 0)) == capi.OK
            assert e.register(k, isl, cond=cond.ravel()) == capi.OK
        assert e.tick_dbb(ids, capacity=1 << 16)[0] == capi.OK
        for k, (isl, cond) in zip(ids, cases):
            phi = e.phi(k)
            r, st = e.status(k)
            assert st.operator_mode == 1 and st.last_solve.converged == 1, k
            assert st.last_solve.rel_residual <= 1e-6, k
            assert np.array_equal(phi.view(np.uint32), alone(isl, cond).view(np.uint32)), k
            ref = O.direct_solve_soa6(isl, cond).ravel()
            assert np.linalg.norm(phi - ref) <= 1e-5 * np.linalg.norm(ref), k
        # island 2's load doubled in its registered buffer (uploaded with
        # every tick); every island dirty again (warm starts)
        e._keep[2]["src"] *= np.float32(2)
        for k in ids:
            e.lib.pfsf_mark_full_rebuild(e.h, k)
        assert e.tick_dbb(ids, capacity=1 << 16)[0] == capi.OK
        for k, (isl, cond) in zip(ids, cases):
            r, st = e.status(k)
            assert st.last_solve.converged == 1 and st.last_solve.rel_residual <= 1e-6, k
            ref = O.direct_solve_soa6(isl, cond, b=e._keep[k]["src"]).ravel()
            assert np.linalg.norm(e.phi(k) - ref) <= 1e-5 * np.linalg.norm(ref), k
