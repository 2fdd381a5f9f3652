"""The drop-in tick path (pfsf.h) on the GPU, driven the way the Java runtime
drives it (NativePFSFRuntime.java:267-383): add island, register host
buffers (phi doubles as the stress readback), tick_dbb, decode failures."""
import dataclasses

import numpy as np
import pytest

import domains
from oracle import oracle as O
from pfsf_amd import capi
from pfsf_amd import workloads as W
from pfsf_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _expected_events(isl, phi, origin):
    code = O.failure_scan(isl, phi)
    flat = np.flatnonzero(code.ravel())
    lx, ly = isl.lx, isl.ly
    return np.stack([flat % lx + origin[0], (flat // lx) % ly + origin[1],
                     flat // (lx * ly) + origin[2], code.ravel()[flat]], 1).astype(np.int32)


def test_tick_dbb_solves_scans_and_reports(cuda):
    isl = W.small_box_union(2, L=16)
    isl.rcomp[...] = np.float32(0.4)  # low strengths: some voxels fail
    isl.rtens[...] = np.float32(0.2)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(7, isl.lx, isl.ly, isl.lz, origin=(100, -20, 64)) == capi.OK
        assert e.register(7, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb([7], capacity=4096)
        assert r == capi.OK
        phi = e.phi(7).reshape(isl.vtype.shape)  # stress readback == phi buffer
        ref = O.direct_solve(isl)
        assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= 1e-5
        want = _expected_events(isl, phi, (100, -20, 64))
        assert len(want) > 0
        assert np.array_equal(ev, want)
        r, st = e.read_stress(7, isl.n)
        assert r == capi.OK and np.array_equal(st, phi.ravel())
        # history: H = max(0, psi(phi))
        r, H = e.read_history(7, isl.n)
        assert np.array_equal(H, O.psi(isl, phi).ravel())
        # a clean island is skipped; a second tick after mark_full_rebuild warm-starts
        r, ev2 = e.tick_dbb([7])
        assert r == capi.OK and len(ev2) == 0
        e.lib.pfsf_mark_full_rebuild(e.h, 7)
        r, ev3 = e.tick_dbb([7], capacity=4096)
        assert r == capi.OK and np.array_equal(ev3, want)


def test_tick_legacy_result_and_capacity(cuda):
    isl = W.small_box_union(4, L=12)
    isl.rcomp[...] = np.float32(0.3)
    isl.rtens[...] = np.float32(0.1)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(1, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(1, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick([1], capacity=5)
        assert r == capi.OK and len(ev) == 5
        phi = e.phi(1).reshape(isl.vtype.shape)
        assert np.array_equal(ev, _expected_events(isl, phi, (0, 0, 0))[:5])


def test_faces_only_registration_solves_the_soa6_operator(cuda):
    """Without a registered kappa the island's faces define the operator
    (pfsf_hip_soa6_*): isotropic faces with some halved still solve the
    SoA-6 system, not the per-voxel maximum."""
    isl = W.small_box_union(1, L=10)
    rng = np.random.default_rng(0)
    cond = np.tile(isl.kappa.ravel(), 6).astype(np.float32)
    cond[rng.integers(0, cond.size, 50)] *= 0.5
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(3, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(3, isl, cond=cond) == capi.OK
        assert e.tick_dbb([3])[0] == capi.OK
        c6 = cond.reshape((6,) + isl.vtype.shape)
        ref = O.direct_solve_soa6(isl, c6)
        phi = e.phi(3).reshape(isl.vtype.shape)
        assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= 1e-5
        r, st = e.status(3)
        assert r == capi.OK and st.operator_mode == 1 and st.last_solve.converged == 1


def test_orphan_cluster_reports_no_support_only_for_isolated(cuda):
    isl = domains.pillar_and_cluster()
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(2, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(2, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb([2])
        assert r == capi.OK
        phi = e.phi(2).reshape(isl.vtype.shape)
        assert np.array_equal(ev, _expected_events(isl, phi, (0, 0, 0)))


def test_argument_errors(cuda):
    with Engine(max_island_size=1000) as e:
        assert e.init() == capi.OK
        assert e.add_island(1, 20, 20, 20) == capi.ISLAND_FULL
        assert e.add_island(1, 0, 2, 2) == capi.ISLAND_FULL
        assert e.add_island(1, 10, 10, 10) == capi.OK
        small = W.dense_block(1, 8)  # 512 voxels, island 1 has 1000
        # undersized buffers -> INVALID_ARG
        assert e.register(1, small) == capi.INVALID_ARG
        assert e.register(99, small) == capi.INVALID_ARG
        r, _ = e.tick_dbb([1])  # registered nothing: nothing to solve, but OK
        assert r == capi.OK
        assert e.lib.pfsf_tick_dbb(e.h, None, 1, 0, None, 0) == capi.INVALID_ARG
        assert e.read_stress(42, 10)[0] == capi.INVALID_ARG


def test_budget_zero_leaves_island_dirty(cuda):
    isl = W.small_box_union(0, L=8)
    with Engine(max_island_size=10**7, tick_budget_ms=0) as e:
        assert e.init() == capi.OK
        assert e.add_island(5, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(5, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb([5])
        assert r == capi.OK and len(ev) == 0
        assert np.all(e.phi(5) == 0)  # nothing dispatched, nothing read back


def _fresh_tick(isls, phis):
    """One tick of a new engine on the islands, warm-started from phis."""
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, (isl, phi0) in enumerate(zip(isls, phis)):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=(40 * k, 0, 0)) == capi.OK
            assert e.register(k + 1, isl, phi=phi0.copy(), kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb(list(range(1, len(isls) + 1)), capacity=1 << 14)
        assert r == capi.OK
        return [e.phi(k + 1).copy() for k in range(len(isls))], ev


@pytest.mark.parametrize("n_islands", [1, 2])
def test_reupload_keeps_or_rebuilds_the_hierarchy(cuda, n_islands):
    """Each tick re-uploads every registered host array.  With type and kappa
    unchanged an island keeps its multigrid hierarchy; a changed kappa or type
    rebuilds it.  Either way the tick equals, bit for bit, the tick of a new
    engine (which builds every hierarchy) from the same warm starts, and each
    re-solve iterates (the sources change every tick).  With two islands the
    tick takes the batched solve and only the first island changes."""
    isls = [W.small_box_union(3, L=20), W.small_box_union(8, L=14)][:n_islands]
    ids = list(range(1, n_islands + 1))
    for isl in isls:
        isl.rcomp[...] = np.float32(0.5)
        isl.rtens[...] = np.float32(0.25)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, isl in enumerate(isls):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=(40 * k, 0, 0)) == capi.OK
            assert e.register(k + 1, isl, kappa=isl.kappa) == capi.OK
        assert e.tick_dbb(ids, capacity=1 << 14)[0] == capi.OK
        keep = e._keep[1]
        rng = np.random.default_rng(5)
        solid = np.flatnonzero(keep["vt"] == 1)
        for step in ("same", "kappa", "type", "same"):
            phis = [e.phi(i).copy() for i in ids]
            for i in ids:
                e._keep[i]["src"] *= np.float32(1.25)
            if step == "kappa":
                keep["kappa"][rng.choice(solid, 40, replace=False)] *= np.float32(3.0)
            elif step == "type":
                keep["vt"][rng.choice(solid, 5, replace=False)] = 0
            cur = [dataclasses.replace(
                isl, vtype=e._keep[i]["vt"].reshape(isl.vtype.shape).copy(),
                kappa=e._keep[i]["kappa"].reshape(isl.vtype.shape).copy(),
                source=e._keep[i]["src"].reshape(isl.vtype.shape).copy())
                for i, isl in zip(ids, isls)]
            for i in ids:
                e.lib.pfsf_mark_full_rebuild(e.h, i)
            r, ev = e.tick_dbb(ids, capacity=1 << 14)
            assert r == capi.OK
            want_phis, want_ev = _fresh_tick(cur, phis)
            for i, want, phi0 in zip(ids, want_phis, phis):
                assert not np.array_equal(want, phi0), (step, i)
                assert np.array_equal(e.phi(i), want), (step, i)
            assert np.array_equal(ev, want_ev), step


def test_multi_island_tick_is_batched_and_matches_oracle(cuda):
    """Several dirty islands in one tick take the batched solve; each island's
    phi, history and failure records match the oracle as a lone tick does."""
    isls = [W.small_box_union(s, L=L) for s, L in ((11, 12), (12, 16), (13, 10), (14, 14))]
    origins = [(0, 0, 0), (40, 0, 0), (0, 40, 0), (40, 40, 8)]
    for isl in isls:
        isl.rcomp[...] = np.float32(0.4)
        isl.rtens[...] = np.float32(0.2)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, (isl, o) in enumerate(zip(isls, origins)):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=o) == capi.OK
            assert e.register(k + 1, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb([1, 2, 3, 4], capacity=1 << 14)
        assert r == capi.OK
        want = []
        for k, (isl, o) in enumerate(zip(isls, origins)):
            phi = e.phi(k + 1).reshape(isl.vtype.shape)
            ref = O.direct_solve(isl)
            assert np.linalg.norm(phi - ref) / np.linalg.norm(ref) <= 1e-5, k
            r, H = e.read_history(k + 1, isl.n)
            assert np.array_equal(H, O.psi(isl, phi).ravel())
            want.append(_expected_events(isl, phi, o))
        assert np.array_equal(ev, np.concatenate(want))


def test_island_components_match_oracle_and_follow_host_types(cuda):
    # pillar + detached cluster: one anchored component, one orphan
    isl = domains.pillar_and_cluster()
    with Engine(max_island_size=10**6, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        assert e.add_island(3, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(3, isl) == capi.OK
        r, rec, num, ov, lab = e.components(3, 64, labels=True)
        assert r == capi.OK
        want_ids, _ = O.label_sv(isl.vtype, 26)
        wrec, wnum, wov = O.label_summarise(isl.vtype, want_ids, 64)
        assert np.array_equal(lab.reshape(isl.vtype.shape), want_ids)
        assert (num, ov) == (wnum, wov) == (2, False)
        assert np.array_equal(rec, wrec)
        assert sorted(rec[:num, 2].tolist()) == [0, 1]
        # the registered host buffer is re-read: cut the pillar off the ground
        vt = e._keep[3]["vt"].reshape(isl.vtype.shape)
        vt[1, 1, 1] = 0
        r, rec2, num2, _, _ = e.components(3, 64)
        assert r == capi.OK
        t2 = vt.copy()
        wrec2, wnum2, _ = O.label_summarise(t2, O.label_sv(t2, 26)[0], 64)
        assert num2 == wnum2 and np.array_equal(rec2, wrec2)
        assert e.lib.pfsf_island_components(e.h, 99, 64, None, None, None) != capi.OK


def test_reference_fixtures_in_one_tick(cuda):
    """The reference's 20 geometry fixtures as 20 islands of one tick (the
    batched solve, including an island of anchors only): every island's phi
    matches its direct solve, its history the oracle's psi, and the failure
    records (strengths lowered so that voxels fail) the oracle's scan."""
    isls = [domains.reference_fixture(n)[0] for n in domains.fixture_names()]
    origins = [(16 * k, -8, 3 * k) for k in range(len(isls))]
    for isl in isls:
        isl.rcomp[...] = np.where(isl.vtype != 0, np.float32(0.01), np.float32(0))
        isl.rtens[...] = np.where(isl.vtype != 0, np.float32(0.003), np.float32(0))
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, (isl, o) in enumerate(zip(isls, origins)):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=o) == capi.OK
            assert e.register(k + 1, isl, kappa=isl.kappa) == capi.OK
        r, ev = e.tick_dbb(list(range(1, len(isls) + 1)), capacity=1 << 14)
        assert r == capi.OK
        want = []
        for k, (isl, o) in enumerate(zip(isls, origins)):
            phi = e.phi(k + 1).reshape(isl.vtype.shape)
            ref = O.direct_solve(isl)
            nref = np.linalg.norm(ref)
            if nref == 0.0:
                assert np.all(phi == 0), isl.name
            else:
                assert np.linalg.norm(phi - ref) / nref <= 1e-5, isl.name
            r, H = e.read_history(k + 1, isl.n)
            assert np.array_equal(H, O.psi(isl, phi).ravel()), isl.name
            want.append(_expected_events(isl, phi, o))
        want = np.concatenate(want)
        assert len(want) > 0
        assert np.array_equal(ev, want)


def _status(e, i):
    r, st = e.status(i)
    assert r == capi.OK
    return st


@pytest.mark.parametrize("n_islands", [1, 2])
def test_reupload_rebuild_counts(cuda, n_islands):
    """Re-uploads with unchanged type and kappa keep the hierarchy (no build);
    a changed kappa or type rebuilds it once, for the single and the batched
    tick alike."""
    isls = [W.small_box_union(3, L=14), W.small_box_union(8, L=12)][:n_islands]
    ids = list(range(1, n_islands + 1))
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, isl in enumerate(isls):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz) == capi.OK
            assert e.register(k + 1, isl, kappa=isl.kappa) == capi.OK
        assert e.tick_dbb(ids)[0] == capi.OK
        assert [_status(e, i).solver_builds for i in ids] == [1] * n_islands
        keep = e._keep[1]
        solid = np.flatnonzero(keep["vt"] == 1)
        want = [1] * n_islands
        for step in ("same", "kappa", "same", "type", "same"):
            if step == "kappa":
                keep["kappa"][solid[:7]] *= np.float32(2.0)
            elif step == "type":
                keep["vt"][solid[-3:]] = 0
            for i in ids:
                e.lib.pfsf_mark_full_rebuild(e.h, i)
            assert e.tick_dbb(ids)[0] == capi.OK
            if step != "same":
                want[0] += 1
            assert [_status(e, i).solver_builds for i in ids] == want, step
            assert all(_status(e, i).dirty == 0 for i in ids)


def test_new_tolerance_then_changed_kappa_matches_a_fresh_engine(cuda):
    """pfsf_set_solve_tolerance drops every solver; the next tick with a
    changed kappa builds new ones from the uploaded arrays (the build waits
    for the upload), so it equals a new engine's tick bit for bit."""
    isls = [W.small_box_union(4, L=16), W.small_box_union(6, L=13)]
    ids = [1, 2]
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        for k, isl in enumerate(isls):
            assert e.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=(40 * k, 0, 0)) == capi.OK
            assert e.register(k + 1, isl, kappa=isl.kappa) == capi.OK
        assert e.tick_dbb(ids)[0] == capi.OK
        e.set_tolerance(1e-7, 800)
        phis = [e.phi(i).copy() for i in ids]
        rng = np.random.default_rng(1)
        for i in ids:
            k = e._keep[i]["kappa"]
            k[rng.choice(np.flatnonzero(k > 0), 30, replace=False)] *= np.float32(4.0)
            e.lib.pfsf_mark_full_rebuild(e.h, i)
        r, ev = e.tick_dbb(ids, capacity=1 << 14)
        assert r == capi.OK
        cur = [dataclasses.replace(isl, kappa=e._keep[i]["kappa"].reshape(isl.vtype.shape).copy())
               for i, isl in zip(ids, isls)]
        with Engine(max_island_size=10**7, tick_budget_ms=10_000) as f:
            assert f.init() == capi.OK
            f.set_tolerance(1e-7, 800)
            for k, (isl, phi0) in enumerate(zip(cur, phis)):
                assert f.add_island(k + 1, isl.lx, isl.ly, isl.lz, origin=(40 * k, 0, 0)) == 0
                assert f.register(k + 1, isl, phi=phi0.copy(), kappa=isl.kappa) == capi.OK
            r2, ev2 = f.tick_dbb(ids, capacity=1 << 14)
            assert r2 == capi.OK
            for i in ids:
                assert np.array_equal(e.phi(i), f.phi(i)), i
            assert np.array_equal(ev, ev2)
        for i in ids:
            assert _status(e, i).last_solve.rel_residual <= 1e-7


@pytest.mark.parametrize("soa6", [False, True])
def test_non_finite_inputs_count_as_zero(cuda, soa6):
    """NaN / Inf in kappa (or a face), in the source and in the warm start:
    the tick ends, and equals the tick of the same island with those entries
    set to 0, bit for bit (phi, history and failure records)."""
    isl = W.small_box_union(7, L=12)
    isl.rcomp[...] = np.where(isl.vtype != 0, np.float32(0.3), np.float32(0))
    isl.rtens[...] = np.where(isl.vtype != 0, np.float32(0.1), np.float32(0))
    rng = np.random.default_rng(2)
    solid = np.flatnonzero(isl.vtype.ravel() == 1)
    bad = rng.choice(solid, 12, replace=False)
    cond = np.tile(isl.kappa.ravel(), 6).astype(np.float32)
    cond *= rng.uniform(0.5, 1.5, cond.size).astype(np.float32)
    phi0 = np.zeros(isl.n, np.float32)

    def run(kap, src, c, p0):
        with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
            assert e.init() == capi.OK
            assert e.add_island(1, isl.lx, isl.ly, isl.lz) == capi.OK
            i2 = dataclasses.replace(isl, source=src.reshape(isl.vtype.shape))
            assert e.register(1, i2, phi=p0.copy(), cond=c,
                              kappa=None if soa6 else kap) == capi.OK
            r, ev = e.tick_dbb([1], capacity=1 << 14)
            assert r == capi.OK
            r, H = e.read_history(1, isl.n)
            return e.phi(1).copy(), ev, H, _status(e, 1)

    kap, src, c, p0 = (isl.kappa.ravel().copy(), isl.source.ravel().copy(), cond.copy(),
                       phi0.copy())
    kap[bad[:4]] = np.nan
    kap[bad[4:6]] = np.inf
    c[bad[:3]] = np.nan
    c[5 * isl.n + bad[3:6]] = np.inf
    src[bad[6:9]] = np.nan
    src[bad[9:]] = -np.inf
    p0[bad] = np.nan
    got = run(kap, src, c, p0)
    kz, sz, cz = kap.copy(), src.copy(), c.copy()
    kz[~np.isfinite(kz)] = 0
    sz[~np.isfinite(sz)] = 0
    cz[~np.isfinite(cz)] = 0
    want = run(kz, sz, cz, np.zeros(isl.n, np.float32))
    assert np.all(np.isfinite(got[0]))
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2])
    assert got[3].last_solve.converged == 1


@pytest.mark.parametrize("soa6", [False, True])
def test_unconverged_solve_leaves_the_island_dirty(cuda, soa6):
    """A solve capped before the tolerance reports converged = 0 and keeps the
    island dirty; following ticks continue from its phi until it converges."""
    isl = W.small_box_union(2, L=20)
    with Engine(max_island_size=10**7, tick_budget_ms=10_000) as e:
        assert e.init() == capi.OK
        e.set_tolerance(1e-6, 3)
        assert e.add_island(1, isl.lx, isl.ly, isl.lz) == capi.OK
        assert e.register(1, isl, kappa=None if soa6 else isl.kappa) == capi.OK
        rels = []
        for _ in range(40):
            assert e.tick_dbb([1])[0] == capi.OK
            st = _status(e, 1)
            rels.append(st.last_solve.rel_residual)
            if not st.dirty:
                break
        assert len(rels) > 2 and rels[0] > 1e-6
        assert st.dirty == 0 and st.last_solve.converged == 1 and rels[-1] <= 1e-6
        # a clean island is not solved again
        assert e.tick_dbb([1])[0] == capi.OK and _status(e, 1).last_solve.rel_residual == rels[-1]
