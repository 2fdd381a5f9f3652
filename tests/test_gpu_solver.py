"""The MG-preconditioned PCG solve on the GPU against the FP64 oracle.

Bar: the FP64 residual ||b - A x|| / ||b|| (recomputed by the oracle from the
returned x) <= 1e-6, and phi within 1e-5 relative L2 of the FP64 direct solve
where the direct solve is affordable."""
import dataclasses
import os

import numpy as np
import pytest
import torch

import domains
from oracle import oracle as O
from pfsf_amd import capi
from pfsf_amd import device as D
from pfsf_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL_PHI = 1e-5  # relative L2 of phi vs the FP64 oracle


def _solve(isl, cuda, **kw):
    vt = torch.from_numpy(isl.vtype).to(cuda)
    kp = torch.from_numpy(isl.kappa).to(cuda)
    b = torch.from_numpy(isl.source).to(cuda)
    s = D.Solver(vt, kp, **kw)
    x64 = torch.zeros(isl.vtype.shape, dtype=torch.float64, device=cuda)
    phi = torch.empty(isl.vtype.shape, dtype=torch.float32, device=cuda)
    info = s.solve(b, x64, phi)
    s.close()
    return info, x64.cpu().numpy(), phi.cpu().numpy()


@pytest.mark.parametrize("levels", [1, 2, 0])
@pytest.mark.parametrize("seed,L", [(0, 12), (1, 16), (2, 20), (4, 17)])
def test_small_islands_match_direct_solve(cuda, seed, L, levels):
    isl = W.small_box_union(seed, L=L)
    info, x, phi = _solve(isl, cuda, mg_levels=levels, tol=1e-8, max_iters=4000)
    ref = O.direct_solve(isl)
    assert info.converged == 1, (info.iterations, info.rel_residual)
    assert O.residual(isl, x) <= 1e-8 * 1.01
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    assert err <= TOL_PHI, err


def test_column_closed_form(cuda):
    L, k, q = 40, 2.0, 0.5
    isl = W.column(L, kappa=k, load=q)
    info, x, phi = _solve(isl, cuda, tol=1e-10, max_iters=2000)
    z = np.arange(1, L)
    exact = (np.float32(q) / k) * (z * L - z * (z + 1) / 2)
    assert np.allclose(x.ravel()[1:], exact, rtol=1e-8)


def test_orphans_and_isolated_voxels_are_excluded(cuda):
    isl = domains.pillar_and_cluster()
    info, x, phi = _solve(isl, cuda, tol=1e-9)
    assert info.converged == 1
    assert np.all(phi[3:6, 6, 6] == 0)  # detached 3-voxel column: not solved
    assert phi[1, 1, 1] > 0 and phi[2, 1, 1] > phi[1, 1, 1]


@pytest.mark.parametrize("levels", [2, 0])
def test_zero_kappa_anchors_and_voxels(cuda, levels):
    """Anchors of kappa 0 still couple through the face weight (kappa_i + 0) / 2
    (the streaming kernels code them as -FLT_MIN, not as air), and free
    voxels of kappa 0 couple through faces only: the solve matches the FP64
    direct solve of that system (which differs from the unmodified island's
    by ~150 % relative L2)."""
    isl = W.small_box_union(2, L=16)
    k = isl.kappa.copy()
    k[isl.vtype == 2] = 0.0
    rng = np.random.default_rng(7)
    solid = np.argwhere(isl.vtype == 1)
    pick = solid[rng.choice(len(solid), size=len(solid) // 10, replace=False)]
    k[tuple(pick.T)] = 0.0
    isl = dataclasses.replace(isl, kappa=k)
    info, x, phi = _solve(isl, cuda, mg_levels=levels, tol=1e-8, max_iters=4000)
    ref = O.direct_solve(isl)
    assert info.converged == 1, (info.iterations, info.rel_residual)
    assert O.residual(isl, x) <= 1e-8 * 1.01
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    assert err <= TOL_PHI, err


@pytest.mark.parametrize("prec", [capi.PREC_FP32, capi.PREC_BF16_COARSE])
def test_box_union_64_to_1e6(cuda, prec):
    isl = W.box_union(1, 64, 0.6, 0.5, n_boxes=100000, edge=(2, 20))
    info, x, phi = _solve(isl, cuda, precision=prec, tol=1e-6, max_iters=2000)
    assert info.converged == 1, (info.iterations, info.refinements, info.levels,
                                 info.rel_residual)
    rel = O.residual(isl, x)
    assert rel <= 1e-6 * 1.01, rel
    ref, it, rr = O.pcg(isl, tol=1e-11)
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    assert err <= TOL_PHI, err


def test_workload_128_to_1e6(cuda):
    """The 128^3 box-union workload: FP64 residual recomputed on the CPU."""
    isl = W.workload(1)
    info, x, phi = _solve(isl, cuda, tol=1e-6, max_iters=2000)
    print(f"128^3: iters={info.iterations} passes={info.refinements} levels={info.levels} "
          f"rel={info.rel_residual:.3e} ms={info.solve_ms:.2f}")
    assert info.converged == 1
    rel = O.residual(isl, x)
    assert rel <= 1e-6 * 1.01, rel
    # phi is the FP32 rounding of x
    assert np.array_equal(phi, x.astype(np.float32))
    # iteration budget of the V-cycle with connectivity-masked coarse levels
    # and the [lmax/100, lmax] coarsest Chebyshev interval
    assert info.iterations <= 44, info.iterations
    # phi against the oracle's FP64 Jacobi-PCG solution to 1e-10 (all host
    # threads; the same iterates on any thread count)
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))
    try:
        ref, it, rr = O.pcg(isl, tol=1e-10, maxit=20000)
    finally:
        O.set_threads(1)
    assert rr <= 1e-9, rr
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    print(f"128^3: phi rel-L2 vs FP64 oracle solution {err:.3e} ({it} oracle iterations)")
    assert err <= 1e-5, err


def test_strided_column_walk_on_a_ragged_grid(cuda):
    """150^3 (~1,750 finest columns, over the one-column-per-work-group
    limit of 1,024): the finest kernels walk the column list in strides,
    with columns cut by the grid in x, y and z and an empty slot in the
    XCD-grouped list.  The 1e-6 solve meets the FP64 residual recomputed by
    the oracle and agrees with a 1e-10 solve of the same system."""
    isl = W.box_union(1150, 150, 0.6, 0.5, n_boxes=100000, edge=(2, 30))
    info, x, phi = _solve(isl, cuda, tol=1e-6, max_iters=2000)
    assert info.converged == 1, (info.iterations, info.rel_residual)
    assert O.residual(isl, x) <= 1e-6 * 1.01
    ref_info, xr, pr = _solve(isl, cuda, tol=1e-10, max_iters=4000)
    assert ref_info.converged == 1 and O.residual(isl, xr) <= 1e-10 * 1.01
    err = np.linalg.norm(phi - pr) / np.linalg.norm(pr)
    assert err <= TOL_PHI, err


def test_workload_256_mixed_precision(cuda):
    """256^3, kappa lognormal sigma 1.5, ~70 % occupied: FP32 and BF16-coarse
    solves reach 1e-6 (FP64 residual recomputed on the CPU); phi agrees with a
    1e-11 solve of the same system and the failure codes of that solution,
    computed by the oracle's scan, match the GPU scan of the 1e-6 solutions
    on all but a handful of voxels sitting on a threshold."""
    isl = W.workload(3)
    _, xr, pr = _solve(isl, cuda, tol=1e-11, max_iters=4000)
    code_ref = O.failure_scan(isl, pr)
    for prec in (capi.PREC_FP32, capi.PREC_BF16_COARSE):
        info, x, phi = _solve(isl, cuda, precision=prec, tol=1e-6, max_iters=2000)
        assert info.converged == 1, (prec, info.iterations, info.rel_residual)
        assert O.residual(isl, x) <= 1e-6 * 1.01
        err = np.linalg.norm(phi - pr) / np.linalg.norm(pr)
        assert err <= TOL_PHI, (prec, err)
        vt, kp, ph, rc, rt = (torch.from_numpy(a).to(cuda) for a in
                              (isl.vtype, isl.kappa, phi, isl.rcomp, isl.rtens))
        code = torch.empty_like(vt)
        D.failure_scan(vt, kp, ph, rc, rt, code)
        differ = int((code.cpu().numpy() != code_ref).sum())
        print(f"256^3 prec={prec}: iters={info.iterations} passes={info.refinements} "
              f"ms={info.solve_ms:.2f} phi_err={err:.2e} codes_differ={differ}")
        assert differ <= 1e-5 * isl.n_active, differ


def test_solve_is_deterministic(cuda):
    isl = W.small_box_union(6, L=20)
    a = _solve(isl, cuda, tol=1e-7)[1]
    b = _solve(isl, cuda, tol=1e-7)[1]
    assert np.array_equal(a, b)


def test_warm_start_converged_input_needs_no_iterations(cuda):
    isl = W.small_box_union(5, L=16)
    vt = torch.from_numpy(isl.vtype).to(cuda)
    kp = torch.from_numpy(isl.kappa).to(cuda)
    b = torch.from_numpy(isl.source).to(cuda)
    s = D.Solver(vt, kp, tol=1e-7)
    x64 = torch.zeros(isl.vtype.shape, dtype=torch.float64, device=cuda)
    i1 = s.solve(b, x64)
    x1 = x64.cpu().numpy()
    i2 = s.solve(b, x64)
    # a zero guess again (its first residual is b without the stencil) must
    # repeat the cold solve exactly
    x64.zero_()
    i3 = s.solve(b, x64)
    x3 = x64.cpu().numpy()
    # no guess at all: the same arithmetic
    i4 = s.solve(b, None)
    s.close()
    assert i1.iterations > 0 and i2.iterations == 0 and i2.converged == 1
    assert i3.iterations == i1.iterations and np.array_equal(x3, x1)
    assert i4.iterations == i1.iterations and i4.rel_residual == i1.rel_residual


def test_iteration_cap_stops_the_solve(cuda):
    """The device stops a pass at max_iters; the host queue never runs the
    solve past it and reports it unconverged."""
    isl = W.workload(1)
    for cap in (1, 3, 7):
        info, x, _ = _solve(isl, cuda, tol=1e-6, max_iters=cap)
        assert info.iterations <= cap and info.converged == 0, (cap, info.iterations)
        assert np.isfinite(x).all()
