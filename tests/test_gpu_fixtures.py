"""The reference's 20 geometry fixtures (tests/golden/pfsf-fixtures) through
the HIP path against the oracle on the same inputs.

The fixtures are small and thin (1-voxel-deep walls and slabs, a 5x1x1 row
of anchors with nothing to solve, 2x2x1 smoke cases), so they exercise the
ragged-grid edges of every kernel.  Bars: the MG-PCG solve within 1e-5
relative L2 of the FP64 direct solve with an FP64 residual <= 1e-8; RBGS
(the fixture's tick count, at most 200 sweeps), the failure scan (on the
direct solution scaled by 1, 30 and 300, so crushing and tension codes
occur), the history update and the component labels bit for bit."""
import numpy as np
import pytest
import torch

import domains
from oracle import oracle as O
from pfsf_amd import device as D

pytestmark = pytest.mark.gpu

NAMES = domains.fixture_names()


def _dev(isl, cuda):
    return {k: torch.from_numpy(np.ascontiguousarray(getattr(isl, k))).to(cuda)
            for k in ("vtype", "kappa", "source", "rcomp", "rtens")}


@pytest.mark.parametrize("name", NAMES)
def test_fixture_solve_matches_direct(cuda, name):
    isl, _ = domains.reference_fixture(name)
    d = _dev(isl, cuda)
    s = D.Solver(d["vtype"], d["kappa"], tol=1e-8, max_iters=2000)
    x64 = torch.zeros(isl.vtype.shape, dtype=torch.float64, device=cuda)
    phi = torch.empty(isl.vtype.shape, dtype=torch.float32, device=cuda)
    info = s.solve(d["source"], x64, phi)
    s.close()
    ref = O.direct_solve(isl)
    x = x64.cpu().numpy()
    assert info.converged == 1, (name, info.iterations, info.rel_residual)
    assert O.residual(isl, x) <= 1e-8 * 1.01, name
    nref = np.linalg.norm(ref)
    if nref == 0.0:  # nothing to solve (anchors only): phi stays 0
        assert np.all(phi.cpu().numpy() == 0), name
    else:
        err = np.linalg.norm(phi.cpu().numpy() - ref) / nref
        assert err <= 1e-5, (name, err)


@pytest.mark.parametrize("name", NAMES)
def test_fixture_rbgs_bit_exact(cuda, name):
    isl, ticks = domains.reference_fixture(name)
    sweeps = min(ticks, 200)
    d = _dev(isl, cuda)
    phi = torch.zeros_like(d["kappa"])
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, sweeps)
    want = O.rbgs(isl, sweeps=sweeps)
    assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32)), name


@pytest.mark.parametrize("name", NAMES)
def test_fixture_scan_history_labels_bit_exact(cuda, name):
    isl, _ = domains.reference_fixture(name)
    d = _dev(isl, cuda)
    ref = O.direct_solve(isl).astype(np.float32)
    for scale in (1, 30, 300):
        p = (ref * np.float32(scale)).astype(np.float32)
        pg = torch.from_numpy(p).to(cuda)
        code = torch.empty_like(d["vtype"])
        D.failure_scan(d["vtype"], d["kappa"], pg, d["rcomp"], d["rtens"], code)
        assert np.array_equal(code.cpu().numpy(), O.failure_scan(isl, p)), (name, scale)
        h0 = np.full(isl.vtype.shape, np.float32(1e-9))
        H = torch.from_numpy(h0).to(cuda)
        D.history_update(d["vtype"], d["kappa"], pg, H)
        want = O.history_update(isl, p, h0)
        assert np.array_equal(H.cpu().numpy().view(np.uint32), want.view(np.uint32)), (name, scale)
    ids, rec, num, ov = D.label_components(d["vtype"], 26, 64)
    want = O.label_sv(isl.vtype, 26)[0]
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), want), name
    wrec, wnum, wov = O.label_summarise(isl.vtype, want, 64)
    assert (num, ov) == (wnum, wov), name
    assert np.array_equal(rec.cpu().numpy().view(np.uint32), wrec), name
