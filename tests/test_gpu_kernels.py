"""HIP kernels through the C ABI against the CPU oracle on identical inputs.

Bar: bit-exact for Jacobi/RBGS phi, failure codes, psi/H, compaction and
downsample; FP64 apply within 1e-12 relative (summation order is the same, so
it is in practice exact)."""
import numpy as np
import pytest
import torch

import domains
from oracle import oracle as O
from pfsf_amd import device as D
from pfsf_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [("dense8", lambda: W.dense_block(7, 8)),
         ("box12", lambda: W.small_box_union(0, L=12)),
         ("box16", lambda: W.small_box_union(1, L=16)),
         ("ragged", lambda: W.box_union(1100, 13, 0.5, 0.5, 40, (2, 7), True)),
         ("column", lambda: W.column(9, 2.0, 1.0)),
         ("floating", domains.floating_island)]


def _isl_dev(isl, cuda):
    return {k: torch.from_numpy(np.ascontiguousarray(getattr(isl, k))).to(cuda)
            for k in ("vtype", "kappa", "source", "rcomp", "rtens")}


@pytest.mark.parametrize("name,mk", CASES)
def test_jacobi_step_bit_exact(cuda, name, mk):
    isl = mk()
    d = _isl_dev(isl, cuda)
    rng = np.random.default_rng(3)
    prev = (rng.random(isl.vtype.shape) * 3).astype(np.float32)
    for omega in (1.0, 0.8):
        out = torch.empty_like(d["kappa"])
        D.jacobi_step(d["vtype"], d["kappa"], d["source"], torch.from_numpy(prev).to(cuda), out,
                      omega)
        want = O.jacobi_step(isl, prev, omega)
        assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32)), name


@pytest.mark.parametrize("name,mk", CASES)
def test_rbgs_bit_exact(cuda, name, mk):
    isl = mk()
    d = _isl_dev(isl, cuda)
    phi = torch.zeros_like(d["kappa"])
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, 7)
    want = O.rbgs(isl, sweeps=7)
    assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32)), name


def test_rbgs_graph_replays_bit_exact(cuda):
    """Runs of 8 sweeps replay a captured graph: 13 sweeps (one replay + 5
    direct) then 16 more on the same buffers equal 29 oracle sweeps; ten
    buffer sets of one grid (more than the graph cache keeps) each equal
    their oracle run."""
    isl = W.box_union(1100, 13, 0.5, 0.5, 40, (2, 7), True)
    d = _isl_dev(isl, cuda)
    phi = torch.zeros_like(d["kappa"])
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, 13)
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, 16)
    want = O.rbgs(isl, sweeps=29)
    assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32))
    small = W.dense_block(7, 8)
    s = _isl_dev(small, cuda)
    rng = np.random.default_rng(5)
    starts = [(rng.random(small.vtype.shape) * 2).astype(np.float32) for _ in range(10)]
    phis = [torch.from_numpy(p0).to(cuda) for p0 in starts]
    for p in phis:
        D.rbgs(s["vtype"], s["kappa"], s["source"], p, 8)
    for p, p0 in zip(phis, starts):
        want = O.rbgs(small, phi=p0, sweeps=8)
        assert np.array_equal(p.cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_dense32_rbgs_200_sweeps_and_scan_bit_exact(cuda):
    """The dense 32^3 workload: 200 RBGS sweeps then the failure scan."""
    isl = W.workload(0)
    d = _isl_dev(isl, cuda)
    phi = torch.zeros_like(d["kappa"])
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, 200)
    want = O.rbgs(isl, sweeps=200)
    got = phi.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    code = torch.empty_like(d["vtype"])
    D.failure_scan(d["vtype"], d["kappa"], phi, d["rcomp"], d["rtens"], code)
    assert np.array_equal(code.cpu().numpy(), O.failure_scan(isl, want))


def _random_island(shape, seed, air=0.2):
    """Random (lz, ly, lx) grid: air, anchors on part of z = 0, kappa with
    zeros (and non-zero values on air, which no coupling may use), a source,
    and a random starting phi everywhere (NaN on some air voxels), which the
    sweeps must skip or overwrite."""
    rng = np.random.default_rng(seed)
    vt = np.where(rng.random(shape) < air, W.AIR, W.SOLID).astype(np.uint8)
    vt[0][rng.random(shape[1:]) < 0.5] = W.ANCHOR
    zeros = np.zeros(shape, np.float32)
    isl = W._finish(vt, zeros, rng.random(shape).astype(np.float32), zeros, zeros, "random")
    kappa = rng.uniform(0.5, 40.0, shape).astype(np.float32)
    kappa[rng.random(shape) < 0.03] = 0.0
    isl.kappa = kappa
    phi0 = (rng.random(shape) * 2 - 0.5).astype(np.float32)
    phi0[(vt == W.AIR) & (rng.random(shape) < 0.3)] = np.nan
    return isl, phi0


# (lz, ly, lx) and sweeps, named for the form pfsf_hip_rbgs runs them in
RBGS_FORMS = [("planes-odd", (9, 17, 31), 11),  # planes of at most 32 x 32
              ("planes-flat", (1, 20, 20), 9),
              ("planes-thin", (5, 32, 1), 9),
              ("planes-deep", (40, 1, 32), 3),
              ("table-wide", (10, 12, 40), 9),  # planes wider than 32
              ("table-tall", (8, 33, 12), 9),
              ("table-many", (300, 16, 16), 5),  # more planes than CUs
              ("graph", (130, 128, 128), 9)]  # beyond 2^21 voxels: replays + direct


@pytest.mark.parametrize("name,shape,sweeps", RBGS_FORMS)
def test_rbgs_forms_bit_exact(cuda, name, shape, sweeps):
    isl, phi0 = _random_island(shape, sum(shape))
    d = _isl_dev(isl, cuda)
    phi = torch.from_numpy(phi0).to(cuda)
    D.rbgs(d["vtype"], d["kappa"], d["source"], phi, sweeps)
    O.set_threads(8)
    try:
        want = O.rbgs(isl, phi=phi0, sweeps=sweeps)
    finally:
        O.set_threads(1)
    assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32)), name


def test_rbgs_under_stream_capture_bit_exact(cuda):
    """On a stream under capture the sweeps are recorded as direct colour
    launches; every replay runs them again on the same buffers."""
    isl, phi0 = _random_island((6, 10, 14), 21)
    d = _isl_dev(isl, cuda)
    phi = torch.from_numpy(phi0).to(cuda)
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        D.rbgs(d["vtype"], d["kappa"], d["source"], phi, 3)
    want = phi0
    for rep in range(3):
        g.replay()
        torch.cuda.synchronize()
        want = O.rbgs(isl, phi=want, sweeps=3)
        assert np.array_equal(phi.cpu().numpy().view(np.uint32), want.view(np.uint32)), rep
        d["source"].mul_(1.0)  # eager work between replays


@pytest.mark.parametrize("name,mk", CASES)
def test_failure_scan_bit_exact(cuda, name, mk):
    isl = mk()
    rng = np.random.default_rng(11)
    phi = (rng.random(isl.vtype.shape) * 3).astype(np.float32)
    isl.rcomp[...] = rng.uniform(0.0, 2.0, phi.shape).astype(np.float32)
    isl.rtens[...] = rng.uniform(0.0, 1.0, phi.shape).astype(np.float32)
    phi.ravel()[::97] = 2e6  # some orphan-threshold hits
    d = _isl_dev(isl, cuda)
    code = torch.empty_like(d["vtype"])
    D.failure_scan(d["vtype"], d["kappa"], torch.from_numpy(phi).to(cuda), d["rcomp"], d["rtens"],
                   code)
    assert np.array_equal(code.cpu().numpy(), O.failure_scan(isl, phi)), name


@pytest.mark.parametrize("name,mk", CASES)
def test_history_bit_exact(cuda, name, mk):
    isl = mk()
    rng = np.random.default_rng(5)
    phi = (rng.random(isl.vtype.shape) * 4).astype(np.float32)
    H0 = (rng.random(isl.vtype.shape) * 2).astype(np.float32)
    d = _isl_dev(isl, cuda)
    H = torch.from_numpy(H0).to(cuda)
    psi = torch.empty_like(H)
    D.history_update(d["vtype"], d["kappa"], torch.from_numpy(phi).to(cuda), H, psi)
    assert np.array_equal(psi.cpu().numpy().view(np.uint32), O.psi(isl, phi).view(np.uint32))
    assert np.array_equal(H.cpu().numpy().view(np.uint32),
                          O.history_update(isl, phi, H0).view(np.uint32))


@pytest.mark.parametrize("name,mk", CASES)
def test_apply_f64(cuda, name, mk):
    isl = mk()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(isl.vtype.shape)
    d = _isl_dev(isl, cuda)
    y = torch.empty(isl.vtype.shape, dtype=torch.float64, device=cuda)
    D.apply_f64(d["vtype"], d["kappa"], torch.from_numpy(x).to(cuda), y)
    want = O.apply(isl, x)
    assert np.allclose(y.cpu().numpy(), want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 7, 3), (33, 17, 64), (40, 40, 40)])
def test_compaction_order_and_capacity(cuda, shape):
    rng = np.random.default_rng(sum(shape))
    code = rng.choice([0, 0, 0, 0, 2, 3, 4], size=shape).astype(np.uint8)
    flat = np.flatnonzero(code.ravel())
    lz, ly, lx = shape
    want = np.stack([flat % lx + 10, (flat // lx) % ly - 5, flat // (lx * ly) + 7,
                     code.ravel()[flat]], 1).astype(np.int32)
    for cap in (0, 1, max(1, len(flat) // 2), len(flat) + 3):
        out = torch.full((1 + 4 * max(cap, 1),), -1, dtype=torch.int32, device=cuda)
        D.compact_failures(torch.from_numpy(code).to(cuda), out, cap, origin=(10, -5, 7))
        o = out.cpu().numpy()
        n = min(cap, len(flat))
        assert o[0] == n
        assert np.array_equal(o[1:1 + 4 * n].reshape(n, 4), want[:n])


@pytest.mark.parametrize("shape", [(8, 8, 8), (7, 5, 9), (3, 1, 4), (64, 32, 16)])
def test_downsample_bit_exact(cuda, shape):
    rng = np.random.default_rng(7)
    f = rng.random(shape).astype(np.float32)
    ft = rng.integers(0, 3, shape).astype(np.uint8)
    c, ct = O.downsample_2to1(f, ft, shape[2], shape[1], shape[0])
    gc = torch.empty(c.shape, dtype=torch.float32, device=cuda)
    gt = torch.empty(ct.shape, dtype=torch.uint8, device=cuda)
    D.downsample_2to1(torch.from_numpy(f).to(cuda), torch.from_numpy(ft).to(cuda), gc, gt)
    assert np.array_equal(gc.cpu().numpy(), c) and np.array_equal(gt.cpu().numpy(), ct)


def test_normalize_and_kappa_from_soa6(cuda):
    rng = np.random.default_rng(9)
    n = 5000
    for scale in (0.5, 40.0):
        s, r, t = (rng.random(n).astype(np.float32) * 10 for _ in range(3))
        c = (rng.random(6 * n) * scale).astype(np.float32)
        want = O.normalize_soa6(s, r, t, c)
        g = [torch.from_numpy(a.copy()).to(cuda) for a in (s, r, t, c)]
        smax = D.normalize_soa6(*g)
        assert smax == want[4]
        for a, b in zip(g, want[:4]):
            assert np.array_equal(a.cpu().numpy(), b)
    cond = (rng.random(6 * n)).astype(np.float32)
    k = torch.empty(n, dtype=torch.float32, device=cuda)
    D.kappa_from_soa6(torch.from_numpy(cond).to(cuda), k)
    assert np.array_equal(k.cpu().numpy(), cond.reshape(6, n).max(0))


def test_anchored_mask_matches_bfs(cuda):
    for isl in (domains.floating_island(), domains.pillar_and_cluster(), W.small_box_union(3, 16)):
        reach = torch.empty_like(torch.from_numpy(isl.vtype)).to(cuda)
        D.anchored_mask(torch.from_numpy(isl.vtype).to(cuda), reach)
        r, _ = O.anchored(isl)
        want = (r.astype(bool) & (isl.vtype == 1)).astype(np.uint8)
        assert np.array_equal(reach.cpu().numpy(), want)


def test_workload_128_scan_and_history_bit_exact(cuda):
    """Full size: the 128^3 island solved on the GPU, then the failure scan and
    H = max(H, psi_e) of that phi on the GPU and in the oracle agree bit for
    bit (the solution carries both crushing and tension codes)."""
    isl = W.workload(1)
    d = _isl_dev(isl, cuda)
    s = D.Solver(d["vtype"], d["kappa"], tol=1e-6, max_iters=2000)
    phi = torch.empty_like(d["kappa"])
    info = s.solve(d["source"], None, phi)
    s.close()
    assert info.converged == 1
    p = phi.cpu().numpy()
    code = torch.empty_like(d["vtype"])
    D.failure_scan(d["vtype"], d["kappa"], phi, d["rcomp"], d["rtens"], code)
    got = code.cpu().numpy()
    want = O.failure_scan(isl, p)
    hist = np.bincount(want.ravel(), minlength=5)
    assert hist[2] > 0 and hist[4] > 0, hist
    assert np.array_equal(got, want)
    h0 = np.zeros(isl.vtype.shape, np.float32)
    H = torch.from_numpy(h0).to(cuda)
    D.history_update(d["vtype"], d["kappa"], phi, H)
    assert np.array_equal(H.cpu().numpy().view(np.uint32),
                          O.history_update(isl, p, h0).view(np.uint32))


def _material_table(rows):
    """capi.MATERIAL_DTYPE table and the oracle's {density, rcomp, rtens,
    indestructible} rows from (density, rcomp, rtens, indestructible)."""
    from pfsf_amd import capi

    t = np.zeros(len(rows), dtype=capi.MATERIAL_DTYPE)
    for k, (dens, rc, rt, _) in enumerate(rows):
        t[k]["density"], t[k]["rcomp"], t[k]["rtens"] = dens, rc, rt
    ind = np.array([r[3] for r in rows], np.uint8)
    return t, ind, np.array(rows, np.float32)


def _check_build(cuda, mi, rows, anchor=None, curing=None, fill=None):
    table, ind, orows = _material_table(rows)
    dev = lambda a: None if a is None else torch.from_numpy(a).to(cuda)  # noqa: E731
    got = D.build_island_data(dev(mi), table, ind, dev(anchor), dev(curing), dev(fill))
    want = O.build_island_data(mi, orows, anchor, curing, fill)
    for k in ("type", "source", "cond", "rcomp", "rtens", "max_phi"):
        g = got[k].cpu().numpy()
        w = want[k]
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), k


def test_island_data_prep_bit_exact_on_fixtures(cuda):
    """The reference fixtures' material grids and registries through the
    device data prep and the oracle: every array bit for bit, with random
    curing and fill fields."""
    import base64
    import json

    rng = np.random.default_rng(21)
    for name in domains.fixture_names():
        d = json.loads((domains.FIXTURE_DIR / f"{name}.json").read_text())
        lx, ly, lz = d["dims"]["lx"], d["dims"]["ly"], d["dims"]["lz"]
        ids = np.frombuffer(base64.b64decode(d["materials"]["voxels"]), "<i4").reshape(lz, ly, lx)
        reg = d["materials"]["registry"]
        slot = {m["id"]: k for k, m in enumerate(reg)}
        mi = np.vectorize(lambda v: slot.get(int(v), -1))(ids).astype(np.int32)
        rows = [(m["density"], m["rcomp"], m["rtens"], 0) for m in reg]
        anchor = np.zeros(mi.shape, np.uint8)
        for x, y, z in d["anchors"]:
            anchor[z, y, x] = 1
        curing = rng.uniform(-0.2, 1.2, mi.shape).astype(np.float32)
        fill = rng.uniform(0.3, 1.0, mi.shape).astype(np.float32)
        _check_build(cuda, mi, rows, anchor, curing, fill)
        _check_build(cuda, mi, rows)


def test_island_data_prep_bit_exact_on_random_grids(cuda):
    rng = np.random.default_rng(22)
    rows = [(2400, 30, 3, 0), (7850, 250, 400, 0), (600, 8, 0.5, 0), (2600, 5, 1, 1),
            (1500, 0, 0, 0), (1800, 12.5, 2.25, 0)]
    for shape in [(1, 1, 1), (3, 5, 7), (17, 9, 33), (40, 40, 40)]:
        mi = rng.integers(-1, len(rows) + 1, shape).astype(np.int32)  # -1 and len(rows) are air
        anchor = (rng.random(shape) < 0.1).astype(np.uint8)
        curing = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        _check_build(cuda, mi, rows, anchor, curing, None)


@pytest.mark.parametrize("seed,L", [(0, 20), (1, 33)])
def test_macro_residual_map_bit_exact(cuda, seed, L):
    """pfsf_hip_macro_residual: 8^3 block maxima of |b - A phi| equal the
    oracle's, on a perturbed solution and on ragged block edges."""
    isl = W.small_box_union(seed, L=L)
    rng = np.random.default_rng(seed)
    phi = (O.direct_solve(isl) + rng.normal(0, 1e-3, isl.vtype.shape)).astype(np.float32)
    d = _isl_dev(isl, cuda)
    out = torch.full(((isl.lz + 7) // 8, (isl.ly + 7) // 8, (isl.lx + 7) // 8), -1.0,
                     dtype=torch.float32, device=cuda)
    D.macro_residual(d["vtype"], d["kappa"], torch.from_numpy(phi).to(cuda), d["source"], out)
    want = O.macro_residual(isl, phi)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))
