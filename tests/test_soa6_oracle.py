"""The oracle's SoA-6 operator (the drop-in tick's operator over six face
conductivities per voxel), pinned without the GPU.

The reference's shader form (SH/pcg_matvec.comp.glsl:42-128) weights a pair
from voxel i's faces alone; the oracle's operator averages the two endpoints'
half weights.  Pins: hand-derived couplings (face, edge, corner, air), the
cube root against numpy, symmetry and equality with an independent numpy
assembly, reduction to the per-voxel kappa operator on uniform faces,
equality with the shader's weights wherever both endpoints agree, the
1 x 1 x L column closed form with unequal faces, the failure-scan precedence,
the anisotropic energy density on linear fields and the RBGS fixed point."""
import numpy as np
import pytest

import domains
from oracle import oracle as O
from pfsf_amd import workloads as W

AIR, SOLID, ANCHOR = 0, 1, 2


def _cond(shape, fill=0.0):
    return np.full((6,) + shape, np.float32(fill), np.float32)


def test_cbrt_within_an_ulp_and_exact_on_cubes():
    for v, r in ((8.0, 2.0), (27.0, 3.0), (0.125, 0.5), (216.0, 6.0), (64.0, 4.0), (1.0, 1.0)):
        assert O.cbrt(v) == r
    assert O.cbrt(0.0) == 0.0 and O.cbrt(-1.0) == 0.0 and O.cbrt(float("nan")) == 0.0
    vals = np.exp(np.random.default_rng(0).uniform(-300, 300, 4000))
    got = np.array([O.cbrt(v) for v in vals])
    want = np.cbrt(vals)
    assert np.all(np.abs(got - want) <= np.spacing(want))


def test_hand_couplings_face_edge_corner_and_air():
    # face: two voxels along x, c0[+x] = 4, c1[-x] = 2 -> 0.5 (4 + 2) = 3
    t = np.array([[[ANCHOR, SOLID]]], np.uint8)
    c = _cond(t.shape)
    c[1, 0, 0, 0] = 4.0   # +x of voxel 0
    c[0, 0, 0, 1] = 2.0   # -x of voxel 1
    isl = domains.make(t)
    cw = O.soa6_couplings(isl, c)
    assert cw[0, 0, 0, 0] == 3.0 and np.count_nonzero(cw) == 1
    # edge (1,1,0): voxel (0,0) half sqrt(2 * 8) = 4, voxel (1,1) half sqrt(1 * 9) = 3
    t = np.array([[[SOLID, AIR], [AIR, SOLID]]], np.uint8)
    c = _cond(t.shape)
    c[1, 0, 0, 0], c[3, 0, 0, 0] = 2.0, 8.0   # +x, +y of (0,0)
    c[0, 0, 1, 1], c[2, 0, 1, 1] = 1.0, 9.0   # -x, -y of (1,1)
    cw = O.soa6_couplings(domains.make(t), c)
    assert cw[3, 0, 0, 0] == 0.5 * 0.5 * (4.0 + 3.0)  # slot of offset k = 6
    assert np.count_nonzero(cw) == 1
    # corner (1,1,1): halves cbrt(1 * 8 * 27) = 6 and cbrt(2 * 4 * 8) = 4
    t = np.zeros((2, 2, 2), np.uint8)
    t[0, 0, 0] = t[1, 1, 1] = SOLID
    c = _cond(t.shape)
    c[1, 0, 0, 0], c[3, 0, 0, 0], c[5, 0, 0, 0] = 1.0, 8.0, 27.0
    c[0, 1, 1, 1], c[2, 1, 1, 1], c[4, 1, 1, 1] = 2.0, 4.0, 8.0
    cw = O.soa6_couplings(domains.make(t), c)
    assert cw[9, 0, 0, 0] == float(np.float32(0.1666667)) * (0.5 * 10.0)  # k = 18
    # an air endpoint has no coupling whatever its faces say
    t[1, 1, 1] = AIR
    assert np.count_nonzero(O.soa6_couplings(domains.make(t), c)) == 0
    # non-finite and negative faces count 0
    t = np.array([[[SOLID, SOLID]]], np.uint8)
    c = _cond(t.shape)
    c[1, 0, 0, 0], c[0, 0, 0, 1] = np.inf, -5.0
    assert np.count_nonzero(O.soa6_couplings(domains.make(t), c)) == 0


def _random_case(seed, L=9):
    isl = W.small_box_union(seed, L=L)
    rng = np.random.default_rng(seed)
    cond = rng.uniform(0.2, 5.0, (6,) + isl.vtype.shape).astype(np.float32)
    cond[:, rng.random(isl.vtype.shape) < 0.1] = 0.0  # some voxels uncoupled on their side
    return isl, cond


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_symmetric_and_equal_to_numpy_assembly(seed):
    isl, cond = _random_case(seed)
    cw = O.soa6_couplings(isl, cond)
    rng = np.random.default_rng(seed + 10)
    x = rng.standard_normal(isl.vtype.shape)
    y = rng.standard_normal(isl.vtype.shape)
    solid = isl.vtype == SOLID
    x, y = np.where(solid, x, 0.0), np.where(solid, y, 0.0)
    ax, ay = O.soa6_apply(isl, cw, x), O.soa6_apply(isl, cw, y)
    assert abs(np.sum(x * ay) - np.sum(y * ax)) <= 1e-11 * np.sum(np.abs(x * ay))
    A, f = O.assemble_soa6(isl, cond)
    xf = np.zeros(isl.n)
    xf[f] = x.ravel()[f]
    got = O.soa6_apply(isl, cw, xf.reshape(isl.vtype.shape)).ravel()[f]
    assert np.max(np.abs(got - A @ xf[f])) <= 1e-12 * np.max(np.abs(got))
    d = O.soa6_diag(isl, cw).ravel()[f]
    assert np.allclose(d, A.diagonal(), rtol=1e-14, atol=0)


def test_uniform_faces_reduce_to_the_kappa_operator():
    isl = W.small_box_union(3, L=8)
    kap = np.float32(2.5)
    isl.kappa = np.where(isl.vtype != AIR, kap, np.float32(0)).astype(np.float32)
    cond = np.where(isl.vtype != AIR, kap, np.float32(0))[None].repeat(6, 0).astype(np.float32)
    A6, f6 = O.assemble_soa6(isl, cond)
    Ak, sk = O.assemble(isl)
    keep = np.isin(sk, f6)
    Ak = Ak[keep][:, keep]
    assert np.array_equal(sk[keep], f6)
    assert abs(A6 - Ak).max() <= 1e-14 * abs(Ak).max()


def _shader_weight(c, t, z, y, x, o):
    """SH/pcg_matvec.comp.glsl:79-123 weight of voxel (z,y,x) toward offset o,
    from its own faces in FP32."""
    dirs = [2 * a + (1 if o[a] > 0 else 0) for a in range(3) if o[a] != 0]
    v = [np.float32(c[d, z, y, x]) for d in dirs]
    if len(v) == 1:
        return v[0]
    if len(v) == 2:
        return np.sqrt(np.maximum(v[0] * v[1], np.float32(0))) * np.float32(0.5)
    p = np.maximum(v[0] * v[1] * v[2], np.float32(0))
    return np.float32(np.power(p, np.float32(1.0 / 3.0))) * np.float32(0.1666667)


@pytest.mark.parametrize("materials", [1, 3])
def test_equal_to_shader_weights_where_the_endpoints_agree(materials):
    rng = np.random.default_rng(materials)
    shape = (5, 6, 7)
    mi = rng.integers(0, materials, shape).astype(np.int32)
    rows = [(2400, 30, 3, 0), (7850, 250, 400, 0), (600, 8, 0.5, 0)][:materials]
    a = O.build_island_data(mi, rows)
    isl = domains.make(a["type"])
    c = a["cond"]
    cw = O.soa6_couplings(isl, c)
    lz, ly, lx = shape
    agree = 0
    for k in range(0, 26, 2):
        o = O.OFFSETS[k]
        for z in range(lz):
            for y in range(ly):
                for x in range(lx):
                    j = (z + o[2], y + o[1], x + o[0])
                    if not (0 <= j[0] < lz and 0 <= j[1] < ly and 0 <= j[2] < lx):
                        continue
                    wi = _shader_weight(c, a["type"], z, y, x, o)
                    wj = _shader_weight(c, a["type"], *j, -o)
                    if wi != wj:
                        continue
                    agree += 1
                    assert cw[k // 2, z, y, x] == pytest.approx(float(wi), rel=3e-7, abs=0)
    assert agree > 0
    if materials == 1:  # one material: every pair agrees
        assert agree == sum(
            max(0, lx - abs(o[0])) * max(0, ly - abs(o[1])) * max(0, lz - abs(o[2]))
            for o in O.OFFSETS[::2])


def test_column_closed_form_with_unequal_faces():
    """1 x 1 x L column, z = 0 anchored, load q on z >= 1: the face between
    m-1 and m carries (L - m) q, so phi_k = sum_{m<=k} q (L - m) / a_m with
    a_m = 0.5 (c_{m-1}[+z] + c_m[-z])."""
    L, q = 12, 0.75
    t = np.full((L, 1, 1), SOLID, np.uint8)
    t[0] = ANCHOR
    rng = np.random.default_rng(5)
    c = rng.uniform(0.5, 4.0, (6, L, 1, 1)).astype(np.float32)
    src = np.where(t == SOLID, np.float32(q), np.float32(0)).astype(np.float32)
    isl = domains.make(t, 1.0, src)
    phi = O.direct_solve_soa6(isl, c).ravel()
    a = 0.5 * (c[5, :-1, 0, 0].astype(np.float64) + c[4, 1:, 0, 0].astype(np.float64))
    want = np.concatenate([[0.0], np.cumsum(q * (L - np.arange(1, L)) / a)])
    assert np.allclose(phi, want, rtol=1e-12)
    cw = O.soa6_couplings(isl, c)
    assert O.soa6_residual(isl, cw, phi.reshape(t.shape)) <= 1e-13


def test_failure_scan_precedence():
    t = np.zeros((1, 1, 5), np.uint8)
    t[0, 0, :] = [ANCHOR, SOLID, SOLID, AIR, SOLID]  # voxel 4: no coupled neighbour
    c = _cond(t.shape, 1.0)
    isl = domains.make(t, 1.0, None, rcomp=np.full(t.shape, 0.5, np.float32),
                       rtens=np.full(t.shape, 0.5, np.float32))
    cw = O.soa6_couplings(isl, c)
    phi = np.array([[[0.0, 1.0, 1.2, 0.0, 5.0]]], np.float32)
    code = O.soa6_failure_scan(isl, cw, phi)
    # voxel 1: flux_in from 2 = 0.2 (<= 0.5), flux_out to 0 = 1.0 > 0.5 -> tension
    # voxel 2: flux_out 0.2 -> 0; voxel 4 isolated -> no support; anchor -> 0
    assert code.ravel().tolist() == [0, 4, 0, 0, 3]
    isl.rcomp[...] = 0.1
    isl.rtens[...] = 0.1
    code = O.soa6_failure_scan(isl, cw, phi)
    # voxel 1: flux_in 0.2 > 0.1 wins over its tension; voxel 2: flux_out 0.2
    assert code.ravel().tolist() == [0, 2, 4, 0, 3]
    phi[0, 0, 2] = 2e6
    assert O.soa6_failure_scan(isl, cw, phi).ravel()[2] == 3  # phi > 1e6


def test_energy_density_exact_on_linear_fields():
    shape = (4, 5, 6)
    t = np.full(shape, SOLID, np.uint8)
    rng = np.random.default_rng(7)
    c = rng.uniform(0.5, 3.0, (6,) + shape).astype(np.float32)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float32) for s in shape], indexing="ij")
    phi = (np.float32(0.5) * x + np.float32(0.25) * y - np.float32(1.0) * z).astype(np.float32)
    psi = O.soa6_psi(domains.make(t), c, phi)
    k = [np.float32(0.5) * (c[2 * a] + c[2 * a + 1]) for a in range(3)]
    want = np.float32(0.5) * (k[0] * np.float32(0.25) + k[1] * np.float32(0.0625) + k[2] * 1)
    assert np.allclose(psi, want, rtol=1e-6)
    H = O.soa6_history_update(domains.make(t), c, phi, np.full(shape, 1e9, np.float32))
    assert np.all(H == 1e9)  # monotone: never below the old history


def test_rbgs_fixed_point_is_the_direct_solve():
    isl, cond = domains.soa6_island("l_bracket_5x5x5")
    cw = O.soa6_couplings(isl, cond)
    ref = O.direct_solve_soa6(isl, cond)
    phi = O.soa6_rbgs(isl, cw, sweeps=3000)
    assert np.linalg.norm(phi - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.parametrize("name", domains.fixture_names())
def test_fixture_direct_solves_have_small_residuals(name):
    isl, cond = domains.soa6_island(name)
    phi = O.direct_solve_soa6(isl, cond)
    cw = O.soa6_couplings(isl, cond)
    assert O.soa6_residual(isl, cw, phi) <= 1e-12
