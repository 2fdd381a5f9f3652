"""The oracle's phase-field damage step (SH/phase_field_evolve.comp.glsl:
50-168), pinned without the GPU: the closed form of one step on a uniform
field, the frozen and inert voxels, monotonicity, the fracture threshold over
repeated steps, the spectral split and the hydration scaling."""
import numpy as np

import domains
from oracle import oracle as O

AIR, SOLID, ANCHOR = 0, 1, 2
f = np.float32


def _uniform(shape=(3, 4, 5), d0=0.0, H=1.0):
    t = np.full(shape, SOLID, np.uint8)
    isl = domains.make(t)
    cond = np.full((6,) + shape, f(1.0), f)
    phi = np.zeros(shape, f)
    return isl, cond, phi, np.full(shape, f(H), f), np.full(shape, f(d0), f)


def _step_closed_form(d, H, gc=f(O.PF_GC), l0=f(O.PF_L0), relax=f(O.PF_RELAX)):
    """One step with lap_d = 0 and H_eff = H, FP32 in the shader's order."""
    num = f(2.0) * H + gc * l0 * f(0.0)
    den = f(2.0) * H + gc / l0
    q = min(max(num / den, d), f(1.0))
    return max(min(max(d * (f(1.0) - relax) + q * relax, f(0.0)), f(1.0)), d)


def test_one_step_closed_form_on_a_uniform_field():
    for H, d0 in ((1.0, 0.0), (40.0, 0.2), (1e4, 0.5), (0.25, 0.9)):
        isl, cond, phi, Hf, d = _uniform(d0=d0, H=H)
        code = np.zeros(isl.vtype.shape, np.uint8)
        dn, frac = O.phase_field(isl, cond, phi, Hf, d, code)
        want = _step_closed_form(f(d0), f(H))
        assert np.all(dn == want), (H, d0, dn.ravel()[:3], want)
        assert np.array_equal(frac, (dn > f(0.95)).astype(np.uint8))


def test_inert_and_frozen_voxels_keep_their_damage():
    isl, cond, phi, H, d = _uniform(d0=0.1)
    t = isl.vtype
    t[0, 0, 0], t[0, 0, 1] = AIR, ANCHOR
    H[0, 1, 0] = 0.0      # no strain energy: d does not evolve
    H[0, 1, 1] = np.nan
    code = np.zeros(t.shape, np.uint8)
    code[1, 1, 1] = 2     # failed last tick: frozen
    dn, frac = O.phase_field(isl, cond, phi, H, d, code)
    for z, y, x in ((0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 1, 1)):
        assert dn[z, y, x] == f(0.1) and frac[z, y, x] == 0
    assert np.all(dn >= d)


def test_damage_is_monotone_and_fractures_past_095():
    rng = np.random.default_rng(0)
    isl = domains.make(np.full((6, 7, 8), SOLID, np.uint8))
    sh = isl.vtype.shape
    cond = rng.uniform(0.1, 2.0, (6,) + sh).astype(f)
    phi = rng.uniform(0, 3, sh).astype(f)
    H = rng.uniform(0, 5000, sh).astype(f)
    d = np.zeros(sh, f)
    code = np.zeros(sh, np.uint8)
    seen = False
    for _ in range(30):
        dn, frac = O.phase_field(isl, cond, phi, H, d, code)
        assert np.all(dn >= d) and np.all(dn <= 1)
        # frozen voxels (a code from an earlier step) report nothing new
        assert np.array_equal(frac, ((dn > f(0.95)) & (code == 0)).astype(np.uint8))
        assert np.all(dn[code != 0] == d[code != 0])
        seen |= bool(frac.any())
        code = np.maximum(code, frac)  # fractured voxels are frozen next step
        d = dn
    assert seen
    # uniform field: the closed form iterated reaches 0.95 exactly when predicted
    isl, cond, phi, Hf, d = _uniform(H=1e6)
    code = np.zeros(isl.vtype.shape, np.uint8)
    dk = f(0.0)
    for k in range(20):
        d, frac = O.phase_field(isl, cond, phi, Hf, d, code)
        dk = _step_closed_form(dk, f(1e6))
        assert np.all(d == dk) and bool(frac.all()) == bool(dk > f(0.95))


def test_spectral_split_takes_the_outward_flux():
    """A voxel below its neighbours (lap_phi > 0) with small H: H_eff =
    lap_phi / sumSigma."""
    isl, cond, phi, H, d = _uniform(shape=(3, 3, 3), H=1e-3)
    phi[...] = 2.0
    phi[1, 1, 1] = 1.0  # six face neighbours 1 above: lap_phi = 6, sumSigma = 6
    code = np.zeros(isl.vtype.shape, np.uint8)
    dn, _ = O.phase_field(isl, cond, phi, H, d, code)
    want = _step_closed_form(f(0.0), f(1.0))  # H_eff = 6 / 6
    assert dn[1, 1, 1] == want
    dn2, _ = O.phase_field(isl, cond, phi, H, d, code, split=False)
    assert dn2[1, 1, 1] == _step_closed_form(f(0.0), f(1e-3))


def test_hydration_scales_the_fracture_energy():
    isl, cond, phi, H, d = _uniform(H=10.0)
    code = np.zeros(isl.vtype.shape, np.uint8)
    hy = np.full(isl.vtype.shape, f(0.25), f)
    dn, _ = O.phase_field(isl, cond, phi, H, d, code, hydration=hy)
    gc = f(O.PF_GC) * (f(0.25) * f(0.5))  # G_c h^1.5
    assert np.all(dn == _step_closed_form(f(0.0), f(10.0), gc=gc))
    hy[...] = 0.0  # clamped to 0.01
    dn, _ = O.phase_field(isl, cond, phi, H, d, code, hydration=hy)
    gc = f(O.PF_GC) * (f(0.01) * f(np.sqrt(np.float64(f(0.01)))))
    assert np.all(dn == _step_closed_form(f(0.0), f(10.0), gc=gc))
