"""Load-path loss and adjoint gradient dL/dkappa on the GPU.

Bars: on small islands the GPU loss and gradient match the oracle's exact
FP64 load path (sparse direct solves, combined adjoint; the oracle's pieces
are pinned by finite differences in test_oracle.py) -- loss within 1e-6
relative (FP64 variant) / 1e-5 (FP32 forward), gradient within 1e-4 relative
L2; on the 96^3 island central differences of the FP64 loss (h = 1e-3 kappa_i)
agree with the gradient on 64 random voxels within 2e-3 of the largest
gradient entry; and the torch.autograd front end returns the same gradient.
The damage-coupled path (damage degrades the next state's kappa) meets the
same loss and gradient bars against the oracle's FP64 damage path."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pfsf_amd import adjoint as A
from pfsf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _loaded(seed, L, load=40.0):
    isl = W.small_box_union(seed, L=L)
    isl.source = np.ascontiguousarray(isl.source * np.float32(load))
    return isl


def _dev(isl, cuda):
    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
                 for a in (isl.vtype, isl.kappa, isl.source, isl.rcomp))


@pytest.mark.parametrize("seed,L,steps,every", [(21, 7, 10, 3), (22, 12, 20, 10)])
def test_gradient_matches_exact_fp64(cuda, seed, L, steps, every):
    isl = _loaded(seed, L)
    ref_L, ref_g = O.load_path(isl, steps=steps)
    vt, kp, b, rc = _dev(isl, cuda)
    i64 = A.load_path_loss(vt, kp, b, rc, steps, every, tol=1e-10, fp64=True)
    assert abs(i64.loss - ref_L) <= 1e-6 * ref_L, (i64.loss, ref_L)
    info, g = A.load_path_grad(vt, kp, b, rc, steps, every, tol=1e-10)
    assert abs(info.loss - ref_L) <= 1e-5 * ref_L, (info.loss, ref_L)
    assert info.forward_solves == steps and info.recompute_solves == steps
    assert info.adjoint_solves == steps and info.checkpoints == (steps + every - 1) // every
    g = g.cpu().numpy()
    err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    assert err <= 1e-4, err


def test_finite_differences_96(cuda):
    isl = W.workload(4)
    vt, kp, b, rc = _dev(isl, cuda)
    info, g = A.load_path_grad(vt, kp, b, rc, 50, 10, tol=1e-10)
    g = g.cpu().numpy().ravel()
    print(f"96^3: loss={info.loss:.6e} solves={info.forward_solves}+{info.recompute_solves}+"
          f"{info.adjoint_solves} pcg={info.pcg_iterations} ms={info.ms:.1f}")
    rng = np.random.Generator(np.random.PCG64(96))
    free = np.flatnonzero(isl.vtype.ravel() == W.SOLID)
    gmax = np.abs(g).max()
    worst = 0.0
    for m in rng.choice(free, size=64, replace=False):
        k0 = kp.view(-1)[m].item()
        vals, ks = [], []
        for sgn in (1, -1):
            kp.view(-1)[m] = k0 * (1 + sgn * 1e-3)
            ks.append(kp.view(-1)[m].item())
            vals.append(A.load_path_loss(vt, kp, b, rc, 50, 10, tol=1e-13, fp64=True).loss)
        kp.view(-1)[m] = k0
        fd = (vals[0] - vals[1]) / (ks[0] - ks[1])
        worst = max(worst, abs(fd - g[m]) / gmax)
    print(f"96^3 FD: worst |fd - grad| / max|grad| = {worst:.2e} over 64 voxels")
    assert worst <= 2e-3, worst


def test_autograd_function(cuda):
    isl = _loaded(23, 10)
    vt, kp, b, rc = _dev(isl, cuda)
    k = kp.clone().requires_grad_(True)
    L = A.loss(k, vt, b, rc, steps=8, checkpoint_every=4, tol=1e-9)
    L.backward()
    _, g = A.load_path_grad(vt, kp, b, rc, 8, 4, tol=1e-9)
    assert torch.allclose(k.grad.double(), g.to(torch.float32).double(), rtol=1e-6, atol=0)
    assert abs(L.item() - A.load_path_loss(vt, kp, b, rc, 8, 4, tol=1e-9).loss) == 0.0


@pytest.mark.parametrize("seed,L,steps,every", [(21, 7, 4, 3), (24, 10, 8, 3)])
def test_damage_gradient_matches_exact_fp64(cuda, seed, L, steps, every):
    isl = _loaded(seed, L, load=400.0)
    ref_L, ref_g = O.load_path_damage(isl, steps=steps)
    und_L, und_g = O.load_path(isl, steps=steps)
    # the damage changes the path: the gradient moves far beyond the bar below
    assert np.linalg.norm(ref_g - und_g) >= 0.1 * np.linalg.norm(und_g)
    vt, kp, b, rc = _dev(isl, cuda)
    i64 = A.load_path_loss(vt, kp, b, rc, steps, every, tol=1e-10, fp64=True, damage=True)
    assert abs(i64.loss - ref_L) <= 1e-6 * ref_L, (i64.loss, ref_L)
    info, g = A.load_path_grad(vt, kp, b, rc, steps, every, tol=1e-10, damage=True)
    assert abs(info.loss - ref_L) <= 1e-5 * ref_L, (info.loss, ref_L)
    assert info.forward_solves == steps and info.recompute_solves == steps
    assert info.adjoint_solves == steps and info.checkpoints == (steps + every - 1) // every
    g = g.cpu().numpy()
    err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    print(f"damage seed {seed}: loss {info.loss:.6e} (undamaged {und_L:.6e}) grad rel-L2 {err:.2e}")
    assert err <= 1e-4, err


def test_damage_autograd_function(cuda):
    isl = _loaded(23, 10, load=400.0)
    vt, kp, b, rc = _dev(isl, cuda)
    k = kp.clone().requires_grad_(True)
    L = A.loss(k, vt, b, rc, steps=6, checkpoint_every=4, tol=1e-9, damage=True)
    L.backward()
    _, g = A.load_path_grad(vt, kp, b, rc, 6, 4, tol=1e-9, damage=True)
    assert torch.allclose(k.grad.double(), g.to(torch.float32).double(), rtol=1e-6, atol=0)
