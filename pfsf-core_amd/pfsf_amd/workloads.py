"""Seeded synthetic voxel structures for the structural solve.

Arrays use the engine's dense layout: shape (lz, ly, lx), flat index
``x + lx * (y + ly * z)`` (NP/src/core/island_buffer.cpp:24-28,
VoxelPhysicsCpuReference.java:45).  Voxel types: 0 air, 1 solid, 2 anchor
(pfsf_types.h:44-48).

Units (one convention for every workload): kappa = Young's modulus in GPa,
b = rho * g * 1 m^3 * 1e-6 in MN, rcomp / rtens in MPa acting on a 1 m^2 face
(so flux and strength are both MN).  Every generator draws from numpy PCG64
seeded ``0x9F5F0000 + workload_index``.
"""
from __future__ import annotations

import dataclasses

import numpy as np

SEED_BASE = 0x9F5F0000
GRAVITY = 9.81

AIR, SOLID, ANCHOR = 0, 1, 2


@dataclasses.dataclass
class Island:
    """One voxel island on a dense lx*ly*lz grid."""

    lx: int
    ly: int
    lz: int
    vtype: np.ndarray  # uint8 (lz, ly, lx)
    kappa: np.ndarray  # float32
    source: np.ndarray  # float32, 0 on non-solid voxels
    rcomp: np.ndarray  # float32
    rtens: np.ndarray  # float32
    name: str = ""

    @property
    def n(self) -> int:
        return self.lx * self.ly * self.lz

    @property
    def n_solid(self) -> int:
        return int((self.vtype == SOLID).sum())

    @property
    def n_active(self) -> int:
        """Occupied voxels (solid + anchor)."""
        return int((self.vtype != AIR).sum())


def _rng(index: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED_BASE + index))


def _materials(rng, shape, sigma_logn: float):
    """E ~ 30 GPa * exp(sigma * N(0,1)) -> kappa (GPa); rho ~ U[1500,7800]."""
    kappa = (30.0 * np.exp(sigma_logn * rng.standard_normal(shape))).astype(np.float32)
    rho = rng.uniform(1500.0, 7800.0, shape)
    source = (rho * GRAVITY * 1e-6).astype(np.float32)
    rcomp = rng.uniform(5.0, 60.0, shape).astype(np.float32)
    rtens = (np.float32(0.1) * rcomp).astype(np.float32)
    return kappa, source, rcomp, rtens


def _finish(vtype, kappa, source, rcomp, rtens, name) -> Island:
    air = vtype == AIR
    kappa = np.where(air, np.float32(0), kappa).astype(np.float32)
    source = np.where(vtype == SOLID, source, np.float32(0)).astype(np.float32)
    rcomp = np.where(air, np.float32(0), rcomp).astype(np.float32)
    rtens = np.where(air, np.float32(0), rtens).astype(np.float32)
    lz, ly, lx = vtype.shape
    return Island(lx, ly, lz, np.ascontiguousarray(vtype, dtype=np.uint8),
                  np.ascontiguousarray(kappa), np.ascontiguousarray(source),
                  np.ascontiguousarray(rcomp), np.ascontiguousarray(rtens), name)


def anchored_component(occ: np.ndarray, anchor: np.ndarray) -> np.ndarray:
    """Occupied voxels 26-connected to an anchor (VoxelPhysicsCpuReference.findOrphans)."""
    from scipy import ndimage

    lab, _ = ndimage.label(occ, structure=np.ones((3, 3, 3), dtype=bool))
    keep = np.unique(lab[anchor & occ])
    keep = keep[keep > 0]
    return np.isin(lab, keep)


def dense_block(index: int = 0, L: int = 32, sigma_logn: float = 0.5) -> Island:
    """Dense L^3 island, every voxel occupied, plane z=0 anchored."""
    rng = _rng(index)
    shape = (L, L, L)
    kappa, source, rcomp, rtens = _materials(rng, shape, sigma_logn)
    vtype = np.full(shape, SOLID, dtype=np.uint8)
    vtype[0] = ANCHOR
    return _finish(vtype, kappa, source, rcomp, rtens, f"dense{L}")


def box_union(index: int, L: int, occupancy: float, sigma_logn: float, n_boxes: int = 200,
              edge=(4, 40), grounded_first: bool = False) -> Island:
    """Union of seed-placed axis-aligned boxes over an anchored z=0 slab.

    Boxes have edges U{edge} per axis and uniformly random positions inside the
    grid; they are added until the union reaches the target occupancy or
    n_boxes have been placed.  The z=0 plane is a full anchored slab, and the
    result keeps only the occupied voxels 26-connected to it.  With
    grounded_first the first box starts at z=0, so small grids are never empty.
    """
    rng = _rng(index)
    shape = (L, L, L)
    occ = np.zeros(shape, dtype=bool)
    occ[0] = True
    target = occupancy * occ.size
    count = int(occ.sum())
    for b in range(n_boxes):
        ex, ey, ez = (int(min(v, L)) for v in rng.integers(edge[0], edge[1] + 1, size=3))
        x0 = int(rng.integers(0, L - ex + 1))
        y0 = int(rng.integers(0, L - ey + 1))
        z0 = int(rng.integers(0, L - ez + 1))
        if grounded_first and b == 0:
            z0 = 0
        blk = occ[z0:z0 + ez, y0:y0 + ey, x0:x0 + ex]
        count += int((~blk).sum())
        blk[...] = True
        if count >= target:
            break
    anchor = np.zeros(shape, dtype=bool)
    anchor[0] = True
    occ = anchored_component(occ, anchor)
    vtype = np.where(occ, SOLID, AIR).astype(np.uint8)
    vtype[0][occ[0]] = ANCHOR
    kappa, source, rcomp, rtens = _materials(rng, shape, sigma_logn)
    return _finish(vtype, kappa, source, rcomp, rtens, f"boxunion{L}")


def workload(index: int) -> Island:
    """The single-island workloads, by index.

    0: dense 32^3, z=0 anchored, kappa lognormal sigma 0.5 (RBGS 200 sweeps).
    1: 128^3 60 % box union on an anchored slab, sigma 0.5 (PCG to 1e-6);
       boxes are placed until the union holds 60 % of the grid.
    3: 256^3 ~70 % box union, sigma 1.5 (mixed precision).
    4: 96^3 box union (config-1 recipe) for the phase-field adjoint.
    """
    if index == 0:
        return dense_block(0, 32, 0.5)
    if index == 1:
        return box_union(1, 128, 0.60, 0.5, n_boxes=100000)
    if index == 3:
        return box_union(3, 256, 0.70, 1.5, n_boxes=100000)
    if index == 4:
        return box_union(4, 96, 0.60, 0.5, n_boxes=100000)
    raise ValueError(f"no single-island workload {index}")


def grown_island(rng: np.random.Generator, n_target: int, sigma_logn: float = 0.5,
                 max_anchors: int = 8, name: str = "") -> Island:
    """Random box union grown to about n_target voxels with 1..max_anchors
    anchored base voxels at z=0.

    The domain is a cube holding ~2.5 n_target cells.  The first box stands on
    z=0; every later box contains a randomly chosen occupied voxel, so the
    union stays 26-connected and every voxel is kept by the anchored-component
    filter.  Box edges are U{2..max(2, L/2)} per axis.  Anchors are distinct
    occupied voxels of the z=0 plane; the island is cropped to its bounding box.
    """
    L = max(4, int(round((2.5 * n_target) ** (1.0 / 3.0))))
    occ = np.zeros((L, L, L), dtype=bool)
    hi = max(2, L // 2)
    count = 0
    nb = 0
    while count < n_target:
        # random draws in blocks: box edges, placement fractions, candidate voxels
        if nb == 0 or k >= nb:
            nb = 64
            edges = rng.integers(2, hi + 1, size=(nb, 3)).tolist()
            frac = rng.random((nb, 3)).tolist()
            cand = rng.integers(0, L, size=(nb * 16, 3)).tolist()
            k = ci = 0
            if count == 0:  # the first box stands on z=0
                ex, ey, ez = edges[0]
                x0 = int(frac[0][0] * (L - ex + 1))
                y0 = int(frac[0][1] * (L - ey + 1))
                z0 = 0
                k = 1
                occ[z0:z0 + ez, y0:y0 + ey, x0:x0 + ex] = True
                count = ex * ey * ez
                continue
        ex, ey, ez = edges[k]
        fx, fy, fz = frac[k]
        k += 1
        while True:  # a random occupied voxel the box must contain
            if ci >= len(cand):
                cand = rng.integers(0, L, size=(nb * 16, 3)).tolist()
                ci = 0
            cx, cy, cz = cand[ci]
            ci += 1
            if occ[cz, cy, cx]:
                break
        x0 = min(max(cx - int(fx * ex), 0), L - ex)
        y0 = min(max(cy - int(fy * ey), 0), L - ey)
        z0 = min(max(cz - int(fz * ez), 0), L - ez)
        blk = occ[z0:z0 + ez, y0:y0 + ey, x0:x0 + ex]
        count += blk.size - int(np.count_nonzero(blk))
        blk[...] = True
    zs, ys, xs = np.nonzero(occ)
    occ = occ[:zs.max() + 1, ys.min():ys.max() + 1, xs.min():xs.max() + 1]
    base = np.flatnonzero(occ[0])
    k = int(rng.integers(1, max_anchors + 1))
    pick = rng.choice(base, size=min(k, base.size), replace=False)
    vtype = np.where(occ, SOLID, AIR).astype(np.uint8)
    vtype[0].reshape(-1)[pick] = ANCHOR
    kappa, source, rcomp, rtens = _materials32(rng, occ.shape, sigma_logn)
    return _finish(vtype, kappa, source, rcomp, rtens, name)


def _materials32(rng, shape, sigma_logn: float):
    """_materials' distributions drawn in float32 (the batch generator's volume)."""
    f = np.float32
    kappa = f(30.0) * np.exp(f(sigma_logn) * rng.standard_normal(shape, dtype=f))
    rho = f(1500.0) + f(6300.0) * rng.random(shape, dtype=f)
    source = rho * f(GRAVITY * 1e-6)
    rcomp = f(5.0) + f(55.0) * rng.random(shape, dtype=f)
    rtens = f(0.1) * rcomp
    return kappa, source, rcomp, rtens


def island_batch(count: int = 4096, index: int = 2, n_min: int = 64, n_max: int = 262144,
                 sigma_logn: float = 0.5) -> list:
    """The multi-island batch: voxel counts log-uniform on [n_min, n_max]
    (4^3 .. 64^3), each island grown by grown_island, seed 0x9F5F0000 + index."""
    rng = _rng(index)
    sizes = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=count))).astype(int)
    return [grown_island(rng, int(n), sigma_logn, name=f"batch{index}.{i}")
            for i, n in enumerate(sizes)]


def small_box_union(seed: int, L: int = 16, occupancy: float = 0.5, sigma_logn: float = 0.5,
                    n_boxes: int = 40, edge=(2, 8)) -> Island:
    """Small random box union (test-sized), same recipe as the 128^3 workload."""
    return box_union(1000 + seed, L, occupancy, sigma_logn, n_boxes=n_boxes, edge=edge,
                     grounded_first=True)


def column(L: int, kappa: float = 1.0, load: float = 1.0) -> Island:
    """1x1xL column, z=0 anchored (VoxelPhysicsCpuReference.buildCantilever, java:179-194)."""
    vtype = np.full((L, 1, 1), SOLID, dtype=np.uint8)
    vtype[0] = ANCHOR
    kap = np.full((L, 1, 1), kappa, dtype=np.float32)
    src = np.full((L, 1, 1), load, dtype=np.float32)
    src[0] = 0
    big = np.full((L, 1, 1), 1e9, dtype=np.float32)
    return _finish(vtype, kap, src, big, big.copy(), f"column{L}")
