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
    1: 128^3 ~60 % box union on an anchored slab, sigma 0.5 (PCG to 1e-6).
    3: 256^3 ~70 % box union, sigma 1.5 (mixed precision).
    4: 96^3 box union (config-1 recipe) for the phase-field adjoint.
    """
    if index == 0:
        return dense_block(0, 32, 0.5)
    if index == 1:
        return box_union(1, 128, 0.60, 0.5)
    if index == 3:
        return box_union(3, 256, 0.70, 1.5, n_boxes=100000)
    if index == 4:
        return box_union(4, 96, 0.60, 0.5)
    raise ValueError(f"no single-island workload {index}")


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
