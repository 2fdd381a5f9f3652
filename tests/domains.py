"""Test geometries restated from the reference's Java test builders."""
import numpy as np

from pfsf_amd.workloads import ANCHOR, SOLID, Island


def make(vtype, kappa=1.0, source=None, rcomp=None, rtens=None, name="t"):
    vtype = np.ascontiguousarray(vtype, np.uint8)
    sh = vtype.shape
    kap = np.ascontiguousarray(np.broadcast_to(np.float32(kappa), sh), np.float32) \
        if np.isscalar(kappa) else np.ascontiguousarray(kappa, np.float32)
    src = np.zeros(sh, np.float32) if source is None else np.ascontiguousarray(source, np.float32)
    rc = np.full(sh, 1e9, np.float32) if rcomp is None else np.ascontiguousarray(rcomp, np.float32)
    rt = np.full(sh, 1e9, np.float32) if rtens is None else np.ascontiguousarray(rtens, np.float32)
    lz, ly, lx = sh
    return Island(lx, ly, lz, vtype, kap, src, rc, rt, name)


def floating_island(W=11, g=5):
    """VoxelPhysicsCpuReference.buildFloatingIsland (java:200-231)."""
    t = np.zeros((W, W, W), np.uint8)
    t[0] = ANCHOR
    o = (W - g) // 2
    t[2:2 + g, o:o + g, o:o + g] = SOLID
    src = np.where(t == SOLID, 1.0, 0.0)
    return make(t, 1.0, src, name="floating")


def pillar_and_cluster():
    """FloatingIslandTest.java:120-147: anchored ground, 2-voxel pillar, 3-voxel detached column."""
    L = 8
    t = np.zeros((L, L, L), np.uint8)
    t[0] = ANCHOR
    t[1:3, 1, 1] = SOLID
    t[3:6, 6, 6] = SOLID
    return make(t, 1.0, np.where(t == SOLID, 1.0, 0.0), name="pillar")


def cantilever(N, pad=0, load=1.0):
    """VoxelPhysicsCpuReference.buildCantilever (java:179-194)."""
    L = 1 + 2 * pad
    t = np.zeros((N, L, L), np.uint8)
    t[:, pad, pad] = SOLID
    t[0, pad, pad] = ANCHOR
    src = np.zeros(t.shape, np.float32)
    src[1:, pad, pad] = load
    return make(t, np.where(t != 0, 1.0, 0.0), src, name=f"cantilever{N}")
