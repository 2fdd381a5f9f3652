"""Test geometries restated from the reference's Java test builders, and the
reference's geometry fixtures (tests/golden/pfsf-fixtures) as islands."""
import base64
import json
import pathlib

import numpy as np

from pfsf_amd.workloads import ANCHOR, GRAVITY, SOLID, Island


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


FIXTURE_DIR = pathlib.Path(__file__).resolve().parent / "golden" / "pfsf-fixtures"


def fixture_names():
    return sorted(p.stem for p in FIXTURE_DIR.glob("*.json"))


def reference_fixture(name):
    """One of the reference's geometry fixtures (schema v1 of
    Block Reality/api/src/test/resources/pfsf-fixtures/README.md): dims, the
    anchor list, per-voxel material ids (base64 little-endian int32, flat
    index (z*ly + y)*lx + x, 0 = air) and the material registry.  Mapped with
    the workloads' units: kappa = Young's modulus in GPa, b = density * g *
    1 m^3 * 1e-6 (MN) on solid voxels, rcomp / rtens in MPa; a voxel is an
    anchor when listed in `anchors` or its material is an anchor material.
    The fixtures hold no expected outputs: they are inputs only."""
    d = json.loads((FIXTURE_DIR / f"{name}.json").read_text())
    lx, ly, lz = d["dims"]["lx"], d["dims"]["ly"], d["dims"]["lz"]
    ids = np.frombuffer(base64.b64decode(d["materials"]["voxels"]), dtype="<i4")
    assert ids.size == lx * ly * lz, name
    ids = ids.reshape(lz, ly, lx)
    reg = {m["id"]: m for m in d["materials"]["registry"]}
    vtype = np.zeros(ids.shape, np.uint8)
    kappa = np.zeros(ids.shape, np.float32)
    rho = np.zeros(ids.shape, np.float64)
    rc = np.zeros(ids.shape, np.float32)
    rt = np.zeros(ids.shape, np.float32)
    for mid, m in reg.items():
        sel = ids == mid
        vtype[sel] = ANCHOR if m.get("is_anchor", False) else SOLID
        kappa[sel] = np.float32(m["youngs_gpa"])
        rho[sel] = m["density"]
        rc[sel] = np.float32(m["rcomp"])
        rt[sel] = np.float32(m["rtens"])
    for x, y, z in d["anchors"]:
        vtype[z, y, x] = ANCHOR
    src = np.where(vtype == SOLID, rho * GRAVITY * 1e-6, 0.0).astype(np.float32)
    isl = make(vtype, kappa, src, rc, rt, name=name)
    return isl, int(d.get("ticks", 0))


def fixture_material_grid(name):
    """A fixture as the Java data builder sees it: the material index grid
    (lz, ly, lx) int32 into the fixture's registry (-1 = air), the registry
    rows (density, rcomp, rtens, indestructible) and the anchor flags (uint8)
    from the fixture's anchor list."""
    d = json.loads((FIXTURE_DIR / f"{name}.json").read_text())
    lx, ly, lz = d["dims"]["lx"], d["dims"]["ly"], d["dims"]["lz"]
    ids = np.frombuffer(base64.b64decode(d["materials"]["voxels"]), "<i4").reshape(lz, ly, lx)
    reg = d["materials"]["registry"]
    slot = {m["id"]: k for k, m in enumerate(reg)}
    mi = np.vectorize(lambda v: slot.get(int(v), -1))(ids).astype(np.int32)
    rows = [(m["density"], m["rcomp"], m["rtens"], 0) for m in reg]
    anchor = np.zeros(mi.shape, np.uint8)
    for x, y, z in d["anchors"]:
        anchor[z, y, x] = 1
    return mi, rows, anchor


def soa6_island(name, normalize=True):
    """A fixture through the Java data-prep rule (the oracle's restatement of
    PFSFDataBuilder / PFSFConductivity, no curing) and, as the Java caller
    does before registering, the sigma-max normalisation: an Island (kappa =
    largest face, for reference only) and its SoA-6 faces (6, lz, ly, lx)."""
    from oracle import oracle as O

    mi, rows, anchor = fixture_material_grid(name)
    a = O.build_island_data(mi, rows, anchor)
    src, rc, rt, cond = a["source"], a["rcomp"], a["rtens"], a["cond"]
    if normalize:
        src, rc, rt, cond, _ = O.normalize_soa6(src, rc, rt, cond)
    src = np.where(a["type"] == SOLID, src, np.float32(0)).astype(np.float32)
    kap = cond.max(axis=0).astype(np.float32)
    isl = make(a["type"], kap, src, rc, rt, name=name)
    return isl, np.ascontiguousarray(cond, np.float32)
