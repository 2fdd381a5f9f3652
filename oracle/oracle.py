"""Python face of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the product.

Two layers:
  * ctypes bindings of oracle/_build/liboracle.so, our C restatement of the
    reference's Java CPU solver (VoxelPhysicsCpuReference.java:66-166) and of
    the failure-scan / history rules (SH/failure_scan.comp.glsl:36-196);
  * an FP64 scipy formulation of the same operator (sparse assembly + direct
    solve) used to cross-check the C restatement and as the phi reference.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_PATH = HERE / "_ref" / "libpfsf_compute_ref.so"

EDGE_P = np.float32(0.5)
CORNER_P = np.float32(0.1666667)
# PFSFStencil.java:70-84
OFFSETS = np.array([
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0),
    (1, 0, 1), (1, 0, -1), (-1, 0, 1), (-1, 0, -1),
    (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1),
    (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1),
    (-1, 1, 1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1),
], dtype=np.int64)

_lib = None


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        I = C.c_int
        L.or_jacobi_step.argtypes = [I, I, I, P, P, P, P, P, C.c_float]
        L.or_rbgs.argtypes = [I, I, I, P, P, P, P, I]
        L.or_diag.argtypes = [I, I, I, P, P, P]
        L.or_apply.argtypes = [I, I, I, P, P, P, P]
        L.or_residual.argtypes = [I, I, I, P, P, P, P, P]
        L.or_residual.restype = C.c_double
        L.or_pcg.argtypes = [I, I, I, P, P, P, P, C.c_double, I, P]
        L.or_pcg.restype = I
        L.or_failure_scan.argtypes = [I, I, I, P, P, P, P, P, C.c_float, P]
        L.or_psi.argtypes = [I, I, I, P, P, P, P]
        L.or_history_update.argtypes = [I, I, I, P, P, P, P]
        L.or_anchored.argtypes = [I, I, I, P, P]
        L.or_anchored.restype = C.c_int64
        L.or_normalize_soa6.argtypes = [P, P, P, P, C.c_int32]
        L.or_normalize_soa6.restype = C.c_float
        L.or_chebyshev_omega.argtypes = [C.c_int32, C.c_float]
        L.or_chebyshev_omega.restype = C.c_float
        L.or_downsample_2to1.argtypes = [P, P, I, I, I, P, P]
        L.or_energy_grad_kappa.argtypes = [I, I, I, P, P, P, P, P]
        L.or_flux_loss.argtypes = [I, I, I, P, P, P, P, P, P]
        L.or_flux_loss.restype = C.c_double
        L.or_flux_grad_phi.argtypes = [I, I, I, P, P, P, P, P, P]
        L.or_flux_grad_kappa.argtypes = [I, I, I, P, P, P, P, P, P]
        L.or_label_sv.argtypes = [I, I, I, P, I, P]
        L.or_label_sv.restype = I
        L.or_label_bfs.argtypes = [I, I, I, P, I, P]
        L.or_label_bfs.restype = C.c_int64
        L.or_label_summarise.argtypes = [I, I, I, P, P, I, P, C.POINTER(I)]
        L.or_label_summarise.restype = I
        L.or_build_island_data.argtypes = [I, I, I, P, P, I, P, P, P, P, P, P, P, P, P]
        L.or_cbrt.argtypes = [C.c_double]
        L.or_cbrt.restype = C.c_double
        L.or_soa6_couplings.argtypes = [I, I, I, P, P, P]
        L.or_soa6_apply.argtypes = [I, I, I, P, P, P, P]
        L.or_soa6_diag.argtypes = [I, I, I, P, P, P]
        L.or_soa6_failure_scan.argtypes = [I, I, I, P, P, P, P, P, C.c_float, P]
        L.or_soa6_psi.argtypes = [I, I, I, P, P, P, P]
        L.or_soa6_history_update.argtypes = [I, I, I, P, P, P, P]
        L.or_soa6_rbgs.argtypes = [I, I, I, P, P, P, P, I]
        L.or_phase_field.argtypes = [I, I, I, P, P, P, P, P, P, P, P, P, C.c_float, C.c_float,
                                     C.c_float, I]
        L.or_macro_residual.argtypes = [I, I, I, P, P, P, P, P]
        L.or_set_threads.argtypes = [I]
        L.or_set_threads.restype = I
        L.or_set_threads(1)  # single-threaded unless a caller asks for more
        _lib = L
    return _lib


def build_island_data(mat_idx, table, anchor=None, curing=None, fill=None):
    """PFSFDataBuilder data prep over a material index grid (lz, ly, lx)
    int32; table rows {density, rcomp, rtens, indestructible}.  Returns
    type, source, cond (6, lz, ly, lx), rcomp, rtens, max_phi."""
    mi = np.ascontiguousarray(mat_idx, np.int32)
    tb = np.ascontiguousarray(table, np.float32).reshape(-1, 4)
    lz, ly, lx = mi.shape
    opt = [None if a is None else np.ascontiguousarray(a, dt)
           for a, dt in ((anchor, np.uint8), (curing, np.float32), (fill, np.float32))]
    out = dict(type=np.empty(mi.shape, np.uint8), source=np.empty(mi.shape, np.float32),
               cond=np.empty((6,) + mi.shape, np.float32), rcomp=np.empty(mi.shape, np.float32),
               rtens=np.empty(mi.shape, np.float32), max_phi=np.empty(mi.shape, np.float32))
    lib().or_build_island_data(lx, ly, lz, _p(mi), _p(tb), tb.shape[0],
                               *[None if a is None else _p(a) for a in opt],
                               *[_p(out[k]) for k in ("type", "source", "cond", "rcomp", "rtens",
                                                      "max_phi")])
    return out


def set_threads(n: int) -> int:
    """Threads of the oracle's per-voxel loops (results do not depend on it)."""
    return int(lib().or_set_threads(int(n)))


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _dims(isl):
    return isl.lx, isl.ly, isl.lz


# ---------------------------------------------------------------- C oracle

def jacobi_step(isl, phi_prev, omega=1.0):
    out = np.empty_like(phi_prev, dtype=np.float32)
    lib().or_jacobi_step(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(isl.source),
                         _p(np.ascontiguousarray(phi_prev, np.float32)), _p(out), omega)
    return out


def jacobi_solve(isl, steps):
    """VoxelPhysicsCpuReference.solve (java:121-127): steps Jacobi sweeps from 0."""
    phi = np.zeros(isl.vtype.shape, np.float32)
    for _ in range(steps):
        phi = jacobi_step(isl, phi)
    return phi


def rbgs(isl, phi=None, sweeps=1):
    phi = np.zeros(isl.vtype.shape, np.float32) if phi is None else np.array(phi, np.float32)
    lib().or_rbgs(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(isl.source), _p(phi), sweeps)
    return phi


def diag(isl):
    d = np.empty(isl.vtype.shape, np.float64)
    lib().or_diag(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(d))
    return d


def apply(isl, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(isl.vtype.shape, np.float64)
    lib().or_apply(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(x), _p(y))
    return y


def residual(isl, x, b=None):
    b = isl.source if b is None else np.ascontiguousarray(b, np.float32)
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_residual(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(b), _p(x), None)


def pcg(isl, tol=1e-10, maxit=100000, x0=None, b=None):
    """FP64 Jacobi-PCG; returns (x, iterations, true relative residual)."""
    b = isl.source if b is None else np.ascontiguousarray(b, np.float32)
    x = np.zeros(isl.vtype.shape, np.float64) if x0 is None else np.array(x0, np.float64)
    rel = C.c_double(0)
    it = lib().or_pcg(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(b), _p(x), tol, maxit,
                      C.byref(rel))
    return x, it, rel.value


def failure_scan(isl, phi, phi_orphan=1e6):
    code = np.empty(isl.vtype.shape, np.uint8)
    lib().or_failure_scan(*_dims(isl), _p(isl.vtype), _p(isl.kappa),
                          _p(np.ascontiguousarray(phi, np.float32)), _p(isl.rcomp),
                          _p(isl.rtens), phi_orphan, _p(code))
    return code


def psi(isl, phi):
    out = np.empty(isl.vtype.shape, np.float32)
    lib().or_psi(*_dims(isl), _p(isl.vtype), _p(isl.kappa),
                 _p(np.ascontiguousarray(phi, np.float32)), _p(out))
    return out


def history_update(isl, phi, H):
    H = np.array(H, np.float32)
    lib().or_history_update(*_dims(isl), _p(isl.vtype), _p(isl.kappa),
                            _p(np.ascontiguousarray(phi, np.float32)), _p(H))
    return H


def anchored(isl):
    reach = np.empty(isl.vtype.shape, np.uint8)
    orphans = lib().or_anchored(*_dims(isl), _p(isl.vtype), _p(reach))
    return reach, int(orphans)


def normalize_soa6(source, rcomp, rtens, cond):
    s, r, t, c = (np.array(a, np.float32) for a in (source, rcomp, rtens, cond))
    smax = lib().or_normalize_soa6(_p(s), _p(r), _p(t), _p(c), s.size)
    return s, r, t, c, smax


def chebyshev_omega(it, rho):
    return lib().or_chebyshev_omega(it, rho)


def downsample_2to1(fine, ftype, lx, ly, lz):
    lxc, lyc, lzc = (lx + 1) // 2, (ly + 1) // 2, (lz + 1) // 2
    c = np.empty((lzc, lyc, lxc), np.float32)
    ct = np.empty((lzc, lyc, lxc), np.uint8)
    lib().or_downsample_2to1(_p(np.ascontiguousarray(fine, np.float32)),
                             _p(np.ascontiguousarray(ftype, np.uint8)), lx, ly, lz, _p(c), _p(ct))
    return c, ct


# ------------------------------------------------------- SoA-6 operator

def cbrt(v: float) -> float:
    return float(lib().or_cbrt(float(v)))


def _cond(isl, cond):
    c = np.ascontiguousarray(cond, np.float32).reshape((6,) + isl.vtype.shape)
    return c


def soa6_couplings(isl, cond):
    """Coupling table (13, lz, ly, lx) FP64 of the SoA-6 operator."""
    cw = np.empty((13,) + isl.vtype.shape, np.float64)
    lib().or_soa6_couplings(*_dims(isl), _p(isl.vtype), _p(_cond(isl, cond)), _p(cw))
    return cw


def soa6_apply(isl, cw, x):
    y = np.empty(isl.vtype.shape, np.float64)
    lib().or_soa6_apply(*_dims(isl), _p(isl.vtype), _p(np.ascontiguousarray(cw)),
                        _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def soa6_diag(isl, cw):
    d = np.empty(isl.vtype.shape, np.float64)
    lib().or_soa6_diag(*_dims(isl), _p(isl.vtype), _p(np.ascontiguousarray(cw)), _p(d))
    return d


def soa6_residual(isl, cw, x, b=None):
    """||b - A x|| / ||b|| over solid voxels (FP64)."""
    b = (isl.source if b is None else np.asarray(b)).astype(np.float64)
    r = np.where(isl.vtype == 1, b - soa6_apply(isl, cw, x), 0.0)
    bb = float(np.sum(np.where(isl.vtype == 1, b, 0.0) ** 2))
    return float(np.sqrt(np.sum(r * r) / (bb if bb > 0 else 1.0)))


def soa6_failure_scan(isl, cw, phi, phi_orphan=1e6):
    code = np.empty(isl.vtype.shape, np.uint8)
    lib().or_soa6_failure_scan(*_dims(isl), _p(isl.vtype), _p(np.ascontiguousarray(cw)),
                               _p(np.ascontiguousarray(phi, np.float32)), _p(isl.rcomp),
                               _p(isl.rtens), phi_orphan, _p(code))
    return code


def soa6_psi(isl, cond, phi):
    out = np.empty(isl.vtype.shape, np.float32)
    lib().or_soa6_psi(*_dims(isl), _p(isl.vtype), _p(_cond(isl, cond)),
                      _p(np.ascontiguousarray(phi, np.float32)), _p(out))
    return out


def soa6_history_update(isl, cond, phi, H):
    H = np.array(H, np.float32)
    lib().or_soa6_history_update(*_dims(isl), _p(isl.vtype), _p(_cond(isl, cond)),
                                 _p(np.ascontiguousarray(phi, np.float32)), _p(H))
    return H


def soa6_rbgs(isl, cw, phi=None, sweeps=1):
    phi = np.zeros(isl.vtype.shape, np.float32) if phi is None else np.array(phi, np.float32)
    lib().or_soa6_rbgs(*_dims(isl), _p(isl.vtype), _p(np.ascontiguousarray(cw)),
                       _p(isl.source), _p(phi), sweeps)
    return phi


def soa6_pair_weights(isl, cond):
    """Independent numpy formulation of the SoA-6 couplings (np.sqrt,
    np.cbrt): list over the 26 offsets of (mask, i_flat, j_flat, a)."""
    c = _cond(isl, cond).astype(np.float64)
    c = np.where((c > 0) & (c <= np.finfo(np.float32).max), c, 0.0)
    lz, ly, lx = isl.vtype.shape
    t = isl.vtype
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    out = []
    for k, o in enumerate(OFFSETS):
        nx, ny, nz = x + o[0], y + o[1], z + o[2]
        ok = (nx >= 0) & (nx < lx) & (ny >= 0) & (ny < ly) & (nz >= 0) & (nz < lz) & (t != 0)
        jx, jy, jz = np.clip(nx, 0, lx - 1), np.clip(ny, 0, ly - 1), np.clip(nz, 0, lz - 1)
        ok &= t[jz, jy, jx] != 0
        hi = np.ones(t.shape)
        hj = np.ones(t.shape)
        nd = 0
        for a in range(3):
            if o[a] == 0:
                continue
            nd += 1
            di = 2 * a + (1 if o[a] > 0 else 0)
            dj = 2 * a + (0 if o[a] > 0 else 1)
            hi = hi * c[di]
            hj = hj * c[dj][jz, jy, jx]
        if nd == 2:
            hi, hj, w = np.sqrt(hi), np.sqrt(hj), float(EDGE_P)
        elif nd == 3:
            hi, hj, w = np.cbrt(hi), np.cbrt(hj), float(CORNER_P)
        else:
            w = 1.0
        a = w * 0.5 * (hi + hj)
        out.append((ok, (x + lx * (y + ly * z)).ravel(), (jx + lx * (jy + ly * jz)).ravel(),
                    np.where(ok, a, 0.0)))
    return out


def assemble_soa6(isl, cond):
    """Sparse FP64 matrix of the SoA-6 operator over the free set (solid
    voxels 26-connected to an anchor with a non-zero diagonal).  Returns
    (A, free_flat_indices)."""
    import scipy.sparse as sp

    t = isl.vtype.ravel()
    reach, _ = anchored(isl)
    pw = soa6_pair_weights(isl, cond)
    dg = np.zeros(isl.n)
    for ok, ii, jj, a in pw:
        m = ok.ravel() & (t[ii] == 1)
        np.add.at(dg, ii[m], a.ravel()[m])
    free = (t == 1) & (reach.ravel() != 0) & (dg > 0)
    fidx = np.flatnonzero(free)
    ids = -np.ones(isl.n, np.int64)
    ids[fidx] = np.arange(fidx.size)
    rows, cols, vals = [np.arange(fidx.size)], [np.arange(fidx.size)], [dg[fidx]]
    for ok, ii, jj, a in pw:
        m = ok.ravel() & free[ii] & free[jj]
        rows.append(ids[ii[m]])
        cols.append(ids[jj[m]])
        vals.append(-a.ravel()[m])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(fidx.size, fidx.size))
    return A, fidx


def direct_solve_soa6(isl, cond, b=None):
    """FP64 sparse direct solve of the SoA-6 system on the free set."""
    import scipy.sparse.linalg as spla

    A, fidx = assemble_soa6(isl, cond)
    b = isl.source if b is None else b
    phi = np.zeros(isl.n, np.float64)
    if fidx.size:
        phi[fidx] = spla.spsolve(A.tocsc(), np.asarray(b, np.float64).ravel()[fidx])
    return phi.reshape(isl.vtype.shape)


def macro_residual(isl, phi, b=None):
    """8^3 block maxima of |b - A phi| over solid voxels (FP32), shape
    (ceil(lz/8), ceil(ly/8), ceil(lx/8))."""
    b = isl.source if b is None else np.ascontiguousarray(b, np.float32)
    out = np.empty(((isl.lz + 7) // 8, (isl.ly + 7) // 8, (isl.lx + 7) // 8), np.float32)
    lib().or_macro_residual(*_dims(isl), _p(isl.vtype), _p(isl.kappa),
                            _p(np.ascontiguousarray(phi, np.float32)), _p(b), _p(out))
    return out


PF_L0, PF_GC, PF_RELAX = 2.0, 100.0, 0.3  # dispatcher.cpp:240-248 values


def phase_field(isl, cond, phi, H, d, code, hydration=None, l0=PF_L0, gc=PF_GC,
                relax=PF_RELAX, split=True):
    """One phase-field damage step (SH/phase_field_evolve.comp.glsl): returns
    (d_new, fractured u8).  cond (6, lz, ly, lx); code = the previous failure
    codes (voxels with a code are skipped)."""
    sh = isl.vtype.shape
    d_out = np.empty(sh, np.float32)
    frac = np.empty(sh, np.uint8)
    hy = None if hydration is None else np.ascontiguousarray(hydration, np.float32)
    lib().or_phase_field(*_dims(isl), _p(isl.vtype), _p(_cond(isl, cond)),
                         _p(np.ascontiguousarray(phi, np.float32)),
                         _p(np.ascontiguousarray(H, np.float32)), None if hy is None else _p(hy),
                         _p(np.ascontiguousarray(code, np.uint8)),
                         _p(np.ascontiguousarray(d, np.float32)), _p(d_out), _p(frac), l0, gc,
                         relax, 1 if split else 0)
    return d_out, frac


def sparse_scatter(records, arrays):
    """SH/sparse_scatter.comp.glsl:21-67: each 48-byte record {int32 flat,
    f32 source, u32 type, f32 max_phi, f32 rcomp, f32 rtens, f32 cond[6]}
    overwrites voxel `flat` of the island arrays (dict of flat numpy arrays:
    source, type, max_phi, rcomp, rtens, cond (6n) in SoA order d * n + i).
    A record whose index lies outside [0, n) is skipped (the shader does not
    check; this library defines the skip).  Records must name distinct
    voxels (the device applies them concurrently)."""
    rec = np.ascontiguousarray(records).view(np.uint32).reshape(-1, 12)
    n = arrays["source"].size
    out = {k: np.array(v) for k, v in arrays.items()}
    for r in rec:
        i = int(np.int32(r[0]))
        if i < 0 or i >= n:
            continue
        out["source"][i] = r[1:2].view(np.float32)[0]
        out["type"][i] = np.uint8(r[2] & 0xFF)
        out["max_phi"][i] = r[3:4].view(np.float32)[0]
        out["rcomp"][i] = r[4:5].view(np.float32)[0]
        out["rtens"][i] = r[5:6].view(np.float32)[0]
        for d in range(6):
            out["cond"][d * n + i] = r[6 + d:7 + d].view(np.float32)[0]
    return out


# ------------------------------------------------ reference CPU primitives

def ref_lib():
    """oracle/_ref build of the reference's own normalize/chebyshev/downsample."""
    if not REF_PATH.exists():
        return None
    L = C.CDLL(str(REF_PATH))
    P = C.c_void_p
    L.pfsf_normalize_soa6.argtypes = [P, P, P, P, P, C.c_int32, P]
    L.pfsf_chebyshev_omega.argtypes = [C.c_int32, C.c_float]
    L.pfsf_chebyshev_omega.restype = C.c_float
    L.pfsf_downsample_2to1.argtypes = [P, P, C.c_int32, C.c_int32, C.c_int32, P, P]
    L.pfsf_precompute_omega_table.argtypes = [C.c_float, P, C.c_int32]
    L.pfsf_precompute_omega_table.restype = C.c_int32
    L.pfsf_estimate_spectral_radius.argtypes = [C.c_int32, C.c_float]
    L.pfsf_estimate_spectral_radius.restype = C.c_float
    return L


# ------------------------------------------------------- FP64 scipy oracle

def coupling_f64(k: int, ki: np.ndarray, kj: np.ndarray) -> np.ndarray:
    """Coupling weights in the Java expression order (float sum/product, double scale)."""
    ki = np.maximum(ki, np.float32(0)).astype(np.float32)
    kj = np.maximum(kj, np.float32(0)).astype(np.float32)
    if k < 6:
        return 0.5 * (ki + kj).astype(np.float64)
    r = np.sqrt((ki * kj).astype(np.float64))
    return r * (np.float64(EDGE_P) if k < 18 else np.float64(CORNER_P))


def assemble(isl):
    """Sparse FP64 matrix of the operator restricted to solid voxels.

    Returns (A, solid_flat_indices).  Independent of the C restatement: built
    edge-by-edge with numpy from the same coupling rule.
    """
    import scipy.sparse as sp

    lz, ly, lx = isl.vtype.shape
    t = isl.vtype
    solid = (t != 0) & (t != 2)
    nonair = t != 0
    ids = -np.ones(t.shape, np.int64)
    sidx = np.flatnonzero(solid.ravel())
    ids.ravel()[sidx] = np.arange(sidx.size)
    rows, cols, vals = [], [], []
    dg = np.zeros(sidx.size)
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    for k, (dx, dy, dz) in enumerate(OFFSETS):
        nx, ny, nz = x + dx, y + dy, z + dz
        ok = (nx >= 0) & (nx < lx) & (ny >= 0) & (ny < ly) & (nz >= 0) & (nz < lz) & solid
        ii = np.flatnonzero(ok.ravel())
        jx, jy, jz = nx.ravel()[ii], ny.ravel()[ii], nz.ravel()[ii]
        jj = jx + lx * (jy + ly * jz)
        keep = nonair.ravel()[jj]
        ii, jj = ii[keep], jj[keep]
        a = coupling_f64(k, isl.kappa.ravel()[ii], isl.kappa.ravel()[jj])
        ri = ids.ravel()[ii]
        np.add.at(dg, ri, a)
        sj = solid.ravel()[jj]
        rows.append(ri[sj])
        cols.append(ids.ravel()[jj][sj])
        vals.append(-a[sj])
    rows.append(np.arange(sidx.size))
    cols.append(np.arange(sidx.size))
    vals.append(dg)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(sidx.size, sidx.size))
    return A, sidx


def direct_solve(isl, b=None):
    """FP64 sparse direct solve of A phi = b on solid voxels (0 elsewhere)."""
    import scipy.sparse.linalg as spla

    A, sidx = assemble(isl)
    b = isl.source if b is None else b
    rhs = np.asarray(b, np.float64).ravel()[sidx]
    sol = spla.spsolve(A.tocsc(), rhs)
    phi = np.zeros(isl.n, np.float64)
    phi[sidx] = sol
    return phi.reshape(isl.vtype.shape)


def free_mask(isl):
    """Unknowns of the solve: solid voxels 26-connected to an anchor (u8)."""
    reach, _ = anchored(isl)
    return ((isl.vtype == 1) & (reach != 0)).astype(np.uint8)


def energy_grad_kappa(isl, u, v):
    """sum over couplings of d a_ij / d kappa_i (u_i - u_j)(v_i - v_j) per voxel i."""
    g = np.empty(isl.vtype.shape, np.float64)
    lib().or_energy_grad_kappa(*_dims(isl), _p(isl.vtype), _p(isl.kappa),
                               _p(np.ascontiguousarray(u, np.float64)),
                               _p(np.ascontiguousarray(v, np.float64)), _p(g))
    return g


def flux_loss(isl, fr, phi, rcomp=None):
    """(L, w) of the load-path loss for one load state phi (FP64)."""
    rc = isl.rcomp if rcomp is None else rcomp
    w = np.empty(isl.vtype.shape, np.float64)
    L = lib().or_flux_loss(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(fr),
                           _p(np.ascontiguousarray(phi, np.float64)), _p(rc), _p(w))
    return L, w


def flux_grad_phi(isl, fr, phi, w):
    g = np.empty(isl.vtype.shape, np.float64)
    lib().or_flux_grad_phi(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(fr),
                           _p(np.ascontiguousarray(phi, np.float64)), _p(w), _p(g))
    return g


def flux_grad_kappa(isl, fr, phi, w):
    g = np.empty(isl.vtype.shape, np.float64)
    lib().or_flux_grad_kappa(*_dims(isl), _p(isl.vtype), _p(isl.kappa), _p(fr),
                             _p(np.ascontiguousarray(phi, np.float64)), _p(w), _p(g))
    return g


def load_path(isl, steps=50, b=None):
    """Load path b_t = (t/steps) b, t = 1..steps, solved exactly (sparse direct):
    returns (L, dL/dkappa) in FP64.  A phi_t = b_t is linear in the load, so
    phi_t = s_t phi_1 (s_t = t/steps) and the adjoint states combine into one
    solve: lambda = A^-1 sum_t s_t dL_t/dphi_t."""
    fr = free_mask(isl)
    b = isl.source if b is None else b
    phi1 = direct_solve(isl, b)
    L = 0.0
    gexp = np.zeros(isl.vtype.shape, np.float64)
    G = np.zeros(isl.vtype.shape, np.float64)
    for t in range(1, steps + 1):
        s = t / steps
        ph = s * phi1
        Lt, w = flux_loss(isl, fr, ph)
        L += Lt
        G += s * flux_grad_phi(isl, fr, ph, w)
        gexp += flux_grad_kappa(isl, fr, ph, w)
    lam = direct_solve(isl, G)
    return L, gexp - energy_grad_kappa(isl, lam, phi1)


def energy(isl_all_solid_kappa, phi, lx, ly, lz):
    """E = 1/2 phi^T L phi with every voxel solid (EnergyEvaluatorCPU elastic term)."""
    from types import SimpleNamespace

    vt = np.ones((lz, ly, lx), np.uint8)
    k = np.ascontiguousarray(np.broadcast_to(isl_all_solid_kappa, (lz, ly, lx)), np.float32)
    isl = SimpleNamespace(lx=lx, ly=ly, lz=lz, vtype=vt, kappa=k)
    y = apply(isl, np.ascontiguousarray(phi, np.float64).reshape(lz, ly, lx))
    return 0.5 * float((np.asarray(phi, np.float64).ravel() * y.ravel()).sum())


__all__ = [n for n in dir() if not n.startswith("_")]


# ------------------------------------------------------ connected components

LABEL_NONE = np.uint32(0xFFFFFFFF)


def _grid_dims(vtype):
    lz, ly, lx = vtype.shape
    return lx, ly, lz


def label_sv(vtype, conn=26):
    """LabelPropagation.shiloachVishkin to its fixed point (uint32 labels,
    air = 0xFFFFFFFF).  Returns (labels, rounds)."""
    t = np.ascontiguousarray(vtype, np.uint8)
    ids = np.empty(t.shape, np.uint32)
    rounds = lib().or_label_sv(*_grid_dims(t), _p(t), conn, _p(ids))
    return ids, int(rounds)


def label_bfs(vtype, conn=26):
    """Breadth-first partition labelled 1 + smallest flat index per component.
    Returns (labels, number of components)."""
    t = np.ascontiguousarray(vtype, np.uint8)
    ids = np.empty(t.shape, np.uint32)
    comps = lib().or_label_bfs(*_grid_dims(t), _p(t), conn, _p(ids))
    return ids, int(comps)


def label_summarise(vtype, ids, max_components=64):
    """Component records (max_components x 12 uint32) of converged labels,
    the number of allocated slots and the overflow flag."""
    t = np.ascontiguousarray(vtype, np.uint8)
    ids = np.ascontiguousarray(ids, np.uint32)
    rec = np.empty((max_components, 12), np.uint32)
    ov = C.c_int(0)
    num = lib().or_label_summarise(*_grid_dims(t), _p(t), _p(ids), max_components, _p(rec),
                                   C.byref(ov))
    return rec, int(num), bool(ov.value)


# ── Load path with damage: the phase-field state couples the steps ──────
# Restated in FP64 with sparse direct solves, as the checker of the GPU
# damage-coupled load path (pfsf_load_path_*, desc.damage = 1).  Per step t:
#   kappa_t = kappa * g(d_{t-1}),  g(d) = (1 - K_RES) (1 - d)^2 + K_RES
#   A(kappa_t) phi_t = (t / T) b;  L += sum softplus(f_i(phi_t; kappa_t) - rcomp_i)
#   psi_t = 0.5 kappa_t |grad phi_t|^2 (the per-axis differences of or_psi)
#   H_t = max(H_{t-1}, psi_t)
#   d_t = the phase-field step of or_phase_field (hybrid AT2, split on,
#         hydration 1, faces of conductivity kappa_t) on the free voxels
# and its reverse-mode gradient dL/dkappa.
K_RES = 1e-3
# face neighbours in (axis, step) order: -x, +x, -y, +y, -z, +z on (lz, ly, lx)
_FACES = [(2, -1), (2, 1), (1, -1), (1, 1), (0, -1), (0, 1)]


def _shift(a, axis, step, fill=0.0):
    """out[i] = a[i + step e_axis] (fill outside the grid)."""
    out = np.full_like(a, fill)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    n = a.shape[axis]
    if step > 0:
        src[axis], dst[axis] = slice(step, n), slice(0, n - step)
    else:
        src[axis], dst[axis] = slice(0, n + step), slice(-step, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def damage_g(d):
    return (1.0 - K_RES) * (1.0 - d) ** 2 + K_RES


def damage_dg(d):
    return -2.0 * (1.0 - K_RES) * (1.0 - d)


def psi64(occ, fr, phi, kt):
    """psi = 0.5 kt (gx^2 + gy^2 + gz^2) on free voxels; also the per-axis
    differences and their cases for the reverse pass."""
    g2 = np.zeros(phi.shape)
    parts = []
    for axis in (2, 1, 0):
        pm, pp = _shift(phi, axis, -1), _shift(phi, axis, 1)
        om, op = _shift(occ, axis, -1, False), _shift(occ, axis, 1, False)
        g = np.where(om & op, (pp - pm) * 0.5, np.where(op, pp - phi, np.where(om, phi - pm, 0.0)))
        parts.append((axis, g, om, op))
        g2 += g * g
    return np.where(fr, 0.5 * kt * g2, 0.0), parts


def psi64_vjp(occ, fr, phi, kt, psib):
    """(dphi, dkt) of psi64 against the adjoint psib (free voxels only)."""
    _, parts = psi64(occ, fr, phi, kt)
    psib = np.where(fr, psib, 0.0)
    dkt = np.zeros(phi.shape)
    dphi = np.zeros(phi.shape)
    for axis, g, om, op in parts:
        dkt += 0.5 * g * g * psib
        dg = kt * g * psib
        both, onlyp, onlym = om & op, op & ~om, om & ~op
        # d g / d phi_{i+e} and d g / d phi_{i-e}, scattered to those voxels
        cp = np.where(both, 0.5 * dg, np.where(onlyp, dg, 0.0))
        cm = np.where(both, -0.5 * dg, np.where(onlym, -dg, 0.0))
        dphi += np.where(onlyp, -dg, 0.0) + np.where(onlym, dg, 0.0)  # phi_i in one-sided cases
        dphi += _shift(cp, axis, -1) + _shift(cm, axis, 1)
    return np.where(fr, dphi, 0.0), dkt


def pf64(occ, fr, d, H, phi, kt, l0=PF_L0, gc=PF_GC, relax=PF_RELAX):
    """One phase-field step (or_phase_field's formula in FP64, split on,
    hydration 1, faces of conductivity kt) on free voxels with H > 0, and
    the intermediates of its reverse pass."""
    nf = np.zeros(phi.shape)
    lp = np.zeros(phi.shape)
    ld = np.zeros(phi.shape)
    for axis, step in _FACES:
        o = _shift(occ, axis, step, False)
        nf += o
        lp += np.where(o, _shift(phi, axis, step) - phi, 0.0)
        ld += np.where(o, _shift(d, axis, step) - d, 0.0)
    ssum = nf * kt
    ss = np.maximum(ssum, 1e-12)
    pp = np.maximum(lp, 0.0) / ss
    he = np.maximum(pp, H)
    c1, c2 = gc * l0, gc / l0
    num = 2.0 * he + c1 * ld
    den0 = 2.0 * he + c2
    den = np.maximum(den0, 1e-8)
    q0 = num / den
    q1 = np.maximum(q0, d)
    q = np.minimum(q1, 1.0)
    v0 = d * (1.0 - relax) + q * relax
    v1 = np.clip(v0, 0.0, 1.0)
    v2 = np.maximum(v1, d)
    active = fr.astype(bool) & (H > 0)
    out = np.where(active, v2, d)
    st = dict(nf=nf, lp=lp, ssum=ssum, ss=ss, pp=pp, H=H, he=he, num=num, den0=den0, den=den,
              q0=q0, q1=q1, v0=v0, v1=v1, d=d, active=active, c1=c1, relax=relax)
    return out, st


def pf64_vjp(occ, st, Db):
    """(dd_prev, dH, dphi, dkt) of pf64 against the adjoint Db of its output."""
    a = st["active"]
    r = st["relax"]
    m2 = st["v1"] > st["d"]
    dv1 = Db * m2
    dd = Db * ~m2
    dv0 = dv1 * ((st["v0"] > 0) & (st["v0"] < 1))
    dd = dd + dv0 * (1.0 - r)
    dq = dv0 * r
    dq1 = dq * (st["q1"] < 1.0)
    m0 = st["q0"] > st["d"]
    dq0 = dq1 * m0
    dd = dd + dq1 * ~m0
    dnum = dq0 / st["den"]
    dden = -dq0 * st["q0"] / st["den"] * (st["den0"] > 1e-8)
    dhe = 2.0 * dnum + 2.0 * dden
    dld = st["c1"] * dnum
    mp = st["pp"] > st["H"]
    dpp = dhe * mp
    dH = dhe * ~mp
    ml = st["lp"] > 0
    dlp = dpp / st["ss"] * ml
    dssum = -dpp * st["lp"] / st["ss"] ** 2 * ml * (st["ssum"] > 1e-12)
    dkt = st["nf"] * dssum
    # inactive voxels pass d through and depend on nothing else
    dd = np.where(a, dd, Db)
    dH, dld, dlp, dkt = (np.where(a, v, 0.0) for v in (dH, dld, dlp, dkt))
    # ld and lp: +1 for each occupied face neighbour j, -1 per neighbour for i
    dd = dd - st["nf"] * dld
    dphi = -st["nf"] * dlp
    for axis, step in _FACES:
        o = _shift(occ, axis, step, False)
        # voxel i uses neighbour j = i + step e; j receives i's coefficient
        dd += _shift(np.where(o, dld, 0.0), axis, -step)
        dphi += _shift(np.where(o, dlp, 0.0), axis, -step)
    return dd, dH, dphi, dkt


def _with_kappa(isl, kt32):
    import dataclasses
    return dataclasses.replace(isl, kappa=np.ascontiguousarray(kt32, np.float32))


def _branches(fr, ps, H, st):
    """The branch decisions of one step (max / min / clamp choices), for
    finite-difference checks that must not straddle a kink."""
    a = st["active"]
    return np.concatenate([m.ravel() for m in (
        fr & (ps > H), (st["v1"] > st["d"]) & a, ((st["v0"] > 0) & (st["v0"] < 1)) & a,
        (st["q1"] < 1) & a, (st["q0"] > st["d"]) & a, (st["pp"] > st["H"]) & a,
        (st["lp"] > 0) & a, a)])


def load_path_damage(isl, steps=10, b=None, signature=False):
    """(L, dL/dkappa) of the damage-coupled load path, FP64, direct solves;
    with signature, also every step's branch decisions (_branches)."""
    fr = free_mask(isl).astype(bool)
    occ = isl.vtype != 0
    b = isl.source if b is None else b
    k0 = np.maximum(isl.kappa.astype(np.float64), 0.0)
    d = np.zeros(isl.vtype.shape)
    H = np.zeros(isl.vtype.shape)
    L = 0.0
    tape = []
    for t in range(1, steps + 1):
        kt = np.where(occ, k0 * damage_g(d), 0.0).astype(np.float32)
        it = _with_kappa(isl, kt)
        phi = direct_solve(it, (t / steps) * np.asarray(b, np.float64))
        phi = np.where(fr, phi, 0.0)
        Lt, w = flux_loss(it, fr.astype(np.uint8), phi)
        L += Lt
        ps, _ = psi64(occ, fr, phi, kt.astype(np.float64))
        Hn = np.where(fr, np.maximum(H, ps), H)
        dn, st = pf64(occ, fr, d, Hn, phi, kt.astype(np.float64))
        tape.append((it, kt, phi, w, d, H, ps, st))
        d, H = dn, Hn
    sig = np.concatenate([_branches(fr, ps_, H_, st_) for _, _, _, _, _, H_, ps_, st_ in tape])
    grad = np.zeros(isl.vtype.shape)
    Db = np.zeros(isl.vtype.shape)
    Hb = np.zeros(isl.vtype.shape)
    fr8 = fr.astype(np.uint8)
    for it, kt, phi, w, dprev, Hprev, ps, st in reversed(tape):
        k64 = kt.astype(np.float64)
        dd, dH, dphi_pf, dkt_pf = pf64_vjp(occ, st, Db)
        Hbt = Hb + dH
        mpsi = fr & (ps > Hprev)
        psib = np.where(mpsi, Hbt, 0.0)
        Hb = np.where(mpsi, 0.0, Hbt)
        dphi_psi, dkt_psi = psi64_vjp(occ, fr, phi, k64, psib)
        gphi = flux_grad_phi(it, fr8, phi, w) + np.where(fr, dphi_pf + dphi_psi, 0.0)
        lam = direct_solve(it, gphi)
        lam = np.where(fr, lam, 0.0)
        kbar = (flux_grad_kappa(it, fr8, phi, w) - energy_grad_kappa(it, lam, phi)
                + dkt_pf + dkt_psi)
        kbar = np.where(occ, kbar, 0.0)
        grad += kbar * damage_g(dprev)
        Db = dd + kbar * k0 * damage_dg(dprev)
    grad = np.where(isl.kappa > 0, grad, 0.0)
    return (L, grad, sig) if signature else (L, grad)
