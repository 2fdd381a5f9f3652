"""Device-array front end of libpfsf's pfsf_hip.h surface.

Arguments are torch CUDA tensors (HBM allocations); each call passes their raw
pointers and torch's current HIP stream to the C ABI.  Arrays are contiguous
with shape (lz, ly, lx) or flat N.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


def _lib():
    return capi.load()


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous CUDA tensor"
    return C.c_void_p(t.data_ptr())


def _dims(t):
    lz, ly, lx = t.shape
    return capi.Dims(lx, ly, lz)


def upload(isl, device="cuda"):
    """Island arrays to device tensors (dict)."""
    import torch

    out = {}
    for k in ("vtype", "kappa", "source", "rcomp", "rtens"):
        out[k] = torch.from_numpy(np.ascontiguousarray(getattr(isl, k))).to(device)
    return out


def jacobi_step(vtype, kappa, source, phi_prev, phi_out, omega=1.0):
    capi.check("pfsf_hip_jacobi_step", _lib().pfsf_hip_jacobi_step(
        _dims(vtype), _p(vtype), _p(kappa), _p(source), _p(phi_prev), _p(phi_out), omega,
        _stream()))


def rbgs(vtype, kappa, source, phi, sweeps):
    capi.check("pfsf_hip_rbgs", _lib().pfsf_hip_rbgs(
        _dims(vtype), _p(vtype), _p(kappa), _p(source), _p(phi), sweeps, _stream()))


def failure_scan(vtype, kappa, phi, rcomp, rtens, code, phi_orphan=1e6):
    capi.check("pfsf_hip_failure_scan", _lib().pfsf_hip_failure_scan(
        _dims(vtype), _p(vtype), _p(kappa), _p(phi), _p(rcomp), _p(rtens), phi_orphan,
        _p(code), _stream()))


def history_update(vtype, kappa, phi, H=None, psi=None):
    capi.check("pfsf_hip_history_update", _lib().pfsf_hip_history_update(
        _dims(vtype), _p(vtype), _p(kappa), _p(phi), _p(H) if H is not None else None,
        _p(psi) if psi is not None else None, _stream()))


def apply_f64(vtype, kappa, x, y):
    capi.check("pfsf_hip_apply_f64", _lib().pfsf_hip_apply_f64(
        _dims(vtype), _p(vtype), _p(kappa), _p(x), _p(y), _stream()))


def compact_failures(code, out, capacity, origin=(0, 0, 0)):
    capi.check("pfsf_hip_compact_failures", _lib().pfsf_hip_compact_failures(
        _dims(code), _p(code), capi.Pos(*origin), _p(out), capacity, _stream()))


def downsample_2to1(field, vtype, coarse, ctype):
    capi.check("pfsf_hip_downsample_2to1", _lib().pfsf_hip_downsample_2to1(
        _dims(vtype), _p(field), _p(vtype), _p(coarse), _p(ctype), _stream()))


def normalize_soa6(source, rcomp, rtens, cond):
    smax = C.c_float(0)
    capi.check("pfsf_hip_normalize_soa6", _lib().pfsf_hip_normalize_soa6(
        _p(source), _p(rcomp), _p(rtens), _p(cond), source.numel(), C.byref(smax), _stream()))
    return smax.value


def kappa_from_soa6(cond, kappa):
    capi.check("pfsf_hip_kappa_from_soa6", _lib().pfsf_hip_kappa_from_soa6(
        _p(cond), kappa.numel(), _p(kappa), _stream()))


def build_island_data(material_index, materials, indestructible=None, anchor=None, curing=None,
                      fill=None):
    """Island arrays from a (lz, ly, lx) int32 material index grid on the
    device (pfsf_hip_build_island_data).  materials: a capi.MATERIAL_DTYPE
    array (host); indestructible: per-material uint8 flags (host) or None;
    anchor (uint8), curing and fill (float32): device tensors or None.
    Returns type, source, cond (6, lz, ly, lx), rcomp, rtens, max_phi."""
    import numpy as np
    import torch

    dev = material_index.device
    mats = np.ascontiguousarray(materials, dtype=capi.MATERIAL_DTYPE)
    mt = torch.from_numpy(mats.view(np.uint8).copy()).to(dev)
    ind = None if indestructible is None else torch.from_numpy(
        np.ascontiguousarray(indestructible, np.uint8)).to(dev)
    sh = material_index.shape
    out = dict(type=torch.empty(sh, dtype=torch.uint8, device=dev),
               source=torch.empty(sh, dtype=torch.float32, device=dev),
               cond=torch.empty((6,) + tuple(sh), dtype=torch.float32, device=dev),
               rcomp=torch.empty(sh, dtype=torch.float32, device=dev),
               rtens=torch.empty(sh, dtype=torch.float32, device=dev),
               max_phi=torch.empty(sh, dtype=torch.float32, device=dev))
    opt = [None if a is None else _p(a) for a in (ind, anchor, curing, fill)]
    capi.check("pfsf_hip_build_island_data", _lib().pfsf_hip_build_island_data(
        _dims(material_index), _p(material_index), _p(mt), len(mats), *opt,
        *[_p(out[k]) for k in ("type", "source", "cond", "rcomp", "rtens", "max_phi")],
        _stream()))
    return out


def anchored_mask(vtype, reach):
    capi.check("pfsf_hip_anchored_mask", _lib().pfsf_hip_anchored_mask(
        _dims(vtype), _p(vtype), _p(reach), _stream()))


def label_components(vtype, connectivity=26, max_components=64, island_id=None,
                     components=None):
    """Connected components of the solid/anchor voxels of a (lz, ly, lx) uint8
    grid: island_id (int32 view of the uint32 labels, -1 = air), the component
    records (max_components x 12, int32 view of uint32), the number of records
    and the overflow flag (LabelPropagation.java:199-270 fixed point,
    PFSFLabelPropCpuSimulator.java:164-238 records)."""
    import torch

    if island_id is None:
        island_id = torch.empty(vtype.shape, dtype=torch.int32, device=vtype.device)
    if components is None:
        components = torch.empty((max_components, 12), dtype=torch.int32, device=vtype.device)
    num, ov = C.c_int32(0), C.c_int32(0)
    capi.check("pfsf_hip_label_components", _lib().pfsf_hip_label_components(
        _dims(vtype), _p(vtype), connectivity, _p(island_id), _p(components), max_components,
        C.byref(num), C.byref(ov), _stream()))
    return island_id, components, num.value, bool(ov.value)


class Solver:
    """MG-preconditioned PCG to an FP64 relative-residual tolerance."""

    def __init__(self, vtype, kappa, mg_levels=0, precision=capi.PREC_FP32, tol=1e-6,
                 max_iters=1000, coarse_sweeps=0):
        self._vtype, self._kappa = vtype, kappa  # keep the device arrays alive
        desc = capi.SolverDesc(_dims(vtype), _p(vtype).value, _p(kappa).value, mg_levels,
                               precision, coarse_sweeps, tol, max_iters, _stream().value)
        h = C.c_void_p()
        capi.check("pfsf_solver_create", _lib().pfsf_solver_create(C.byref(desc), C.byref(h)))
        self.h = h

    def solve(self, b, x64=None, phi32=None):
        info = capi.SolveInfo()
        capi.check("pfsf_solver_solve", _lib().pfsf_solver_solve(
            self.h, _p(b), _p(x64) if x64 is not None else None,
            _p(phi32) if phi32 is not None else None, C.byref(info)))
        return info

    def profile(self, which, reps=50):
        """Average device ms of one launch of solver kernel `which` (see pfsf_hip.h)."""
        ms = C.c_double(0)
        capi.check("pfsf_solver_profile",
                   _lib().pfsf_solver_profile(self.h, which, reps, C.byref(ms)))
        return ms.value

    def update(self, vtype, kappa):
        self._vtype, self._kappa = vtype, kappa
        capi.check("pfsf_solver_update", _lib().pfsf_solver_update(self.h, _p(vtype), _p(kappa)))

    def close(self):
        if getattr(self, "h", None):
            _lib().pfsf_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver:
    """Many independent islands solved together (pfsf_batch_*).

    islands: list of (vtype, kappa) device tensors, shapes (lz, ly, lx).
    precision: 0 FP32, 1 BF16 coarse kappa; shape: tile-shape candidate for
    every island (capi.BATCH_SHAPE_CANDIDATES index), or capi.BATCH_AUTOTUNE
    to time the candidates per island-size class (classes() reports them)."""

    def __init__(self, islands, tol=1e-6, max_iters=1000, coarse_sweeps=0, precision=0,
                 shape=0):
        self._keep = list(islands)  # keep the device arrays alive
        n = len(self._keep)
        arr = (capi.BatchIsland * n)()
        for k, (vt, kp) in enumerate(self._keep):
            arr[k] = capi.BatchIsland(_dims(vt), _p(vt).value, _p(kp).value)
        self._arr = arr
        desc = capi.BatchDesc(n, arr, coarse_sweeps, tol, max_iters, _stream().value,
                              precision, shape)
        h = C.c_void_p()
        capi.check("pfsf_batch_create", _lib().pfsf_batch_create(C.byref(desc), C.byref(h)))
        self.h = h
        self.n = n

    def solve(self, bs, x64s=None, phis=None, per_island=False):
        """Solve every island; returns BatchInfo (and per-island SolveInfo list)."""
        n = self.n
        parr = lambda ts: (C.c_void_p * n)(*[_p(t).value if t is not None else None
                                             for t in ts]) if ts is not None else None
        b = parr(bs)
        x = parr(x64s)
        ph = parr(phis)
        info = capi.BatchInfo()
        per = (capi.SolveInfo * n)() if per_island else None
        cast = lambda a: C.cast(a, C.POINTER(C.c_void_p)) if a is not None else None
        capi.check("pfsf_batch_solve", _lib().pfsf_batch_solve(
            self.h, cast(b), cast(x), cast(ph), C.byref(info), per))
        return (info, list(per)) if per_island else info

    def classes(self):
        """Size classes with their tile shapes: list of dicts (cells range,
        islands, sample, chosen depths, per-candidate tuning ms)."""
        out = (capi.BatchClass * capi.BATCH_CLASSES)()
        k = _lib().pfsf_batch_classes(self.h, out, capi.BATCH_CLASSES)
        if k < 0:
            raise RuntimeError("pfsf_batch_classes: no batch")
        return [{"cells": [c.min_cells, c.max_cells], "islands": c.islands, "sample": c.sample,
                 "tz_fine": c.tz_fine, "tz_coarse": c.tz_coarse,
                 "ms": {f"{a}x{b}": round(float(m), 3)
                        for (a, b), m in zip(capi.BATCH_SHAPE_CANDIDATES, c.ms) if m > 0}}
                for c in out[:k]]

    def tuning(self):
        """The tuner's whole-batch decision on its stratified sample."""
        t = capi.BatchTuning()
        capi.check("pfsf_batch_tuning_info", _lib().pfsf_batch_tuning_info(self.h, C.byref(t)))
        return {"sample": t.sample, "mixed": bool(t.mixed),
                "mixed_ms": round(float(t.mixed_ms), 3) if t.mixed_ms > 0 else None,
                "uniform_ms": {f"{a}x{b}": round(float(m), 3)
                               for (a, b), m in zip(capi.BATCH_SHAPE_CANDIDATES, t.uniform_ms)
                               if m > 0}}

    def close(self):
        if getattr(self, "h", None):
            _lib().pfsf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ── Chebyshev-Jacobi and the SoA-6 operator ─────────────────────────────

def chebyshev_jacobi(vtype, kappa, source, phi, tmp, steps, rho):
    capi.check("pfsf_hip_chebyshev_jacobi", _lib().pfsf_hip_chebyshev_jacobi(
        _dims(vtype), _p(vtype), _p(kappa), _p(source), _p(phi), _p(tmp), steps, rho, _stream()))


def soa6_couplings(vtype, cond, cw):
    """cond: (6, lz, ly, lx) float32; cw: (13, lz, ly, lx) float64."""
    capi.check("pfsf_hip_soa6_couplings", _lib().pfsf_hip_soa6_couplings(
        _dims(vtype), _p(vtype), _p(cond), _p(cw), _stream()))


def soa6_apply_f64(vtype, cw, x, y):
    capi.check("pfsf_hip_soa6_apply_f64", _lib().pfsf_hip_soa6_apply_f64(
        _dims(vtype), _p(vtype), _p(cw), _p(x), _p(y), _stream()))


def soa6_failure_scan(vtype, cw, phi, rcomp, rtens, code, phi_orphan=1e6):
    capi.check("pfsf_hip_soa6_failure_scan", _lib().pfsf_hip_soa6_failure_scan(
        _dims(vtype), _p(vtype), _p(cw), _p(phi), _p(rcomp), _p(rtens), phi_orphan, _p(code),
        _stream()))


def soa6_history_update(vtype, cond, phi, H=None, psi=None):
    capi.check("pfsf_hip_soa6_history_update", _lib().pfsf_hip_soa6_history_update(
        _dims(vtype), _p(vtype), _p(cond), _p(phi), _p(H) if H is not None else None,
        _p(psi) if psi is not None else None, _stream()))


def soa6_rbgs(vtype, cw, source, phi, sweeps):
    capi.check("pfsf_hip_soa6_rbgs", _lib().pfsf_hip_soa6_rbgs(
        _dims(vtype), _p(vtype), _p(cw), _p(source), _p(phi), sweeps, _stream()))


def macro_residual(vtype, kappa, phi, b, out):
    """8^3 block maxima of |b - A phi| into out (float32, ceil(l/8) per axis)."""
    capi.check("pfsf_hip_macro_residual", _lib().pfsf_hip_macro_residual(
        _dims(vtype), _p(vtype), _p(kappa), _p(phi), _p(b), _p(out), _stream()))
