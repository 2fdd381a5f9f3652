"""The reference tick-path API (pfsf.h) driven from Python with host buffers.

Mirrors how NativePFSFRuntime.onServerTick (NativePFSFRuntime.java:267-383)
uses the JNI bridge: add the island, register its direct buffers (phi doubles
as the stress readback, which makes it the next tick's warm start), then
tick_dbb with a failure buffer and decode {count}{x,y,z,type}*n.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


def _addr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


class Engine:
    def __init__(self, max_island_size=50_000, tick_budget_ms=8, vram_budget=512 << 20,
                 phase_field=True, multigrid=True, config=True):
        self.lib = capi.load()
        cfg = capi.Config(max_island_size, tick_budget_ms, vram_budget, phase_field, multigrid)
        self.h = self.lib.pfsf_create(C.byref(cfg) if config else None)
        if not self.h:
            raise MemoryError("pfsf_create returned NULL")
        self._keep = {}

    def init(self) -> int:
        return self.lib.pfsf_init(self.h)

    def close(self):
        if self.h:
            self.lib.pfsf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_island(self, island_id, lx, ly, lz, origin=(0, 0, 0)) -> int:
        d = capi.IslandDesc(island_id, capi.Pos(*origin), lx, ly, lz)
        return self.lib.pfsf_add_island(self.h, C.byref(d))

    def register(self, island_id, isl, phi=None, cond=None, max_phi=None, kappa=None,
                 stress_is_phi=True) -> int:
        """Register host buffers of a workloads.Island.  Without `kappa` the
        island is solved with the SoA-6 operator of `cond` (default: kappa on
        all six faces); with `kappa` the per-voxel kappa operator is used."""
        n = isl.n
        phi = np.zeros(n, np.float32) if phi is None else np.ascontiguousarray(phi, np.float32)
        if cond is None:
            cond = np.ascontiguousarray(np.tile(isl.kappa.ravel(), 6), np.float32)
        max_phi = np.full(n, 1e9, np.float32) if max_phi is None else max_phi
        src = np.ascontiguousarray(isl.source.ravel())
        vt = np.ascontiguousarray(isl.vtype.ravel())
        rc = np.ascontiguousarray(isl.rcomp.ravel())
        rt = np.ascontiguousarray(isl.rtens.ravel())
        keep = dict(phi=phi, cond=cond, max_phi=max_phi, src=src, vt=vt, rc=rc, rt=rt)
        b = capi.IslandBuffers(_addr(phi), phi.nbytes, _addr(src), src.nbytes, _addr(cond),
                               cond.nbytes, _addr(vt), vt.nbytes, _addr(rc), rc.nbytes,
                               _addr(rt), rt.nbytes, _addr(max_phi), max_phi.nbytes)
        r = self.lib.pfsf_register_island_buffers(self.h, island_id, C.byref(b))
        if r != capi.OK:
            return r
        if kappa is not None:
            k = np.ascontiguousarray(kappa, np.float32).ravel()
            keep["kappa"] = k
            r = self.lib.pfsf_register_island_kappa(self.h, island_id, _addr(k), k.nbytes)
            if r != capi.OK:
                return r
        if stress_is_phi:
            r = self.lib.pfsf_register_stress_readback(self.h, island_id, _addr(phi), phi.nbytes)
        self._keep[island_id] = keep
        return r

    def phi(self, island_id) -> np.ndarray:
        return self._keep[island_id]["phi"]

    def tick_dbb(self, ids, capacity=1024, epoch=0):
        ids = np.ascontiguousarray(ids, np.int32)
        buf = np.zeros(1 + 4 * capacity, np.int32)
        r = self.lib.pfsf_tick_dbb(self.h, _addr(ids), ids.size, epoch, _addr(buf), buf.nbytes)
        n = int(buf[0])
        return r, buf[1:1 + 4 * n].reshape(n, 4).copy()

    def tick(self, ids, capacity=1024, epoch=0):
        ids = np.ascontiguousarray(ids, np.int32)
        ev = (capi.FailureEvent * capacity)()
        res = capi.TickResult(ev, capacity, 0)
        r = self.lib.pfsf_tick(self.h, _addr(ids), ids.size, epoch, C.byref(res))
        out = np.array([(e.pos.x, e.pos.y, e.pos.z, e.type) for e in ev[:res.count]],
                       np.int32).reshape(-1, 4)
        return r, out

    def read_stress(self, island_id, n):
        out = np.zeros(n, np.float32)
        cnt = C.c_int32(0)
        r = self.lib.pfsf_read_stress(self.h, island_id, _addr(out), n, C.byref(cnt))
        return r, out[:cnt.value]

    def read_damage(self, island_id, n):
        out = np.zeros(n, np.float32)
        cnt = C.c_int32(0)
        r = self.lib.pfsf_read_damage(self.h, island_id, _addr(out), n, C.byref(cnt))
        return r, out[:cnt.value]

    def read_history(self, island_id, n):
        out = np.zeros(n, np.float32)
        cnt = C.c_int32(0)
        r = self.lib.pfsf_read_history(self.h, island_id, _addr(out), n, C.byref(cnt))
        return r, out[:cnt.value]

    def components(self, island_id, max_components=64, labels=False):
        """pfsf_island_components: (result, records [max_components x 12] uint32 of
        which the first `num` are filled, num, overflow, labels or None)."""
        hdr = np.zeros(2, np.uint32)
        rec = np.zeros((max_components, 12), np.uint32)
        keep = self._keep.get(island_id)
        out = None
        if labels and keep is not None:
            out = np.zeros(keep["vt"].size, np.uint32)
        r = self.lib.pfsf_island_components(self.h, island_id, max_components, _addr(hdr),
                                            _addr(rec), _addr(out) if out is not None else None)
        return r, rec, int(hdr[0]), bool(hdr[1]), out

    def stats(self):
        s = capi.Stats()
        r = self.lib.pfsf_get_stats(self.h, C.byref(s))
        return r, s

    def status(self, island_id):
        """pfsf_get_island_status: (result, IslandStatus)."""
        st = capi.IslandStatus()
        r = self.lib.pfsf_get_island_status(self.h, island_id, C.byref(st))
        return r, st

    def set_tolerance(self, tol, max_iters=0):
        self.lib.pfsf_set_solve_tolerance(self.h, tol, max_iters)
