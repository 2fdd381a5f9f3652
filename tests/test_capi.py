"""The C-ABI library loads, exports every symbol its headers declare, and
validates arguments before touching a device (the ABI-gate pattern of
scripts/check_abi.py and NP/abi/pfsf_v1.abi.json)."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from pfsf_amd import capi


def _declared():
    names = set()
    for h in sorted(capi.INCLUDE_DIR.glob("*.h")):
        txt = h.read_text()
        for m in re.finditer(r"PFSF_API\s+[\w\s\*]+?\b(pfsf_\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_headers_declare_the_tick_path():
    names = _declared()
    for fn in ("pfsf_create", "pfsf_init", "pfsf_add_island", "pfsf_register_island_buffers",
               "pfsf_register_stress_readback", "pfsf_tick_dbb", "pfsf_tick", "pfsf_read_stress",
               "pfsf_abi_contract_version", "pfsf_solver_solve", "pfsf_hip_failure_scan"):
        assert fn in names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = _declared() - exported
    assert not missing, missing
    assert set(capi.SIGNATURES) <= exported


def test_enum_values_frozen():
    txt = (capi.INCLUDE_DIR / "pfsf_types.h").read_text()
    for name, val in (("PFSF_ERROR_VULKAN", -1), ("PFSF_ERROR_NO_DEVICE", -2),
                      ("PFSF_ERROR_OUT_OF_VRAM", -3), ("PFSF_ERROR_INVALID_ARG", -4),
                      ("PFSF_ERROR_NOT_INIT", -5), ("PFSF_ERROR_ISLAND_FULL", -6),
                      ("PFSF_VOXEL_ANCHOR", 2), ("PFSF_FAIL_CRUSHING", 2),
                      ("PFSF_FAIL_NO_SUPPORT", 3), ("PFSF_FAIL_TENSION", 4), ("PFSF_DIR_POS_Z", 5)):
        assert re.search(rf"\b{name}\s*=\s*{val}\b", txt), name
    assert C.sizeof(capi.FailureEvent) == 16
    assert C.sizeof(capi.IslandBuffers) == 7 * 16


def test_version_probes():
    L = capi.load()
    assert L.pfsf_version() and L.pfsf_build_info()
    assert L.pfsf_abi_contract_version() == b"1.6.0"
    assert L.pfsf_abi_version() >> 16 == 1
    assert L.pfsf_has_feature(b"solve.mg_pcg") and not L.pfsf_has_feature(b"simd.avx2")
    assert not L.pfsf_has_feature(None)


def test_lifecycle_without_init():
    L = capi.load()
    assert L.pfsf_create_with_vulkan(None, None) is None
    h = L.pfsf_create(None)
    assert h
    assert not L.pfsf_is_available(h)
    d = capi.IslandDesc(1, capi.Pos(0, 0, 0), 4, 4, 4)
    assert L.pfsf_add_island(h, C.byref(d)) == capi.NOT_INIT
    assert L.pfsf_tick_dbb(h, None, 0, 0, None, 0) == capi.NOT_INIT
    s = capi.Stats()
    assert L.pfsf_get_stats(h, C.byref(s)) == capi.NOT_INIT
    assert L.pfsf_drain_callbacks(h, None, 0) == 0
    L.pfsf_shutdown(h)
    L.pfsf_destroy(h)
    L.pfsf_destroy(None)  # no-op
    assert L.pfsf_add_island(None, C.byref(d)) == capi.INVALID_ARG


def test_device_entry_points_reject_bad_arguments_on_the_host():
    L = capi.load()
    bad = capi.Dims(0, 4, 4)
    good = capi.Dims(4, 4, 4)
    fake = C.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert L.pfsf_hip_jacobi_step(bad, fake, fake, fake, fake, C.c_void_p(0x2000), 1.0,
                                  None) == capi.INVALID_ARG
    assert L.pfsf_hip_jacobi_step(good, fake, fake, fake, fake, fake, 1.0,
                                  None) == capi.INVALID_ARG  # in-place jacobi
    assert L.pfsf_hip_rbgs(good, None, fake, fake, fake, 1, None) == capi.INVALID_ARG
    assert L.pfsf_hip_rbgs(good, fake, fake, fake, fake, -1, None) == capi.INVALID_ARG
    assert L.pfsf_hip_failure_scan(bad, fake, fake, fake, fake, fake, 1e6, fake,
                                   None) == capi.INVALID_ARG
    assert L.pfsf_hip_history_update(good, fake, fake, fake, None, None, None) == capi.INVALID_ARG
    assert L.pfsf_hip_apply_f64(good, fake, fake, fake, fake, None) == capi.INVALID_ARG
    assert L.pfsf_hip_compact_failures(good, fake, capi.Pos(), fake, -1, None) == capi.INVALID_ARG
    h = C.c_void_p()
    desc = capi.SolverDesc(bad, 0x1000, 0x1000, 0, 0, 0, 1e-6, 10, None)
    assert L.pfsf_solver_create(C.byref(desc), C.byref(h)) == capi.INVALID_ARG
    desc = capi.SolverDesc(good, 0x1000, 0x1000, 0, 7, 0, 1e-6, 10, None)
    assert L.pfsf_solver_create(C.byref(desc), C.byref(h)) == capi.INVALID_ARG
    assert L.pfsf_solver_solve(None, fake, None, None, None) == capi.INVALID_ARG
    sm = C.c_float(0)
    assert L.pfsf_hip_normalize_soa6(None, None, None, None, 0, C.byref(sm), None) == capi.OK
    assert sm.value == 1.0


def test_engine_wrapper_requires_library():
    from pfsf_amd import engine

    e = engine.Engine()
    assert e.add_island(1, 4, 4, 4) == capi.NOT_INIT
    e.close()
