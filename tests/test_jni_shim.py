"""The JNI shim (pfsf-core_amd/jni/pfsf_jni.c, the natives of
NativePFSFBridge.java:186-600) run against a JNIEnv stand-in
(tests/jni/mock_env.c; this image has no JDK): the marshalling of every
tick-path native (direct buffers, int[] failure tuples, long[] stats,
strings, the sparse upload buffer) and the host primitives, checked
against the C ABI it forwards to.  CPU tests cover what needs no device;
the gpu-marked test drives a whole tick through the shim."""
import ctypes as C
import pathlib
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pfsf_amd import capi
from pfsf_amd import workloads as W

ROOT = pathlib.Path(__file__).resolve().parent.parent
P = C.c_void_p
PRE = "Java_com_blockreality_api_physics_pfsf_NativePFSFBridge_"


@pytest.fixture(scope="module")
def jni(tmp_path_factory):
    # torch first: it carries its own HIP runtime under the same soname, and
    # libpfsf must bind to that one copy (as in every other test process)
    # rather than load the system runtime beside it
    import torch  # noqa: F401

    capi.load()  # libpfsf must be built
    out = tmp_path_factory.mktemp("jni") / "libpfsf_jni_test.so"
    libdir = ROOT / "pfsf-core_amd" / "lib"
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", "-Wall", "-Wextra", "-Werror",
                    f"-I{ROOT / 'tests' / 'jni'}", f"-I{ROOT / 'include'}", "-o", str(out),
                    str(ROOT / "pfsf-core_amd" / "jni" / "pfsf_jni.c"),
                    str(ROOT / "tests" / "jni" / "mock_env.c"), f"-L{libdir}", "-lpfsf",
                    f"-Wl,-rpath,{libdir}"], check=True)
    L = C.CDLL(str(out))
    L.mock_env.restype = P
    L.mock_direct.restype = P
    L.mock_direct.argtypes = [P, C.c_int64]
    L.mock_array.restype = P
    L.mock_array.argtypes = [C.c_int, P, C.c_int64]
    L.mock_string.restype = P
    L.mock_string.argtypes = [C.c_char_p]
    L.mock_data.restype = P
    L.mock_data.argtypes = [P]
    L.mock_len.restype = C.c_int64
    L.mock_len.argtypes = [P]
    L.mock_free.argtypes = [P]
    return L


def nat(L, name, res, *args):
    f = getattr(L, PRE + name)
    f.restype = res
    f.argtypes = [P, P] + list(args)
    return lambda *a: f(L.mock_env(), None, *a)


def arr(L, kind, a):
    return L.mock_array(ord(kind), a.ctypes.data_as(P), a.size)


def direct(L, a):
    return L.mock_direct(a.ctypes.data_as(P), a.nbytes)


def test_shim_exports_the_bridge_natives(jni):
    out = subprocess.run(["nm", "-D", "--defined-only", jni._name], check=True,
                         capture_output=True, text=True).stdout
    names = {ln.split()[-1][len(PRE):] for ln in out.splitlines() if PRE in ln}
    for n in ("nativeCreate", "nativeCreateWithVulkan", "nativeInit", "nativeShutdown",
              "nativeDestroy", "nativeIsAvailable", "nativeGetStats", "nativeSetWind",
              "nativeSetPCGEnabled", "nativeAddIsland", "nativeRemoveIsland",
              "nativeMarkFullRebuild", "nativeNotifyBlockChange", "nativeTick", "nativeReadStress",
              "nativeRegisterIslandBuffers", "nativeRegisterIslandLookups",
              "nativeRegisterStressReadback", "nativeTickDbb", "nativeDrainCallbacks",
              "nativeGetSparseUploadBuffer", "nativeNotifySparseUpdates", "nativeRegisterShader",
              "nativeVersion", "nativeAbiVersion", "nativeHasFeature", "nativeBuildInfo",
              "nativeAbiContractVersion", "nativeNormalizeSoA6", "nativeDownsample2to1",
              "nativeChebyshevOmega", "nativePrecomputeOmegaTable",
              "nativeEstimateSpectralRadius"):
        assert n in names, n


def test_version_and_feature_probes(jni):
    L = jni
    capi_l = capi.load()
    for name, fn in (("nativeAbiContractVersion", capi_l.pfsf_abi_contract_version),
                     ("nativeBuildInfo", capi_l.pfsf_build_info)):
        s = nat(L, name, P)()
        assert C.string_at(L.mock_data(s)) == fn()
        L.mock_free(s)
    assert nat(L, "nativeAbiVersion", C.c_int32)() == capi_l.pfsf_abi_version()
    has = nat(L, "nativeHasFeature", C.c_uint8, P)
    for feat, want in ((b"operator.soa6", 1), (b"solve.batched", 1), (b"simd.avx2", 0)):
        s = L.mock_string(feat)
        assert has(s) == want
        L.mock_free(s)
    assert has(None) == 0
    assert L.mock_pins() == 0


def test_host_primitives_through_the_shim(jni):
    L = jni
    rng = np.random.default_rng(1)
    n = 300
    src, rc, rt = (rng.uniform(0, 40, n).astype(np.float32) for _ in range(3))
    cond = rng.uniform(0, 40, 6 * n).astype(np.float32)
    want = O.normalize_soa6(src, rc, rt, cond)
    got = [a.copy() for a in (src, rc, rt, cond)]
    norm = nat(L, "nativeNormalizeSoA6", C.c_float, P, P, P, P, P, C.c_int32)
    objs = [arr(L, "F", a) for a in got]
    smax = norm(*objs, None, n)
    assert np.float32(smax) == np.float32(want[4])
    for a, b in zip(got, want[:4]):
        assert a.tobytes() == b.tobytes()
    assert norm(None, None, None, None, None, n) == 1.0
    fine = rng.standard_normal((5, 6, 7)).astype(np.float32)
    ft = rng.integers(0, 3, fine.shape).astype(np.uint8)
    coarse = np.zeros((3, 3, 4), np.float32)
    ct = np.zeros((3, 3, 4), np.uint8)
    ds = nat(L, "nativeDownsample2to1", C.c_int32, P, P, C.c_int32, C.c_int32, C.c_int32, P, P)
    assert ds(arr(L, "F", fine), arr(L, "B", ft), 7, 6, 5, arr(L, "F", coarse),
              arr(L, "B", ct)) == capi.OK
    wc, wt = O.downsample_2to1(fine, ft, 7, 6, 5)
    assert coarse.tobytes() == wc.tobytes() and np.array_equal(ct, wt)
    assert ds(None, None, 7, 6, 5, None, None) == capi.INVALID_ARG
    om = nat(L, "nativeChebyshevOmega", C.c_float, C.c_int32, C.c_float)
    tab = nat(L, "nativePrecomputeOmegaTable", C.c_int32, C.c_float, P)
    sr = nat(L, "nativeEstimateSpectralRadius", C.c_float, C.c_int32, C.c_float)
    capi_l = capi.load()
    t = np.zeros(64, np.float32)
    assert tab(0.97, arr(L, "F", t)) == 64
    want_t = np.zeros(64, np.float32)
    capi_l.pfsf_precompute_omega_table(0.97, want_t.ctypes.data_as(P), 64)
    assert t.tobytes() == want_t.tobytes()
    for k in range(10):
        assert np.float32(om(k, 0.97)) == np.float32(capi_l.pfsf_chebyshev_omega(k, 0.97))
    assert np.float32(sr(40, 0.95)) == np.float32(capi_l.pfsf_estimate_spectral_radius(40, 0.95))
    assert L.mock_pins() == 0


def test_null_handles_and_missing_buffers(jni):
    L = jni
    reg = nat(L, "nativeRegisterIslandBuffers", C.c_int32, C.c_int64, C.c_int32, P, P, P, P, P,
              P, P)
    assert reg(0, 1, None, None, None, None, None, None, None) == capi.NOT_INIT
    assert nat(L, "nativeInit", C.c_int32, C.c_int64)(0) == capi.INVALID_ARG
    assert nat(L, "nativeTickDbb", C.c_int32, C.c_int64, P, C.c_int64, P)(0, None, 0,
                                                                          None) == capi.NOT_INIT
    assert nat(L, "nativeGetStats", P, C.c_int64)(0) is None
    vk = nat(L, "nativeCreateWithVulkan", C.c_int64, C.c_int32, C.c_int32, C.c_int64, C.c_uint8,
             C.c_uint8, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int64)
    assert vk(50000, 8, 1 << 29, 1, 1, 1, 2, 3, 0, 4) == 0  # no Vulkan: Java falls back
    h = nat(L, "nativeCreate", C.c_int64, C.c_int32, C.c_int32, C.c_int64, C.c_uint8,
            C.c_uint8)(50000, 8, 1 << 29, 1, 1)
    assert h != 0
    # not initialised yet: the engine answers NOT_INIT through the shim
    assert reg(h, 1, None, None, None, None, None, None, None) == capi.NOT_INIT
    nat(L, "nativeRegisterShader", None, P, P)(None, None)  # accepted, ignored
    nat(L, "nativeDestroy", None, C.c_int64)(h)
    assert L.mock_pins() == 0


@pytest.mark.gpu
def test_tick_through_the_shim_matches_the_c_abi(jni, cuda):
    """Create, init, add, register (faces only), stress readback, tick_dbb,
    tick, read_stress, stats, sparse upload, all through the natives; the
    results equal the same calls on the C ABI (pfsf_amd.engine.Engine)."""
    from pfsf_amd.engine import Engine

    L = jni
    isl = W.small_box_union(6, L=12)
    isl.rcomp[...] = np.where(isl.vtype != 0, np.float32(0.3), np.float32(0))
    isl.rtens[...] = np.where(isl.vtype != 0, np.float32(0.1), np.float32(0))
    rng = np.random.default_rng(2)
    cond = (np.tile(isl.kappa.ravel(), 6) * rng.uniform(0.5, 2, 6 * isl.n)).astype(np.float32)
    n = isl.n
    bufs = dict(phi=np.zeros(n, np.float32), src=isl.source.ravel().copy(), cond=cond.copy(),
                vt=isl.vtype.ravel().copy(), rc=isl.rcomp.ravel().copy(),
                rt=isl.rtens.ravel().copy(), mp=np.full(n, 1e9, np.float32))
    create = nat(L, "nativeCreate", C.c_int64, C.c_int32, C.c_int32, C.c_int64, C.c_uint8,
                 C.c_uint8)
    h = create(1 << 20, 10_000, 1 << 30, 1, 1)
    assert nat(L, "nativeInit", C.c_int32, C.c_int64)(h) == capi.OK
    assert nat(L, "nativeIsAvailable", C.c_uint8, C.c_int64)(h) == 1
    add = nat(L, "nativeAddIsland", C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
              C.c_int32, C.c_int32, C.c_int32, C.c_int32)
    assert add(h, 4, 10, -3, 7, isl.lx, isl.ly, isl.lz) == capi.OK
    reg = nat(L, "nativeRegisterIslandBuffers", C.c_int32, C.c_int64, C.c_int32, P, P, P, P, P,
              P, P)
    d = {k: direct(L, v) for k, v in bufs.items()}
    assert reg(h, 4, d["phi"], d["src"], d["cond"], d["vt"], d["rc"], d["rt"], d["mp"]) == 0
    assert nat(L, "nativeRegisterStressReadback", C.c_int32, C.c_int64, C.c_int32, P)(
        h, 4, d["phi"]) == capi.OK
    fail = np.zeros(1 + 4 * 4096, np.int32)
    ids = np.array([4], np.int32)
    tick_dbb = nat(L, "nativeTickDbb", C.c_int32, C.c_int64, P, C.c_int64, P)
    assert tick_dbb(h, arr(L, "I", ids), 1, direct(L, fail)) == capi.OK
    got_ev = fail[1:1 + 4 * fail[0]].reshape(-1, 4).copy()
    got_phi = bufs["phi"].copy()
    # the same through the C ABI
    with Engine(max_island_size=1 << 20, tick_budget_ms=10_000, vram_budget=1 << 30) as e:
        assert e.init() == capi.OK
        assert e.add_island(4, isl.lx, isl.ly, isl.lz, origin=(10, -3, 7)) == capi.OK
        assert e.register(4, isl, cond=cond.copy()) == capi.OK
        r, ev = e.tick_dbb([4], capacity=4096)
        assert r == capi.OK
        assert np.array_equal(got_phi, e.phi(4)) and np.array_equal(got_ev, ev)
    assert len(got_ev) > 0
    # legacy tick: int[] {count, tuples}
    nat(L, "nativeMarkFullRebuild", None, C.c_int64, C.c_int32)(h, 4)
    out = np.zeros(1 + 4 * 10, np.int32)
    assert nat(L, "nativeTick", C.c_int32, C.c_int64, P, C.c_int64, P)(
        h, arr(L, "I", ids), 2, arr(L, "I", out)) == capi.OK
    assert out[0] == min(10, len(got_ev)) and np.array_equal(out[1:].reshape(-1, 4),
                                                             got_ev[:10])
    st = np.zeros(n, np.float32)
    assert nat(L, "nativeReadStress", C.c_int32, C.c_int64, C.c_int32, P)(
        h, 4, arr(L, "F", st)) == n
    assert np.array_equal(st, bufs["phi"])  # the second tick's phi, also in the stress sink
    s = nat(L, "nativeGetStats", P, C.c_int64)(h)
    stats = np.ctypeslib.as_array(C.cast(L.mock_data(s), C.POINTER(C.c_int64)), (5,))
    assert stats[0] == 1 and stats[1] == n and stats[3] == 1 << 30
    L.mock_free(s)
    sb = nat(L, "nativeGetSparseUploadBuffer", P, C.c_int64, C.c_int32)(h, 4)
    assert sb and L.mock_len(sb) == 512 * 48
    L.mock_free(sb)
    assert nat(L, "nativeNotifySparseUpdates", C.c_int32, C.c_int64, C.c_int32, C.c_int32)(
        h, 4, 0) == capi.OK
    assert nat(L, "nativeDrainCallbacks", C.c_int32, C.c_int64, P)(h, arr(L, "I", out)) == 0
    nat(L, "nativeRemoveIsland", None, C.c_int64, C.c_int32)(h, 4)
    nat(L, "nativeShutdown", None, C.c_int64)(h)
    nat(L, "nativeDestroy", None, C.c_int64)(h)
    assert L.mock_pins() == 0
