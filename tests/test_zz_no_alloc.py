"""pfsf_hip_compact_failures_noalloc (the compaction entry point for callers
that cannot allocate device memory) against pfsf_hip_compact_failures, the
entry point it varies: the same bytes on two calls in a row on one GPU, the
second on a smaller grid (the shared static scratch keeps the first call's
contents), and the size limit checked on the host."""
import ctypes as C

import numpy as np
import pytest
import torch

from pfsf_amd import capi

pytestmark = pytest.mark.gpu


def _compact(fn, code, cap, origin):
    out = torch.full((1 + 4 * cap,), -7, dtype=torch.int32, device=code.device)
    lz, ly, lx = code.shape
    r = fn(capi.Dims(lx, ly, lz), C.c_void_p(code.data_ptr()), capi.Pos(*origin),
           C.c_void_p(out.data_ptr()), cap, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert r == capi.OK
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_compaction_matches_the_allocating_entry_point(cuda):
    L = capi.load()
    rng = np.random.default_rng(4)
    for shape, frac, cap, origin in (((96, 80, 130), 0.3, 1 << 20, (5, -9, 2)),
                                     ((17, 9, 5), 0.5, 300, (0, 0, 0)),
                                     ((17, 9, 5), 0.5, 40, (1, 2, 3))):
        code = np.where(rng.random(shape) < frac, rng.integers(1, 5, shape), 0).astype(np.uint8)
        dc = torch.from_numpy(code).to(cuda)
        want = _compact(L.pfsf_hip_compact_failures, dc, cap, origin)
        got = _compact(L.pfsf_hip_compact_failures_noalloc, dc, cap, origin)
        n = min(int(np.count_nonzero(code)), cap)
        assert want[0] == n
        assert np.array_equal(got, want), shape


def test_noalloc_compaction_rejects_grids_beyond_its_scratch(cuda):
    L = capi.load()
    fake = C.c_void_p(0x1000)  # never dereferenced: the size check fails first
    big = capi.Dims(1024, 1024, 257)  # 269,484,032 voxels > 268,433,408
    assert L.pfsf_hip_compact_failures_noalloc(big, fake, capi.Pos(), fake, 16,
                                               None) == capi.INVALID_ARG
    assert L.pfsf_hip_compact_failures_noalloc(capi.Dims(0, 4, 4), fake, capi.Pos(), fake, 16,
                                               None) == capi.INVALID_ARG
