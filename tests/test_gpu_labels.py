"""Connected-component labels and records on the GPU against the oracle.

Bar: bit-exact.  island_id equals the oracle's Shiloach-Vishkin fixed point
(small grids) or its BFS minimum-index labelling (large grids, where the two
are pinned equal on the CPU); the records, their count and the overflow flag
equal the oracle's summarise pass; the anchored mask built on the same
union-find equals the oracle's BFS reachability."""
import numpy as np
import pytest
import torch

import domains
from oracle import oracle as O
from pfsf_amd import device as D
from pfsf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _grid(rng, shape, p_solid, p_anchor):
    r = rng.random(shape)
    return np.where(r < p_solid, 1, np.where(r < p_solid + p_anchor, 2, 0)).astype(np.uint8)


def _check(cuda, t, conn, maxc, exact_sv):
    tg = torch.from_numpy(np.ascontiguousarray(t)).to(cuda)
    ids, rec, num, ov = D.label_components(tg, conn, maxc)
    got = ids.cpu().numpy().view(np.uint32)
    want = O.label_sv(t, conn)[0] if exact_sv else O.label_bfs(t, conn)[0]
    assert np.array_equal(got, want), (t.shape, conn)
    wrec, wnum, wov = O.label_summarise(t, want, maxc)
    assert (num, ov) == (wnum, wov), (t.shape, conn, num, wnum)
    assert np.array_equal(rec.cpu().numpy().view(np.uint32), wrec), (t.shape, conn)


@pytest.mark.parametrize("conn", [26, 6])
def test_random_small_grids_match_sv(cuda, conn):
    rng = np.random.default_rng(20260101 + conn)
    for trial in range(40):
        shape = tuple(int(v) for v in rng.integers(1, 40, 3))
        t = _grid(rng, shape, rng.uniform(0.2, 0.7), 0.1)
        print(f"labels trial={trial} shape={shape} conn={conn}", flush=True)
        _check(cuda, t, conn, int(rng.choice([1, 64, 1024])), True)


@pytest.mark.parametrize("shape,p", [((1, 1, 1), 1.0), ((4, 9, 33), 0.0), ((8, 8, 32), 1.0),
                                     ((37, 11, 9), 0.45), ((64, 64, 64), 0.3),
                                     ((33, 65, 129), 0.55)])
def test_edge_shapes(cuda, shape, p):
    rng = np.random.default_rng(7)
    t = _grid(rng, shape, p * 0.9, p * 0.1)
    print(f"labels shape={shape} p={p}", flush=True)
    _check(cuda, t, 26, 64, shape[0] * shape[1] * shape[2] <= 40000)


def test_overflow_and_isolated_voxels(cuda):
    L = 10
    t = np.zeros((L, L, L), np.uint8)
    t[::2, ::2, ::2] = 1  # 125 isolated voxels
    _check(cuda, t, 26, 64, True)
    _check(cuda, t, 26, 1024, True)


def test_percolating_grid_many_components(cuda):
    # near the 26-connected percolation threshold: components of every size
    rng = np.random.default_rng(11)
    t = _grid(rng, (64, 256, 256), 0.09, 0.01)
    _, comps = O.label_bfs(t, 26)
    assert comps > 10000
    _check(cuda, t, 26, 1024, False)
    t6 = _grid(rng, (64, 256, 256), 0.28, 0.03)
    _check(cuda, t6, 6, 1024, False)


def test_workload_128_single_component(cuda):
    isl = W.workload(1)
    _check(cuda, isl.vtype, 26, 64, False)
    t = np.ones((128, 128, 128), np.uint8)
    tg = torch.from_numpy(t).to(cuda)
    ids, rec, num, ov = D.label_components(tg)
    assert num == 1 and not ov
    assert bool((ids == 1).all())
    assert rec[0].cpu().tolist() == [1, 128 ** 3, 0, 0, 0, 0, 0, 0, 127, 127, 127, 0]


def test_repeated_calls_are_identical(cuda):
    rng = np.random.default_rng(5)
    tg = torch.from_numpy(_grid(rng, (48, 96, 96), 0.25, 0.02)).to(cuda)
    first = D.label_components(tg, 26, 256)
    for _ in range(5):
        again = D.label_components(tg, 26, 256)
        assert torch.equal(first[0], again[0]) and torch.equal(first[1], again[1])
        assert first[2:] == again[2:]


def test_anchored_mask_on_random_and_workload_grids(cuda):
    rng = np.random.default_rng(3)
    grids = [_grid(rng, (30, 41, 53), 0.5, 0.02), _grid(rng, (64, 128, 128), 0.12, 0.005),
             W.workload(1).vtype, domains.floating_island().vtype]
    for t in grids:
        isl = domains.make(t)
        tg = torch.from_numpy(np.ascontiguousarray(t)).to(cuda)
        reach = torch.empty_like(tg)
        D.anchored_mask(tg, reach)
        r, _ = O.anchored(isl)
        want = (r.astype(bool) & (t == 1)).astype(np.uint8)
        assert np.array_equal(reach.cpu().numpy(), want), t.shape
