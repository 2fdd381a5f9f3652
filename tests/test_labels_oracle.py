"""Pins the connected-component oracle (label propagation) on the CPU.

Sources of the checks:
  LabelPropagationTest.java:33-75          SV partition == BFS partition on
                                           100 random FACE_6 and FULL_26 domains
  LabelPropagationTest.java:104-170        two-tower bridge, floating fragment
  LabelPropagationTest.java:197-236        2x2x2 label semantics (FACE_6)
  GpuLabelPropSimulatorTest.java:89-134    records == independent re-scan
  GpuLabelPropSimulatorTest.java:157-203   overflow past MAX_COMPONENTS = 64
  FloatingIslandTest.java:29-33, 120-147   orphan counts 125 and 3
Mathematics: the Shiloach-Vishkin fixed point labels every voxel with
1 + the smallest flat index of its component, which the BFS produces
directly (its first visit in flat order is the minimum).
"""
import numpy as np
import pytest

import domains
from oracle import oracle as O

NONE = 0xFFFFFFFF


def random_domain(rng, lo=4, hi=12, p_solid=0.55, p_anchor=0.10):
    lx, ly, lz = (int(v) for v in rng.integers(lo, hi, 3))
    r = rng.random((lz, ly, lx))
    t = np.where(r < p_solid, 1, np.where(r < p_solid + p_anchor, 2, 0)).astype(np.uint8)
    return t


def partition(ids, vtype):
    """{frozenset of flat indices: anchored} of a labelling."""
    groups = {}
    flat, typ = ids.ravel(), vtype.ravel()
    for i in np.flatnonzero(flat != NONE):
        groups.setdefault(int(flat[i]), []).append(int(i))
    return {frozenset(m): bool((typ[m] == 2).any()) for m in groups.values()}


@pytest.mark.parametrize("conn,seed", [(6, 20251215), (26, 20251216)])
def test_sv_partition_equals_bfs_on_random_domains(conn, seed):
    rng = np.random.default_rng(seed)
    for trial in range(100):
        t = random_domain(rng)
        sv, rounds = O.label_sv(t, conn)
        bfs, comps = O.label_bfs(t, conn)
        assert partition(sv, t) == partition(bfs, t), (trial, conn)
        # the fixed point is the minimum-index labelling exactly
        assert np.array_equal(sv, bfs), (trial, conn)
        assert len(partition(sv, t)) == comps
        assert rounds >= 1


def test_sv_label_semantics_2x2x2_face6():
    t = np.array([2, 1, 1, 0, 0, 0, 0, 1], np.uint8).reshape(2, 2, 2)
    ids, _ = O.label_sv(t, 6)
    assert ids.ravel().tolist() == [1, 1, 1, NONE, NONE, NONE, NONE, 8]
    part = partition(ids, t)
    assert sorted(part.values()) == [False, True]
    # under 26-connectivity voxel 7 touches 1 through an edge
    ids26, _ = O.label_sv(t, 26)
    assert ids26.ravel().tolist() == [1, 1, 1, NONE, NONE, NONE, NONE, 1]


def _place(shape, solids, anchors, shift):
    t = np.zeros(shape, np.uint8)
    for (x, y, z) in solids:
        t[z + shift[2], y + shift[1], x + shift[0]] = 1
    for (x, y, z) in anchors:
        t[z + shift[2], y + shift[1], x + shift[0]] = 2
    return t


def test_two_tower_bridge_both_stubs_anchored():
    solids = [(0, y, 0) for y in range(4)] + [(6, y, 0) for y in range(4)]
    solids += [(1, 3, 0), (2, 3, 0), (5, 3, 0)]
    t = _place((1, 5, 7), solids, [(0, -1, 0), (6, -1, 0)], (0, 1, 0))
    ids, _ = O.label_sv(t, 6)
    part = partition(ids, t)
    assert len(part) == 2 and all(part.values())


def test_floating_fragment_is_orphan():
    solids = [(0, y, 0) for y in range(5)] + [(4, 8, 4), (5, 8, 4), (5, 8, 5)]
    t = _place((6, 10, 6), solids, [(0, -1, 0)], (0, 1, 0))
    ids, _ = O.label_sv(t, 6)
    part = partition(ids, t)
    assert sorted(part.values()) == [False, True]
    orphan = [m for m, a in part.items() if not a]
    assert len(orphan) == 1 and len(orphan[0]) == 3


def test_records_match_independent_rescan():
    rng = np.random.default_rng(20260102)
    for trial in range(50):
        L = int(rng.integers(6, 12))
        r = rng.random((L, L, L))
        t = np.where(r < 0.5, 1, np.where(r < 0.6, 2, 0)).astype(np.uint8)
        ids, _ = O.label_sv(t, 26)
        rec, num, ov = O.label_summarise(t, ids, 64)
        roots = sorted(set(ids.ravel().tolist()) - {NONE})
        assert num == min(len(roots), 64) and ov == (len(roots) > 64)
        assert np.all(rec[num:] == 0)
        for s in range(num):
            lab = roots[s]  # slots follow ascending root order
            z, y, x = np.nonzero(ids == lab)
            want = [lab, len(x), int((t[ids == lab] == 2).any()), 0, x.min(), y.min(), z.min(), 0,
                    x.max(), y.max(), z.max(), 0]
            assert rec[s].tolist() == want, (trial, s)


def test_overflow_past_64_components():
    L = 10
    t = np.zeros((L, L, L), np.uint8)
    cnt = 0
    for z in range(0, L, 2):
        for y in range(0, L, 2):
            for x in range(0, L, 2):
                if cnt < 72:
                    t[z, y, x] = 1
                    cnt += 1
    ids, _ = O.label_sv(t, 26)
    rec, num, ov = O.label_summarise(t, ids, 64)
    assert ov and num == 64
    rec_all, num_all, ov_all = O.label_summarise(t, ids, 128)
    assert not ov_all and num_all == 72
    assert np.array_equal(rec_all[:64], rec)
    assert np.all(rec[:, 1] == 1)


def test_anchored_components_reproduce_orphan_counts():
    for isl, orphans in ((domains.floating_island(), 125), (domains.pillar_and_cluster(), 3)):
        ids, _ = O.label_bfs(isl.vtype, 26)
        part = partition(ids, isl.vtype)
        n = sum(sum(1 for i in m if isl.vtype.ravel()[i] == 1) for m, a in part.items() if not a)
        assert n == orphans
        reach, _ = O.anchored(isl)
        anchored_solid = np.zeros(isl.vtype.size, bool)
        for m, a in part.items():
            if a:
                anchored_solid[list(m)] = True
        anchored_solid &= isl.vtype.ravel() == 1
        assert np.array_equal(anchored_solid, (reach.ravel() != 0) & (isl.vtype.ravel() == 1))
