"""The reference's geometry fixtures (tests/golden/pfsf-fixtures, schema v1)
through the oracle, on the CPU.

The fixtures carry no expected stress or failures (their README and
GoldenParityTest.java:109-162 check only that they decode), so here they
pin the fixture decoding and serve as inputs: every fixture decodes to
lx*ly*lz voxels whose material ids are in its registry, every listed anchor
lies in the grid, and the oracle's FP64 operator on each is symmetric with a
direct solve whose residual is at rounding level."""
import base64
import json

import numpy as np
import pytest

import domains
from oracle import oracle as O

NAMES = domains.fixture_names()


def test_twenty_fixtures_present():
    assert len(NAMES) == 20


@pytest.mark.parametrize("name", NAMES)
def test_fixture_decodes(name):
    """GoldenParityTest.java:109-162: voxels decode to lx*ly*lz int32."""
    d = json.loads((domains.FIXTURE_DIR / f"{name}.json").read_text())
    assert d["schema_version"] == 1
    lx, ly, lz = d["dims"]["lx"], d["dims"]["ly"], d["dims"]["lz"]
    raw = base64.b64decode(d["materials"]["voxels"])
    assert len(raw) == 4 * lx * ly * lz
    ids = np.frombuffer(raw, dtype="<i4")
    reg = {m["id"] for m in d["materials"]["registry"]}
    assert set(np.unique(ids[ids != 0]).tolist()) <= reg
    for x, y, z in d["anchors"]:
        assert 0 <= x < lx and 0 <= y < ly and 0 <= z < lz
    isl, ticks = domains.reference_fixture(name)
    assert (isl.lx, isl.ly, isl.lz) == (lx, ly, lz) and ticks == d["ticks"]
    # flat index (z*ly + y)*lx + x: the listed anchors land on anchor voxels
    flat = isl.vtype.ravel()
    for x, y, z in d["anchors"]:
        assert flat[(z * ly + y) * lx + x] == 2


@pytest.mark.parametrize("name", NAMES)
def test_oracle_on_fixture(name):
    isl, _ = domains.reference_fixture(name)
    _, orphans = O.anchored(isl)
    assert orphans == 0
    rng = np.random.default_rng(len(name))
    x = rng.random(isl.vtype.shape)
    y = rng.random(isl.vtype.shape)
    fr = isl.vtype == 1
    x, y = np.where(fr, x, 0.0), np.where(fr, y, 0.0)
    xay = float(np.sum(x * O.apply(isl, y)))
    yax = float(np.sum(y * O.apply(isl, x)))
    assert abs(xay - yax) <= 1e-12 * max(1.0, abs(xay))
    phi = O.direct_solve(isl)
    assert O.residual(isl, phi) <= 1e-12
    assert np.all(phi[~fr] == 0)
