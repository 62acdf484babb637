"""The reference's binary output formats (solaris_amd/formats.py).

Pinned against files the reference program wrote (tests/golden/scenario_raw,
from oracle/_ref/solaris_ref via tests/golden/make_scenario_golden.py): the
readers parse them, and the writers, fed what was read, reproduce them byte
for byte.
"""
import os

import numpy as np
import pytest

from solaris_amd import formats as F

RAW = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenario_raw")


@pytest.mark.parametrize("name", ["EjectionTest", "CollisionMerge"])
def test_writers_reproduce_reference_files(name, tmp_path):
    ph = F.read_phases(os.path.join(RAW, f"{name}_Phases.dat"))
    out = tmp_path / "Phases.dat"
    for r in ph.records:
        F.write_phases(str(out), r.time, r.ids, r.y)
    assert out.read_bytes() == open(os.path.join(RAW, f"{name}_Phases.dat"), "rb").read()

    header, rows = F.read_integrals(os.path.join(RAW, f"{name}_Integrals.dat"))
    assert header == F.INTEGRALS_HEADER
    assert rows.shape[1] == 17
    out = tmp_path / "Integrals.dat"
    for k, row in enumerate(rows):
        F.write_integrals(str(out), row[0], row[1:], first=(k == 0))
    assert out.read_bytes() == open(os.path.join(RAW, f"{name}_Integrals.dat"), "rb").read()

    ev = F.read_two_body_affairs(os.path.join(RAW, f"{name}_TwoBodyAffair.dat"))
    out = tmp_path / "TwoBodyAffair.dat"
    for e in ev:
        F.write_two_body_affair(str(out), int(e["id"]), int(e["type"]), int(e["body1"]), int(e["body2"]),
                                e["phase1"], e["phase2"], float(e["time"]))
    assert out.read_bytes() == open(os.path.join(RAW, f"{name}_TwoBodyAffair.dat"), "rb").read()


def test_reference_files_content():
    # EjectionTest: the star and four bodies beyond 3 AU; all four are
    # ejected at the first step, after which the star is alone
    ph = F.read_phases(os.path.join(RAW, "EjectionTest_Phases.dat"))
    assert [len(r.ids) for r in ph.records] == [5, 1]
    assert ph.records[0].time == 0.0
    assert np.all(ph.records[1].y == 0.0)
    ev = F.read_two_body_affairs(os.path.join(RAW, "EjectionTest_TwoBodyAffair.dat"))
    assert list(ev["type"]) == [F.EJECTION] * 4
    assert list(ev["body2"]) == [1, 2, 3, 4]
    assert np.all(ev["time"] == ph.records[1].time)
    # CollisionMerge: the two giants merge into the heavier one at the first step
    ev = F.read_two_body_affairs(os.path.join(RAW, "CollisionMerge_TwoBodyAffair.dat"))
    assert list(ev["type"]) == [F.COLLISION] and (ev["body1"][0], ev["body2"][0]) == (1, 2)
    ph = F.read_phases(os.path.join(RAW, "CollisionMerge_Phases.dat"))
    assert [len(r.ids) for r in ph.records[:2]] == [3, 2]
    assert list(ph.records[-1].ids) == [0, 1]


def test_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    p = str(tmp_path / "p.dat")
    recs = [(0.0, [0, 3, 7], rng.normal(size=(3, 6))), (12.5, [0, 7], rng.normal(size=(2, 6))), (30.0, [], np.zeros((0, 6)))]
    for t, ids, y in recs:
        F.write_phases(p, t, ids, y)
    ph = F.read_phases(p)
    assert len(ph.records) == 3
    for (t, ids, y), r in zip(recs, ph.records):
        assert r.time == t and list(r.ids) == ids and np.array_equal(r.y, y)
    # a truncated file is an error, not a short read
    with open(p, "rb") as f:
        data = f.read()
    with open(p, "wb") as f:
        f.write(data[:-5])
    with pytest.raises(F.FormatError):
        F.read_phases(p)
    q = str(tmp_path / "e.dat")
    with open(q, "wb") as f:
        f.write(b"\0" * 119)
    with pytest.raises(F.FormatError):
        F.read_two_body_affairs(q)
    i = str(tmp_path / "i.dat")
    F.write_integrals(i, 1.0, np.arange(16.0), first=True)
    F.write_integrals(i, 2.0, np.arange(16.0) * 2, first=False)
    h, rows = F.read_integrals(i)
    assert h == F.INTEGRALS_HEADER and rows.shape == (2, 17)
    assert rows[1, 0] == 2.0 and np.array_equal(rows[1, 1:], np.arange(16.0) * 2)
