"""The reference's binary output files, read and written byte for byte.

Solaris writes three binary streams during a run (BinaryFileAdapter,
little-endian as on the platforms it runs on):

* ``Phases.dat`` (``Solaris/BinaryFileAdapter.cpp:107-183``): per output
  time, ``f64 t, i32 n``, then n records ``i32 id, 6 x f64 (x y z vx vy vz)``.
* ``Integrals.dat`` (``:186-224``): once per file an ``i32`` length and that
  many bytes of a text header; then per output time ``i32 17, f64 t`` and the
  16 integrals of ``Calculate::Integrals`` (``Solaris/Calculate.cpp:43-63``):
  mass, barycentre position (3) and velocity (3), their lengths, angular
  momentum (3) and its length, kinetic, potential and total energy.
* ``TwoBodyAffair.dat`` (``:226-261``): per event ``i32 id, i32 type,
  i32 body1Id, i32 body2Id``, the two bodies' phases (``6 x f64`` each) and
  ``f64 time``; type is ``Solaris/TwoBodyAffair.h:6-13`` (Ejection 0,
  HitCentrum 1, CloseEncounter 2, Collision 3, WeakCapture 4).

The writers produce files the reference's own readers and post-processing
accept; the readers take the reference's files (tests read the output of
the reference program built from its sources).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List

import numpy as np

EJECTION, HIT_CENTRUM, CLOSE_ENCOUNTER, COLLISION, WEAK_CAPTURE = 0, 1, 2, 3, 4
EVENT_NAMES = {EJECTION: "Ejection", HIT_CENTRUM: "HitCentrum", CLOSE_ENCOUNTER: "CloseEncounter",
               COLLISION: "Collision", WEAK_CAPTURE: "WeakCapture"}

# The text header the reference writes at the start of Integrals.dat
# (Solaris/BinaryFileAdapter.cpp:195); readers skip it by its length.
INTEGRALS_HEADER = (b"time [day], mass [Solar], bary center position r: x [AU], y [AU], z [AU], "
                    b"bary center velocity v: x [AU/day], y [AU/day], z [AU/day], length of r [AU], "
                    b"length of v [AU/day], angular momentum vector c: x, y, z, length of c, "
                    b"kinetic energy, potential energy, total energy")

_PHASE_REC = np.dtype([("id", "<i4"), ("y", "<f8", (6,))])
_AFFAIR_REC = np.dtype([("id", "<i4"), ("type", "<i4"), ("body1", "<i4"), ("body2", "<i4"),
                        ("phase1", "<f8", (6,)), ("phase2", "<f8", (6,)), ("time", "<f8")])


class FormatError(ValueError):
    pass


@dataclass
class PhaseRecord:
    time: float
    ids: np.ndarray          # (n,) int32
    y: np.ndarray            # (n, 6) float64


@dataclass
class Phases:
    records: List[PhaseRecord] = field(default_factory=list)

    @property
    def times(self) -> np.ndarray:
        return np.array([r.time for r in self.records])


def read_phases(path: str) -> Phases:
    """Parses a Phases.dat file; raises FormatError on a truncated record."""
    with open(path, "rb") as f:
        buf = f.read()
    out = Phases()
    off = 0
    while off < len(buf):
        if off + 12 > len(buf):
            raise FormatError(f"{path}: truncated record header at byte {off}")
        t, n = struct.unpack_from("<di", buf, off)
        off += 12
        if n < 0 or off + n * _PHASE_REC.itemsize > len(buf):
            raise FormatError(f"{path}: record at t={t} claims {n} bodies past the end of the file")
        rec = np.frombuffer(buf, dtype=_PHASE_REC, count=n, offset=off)
        off += n * _PHASE_REC.itemsize
        out.records.append(PhaseRecord(t, rec["id"].copy(), rec["y"].copy()))
    return out


def write_phases(path: str, time: float, ids, y, append: bool = True) -> None:
    """Appends one output record, as BinaryFileAdapter::SavePhases does."""
    ids = np.ascontiguousarray(ids, dtype="<i4")
    y = np.ascontiguousarray(y, dtype="<f8").reshape(-1, 6)
    if len(ids) != len(y):
        raise ValueError("ids and y differ in length")
    rec = np.empty(len(ids), dtype=_PHASE_REC)
    rec["id"] = ids
    rec["y"] = y
    with open(path, "ab" if append else "wb") as f:
        f.write(struct.pack("<di", float(time), len(ids)))
        f.write(rec.tobytes())


def read_integrals(path: str):
    """Returns (header bytes, array (records, 17): t and the 16 integrals)."""
    with open(path, "rb") as f:
        buf = f.read()
    if not buf:
        return b"", np.zeros((0, 17))
    (hlen,) = struct.unpack_from("<i", buf, 0)
    if hlen < 0 or 4 + hlen > len(buf):
        raise FormatError(f"{path}: bad header length {hlen}")
    header = buf[4:4 + hlen]
    off = 4 + hlen
    rows = []
    while off < len(buf):
        if off + 4 > len(buf):
            raise FormatError(f"{path}: truncated record at byte {off}")
        (ne,) = struct.unpack_from("<i", buf, off)
        off += 4
        if ne < 1 or off + 8 * ne > len(buf):
            raise FormatError(f"{path}: record with {ne} elements past the end of the file")
        rows.append(np.frombuffer(buf, dtype="<f8", count=ne, offset=off).copy())
        off += 8 * ne
    if rows and len({len(r) for r in rows}) != 1:
        raise FormatError(f"{path}: records of different lengths")
    return header, (np.array(rows) if rows else np.zeros((0, 17)))


def write_integrals(path: str, time: float, integrals, first: bool) -> None:
    """Appends one record (with the header when first), as
    BinaryFileAdapter::SaveIntegrals does."""
    v = np.ascontiguousarray(integrals, dtype="<f8").reshape(-1)
    with open(path, "wb" if first else "ab") as f:
        if first:
            f.write(struct.pack("<i", len(INTEGRALS_HEADER)))
            f.write(INTEGRALS_HEADER)
        f.write(struct.pack("<id", len(v) + 1, float(time)))
        f.write(v.tobytes())


def read_two_body_affairs(path: str) -> np.ndarray:
    """Structured array of the events in a TwoBodyAffair.dat file."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) % _AFFAIR_REC.itemsize:
        raise FormatError(f"{path}: {len(buf)} bytes is not a whole number of {_AFFAIR_REC.itemsize}-byte events")
    return np.frombuffer(buf, dtype=_AFFAIR_REC).copy()


def write_two_body_affair(path: str, event_id: int, event_type: int, body1: int, body2: int,
                          phase1, phase2, time: float) -> None:
    """Appends one event, as BinaryFileAdapter::SaveTwoBodyAffair does."""
    rec = np.zeros(1, dtype=_AFFAIR_REC)
    rec["id"], rec["type"], rec["body1"], rec["body2"] = event_id, event_type, body1, body2
    rec["phase1"] = np.asarray(phase1, dtype=np.float64)
    rec["phase2"] = np.asarray(phase2, dtype=np.float64)
    rec["time"] = time
    with open(path, "ab") as f:
        f.write(rec.tobytes())
