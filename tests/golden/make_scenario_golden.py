"""Expected outputs of whole Solaris scenarios, from the reference program.

Runs ``oracle/_ref/solaris_ref`` (the reference's own Solaris.cpp and
Solaris/*.cpp, built from its sources by ``make -C oracle programs``) on
every scenario under ``tests/golden/scenarios/`` in a scratch directory,
reads its Phases.dat, Integrals.dat and TwoBodyAffair.dat with
``solaris_amd.formats`` and stores them as ``scenario_<name>.npz``. The
output files of the smallest scenarios are also kept byte for byte under
``scenario_raw/`` to pin the format writers.

The scenarios are the reference's ``TestCases`` inputs, some shortened or
adjusted so that they finish in seconds and exercise one event each:

* ``EjectionTest``, ``HitCentrumTest``: as in TestCases (events at the
  first step);
* ``SunJupiter``: TimeLine shortened to 5,000 yr;
* ``SunJupiterSaturn``: TimeLine length (an end time: the bodies' epoch is
  JD 2451544.5) shortened to 12,000 yr, about 5,300 yr of integration;
* ``CollisionTest``: TimeLine shortened to 200 yr (no collision: the
  reference reads no radius from its ``VolumeDensity`` tags);
* ``CollisionMerge``: CollisionTest with ``Density`` tags (which the reader
  does take, so the radii are set), collision factor 3000 and 200 yr: the
  two giants merge at the first step;
* ``PlanetesimalWithDrag``: TimeLine shortened to 500 yr;
* ``MigratingJupiterHitCentrum``: MigratingJupiter (Type II migration in a
  nebula) with HitCentrum at 4.955 AU and 100 yr, so that Jupiter is
  removed near its first perihelion.

Usage: ``python tests/golden/make_scenario_golden.py`` (needs the reference
tree to build ``oracle/_ref``; the committed .npz files need nothing).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from solaris_amd import formats as F  # noqa: E402

SCENARIOS = os.path.join(HERE, "scenarios")
RAW_KEEP = ("EjectionTest", "CollisionMerge")


def run_program(exe: str, name: str, workdir: str, timeout: float = 600.0) -> str:
    """Runs one scenario with `exe` in a fresh directory under `workdir`;
    returns that directory (the program writes its outputs next to the
    input file)."""
    d = os.path.join(workdir, name)
    os.makedirs(d, exist_ok=True)
    xml = os.path.join(d, name + ".xml")
    shutil.copy(os.path.join(SCENARIOS, name + ".xml"), xml)
    env = dict(os.environ, OSTYPE="linux")
    r = subprocess.run([exe, "-i", xml], cwd=d, env=env, stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=timeout, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} {name}: exit {r.returncode}\n{r.stdout[-2000:]}")
    return d


def load_outputs(d: str) -> dict:
    ph = F.read_phases(os.path.join(d, "Phases.dat"))
    _, integ = F.read_integrals(os.path.join(d, "Integrals.dat"))
    p = os.path.join(d, "TwoBodyAffair.dat")
    aff = F.read_two_body_affairs(p) if os.path.exists(p) else np.zeros(0, dtype=F._AFFAIR_REC)
    return dict(
        phase_times=ph.times,
        phase_n=np.array([len(r.ids) for r in ph.records], dtype=np.int32),
        phase_ids=np.concatenate([r.ids for r in ph.records]).astype(np.int32),
        phase_y=np.concatenate([r.y for r in ph.records]),
        integrals=integ,
        affair_type=aff["type"].astype(np.int32), affair_body1=aff["body1"].astype(np.int32),
        affair_body2=aff["body2"].astype(np.int32), affair_time=aff["time"].copy(),
        affair_phase1=aff["phase1"].copy(), affair_phase2=aff["phase2"].copy(),
    )


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "solaris_ref")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "programs"], check=True)
    raw_dir = os.path.join(HERE, "scenario_raw")
    os.makedirs(raw_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        for f in sorted(os.listdir(SCENARIOS)):
            name = f[:-4]
            d = run_program(exe, name, tmp)
            out = load_outputs(d)
            np.savez_compressed(os.path.join(HERE, f"scenario_{name}.npz"), **out)
            print(f"{name}: {len(out['phase_times'])} phase records, {len(out['affair_type'])} events")
            if name in RAW_KEEP:
                for fn in ("Phases.dat", "Integrals.dat", "TwoBodyAffair.dat"):
                    shutil.copy(os.path.join(d, fn), os.path.join(raw_dir, f"{name}_{fn}"))


if __name__ == "__main__":
    main()
