"""Loads the golden fixtures of tests/golden/ (written by make_golden.py
from the compiled reference)."""
import glob
import json
import os

import numpy as np

from solaris_amd import scenarios as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(prefix=""):
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load(name):
    """Returns (Scenario, dict of outputs)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    counts = json.loads(str(z["counts"]))
    gas = json.loads(str(z["gas"]))
    if gas is not None:
        for k in ("eta", "tau", "scale_height", "density", "mean_free_path"):
            gas[k] = tuple(gas[k])
    meta = json.loads(str(z["meta"]))
    scn = S.Scenario(name=meta["name"], counts=counts, mass=z["mass"], radius=z["radius"],
                     density=z["density"], cd=z["cd"], gamma_stokes=z["gamma_stokes"],
                     gamma_epstein=z["gamma_epstein"], mig_stop_at=z["mig_stop_at"], type=z["type"],
                     mig_type=z["mig_type"], y0=z["y0"], gas=gas, h_next=meta["h_next"],
                     epsilon=meta["epsilon"])
    out = {k: z[k] for k in z.files if k not in ("counts", "gas", "meta")}
    return scn, out


COMPUTE = [n for n in names("ref_") if "compute" in n or n in
           ("ref_drag_regimes", "ref_migration_ac", "ref_migration_bc", "ref_tp_only")]
DRIVERS = [n for n in names("ref_") if n not in COMPUTE]
