"""The C++ drop-in classes (solaris_amd/dropin): the reference's
Acceleration / RungeKuttaFehlberg78 / RungeKutta4 / DormandPrince /
BodyData interfaces over the C ABI.

CPU: the drop-in set compiles against the reference's own headers (where
the reference tree is present) and against solaris_amd/dropin/include.
GPU: tests/cpp/dropin_driver, which calls Driver() in a loop as
Simulator::Integrate does, gives bit-identical states and step sizes to the
same run through the C ABI directly, and its final Compute() matches the
oracle within 1e-12.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.cpu_ref import Oracle
from solaris_amd import _abi, build, runtime, scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


@pytest.mark.skipif(not os.path.isdir("/root/reference/Solaris"), reason="reference headers not present")
def test_dropin_compiles_against_reference_headers():
    build.check_dropin_against("/root/reference/Solaris")


def test_dropin_library_exports_reference_classes():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(ROOT, "solaris_amd", "libsolaris_dropin.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    for sym in ("Acceleration::Compute(double, double*, double*)",
                "RungeKuttaFehlberg78::Driver(BodyData*, Acceleration*, TimeLine*)",
                "RungeKuttaFehlberg78::Step(BodyData*, Acceleration*)",
                "RungeKuttaFehlberg78::GetErrorMax(int, double const*, double const*)",
                "RungeKutta4::Driver(BodyData*, Acceleration*, TimeLine*)",
                "DormandPrince::Driver(BodyData*, Acceleration*, TimeLine*)",
                "DormandPrince::Step2(BodyData*, Acceleration*)",
                "DormandPrince::GetErrorMax(int, double const*)",
                "BodyData::Allocate()"):
        assert sym in out, sym


def write_input(path, scn, h0):
    with open(path, "wb") as f:
        c = scn.counts
        f.write(struct.pack("<8i", c["central_body"], c["giant_planet"], c["rocky_planet"], c["proto_planet"],
                            c["super_planetesimal"], c["planetesimal"], c["test_particle"], c["total"]))
        for k in ("mass", "radius", "density", "cd", "gamma_stokes", "gamma_epstein", "mig_stop_at"):
            f.write(np.ascontiguousarray(getattr(scn, k), dtype="<f8").tobytes())
        for k in ("type", "mig_type"):
            f.write(np.ascontiguousarray(getattr(scn, k), dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(scn.y0, dtype="<f8").tobytes())
        g = scn.gas
        gv = np.zeros(21)
        if g is not None:
            gv[:9] = [g["alpha"], g["mean_molecular_weight"], g["particle_diameter"], g["type"],
                      g["time_scale"], g["t0"], g["t1"], g["inner_edge"], g["a"]]
            k = 9
            for name in ("eta", "tau", "scale_height", "density", "mean_free_path"):
                gv[k:k + 2] = g[name]
                k += 2
        f.write(struct.pack("<i", 1 if g is not None else 0))
        f.write(gv.astype("<f8").tobytes())
        f.write(struct.pack("<2d", h0, scn.epsilon))


def read_output(path, n):
    b = open(path, "rb").read()
    status, done = struct.unpack_from("<2i", b, 0)
    o = 8
    hdid = np.frombuffer(b, "<f8", done, o); o += 8 * done
    t = struct.unpack_from("<d", b, o)[0]; o += 8
    y0 = np.frombuffer(b, "<f8", 6 * n, o).reshape(n, 6); o += 48 * n
    acc = np.frombuffer(b, "<f8", 6 * n, o).reshape(n, 6); o += 48 * n
    rm3 = np.frombuffer(b, "<f8", n, o); o += 8 * n
    nn = np.frombuffer(b, "<i4", n, o); o += 4 * n
    mig = np.frombuffer(b, "<i4", n, o); o += 4 * n
    dydt = np.frombuffer(b, "<f8", 6 * n, o).reshape(n, 6)
    return dict(status=status, hdid=hdid, time=t, y0=y0, accel=acc, rm3=rm3, nn=nn, mig=mig, dydt=dydt)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["c1_rkf78", "c2_rkf78", "c5_rk4", "c2_dp"])
def test_dropin_driver_matches_c_abi_run(case):
    scn = {"c1_rkf78": lambda: S.make_c1(n_tp=96), "c2_rkf78": lambda: S.make_c2(n_pl=300),
           "c5_rk4": lambda: S.make_c5(n_tp=100), "c2_dp": lambda: S.make_c2(n_pl=200)}[case]()
    integ = {"c1_rkf78": 3, "c2_rkf78": 3, "c5_rk4": 1, "c2_dp": 0}[case]
    h0 = scn.h_next if integ == 3 else (1.0 if integ == 1 else 5.0)
    steps = 25
    print(f"[dropin {case}]", flush=True)
    with tempfile.TemporaryDirectory() as d:
        inp, outp = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(inp, scn, h0)
        r = subprocess.run([DRIVER, inp, outp, str(steps), str(integ)], timeout=120)
        assert r.returncode == 0
        got = read_output(outp, scn.n)
    ctx = runtime.Context(scn)
    ctx.set_state(scn.y0)
    tl = _abi.timeline(0.0, h0, 1e30)
    hd = []
    for _ in range(steps):
        rr, _ = ctx.driver(integ, scn.epsilon, tl)
        assert rr == 0
        hd.append(tl.hDid)
    assert np.array_equal(got["hdid"], np.array(hd))
    assert got["time"] == tl.time
    assert np.array_equal(got["y0"], ctx.get_array(runtime.Y0))
    assert np.array_equal(got["accel"], ctx.get_array(runtime.ACCEL))
    rm3, nn, _, mig = ctx.side_outputs()
    assert np.array_equal(got["rm3"], rm3) and np.array_equal(got["nn"], nn) and np.array_equal(got["mig"], mig)
    # the final Compute() of the drop-in against the oracle.  After a Driver
    # call the migration switches are off, so Compute re-adds the migration
    # cached at the last call's first evaluation (time - hDid, the previous
    # state, which the device Y holds after the y0/y swap): prime the
    # oracle's cache there first.
    o = Oracle(scn)
    o.set_mig_type(got["mig"])
    o.set_evaluate(1, 1, 1)
    o.compute(tl.time - tl.hDid, ctx.get_array(runtime.Y))
    o.set_evaluate(1, 0, 0)
    ref = o.compute(got["time"], got["y0"])
    num = np.linalg.norm(got["dydt"][:, 3:] - ref[:, 3:], axis=1)
    den = np.linalg.norm(ref[:, 3:], axis=1)
    assert (num[1:] / den[1:]).max() <= 1e-12
