"""solaris_amd -- MI355X-native force-evaluation and integration path of
the Solaris N-body simulator.

The product is the shared library solaris_amd/libsolaris_amd.so (HIP
kernels for gfx950 behind the C ABI in include/solaris_amd.h) and the C++
classes of solaris_amd/dropin that keep the reference's Acceleration /
integrator / BodyData interfaces.  This package gives Python access to the
same library (runtime.Context) and the seeded synthetic inputs
(scenarios)."""

from . import scenarios  # noqa: F401

__all__ = ["scenarios", "runtime"]
