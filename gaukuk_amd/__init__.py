"""gaukuk_amd: an MI355X-native implementation of Gaukuk's per-timestep hydro update.

The compute path is the HIP library ``gaukuk_amd/libgaukuk_hip.so`` (C ABI declared in
include/gaukuk.h), reached from Python through :mod:`gaukuk_amd.hip`.  Importing the package
does not load the library; the first use of :mod:`gaukuk_amd.hip` does, and raises if it is
missing.  The C++ host driver that mirrors the reference's Sim lives in gaukuk_amd/host/.
"""
from . import config, gkio, problems  # noqa: F401

__all__ = ["config", "gkio", "problems"]
