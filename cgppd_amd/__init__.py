"""cgppd_amd -- MI355X-native REMC energy hot path of cgppd.

The product is native: libcgppd_hip.so (gfx950 kernels behind the C ABI in
include/cgppd_hip.h) and libcgppd_host.so (C++ Replica/Simulation API).
This Python package holds the in-tree build, a ctypes handle over the C ABI,
the reference data-format readers and the seeded synthetic systems used by
tests and the benchmark.
"""
__all__ = ["build", "engine", "formats", "synth"]
