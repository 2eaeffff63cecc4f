"""cgppd_energy_all_noalloc, the allocation-free variant of cgppd_energy_all,
through the C ABI: the energies and molecule-pair matrices it writes into
caller-owned device buffers equal cgppd_energy_all's on two calls in a row,
the second on a smaller system with fewer replicas (the library's scratch
array keeps the first call's sums, so a missing clear would show).

Both entry points run the same FP32 pair kernels; only the order of the
FP64 atomic additions differs, so the bar is 1e-10 relative."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-10, 1e-12


class DeviceBuffer:
    """Caller-owned device memory through the HIP runtime (the test's own
    allocation; the entry point under test allocates nothing)."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        # poison: NaN bytes, so an output the call leaves unwritten shows
        assert hip.hipMemset(self.ptr, 0xFF, C.c_size_t(8 * n)) == 0
        assert hip.hipDeviceSynchronize() == 0

    def read(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    return lib


def check_call(e, hip):
    R, M = e.n_rep, e.n_mol
    lj, dh, tot, mat = (DeviceBuffer(hip, n) for n in (R, R, R, R * M * M))
    rc = e.lib.cgppd_energy_all_noalloc(e.h, lj.ptr, dh.ptr, tot.ptr, mat.ptr)
    assert rc == 0, e.lib.cgppd_last_error(e.h)
    e.synchronize()
    got = [b.read() for b in (lj, dh, tot, mat)]
    for b in (lj, dh, tot, mat):
        b.free()
    want = e.energy_all(matrix=True)
    for name, g, w in zip(("lj", "dh", "total", "matrix"), got, want):
        w = np.asarray(w).reshape(-1)
        print(f"R={R} M={M} {name}: max rel diff {np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-30)):.2e}",
              flush=True)
        np.testing.assert_allclose(g, w, rtol=RTOL, atol=ATOL)
    return got


def test_noalloc_energy_matches_energy_all(gpu_engine_factory, hip):
    from cgppd_amd import synth
    # a flexible molecule of linked domains and two rigid chains (the
    # intra-molecular kernel also adds into the scratch sums), 6 replicas
    big = synth.make_linked_system("noalloc_a", 2, 30, 6, 0.2, 120.0, 6, 288.0, 450.0, 41, extra_rigid=(25, 40))
    e = gpu_engine_factory()
    e.load_spec(big, seed=3)
    e.mc_prepare()
    e.mc_steps(40)  # replicas differ from each other
    first = check_call(e, hip)
    e.close()
    # second call: a smaller rigid system with fewer replicas
    small = synth.make_system("noalloc_b", [22, 31, 9], 70.0, 3, 288.0, 450.0, 5, 2.0, 0.2)
    f = gpu_engine_factory()
    f.load_spec(small, seed=8)
    f.mc_prepare()
    f.mc_steps(25)
    second = check_call(f, hip)
    assert len(second[0]) == 3 and len(first[0]) == 6
    # the running totals of the MC state agree too
    np.testing.assert_allclose(second[2], f.stats()["energy"], rtol=1e-6, atol=1e-9)
    f.close()


def test_noalloc_rejects_bad_state(gpu_engine_factory):
    e = gpu_engine_factory()
    assert e.lib.cgppd_energy_all_noalloc(e.h, None, None, None, None) == -4  # nothing set up
    e.close()
