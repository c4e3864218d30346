"""mdx_kinetics_noalloc: the kinetics reductions over caller-owned device
arrays with the library's partial sums in a file-scope device array (no
allocation in the call).  Its output must equal mdx_kinetics's on the same
state bit for bit (same kernels, same block partition, same order), on two
calls in a row, the second with fewer atoms; sizes beyond its limit and bad
arguments are rejected with an error.  The test's device memory comes from
the HIP runtime libmdx.so itself links (by soname, so the loader hands back
that same copy)."""
import ctypes as C

import numpy as np
import pytest

from moleculardynamics_amd import BOND_REFERENCE, BOX_WALLED, LJ_REFERENCE, PREC_FP64, System, _lib

pytestmark = pytest.mark.gpu

MDX_ERR_ARG, MDX_ERR_CAPACITY = -1, -5
KB = float(np.float32(8.314e-7))
# the reference's masses by species (common/Constants.h:13-15), as the engine holds them
MASS4 = np.array([np.float32(14.00674), np.float32(15.9994), np.float32(20.1797), np.float32(20.1797)],
                 dtype=np.float64)


def entry():
    lib = _lib.load()
    f = lib.mdx_kinetics_noalloc
    P = C.c_void_p
    f.argtypes = [P, P, P, C.c_int32, C.c_int64, P, C.c_double, P, P]
    f.restype = C.c_int
    lib.mdx_kinetics_noalloc_last_error.argtypes = []
    lib.mdx_kinetics_noalloc_last_error.restype = C.c_char_p
    return f, lib.mdx_kinetics_noalloc_last_error


def state(n, seed):
    rng = np.random.default_rng(seed)
    pos4 = np.c_[rng.uniform(-20, 20, (n, 3)), rng.integers(0, 3, n)]
    vel = rng.normal(0, 0.01, (n, 3))
    return pos4, vel


def reference(pos4, vel):
    n = len(pos4)
    with System(n_atoms=n, n_bond=0, box=BOX_WALLED, precision=PREC_FP64, use_lj=1, use_bond=0,
                lj_table=LJ_REFERENCE, bond_form=BOND_REFERENCE, dt=1.0, lj_rs=7.6, lj_rcut=6.8,
                wall=100.0) as s:
        s.set_state(pos4, vel)
        return s.kinetics(KB)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libmdx.so uses."""

    def __init__(self):
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.held = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0
        self.held.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # hipMemcpyHostToDevice
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
        return out

    def close(self):
        for p in self.held:
            self.h.hipFree(p)
        self.held = []


def test_kinetics_noalloc_equals_kinetics():
    f, last_error = entry()
    mass = np.ascontiguousarray(MASS4)
    hip = Hip()
    try:
        for call, n in enumerate((5000, 1234)):   # the second call with a smaller input
            pos4, vel = state(n, seed=call)
            want = reference(pos4, vel)
            stride = (n + 63) // 64 * 64          # the engine's component stride
            pos = np.zeros((3, stride))
            v = np.zeros((3, stride))
            pos[:, :n] = pos4[:, :3].T
            v[:, :n] = vel.T
            dpos, dv = hip.upload(pos), hip.upload(v)
            dtyp = hip.upload(pos4[:, 3].astype(np.int32))
            nan8 = np.full(8, np.nan)
            dout = hip.upload(nan8)
            print(f"mdx_kinetics_noalloc call {call + 1}: {n} atoms", flush=True)
            rc = f(dpos, dv, dtyp, n, stride, mass.ctypes.data, KB, dout, None)
            assert rc == 0, last_error().decode()
            assert hip.h.hipDeviceSynchronize() == 0
            np.testing.assert_array_equal(hip.download(dout, nan8), want)
    finally:
        hip.close()


def test_kinetics_noalloc_rejects_bad_sizes():
    f, last_error = entry()
    mass = np.ascontiguousarray(MASS4)
    dummy = C.c_void_p(64)   # never dereferenced: the arguments are rejected first
    rc = f(dummy, dummy, dummy, (1 << 28) + 1, (1 << 28) + 1, mass.ctypes.data, KB, dummy, None)
    assert rc == MDX_ERR_CAPACITY and b"exceeds the limit" in last_error()
    rc = f(dummy, dummy, dummy, 0, 64, mass.ctypes.data, KB, dummy, None)
    assert rc == MDX_ERR_ARG
    rc = f(dummy, dummy, dummy, 100, 99, mass.ctypes.data, KB, dummy, None)
    assert rc == MDX_ERR_ARG and b"stride" in last_error()
    bad = np.array([1.0, 0.0, 1.0, 1.0])
    rc = f(dummy, dummy, dummy, 100, 128, bad.ctypes.data, KB, dummy, None)
    assert rc == MDX_ERR_ARG and b"masses" in last_error()
