"""Independent numpy/scipy references used by the tests (assembled sparse
operators and direct solves).  They restate the mathematics, not the oracle."""
import ctypes as C
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import assemble, elem_nodes, filter_direct, helm_exact, ke_exact  # noqa: E402,F401


def cantilever_arrays(nx, ny, nz, seed, lo=0.001):
    from v119_2025_12_19_hexatop_aws_amd import problems as P
    p = P.cantilever(nx, ny, nz, seed=seed, lo=lo)
    return p.fixed, p.f[0], p.rho


def direct_solve(nx, ny, nz, Ke, factor, fixed, f, shift_factor=1e-10):
    K = assemble(nx, ny, nz, Ke, factor)
    free = np.flatnonzero(np.asarray(fixed) == 0)
    Kff = K[free][:, free]
    s = shift_factor * Kff.diagonal().max()
    u = np.zeros(K.shape[0])
    u[free] = spla.spsolve((Kff + s * sp.identity(len(free))).tocsc(), np.asarray(f)[free])
    return u, s


def hip_runtime():
    """The HIP runtime through ctypes (device buffers without torch)."""
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            L = C.CDLL(name)
            break
        except OSError:
            continue
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DevBuf:
    """A device buffer holding a copy of a host array."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, a
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), a.nbytes) == 0
        assert hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def read(self):
        out = np.empty_like(self.a)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)
