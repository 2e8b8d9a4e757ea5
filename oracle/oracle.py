"""ctypes front end of the CPU oracle (``hexa_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
and by the ``cpu_baseline`` leg of ``bench.py``, never by the package.  Each
wrapper names the reference function it restates; the C source carries the
file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhexa_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class UpdateParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "E", "l1_allow", "min_density", "max_density", "max_domain_dim", "filter_tol",
        "R_init_perc", "R_interm_perc", "R_final_perc", "R_interm_iter_perc",
        "min_radius_factor", "final_threshold", "max_culling_ratio")] + [
        ("iter", C.c_int), ("n_iters", C.c_int), ("hard", C.c_int), ("bug_compat", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH) or (
                os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "hexa_oracle.c"))):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_splitmix64.restype = C.c_uint64
        L.or_splitmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.or_ke_hex8.restype = C.c_int
        L.or_ke_hex8.argtypes = [C.c_double] * 4 + [C.c_int, _dp]
        L.or_helm_mats.restype = C.c_int
        L.or_helm_mats.argtypes = [C.c_double] * 3 + [C.c_int, _dp, _dp]
        L.or_stiffness_factor.restype = None
        L.or_stiffness_factor.argtypes = [C.c_int64, _dp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_void_p, _dp]
        L.or_apply_stiffness.restype = None
        L.or_apply_stiffness.argtypes = [C.c_int] * 3 + [_dp, _dp, _dp, _dp, C.c_int]
        L.or_diag.restype = None
        L.or_diag.argtypes = [C.c_int] * 3 + [_dp, _dp, C.c_int, _dp]
        L.or_pcg_cpu.restype = C.c_int
        L.or_pcg_cpu.argtypes = [C.c_int] * 3 + [_dp, _dp, _u8, _dp, C.c_double, C.c_int,
                                                 C.c_double, _dp, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.c_int]
        L.or_pcg_gpu_semantics.restype = C.c_int
        L.or_pcg_gpu_semantics.argtypes = [C.c_int] * 3 + [
            _dp, _dp, _u8, _dp, C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int,
            C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p, _dp, _dp, C.c_int]
        L.or_regrid.restype = None
        L.or_regrid.argtypes = [C.c_int] * 6 + [C.c_float] * 6 + [C.c_int, _fp, C.c_void_p, _fp, C.c_void_p]
        L.or_refmg_apply.restype = None
        L.or_refmg_apply.argtypes = [C.c_int] * 3 + [_dp, C.c_void_p, _u8, _dp, _dp, _dp]
        L.or_filter.restype = C.c_int
        L.or_filter.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [
            C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.or_filter_reg.restype = C.c_int
        L.or_filter_reg.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [
            C.c_int, C.c_double, _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.or_stress.restype = None
        L.or_stress.argtypes = [C.c_int] * 3 + [C.c_double] * 5 + [_dp, _dp] + [C.c_void_p] * 6
        L.or_update_density.restype = C.c_int
        L.or_update_density.argtypes = [_fp, _fp, _u8, _fp, C.POINTER(UpdateParams), _dp]
        L.or_simp_sensitivity.restype = C.c_double
        L.or_simp_sensitivity.argtypes = [C.c_int] * 3 + [_dp, _dp, C.c_double, C.c_double,
                                                          C.c_double, _dp, _dp]
        L.or_oc_update.restype = C.c_int
        L.or_oc_update.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                   C.c_double, _dp, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ndof(nx, ny, nz):
    return 3 * (nx + 1) * (ny + 1) * (nz + 1)


def ke_hex8(dx=1.0, dy=1.0, dz=1.0, nu=0.3, f32=False):
    """Element.get_canonical_stiffness (Element.jl:104-111)."""
    Ke = np.zeros(576)
    if lib().or_ke_hex8(dx, dy, dz, nu, int(f32), Ke) != 0:
        raise ValueError("invalid element")
    return Ke.reshape(24, 24)


def helm_mats(dx=1.0, dy=1.0, dz=1.0, bug_compat=True):
    """Element.get_scalar_canonical_matrices (Element.jl:113-134)."""
    K8, M8 = np.zeros(64), np.zeros(64)
    lib().or_helm_mats(dx, dy, dz, int(bug_compat), K8, M8)
    return K8.reshape(8, 8), M8.reshape(8, 8)


def stiffness_factor(rho, E0=1.0, Emin=0.0, penal=1.0, threshold=-np.inf, active=None):
    rho = _f64(rho)
    out = np.zeros_like(rho)
    act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    lib().or_stiffness_factor(rho.size, rho, E0, Emin, penal, threshold, _ptr(act), out)
    return out


def apply_stiffness(grid, Ke, factor, x, nchunks=0):
    """CPUSolver.apply_stiffness (CPUSolver.jl:56-122)."""
    nx, ny, nz = grid
    y = np.zeros(ndof(nx, ny, nz))
    lib().or_apply_stiffness(nx, ny, nz, _f64(Ke).ravel(), _f64(factor), _f64(x), y, nchunks)
    return y


def diag(grid, Ke, factor, mode=0):
    nx, ny, nz = grid
    d = np.zeros(ndof(nx, ny, nz))
    lib().or_diag(nx, ny, nz, _f64(Ke).ravel(), _f64(factor), mode, d)
    return d


def pcg_cpu(grid, Ke, factor, fixed, f, tol=1e-8, max_iter=10000, shift_factor=0.0, nchunks=0):
    """CPUSolver.matrix_free_cg_solve (CPUSolver.jl:192-255) in FP64."""
    nx, ny, nz = grid
    x = np.zeros(ndof(nx, ny, nz))
    rel, sh = C.c_double(), C.c_double()
    it = lib().or_pcg_cpu(nx, ny, nz, _f64(Ke).ravel(), _f64(factor),
                          np.ascontiguousarray(fixed, dtype=np.uint8), _f64(f), tol, max_iter,
                          shift_factor, x, C.byref(rel), C.byref(sh), nchunks)
    return x, rel.value, it, sh.value


def regrid(old_dims, new_dims, new_h, lengths, rho, hard=False, alpha=None):
    """MeshRefiner density remap (MeshRefiner.jl:185-237), Float32."""
    out = np.zeros(int(np.prod(new_dims)), dtype=np.float32)
    aout = np.zeros_like(out) if alpha is not None else None
    a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float32)
    lib().or_regrid(*[int(v) for v in old_dims], *[int(v) for v in new_dims], *[float(v) for v in new_h],
                    *[float(v) for v in lengths], int(bool(hard)), np.ascontiguousarray(rho, dtype=np.float32),
                    _ptr(a), out, _ptr(aout))
    return out if alpha is None else (out, aout)


def refmg_kd8(dx=1.0, dy=1.0, dz=1.0):
    """x-entry diagonal of the coarse brick's Ke (twice the fine size, E = 1,
    nu = 0.3, Float32 evaluation) at the 8 corners
    (R/src/Solvers/GPUGeometricMultigrid.jl:299-302, Element.jl:104-111)."""
    Ke = ke_hex8(2.0 * dx, 2.0 * dy, 2.0 * dz, 0.3, f32=True)
    return np.array([Ke[3 * a, 3 * a] for a in range(8)])


def refmg_apply(grid, kd8, fixed, minv, r, present=None):
    """The reference's additive 2-level preconditioner z = M r
    (GPUGeometricMultigrid.jl:253-346)."""
    nx, ny, nz = grid
    z = np.zeros(ndof(nx, ny, nz))
    pr = None if present is None else np.ascontiguousarray(present, dtype=np.uint8)
    lib().or_refmg_apply(nx, ny, nz, _f64(kd8), _ptr(pr), np.ascontiguousarray(fixed, dtype=np.uint8),
                         _f64(minv), _f64(r), z)
    return z


def pcg_gpu_semantics(grid, Ke, factor, fixed, f, u_guess=None, tol=1e-8, max_iter=10000,
                      shift_factor=1e-10, max_attempts=3, shift_multiplier=10.0,
                      stagnation_tol=0.0, diag_mode=1, refmg_kd8=None, present=None, nchunks=0):
    """GPUSolver.gpu_matrix_free_cg_solve control flow, FP64: Jacobi, or the
    reference 2-level preconditioner (with its Jacobi fallback) when
    ``refmg_kd8`` is given."""
    nx, ny, nz = grid
    x = np.zeros(ndof(nx, ny, nz))
    info = np.zeros(5)
    g = None if u_guess is None else _f64(u_guess)
    kd = None if refmg_kd8 is None else _f64(refmg_kd8)
    pr = None if present is None else np.ascontiguousarray(present, dtype=np.uint8)
    st = lib().or_pcg_gpu_semantics(nx, ny, nz, _f64(Ke).ravel(), _f64(factor),
                                    np.ascontiguousarray(fixed, dtype=np.uint8), _f64(f), _ptr(g),
                                    tol, max_iter, shift_factor, max_attempts, shift_multiplier,
                                    stagnation_tol, diag_mode, _ptr(kd), _ptr(pr), x, info, nchunks)
    return x, st, dict(rel_res=info[0], iterations=int(info[1]), attempt=int(info[2]),
                       shift=info[3], mg_fallback=bool(info[4]))


def filter_density(grid, h, radius, rho, bug_compat=True, tol=1e-6, max_iter=200, check_every=10):
    """GPUHelmholtz.solve_helmholtz_on_gpu semantics (GPUHelmholtz.jl:190-266) in FP64."""
    nx, ny, nz = grid
    out = np.zeros(nx * ny * nz)
    rel = C.c_double()
    it = lib().or_filter(nx, ny, nz, h[0], h[1], h[2], radius, int(bug_compat), _f64(rho), out,
                         tol, max_iter, check_every, C.byref(rel))
    return out, it, rel.value


def filter_density_reg(grid, h, radius, rho, reg, bug_compat=True, tol=1e-13, max_iter=20000, check_every=1):
    """The filter with reg I added to the operator: the reference's CPU
    filter, R^2 K + M + 1e-9 I (R/src/Optimization/TopOpt.jl:32-127)."""
    nx, ny, nz = grid
    out = np.zeros(nx * ny * nz)
    rel = C.c_double()
    it = lib().or_filter_reg(nx, ny, nz, h[0], h[1], h[2], radius, int(bug_compat), float(reg), _f64(rho), out,
                             tol, max_iter, check_every, C.byref(rel))
    return out, it, rel.value


def filter_blocked(grid, h, radius, rho, bug_compat=True, target_block=50000, tol=1e-13):
    """apply_domain_decomposed_cpu_filter! (R/src/Optimization/TopOpt.jl:129-209):
    ceil(n / cbrt(target)) blocks per axis of ceil(n / chunks) elements, each
    filtered alone (R^2 K + M + 1e-9 I, exact solve) on its box widened by
    max(2, ceil(3R / min h)) cells and clipped to the grid; the block cores
    are copied back."""
    nx, ny, nz = grid
    n = (nx, ny, nz)
    rho = _f64(rho).reshape(nz, ny, nx)
    cube = float(np.cbrt(target_block))
    chunks = [max(1, int(np.ceil(v / cube))) for v in n]
    blk = [-(-v // c) for v, c in zip(n, chunks)]
    halo = max(2, int(np.ceil(3.0 * radius / min(h))))
    out = np.zeros((nz, ny, nx))
    for cz in range(chunks[2]):
        for cy in range(chunks[1]):
            for cx in range(chunks[0]):
                c = (cx, cy, cz)
                s0 = [c[a] * blk[a] for a in range(3)]
                s1 = [min((c[a] + 1) * blk[a], n[a]) for a in range(3)]
                if any(s0[a] >= s1[a] for a in range(3)):
                    continue
                h0 = [max(0, s0[a] - halo) for a in range(3)]
                h1 = [min(n[a], s1[a] + halo) for a in range(3)]
                loc = rho[h0[2]:h1[2], h0[1]:h1[1], h0[0]:h1[0]]
                lg = (h1[0] - h0[0], h1[1] - h0[1], h1[2] - h0[2])
                res, _, _ = filter_density_reg(lg, h, radius, loc.ravel(), 1e-9, bug_compat, tol=tol)
                res = res.reshape(lg[2], lg[1], lg[0])
                out[s0[2]:s1[2], s0[1]:s1[1], s0[0]:s1[0]] = res[s0[2] - h0[2]:s1[2] - h0[2],
                                                                  s0[1] - h0[1]:s1[1] - h0[1],
                                                                  s0[0] - h0[0]:s1[0] - h0[0]]
    return out.ravel()


def stress(grid, h, E, nu, U, rho):
    """Stress.stress_kernel! (Stress.jl:159-285) in FP64."""
    nx, ny, nz = grid
    nE = nx * ny * nz
    pr, vm, l1 = np.zeros(3 * nE), np.zeros(nE), np.zeros(nE)
    dmax, dmin, vo = np.zeros(3 * nE), np.zeros(3 * nE), np.zeros(6 * nE)
    lib().or_stress(nx, ny, nz, h[0], h[1], h[2], E, nu, _f64(U), _f64(rho),
                    _ptr(pr), _ptr(vm), _ptr(l1), _ptr(dmax), _ptr(dmin), _ptr(vo))
    return dict(principal=pr.reshape(nE, 3), vonmises=vm, l1=l1, dir_max=dmax.reshape(nE, 3),
                dir_min=dmin.reshape(nE, 3), voigt=vo.reshape(nE, 6))


def update_density(density, l1, protected, original, params: dict):
    """TopologyOptimization.update_density! (TopOpt.jl:211-482)."""
    d = np.ascontiguousarray(density, dtype=np.float32).copy()
    P = UpdateParams(**params)
    out = np.zeros(6)
    skipped = lib().or_update_density(d, np.ascontiguousarray(l1, dtype=np.float32),
                                      np.ascontiguousarray(protected, dtype=np.uint8),
                                      np.ascontiguousarray(original, dtype=np.float32),
                                      C.byref(P), out)
    return d, dict(mean_change=out[0], R_length=out[1], threshold=out[2],
                   filter_iters=int(out[3]), filter_res=out[4], culled=int(out[5]),
                   skipped=bool(skipped))


def simp_sensitivity(grid, Ke, rho_phys, E0, Emin, penal, U):
    nx, ny, nz = grid
    dc = np.zeros(nx * ny * nz)
    c = lib().or_simp_sensitivity(nx, ny, nz, _f64(Ke).ravel(), _f64(rho_phys), E0, Emin, penal,
                                  _f64(U), dc)
    return c, dc


def oc_update(x, dc, dv, w, volfrac, move, bisect_tol=1e-3):
    x = _f64(x)
    xn = np.zeros_like(x)
    lam = C.c_double()
    steps = lib().or_oc_update(x.size, x, _f64(dc), _f64(dv), _f64(w), volfrac, move, bisect_tol,
                               xn, C.byref(lam))
    return xn, lam.value, steps
