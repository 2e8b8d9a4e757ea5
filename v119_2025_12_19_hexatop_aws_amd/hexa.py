"""Python host side of libhexa: ctypes bindings over ``include/hexa.h`` and a
mirror of the reference's solver entry points.

The reference is Julia (no toolchain here); its GPU entry points are mirrored
with the same names, argument meaning and return conventions:

* ``solve_system_gpu``  -- GPUSolver.solve_system_gpu
  (R/src/Solvers/GPUSolver.jl:467-477): returns ``(U, rel_res, method)``;
  a NaN/Inf diagonal returns ``(zeros, nan, "Failed_Corrupt_Input")``
  (GPUSolver.jl:310-315); a non-converged solve returns the best iterate
  (GPUSolver.jl:439-443).
* ``solve_system``      -- Solver.solve_system (R/src/Solvers/Solver.jl:39-160):
  tolerance/shift read from ``config["solver_parameters"]``, void pruning,
  dispatch.

There is no CPU fallback: if ``libhexa.so`` is missing the import fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HEXA_LIB_PATH") or os.path.join(_HERE, "libhexa.so")

OK, E_INVALID, E_OOM, E_NAN_INPUT, E_NOT_CONVERGED, E_DEVICE = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_MIXED = 0, 1
KE_FP64, KE_F32_WIDENED = 0, 1
HELM_BUG_COMPAT, HELM_CORRECTED = 0, 1
PRECOND_JACOBI, PRECOND_MG, PRECOND_REF_MG = 0, 1, 2
MATVEC_AUTO, MATVEC_DENSE, MATVEC_FACTORED = 0, 1, 2
SEMANTICS_GPU, SEMANTICS_CPU = 0, 1
DIAG_EXACT, DIAG_X_ENTRY = 0, 1
RHS_AUTO, RHS_SEQUENTIAL = 0, 1
CYCLE_V, CYCLE_W = 0, 1
METHOD_NAMES = {0: "Jacobi_GPU", 1: "GMG_GPU", 2: "Failed_Corrupt_Input", 3: "Jacobi_CPU", 4: "None"}


class HexaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libhexa status {status}: {msg}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class Options(C.Structure):
    _fields_ = [("E0", C.c_double), ("Emin", C.c_double), ("penal", C.c_double), ("nu", C.c_double),
                ("ke_mode", C.c_int32), ("helm_element", C.c_int32), ("precision", C.c_int32),
                ("matvec_kernel", C.c_int32), ("mg_levels", C.c_int32), ("mg_pre", C.c_int32),
                ("mg_post", C.c_int32), ("mg_omega", C.c_double), ("nrhs", C.c_int32),
                ("diag_mode", C.c_int32), ("mg_cycle", C.c_int32), ("mg_coarse_dofs", C.c_int32),
                ("reserved", C.c_int64 * 7)]


class SolveParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("shift_factor", C.c_double),
                ("max_shift_attempts", C.c_int32), ("shift_multiplier", C.c_double),
                ("stagnation_tol", C.c_double), ("precond", C.c_int32), ("semantics", C.c_int32),
                ("warm_start", C.c_int32), ("check_every", C.c_int32), ("rhs_mode", C.c_int32),
                ("pad", C.c_int32), ("reserved", C.c_int64 * 3)]


class SolveReport(C.Structure):
    _fields_ = [("rel_res", C.c_double), ("true_rel_res", C.c_double), ("shift_used", C.c_double),
                ("compliance", C.c_double), ("setup_ms", C.c_double), ("solve_ms", C.c_double),
                ("iterations", C.c_int32), ("attempt", C.c_int32), ("method", C.c_int32),
                ("converged", C.c_int32), ("mg_levels", C.c_int32), ("pad", C.c_int32),
                ("matvecs", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}
        d["method_name"] = METHOD_NAMES.get(self.method, str(self.method))
        return d


class FilterReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("pad", C.c_int32), ("rel_res", C.c_double), ("ms", C.c_double)]


class OcParams(C.Structure):
    _fields_ = [("volfrac", C.c_double), ("move", C.c_double), ("rmin", C.c_double), ("filter_tol", C.c_double),
                ("bisect_tol", C.c_double), ("filter_max_iter", C.c_int32), ("pad", C.c_int32),
                ("solve", SolveParams)]


class OcReport(C.Structure):
    _fields_ = [("compliance", C.c_double), ("volume", C.c_double), ("change", C.c_double), ("lambda_", C.c_double),
                ("solve_ms", C.c_double), ("step_ms", C.c_double), ("solve_iterations", C.c_int32),
                ("filter_iterations", C.c_int32), ("status", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {k.rstrip("_"): getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class UpdateParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "E", "l1_allow", "min_density", "max_density", "max_domain_dim", "filter_tol", "R_init_perc",
        "R_interm_perc", "R_final_perc", "R_interm_iter_perc", "min_radius_factor", "final_threshold",
        "max_culling_ratio")] + [("iter", C.c_int32), ("n_iters", C.c_int32), ("hard", C.c_int32),
                                 ("pad", C.c_int32)]


class UpdateReport(C.Structure):
    _fields_ = [("mean_change", C.c_double), ("R_length", C.c_double), ("threshold", C.c_double),
                ("filter_res", C.c_double), ("filter_iters", C.c_int32), ("skipped", C.c_int32),
                ("culled", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("level", C.c_int32), ("launches", C.c_int32), ("ms", C.c_double),
                ("bytes", C.c_double), ("gbps", C.c_double)]

    def as_dict(self):
        return dict(name=self.name.decode(), level=self.level, launches=self.launches, ms=self.ms,
                    bytes=self.bytes, GBps=self.gbps)


class Job(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("support_face", C.c_int32),
                ("n_loads", C.c_int32), ("oc_iters", C.c_int32), ("precision", C.c_int32), ("pad", C.c_int32),
                ("volfrac", C.c_double), ("move", C.c_double), ("rmin", C.c_double), ("tol", C.c_double),
                ("seed", C.c_uint64)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("oc_iters_done", C.c_int32),
                ("solve_iterations", C.c_int32), ("pad", C.c_int32), ("compliance", C.c_double),
                ("volume", C.c_double), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_end_ms", C.c_double), ("est_bytes", C.c_int64), ("device_bytes", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


_H = C.c_void_p
_P = C.c_void_p

# symbol -> (restype, argtypes); the list of everything include/hexa.h declares
SIGNATURES = {
    "hexa_last_error": (C.c_char_p, []),
    "hexa_version": (C.c_int32, []),
    "hexa_default_options": (None, [C.POINTER(Options)]),
    "hexa_default_solve_params": (None, [C.POINTER(SolveParams)]),
    "hexa_device_info": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "hexa_create": (C.c_int32, [C.POINTER(Grid), C.POINTER(Options), C.POINTER(_H)]),
    "hexa_destroy": (C.c_int32, [_H]),
    "hexa_set_problem": (C.c_int32, [_H, _P, _P, _P]),
    "hexa_set_density": (C.c_int32, [_H, _P]),
    "hexa_set_density_dev": (C.c_int32, [_H, _P]),
    "hexa_prune_voids": (C.c_int32, [_H, C.c_double, C.POINTER(C.c_int64)]),
    "hexa_solve": (C.c_int32, [_H, _P, C.POINTER(SolveParams), _P]),
    "hexa_get_displacement": (C.c_int32, [_H, _P]),
    "hexa_displacement_dev": (C.c_int32, [_H, C.POINTER(C.c_void_p)]),
    "hexa_matvec": (C.c_int32, [_H, _P, _P]),
    "hexa_matvec_dev": (C.c_int32, [_H, _P, _P, C.c_int32, C.c_int32]),
    "hexa_apply_precond": (C.c_int32, [_H, _P, _P]),
    "hexa_last_kernel_ms": (C.c_int32, [_H, C.POINTER(C.c_double)]),
    "hexa_device_ceilings": (C.c_int32, [C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
    "hexa_kernel_report": (C.c_int32, [_H, C.c_int32, C.POINTER(KernelStat), C.c_int32, C.POINTER(C.c_int32)]),
    "hexa_stream_sync": (C.c_int32, [_H]),
    "hexa_filter": (C.c_int32, [_H, _P, _P, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                C.POINTER(FilterReport)]),
    "hexa_filter_fallback": (C.c_int32, [_H, _P, _P, C.c_double, C.c_int64, C.c_int32, C.POINTER(FilterReport)]),
    "hexa_stress": (C.c_int32, [_H, _P, _P, C.c_double, _P, _P, _P, _P, _P, _P]),
    "hexa_sensitivity": (C.c_int32, [_H, _P, C.POINTER(C.c_double)]),
    "hexa_oc_update": (C.c_int32, [_H, _P, _P, _P, _P, C.c_double, C.c_double, C.c_double,
                                   C.POINTER(C.c_double)]),
    "hexa_oc_update_noalloc": (C.c_int32, [C.c_int64, _P, _P, _P, _P, C.c_double, C.c_double, C.c_double, _P,
                                           C.POINTER(C.c_double)]),
    "hexa_oc_step": (C.c_int32, [_H, _P, C.POINTER(OcParams), C.POINTER(OcReport)]),
    "hexa_oc_step_dev": (C.c_int32, [_H, C.c_void_p, C.POINTER(OcParams), C.POINTER(OcReport)]),
    "hexa_update_density": (C.c_int32, [_H, _P, _P, _P, _P, C.POINTER(UpdateParams), C.POINTER(UpdateReport)]),
    "hexa_set_update_fields": (C.c_int32, [_H, _P, _P]),
    "hexa_write_vti": (C.c_int32, [_H, C.c_char_p, _P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]),
    "hexa_write_bin": (C.c_int32, [_H, C.c_char_p, _P, C.c_double, C.c_int32, C.c_float, C.c_float, C.c_int64,
                                   C.c_int32, C.c_char_p]),
    "hexa_regrid_density": (C.c_int32, [C.POINTER(Grid), _P, _P, C.POINTER(Grid), C.c_double, C.c_double,
                                        C.c_double, C.c_int32, _P, _P, _P]),
    "hexa_regrid_density_host": (C.c_int32, [C.POINTER(Grid), _P, _P, C.POINTER(Grid), C.c_double, C.c_double,
                                             C.c_double, C.c_int32, _P, _P]),
    "hexa_ref_step": (C.c_int32, [_H, C.POINTER(SolveParams), C.c_double, C.c_double, C.POINTER(UpdateParams),
                                  C.POINTER(SolveReport), C.POINTER(UpdateReport), _P]),
    "hexa_estimate_bytes": (C.c_int64, [C.POINTER(Grid), C.POINTER(Options)]),
    "hexa_server_create": (C.c_int32, [C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]),
    "hexa_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_int64)]),
    "hexa_server_wait": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(JobResult), _P]),
    "hexa_server_clock_ms": (C.c_double, [C.c_void_p]),
    "hexa_server_destroy": (C.c_int32, [C.c_void_p]),
}

_lib = None


def lib():
    """Load libhexa.so (built in-tree); fails loudly when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                              "or python -m v119_2025_12_19_hexatop_aws_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().hexa_last_error().decode(errors="replace")


def _check(status, allow=(OK,)):
    if status not in allow:
        raise HexaError(status, last_error())
    return status


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def device_info():
    name = C.create_string_buffer(256)
    tot, ncu = C.c_int64(), C.c_int32()
    _check(lib().hexa_device_info(name, 256, C.byref(tot), C.byref(ncu)))
    return dict(name=name.value.decode(), total_bytes=tot.value, n_cu=ncu.value)


def device_ceilings(nbytes=1 << 30):
    """Measured copy / triad GB/s and FP64 FMA TFLOP/s of the current GPU."""
    c, t, f = C.c_double(), C.c_double(), C.c_double()
    _check(lib().hexa_device_ceilings(int(nbytes), C.byref(c), C.byref(t), C.byref(f)))
    return dict(copy_GBps=c.value, triad_GBps=t.value, fp64_fma_TFLOPs=f.value, array_bytes=int(nbytes))


def default_options(**kw) -> Options:
    o = Options()
    lib().hexa_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_params(**kw) -> SolveParams:
    p = SolveParams()
    lib().hexa_default_solve_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Solver:
    """One libhexa handle: a structured grid with device-resident state."""

    def __init__(self, nx, ny, nz, h=(1.0, 1.0, 1.0), **options):
        self.grid = (int(nx), int(ny), int(nz))
        self.h = tuple(float(v) for v in h)
        self.nn = (nx + 1) * (ny + 1) * (nz + 1)
        self.ne = nx * ny * nz
        self.options = default_options(**options)
        g = Grid(nx, ny, nz, *self.h)
        hd = C.c_void_p()
        _check(lib().hexa_create(C.byref(g), C.byref(self.options), C.byref(hd)))
        self._h = hd

    @property
    def ndof(self):
        return 3 * self.nn

    def close(self):
        if self._h:
            lib().hexa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_problem(self, fixed, f, active=None):
        fixed = np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if fixed.size != self.ndof or f.size != self.ndof * self.options.nrhs:
            raise ValueError("fixed/f sizes do not match the grid")
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8).reshape(-1)
        if act is not None and act.size != self.ne:
            raise ValueError("active mask size does not match the grid")
        _check(lib().hexa_set_problem(self._h, _ptr(fixed), _ptr(f), _ptr(act)))

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        if rho.size != self.ne:
            raise ValueError("density size does not match the grid")
        _check(lib().hexa_set_density(self._h, _ptr(rho)))

    def set_density_dev(self, ptr: int):
        _check(lib().hexa_set_density_dev(self._h, C.c_void_p(ptr)))

    def solve(self, u0=None, **params):
        """Returns (u [nrhs, 3nN], reports list, status)."""
        prm = default_params(**params)
        nrhs = self.options.nrhs
        u = np.zeros((nrhs, self.ndof)) if u0 is None else np.array(u0, dtype=np.float64).reshape(nrhs, -1)
        reps = (SolveReport * nrhs)()
        st = lib().hexa_solve(self._h, _ptr(u), C.byref(prm), C.cast(reps, C.c_void_p))
        _check(st, allow=(OK, E_NOT_CONVERGED, E_NAN_INPUT))
        return u, [r.as_dict() for r in reps], st

    def prune_voids(self, min_stiffness_threshold):
        """Device-side void pruning (Solver.jl:85-116); returns the number of
        active elements (nE when nothing was pruned); a negative threshold
        makes every element active again."""
        n = C.c_int64()
        _check(lib().hexa_prune_voids(self._h, float(min_stiffness_threshold), C.byref(n)))
        return n.value

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.zeros_like(x)
        _check(lib().hexa_matvec(self._h, _ptr(x), _ptr(y)))
        return y

    def matvec_dev(self, xptr: int, yptr: int, precision=PREC_FP64, reps=1):
        _check(lib().hexa_matvec_dev(self._h, C.c_void_p(xptr), C.c_void_p(yptr), precision, reps))
        ms = C.c_double()
        _check(lib().hexa_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def kernel_report(self, reps=10):
        """Per-kernel GPU time and algorithmic HBM rate on this handle's grid
        (after a multigrid solve); a list of dicts."""
        out = (KernelStat * 32)()
        n = C.c_int32()
        _check(lib().hexa_kernel_report(self._h, int(reps), out, 32, C.byref(n)))
        return [out[i].as_dict() for i in range(n.value)]

    def apply_precond(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        z = np.zeros_like(r)
        _check(lib().hexa_apply_precond(self._h, _ptr(r), _ptr(z)))
        return z

    def filter(self, rho, radius, tol=1e-6, max_iter=200, check_every=10):
        """Helmholtz filter of an element field (GPUHelmholtz.jl:190-266 semantics)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        out = np.zeros_like(rho)
        rep = FilterReport()
        _check(lib().hexa_filter(self._h, _ptr(rho), _ptr(out), float(radius), float(tol), int(max_iter),
                                 int(check_every), C.byref(rep)))
        return out, rep.iterations, rep.rel_res

    def filter_fallback(self, rho, radius, min_blocked=0, target_block=0):
        """The reference's CPU filter semantics (TopOpt.jl:319-351): global
        regularised solve below min_blocked elements, halo-blocked domain
        decomposition from there on; returns (filtered, iterations, rel)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        out = np.zeros_like(rho)
        rep = FilterReport()
        _check(lib().hexa_filter_fallback(self._h, _ptr(rho), _ptr(out), float(radius), int(min_blocked),
                                          int(target_block), C.byref(rep)))
        return out, rep.iterations, rep.rel_res

    def stress(self, u=None, rho=None, E=1.0):
        ne = self.ne
        out = dict(principal=np.zeros(3 * ne, np.float32), vonmises=np.zeros(ne, np.float32),
                   l1=np.zeros(ne, np.float32), dir_max=np.zeros(3 * ne, np.float32),
                   dir_min=np.zeros(3 * ne, np.float32), voigt=np.zeros(6 * ne, np.float32))
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        rr = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        _check(lib().hexa_stress(self._h, _ptr(uu), _ptr(rr), float(E), _ptr(out["principal"]),
                                 _ptr(out["vonmises"]), _ptr(out["l1"]), _ptr(out["dir_max"]), _ptr(out["dir_min"]),
                                 _ptr(out["voigt"])))
        for k in ("principal", "dir_max", "dir_min"):
            out[k] = out[k].reshape(ne, 3)
        out["voigt"] = out["voigt"].reshape(ne, 6)
        return out

    def sensitivity(self):
        dc = np.zeros(self.ne)
        c = C.c_double()
        _check(lib().hexa_sensitivity(self._h, _ptr(dc), C.byref(c)))
        return c.value, dc

    def oc_update(self, x, dc, dv, w, volfrac, move, bisect_tol=1e-3):
        x = np.array(x, dtype=np.float64).reshape(-1)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        dc, dv, w = f(dc), f(dv), f(w)
        lam = C.c_double()
        _check(lib().hexa_oc_update(self._h, _ptr(x), _ptr(dc), _ptr(dv), _ptr(w), float(volfrac), float(move),
                                    float(bisect_tol), C.byref(lam)))
        return x, lam.value

    def oc_step(self, x, volfrac, move, rmin, filter_tol=1e-10, filter_max_iter=500, bisect_tol=1e-12, **solve):
        x = np.array(x, dtype=np.float64).reshape(-1)
        p = OcParams(volfrac=volfrac, move=move, rmin=rmin, filter_tol=filter_tol, bisect_tol=bisect_tol,
                     filter_max_iter=filter_max_iter, solve=default_params(**solve))
        rep = OcReport()
        st = lib().hexa_oc_step(self._h, _ptr(x), C.byref(p), C.byref(rep))
        _check(st, allow=(OK, E_NOT_CONVERGED))
        return x, rep.as_dict()

    def oc_step_dev(self, d_x, volfrac, move, rmin, filter_tol=1e-10, filter_max_iter=500, bisect_tol=1e-12,
                    **solve):
        """oc_step on a design in device memory (d_x: device pointer to nE doubles, updated in place)."""
        p = OcParams(volfrac=volfrac, move=move, rmin=rmin, filter_tol=filter_tol, bisect_tol=bisect_tol,
                     filter_max_iter=filter_max_iter, solve=default_params(**solve))
        rep = OcReport()
        st = lib().hexa_oc_step_dev(self._h, C.c_void_p(int(d_x)), C.byref(p), C.byref(rep))
        _check(st, allow=(OK, E_NOT_CONVERGED))
        return rep.as_dict()

    def update_density(self, density, l1, protected, original, **params):
        d = np.array(density, dtype=np.float32).reshape(-1)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        p = UpdateParams(**params)
        rep = UpdateReport()
        _check(lib().hexa_update_density(self._h, _ptr(d), _ptr(f32(l1)),
                                         _ptr(np.ascontiguousarray(protected, dtype=np.uint8).reshape(-1)),
                                         _ptr(f32(original)), C.byref(p), C.byref(rep)))
        return d, rep.as_dict()

    def set_update_fields(self, protected, original):
        """Protected flags and original densities of the reference update,
        resident on the device for ref_step."""
        self._upd = (np.ascontiguousarray(protected, dtype=np.uint8).reshape(-1),
                     np.ascontiguousarray(original, dtype=np.float32).reshape(-1))
        _check(lib().hexa_set_update_fields(self._h, _ptr(self._upd[0]), _ptr(self._upd[1])))

    def ref_step(self, E, update, prune_threshold=-1.0, want_density=True, **solve_params):
        """One outer iteration of the reference loop on the device: solve,
        stress l1, update_density!.  Returns (density or None, solve report,
        update report, status)."""
        sp = default_params(**solve_params)
        up = UpdateParams(**update)
        sr, ur = SolveReport(), UpdateReport()
        d = np.zeros(self.ne, dtype=np.float32) if want_density else None
        st = lib().hexa_ref_step(self._h, C.byref(sp), float(E), float(prune_threshold), C.byref(up),
                                 C.byref(sr), C.byref(ur), _ptr(d))
        _check(st, allow=(OK, E_NOT_CONVERGED))
        return d, sr.as_dict(), ur.as_dict(), st

    def write_vti(self, path, E=1.0, density=None, origin=(0.0, 0.0, 0.0), write_vectors=False):
        d = None if density is None else np.ascontiguousarray(density, dtype=np.float32).reshape(-1)
        _check(lib().hexa_write_vti(self._h, str(path).encode(), _ptr(d), float(E), *[float(v) for v in origin],
                                    int(bool(write_vectors))))

    def write_bin(self, path, E=1.0, density=None, iter=0, radius=0.0, threshold=0.5, max_export_cells=0,
                  full_field=False, json="{}"):
        d = None if density is None else np.ascontiguousarray(density, dtype=np.float32).reshape(-1)
        _check(lib().hexa_write_bin(self._h, str(path).encode(), _ptr(d), float(E), int(iter), float(radius),
                                    float(threshold), int(max_export_cells), int(bool(full_field)),
                                    json.encode()))

    def displacement(self):
        u = np.zeros(self.ndof)
        _check(lib().hexa_get_displacement(self._h, _ptr(u)))
        return u

    def displacement_dev(self) -> int:
        p = C.c_void_p()
        _check(lib().hexa_displacement_dev(self._h, C.byref(p)))
        return p.value


def regrid_density_dev(old_grid, d_old, new_grid, lengths, hard=False, d_new=None, d_old_alpha=None,
                       d_new_alpha=None):
    """Density remap onto a refined grid on the device (MeshRefiner.jl:185-237).
    Grids are (nx, ny, nz, dx, dy, dz); arrays are device pointers (ints)."""
    og, ng = Grid(*old_grid), Grid(*new_grid)
    _check(lib().hexa_regrid_density(C.byref(og), C.c_void_p(d_old), C.c_void_p(d_old_alpha or 0), C.byref(ng),
                                     float(lengths[0]), float(lengths[1]), float(lengths[2]), int(bool(hard)),
                                     C.c_void_p(d_new), C.c_void_p(d_new_alpha or 0), None))


def regrid_density(old_grid, rho, new_grid, lengths, hard=False, alpha=None):
    """Host-array form of regrid_density_dev; returns the new density (and the
    new alpha when alpha is given)."""
    og, ng = Grid(*old_grid), Grid(*new_grid)
    r = np.ascontiguousarray(rho, dtype=np.float32).reshape(-1)
    out = np.zeros(ng.nx * ng.ny * ng.nz, dtype=np.float32)
    a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float32).reshape(-1)
    aout = None if alpha is None else np.zeros_like(out)
    _check(lib().hexa_regrid_density_host(C.byref(og), _ptr(r), _ptr(a), C.byref(ng), float(lengths[0]),
                                          float(lengths[1]), float(lengths[2]), int(bool(hard)), _ptr(out),
                                          _ptr(aout)))
    return out if alpha is None else (out, aout)


class Server:
    """Device-resident job server (include/hexa.h, hexa_server_*)."""

    def __init__(self, workers=8, mem_budget_bytes=200 << 30):
        h = C.c_void_p()
        _check(lib().hexa_server_create(int(workers), int(mem_budget_bytes), C.byref(h)))
        self._s = h
        self._ne = {}

    def clock_ms(self):
        return lib().hexa_server_clock_ms(self._s)

    def submit(self, nx, ny, nz, support_face=0, n_loads=1, oc_iters=5, precision=PREC_MIXED, volfrac=0.3,
               move=0.2, rmin=1.5, tol=1e-8, seed=0):
        j = Job(nx=nx, ny=ny, nz=nz, support_face=support_face, n_loads=n_loads, oc_iters=oc_iters,
                precision=precision, volfrac=volfrac, move=move, rmin=rmin, tol=tol, seed=seed)
        i = C.c_int64()
        _check(lib().hexa_server_submit(self._s, C.byref(j), C.byref(i)))
        self._ne[i.value] = nx * ny * nz
        return i.value

    def wait(self, job_id, want_x=False):
        r = JobResult()
        x = np.zeros(self._ne.pop(job_id)) if want_x else None
        st = lib().hexa_server_wait(self._s, int(job_id), C.byref(r), _ptr(x))
        d = r.as_dict()
        if st != OK:
            d["error"] = lib().hexa_last_error().decode(errors="replace")
        return (d, x) if want_x else d

    def close(self):
        if self._s:
            lib().hexa_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---------------------------------------------------------------- mirrors
def _geometry(config):
    g = config["geometry"]
    nx, ny, nz = int(g["nElem_x_computed"]), int(g["nElem_y_computed"]), int(g["nElem_z_computed"])
    h = (float(g.get("dx_computed", 1.0)), float(g.get("dy_computed", 1.0)), float(g.get("dz_computed", 1.0)))
    return nx, ny, nz, h


def _lattice(nodes, elements, nx, ny, nz, h):
    """Map the reference's (possibly pruned) node/element arrays onto the
    structured box: node ids by rounding coordinates to the lattice, element
    ids by their first node (n1 = corner (i,j,k), Mesh.jl:44-55)."""
    ijk = np.rint(np.asarray(nodes, dtype=np.float64) / np.asarray(h)).astype(np.int64)
    nid = ijk[:, 0] + (nx + 1) * (ijk[:, 1] + (ny + 1) * ijk[:, 2])
    first = ijk[np.asarray(elements)[:, 0] - 1]
    eid = first[:, 0] + nx * (first[:, 1] + ny * first[:, 2])
    return nid, eid


def solve_system_gpu(nodes, elements, E, nu, bc, f, density, *, max_iter=40000, tol=1e-6,
                     method="native", solver="cg", use_precond=True, shift_factor=1.0e-4,
                     min_stiffness_threshold=1.0e-3, u_guess=(), config=None):
    """Mirror of GPUSolver.solve_system_gpu (GPUSolver.jl:467-477).

    ``nodes`` (nN x 3), ``elements`` (nE x 8, 1-based), ``bc`` (nN x 3, >0 =
    held), ``f`` (3nN), ``density`` (nE) as in the reference, possibly a
    pruned subset of the box given by ``config["geometry"]``.  Precision
    follows the reference's profile switch (GPUSolver.jl:187-189) extended
    with ``solver_parameters.precision``; the preconditioner follows
    ``solver_parameters.preconditioner`` (GPUSolver.jl:231).

    The element matrix is the exact brick Ke unless
    ``solver_parameters.ke_mode`` is ``"f32_widened"`` (the reference's
    Float32 Gauss-quadrature Ke widened to FP64, GPUSolver.jl:248-249).  The
    Float32 rounding makes that matrix indefinite on the rigid-body modes
    (eigenvalues down to -1.6e-8 against 0.064 for the smallest elastic one);
    on a 512x256x256 cantilever it moves the compliance by 2%, and the
    reference's own rounding pattern (its BLAS summation order) cannot be
    reproduced, so no implementation matches its Float32 results closer than
    that at such sizes.  The exact Ke also runs on the factored kernel."""
    config = config or {}
    sp = config.get("solver_parameters", {})
    nx, ny, nz, h = _geometry(config)
    nid, eid = _lattice(nodes, elements, nx, ny, nz, h)
    nN = (nx + 1) * (ny + 1) * (nz + 1)
    nE = nx * ny * nz
    fixed = np.ones((nN, 3), dtype=np.uint8)           # nodes outside the input: excluded
    fixed[nid] = (np.asarray(bc) > 0).astype(np.uint8)
    ff = np.zeros((nN, 3))
    ff[nid] = np.asarray(f, dtype=np.float64).reshape(-1, 3)
    rho = np.zeros(nE)
    rho[eid] = np.asarray(density, dtype=np.float64)
    active = np.zeros(nE, dtype=np.uint8)
    active[eid] = 1
    want_mg = str(sp.get("preconditioner", "jacobi")) == "multigrid"
    # "multigrid" is the reference's two-level preconditioner unless
    # solver_parameters.mg_mode asks for this library's V-cycle
    mg_precond = PRECOND_MG if str(sp.get("mg_mode", "reference")) == "vcycle" else PRECOND_REF_MG
    prec = PREC_MIXED if str(sp.get("precision", "fp64")) == "mixed" else PREC_FP64
    opts = dict(E0=float(E), Emin=0.0, penal=1.0, nu=float(nu), precision=prec,
                ke_mode=KE_F32_WIDENED if str(sp.get("ke_mode", "fp64")) == "f32_widened" else KE_FP64,
                diag_mode=DIAG_X_ENTRY if str(sp.get("diag_mode", "x_entry")) == "x_entry" else DIAG_EXACT,
                mg_levels=int(sp.get("mg_levels", 0)), mg_pre=int(sp.get("mg_smoother_sweeps", 2)),
                mg_post=int(sp.get("mg_smoother_sweeps", 2)), mg_omega=float(sp.get("mg_omega", 0.6)),
                mg_coarse_dofs=int(sp.get("mg_coarse_dofs", 0)),
                mg_cycle=CYCLE_W if str(sp.get("mg_cycle", "V")).upper() == "W" else CYCLE_V)
    with Solver(nx, ny, nz, h, **opts) as s:
        s.set_problem(fixed.reshape(-1), ff.reshape(-1), active if len(eid) < nE else None)
        s.set_density(rho)
        warm = len(u_guess) == 3 * len(nid)
        u0 = None
        if warm:
            u0 = np.zeros((nN, 3))
            u0[nid] = np.asarray(u_guess, dtype=np.float64).reshape(-1, 3)
        u, reps, st = s.solve(u0, tol=float(np.float32(tol)), max_iter=int(max_iter),
                              shift_factor=float(shift_factor),
                              max_shift_attempts=int(sp.get("max_shift_attempts", 3)),
                              shift_multiplier=float(sp.get("shift_multiplier", 10.0)),
                              stagnation_tol=float(sp.get("stagnation_tolerance", 0.0)),
                              precond=mg_precond if want_mg else PRECOND_JACOBI, warm_start=int(warm))
    rep = reps[0]
    U = u[0].reshape(nN, 3)[nid].reshape(-1)
    if st == E_NAN_INPUT:
        return np.zeros(3 * len(nid)), float("nan"), "Failed_Corrupt_Input"
    return U, rep["rel_res"], rep["method_name"]


def solve_system(nodes, elements, E, nu, bc, F, *, density=(), config, min_stiffness_threshold=1.0e-3,
                 prune_voids=True, u_prev=()):
    """Mirror of Solver.solve_system for ``solver_type: gpu``
    (Solver.jl:39-160): Float32 tolerance cast (:57), pruning skipped before
    outer iteration 3 (:67-72), pruning of elements with rho <= 1.01 *
    threshold when fewer than 99% are active (:85-116), solve, un-prune
    (:154-156)."""
    sp = config["solver_parameters"]
    tol = float(np.float32(sp.get("tolerance", 1.0e-6)))
    max_iter = int(sp.get("max_iterations", 1000))
    shift = float(np.float32(sp.get("diagonal_shift_factor", 1.0e-6)))
    it = config.get("current_outer_iter", 1)
    if isinstance(it, int) and it < 3:
        prune_voids = False
    density = np.asarray(density, dtype=np.float32)
    nodes = np.asarray(nodes)
    elements = np.asarray(elements)
    keep_e = np.arange(len(elements))
    if prune_voids and density.size:
        thr = np.float32(min_stiffness_threshold) * np.float32(1.01)
        act = density > thr
        if act.sum() < len(elements) * 0.99:
            if not act.any():
                raise ValueError("MeshPruner: No active elements found")
            keep_e = np.flatnonzero(act)
    used = np.zeros(len(nodes), dtype=bool)
    used[elements[keep_e].ravel() - 1] = True
    new_to_old = np.flatnonzero(used)
    old_to_new = np.zeros(len(nodes), dtype=np.int64)
    old_to_new[new_to_old] = np.arange(1, len(new_to_old) + 1)
    sub_el = old_to_new[elements[keep_e] - 1]
    F = np.asarray(F, dtype=np.float64).reshape(-1, 3)
    guess = ()
    if len(u_prev) == F.size:
        guess = np.asarray(u_prev, dtype=np.float64).reshape(-1, 3)[new_to_old].reshape(-1)
    U, res, meth = solve_system_gpu(nodes[new_to_old], sub_el, E, nu, np.asarray(bc)[new_to_old],
                                    F[new_to_old].reshape(-1), density[keep_e], max_iter=max_iter, tol=tol,
                                    shift_factor=shift, min_stiffness_threshold=min_stiffness_threshold,
                                    u_guess=guess, config=config)
    full = np.zeros((len(nodes), 3))
    full[new_to_old] = np.asarray(U).reshape(-1, 3)
    return full.reshape(-1), res, meth


def _grid_from_nodes(nodes, elements):
    """Structured box and element size from generate_mesh-style arrays
    (element size from element 1's edges, as CPUSolver.jl:41-42)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    el = np.asarray(elements)
    n1, n2, n4, n5 = (nodes[el[0, i] - 1] for i in (0, 1, 3, 4))
    h = (float(np.linalg.norm(n2 - n1)), float(np.linalg.norm(n4 - n1)), float(np.linalg.norm(n5 - n1)))
    ext = nodes.max(axis=0) - nodes.min(axis=0)
    return tuple(int(round(ext[d] / h[d])) for d in range(3)), h


_FILTER_CACHE = {}


def apply_gpu_filter(density, nElem_x, nElem_y, nElem_z, dx, dy, dz, radius, elements=None, tol=1.0e-5,
                     helm_element=HELM_BUG_COMPAT):
    """Mirror of GPUHelmholtz.apply_gpu_filter! (GPUHelmholtz.jl:422-443):
    returns (filtered Float32 vector, time_s, iterations, rel_res); an empty
    vector signals failure (the caller's CPU fallback, TopOpt.jl:319-328)."""
    try:
        key = (int(nElem_x), int(nElem_y), int(nElem_z), float(dx), float(dy), float(dz), helm_element)
        s = _FILTER_CACHE.get(key)
        if s is None:
            s = Solver(key[0], key[1], key[2], key[3:6], helm_element=helm_element)
            _FILTER_CACHE.clear()
            _FILTER_CACHE[key] = s
        out, it, rel = s.filter(np.asarray(density, dtype=np.float64), radius, tol=float(np.float32(tol)),
                                max_iter=200, check_every=10)
        return out.astype(np.float32), 0.0, it, rel
    except (HexaError, ImportError, OSError) as e:  # reference: any exception -> Float32[]
        print(f"  [GPU Filter] CRITICAL ERROR: {e}. Falling back to CPU Filter...")
        return np.zeros(0, np.float32), 0.0, 0, 0.0


def compute_stress_field(nodes, elements, U, E, nu, density, return_voigt=True):
    """Mirror of Stress.compute_stress_field (Stress.jl:287-374) on the
    structured box: returns (principal 3 x nE, von Mises, voigt 6 x nE (or
    0 x 0), l1, max direction 3 x nE, min direction 3 x nE), Float32."""
    (nx, ny, nz), h = _grid_from_nodes(nodes, elements)
    nid, eid = _lattice(nodes, elements, nx, ny, nz, h)
    nN = (nx + 1) * (ny + 1) * (nz + 1)
    u = np.zeros((nN, 3))
    u[nid] = np.asarray(U, dtype=np.float64).reshape(-1, 3)
    rho = np.zeros(nx * ny * nz)
    rho[eid] = np.asarray(density, dtype=np.float64)
    with Solver(nx, ny, nz, h, nu=float(nu)) as s:
        o = s.stress(u.reshape(-1), rho, E=float(E))
    pick = lambda a: np.ascontiguousarray(a[eid].T)
    voigt = pick(o["voigt"]) if return_voigt else np.zeros((0, 0), np.float32)
    return pick(o["principal"]), o["vonmises"][eid], voigt, o["l1"][eid], pick(o["dir_max"]), pick(o["dir_min"])


def update_density(density, l1_stress_norm_field, protected_elements_mask, E, l1_stress_allowable, iter,
                   number_of_iterations, original_density, min_density, max_density, config, elements=None,
                   is_annealing=False, helm_element=HELM_BUG_COMPAT):
    """Mirror of TopologyOptimization.update_density! (TopOpt.jl:211-482):
    mutates ``density`` (Float32) and returns (mean_change, R_length,
    threshold, filter_time, filter_iters, filter_res)."""
    g = config["geometry"]
    op = config["optimization_parameters"]
    sp = config.get("solver_parameters", {})
    nx, ny, nz = int(g["nElem_x_computed"]), int(g["nElem_y_computed"]), int(g["nElem_z_computed"])
    h = (float(g["dx_computed"]), float(g["dy_computed"]), float(g["dz_computed"]))
    prm = dict(E=float(E), l1_allow=float(l1_stress_allowable), min_density=float(min_density),
               max_density=float(max_density), max_domain_dim=float(g["max_domain_dim"]),
               filter_tol=float(np.float32(sp.get("filter_tolerance", 1.0e-5))),
               R_init_perc=float(op.get("filter_R_init_perc", 0.0)),
               R_interm_perc=float(op.get("filter_R_interm_perc", 0.0)),
               R_final_perc=float(op.get("filter_R_final_perc", 0.0)),
               R_interm_iter_perc=float(op.get("filter_R_interm_iter_perc", 50.0)),
               min_radius_factor=float(op.get("minimum_filter_radius_as_elements", 2.5)),
               final_threshold=float(op.get("final_density_threshold", 0.95)),
               max_culling_ratio=float(op.get("max_culling_ratio", 0.1)), iter=int(iter),
               n_iters=int(number_of_iterations), hard=int(op.get("density_update_method", "soft") == "hard"))
    with Solver(nx, ny, nz, h, helm_element=helm_element) as s:
        d, rep = s.update_density(density, l1_stress_norm_field, np.asarray(protected_elements_mask, dtype=np.uint8),
                                  original_density, **prm)
    if rep["skipped"]:
        return 0.0, 0.0, 0.0, 0.0, 0, 0.0
    density[:] = d
    return rep["mean_change"], rep["R_length"], rep["threshold"], 0.0, rep["filter_iters"], rep["filter_res"]
