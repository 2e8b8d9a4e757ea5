"""The large-grid fallback kernels, forced on small grids (HEXA_FALLBACK_KERNELS=1):
the factored tile kernel (k_elastic_fact, used when a node plane exceeds the
row-wave kernel's 2 GB buffer descriptors) and the node-centric dense kernel
(k_elastic_dense, used at 2^31 nodes or more in place of the matrix cores).
Both against the oracle, and the V-cycle built on them against the golden
direct solves."""
import os
import sys
from contextlib import contextmanager

import numpy as np
import pytest

import oracle
from helpers import DevBuf, cantilever_arrays, hip_runtime
from v119_2025_12_19_hexatop_aws_amd import hexa

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz"))


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@contextmanager
def fallback():
    os.environ["HEXA_FALLBACK_KERNELS"] = "1"
    try:
        yield
    finally:
        os.environ.pop("HEXA_FALLBACK_KERNELS", None)


@pytest.mark.parametrize("ke_mode", [hexa.KE_FP64, hexa.KE_F32_WIDENED])
@pytest.mark.parametrize("grid,h", [((4, 3, 2), (1.0, 1.0, 1.0)), ((33, 9, 5), (1.0, 2.0, 0.5)),
                                    ((20, 31, 6), (1.0, 1.0, 1.0)), ((70, 15, 17), (0.5, 1.0, 1.5))])
def test_fallback_matvec_matches_oracle(grid, h, ke_mode):
    say("fallback matvec", grid, h, ke_mode)
    nx, ny, nz = grid
    rng = np.random.default_rng(17)
    rho = rng.uniform(0.001, 1.0, nx * ny * nz)
    x = rng.standard_normal(3 * (nx + 1) * (ny + 1) * (nz + 1))
    Ke = oracle.ke_hex8(*h, nu=0.3, f32=(ke_mode == hexa.KE_F32_WIDENED))
    ref = oracle.apply_stiffness(grid, Ke, oracle.stiffness_factor(rho, 1.0, 1e-9, 3.0), x)
    hip = hip_runtime()
    dx = DevBuf(hip, x.astype(np.float32))
    dy = DevBuf(hip, np.zeros(x.size, np.float32))
    with fallback(), hexa.Solver(nx, ny, nz, h, E0=1.0, Emin=1e-9, penal=3.0, ke_mode=ke_mode) as s:
        s.set_density(rho)
        y = s.matvec(x)
        s.matvec_dev(dx.p.value, dy.p.value, hexa.PREC_MIXED, reps=1)
        y32 = dy.read()
    dx.free()
    dy.free()
    tol = 1e-13 if ke_mode == hexa.KE_FP64 else 1e-6
    assert np.abs(y - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(y32 - ref).max() <= 2e-5 * np.abs(ref).max()


@pytest.mark.parametrize("kernel", [hexa.MATVEC_FACTORED, hexa.MATVEC_DENSE])
@pytest.mark.parametrize("precision", [hexa.PREC_FP64, hexa.PREC_MIXED])
def test_fallback_vcycle_solve_matches_golden(precision, kernel):
    say("fallback solve", precision, kernel)
    nx, ny, nz, seed = (int(v) for v in GOLD["solve_b_grid"])
    fixed, f, _ = cantilever_arrays(nx, ny, nz, seed)
    rho, ug = GOLD["solve_b_rho"], GOLD["solve_b_u"]
    with fallback(), hexa.Solver(nx, ny, nz, E0=1.0, Emin=1e-9, penal=3.0, precision=precision,
                                 matvec_kernel=kernel) as s:
        s.set_problem(fixed, f)
        s.set_density(rho)
        u, reps, st = s.solve(tol=1e-12, precond=hexa.PRECOND_MG, semantics=hexa.SEMANTICS_CPU, shift_factor=1e-10,
                              max_iter=20000)
    assert st == hexa.OK, reps[0]
    assert np.linalg.norm(u[0] - ug) / np.linalg.norm(ug) < 1e-9
