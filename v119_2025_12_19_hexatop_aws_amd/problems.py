"""Seeded synthetic inputs on the structured hex8 grid.

The reference builds its problems from YAML (``R/configs/default.yaml``):
a box of ``nElem_x x nElem_y x nElem_z`` unit bricks (``R/src/Mesh/Mesh.jl:23-60``),
supports selected by location (``R/src/Core/Boundary.jl:79-160``) and loads
spread evenly over the selected nodes (``Boundary.jl:232-291``).  This module
produces the same arrays for the cantilever workloads directly, seeded by
SplitMix64 so that C, Python and Julia draw identical numbers:

    z = seed + (index + 1) * 0x9E3779B97F4A7C15
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB
    u = ((z ^ (z >> 31)) >> 11) * 2^-53

Layouts (0-based): node (i,j,k) -> i + (nx+1)*(j + (ny+1)*k); element likewise
with nx, ny; DOF 3*node + component; ``fixed`` is a uint8 per DOF (1 = held).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, index) -> np.ndarray:
    idx = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, index) -> np.ndarray:
    """U[0,1) with 53 random bits."""
    return (splitmix64(seed, index) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def n_nodes(nx, ny, nz):
    return (nx + 1) * (ny + 1) * (nz + 1)


def node_index(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


@dataclass
class Problem:
    """One linear-elastic solve on the structured grid."""
    nx: int
    ny: int
    nz: int
    h: tuple = (1.0, 1.0, 1.0)
    fixed: np.ndarray = None        # uint8, 3*nN
    f: np.ndarray = None            # float64, nrhs x 3*nN
    rho: np.ndarray = None          # float64, nE (design or physical density)
    E0: float = 1.0
    Emin: float = 1e-9
    penal: float = 3.0
    nu: float = 0.3
    meta: dict = field(default_factory=dict)

    @property
    def grid(self):
        return (self.nx, self.ny, self.nz)

    @property
    def n_elem(self):
        return self.nx * self.ny * self.nz

    @property
    def n_dof(self):
        return 3 * n_nodes(self.nx, self.ny, self.nz)


def fixed_face_x0(nx, ny, nz) -> np.ndarray:
    """All DOFs of the x=0 face held (``location: [0, ":", ":"]``)."""
    fx = np.zeros((nz + 1, ny + 1, nx + 1, 3), dtype=np.uint8)
    fx[:, :, 0, :] = 1
    return fx.reshape(-1)


def fixed_face(nx, ny, nz, face: int) -> np.ndarray:
    """face 0..5 = x=0, x=max, y=0, y=max, z=0, z=max."""
    fx = np.zeros((nz + 1, ny + 1, nx + 1, 3), dtype=np.uint8)
    sl = [slice(None)] * 3
    axis = {0: 2, 1: 2, 2: 1, 3: 1, 4: 0, 5: 0}[face]
    sl[axis] = 0 if face % 2 == 0 else -1
    fx[tuple(sl) + (slice(None),)] = 1
    return fx.reshape(-1)


def load_end_bottom_edge(nx, ny, nz, total=(0.0, -1.0, 0.0)) -> np.ndarray:
    """Total force spread 1/n over the nodes of the edge {x=max, y=0}
    (``location: ["100%", 0, ":"]``, Boundary.jl:232-291)."""
    f = np.zeros(3 * n_nodes(nx, ny, nz))
    ks = np.arange(nz + 1)
    nodes = node_index(nx, 0, ks, nx, ny)
    for c in range(3):
        f[3 * nodes + c] += total[c] / len(ks)
    return f


def box_smooth(field: np.ndarray, shape, passes: int) -> np.ndarray:
    """``passes`` rounds of the 27-point box average over existing
    neighbours (separable: 3-tap mean along each axis)."""
    a = field.reshape(shape)  # (nz, ny, nx)
    for _ in range(passes):
        for ax in range(3):
            s = a.copy()
            cnt = np.ones(a.shape[ax])
            idx = [slice(None)] * 3
            lo = list(idx); lo[ax] = slice(1, None)
            hi = list(idx); hi[ax] = slice(None, -1)
            s[tuple(lo)] += a[tuple(hi)]
            s[tuple(hi)] += a[tuple(lo)]
            cnt[1:] += 1
            cnt[:-1] += 1
            bshape = [1, 1, 1]; bshape[ax] = -1
            a = s / cnt.reshape(bshape)
    return a.reshape(-1)


def cantilever_c1() -> Problem:
    """60x20x20 cantilever: x=0 face fixed, unit downward load on the bottom
    end edge, rho0 = 0.3 + U(-0.01, 0.01), seed 1."""
    nx, ny, nz = 60, 20, 20
    nE = nx * ny * nz
    rho = 0.3 + (2.0 * uniform(1, np.arange(nE)) - 1.0) * 0.01
    return Problem(nx, ny, nz, fixed=fixed_face_x0(nx, ny, nz),
                   f=load_end_bottom_edge(nx, ny, nz)[None, :], rho=rho,
                   meta=dict(name="cantilever 60x20x20", volfrac=0.3, rmin=1.5, move=0.2,
                             oc_iters=20, tol=1e-8))


def cantilever(nx, ny, nz, seed=2, smooth_passes=0, lo=0.001) -> Problem:
    """Cantilever of any size with rho = U(lo, 1) (seeded), optionally box
    smoothed ``smooth_passes`` times and affinely re-mapped to [lo, 1]."""
    nE = nx * ny * nz
    u = uniform(seed, np.arange(nE, dtype=np.uint64))
    if smooth_passes:
        u = box_smooth(u, (nz, ny, nx), smooth_passes)
        u = (u - u.min()) / max(u.max() - u.min(), 1e-300)
        rho = lo + (1.0 - lo) * u
    else:
        rho = lo + (1.0 - lo) * u
    return Problem(nx, ny, nz, fixed=fixed_face_x0(nx, ny, nz),
                   f=load_end_bottom_edge(nx, ny, nz)[None, :], rho=rho,
                   meta=dict(name=f"cantilever {nx}x{ny}x{nz}", seed=seed,
                             smooth_passes=smooth_passes))


def cantilever_c2() -> Problem:
    p = cantilever(192, 64, 64, seed=2)
    p.meta.update(mg_levels=4, oc_iters=10, rmin=1.5, volfrac=0.3, move=0.2)
    return p


def cantilever_c3() -> Problem:
    p = cantilever(512, 256, 256, seed=3, smooth_passes=3)
    # 6 levels leave a 16x8x8 coarsest grid (4131 DOFs); automatic level
    # selection goes one level further (8x4x4, 675 DOFs)
    p.meta.update(mg_levels_config=6)
    return p


def point_loads(nx, ny, nz, nrhs, seed=4) -> np.ndarray:
    """Unit point loads at nodes drawn uniformly from the top (y=max) and
    end (x=max) faces, direction uniform on the sphere."""
    top = [node_index(i, ny, k, nx, ny) for k in range(nz + 1) for i in range(nx + 1)]
    end = [node_index(nx, j, k, nx, ny) for k in range(nz + 1) for j in range(ny)]
    cand = np.array(top + end, dtype=np.int64)
    f = np.zeros((nrhs, 3 * n_nodes(nx, ny, nz)))
    for r in range(nrhs):
        u = uniform(seed, np.arange(3 * r, 3 * r + 3, dtype=np.uint64))
        n = cand[min(int(u[0] * len(cand)), len(cand) - 1)]
        z = 2.0 * u[1] - 1.0
        phi = 2.0 * np.pi * u[2]
        s = np.sqrt(max(0.0, 1.0 - z * z))
        f[r, 3 * n:3 * n + 3] = (s * np.cos(phi), s * np.sin(phi), z)
    return f


def cantilever_c4(nrhs=8) -> Problem:
    # density as the 512x256x256 case: box-smoothed U(0,1) mapped to [0.001, 1]
    p = cantilever(256, 128, 128, seed=4, smooth_passes=3)
    p.f = point_loads(256, 128, 128, nrhs, seed=4)
    p.meta.update(name="cantilever 256x128x128 x8 load cases")
    return p


def serving_stream(n_jobs, seed=5, rate=20.0, min_elems=32 ** 3, max_elems=4194304, oc_iters=5, multiple=8):
    """Seeded request stream: Poisson arrivals (``rate`` per second), element
    counts log-uniform in [min_elems, max_elems], aspect nx:ny:nz = a:1:1 with
    a in {1,2,4}, dimensions rounded to multiples of ``multiple`` (for the
    multigrid hierarchy), support on a random face, 1-4 random point loads,
    volfrac U(0.2, 0.4)."""
    jobs, t = [], 0.0
    for j in range(n_jobs):
        u = uniform(seed, np.arange(16 * j, 16 * j + 8, dtype=np.uint64))
        t += -np.log(1.0 - u[0]) / rate
        ne = float(np.exp(np.log(min_elems) + u[1] * (np.log(max_elems) - np.log(min_elems))))
        a = (1, 2, 4)[min(2, int(u[2] * 3))]
        m = (ne / a) ** (1.0 / 3.0)
        m = max(multiple, int(round(m / multiple)) * multiple)
        jobs.append(dict(arrival_s=t, nx=a * m, ny=m, nz=m, support_face=min(5, int(u[3] * 6)),
                         n_loads=1 + min(3, int(u[4] * 4)), volfrac=0.2 + 0.2 * u[5], oc_iters=oc_iters,
                         seed=int(splitmix64(seed + 1000, j))))
    return jobs
