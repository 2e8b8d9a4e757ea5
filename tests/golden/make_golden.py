"""Writes tests/golden/golden.npz: known answers for the hot path, computed
here by an independent numpy/scipy restatement of the mathematics.

The reference ships no tests or fixtures (it is Julia + CUDA.jl and cannot
run here), so these vectors come from closed forms and sparse direct solves,
not from executing the reference:

* ``ke_*``: the hex8 brick stiffness from its exact Kronecker form
  K_(a,i),(b,j) = lam A_ij + mu delta_ij sum_k A_kk + mu A_ji with
  A_kl = int dN_a/dx_k dN_b/dx_l as products of 1-D 2x2 integrals
  (mass h/6[2 1;1 2], stiffness 1/h[1 -1;-1 1], mixed [-1/2 -1/2; 1/2 1/2]).
  The reference integrates the same bilinear form with 2x2x2 Gauss points
  (R/src/Core/Element.jl:73-111), which is exact on a brick.
* ``helm_*``: the scalar Laplacian and mass of the Helmholtz filter element
  (Element.jl:113-134) for the reference's node table (top face at z=+dz,
  i.e. a brick of height 1.5 dz) and for the true brick.
* ``solve_*``: u = (K_ff + s I)^-1 f_f on small cantilevers by scipy's
  sparse LU, s = 1e-10 max diag(K_ff) (the FP64 profile's shift factor,
  R/src/Main.jl:108 with CPUSolver.jl:204-205 / GPUSolver.jl:343).
* ``filter_*``: the PDE filter (R^2 K + M) x = b, b_n = sum rho_e V/8,
  rho~_e = mean of 8 nodal x (GPUHelmholtz.jl:59-110) by sparse LU.

Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

CORNER = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                   [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def splitmix_uniform(seed, n):
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def one_d(h):
    M = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
    K = 1.0 / h * np.array([[1.0, -1.0], [-1.0, 1.0]])
    G = np.array([[-0.5, -0.5], [0.5, 0.5]])          # int phi'_a phi_b
    return M, K, G


def grad_pair(h, k, l):
    """A_kl[a,b] = int d_k N_a d_l N_b over the brick h=(hx,hy,hz)."""
    A = np.ones((8, 8))
    for d in range(3):
        M, K, G = one_d(h[d])
        if d == k and d == l:
            F = K
        elif d == k:
            F = G
        elif d == l:
            F = G.T
        else:
            F = M
        A *= F[np.ix_(CORNER[:, d], CORNER[:, d])]
    return A


def ke_exact(h, nu, E=1.0):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    A = [[grad_pair(h, k, l) for l in range(3)] for k in range(3)]
    lap = A[0][0] + A[1][1] + A[2][2]
    Ke = np.zeros((24, 24))
    for i in range(3):
        for j in range(3):
            blk = lam * A[i][j] + mu * A[j][i] + (mu * lap if i == j else 0.0)
            Ke[i::3, j::3] = blk
    return Ke


def helm_exact(h, bug_compat):
    hz = 1.5 * h[2] if bug_compat else h[2]
    hh = (h[0], h[1], hz)
    K8 = sum(grad_pair(hh, k, k) for k in range(3))
    M8 = np.ones((8, 8))
    for d in range(3):
        M8 *= one_d(hh[d])[0][np.ix_(CORNER[:, d], CORNER[:, d])]
    return K8, M8


def elem_nodes(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    return np.stack([(i + c[0]) + (nx + 1) * ((j + c[1]) + (ny + 1) * (k + c[2])) for c in CORNER], 1)


def assemble(nx, ny, nz, Ke, factor):
    en = elem_nodes(nx, ny, nz)
    dofs = (3 * en[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (factor[:, None] * Ke.ravel()[None, :]).ravel()
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def cantilever_solve(nx, ny, nz, seed, penal=3.0, Emin=1e-9, shift_factor=1e-10):
    nE = nx * ny * nz
    rho = 0.001 + 0.999 * splitmix_uniform(seed, nE)
    factor = Emin + rho ** penal * (1.0 - Emin)
    Ke = ke_exact((1.0, 1.0, 1.0), 0.3)
    K = assemble(nx, ny, nz, Ke, factor)
    nN = (nx + 1) * (ny + 1) * (nz + 1)
    fixed = np.zeros((nz + 1, ny + 1, nx + 1, 3), dtype=np.uint8)
    fixed[:, :, 0, :] = 1
    fixed = fixed.ravel()
    f = np.zeros(3 * nN)
    for k in range(nz + 1):
        f[3 * (nx + (nx + 1) * (0 + (ny + 1) * k)) + 1] -= 1.0 / (nz + 1)
    free = np.flatnonzero(fixed == 0)
    Kff = K[free][:, free]
    s = shift_factor * Kff.diagonal().max()
    A = (Kff + s * sp.identity(len(free))).tocsc()
    u = np.zeros(3 * nN)
    u[free] = spla.spsolve(A, f[free])
    return rho, fixed, f, u, float(f @ u), s


def filter_direct(nx, ny, nz, h, radius, rho, bug_compat, reg=0.0):
    """(R^2 K + M + reg I) x = b by sparse LU; reg = 1e-9 is the reference's
    CPU filter (TopOpt.jl:72-77), 0 its GPU filter."""
    K8, M8 = helm_exact(h, bug_compat)
    Ae = radius ** 2 * K8 + M8
    en = elem_nodes(nx, ny, nz)
    nN = (nx + 1) * (ny + 1) * (nz + 1)
    rows = np.repeat(en, 8, axis=1).ravel()
    cols = np.tile(en, (1, 8)).ravel()
    A = sp.csr_matrix((np.tile(Ae.ravel(), len(en)), (rows, cols)), shape=(nN, nN))
    A = (A + reg * sp.identity(nN)).tocsc()
    b = np.zeros(nN)
    np.add.at(b, en.ravel(), np.repeat(rho * h[0] * h[1] * h[2] / 8.0, 8))
    x = spla.spsolve(A, b)
    return x[en].mean(axis=1)


def main():
    out = {}
    for tag, h, nu in (("unit", (1.0, 1.0, 1.0), 0.3), ("aniso", (1.0, 2.0, 0.5), 0.3),
                       ("h2nu25", (2.0, 2.0, 2.0), 0.25)):
        out[f"ke_{tag}"] = ke_exact(h, nu)
        out[f"ke_{tag}_params"] = np.array(list(h) + [nu])
    for bc in (True, False):
        K8, M8 = helm_exact((1.0, 1.0, 1.0), bc)
        tag = "bug" if bc else "corr"
        out[f"helm_K8_{tag}"] = K8
        out[f"helm_M8_{tag}"] = M8
    for tag, (nx, ny, nz, seed) in (("a", (4, 3, 2, 11)), ("b", (8, 4, 4, 12))):
        rho, fixed, f, u, c, s = cantilever_solve(nx, ny, nz, seed)
        out[f"solve_{tag}_grid"] = np.array([nx, ny, nz, seed])
        out[f"solve_{tag}_rho"] = rho
        out[f"solve_{tag}_u"] = u
        out[f"solve_{tag}_compliance"] = np.array([c])
        out[f"solve_{tag}_shift"] = np.array([s])
    nx, ny, nz = 5, 4, 3
    rho = splitmix_uniform(13, nx * ny * nz)
    out["filter_grid"] = np.array([nx, ny, nz])
    out["filter_rho"] = rho
    for bc in (True, False):
        out[f"filter_{'bug' if bc else 'corr'}_R15"] = filter_direct(nx, ny, nz, (1.0, 1.0, 1.0), 1.5,
                                                                     rho, bc)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
