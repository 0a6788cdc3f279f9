"""ORACLE — test infrastructure only (imported by tests/).  A dense float64 restatement of the multigrid
V-cycle of the mixed-precision induction solve (mpcns_uns_amd/csrc/kernels/multigrid.hip), built from matrices
instead of kernels, so that one application of the device cycle can be checked against the product of its stage
matrices, and the hierarchy's Galerkin operators against dense triple products.

The cycle has no counterpart in the reference (whose implicit solve is preconditioned by Jacobi or a block-Jacobi
ILU(0), Common/induction/ImplicitInduction.cpp:209-317); the coarse operators restate the resistive part of the
reference's apply_F (ImplicitInduction.cpp:127-183) on the agglomerated levels.
"""
import numpy as np

NONE = np.iinfo(np.int32).min


def d1_matrix(mesh):
    """Signed face <- edge incidence (d1) as a dense matrix."""
    ptr, idx, sgn = mesh.array("f2e.ptr"), mesh.array("f2e.idx"), mesh.array("f2e.sgn")
    D = np.zeros((mesh.n_faces, mesh.n_edges))
    rows = np.repeat(np.arange(mesh.n_faces), np.diff(ptr))
    np.add.at(D, (rows, idx), sgn)
    return D


def operator_F(mesh, apply_F):
    """Dense matrix of the edge operator x -> F(d1 x), F the induction operator of `apply_F` (a callable on face
    fluxes, e.g. the oracle's ind_apply_F), assembled column by column."""
    D1 = d1_matrix(mesh)
    return np.stack([apply_F(D1[:, j]) for j in range(mesh.n_edges)], axis=1)


def prolongation(level, n_coarse):
    """Dense P of a level (rows: finer unknowns) from its signed map."""
    m = level["map"]
    P = np.zeros((m.size, n_coarse))
    ok = m != NONE
    i = np.nonzero(ok)[0]
    v = m[ok]
    P[i, np.where(v >= 0, v, ~v)] = np.where(v >= 0, 1.0, -1.0)
    return P


def csr_dense(level):
    n = level["n"]
    K = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(level["k_ptr"]))
    K[rows, level["k_col"]] = level["k_val"]
    return K


def vcycle(A0, invdiag0, levels, dt, v, dirichlet, omega=0.7, pre=2, post=2, coarse=4):
    """z = M v: A0 the dense fine operator (rows and columns of Dirichlet edges zero but for no identity: they are
    masked), invdiag0 its Jacobi diagonal inverse; levels from Mesh.mg_hierarchy; the cycle of multigrid.hip."""
    free = ~np.asarray(dirichlet, dtype=bool)
    Ps = [prolongation(levels[l], levels[l]["n"]) for l in range(1, len(levels))]
    As = [A0] + [np.diag(levels[l]["mass"]) + dt * csr_dense(levels[l]) for l in range(1, len(levels))]
    Dinv = [np.where(free, invdiag0, 0.0)]
    for l in range(1, len(levels)):
        dg = np.diag(As[l])
        Dinv.append(np.where(dg != 0, 1.0 / np.where(dg != 0, dg, 1.0), 0.0))

    def cycle(l, b):
        A, Di = As[l], Dinv[l]
        last = l == len(levels) - 1
        x = omega * Di * b
        for _ in range((coarse if last else pre) - 1):
            x = x + omega * Di * (b - A @ x)
        if last:
            return x
        r = b - A @ x
        x = x + Ps[l] @ cycle(l + 1, Ps[l].T @ r)
        for _ in range(post):
            x = x + omega * Di * (b - A @ x)
        return x

    return cycle(0, np.where(free, v, 0.0))
