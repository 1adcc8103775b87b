"""Pins of the oracle's quadratic Dirichlet boundary rule (BC type 2: DiriBC with
order 2, BCFunc.cpp quadraticInterp, ghost = (8/3) v + (1/3) phi(2 in) - 2 phi(1 in)):

* exactness: the rule extrapolates quadratics exactly, so with the face values of
  a quadratic f(x) on the x faces (and homogeneous Neumann faces where phi does not
  vary), the 7-point operator gives alpha f + beta f'' on every cell, boundary cells
  included (the linear rule, type 1, does not);
* independent assembly: the homogeneous operator equals a sparse matrix assembled
  with the boundary-row stencil (4/3 phi1 - 4 phi0) / dx^2 in x, and the multigrid
  solve of the inhomogeneous problem equals a sparse direct solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import hierarchy as H
from oracle import Oracle


def _single(n, box, alpha, beta, lo_type, hi_type, lo_val=(0.0, 0.0, 0.0), hi_val=(0.0, 0.0, 0.0)):
    dom = H.box((0, 0, 0), (n - 1, n - 1, n - 1))
    return H.Hierarchy(dom, [H.domain_split(dom, box, 8)], 1.0 / n, alpha, beta, bc_type_lo=lo_type,
                       bc_type_hi=hi_type, bc_val_lo=lo_val, bc_val_hi=hi_val)


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (0.5, 1.3)])
def test_quadratic_dirichlet_is_exact_on_quadratics(alpha, beta):
    n = 16
    f = lambda x: 0.4 + 1.5 * x - 2.0 * x * x
    h = _single(n, 8, alpha, beta, (2, 0, 0), (2, 0, 0), lo_val=(f(0.0), 0.0, 0.0), hi_val=(f(1.0), 0.0, 0.0))
    o = Oracle(h)
    phi = h.pack(lambda x, y, z, l: f(x))
    L = o.level_apply(0, phi, homog=False)
    assert np.abs(L - (alpha * phi + beta * (-4.0))).max() < 1e-9
    # the linear rule is not exact at the boundary cells
    h1 = _single(n, 8, alpha, beta, (1, 0, 0), (1, 0, 0), lo_val=(f(0.0), 0.0, 0.0), hi_val=(f(1.0), 0.0, 0.0))
    L1 = Oracle(h1).level_apply(0, phi, homog=False)
    assert np.abs(L1 - (alpha * phi + beta * (-4.0))).max() > 0.5


def _assemble(n, alpha, beta):
    """alpha I + beta Laplacian, quadratic Dirichlet (homogeneous) on the x faces,
    linear Dirichlet on y and z."""
    dx = 1.0 / n
    s = beta / dx ** 2
    idx = lambda i, j, k: i + n * (j + n * k)
    rows, cols, vals = [], [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r = idx(i, j, k)
                diag = alpha - 6 * s
                for d, (di, dj, dk) in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                                                  (0, 0, -1)]):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < n and 0 <= jj < n and 0 <= kk < n:
                        rows.append(r); cols.append(idx(ii, jj, kk)); vals.append(s)
                    elif d < 2:
                        # ghost = (1/3) phi(2 in) - 2 phi(1 in): phi(1 in) is this cell
                        ii2 = i - di
                        rows.append(r); cols.append(idx(ii2, j, k)); vals.append(s / 3.0)
                        diag -= 2 * s
                    else:
                        diag -= s                                   # ghost = -phi(1 in)
                rows.append(r); cols.append(r); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n ** 3, n ** 3))


def test_quadratic_dirichlet_operator_and_solve_match_assembly():
    n, alpha, beta = 16, 0.3, 1.0
    h = _single(n, 8, alpha, beta, (2, 1, 1), (2, 1, 1))
    o = Oracle(h)
    A = _assemble(n, alpha, beta)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((n, n, n))
    L = h.to_global(o.level_apply(0, h.from_global([g]), homog=True), 0)
    ref = (A @ g.ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(L - ref).max() < 1e-10 * np.abs(ref).max()
    rhs = H.sinusoid_rhs(h, 1301)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(A.tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-9
