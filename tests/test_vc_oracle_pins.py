"""Pins of the oracle's variable-coefficient operator (VCAMRPoissonOp2:
L phi = alpha a phi - beta div(b grad phi)) against what fixes its answers
independently of its code:

* reduction: a = b = 1 with beta -> -beta is the constant-coefficient operator
  (AMRPoissonOp: alpha phi + beta lap phi), whose oracle is pinned separately;
  operator, residual, relaxation, restriction and the multigrid solve agree to
  rounding;
* closed form: with b linear and phi quadratic the flux-difference operator is
  exact (the centred difference of a quadratic is exact at the half point, and
  the difference of the quadratic flux b phi_x is exact at the centre);
* independent assembly: the operator equals a sparse matrix assembled face by
  face with the homogeneous Dirichlet ghost -phi (BCFunc.cpp linear
  extrapolation), and the multigrid solve equals a sparse direct solve;
* coefficient coarsening: arithmetic averages of linear a and b are their centre
  values, so the lambda of every multigrid depth has a closed form;
* algebra: GSRB fixed point, restriction = mean of fine residuals;
* AMR (two levels, refluxing with -b * gradient fluxes): with b_d depending on
  the normal coordinate only and quadratic phi every fine face flux is exact and
  so is its average, so the composite operator is exact on uncovered interior
  cells; the refluxed composite operator is conservative (its integral is the
  b-weighted boundary flux); the AMR multigrid solve equals a dense solve of the
  column-probed composite system.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import hierarchy as H
from helpers import small_amr, small_single, uncovered_mask
from oracle import Oracle


def _faces_global(h, B, d):
    """Per box face-packed layout of a global face array B[d] of shape n + e_d."""
    lo0 = h.domain[0]
    parts = []
    for (lo, hi) in h.levels[0]:
        sl = tuple(slice(lo[q] - lo0[q], hi[q] + 1 - lo0[q] + (1 if q == d else 0)) for q in range(3))
        parts.append(B[sl].ravel(order="F"))
    return np.concatenate(parts)


def _interior(h, width=1):
    dom = h.level_domain(0)
    parts = []
    for (l, hh) in h.levels[0]:
        I = np.meshgrid(*[np.arange(l[d], hh[d] + 1) for d in range(3)], indexing="ij")
        m = np.ones(I[0].shape, bool)
        for d in range(3):
            m &= (I[d] >= dom[0][d] + width) & (I[d] <= dom[1][d] - width)
        parts.append(m.ravel(order="F"))
    return np.concatenate(parts)


def _unit(h, o):
    o.vc_set(np.ones(h.cells(0)), [np.ones(o.face_cells(0, d)) for d in range(3)])


def test_face_packing_counts_match_oracle():
    h = small_single(n=16, box=8)
    o = Oracle(h)
    a, b = H.vc_coefficients(h)
    assert a.size == h.cells(0)
    for d in range(3):
        assert b[d].size == o.face_cells(0, d) == 8 * 9 * 8 * 8


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (0.6, 1.7)])
def test_unit_coefficients_reduce_to_constant_operator(alpha, beta):
    hc = small_single(n=16, box=8, alpha=alpha, beta=beta)
    hv = small_single(n=16, box=8, alpha=alpha, beta=-beta)
    oc, ov = Oracle(hc), Oracle(hv)
    _unit(hv, ov)
    rng = np.random.default_rng(11)
    phi, rhs = rng.standard_normal(hc.cells(0)), rng.standard_normal(hc.cells(0))
    scale = 6 * beta * 16 ** 2 * np.abs(phi).max()
    for homog in (True, False):
        assert np.abs(ov.level_apply(0, phi, homog) - oc.level_apply(0, phi, homog)).max() < 1e-13 * scale
        assert np.abs(ov.level_residual(0, phi, rhs, homog) - oc.level_residual(0, phi, rhs, homog)).max() < 1e-13 * scale
    assert np.abs(ov.level_relax(0, phi, rhs, 2) - oc.level_relax(0, phi, rhs, 2)).max() < 1e-12 * np.abs(phi).max()
    rv, rc = ov.level_restrict(0, phi, rhs), oc.level_restrict(0, phi, rhs)
    assert np.abs(rv - rc).max() < 1e-13 * scale
    # the relaxation coefficient of each depth: 1 / (alpha - 6 beta / dx_d^2)
    for d in range(ov.mg_depths):
        dx = (1.0 / 16) * 2 ** d
        lam = ov.vc_lambda(d)
        assert np.allclose(lam, 1.0 / (alpha - 6 * beta / dx ** 2), rtol=1e-14, atol=0)
    rhs = H.sinusoid_rhs(hc, 5)
    sp_, rp = H.solver_params(eps=1e-11)
    pc, hcst, _ = oc.solve(rhs, sp_, rp)
    pv, hvc, _ = ov.solve(rhs, sp_, rp)
    assert len(hcst) == len(hvc)
    assert np.abs(pv - pc).max() < 1e-9 * np.abs(pc).max()


def test_operator_exact_for_linear_b_and_quadratic_phi():
    alpha, beta = 0.7, 1.3
    h = small_single(n=32, box=16, alpha=alpha, beta=beta)
    o = Oracle(h)
    c = [(1.0, 0.5, -0.3, 0.2), (1.5, -0.2, 0.4, 0.1), (0.8, 0.3, 0.1, -0.25)]
    a = h.pack(lambda x, y, z, l: 1.0 + 0.5 * x * y - 0.3 * z)
    b = [h.pack_faces(lambda x, y, z, cc=cc: cc[0] + cc[1] * x + cc[2] * y + cc[3] * z, d) for d, cc in enumerate(c)]
    o.vc_set(a, b)
    q = lambda x, y, z, l: 1 + 0.3 * x - 0.2 * y + x * x + 2 * y * y + 0.5 * z * z + 0.7 * x * y - 0.4 * y * z + 0.6 * x * z
    phi = h.pack(q)
    X = h.pack(lambda x, y, z, l: x)
    Y = h.pack(lambda x, y, z, l: y)
    Z = h.pack(lambda x, y, z, l: z)
    grad = [0.3 + 2 * X + 0.7 * Y + 0.6 * Z, -0.2 + 4 * Y + 0.7 * X - 0.4 * Z, Z - 0.4 * Y + 0.6 * X]
    hess = [2.0, 4.0, 1.0]
    div = 0
    for d, cc in enumerate(c):
        bd = cc[0] + cc[1] * X + cc[2] * Y + cc[3] * Z
        div = div + cc[1 + d] * grad[d] + bd * hess[d]
    expect = alpha * a * phi - beta * div
    L = o.level_apply(0, phi, homog=False)
    m = _interior(h)
    assert np.abs(L - expect)[m].max() < 1e-9 * np.abs(expect).max()


def _assemble_vc(n, a, B, alpha, beta):
    """alpha a - beta div(b grad) on n^3 cells, faces B[d] of shape n + e_d, with the
    homogeneous Dirichlet ghost -phi (so a boundary face adds 2 beta b / dx^2)."""
    dx = 1.0 / n
    s = beta / dx ** 2
    idx = lambda i, j, k: i + n * (j + n * k)
    rows, cols, vals = [], [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r = idx(i, j, k)
                diag = alpha * a[i, j, k]
                for d in range(3):
                    e = [0, 0, 0]
                    e[d] = 1
                    for sgn in (1, -1):
                        f = (i + e[0], j + e[1], k + e[2]) if sgn > 0 else (i, j, k)
                        bf = B[d][f]
                        nb = (i + sgn * e[0], j + sgn * e[1], k + sgn * e[2])
                        if all(0 <= v < n for v in nb):
                            rows.append(r); cols.append(idx(*nb)); vals.append(-s * bf)
                            diag += s * bf
                        else:
                            diag += 2 * s * bf
                rows.append(r); cols.append(r); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n ** 3, n ** 3))


def _random_vc(n, rng):
    a = 0.5 + rng.uniform(0, 1, (n, n, n))
    B = [0.5 + 3 * rng.uniform(0, 1, tuple(n + (1 if q == d else 0) for q in range(3))) for d in range(3)]
    return a, B


def test_operator_matches_independent_assembly():
    n = 16
    alpha, beta = 0.4, 1.2
    h = small_single(n=n, box=8, alpha=alpha, beta=beta)
    o = Oracle(h)
    rng = np.random.default_rng(21)
    a, B = _random_vc(n, rng)
    o.vc_set(h.from_global([a]), [_faces_global(h, B[d], d) for d in range(3)])
    A = _assemble_vc(n, a, B, alpha, beta)
    g = rng.standard_normal((n, n, n))
    L = h.to_global(o.level_apply(0, h.from_global([g]), homog=True), 0)
    ref = (A @ g.ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(L - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (2.0, 0.5)])
def test_multigrid_solve_matches_sparse_direct_solve(alpha, beta):
    n = 16
    h = small_single(n=n, box=8, alpha=alpha, beta=beta)
    o = Oracle(h)
    a, b = H.vc_coefficients(h, seed=3)
    o.vc_set(a, b)
    # the same coefficients on global arrays, for the assembly
    ag = h.to_global(a, 0)
    # the face fields evaluated on one box covering the domain
    hd = H.Hierarchy(h.domain, [[h.domain]], 1.0 / n, alpha, beta)
    _, bd = H.vc_coefficients(hd, seed=3)
    B = [bd[d].reshape(tuple(n + (1 if q == d else 0) for q in range(3)), order="F") for d in range(3)]
    for d in range(3):
        assert np.array_equal(_faces_global(h, B[d], d), b[d])
    A = _assemble_vc(n, ag, B, alpha, beta)
    rhs = H.sinusoid_rhs(h, 1001)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(A.tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-9


def test_coarsened_lambda_of_linear_coefficients():
    alpha, beta = 0.9, 1.1
    n = 32
    h = small_single(n=n, box=16, alpha=alpha, beta=beta)
    o = Oracle(h)
    fa = lambda x, y, z: 1.0 + 0.5 * x - 0.25 * y + 0.125 * z
    fb = [lambda x, y, z: 2.0 + x + 0.5 * y - 0.5 * z, lambda x, y, z: 1.5 - 0.5 * x + y + 0.25 * z,
          lambda x, y, z: 1.0 + 0.25 * x + 0.25 * y + z]
    o.vc_set(h.pack(lambda x, y, z, l: fa(x, y, z)), [h.pack_faces(fb[d], d) for d in range(3)])
    assert o.mg_depths >= 3
    for dep in range(o.mg_depths):
        r = 2 ** dep
        hc = H.Hierarchy(H.box_coarsen(h.domain, r), [[H.box_coarsen(b, r) for b in h.levels[0]]], h.dx(0) * r)
        dx = h.dx(0) * r
        X = hc.pack(lambda x, y, z, l: x)
        Y = hc.pack(lambda x, y, z, l: y)
        Z = hc.pack(lambda x, y, z, l: z)
        s = beta / dx ** 2
        den = alpha * fa(X, Y, Z)
        for d, (fx, e) in enumerate(zip(fb, [(1, 0, 0), (0, 1, 0), (0, 0, 1)])):
            den = den + s * (fx(X + 0.5 * dx * e[0], Y + 0.5 * dx * e[1], Z + 0.5 * dx * e[2])
                             + fx(X - 0.5 * dx * e[0], Y - 0.5 * dx * e[1], Z - 0.5 * dx * e[2]))
        lam = o.vc_lambda(dep)
        assert np.abs(lam * den - 1).max() < 1e-13


def test_gsrb_fixed_point_and_restriction_mean():
    h = small_single(n=16, box=8, lo=-8, alpha=0.3, beta=1.0)
    o = Oracle(h)
    rng = np.random.default_rng(31)
    a, b = H.vc_coefficients(h, seed=9)
    o.vc_set(a, b)
    phi = rng.standard_normal(h.cells(0))
    f = o.level_apply(0, phi, homog=True)
    out = o.level_relax(0, phi, f, 3)
    assert np.abs(out - phi).max() < 1e-11 * np.abs(phi).max()
    rhs = rng.standard_normal(h.cells(0))
    r = h.to_global(o.level_residual(0, phi, rhs, homog=True), 0)
    rc = o.level_restrict(0, phi, rhs)
    hc = H.Hierarchy(H.box((-4,) * 3, (3,) * 3), [[H.box_coarsen(bx, 2) for bx in h.levels[0]]], 1 / 8)
    mean = r.reshape(2, 8, 2, 8, 2, 8, order="F").mean(axis=(0, 2, 4))
    assert np.abs(hc.to_global(rc, 0) - mean).max() < 1e-12 * np.abs(mean).max()


def _amr_interior(h, lev, width=1):
    dom = h.level_domain(lev)
    parts = []
    for (l, hh) in h.levels[lev]:
        I = np.meshgrid(*[np.arange(l[d], hh[d] + 1) for d in range(3)], indexing="ij")
        m = np.ones(I[0].shape, bool)
        for d in range(3):
            m &= (I[d] >= dom[0][d] + width) & (I[d] <= dom[1][d] - width)
        parts.append(m.ravel(order="F"))
    return np.concatenate(parts)


def _amr_faces(h, fns):
    return [np.concatenate([h.pack_faces(fns[d], d, lev) for lev in range(h.nlev)]) for d in range(3)]


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (0.4, 1.3)])
def test_amr_composite_operator_exact_for_normal_linear_b(alpha, beta):
    h = small_amr(alpha=alpha, beta=beta)
    o = Oracle(h)
    c = [(1.0, 0.5), (1.5, -0.4), (0.8, 0.3)]
    fns = [lambda x, y, z: c[0][0] + c[0][1] * x, lambda x, y, z: c[1][0] + c[1][1] * y,
           lambda x, y, z: c[2][0] + c[2][1] * z]
    a = h.pack(lambda x, y, z, l: 1.0 + 0.25 * x - 0.5 * y * z)
    o.vc_set(a, _amr_faces(h, fns))
    q = lambda x, y, z, l: 1 + 0.3 * x + x * x + 2 * y * y + 0.5 * z * z + 0.7 * x * y - 0.4 * y * z + 0.25 * x * z
    phi = h.pack(q)
    X, Y, Z = (h.pack(lambda x, y, z, l, t=t: [x, y, z][t]) for t in range(3))
    grad = [0.3 + 2 * X + 0.7 * Y + 0.25 * Z, 4 * Y + 0.7 * X - 0.4 * Z, Z - 0.4 * Y + 0.25 * X]
    hess = [2.0, 4.0, 1.0]
    coord = [X, Y, Z]
    div = sum(c[d][1] * grad[d] + (c[d][0] + c[d][1] * coord[d]) * hess[d] for d in range(3))
    expect = alpha * a * phi - beta * div
    L = o.amr_operator(phi, homog=False)
    offs = h.level_offsets()
    checked = 0
    for lev in range(h.nlev):
        m = uncovered_mask(h, lev) & _amr_interior(h, lev)
        sl = slice(offs[lev], offs[lev + 1])
        assert np.abs(L[sl] - expect[sl])[m].max() < 1e-7 * np.abs(expect).max(), lev
        checked += m.sum()
    assert checked > 0
    # the coarse cells next to the fine level see the fine level's b through the
    # refluxed fluxes: other fine coefficients change them
    o2 = Oracle(h)
    o2.vc_set(a, [np.concatenate([h.pack_faces(fns[d], d, 0)] + [np.ones(h.pack_faces(fns[d], d, l).size)
                                                                 for l in range(1, h.nlev)]) for d in range(3)])
    L2 = o2.amr_operator(phi, homog=False)
    m0 = uncovered_mask(h, 0) & _amr_interior(h, 0)
    assert np.abs(L2 - L)[offs[0]:offs[1]][m0].max() > 1e-3 * np.abs(expect).max()


def test_amr_composite_operator_is_conservative():
    h = small_amr(alpha=0.0, beta=1.0)
    o = Oracle(h)
    fb = lambda x, y, z: 1.0 + 0.5 * np.sin(3 * x) * np.cos(2 * y) + 0.25 * z
    a = h.pack(lambda x, y, z, l: 1.0 + 0.0 * x)
    o.vc_set(a, _amr_faces(h, [fb, fb, fb]))
    rng = np.random.default_rng(9)
    phi = rng.standard_normal(h.total_cells())
    L = o.amr_operator(phi, homog=True)
    offs = h.level_offsets()
    total = 0.0
    for lev in range(h.nlev):
        m = uncovered_mask(h, lev)
        total += L[offs[lev]:offs[lev + 1]][m].sum() * h.dx(lev) ** 3
    # level 0 alone touches the boundary: face term -beta * b * (ghost - near) * dx with
    # ghost = -near, i.e. +2 beta b near dx per boundary face
    g0 = h.to_global(phi[offs[0]:offs[1]], 0)
    n = g0.shape[0]
    dx = h.dx(0)
    cen = (np.arange(n) + 0.5) * dx
    flux = 0.0
    for d in range(3):
        for side, pos in ((0, 0.0), (1, 1.0)):
            near = np.take(g0, 0 if side == 0 else -1, axis=d)
            t = [cen, cen, cen]
            t[d] = np.array([pos])
            G = np.meshgrid(*t, indexing="ij")
            bf = np.squeeze(fb(*G), axis=d)
            flux += (2.0 * bf * near).sum() * dx
    assert abs(total - flux) < 1e-9 * np.abs(L).max()


def test_amr_solve_matches_dense_solve_of_composite_system():
    coarse = H.domain_split(H.box((0, 0, 0), (7, 7, 7)), 4, 4)
    fine = H.domain_split(H.box((4, 4, 4), (11, 11, 11)), 4, 4)
    h = H.Hierarchy(H.box((0, 0, 0), (7, 7, 7)), [coarse, fine], 1.0 / 8, 0.5, 1.0)
    o = Oracle(h)
    a, b = H.vc_coefficients(h, seed=4)
    o.vc_set(a, b)
    N = h.total_cells()
    A = np.zeros((N, N))
    e = np.zeros(N)
    for col in range(N):
        e[:] = 0
        e[col] = 1
        A[:, col] = o.amr_operator(e, homog=True)
    keep = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    rhs = H.bump_rhs(h, 1002)
    rhs_k = np.where(keep, rhs, 0.0)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs_k, sp_, rp)
    assert info["exit_status"] & 1
    x = np.linalg.solve(A[np.ix_(keep, keep)], rhs_k[keep])
    assert np.abs(phi[keep] - x).max() / np.abs(x).max() < 1e-9


def test_vc_jacobi_sweep_unit_coefficients():
    """VCAMRPoissonOp2::levelJacobi with a = b = 1, beta -> -beta: one sweep on a
    Dirichlet sine mode with rhs 0 multiplies it by 1 - 0.5 (alpha + beta lam) /
    (alpha - 6 beta / dx^2) (lambda = 1 / (alpha - 6 beta / dx^2))."""
    n, k, alpha, beta = 16, (2, 1, 1), 0.5, 1.1
    h = small_single(n=n, box=8, alpha=alpha, beta=-beta)
    o = Oracle(h)
    _unit(h, o)
    o.set_relax_mode(4)
    dx = 1.0 / n
    e = h.pack(lambda x, y, z, l: np.sin(np.pi * k[0] * x) * np.sin(np.pi * k[1] * y) * np.sin(np.pi * k[2] * z))
    lam = sum((2 / dx ** 2) * (np.cos(np.pi * kk * dx) - 1) for kk in k)
    out = o.level_relax(0, e, np.zeros_like(e), 1)
    factor = 1 - 0.5 * (alpha + beta * lam) / (alpha - 6 * beta / dx ** 2)
    assert np.abs(out - factor * e).max() < 1e-12 * np.abs(e).max()


def test_harmonic_coarsening_of_reciprocal_linear_coefficients():
    """Harmonic averaging (VCAMRPoissonOp2Factory with CoarseAverage::harmonic): with a
    and b the reciprocals of linear functions, the mean of 1/a over the children (of
    1/b over the fine faces) is the centre value of the linear function, so every
    depth's coefficients and lambda have a closed form; the arithmetic average of the
    same data does not."""
    alpha, beta = 0.9, 1.1
    n = 32
    h = small_single(n=n, box=16, alpha=alpha, beta=beta)
    la = lambda x, y, z: 1.0 + 0.5 * x - 0.25 * y + 0.125 * z
    lb = [lambda x, y, z: 2.0 + x + 0.5 * y - 0.5 * z, lambda x, y, z: 1.5 - 0.5 * x + y + 0.25 * z,
          lambda x, y, z: 1.0 + 0.25 * x + 0.25 * y + z]
    a = h.pack(lambda x, y, z, l: 1.0 / la(x, y, z))
    b = [h.pack_faces(lambda x, y, z, f=lb[d]: 1.0 / f(x, y, z), d) for d in range(3)]
    errs = {}
    for harmonic in (True, False):
        o = Oracle(h)
        o.set_vc_average(harmonic)
        o.vc_set(a, b)
        assert o.mg_depths >= 3
        worst = 0.0
        for dep in range(1, o.mg_depths):
            r = 2 ** dep
            hc = H.Hierarchy(H.box_coarsen(h.domain, r), [[H.box_coarsen(bx, r) for bx in h.levels[0]]], h.dx(0) * r)
            dx = h.dx(0) * r
            X, Y, Z = (hc.pack(lambda x, y, z, l, q=q: (x, y, z)[q]) for q in range(3))
            s = beta / dx ** 2
            den = alpha / la(X, Y, Z)
            for d, e in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
                den = den + s * (1.0 / lb[d](X + 0.5 * dx * e[0], Y + 0.5 * dx * e[1], Z + 0.5 * dx * e[2])
                                 + 1.0 / lb[d](X - 0.5 * dx * e[0], Y - 0.5 * dx * e[1], Z - 0.5 * dx * e[2]))
            worst = max(worst, np.abs(o.vc_lambda(dep) * den - 1).max())
        errs[harmonic] = worst
    assert errs[True] < 1e-13
    assert errs[False] > 1e-6
