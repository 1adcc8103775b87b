"""Pins of the oracle (oracle/amr_oracle.c) against what fixes its answers
independently of its own code:

* closed forms: the 7-point operator is exact on quadratics; sin / cos products
  are discrete eigenfunctions under homogeneous Dirichlet / Neumann ghosts with
  eigenvalue sum_d (2/h^2)(cos(pi k_d h) - 1); AVERAGE of a linear field is its
  centre value; PROLONG adds the parent value;
* the reference's own known-answer test lib/test/AMRTools/quadCFInterpTest.cpp
  (quadratic data reproduced at coarse-fine ghosts to 1e-12 relative, L-shaped
  hierarchy) and the quadratic-exactness of the composite AMR operator with
  refluxing (the composite Laplacian of a full quadratic is its constant);
* algebra: residual = rhs - L(phi) (the opposite sign to the comment at
  LinearSolver.H:38, as OPERATORLAPRES computes), GSRB fixed points and colour
  parity with negative indices, discrete conservation of the refluxed composite
  operator, restriction = mean of fine residuals;
* direct solves: multigrid and AMR multigrid solutions against sparse / dense
  direct solves of independently assembled (single level) or column-probed
  (composite) systems.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import hierarchy as H
from helpers import (ghost_dir, ghost_masks, ghosted_from_fn, lshape_amr, lshape_qcfi, small_amr, small_single,
                     uncovered_mask)
from oracle import Oracle


def test_domain_split_matches_reference_recipe():
    dom = H.box((0, 0, 0), (255, 255, 255))
    assert len(H.domain_split(dom, 64, 8)) == 64
    h = H.amr_hierarchy(256)
    assert [len(l) for l in h.levels] == [64, 125, 343]
    assert [h.cells(l) for l in range(3)] == [256 ** 3, 320 ** 3, 400 ** 3]
    widths = sorted({H.box_shape(b)[0] for b in h.levels[2]})
    assert widths == [16, 64]
    # sorted like Box::operator<: smallEnd lexicographic, dim 0 most significant
    lv = h.levels[1]
    assert all(H.sort_key(lv[i]) < H.sort_key(lv[i + 1]) for i in range(len(lv) - 1))


def _interior(h, lev, width=1):
    dom = h.level_domain(lev)
    parts = []
    for b in h.levels[lev]:
        (l, hh) = b
        idx = [np.arange(l[d], hh[d] + 1) for d in range(3)]
        I = np.meshgrid(*idx, indexing="ij")
        m = np.ones(I[0].shape, bool)
        for d in range(3):
            m &= (I[d] >= dom[0][d] + width) & (I[d] <= dom[1][d] - width)
        parts.append(m.ravel(order="F"))
    return np.concatenate(parts)


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (0.6, 1.7)])
def test_operator_exact_on_quadratic(alpha, beta):
    h = small_single(n=32, box=16, alpha=alpha, beta=beta)
    o = Oracle(h)
    q = lambda x, y, z, l: 1 + 0.3 * x - 0.2 * y + x * x + 2 * y * y + 0.5 * z * z + 0.7 * x * y - 0.4 * y * z
    phi = h.pack(q)
    L = o.level_apply(0, phi, homog=False)
    m = _interior(h, 0)
    expect = alpha * phi + beta * 7.0
    assert np.abs(L - expect)[m].max() < 1e-9


@pytest.mark.parametrize("k", [(1, 1, 1), (2, 3, 1), (3, 1, 2)])
def test_dirichlet_sine_eigenfunctions(k):
    n = 32
    h = small_single(n=n, box=16, lo=0)
    o = Oracle(h)
    dx = 1.0 / n
    f = h.pack(lambda x, y, z, l: np.sin(np.pi * k[0] * x) * np.sin(np.pi * k[1] * y) * np.sin(np.pi * k[2] * z))
    lam = sum((2 / dx ** 2) * (np.cos(np.pi * kk * dx) - 1) for kk in k)
    L = o.level_apply(0, f, homog=True)
    assert np.abs(L - lam * f).max() < 1e-12 * np.abs(lam)


def test_neumann_cosine_eigenfunctions_helmholtz():
    n, k = 16, (1, 2, 0)
    h = small_single(n=n, box=8, bc=(0, 0, 0), alpha=0.5, beta=2.0)
    o = Oracle(h)
    dx = 1.0 / n
    f = h.pack(lambda x, y, z, l: np.cos(np.pi * k[0] * x) * np.cos(np.pi * k[1] * y) * np.cos(np.pi * k[2] * z))
    lam = sum((2 / dx ** 2) * (np.cos(np.pi * kk * dx) - 1) for kk in k)
    L = o.level_apply(0, f, homog=True)
    assert np.abs(L - (0.5 + 2.0 * lam) * f).max() < 1e-11 * abs(lam)


def test_residual_sign_is_rhs_minus_operator():
    h = small_single(n=16, box=8)
    o = Oracle(h)
    rng = np.random.default_rng(0)
    phi, rhs = rng.standard_normal(h.cells(0)), rng.standard_normal(h.cells(0))
    r = o.level_residual(0, phi, rhs, homog=True)
    L = o.level_apply(0, phi, homog=True)
    assert np.abs(r - (rhs - L)).max() < 1e-12 * np.abs(L).max()


def _assemble_dirichlet(n, alpha=0.0, beta=1.0):
    """Independent sparse assembly of alpha*I + beta*Laplacian on n^3 cells with
    homogeneous Dirichlet ghost = -near (linearInterp of BCFunc.cpp)."""
    dx = 1.0 / n
    N = n ** 3
    idx = lambda i, j, k: i + n * (j + n * k)
    rows, cols, vals = [], [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r = idx(i, j, k)
                diag = alpha - 6 * beta / dx ** 2
                for d, (a, b) in enumerate([(1, 0), (-1, 0), (0, 1), (0, -1), (0, 0), (0, 0)]):
                    pass
                for (di, dj, dk) in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < n and 0 <= jj < n and 0 <= kk < n:
                        rows.append(r); cols.append(idx(ii, jj, kk)); vals.append(beta / dx ** 2)
                    else:
                        diag -= beta / dx ** 2
                rows.append(r); cols.append(r); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def test_operator_matches_independent_assembly():
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    A = _assemble_dirichlet(n)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((n, n, n))
    phi = h.from_global([g])
    L = h.to_global(o.level_apply(0, phi, homog=True), 0)
    ref = (A @ g.ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(L - ref).max() < 1e-10 * np.abs(ref).max()


def test_gsrb_fixed_point_and_colour_parity():
    h = small_single(n=16, box=8, lo=-8)
    o = Oracle(h)
    rng = np.random.default_rng(2)
    phi = rng.standard_normal(h.cells(0))
    f = o.level_apply(0, phi, homog=True)
    out = o.level_relax(0, phi, f, 3)
    assert np.abs(out - phi).max() < 1e-11 * np.abs(phi).max()
    # one colour pass changes exactly the cells with (i+j+k+color) even (global indices)
    from helpers import ghosted_random
    pg = ghosted_random(h, 0, rng)
    rhs = rng.standard_normal(h.cells(0))
    valid, _, _ = ghost_masks(h, 0)
    parity = []
    for b in h.levels[0]:
        g = H.box_grow(b, 1)
        idx = [np.arange(g[0][d], g[1][d] + 1) for d in range(3)]
        I = np.meshgrid(*idx, indexing="ij")
        parity.append(((I[0] + I[1] + I[2]) % 2).ravel(order="F"))
    parity = np.concatenate(parity)
    for color in (0, 1):
        out = o.kernel_gsrb(0, pg, rhs, color)
        changed = out != pg
        assert not changed[~valid].any()
        assert (changed[valid] == (parity[valid] == color)).all()


def test_multigrid_solve_matches_sparse_direct_solve():
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    rhs = H.sinusoid_rhs(h, 1001)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    A = _assemble_dirichlet(n)
    x = spla.spsolve(A.tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n), order="F")
    err = np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max()
    assert err < 1e-10


def test_helmholtz_solve_matches_sparse_direct_solve():
    n = 16
    h = small_single(n=n, box=8, alpha=0.8, beta=1.4)
    o = Oracle(h)
    rhs = H.sinusoid_rhs(h, 7, k=(2, 1, 3))
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs, sp_, rp)
    A = _assemble_dirichlet(n, 0.8, 1.4)
    x = spla.spsolve(A.tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n), order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-10


def test_restriction_is_mean_of_fine_residuals_and_prolongation_adds_parent():
    h = small_single(n=16, box=8)
    o = Oracle(h)
    rng = np.random.default_rng(4)
    phi, rhs = rng.standard_normal(h.cells(0)), rng.standard_normal(h.cells(0))
    r = h.to_global(o.level_residual(0, phi, rhs, homog=True), 0)
    rc = o.level_restrict(0, phi, rhs)
    hc = H.Hierarchy(H.box((0, 0, 0), (7, 7, 7)), [[H.box_coarsen(b, 2) for b in h.levels[0]]], 1 / 8)
    mean = r.reshape(2, 8, 2, 8, 2, 8, order="F").mean(axis=(0, 2, 4))
    assert np.abs(hc.to_global(rc, 0) - mean).max() < 1e-12 * np.abs(mean).max()
    ec = rng.standard_normal(hc.cells(0))
    out = o.level_prolong(0, phi, ec)
    up = np.repeat(np.repeat(np.repeat(hc.to_global(ec, 0), 2, 0), 2, 1), 2, 2)
    assert np.abs(h.to_global(out, 0) - (h.to_global(phi, 0) + up)).max() < 1e-14


def test_average_of_linear_field_is_centre_value():
    h = small_amr()
    o = Oracle(h)
    lin = lambda x, y, z, l: 0.5 + 2 * x - 3 * y + 0.25 * z
    fine = h.pack(lin)[h.level_offsets()[1]:h.level_offsets()[2]]
    avg = o.average(1, fine)
    hc = H.Hierarchy(h.domain, [[H.box_coarsen(b, 2) for b in h.levels[1]]], h.dx(0))
    assert np.abs(avg - hc.pack(lin)).max() < 1e-13


def test_quadcfinterp_reference_known_answer():
    """lib/test/AMRTools/quadCFInterpTest.cpp, 3-D branch, tolerance 1e-12."""
    h = lshape_qcfi()
    o = Oracle(h)
    dxf = h.dx(1)
    A, B, Cc = 2.23, -7.389, 3.98
    valid, face, cf = ghost_masks(h, 1)
    for idir in range(3):
        q = lambda *x: A * x[idir] ** 2 + B * x[idir] + Cc
        coarse = h.pack(lambda x, y, z, l: q(x, y, z))[:h.cells(0)]
        exact = ghosted_from_fn(h, 1, q, dxf)
        out = o.quad_cf(1, np.where(cf, 0.0, exact), coarse)
        m = cf & (ghost_dir(h, 1) == idir)   # the reference test checks ghosts normal to idir
        if m.any():  # the fine region spans the whole domain in z: no CF faces normal to z
            assert (np.abs(out[m] - exact[m]) / np.abs(exact[m])).max() < 1e-12


def test_homogeneous_cf_is_quadratic_through_zero_coarse_value():
    """INTERPHOMO: ghost = quadratic through the two inner fine values and zero at
    the coarse cell centre, evaluated at the ghost centre."""
    h = small_amr()
    o = Oracle(h)
    rng = np.random.default_rng(6)
    from helpers import ghosted_random
    lev = 1
    pg = ghosted_random(h, lev, rng)
    out = o.homog_cf(lev, pg)
    hf, hc = h.dx(lev), h.dx(lev - 1)
    # Lagrange weights at x = h/2 through nodes -3h/2 (phi2), -h/2 (phi1), hc/2 (0)
    xs = np.array([-1.5 * hf, -0.5 * hf, 0.5 * hc])
    w = []
    for a in range(2):
        others = [xs[b] for b in range(3) if b != a]
        w.append(np.prod([(0.5 * hf - xo) / (xs[a] - xo) for xo in others]))
    c2, c1 = w  # weight of phi2 (two inward), phi1 (one inward)
    # check on one face: first box's CF ghosts in x-lo (if any)
    valid, face, cf = ghost_masks(h, lev)
    b = h.levels[lev][0]
    g = H.box_grow(b, 1)
    shp = H.box_shape(g)
    n = int(np.prod(shp))
    a = pg[:n].reshape(shp, order="F")
    r = out[:n].reshape(shp, order="F")
    m = cf[:n].reshape(shp, order="F")
    checked = 0
    for i in range(shp[0]):
        for j in range(shp[1]):
            for k in range(shp[2]):
                if not m[i, j, k]:
                    continue
                for d in range(3):
                    p = [i, j, k]
                    if p[d] == 0:
                        s = 1
                    elif p[d] == shp[d] - 1:
                        s = -1
                    else:
                        continue
                    p1 = list(p); p1[d] += s
                    p2 = list(p); p2[d] += 2 * s
                    exp = c1 * a[tuple(p1)] + c2 * a[tuple(p2)]
                    assert abs(r[i, j, k] - exp) < 1e-13 * (abs(a[tuple(p1)]) + abs(a[tuple(p2)]))
                    checked += 1
    assert checked > 0


def test_composite_operator_exact_on_quadratic():
    """With quadratic data everywhere, CF interpolation, the fine operator and the
    refluxed coarse operator are exact: L(q) = alpha*q + beta*lap(q) on every
    uncovered cell away from the domain boundary."""
    for alpha, beta in [(0.0, 1.0), (0.3, 1.2)]:
        h = small_amr(alpha=alpha, beta=beta)
        o = Oracle(h)
        q = lambda x, y, z, l: 1 + 0.3 * x + x * x + 2 * y * y + 0.5 * z * z + 0.7 * x * y - 0.4 * y * z + 0.25 * x * z
        phi = h.pack(q)
        L = o.amr_operator(phi, homog=False)
        offs = h.level_offsets()
        for lev in range(h.nlev):
            m = uncovered_mask(h, lev) & _interior(h, lev)
            Ll = L[offs[lev]:offs[lev + 1]]
            expect = alpha * phi[offs[lev]:offs[lev + 1]] + beta * 7.0
            assert np.abs(Ll - expect)[m].max() < 1e-7, (alpha, lev)


def test_composite_operator_is_conservative():
    """Sum over uncovered cells of L(phi) dV equals the net flux through the domain
    boundary: refluxing makes coarse-fine face fluxes cancel."""
    h = small_amr()
    o = Oracle(h)
    rng = np.random.default_rng(8)
    phi = rng.standard_normal(h.total_cells())
    L = o.amr_operator(phi, homog=True)
    offs = h.level_offsets()
    total = 0.0
    for lev in range(h.nlev):
        m = uncovered_mask(h, lev)
        total += L[offs[lev]:offs[lev + 1]][m].sum() * h.dx(lev) ** 3
    # boundary flux on level 0 (the only level touching the boundary): ghost = -near
    g0 = h.to_global(phi[offs[0]:offs[1]], 0)
    dx = h.dx(0)
    flux = 0.0
    for d in range(3):
        lo = np.take(g0, 0, axis=d)
        hi = np.take(g0, -1, axis=d)
        flux += ((-hi - hi) - (lo - (-lo))).sum() / dx * dx * dx
    assert abs(total - flux) < 1e-9 * np.abs(L).max()


def _tiny_amr():
    coarse = H.domain_split(H.box((0, 0, 0), (7, 7, 7)), 4, 4)
    fine = H.domain_split(H.box((4, 4, 4), (11, 11, 11)), 4, 4)
    return H.Hierarchy(H.box((0, 0, 0), (7, 7, 7)), [coarse, fine], 1.0 / 8)


def test_amr_solve_matches_dense_solve_of_composite_system():
    h = _tiny_amr()
    o = Oracle(h)
    N = h.total_cells()
    A = np.zeros((N, N))
    e = np.zeros(N)
    for c in range(N):
        e[:] = 0
        e[c] = 1
        A[:, c] = o.amr_operator(e, homog=True)
    # rows of covered coarse cells are not part of the composite system
    offs = h.level_offsets()
    keep = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    rhs = H.bump_rhs(h, 1002)
    rhs_k = np.where(keep, rhs, 0.0)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs_k, sp_, rp)
    Ak = A[np.ix_(keep, keep)]
    x = np.linalg.solve(Ak, rhs_k[keep])
    assert np.abs(phi[keep] - x).max() / np.abs(x).max() < 1e-9


def test_bottom_bicgstab_solves_bottom_system():
    h = small_single(n=64, box=32)
    o = Oracle(h)
    d = o.mg_depths - 1
    rng = np.random.default_rng(3)
    res = rng.standard_normal(o.mg_cells(d))
    corr, status = o.bottom_solve(np.zeros_like(res), res, 2, float(np.abs(res).max()), 1e-12)
    assert status == 1
    r = o.level_residual(d, corr, res, homog=True)
    assert np.abs(r).max() < 1e-9 * np.abs(res).max()


def test_level_norms_match_numpy():
    """norm(BoxLayoutData, p) for p = 0, 1, 2 equals numpy's max |x|, sum |x| and
    sqrt(sum x^2) to rounding, on every multigrid depth (box-ordered sums)."""
    h = small_single(n=32, box=16, lo=-8)
    o = Oracle(h)
    rng = np.random.default_rng(12)
    for d in range(o.mg_depths):
        x = rng.standard_normal(o.mg_cells(d))
        assert o.level_norm(d, x, 0) == np.abs(x).max()
        assert abs(o.level_norm(d, x, 1) - np.abs(x).sum()) < 1e-12 * np.abs(x).sum()
        assert abs(o.level_norm(d, x, 2) - np.sqrt((x * x).sum())) < 1e-12 * np.sqrt((x * x).sum())


@pytest.mark.parametrize("numMG", [-1, -2])
def test_fcycle_solve_matches_sparse_direct_solve(numMG):
    """F cycles (MultiGrid::cycle with m_cycle < 0) on one level: the solve converges
    to the sparse direct solution, in no more outer iterations than V-cycles."""
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    rhs = H.sinusoid_rhs(h, 1001)
    sp_, rp = H.solver_params(eps=1e-12, numMG=numMG)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(_assemble_dirichlet(n).tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n),
                                                                                                  order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-10
    _, hist_v, info_v = o.solve(rhs, *H.solver_params(eps=1e-12))
    assert info["iterations"] <= info_v["iterations"]


@pytest.mark.parametrize("alpha,beta", [(0.0, 1.0), (0.7, 1.3)])
def test_jacobi_sweep_on_sine_mode(alpha, beta):
    """levelJacobi (s_relaxMode 4): one sweep on a Dirichlet sine mode e with rhs 0
    gives e (1 - 0.666 (alpha + beta lam) / (alpha - 6 beta / dx^2)), lam the
    discrete eigenvalue of the Laplacian."""
    n, k = 16, (1, 2, 3)
    h = small_single(n=n, box=8, alpha=alpha, beta=beta)
    o = Oracle(h)
    o.set_relax_mode(4)
    dx = 1.0 / n
    e = h.pack(lambda x, y, z, l: np.sin(np.pi * k[0] * x) * np.sin(np.pi * k[1] * y) * np.sin(np.pi * k[2] * z))
    lam = sum((2 / dx ** 2) * (np.cos(np.pi * kk * dx) - 1) for kk in k)
    out = o.level_relax(0, e, np.zeros_like(e), 1)
    factor = 1 - 0.666 * (alpha + beta * lam) / (alpha - 6 * beta / dx ** 2)
    assert np.abs(out - factor * e).max() < 1e-12 * np.abs(e).max()
    # fixed point: rhs = L(phi) leaves phi unchanged
    rng = np.random.default_rng(14)
    phi = rng.standard_normal(h.cells(0))
    f = o.level_apply(0, phi, homog=True)
    assert np.abs(o.level_relax(0, phi, f, 3) - phi).max() < 1e-11 * np.abs(phi).max()


def test_jacobi_multigrid_solve_matches_sparse_direct_solve():
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    o.set_relax_mode(4)
    rhs = H.sinusoid_rhs(h, 1001)
    sp_, rp = H.solver_params(eps=1e-12, pre=4, post=4)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(_assemble_dirichlet(n).tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n),
                                                                                                  order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-10


def _colour_sweep(A, L, x, b, colours, lo, n, d):
    """A brute-force eight-colour sweep on assembled matrices: for each colour in
    order, x += (b - L x) / d on the cells whose absolute index parities are that
    colour (L is the full operator for levelGSRBLazy, the Laplacian alone for
    levelMultiColor's GSMCAMRPOP)."""
    i = np.arange(n) + lo
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    par = [(I & 1).ravel(order="F"), (J & 1).ravel(order="F"), (K & 1).ravel(order="F")]
    x = x.copy()
    for c in colours:
        m = (par[0] == c[0]) & (par[1] == c[1]) & (par[2] == c[2])
        r = b - L @ x
        x[m] += r[m] / d
    return x


LAZY_COLOURS = [(c & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)]
MULTI_COLOURS = [(0, 0, 0), (1, 1, 0), (0, 1, 1), (1, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, 1), (1, 1, 1)]


@pytest.mark.parametrize("mode", [3, 5])
@pytest.mark.parametrize("alpha,beta,lo", [(0.0, 1.0, 0), (0.7, 1.3, -8), (0.0, 1.0, -5)])
def test_colour_relax_matches_assembled_sweep(mode, alpha, beta, lo):
    """levelGSRBLazy (mode 3) and levelMultiColor (mode 5): two sweeps equal two
    brute-force sweeps over an independently assembled Dirichlet operator with the
    colours in the reference's orders (nextColorLoc, x fastest; amrpgetMultiColors);
    the relaxation weight is 1 / (alpha - 6 beta / dx^2) on every cell, boundary
    cells included, and GSMCAMRPOP uses the Laplacian alone."""
    n = 16
    h = small_single(n=n, box=8, block=1 if lo % 8 else 8, lo=lo, alpha=alpha, beta=beta)
    o = Oracle(h)
    o.set_relax_mode(mode)
    rng = np.random.default_rng(21 + mode)
    phi = rng.standard_normal(h.cells(0))
    rhs = rng.standard_normal(h.cells(0))
    out = h.to_global(o.level_relax(0, phi, rhs, 2), 0).ravel(order="F")
    dx = 1.0 / n
    A = _assemble_dirichlet(n, alpha, beta)
    L = A if mode == 3 else _assemble_dirichlet(n, 0.0, 1.0)
    x = h.to_global(phi, 0).ravel(order="F")
    b = h.to_global(rhs, 0).ravel(order="F")
    for _ in range(2):
        x = _colour_sweep(A, L, x, b, LAZY_COLOURS if mode == 3 else MULTI_COLOURS, lo, n, alpha - 6 * beta / dx ** 2)
    assert np.abs(out - x).max() < 1e-12 * np.abs(x).max()
    # the other colour order (or the other mode's update) is a different sweep
    y = h.to_global(phi, 0).ravel(order="F")
    for _ in range(2):
        y = _colour_sweep(A, L, y, b, MULTI_COLOURS if mode == 3 else LAZY_COLOURS, lo, n,
                          alpha - 6 * beta / dx ** 2)
    assert np.abs(out - y).max() > 1e-6 * np.abs(y).max()


def test_colour_relax_interior_cells_are_gauss_seidel():
    """Away from the boundary one colour step of levelGSRBLazy zeroes the residual on
    that colour's cells (exact point Gauss-Seidel): after a whole sweep the last
    colour's interior residual is zero."""
    n = 16
    h = small_single(n=n, box=8, alpha=0.4, beta=1.0)
    o = Oracle(h)
    o.set_relax_mode(3)
    rng = np.random.default_rng(31)
    phi = rng.standard_normal(h.cells(0))
    rhs = rng.standard_normal(h.cells(0))
    out = o.level_relax(0, phi, rhs, 1)
    r = h.to_global(rhs - o.level_apply(0, out, homog=True), 0)
    inner = r[1:-1, 1:-1, 1:-1]
    last = inner[0::2, 0::2, 0::2]     # absolute parity (1, 1, 1) starts at index 1
    assert np.abs(last).max() < 1e-10 * np.abs(r).max()
    assert np.abs(inner[1::2, 1::2, 1::2]).max() > 1e-3 * np.abs(r).max()


@pytest.mark.parametrize("mode", [3, 5])
def test_colour_relax_multigrid_solve_matches_sparse_direct_solve(mode):
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    o.set_relax_mode(mode)
    rhs = H.sinusoid_rhs(h, 1003)
    sp_, rp = H.solver_params(eps=1e-12)
    phi, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(_assemble_dirichlet(n).tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n),
                                                                                                  order="F")
    assert np.abs(h.to_global(phi, 0) - x).max() / np.abs(x).max() < 1e-10


def test_colour_relax_modes_refused_with_variable_coefficients():
    h = small_single(n=16, box=8)
    o = Oracle(h)
    a, b = H.vc_coefficients(h, seed=3)
    o.vc_set(a, b)
    for mode in (3, 5):
        with pytest.raises(RuntimeError):
            o.set_relax_mode(mode)
    for mode in (0, 2, 6):
        with pytest.raises(RuntimeError):
            o.set_relax_mode(mode)


@pytest.mark.parametrize("case,lev", [("qcfi", 1), ("lshape", 1), ("lshape", 2), ("amr", 1), ("amr", 2)])
def test_quadcfinterp_never_reads_covered_coarse_cells(case, lev):
    """Every QuadCFStencil branch (centred, shifted, dropped order, mixed) may only
    use coarse cells that no finer box covers (ivsAllGood, QuadCFStencil.cpp:578-654):
    coarse values under the fine level are never read, so poisoning them with 1e300
    leaves every CF ghost bitwise unchanged.  A centred stencil taken where a
    neighbour is covered reads a poisoned cell."""
    h = {"qcfi": lshape_qcfi, "lshape": lshape_amr, "amr": small_amr}[case]()
    o = Oracle(h)
    rng = np.random.default_rng(40 + lev)
    offs = h.level_offsets()
    coarse = rng.standard_normal(h.cells(lev - 1))
    covered = ~uncovered_mask(h, lev - 1)
    assert covered.any()
    poisoned = np.where(covered, 1e300, coarse)
    from helpers import ghosted_random
    pg = ghosted_random(h, lev, rng)
    _, _, cf = ghost_masks(h, lev)
    a = o.quad_cf(lev, pg, coarse)
    b = o.quad_cf(lev, pg, poisoned)
    assert np.isfinite(b[cf]).all() and np.abs(b[cf]).max() < 1e3
    np.testing.assert_array_equal(a[cf], b[cf])
    # the poison is not inert: the same coarse cells are read when uncovered
    c = o.quad_cf(lev, pg, np.where(~covered, 1e300, coarse))
    assert np.abs(c[cf]).max() > 1e200


def _box_sweeps(h, n, x0, b, alpha, beta, steps):
    """Brute-force GSRB sweeps with per-box ghost copies on one Dirichlet level: a
    step (region mask, colour, lap) updates the cells of that colour in the
    region; a neighbour inside the cell's own box is read live, one in another
    box from the snapshot taken at the last refresh, and a domain ghost is the
    homogeneous Dirichlet -phi(cell) of the snapshot.  steps holds the strings
    "refresh" and (mask, colour, lap) triples."""
    dx = 1.0 / n
    idx = np.arange(n)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    owner = np.zeros((n, n, n), dtype=int)
    for q, (lo, hi) in enumerate(h.levels[0]):
        owner[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = q
    x = x0.copy()
    snap = x.copy()
    for st in steps:
        if isinstance(st, str):
            snap = x.copy()
            continue
        mask, colour, lap = st
        s6 = np.zeros_like(x)
        for d in range(3):
            for sgn in (1, -1):
                sh = np.roll(x, -sgn, axis=d)
                ssh = np.roll(snap, -sgn, axis=d)
                osh = np.roll(owner, -sgn, axis=d)
                coord = (I, J, K)[d] + sgn
                inside = (coord >= 0) & (coord < n)
                s6 += np.where(inside, np.where(osh == owner, sh, ssh), -snap)
        sel = mask & (((I + J + K + colour) & 1) == 0)
        if lap or (alpha == 0.0 and beta == 1.0):
            new = -(dx * dx / 6.0) * (b - s6 / dx ** 2)
        else:
            lam = -1.0 / (alpha - 6.0 * beta / dx ** 2)
            new = x + lam * (alpha * x + beta * (s6 - 6.0 * x) / dx ** 2 - b)
        x = np.where(sel, new, x)
    return x


def _loose_steps(full):
    return ["refresh", (full, 0, False), (full, 1, False)]


def _overlap_steps(h, n, alpha, beta):
    full = np.ones((n, n, n), dtype=bool)
    steps = _loose_steps(full) if (len(h.levels[0]) == 1 or alpha != 1 or beta != 0) else []
    steps.append("refresh")
    inner = np.zeros((n, n, n), dtype=bool)
    for (lo, hi) in h.levels[0]:
        inner[lo[0] + 1:hi[0], lo[1] + 1:hi[1], lo[2] + 1:hi[2]] = True
    steps += [(inner, 0, True), (inner, 1, True)]
    for dir in range(3):
        for side in range(2):
            m = np.zeros((n, n, n), dtype=bool)
            for (lo, hi) in h.levels[0]:
                sl = [slice(lo[e], hi[e] + 1) for e in range(3)]
                for e in range(dir):
                    sl[e] = slice(lo[e] + 1, hi[e])
                sl[dir] = slice(lo[dir], lo[dir] + 1) if side == 0 else slice(hi[dir], hi[dir] + 1)
                m[tuple(sl)] = True
            # every box's face layer at once: boxes read each other only from the snapshot
            steps += [(m, 0, True), (m, 1, True)]
    return steps


@pytest.mark.parametrize("mode,alpha,beta", [(0, 0.0, 1.0), (0, 0.7, 1.3), (2, 0.0, 1.0), (2, 0.7, 1.3)])
def test_loose_and_overlap_relax_match_box_sweeps(mode, alpha, beta):
    """looseGSRB (mode 0) and overlapGSRB (mode 2): one relaxation equals a
    brute-force sweep in which each box sees its neighbours and its Dirichlet
    ghosts as they were at the last ghost refresh (per-box FAB copies), with
    overlapGSRB's interior / face-layer order and its Laplacian-only passes."""
    n = 16
    h = small_single(n=n, box=8, alpha=alpha, beta=beta)
    o = Oracle(h)
    o.set_relax_mode(mode)
    rng = np.random.default_rng(50 + mode)
    phi = rng.standard_normal(h.cells(0))
    rhs = rng.standard_normal(h.cells(0))
    out = h.to_global(o.level_relax(0, phi, rhs, 1), 0)
    x0, b = h.to_global(phi, 0), h.to_global(rhs, 0)
    full = np.ones((n, n, n), dtype=bool)
    steps = _loose_steps(full) if mode == 0 else _overlap_steps(h, n, alpha, beta)
    want = _box_sweeps(h, n, x0, b, alpha, beta, steps)
    assert np.abs(out - want).max() < 1e-12 * np.abs(want).max()
    # a refresh between the colours (levelGSRB) is a different sweep
    ref1 = _box_sweeps(h, n, x0, b, alpha, beta, ["refresh", (full, 0, False), "refresh", (full, 1, False)])
    assert np.abs(out - ref1).max() > 1e-6 * np.abs(ref1).max()


@pytest.mark.parametrize("mode", [0, 2])
def test_loose_and_overlap_relax_fixed_point_and_solve(mode):
    """A manufactured discrete solution is a fixed point, and the multigrid solve
    equals a sparse direct solve."""
    n = 16
    h = small_single(n=n, box=8)
    o = Oracle(h)
    o.set_relax_mode(mode)
    rng = np.random.default_rng(60 + mode)
    phi = rng.standard_normal(h.cells(0))
    f = o.level_apply(0, phi, homog=True)
    assert np.abs(o.level_relax(0, phi, f, 2) - phi).max() < 1e-11 * np.abs(phi).max()
    rhs = H.sinusoid_rhs(h, 1005)
    sp_, rp = H.solver_params(eps=1e-12)
    sol, hist, info = o.solve(rhs, sp_, rp)
    assert info["exit_status"] & 1
    x = spla.spsolve(_assemble_dirichlet(n).tocsc(), h.to_global(rhs, 0).ravel(order="F")).reshape((n, n, n),
                                                                                                  order="F")
    assert np.abs(h.to_global(sol, 0) - x).max() / np.abs(x).max() < 1e-10
