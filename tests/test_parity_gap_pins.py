"""Pins of the oracle branches that closed forms on quadratics leave open.

1. QuadCFStencil's stencil choice (lib/src/AMRTools/QuadCFStencil.cpp:662-918):
   centred, shifted toward +, shifted toward -, or dropped order, per coarse cell and
   tangential direction.  Every choice is exact on quadratics, so the quadratic
   known-answer test cannot tell them apart.  With cubic data they differ (the
   centred second difference of u^3 is exact, the + shifted one errs by +6h, the -
   shifted one by -6h), and the choice rule is mirror-symmetric: reflecting the
   hierarchy and the data must reflect every coarse-fine ghost.  Preferring a
   shifted stencil over the centred one, a one-sided stencil with a wrong weight or
   sign, or an asymmetric fallback breaks the reflection on the hierarchies below,
   whose stencils take every branch (counted by oracle_cf_branch_counts).  The
   order between the + and - shifted stencils cannot change any output: both need
   the cell itself, and a cell whose two neighbours are both admissible takes the
   centred stencil first.
2. BiCGStab restarts (BiCGStabSolver.H:432-488): forced with hang = 1 (every second
   iteration fails the hang test), the number of iterations before exit status 3 is
   2 (numRestarts + 1); the iterate after a restart equals a fresh BiCGStab started
   from the restart's phi.
3. INTERPHOMOLINEAR (InterpF.ChF:466-494; AMRPoissonOp.cpp:1642, 1714): on a level
   with boxes one cell wide, the homogeneous coarse-fine ghost across such a box is
   the linear interpolant through the inner cell and a zero coarse value, packed
   face and irregular face alike.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import ghost_dir, ghost_masks, lshape_qcfi, small_single
from oracle import Oracle


# ---------------------------------------------------------------------------
# 1. QuadCFStencil: reflection equivariance with cubic data
# ---------------------------------------------------------------------------
def lshape_pow2():
    """The quadCFInterpTest hierarchy with dx0 = 1/16 (exact cell centres)."""
    h = lshape_qcfi()
    return H.Hierarchy(h.domain, h.levels, 1.0 / 16)


def notch():
    """Three fine boxes: a notch, a box touching the domain's lo-z face, a narrow block."""
    dom = H.box((0, 0, 0), (15, 15, 15))
    fine = [((4, 2, 0), (11, 9, 5)), ((4, 10, 0), (7, 13, 5)), ((12, 2, 2), (13, 5, 5))]
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), [H.box_refine(b, 2) for b in fine]], 1.0 / 16)


def mirror(h, a):
    """The hierarchy reflected along axis a (each level about its domain's centre)."""
    levels = []
    for l, lv in enumerate(h.levels):
        D = h.level_domain(l)
        s = D[0][a] + D[1][a]
        out = []
        for (lo, hi) in lv:
            lo2, hi2 = list(lo), list(hi)
            lo2[a], hi2[a] = s - hi[a], s - lo[a]
            out.append(H.box(lo2, hi2))
        levels.append(out)
    return H.Hierarchy(h.domain, levels, h.dx0)


def mirror_ghosted(h, hm, lev, g, a):
    """Ghosted-packed data of level lev of h, reflected onto hm."""
    D = h.level_domain(lev)
    s = D[0][a] + D[1][a]
    chunks, o = {}, 0
    for (lo, hi) in h.levels[lev]:
        shp = tuple(hi[d] - lo[d] + 3 for d in range(3))
        n = int(np.prod(shp))
        arr = np.flip(g[o:o + n].reshape(shp, order="F"), axis=a)
        lo2, hi2 = list(lo), list(hi)
        lo2[a], hi2[a] = s - hi[a], s - lo[a]
        chunks[(tuple(lo2), tuple(hi2))] = arr.ravel(order="F")
        o += n
    return np.concatenate([chunks[b] for b in hm.levels[lev]])


def mirror_valid(h, hm, lev, v, a):
    g = h.to_global(np.concatenate([np.zeros(h.level_offsets()[lev]), v]), lev)
    gm = np.flip(g, axis=a)
    arrays = [np.zeros(H.box_shape(hm.level_domain(l))) for l in range(hm.nlev)]
    arrays[lev] = gm
    offs = hm.level_offsets()
    return hm.from_global(arrays)[offs[lev]:offs[lev + 1]]


def centred_coords(h, lev):
    """Per level: cell-centre coordinates measured from the level domain's centre (exact)."""
    D = h.level_domain(lev)
    dx = h.dx(lev)
    c = [0.5 * (D[0][d] + D[1][d] + 1) * dx for d in range(3)]
    return lambda x, y, z: (x - c[0], y - c[1], z - c[2])


def data_on(h, lev, fn):
    """fn(u, v, w) of centred coordinates on the valid cells of level lev (valid-packed)
    and on the ghosted FABs (ghosted-packed)."""
    cc = centred_coords(h, lev)
    offs = h.level_offsets()
    valid = h.pack(lambda x, y, z, l: fn(*cc(x, y, z)) if l == lev else np.zeros_like(x))[offs[lev]:offs[lev + 1]]
    from helpers import ghosted_from_fn
    ghosted = ghosted_from_fn(h, lev, lambda x, y, z: fn(*cc(x, y, z)), h.dx(lev))
    return valid, ghosted


def cf_ghosts(o, h, fn):
    """QuadCFInterp ghosts of level 1 for data fn on both levels (CF ghosts only)."""
    coarse, _ = data_on(h, 0, fn)
    _, fine_g = data_on(h, 1, fn)
    valid, face, cf = ghost_masks(h, 1)
    start = np.where(cf, 0.0, fine_g)
    return o.quad_cf(1, start, coarse), cf


CUBICS = {
    # odd in the mirror axis a (sign -1) and even (sign +1); b, c are the other axes
    "u_a^3": (lambda u, a, b, c: u[a] ** 3, -1),
    "u_a^3 u_b^2": (lambda u, a, b, c: u[a] ** 3 * u[b] ** 2, -1),
    "u_a^2 u_b": (lambda u, a, b, c: u[a] ** 2 * u[b], +1),
    "u_a u_b u_c": (lambda u, a, b, c: u[a] * u[b] * u[c], -1),
    "u_b^3 + u_a^2 u_c": (lambda u, a, b, c: u[b] ** 3 + u[a] ** 2 * u[c], +1),
}


@pytest.mark.parametrize("hname", ["lshape", "notch"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_quadcf_stencil_choice_reflection(hname, axis):
    h = lshape_pow2() if hname == "lshape" else notch()
    hm = mirror(h, axis)
    o, om = Oracle(h), Oracle(hm)
    b, c = [d for d in range(3) if d != axis]
    for name, (f, sign) in CUBICS.items():
        fn = lambda x, y, z: f((x, y, z), axis, b, c)
        g, cf = cf_ghosts(o, h, fn)
        gm, cfm = cf_ghosts(om, hm, fn)
        back = mirror_ghosted(hm, h, 1, gm, axis)
        assert np.array_equal(mirror_ghosted(hm, h, 1, cfm.astype(float), axis) > 0, cf)
        scale = np.abs(g[cf]).max()
        diff = np.abs(back[cf] - sign * g[cf]).max()
        assert diff <= 1e-13 * scale, (name, diff, scale)


def test_quadcf_stencil_hierarchies_reach_every_branch():
    """The hierarchies of the reflection test take every stencil branch that the
    choice rule has (drop-order with a zero first derivative included)."""
    c1 = Oracle(lshape_pow2()).cf_branch_counts(1)
    c2 = Oracle(notch()).cf_branch_counts(1)
    for k in ("centred", "shift_plus", "shift_minus", "drop_plus", "drop_minus", "drop_zero", "mixed1", "mixed2"):
        assert c1[k] > 0, k
    assert c2["mixed3"] > 0


def test_quadcf_cubic_data_separates_the_stencils():
    """On cubic data the coarse-fine ghosts are not exact (the quadratic test's
    guarantee ends there) and a shifted stencil's error differs from the centred
    one's: the errors at the ghosts of special cells differ from those at standard
    cells with the same offsets, so the reflection test compares different values."""
    h = lshape_pow2()
    o = Oracle(h)
    fn = lambda x, y, z: (x - 0.5) ** 3
    g, cf = cf_ghosts(o, h, fn)
    from helpers import ghosted_from_fn
    exact = ghosted_from_fn(h, 1, lambda x, y, z: (x - 0.5) ** 3, h.dx(1))
    err = (g - exact)[cf & (ghost_dir(h, 1) != 0)]
    # distinct error values: the branches leave distinct signatures
    assert np.unique(np.round(err / np.abs(err).max(), 12)).size >= 4
    assert np.abs(err).max() > 1e-7


# ---------------------------------------------------------------------------
# 2. BiCGStab restarts
# ---------------------------------------------------------------------------
def _bottom_case():
    h = small_single(n=40, box=40, block=8)   # multigrid bottom: one 10^3 box
    o = Oracle(h)
    d = o.mg_depths - 1
    rng = np.random.default_rng(21)
    res = rng.standard_normal(o.mg_cells(d))
    return o, res


@pytest.mark.parametrize("restarts", [0, 1, 3])
def test_bicgstab_restart_exhaustion(restarts):
    """hang = 1: every iteration fails the hang test, so recount makes every second
    iteration restart; after numRestarts restarts the next one ends the solve with
    exit status 3 at iteration 2 (numRestarts + 1)."""
    o, res = _bottom_case()
    o.set_bottom_params(imax=200, numRestarts=restarts, reps=1e-300, hang=1.0)
    c, st, it = o.bottom_solve(np.zeros_like(res), res, 2, 1.0, 1e-300, iterations=True)
    assert st == 3
    assert it == 2 * (restarts + 1)


def test_bicgstab_iterate_after_restart_equals_fresh_solve():
    """A solve with two restarts equals: a solve stopped at its first restart (phi
    there is the restart's phi), then a fresh BiCGStab from that phi with one restart
    left and no bottom relaxation -- bit for bit.  Dropping the correction's reset at
    the restart, the new shadow residual or the recomputed residual breaks this."""
    o, res = _bottom_case()
    z = np.zeros_like(res)
    o.set_bottom_params(imax=200, numRestarts=2, reps=1e-300, hang=1.0)
    full, st, it = o.bottom_solve(z, res, 2, 1.0, 1e-300, iterations=True)
    assert (st, it) == (3, 6)
    o.set_bottom_params(imax=200, numRestarts=0, reps=1e-300, hang=1.0)
    first, st, it = o.bottom_solve(z, res, 2, 1.0, 1e-300, iterations=True)
    assert (st, it) == (3, 2)
    o.set_bottom_params(imax=200, numRestarts=1, reps=1e-300, hang=1.0)
    rest, st, it = o.bottom_solve(first, res, 0, 1.0, 1e-300, iterations=True)
    assert (st, it) == (3, 4)
    np.testing.assert_array_equal(rest, full)
    # without restarts the same six iterations give another iterate
    o.set_bottom_params(imax=6, numRestarts=2, reps=1e-300, hang=0.0)
    plain, st, it = o.bottom_solve(z, res, 2, 1.0, 1e-300, iterations=True)
    assert it == 6 and np.abs(plain - full).max() > 1e-6 * np.abs(full).max()


def test_bicgstab_restarted_solve_still_solves():
    """Restarting from the current phi keeps solving the same system: with the default
    hang and a converged tolerance the restarted and unrestarted solves both reach it."""
    o, res = _bottom_case()
    z = np.zeros_like(res)
    o.set_bottom_params(imax=400, numRestarts=1000, reps=1e-300, hang=0.5)
    c, st, it = o.bottom_solve(z, res, 0, -1.0, 1e-10, iterations=True)
    assert st == 1
    d = o.mg_depths - 1
    r = o.level_residual(d, c, res, homog=True)
    assert np.sqrt((r ** 2).sum()) <= 1e-10 * np.sqrt((res ** 2).sum()) * 1.0001


# ---------------------------------------------------------------------------
# 3. INTERPHOMOLINEAR: boxes one cell wide
# ---------------------------------------------------------------------------
def thin_boxes():
    """Two levels; level 1 holds boxes one cell wide in x: D (packed CF face at its
    lo-x side) and B (irregular CF face at its hi-x side, partly covered by C)."""
    dom = H.box((0, 0, 0), (31, 31, 31))
    fine = [((15, 16, 16), (15, 31, 31)),    # D: 1 wide in x
            ((16, 16, 16), (31, 31, 31)),    # A
            ((32, 16, 16), (32, 31, 31)),    # B: 1 wide in x
            ((33, 16, 16), (40, 23, 23))]    # C: covers part of B's hi-x face
    return H.Hierarchy(dom, [H.domain_split(dom, 16, 8), fine], 1.0 / 32)


def test_interphomolinear_closed_form():
    """homogeneousCFInterp on level 1: across a box one cell wide the ghost is the
    straight line through the inner cell (at -h/2 from the face) and the coarse cell
    centre (at +H/2, value 0), taken at +h/2: (H - h) / (H + h) phi_in, on the packed
    face (INTERPHOMOLINEAR) and on the irregular face (interpOnIVSHomo).  Across wider
    boxes it is the quadratic through two inner cells and the zero coarse value."""
    h = thin_boxes()
    o = Oracle(h)
    hf, Hc = h.dx(1), h.dx(0)
    rng = np.random.default_rng(4)
    g = rng.standard_normal(h.ghosted_cells(1))
    out = o.homog_cf(1, g)
    valid, face, cf = ghost_masks(h, 1)
    dirs = ghost_dir(h, 1)
    o_ = 0
    nlin = {"packed": 0, "irregular": 0}
    nquad = 0
    for bi, (lo, hi) in enumerate(h.levels[1]):
        shp = tuple(hi[d] - lo[d] + 3 for d in range(3))
        n = int(np.prod(shp))
        G = g[o_:o_ + n].reshape(shp, order="F")
        O = out[o_:o_ + n].reshape(shp, order="F")
        M = cf[o_:o_ + n].reshape(shp, order="F")
        Dr = dirs[o_:o_ + n].reshape(shp, order="F")
        for idx in zip(*np.nonzero(M)):
            d = int(Dr[idx])
            side = 1 if idx[d] == shp[d] - 1 else -1
            inner = list(idx)
            inner[d] -= side
            pa = G[tuple(inner)]
            if shp[d] == 3:                      # one cell wide: linear through (-h/2, pa), (H/2, 0)
                xa, xc, xg = -hf / 2, Hc / 2, hf / 2
                want = pa + (0.0 - pa) * (xg - xa) / (xc - xa)
                np.testing.assert_allclose(O[idx], want, rtol=1e-14, atol=1e-15)
                faceset = "packed" if bi == 0 else "irregular"
                nlin[faceset] += 1
            else:                                # quadratic through (-3h/2, pb), (-h/2, pa), (H/2, 0)
                inner2 = list(inner)
                inner2[d] -= side
                pb = G[tuple(inner2)]
                xs = np.array([-1.5 * hf, -0.5 * hf, Hc / 2])
                ys = np.array([pb, pa, 0.0])
                xg = hf / 2
                want = sum(ys[i] * np.prod([(xg - xs[j]) / (xs[i] - xs[j]) for j in range(3) if j != i])
                           for i in range(3))
                np.testing.assert_allclose(O[idx], want, rtol=1e-12, atol=1e-14)
                nquad += 1
        o_ += n
    assert nlin["packed"] == 16 * 16          # D's lo-x face
    assert nlin["irregular"] == 16 * 16 - 8 * 8   # B's hi-x face minus what C covers
    assert nquad > 0


def test_interphomolinear_in_level_operators():
    """applyOp / residual (homogeneous) and relax on level 1 read the linear ghosts:
    a field that is zero except on the one-cell-wide box D gives D's lo-x neighbour
    term (H - h)/(H + h) phi_D in the operator."""
    h = thin_boxes()
    o = Oracle(h)
    n1 = h.cells(1)
    phi = np.zeros(n1)
    # D is box 0: valid cells 0 .. 255 (1 x 16 x 16)
    phi[:256] = 1.0
    lphi = o.amrlevel_apply(1, phi, homog=True)
    hf, Hc = h.dx(1), h.dx(0)
    lin = (Hc - hf) / (Hc + hf)
    # an interior cell of D (not on D's y/z faces): x-neighbours are the CF ghost (lin) and A's cell (0);
    # y/z neighbours are D's own cells (1)
    D = lphi[:256].reshape((1, 16, 16), order="F")
    want = ((lin + 0.0 + 4 * 1.0) - 6 * 1.0) / hf ** 2
    np.testing.assert_allclose(D[0, 5:11, 5:11], want, rtol=1e-13)
