"""The operator-level ABI (levels, fields, AMRPoissonOp virtuals, copiers) on the GPU.

A host solver written against the reference's operator virtuals (amrmultigrid.py:
AMRMultiGrid, MultiGrid and BiCGStabSolver in the reference's control flow) drives the
device operators one call per virtual and must reproduce amrmg_solver_solve -- the
library's own fused solve -- bit for bit, history and solution.  Single operator
applications are compared with the oracle.
"""
import numpy as np
import pytest

import amrmg
import amrmultigrid as AMG
import hierarchy as H
import levelops as LO
from helpers import ghosted_random, lshape_amr, mixed_ratio_amr, small_amr, small_single
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _bc(h):
    return (tuple(h.bc_type_lo) + tuple(h.bc_type_hi), tuple(h.bc_val_lo) + tuple(h.bc_val_hi))


def build_ops(ctx, h, merge=True):
    """Levels and AMRPoissonOps of a hierarchy, linked as AMRnewOp links them."""
    levels, ops = [], []
    for l in range(h.nlev):
        L = LO.Level(ctx, h.levels[l], h.level_domain(l), merge=merge)
        op = LO.AMRPoissonOp(L, h.dx(l), h.dx(l - 1) if l > 0 else -1.0, h.alpha, h.beta, _bc(h),
                             coarser=ops[-1] if ops else None)
        levels.append(L)
        ops.append(op)
    return levels, ops


def close_all(fields, ops, levels):
    for f in fields:
        f.close()
    for op in reversed(ops):
        op.close()
    for L in levels:
        L.close()


def level_slices(h):
    offs = h.level_offsets()
    return [slice(offs[l], offs[l + 1]) for l in range(h.nlev)]


def handle_solve(ctx, h, rhs, sp, rp, merge=True, l_base=0, phi0=None, bottom=None):
    """AMRMultiGrid::solve through the operator-level ABI; returns (phi, history, iterations, exit status)."""
    levels, ops = build_ops(ctx, h, merge)
    sl = level_slices(h)
    phi = [LO.Field(L) for L in levels]
    rh = [LO.Field(L) for L in levels]
    for l in range(h.nlev):
        rh[l].upload(rhs[sl[l]])
        if phi0 is not None:
            phi[l].upload(phi0[sl[l]])
    amg = AMG.AMRMultiGrid()
    bic = AMG.BiCGStabSolver()
    if bottom:
        for k, v in bottom.items():
            setattr(bic, k, v)
    amg.define(ops, [AMG.domain_cells(h.level_domain(l)) for l in range(h.nlev)], bic)
    pre, post, nb, numMG, iterMax, imin = [int(v) for v in sp]
    amg.setSolverParameters(pre, post, nb, numMG, iterMax, float(rp[0]), float(rp[1]), float(rp[2]))
    amg.m_imin = imin
    amg.solve(phi, rh, h.nlev - 1, l_base, zeroPhi=phi0 is None)
    out = np.concatenate([phi[l].download() for l in range(h.nlev)])
    res = (out, np.asarray(amg.m_history), amg.m_iterations, amg.m_exitStatus)
    close_all(phi + rh, ops, levels)
    return res


@pytest.mark.parametrize("case", ["single32", "helm24", "amr32", "lshape", "amr32_w", "ratio42", "ratio24_helm"])
def test_handle_driven_solve_equals_library_solve(gpu_ctx, case):
    """The reference's AMRMultiGrid loop over the device operators equals the library's
    fused solve bit for bit (the fusions, buffer swaps and the device bottom solver
    change no result)."""
    print(f"handle solve {case}", flush=True)
    kw = {}
    if case == "single32":
        h = small_single(n=32, box=16)
        rhs = H.sinusoid_rhs(h, 1001)
    elif case == "helm24":
        h = small_single(n=24, box=8, alpha=0.4, beta=1.7)
        rhs = H.sinusoid_rhs(h, 1004, k=(1, 2, 3))
    elif case == "lshape":
        h = lshape_amr()
        rhs = H.bump_rhs(h, 1002)
    elif case.startswith("ratio"):   # refinement ratios 4 then 2, or 2 then 4
        h = mixed_ratio_amr((4, 2)) if case == "ratio42" else mixed_ratio_amr((2, 4), (0, 1, 2), 0.5, 1.2)
        rhs = H.bump_rhs(h, 1061)
        kw = dict(iterMax=60)
    else:
        h = small_amr()
        rhs = H.bump_rhs(h, 1002)
        if case.endswith("_w"):
            kw = dict(numMG=2, iterMax=6)
    sp, rp = H.solver_params(**kw)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    g.close()
    phi_h, hist_h, it_h, st_h = handle_solve(gpu_ctx, h, rhs, sp, rp, merge=case != "lshape")
    assert it_h == info_g["iterations"] and st_h == info_g["exit_status"]
    np.testing.assert_array_equal(hist_h, hist_g)
    np.testing.assert_array_equal(phi_h, phi_g)


def test_handle_one_outer_iteration_amr64(gpu_ctx):
    """One outer iteration of the AMR V-cycle on the benchmark's hierarchy at base 64
    (64^3, 80^3, 100^3 levels, 32^3 boxes), driven only through the handles, equals
    the library's solve stopped after one iteration."""
    h = H.amr_hierarchy(64, fracs=(0.25, 0.06), max_box=32, block=8)
    rhs = H.bump_rhs(h, 1002)
    sp, rp = H.solver_params(iterMax=1)
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    g.close()
    phi_h, hist_h, it_h, st_h = handle_solve(gpu_ctx, h, rhs, sp, rp)
    assert it_h == 1 and info_g["iterations"] == 1 and st_h == info_g["exit_status"]
    np.testing.assert_array_equal(hist_h, hist_g)
    np.testing.assert_array_equal(phi_h, phi_g)


@pytest.mark.parametrize("merge", [True, False])
def test_handle_level_operators_match_oracle(gpu_ctx, merge):
    """applyOp / residual / relax / restrictResidual / prolongIncrement / norms / dot
    through the handles on every multigrid depth of one level, against the oracle."""
    h = small_single(n=32, box=16, lo=-16)
    o = Oracle(h)
    L0 = LO.Level(gpu_ctx, h.levels[0], h.level_domain(0), merge=merge)
    op0 = LO.AMRPoissonOp(L0, h.dx(0), -1.0, h.alpha, h.beta, _bc(h))
    ops, levels = [op0], [L0]
    while len(ops) < o.mg_depths:
        ops.append(ops[-1].createCoarser())
        levels.append(ops[-1].level)
    rng = np.random.default_rng(31)
    fields = []
    for d in range(o.mg_depths):
        op, L = ops[d], levels[d]
        n = o.mg_cells(d)
        assert L.cells() == n
        phi_h, rhs_h = rng.standard_normal(n), rng.standard_normal(n)
        phi, rhs, out = LO.Field(L), LO.Field(L), LO.Field(L)
        fields += [phi, rhs, out]
        for homog in (False, True):
            phi.upload(phi_h)
            rhs.upload(rhs_h)
            op.applyOp(out, phi, homog)
            np.testing.assert_array_equal(out.download(), o.level_apply(d, phi_h, homog))
            op.residual(out, phi, rhs, homog)
            np.testing.assert_array_equal(out.download(), o.level_residual(d, phi_h, rhs_h, homog))
        phi.upload(phi_h)
        op.relax(phi, rhs, 2)
        np.testing.assert_array_equal(phi.download(), o.level_relax(d, phi_h, rhs_h, 2))
        for p in (0, 1, 2):
            assert op.norm(rhs, p) == o.level_norm(d, rhs_h, p)
        a = LO.Field(L).upload(phi_h)
        fields.append(a)
        # LevelDataOps::dotProduct: per box a serial sum in Fortran order, boxes in layout order
        want, o_ = 0.0, 0
        for (lo, hi) in L.boxes:
            nb = int(np.prod([hi[q] - lo[q] + 1 for q in range(3)]))
            part = 0.0
            for x, y in zip(phi_h[o_:o_ + nb], rhs_h[o_:o_ + nb]):
                part += x * y
            want += part
            o_ += nb
        assert op.dotProduct(a, rhs) == want
        if d + 1 < o.mg_depths:
            rc = LO.Field(levels[d + 1])
            fields.append(rc)
            phi.upload(phi_h)
            op.restrictResidual(rc, phi, rhs)
            np.testing.assert_array_equal(rc.download(), o.level_restrict(d, phi_h, rhs_h))
            ec = rng.standard_normal(o.mg_cells(d + 1))
            rc.upload(ec)
            phi.upload(phi_h)
            op.prolongIncrement(phi, rc)
            np.testing.assert_array_equal(phi.download(), o.level_prolong(d, phi_h, ec))
    for f in fields:
        f.close()
    for op in reversed(ops[1:]):
        op.close()
    op0.close()
    L0.close()


def test_handle_dot_matches_solver_bottom_order(gpu_ctx):
    """dotProduct through the handles is the reference's ordered dot: equal to the
    library's BiCGStab bottom solve driven by the handles (host BiCGStab in Python)."""
    h = small_single(n=64, box=32)
    o = Oracle(h)
    d = o.mg_depths - 1
    rng = np.random.default_rng(3)
    res = rng.standard_normal(o.mg_cells(d))
    co, so = o.bottom_solve(np.zeros_like(res), res, 2, 1.0, 1e-10)
    L0 = LO.Level(gpu_ctx, h.levels[0], h.level_domain(0))
    ops = [LO.AMRPoissonOp(L0, h.dx(0), -1.0, h.alpha, h.beta, _bc(h))]
    while len(ops) < o.mg_depths:
        ops.append(ops[-1].createCoarser())
    Lb = ops[-1].level
    c, r = LO.Field(Lb), LO.Field(Lb).upload(res)
    ops[-1].relax(c, r, 2)
    bic = AMG.BiCGStabSolver()
    bic.define(ops[-1], True)
    bic.setConvergenceMetrics(1.0, 1e-10)
    bic.solve(c, r)
    assert bic.m_exitStatus == so
    np.testing.assert_array_equal(c.download(), co)
    c.close()
    r.close()
    for op in reversed(ops):
        op.close()
    L0.close()


def test_handle_amr_pieces_match_oracle(gpu_ctx):
    """QuadCFInterp (coarseFineInterp), homogeneousCFInterp, AMROperator with reflux
    and AMRResidual through the handles on the three-level hierarchy, against the
    oracle; exchange / copier plans built from explicit MotionItems equal the derived
    ones."""
    h = lshape_amr()
    o = Oracle(h)
    levels, ops = build_ops(gpu_ctx, h, merge=False)
    sl = level_slices(h)
    rng = np.random.default_rng(8)
    phi_h = rng.standard_normal(h.total_cells())
    rhs_h = rng.standard_normal(h.total_cells())
    phi = [LO.Field(L).upload(phi_h[sl[l]]) for l, L in enumerate(levels)]
    rhs = [LO.Field(L).upload(rhs_h[sl[l]]) for l, L in enumerate(levels)]
    out = [LO.Field(L) for L in levels]
    # computeAMROperator / computeAMRResidual level by level (l_base 0)
    lmax = h.nlev - 1
    for homog in (False, True):
        for l in range(h.nlev):
            if l == lmax:
                ops[l].AMROperatorNF(out[l], phi[l], phi[l - 1], homog)
            elif l == 0:
                ops[l].AMROperatorNC(out[l], phi[l + 1], phi[l], homog, ops[l + 1])
            else:
                ops[l].AMROperator(out[l], phi[l + 1], phi[l], phi[l - 1], homog, ops[l + 1])
        got = np.concatenate([f.download() for f in out])
        want = o.amr_operator(phi_h, homog)
        from helpers import uncovered_mask
        for l in range(h.nlev):
            m = uncovered_mask(h, l)
            np.testing.assert_array_equal(got[sl[l]][m], want[sl[l]][m])
        for l in range(h.nlev):
            if l == lmax:
                ops[l].AMRResidualNF(out[l], phi[l], phi[l - 1], rhs[l], homog)
            elif l == 0:
                ops[l].AMRResidualNC(out[l], phi[l + 1], phi[l], rhs[l], homog, ops[l + 1])
            else:
                ops[l].AMRResidual(out[l], phi[l + 1], phi[l], phi[l - 1], rhs[l], homog, ops[l + 1])
            if l != lmax:
                ops[l].zeroCovered(out[l])
        got = np.concatenate([f.download() for f in out])
        want, _ = o.amr_residual(phi_h, rhs_h, homog)
        np.testing.assert_array_equal(got, want)
    # QuadCFInterp and homogeneous CF on the ghosted FABs of level 1 (stored box by box)
    from helpers import ghost_masks
    valid, face, cf = ghost_masks(h, 1)
    g1 = ghosted_random(h, 1, rng)
    coarse = rng.standard_normal(h.cells(0))
    f1 = LO.Field(levels[1]).upload(g1, ghosted=True)
    c0 = LO.Field(levels[0]).upload(coarse)
    ops[1].coarseFineInterp(f1, c0)
    np.testing.assert_array_equal(f1.download(ghosted=True)[cf], o.quad_cf(1, g1, coarse)[cf])
    f1.upload(g1, ghosted=True)
    ops[1].homogeneousCFInterp(f1)
    np.testing.assert_array_equal(f1.download(ghosted=True)[cf], o.homog_cf(1, g1)[cf])
    # exchange: derived plan, explicit MotionItems and the operator's own exchange agree
    f1.upload(g1, ghosted=True)
    ops[1].exchange(f1)
    ex_op = f1.download(ghosted=True)
    boxes = levels[1].boxes
    items = []
    for b, (lo, hi) in enumerate(boxes):
        for d in range(3):
            for side in (0, 1):
                glo, ghi = list(lo), list(hi)
                if side == 0:
                    glo[d] = ghi[d] = lo[d] - 1
                else:
                    glo[d] = ghi[d] = hi[d] + 1
                for n, (nlo, nhi) in enumerate(boxes):
                    if n == b:
                        continue
                    ilo = [max(glo[q], nlo[q]) for q in range(3)]
                    ihi = [min(ghi[q], nhi[q]) for q in range(3)]
                    if all(ilo[q] <= ihi[q] for q in range(3)):
                        items.append([n, b, *ilo, *ihi])
    for its in (None, items):
        cp = LO.Copier(levels[1], items=its, exchange=True)
        f1.upload(g1, ghosted=True)
        cp.copyTo(f1, f1)
        np.testing.assert_array_equal(f1.download(ghosted=True), ex_op)
        cp.close()
    # copyTo / copyToZero between the coarsened level 2 and level 1 (the resC copier)
    cl = ops[2].coarsenedLevel()
    src = LO.Field(cl).upload(rng.standard_normal(cl.cells()))
    dst = LO.Field(levels[1]).upload(phi_h[sl[1]])
    cp = LO.Copier(cl, levels[1])
    cp.copyTo(src, dst)
    d1 = dst.download()
    cp.copyToZero(dst)
    d2 = dst.download()
    m = ~uncovered_mask(h, 1)
    assert np.all(d2[m] == 0) and np.array_equal(d2[~m], phi_h[sl[1]][~m])
    assert np.array_equal(d1[~m], phi_h[sl[1]][~m]) and not np.array_equal(d1[m], phi_h[sl[1]][m])
    cp.close()
    close_all(phi + rhs + out + [f1, c0, src, dst], ops, levels)


def test_handle_errors(gpu_ctx):
    """Operands on the wrong level, destroying a coarser operator first and unsorted
    boxes are refused with an error."""
    h = small_amr()
    levels, ops = build_ops(gpu_ctx, h)
    a = LO.Field(levels[0])
    b = LO.Field(levels[1])
    with pytest.raises(amrmg.AMRMGError):
        ops[0].applyOp(b, a, True)
    with pytest.raises(amrmg.AMRMGError):
        ops[0].close()          # level 1's operator still links to it
    with pytest.raises(amrmg.AMRMGError):
        levels[0].close()       # fields and operators live on it
    with pytest.raises(amrmg.AMRMGError):
        LO.Level(gpu_ctx, list(reversed(h.levels[0])), h.level_domain(0))
    close_all([a, b], ops, levels)
