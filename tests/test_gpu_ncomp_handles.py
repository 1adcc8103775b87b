"""Multi-component fields through the operator-level ABI (amrmg_field_create_ncomp,
amrmg_op_set_ncomp): the reference's AMRMultiGrid loop over device operators and
fields of two components equals the library's two-component solve bit for bit, and
the oracle's; transfers, dot products and norms of such fields."""
import numpy as np
import pytest

import amrmg
import amrmultigrid as AMG
import hierarchy as H
import levelops as LO
from helpers import comps_pack, lshape_amr, small_amr
from oracle import Oracle
from test_gpu_levelops import build_ops, close_all

pytestmark = pytest.mark.gpu


def _handle_solve_nc(ctx, h, rhs, sp, rp, nc, merge=True, bcfn=None):
    levels, ops = build_ops(ctx, h, merge)
    if bcfn is not None:
        for op in ops:
            op.setNComp(nc)
            op.setBCValues(bcfn)
    offs = [o * nc for o in h.level_offsets()]
    phi = [LO.Field(L, nc) for L in levels]
    rh = [LO.Field(L, nc) for L in levels]
    for l in range(h.nlev):
        rh[l].upload(rhs[offs[l]:offs[l + 1]])
    amg = AMG.AMRMultiGrid()
    amg.define(ops, [AMG.domain_cells(h.level_domain(l)) for l in range(h.nlev)], AMG.BiCGStabSolver())
    pre, post, nb, numMG, iterMax, imin = [int(v) for v in sp]
    amg.setSolverParameters(pre, post, nb, numMG, iterMax, float(rp[0]), float(rp[1]), float(rp[2]))
    amg.m_imin = imin
    amg.solve(phi, rh, h.nlev - 1, 0, zeroPhi=True)
    out = np.concatenate([phi[l].download() for l in range(h.nlev)])
    res = (out, np.asarray(amg.m_history), amg.m_iterations, amg.m_exitStatus)
    close_all(phi + rh, ops, levels)
    return res


@pytest.mark.parametrize("case", ["amr32", "lshape", "amr32_bcfn"])
def test_handle_ncomp_solve_equals_library_and_oracle(gpu_ctx, case):
    h = lshape_amr() if case == "lshape" else small_amr()
    nc = 2
    rhs = comps_pack(h, [H.bump_rhs(h, 1301), -0.5 * H.bump_rhs(h, 1303)])
    bcfn = None
    if case.endswith("bcfn"):
        h.bc_type_lo = h.bc_type_hi = (1, 0, 1)

        def bcfn(pos, d, s):
            return (pos[0] - 2 * pos[1] + d, 0.25 * pos[2] - s)
    sp, rp = H.solver_params(iterMax=40)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, ncomp=nc)
    if bcfn:
        g.set_bc_values(bcfn)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    g.close()
    phi_h, hist_h, it_h, st_h = _handle_solve_nc(gpu_ctx, h, rhs, sp, rp, nc, merge=case != "lshape", bcfn=bcfn)
    assert it_h == info_g["iterations"] and st_h == info_g["exit_status"] and st_h & 1
    np.testing.assert_array_equal(hist_h, hist_g)
    np.testing.assert_array_equal(phi_h, phi_g)
    o = Oracle(h)
    o.set_ncomp(nc)
    if bcfn:
        o.set_bc_values(bcfn)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    np.testing.assert_array_equal(phi_h, phi_o)


def test_handle_ncomp_fields_transfer_dot_norm(gpu_ctx):
    h = small_amr()
    L = LO.Level(gpu_ctx, h.levels[0], h.level_domain(0))
    rng = np.random.default_rng(1305)
    nc = 3
    n = h.cells(0) * nc
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    fa, fb = LO.Field(L, nc).upload(a), LO.Field(L, nc).upload(b)
    np.testing.assert_array_equal(fa.download(), a)
    op = LO.AMRPoissonOp(L, h.dx(0), -1.0, 0.0, 1.0, ((1,) * 6, (0.0,) * 6))
    o = Oracle(h)
    o.set_ncomp(nc)
    assert op.dotProduct(fa, fb) == o.level_dot(0, a, b)
    for p in (0, 1, 2):
        assert op.norm(fa, p) == o.level_norm(0, a, p)
    op.incr(fa, fb, 0.5)
    np.testing.assert_array_equal(fa.download(), a + b * 0.5)
    f2 = LO.Field(L, 2)
    with pytest.raises(amrmg.AMRMGError, match="component"):
        op.incr(fa, f2, 1.0)
    o.close()
    for f in (fa, fb, f2):
        f.close()
    op.close()
    L.close()
