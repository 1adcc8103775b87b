"""Heat-equation steppers (BackwardEuler, AMRTGA) on the HIP path against the
oracle's restatement of the same steps (oracle.Oracle.backward_euler_step /
tga_step), several steps in a row: the FP64 solutions are bitwise equal and every
solve takes the same number of iterations; mixed precision stays within 1e-9 of
the FP64 oracle (ten times the solve tolerance)."""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import lshape_amr, small_amr, small_single, uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "single32":
        return small_single(n=32, box=16)
    if name == "amr32":
        return small_amr()
    return lshape_amr()


def _run(gpu_ctx, h, scheme, vc, nsteps, dt, precision=0):
    sp, rp = H.solver_params()
    phi0 = H.bump_rhs(h, 1201)
    src = H.sinusoid_rhs(h, 1202)
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=precision)
    o = Oracle(h)
    if vc:
        # the steppers set beta = -mu dt; VCAMRPoissonOp2's operator is
        # alpha a - beta div(b grad), so d phi/dt = L phi diffuses with -b: pass b < 0
        a, b = H.vc_coefficients(h, seed=41)
        b = [-x for x in b]
        g.set_vc_coefs(a, b)
        o.vc_set(a, b)
    g.setSolverParameters(sp, rp)
    g.set_field(g.PHI, phi0)
    heat = amrmg.HeatSolver(g, scheme)
    heat.set_source(src)
    phi_o = phi0
    out = []
    for n in range(nsteps):
        ig = heat.step(dt)
        if scheme == "be":
            phi_o, _, io = o.backward_euler_step(phi_o, src, dt, sp, rp, zero_phi=False)
            io = (io, None)
        else:
            phi_o, (_, i1), (_, i2) = o.tga_step(phi_o, src, dt, sp, rp)
            io = (i1, i2)
        out.append((g.get_field(g.PHI), phi_o.copy(), ig, io))
    heat.close()
    g.close()
    return out


@pytest.mark.parametrize("scheme", ["be", "tga"])
@pytest.mark.parametrize("case,vc", [("single32", False), ("amr32", False), ("amr32", True), ("lshape", True)])
def test_heat_steps_bitwise(gpu_ctx, scheme, case, vc):
    print(f"heat {scheme} {case} vc={vc}", flush=True)
    h = _case(case)
    for n, (pg, po, ig, io) in enumerate(_run(gpu_ctx, h, scheme, vc, 3, 2e-3)):
        assert ig[0]["iterations"] == io[0]["iterations"], n
        assert ig[0]["exit_status"] == io[0]["exit_status"] and ig[0]["exit_status"] & 1
        if io[1] is not None:
            assert ig[1]["iterations"] == io[1]["iterations"], n
            assert ig[1]["exit_status"] == io[1]["exit_status"] and ig[1]["exit_status"] & 1
        np.testing.assert_array_equal(pg, po)


def test_heat_tga_mixed_precision(gpu_ctx):
    print("heat tga mixed", flush=True)
    h = _case("amr32")
    m = np.concatenate([uncovered_mask(h, lev) for lev in range(h.nlev)])
    for n, (pg, po, ig, io) in enumerate(_run(gpu_ctx, h, "tga", True, 2, 2e-3, precision=1)):
        err = np.abs(pg - po)[m].max() / np.abs(po)[m].max()
        print(f"step {n}: {ig}, err {err:.2e}")
        assert ig[0]["exit_status"] & 1 and ig[1]["exit_status"] & 1
        assert err < 1e-9, err
