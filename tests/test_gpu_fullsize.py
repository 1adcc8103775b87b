"""The benchmark configurations at their stated sizes against the oracle.

C2: 3-level AMR Poisson, base 256^3 (256^3 + 320^3 + 400^3 = 113.5 M cells, 64^3
boxes), Gaussian-bump rhs seed 1002, to 1e-10: the FP64 GPU solve has the oracle's
iterations, exit status and history, and bitwise its solution.
C3: the same with FP32 inner V-cycles: the composite solution within 1e-9 of the
FP64 oracle, same iteration count.
C5: serving requests with 128^3 and 256^3 bases (one and two levels): each result
the same alone and in the mix, bitwise, and a two-level request equal to the oracle.
"""
import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import uncovered_mask
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    print("C2 256^3: building the hierarchy, rhs and the oracle solve", flush=True)
    h = H.amr_hierarchy(256, fracs=(0.25, 0.06), max_box=64, block=8)
    rhs = H.bump_rhs(h, 1002)
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=100)
    o = Oracle(h)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    print(f"C2 oracle: {info_o}", flush=True)
    return h, rhs, sp, rp, phi_o, hist_o, info_o


def test_c2_fullsize_fp64_matches_oracle(gpu_ctx, c2):
    h, rhs, sp, rp, phi_o, hist_o, info_o = c2
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    g.close()
    assert info_g["iterations"] == info_o["iterations"]
    assert info_g["exit_status"] == info_o["exit_status"] and info_g["exit_status"] & 1
    np.testing.assert_allclose(hist_g, hist_o, rtol=1e-9, atol=0)
    err = np.abs(phi_g - phi_o).max() / np.abs(phi_o).max()
    assert err <= 1e-10, err
    np.testing.assert_array_equal(hist_g, hist_o)
    np.testing.assert_array_equal(phi_g, phi_o)


def test_c3_fullsize_mixed_within_1e9_of_oracle(gpu_ctx, c2):
    h, rhs, sp, rp, phi_o, hist_o, info_o = c2
    g = amrmg.AMRMultiGrid(gpu_ctx, h, precision=1)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_g, info_g = g.solve()
    phi_g = g.get_field(g.PHI)
    g.close()
    assert info_g["exit_status"] & 1 and hist_g[-1] <= 1e-10 * hist_g[0]
    assert info_g["iterations"] == info_o["iterations"]
    offs = h.level_offsets()
    mask = np.concatenate([uncovered_mask(h, l) for l in range(h.nlev)])
    assert mask.size == offs[-1]
    err = float(np.abs(phi_g - phi_o)[mask].max() / np.abs(phi_o)[mask].max())
    print(f"C3 composite error vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= 1e-9, err


def test_c5_serving_large_bases(gpu_ctx):
    reqs = H.serve_requests(6, seed=2031, bases=(128, 256))
    assert {h.domain[1][0] + 1 for h, _ in reqs} == {128, 256}
    assert any(h.nlev == 2 for h, _ in reqs)
    sp, rp = H.solver_params(eps=1e-10, hang=1e-15, iterMax=100)
    e1 = amrmg.ServeEngine(0, slots=1, budget_bytes=96 << 30, sparams=sp, rparams=rp)
    alone = [e1.wait(e1.submit(h, rhs)) for h, rhs in reqs]
    e1.close()
    e4 = amrmg.ServeEngine(0, slots=4, budget_bytes=96 << 30, sparams=sp, rparams=rp)
    ids = [e4.submit(h, rhs) for h, rhs in reqs]
    mix = [e4.wait(i) for i in ids]
    st = e4.stats()
    e4.close()
    assert st["max_running"] >= 2
    for q, (a, m) in enumerate(zip(alone, mix)):
        assert a["exit_status"] & 1, q
        assert a["iterations"] == m["iterations"] and a["exit_status"] == m["exit_status"], q
        np.testing.assert_array_equal(a["phi"], m["phi"])
    # a two-level request against the oracle (the smallest one, to bound the CPU time)
    q = min((q for q, (h, _) in enumerate(reqs) if h.nlev == 2), key=lambda q: reqs[q][0].total_cells())
    h, rhs = reqs[q]
    print(f"C5 oracle check: base {h.domain[1][0] + 1}^3, {h.nlev} levels, {h.total_cells()} cells", flush=True)
    o = Oracle(h)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    assert mix[q]["iterations"] == info_o["iterations"]
    err = np.abs(mix[q]["phi"] - phi_o).max() / np.abs(phi_o).max()
    assert err <= 1e-10, err
    np.testing.assert_array_equal(mix[q]["phi"], phi_o)
