"""Serving loop (amrmg_serve_*): continuous admission, overlapping solves on
several HIP streams, device-resident hierarchies reused across requests.

Each request is an ordinary AMRMultiGrid solve on its own stream, so its result
must not depend on what else runs: the test compares every result in the mix
with the same request run alone, bitwise (the stated tolerance, 1e-10 relative,
is looser)."""
import numpy as np
import pytest

import amrmg
import hierarchy as H

pytestmark = pytest.mark.gpu


def _requests(count=12, seed=77):
    return H.serve_requests(count, seed=seed, bases=(32, 64))


def test_serve_results_same_alone_and_in_the_mix(gpu_ctx):
    reqs = _requests()
    sp, rp = H.solver_params()
    alone = []
    e1 = amrmg.ServeEngine(0, slots=1, budget_bytes=32 << 30, sparams=sp, rparams=rp)
    for h, rhs in reqs:
        alone.append(e1.wait(e1.submit(h, rhs)))
    e1.close()
    e4 = amrmg.ServeEngine(0, slots=4, budget_bytes=32 << 30, sparams=sp, rparams=rp)
    ids = [e4.submit(h, rhs) for h, rhs in reqs]
    mix = [e4.wait(i) for i in ids]
    st = e4.stats()
    e4.close()
    assert st["max_running"] >= 2  # solves overlapped
    for q, (a, m) in enumerate(zip(alone, mix)):
        assert a["iterations"] == m["iterations"] and a["exit_status"] == m["exit_status"], q
        np.testing.assert_array_equal(a["phi"], m["phi"])
    # and the same as a direct solve
    h, rhs = reqs[0]
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    g.solve()
    np.testing.assert_array_equal(g.get_field(g.PHI), alone[0]["phi"])
    g.close()


def test_serve_admission_respects_hbm_budget(gpu_ctx):
    reqs = _requests(8, seed=78)
    sp, rp = H.solver_params()
    big = max(h.total_cells() for h, _ in reqs)
    budget = int(1.1 * 8 * 10 * big * 1.6)  # room for about one of the largest requests at a time
    e = amrmg.ServeEngine(0, slots=4, budget_bytes=budget, sparams=sp, rparams=rp)
    ids = [e.submit(h, rhs) for h, rhs in reqs]
    res = [e.wait(i) for i in ids]
    st = e.stats()
    e.close()
    assert st["max_bytes"] <= budget
    for r in res:
        assert r["exit_status"] & 1
        assert r["t_arrival"] <= r["t_admit"] <= r["t_done"]


def test_serve_reuses_resident_hierarchies(gpu_ctx):
    h, rhs = _requests(1, seed=79)[0]
    sp, rp = H.solver_params()
    e = amrmg.ServeEngine(0, slots=2, budget_bytes=16 << 30, sparams=sp, rparams=rp)
    r1 = e.wait(e.submit(h, rhs))
    r2 = e.wait(e.submit(h, 2 * rhs))
    st = e.stats()
    e.close()
    assert st["created"] == 1 and st["reused"] == 1
    np.testing.assert_array_equal(r2["phi"], 2 * r1["phi"])
