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


def test_serve_every_result_matches_oracle(gpu_ctx):
    """Each request's result, taken in a mix of overlapping solves, equals the oracle's
    solve of that request bit for bit."""
    from oracle import Oracle
    reqs = _requests(8, seed=80)
    sp, rp = H.solver_params()
    e = amrmg.ServeEngine(0, slots=4, budget_bytes=32 << 30, sparams=sp, rparams=rp)
    ids = [e.submit(h, rhs) for h, rhs in reqs]
    res = [e.wait(i) for i in ids]
    e.close()
    for q, ((h, rhs), r) in enumerate(zip(reqs, res)):
        o = Oracle(h)
        phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
        o.close()
        assert r["iterations"] == info_o["iterations"] and r["exit_status"] == info_o["exit_status"], q
        np.testing.assert_array_equal(r["phi"], phi_o)


def test_serve_one_waiter_per_request(gpu_ctx):
    """A request is waited on once: of two threads waiting for the same id, one gets
    the result and the other an error."""
    import threading
    h, rhs = _requests(1, seed=81)[0]
    sp, rp = H.solver_params()
    e = amrmg.ServeEngine(0, slots=1, budget_bytes=16 << 30, sparams=sp, rparams=rp)
    rid = e.submit(h, rhs)
    out = []

    def waiter():
        try:
            out.append(("ok", e.wait(rid, want_phi=False)["exit_status"]))
        except amrmg.AMRMGError as x:
            out.append(("err", str(x)))
    ts = [threading.Thread(target=waiter) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    e.close()
    assert sorted(k for k, _ in out) == ["err", "ok"], out


def test_serve_lanes_keep_small_requests_moving(gpu_ctx):
    """With one lane for large requests (at least large_cells cells), a second large
    request is admitted only after the first one is done, while small requests that
    arrived after it are admitted meanwhile; every result equals the request solved
    alone, bit for bit."""
    # 256^3 large requests run for tens of ms, long after the small ones were submitted
    big = [r for r in H.serve_requests(40, seed=82, bases=(256,)) if r[0].nlev == 1][:2]
    small = [r for r in H.serve_requests(40, seed=83, bases=(32,))][:4]
    assert len(big) == 2 and len(small) == 4
    thresh = big[0][0].total_cells()  # 256^3: large; 32^3 requests stay small
    assert all(h.total_cells() < thresh for h, _ in small)
    sp, rp = H.solver_params()
    alone = []
    e1 = amrmg.ServeEngine(0, slots=1, budget_bytes=16 << 30, sparams=sp, rparams=rp)
    for h, rhs in big + small:
        alone.append(e1.wait(e1.submit(h, rhs)))
    e1.close()
    e = amrmg.ServeEngine(0, slots=3, budget_bytes=16 << 30, sparams=sp, rparams=rp, max_large=1,
                          large_cells=thresh)
    ids = [e.submit(h, rhs) for h, rhs in big + small]
    res = [e.wait(i) for i in ids]
    e.close()
    b0, b1 = res[0], res[1]
    assert b1["t_admit"] >= b0["t_done"]  # one large request at a time
    assert any(r["t_admit"] < b1["t_admit"] for r in res[2:])  # small ones went ahead of it
    for q, (a, m) in enumerate(zip(alone, res)):
        assert a["iterations"] == m["iterations"] and a["exit_status"] == m["exit_status"], q
        np.testing.assert_array_equal(a["phi"], m["phi"])


def test_serve_lane_arguments_checked(gpu_ctx):
    e = amrmg.ServeEngine(0, slots=2, budget_bytes=1 << 30)
    with pytest.raises(amrmg.AMRMGError):
        e.set_lanes(0)
    with pytest.raises(amrmg.AMRMGError):
        e.set_lanes(1, -5)
    e.close()
