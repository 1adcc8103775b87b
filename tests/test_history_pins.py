"""Pins of the oracle's outer iteration (AMRMultiGrid::solveNoInitResid,
AMRMultiGrid.H:935-1124) beyond its converged solution:

* every history entry k is the max norm, over cells not covered by a finer level,
  of rhs - L(phi_k) for the k-th iterate phi_k (computeAMRResidual with
  zeroCovered and localMaxNorm, AMRMultiGrid.H:683-736): on one level the residual
  comes from an independently assembled scipy matrix, on three levels from the
  composite residual with the covered cells masked in numpy; the iterate phi_k is
  the solve stopped by iterMax = k, and its history is the prefix of the full one;
* the exit status bits (1 eps reached, 2 iterMax, 4 hang, 8 normThresh;
  AMRMultiGrid.H:1003-1007, 1088-1092) follow from the recorded history under the
  reference's tests, each forced on its own.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import small_amr, small_single, uncovered_mask
from oracle import Oracle
from test_oracle_pins import _assemble_dirichlet


def _iterates(o, rhs, kmax, **kw):
    out = []
    for k in range(kmax + 1):
        sp, rp = H.solver_params(iterMax=k, **kw)
        phi, hist, info = o.solve(rhs, sp, rp)
        out.append((phi, hist, info))
    return out


def test_history_is_residual_norm_of_each_iterate_single_level():
    n = 16
    h = small_single(n=n, box=8, alpha=0.3, beta=1.0)
    o = Oracle(h)
    A = _assemble_dirichlet(n, 0.3, 1.0)
    rhs = H.sinusoid_rhs(h, 1401)
    sp, rp = H.solver_params()
    _, full, info = o.solve(rhs, sp, rp)
    assert info["exit_status"] == 1
    b = h.to_global(rhs, 0).ravel(order="F")
    for k, (phi, hist, ik) in enumerate(_iterates(o, rhs, 5)):
        assert ik["iterations"] == k
        np.testing.assert_array_equal(hist, full[:k + 1])
        r = b - A @ h.to_global(phi, 0).ravel(order="F")
        assert abs(np.abs(r).max() - full[k]) <= 1e-9 * full[k]


def test_history_is_uncovered_residual_norm_amr():
    h = small_amr()
    o = Oracle(h)
    rhs = H.bump_rhs(h, 1402)
    sp, rp = H.solver_params()
    _, full, info = o.solve(rhs, sp, rp)
    assert info["exit_status"] == 1
    mask = np.concatenate([uncovered_mask(h, lev) for lev in range(h.nlev)])
    for k, (phi, hist, ik) in enumerate(_iterates(o, rhs, 3)):
        np.testing.assert_array_equal(hist, full[:k + 1])
        res, _ = o.amr_residual(phi, rhs)
        assert np.abs(res[mask]).max() == full[k]
        # zeroCovered clears the covered cells, whose own residual is not small
        assert np.abs(res[~mask]).max() == 0
        own = rhs - o.amr_operator(phi)
        assert np.abs(own[~mask]).max() > 1e-3 * full[k]


def _status_from_history(hist, eps, iter_max, imin, hang, thresh):
    """The reference's loop tests applied to a recorded history."""
    it = len(hist) - 1
    r, r0 = hist[-1], hist[0]
    last = hist[-2] if it > 0 else 2 * r0
    go_norm, go_redu, go_iter = r > thresh, r > eps * r0, it < iter_max
    go_hang = it < imin or r < (1 - hang) * last
    return (not go_redu) + 2 * (not go_iter) + 4 * (not go_hang) + 8 * (not go_norm)


@pytest.mark.parametrize("kw,want", [
    (dict(), 1),                                   # eps reached
    (dict(iterMax=3), 2),                          # iteration cap
    (dict(hang=0.95, imin=2), 4),                  # too little reduction per cycle
    (dict(normThresh=1e3), 8),                     # initial norm already below threshold
    (dict(eps=0.5, iterMax=1), 3),                 # two tests fail on the same iteration
])
def test_exit_status_follows_from_history(kw, want):
    h = small_single(n=16, box=8)
    o = Oracle(h)
    rhs = H.sinusoid_rhs(h, 1403)
    p = dict(H.DEFAULT_SOLVER)
    p.update(kw)
    sp, rp = H.solver_params(**kw)
    _, hist, info = o.solve(rhs, sp, rp)
    assert info["exit_status"] == want
    assert _status_from_history(hist, p["eps"], p["iterMax"], p["imin"], p["hang"], p["normThresh"]) == want
    # every earlier iteration passed all four tests
    for k in range(1, len(hist) - 1):
        assert _status_from_history(hist[:k + 1], p["eps"], p["iterMax"], p["imin"], p["hang"],
                                    p["normThresh"]) == 0
