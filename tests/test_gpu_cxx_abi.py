"""The C ABI called from C++ (tests/cxx/abi_driver.cpp, built with the library): a
whole solve through amrmg_solver_* and the residual of its result through the
operator-level handles, wrapped the way an AMRLevelOp subclass would wrap them.
Everything it returns must be bitwise what the same calls return through the Python
binding, and the solve bitwise the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import amrmg
import hierarchy as H
from helpers import small_single
from oracle import Oracle

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cxx", "_build", "abi_driver")


def _run_driver(tmp_path, h, rhs, sp, rp):
    isp, rsp = h.spec()
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.asarray([isp.size], np.int32).tofile(f)
        isp.astype(np.int32).tofile(f)
        rsp.astype(np.float64).tofile(f)
        np.asarray(sp, np.int32).tofile(f)
        np.asarray(rp, np.float64).tofile(f)
        np.asarray([rhs.size], np.int64).tofile(f)
        rhs.astype(np.float64).tofile(f)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    info = np.frombuffer(raw[:12], np.int32)
    vals = np.frombuffer(raw[12:], np.float64)
    it, n = int(info[0]), rhs.size
    hist, phi, res, sc = vals[:it + 1], vals[it + 1:it + 1 + n], vals[it + 1 + n:it + 1 + 2 * n], vals[-3:]
    return info, hist, phi, res, sc


@pytest.mark.parametrize("case", ["lap32", "helm_neumann40"])
def test_cxx_driver_matches_python_binding_and_oracle(gpu_ctx, tmp_path, case):
    if not os.path.exists(DRIVER):
        raise RuntimeError("tests/cxx/_build/abi_driver is not built (make -C tests/cxx)")
    if case == "lap32":
        h = small_single(n=32, box=16)
    else:
        h = small_single(n=40, box=16, alpha=0.6, beta=1.4, bc=(0, 1, 1))
        h.bc_val_lo = (0.5, -1.0, 0.25)
    rhs = H.sinusoid_rhs(h, 2101)
    sp, rp = H.solver_params()
    info, hist, phi, res, (dot, n0, n2) = _run_driver(tmp_path, h, rhs, sp, rp)
    assert info[1] & 1
    # the same through the Python binding
    g = amrmg.AMRMultiGrid(gpu_ctx, h)
    g.setSolverParameters(sp, rp)
    g.set_field(g.RHS, rhs)
    hist_p, info_p = g.solve()
    phi_p = g.get_field(g.PHI)
    res_p, n0_p = g.computeAMRResidual()
    g.close()
    assert int(info[0]) == info_p["iterations"] and int(info[1]) == info_p["exit_status"]
    np.testing.assert_array_equal(hist, hist_p)
    np.testing.assert_array_equal(phi, phi_p)
    np.testing.assert_array_equal(res, res_p)
    assert n0 == n0_p
    # dot and 2-norm in the reference's order over the boxes (FArrayBox order each)
    o = Oracle(h)
    assert dot == o.level_dot(0, res, res)
    assert n2 == o.level_norm(0, res, 2)
    phi_o, hist_o, info_o = o.solve(rhs, sp, rp)
    o.close()
    np.testing.assert_array_equal(phi, phi_o)
    np.testing.assert_array_equal(hist, hist_o)
