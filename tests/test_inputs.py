"""ParmParse-style inputs (chombo_3.2_amd/inputs.py) on the reference driver's own
inputs file (tests/golden/amrpoisson_execCell.inputs, a copy of
releasedExamples/AMRPoisson/execCell/inputs): keys, overrides, the hierarchy and the
solver parameters poissonSolve.cpp would build from it, and the oracle's solve of
that problem."""
import math
import os

import numpy as np
import pytest

import hierarchy as H
import inputs as I
from oracle import Oracle

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "amrpoisson_execCell.inputs")


def _pp(*argv):
    return I.ParmParse.from_file(FIXTURE, argv)


def test_parmparse_reads_the_reference_inputs():
    pp = _pp()
    assert pp.getarr("grids.num_cells", int, 3) == [32, 32, 32]
    assert pp.get("grids.max_level", int) == 2
    assert pp.getarr("grids.ref_ratio", int, 2) == [2, 2]
    assert pp.get("solver.tolerance", float) == 1.0e-9
    assert pp.get("grids.read_in_grids", bool) is False
    assert pp.getarr("grids.level_1_box_1_lo", int, 3) == [32, 32, 32]
    assert pp.get("main.iterations", int) == 1
    # commented-out definitions are not read; a missing key is an error, query a default
    assert pp.getarr("bc_lo", int, 3) == [0, 0, 0]
    with pytest.raises(KeyError):
        pp.get("solver.no_such_key")
    assert pp.query("solver.no_such_key", int, 7) == 7


def test_parmparse_overrides_and_errors():
    pp = _pp("grids.max_level=1", "solver.num_smooth = 3")
    assert pp.get("grids.max_level", int) == 1
    assert pp.get("solver.num_smooth", int) == 3
    pp = I.ParmParse("a.b = 1 2 3\na.b = 4  # later wins\nflag = true\n")
    assert pp.getarr("a.b", int, 1) == [4]
    assert pp.get("flag", bool) is True
    with pytest.raises(ValueError):
        pp.getarr("a.b", int, 2)
    with pytest.raises(ValueError):
        I.ParmParse("no equals sign here")


def test_setup_grids_as_the_driver():
    with pytest.raises(ValueError, match="read_in_grids"):
        I.setup_grids(_pp())
    h = I.setup_grids(_pp("grids.read_in_grids = true"))
    assert h.domain == ((0, 0, 0), (31, 31, 31))
    assert h.levels[0] == [((0, 0, 0), (31, 31, 31))]          # max_box_size 10000: one box
    assert h.levels[1] == [((8, 8, 8), (23, 23, 23)), ((32, 32, 32), (47, 47, 47))]
    assert h.levels[2] == [((24, 24, 24), (39, 39, 39))]
    assert h.dx0 == 1.0 / 32 and h.alpha == 0.0 and h.beta == 1.0
    assert h.bc_type_lo == (1, 1, 1) and h.bc_type_hi == (1, 1, 1)
    assert h.bc_val_lo == (0.0, 0.0, 0.0)
    h2 = I.setup_grids(_pp("grids.read_in_grids = true", "grids.is_periodic = 1 0 1", "grids.max_box_size = 16"))
    assert h2.bc_type_lo == (3, 1, 3) and len(h2.levels[0]) == 8
    with pytest.raises(ValueError, match="outside"):
        I.setup_grids(_pp("grids.read_in_grids = true", "grids.level_2_box_0_hi = 39 39 200"))


def test_solver_params_as_the_driver():
    sp, rp = I.solver_params(_pp())
    # num_smooth 4 for all three, then num_pre / num_post 2 and num_bottom 1
    assert list(sp) == [2, 2, 1, 1, 100, 5]
    assert list(rp) == [1.0e-9, 1.0e-10, 1.0e-30]


def test_gaussians_rhs_and_oracle_solve():
    h = I.setup_grids(_pp("grids.read_in_grids = true"))
    rhs = I.gaussians_rhs(h)
    assert rhs.size == h.total_cells()
    # one cell evaluated by hand: level 2, box 0, cell (30, 31, 32)
    dx = h.dx(2)
    loc = [30 * dx + 0.5 * dx, 31 * dx + 0.5 * dx, 32 * dx + 0.5 * dx]
    want = 0.0
    for s, w, c in I.GAUSSIANS:
        r2 = (loc[0] - c[0]) ** 2 + (loc[1] - c[1]) ** 2 + (loc[2] - c[2]) ** 2
        want += s * math.exp(-r2 / w)
    off = h.level_offsets()[2] + (30 - 24) + 16 * ((31 - 24) + 16 * (32 - 24))
    assert abs(rhs[off] - want) <= 1e-15 * abs(want)
    sp, rp = I.solver_params(_pp())
    phi, hist, info = Oracle(h).solve(rhs, sp, rp)
    assert info["exit_status"] & 1 and hist[-1] <= 1e-9 * hist[0]
