"""The reference driver's inputs file (grids read in) solved on the GPU through
chombo_3.2_amd/inputs.run: bitwise the oracle's solve of the same problem."""
import os

import numpy as np
import pytest

import inputs as I
from oracle import Oracle

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "amrpoisson_execCell.inputs")


@pytest.mark.parametrize("argv", [("grids.read_in_grids = true",),
                                  ("grids.read_in_grids = true", "grids.max_box_size = 16", "main.iterations = 2")])
def test_inputs_run_matches_oracle(gpu_ctx, argv):
    pp = I.ParmParse.from_file(FIXTURE, argv)
    h, phi, hist, info = I.run(pp, gpu_ctx)
    sp, rp = I.solver_params(pp)
    phi_o, hist_o, info_o = Oracle(h).solve(I.gaussians_rhs(h), sp, rp)
    assert info["iterations"] == info_o["iterations"] and info["exit_status"] == info_o["exit_status"]
    assert info["exit_status"] & 1
    np.testing.assert_array_equal(hist, hist_o)
    np.testing.assert_array_equal(phi, phi_o)


def test_inputs_plotfile_roundtrip(gpu_ctx, tmp_path):
    """The driver's plotfile (phi, rhs, res) read back: phi and rhs as solved, res the
    composite residual, below the tolerance."""
    import chombo_hdf5 as CH

    pp = I.ParmParse.from_file(FIXTURE, ("grids.read_in_grids = true",))
    out = str(tmp_path / "poissonOut.3d.hdf5")
    h, phi, hist, info = I.run(pp, gpu_ctx, plotfile=out)
    f = CH.read_amr_hierarchy(out)
    assert f.names == ["phi", "rhs", "res"]
    np.testing.assert_array_equal(CH.field_from_file(f, 0), phi)
    np.testing.assert_array_equal(CH.field_from_file(f, 1), I.gaussians_rhs(h))
    res = CH.field_from_file(f, 2)
    assert np.abs(res).max() <= 1e-9 * hist[0]

