"""Host-only construction of every device table (no device): ghost rules,
copiers, coarse-fine stencils and reflux lists build for the test hierarchies and
the full 256^3 three-level workload, and every address a kernel forms from them
is inside its field (amrmg_plan_check)."""
import pytest

import amrmg
import hierarchy as H
from helpers import lshape_qcfi, small_amr, small_single

CASES = {
    "single32": lambda: small_single(32, 16),
    "single32_negidx": lambda: small_single(32, 16, lo=-16),
    "helm24": lambda: small_single(24, 8, alpha=0.7, beta=1.3),
    "neumann16": lambda: small_single(16, 8, bc=(0, 0, 0)),
    "ragged40": lambda: small_single(40, 16),
    "amr32": lambda: small_amr(),
    "lshape": lambda: lshape_qcfi(),
    "c1_64": lambda: H.single_level(64, 32, 8),
}


@pytest.mark.parametrize("case", list(CASES))
def test_plan_check(case):
    h = CASES[case]()
    st = amrmg.plan_check(h)
    assert st["levels"] == h.nlev
    assert st["cells"] == h.total_cells()
    if h.nlev > 1:
        assert st["cf_coarse"] > 0 and st["cf_fine"] == 4 * st["cf_coarse"]
        assert st["reflux_cells"] > 0


def test_plan_check_full_workload():
    h = H.amr_hierarchy(256)
    st = amrmg.plan_check(h)
    assert st["mg_depths"] == 6
    assert st["cells"] == 113545216
    # CF records: 6 faces of the 160^3 and 200^3 coarse regions; four fine ghosts each
    assert st["cf_coarse"] == 6 * 160 ** 2 + 6 * 200 ** 2
    assert st["cf_fine"] == 4 * st["cf_coarse"]


def test_plan_check_rejects_unsorted_layout():
    h = small_single(32, 16)
    h.levels[0] = list(reversed(h.levels[0]))
    with pytest.raises(amrmg.AMRMGError):
        amrmg.plan_check(h)
