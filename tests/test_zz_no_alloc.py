"""amrmg_field_dot_noalloc: the dot product with its partial sums in a fixed device
array (no device allocation) gives bitwise amrmg_field_dot's result, on two calls
in a row (the second on a smaller level), and rejects levels with more boxes than
that array holds."""
import ctypes as C

import numpy as np
import pytest

import hierarchy as H
import levelops as LO

pytestmark = pytest.mark.gpu

MAX_BOXES = 8192


def _level(ctx, n, box, merge=True):
    dom = H.box((0, 0, 0), (n - 1,) * 3)
    boxes = sorted(H.domain_split(dom, box, 1), key=H.sort_key)
    return LO.Level(ctx, boxes, (dom[0], dom[1]), merge=merge)


def _dots(level, seed):
    rng = np.random.default_rng(seed)
    a = LO.Field(level).upload(rng.standard_normal(level.cells()))
    b = LO.Field(level).upload(rng.standard_normal(level.cells()))
    lib = LO._lib()
    ref, new = C.c_double(), C.c_double()
    assert lib.amrmg_field_dot(a.p, b.p, C.byref(ref)) == 0
    rc = lib.amrmg_field_dot_noalloc(a.p, b.p, C.byref(new))
    a.close()
    b.close()
    return rc, ref.value, new.value


@pytest.mark.parametrize("merge", [True, False])
def test_dot_noalloc_equals_dot_on_two_calls(gpu_ctx, merge):
    big = _level(gpu_ctx, 64, 16, merge)      # 64 boxes
    small = _level(gpu_ctx, 24, 8, merge)     # 27 boxes, fewer cells
    for level, seed in ((big, 1), (small, 2)):
        rc, ref, new = _dots(level, seed)
        assert rc == 0
        assert new == ref  # bitwise: the same kernel and order, only the partials' home differs
    big.close()
    small.close()


def test_dot_noalloc_rejects_too_many_boxes(gpu_ctx):
    n = 21  # 21^3 = 9261 one-cell boxes
    assert n ** 3 > MAX_BOXES
    level = _level(gpu_ctx, n, 1, merge=False)
    rc, ref, _ = _dots(level, 3)
    assert rc != 0
    assert b"at most 8192" in LO._lib().amrmg_last_error()
    level.close()
    # the array still serves the next call
    ok = _level(gpu_ctx, 16, 8)
    rc, ref, new = _dots(ok, 4)
    assert rc == 0 and new == ref
    ok.close()
