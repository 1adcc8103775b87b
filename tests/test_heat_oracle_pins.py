"""Pins of the oracle's heat-equation steppers (BackwardEuler, AMRTGA) over the
multigrid solver, d phi/dt = L phi + S:

* a discrete Dirichlet sine mode (eigenvalue lambda of the 7-point operator) is
  multiplied by the scheme's rational amplification factor R(z), z = dt lambda:
  1 / (1 - z) for backward Euler, (1 + mu3 z) / ((1 - mu1 z)(1 - mu2 z)) for TGA;
  a mode source S from phi = 0 gives dt S / (1 - z), resp.
  dt (1 + mu4 z) S / ((1 - mu1 z)(1 - mu2 z));
* the TGA constants themselves: the amplification agrees with exp(z) to second
  order (the error shrinks as z^3) and the source response with (exp(z) - 1) / z
  to second order, which fixes mu1 + mu2 + mu3 = 1, mu1 mu2 and mu4; backward
  Euler is first order only.
"""
import numpy as np
import pytest

import hierarchy as H
from helpers import small_single
from oracle import Oracle


def _mode(h, k):
    return h.pack(lambda x, y, z, l: np.sin(np.pi * k[0] * x) * np.sin(np.pi * k[1] * y) * np.sin(np.pi * k[2] * z))


def _lam(n, k):
    dx = 1.0 / n
    return sum((2 / dx ** 2) * (np.cos(np.pi * kk * dx) - 1) for kk in k)


def _r_tga(z):
    mu1, mu2, mu3, mu4 = Oracle.tga_mu()
    return (1 + mu3 * z) / ((1 - mu1 * z) * (1 - mu2 * z)), (1 + mu4 * z) / ((1 - mu1 * z) * (1 - mu2 * z))


@pytest.mark.parametrize("scheme", ["be", "tga"])
@pytest.mark.parametrize("dt", [1e-4, 3e-3])
def test_sine_mode_amplification(scheme, dt):
    n, k = 16, (1, 2, 1)
    h = small_single(n=n, box=8)
    o = Oracle(h)
    f = _mode(h, k)
    z = dt * _lam(n, k)
    sp, rp = H.solver_params(eps=1e-13)
    zero = np.zeros_like(f)
    if scheme == "be":
        phi, _, info = o.backward_euler_step(f, zero, dt, sp, rp, zero_phi=False)
        src, _, _ = o.backward_euler_step(zero, f, dt, sp, rp)
        amp, samp = 1 / (1 - z), dt / (1 - z)
    else:
        phi, _, _ = o.tga_step(f, zero, dt, sp, rp)
        src, _, _ = o.tga_step(zero, f, dt, sp, rp)
        r, rs = _r_tga(z)
        amp, samp = r, dt * rs
    assert np.abs(phi - amp * f).max() < 1e-10 * np.abs(f).max()
    assert np.abs(src - samp * f).max() < 1e-10 * dt * np.abs(f).max()


def test_tga_constants_are_second_order_and_l_stable():
    mu1, mu2, mu3, mu4 = Oracle.tga_mu()
    errs, serrs = [], []
    for z in (-1e-1, -1e-2, -1e-3):
        r, rs = _r_tga(z)
        errs.append(abs(r - np.exp(z)) / abs(z) ** 3)
        serrs.append(abs(rs - np.expm1(z) / z) / abs(z) ** 2)
    # error / z^3 (resp. / z^2) stays bounded: second order in the amplification
    # and in the source response
    assert max(errs) < 1.0 and max(serrs) < 1.0
    assert errs[2] < 2 * errs[1] and serrs[2] < 2 * serrs[1]
    # backward Euler is only first order: its error / z^2 tends to 1/2
    be = [abs(1 / (1 - z) - np.exp(z)) / z ** 2 for z in (-1e-2, -1e-3)]
    assert abs(be[1] - 0.5) < 0.01
    # stiff decay: |R(z)| < 1 for z < 0 and R(z) -> 0 as z -> -inf
    for z in (-1.0, -10.0, -1e3):
        assert abs(_r_tga(z)[0]) < 1.0
    assert abs(_r_tga(-1e8)[0]) < 1e-6
    assert abs(mu1 + mu2 + mu3 - 1.0) < 1e-15
