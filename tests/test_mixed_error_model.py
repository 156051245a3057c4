"""Error budget of the mixed form on the Kelvin-Helmholtz box (CPU, oracle only).

The mixed form evaluates reconstruction and Riemann fluxes in FP32 (the north star's
precision split).  Its per-field relative L1 against FP64 meets 1e-5 on density, m_x and
energy with two orders of margin, but not on m_z: the KH box's m_z is a field about 5e-6 the
size of the momentum, driven only by dp/dz of the seeded noise, while FP32 rounding of the
O(1) energy and momentum fluxes leaves pressure noise whose z-gradient drives m_z at the
1e-2 level.

orc_set_flux_rounding_model perturbs every face flux of the FP64 oracle by a deterministic
+-1 ulp-of-FP32 relative error (u * 2 * 2^-23, u in [-1/2, 1/2]):
  * mode 1, the error on the flux's own magnitude, reproduces the GPU's measured mixed-form
    L1 (KH 32^3 x 20 steps on the GPU: m_y 1.3e-5, m_z 1.0e-2);
  * mode 2, the error on the flux less the physical flux of the cell left of the face (a
    form that evaluates only a deviation from an FP64 reference flux in FP32), still leaves
    m_z above 1e-5 once the box and the step count grow (64^3 x 50 below), because the
    pressure noise accumulates and m_z integrates its gradient.
So no FP32-flux form reaches 1e-5 on this field at the configs' sizes; the test pins the
model and the numbers the design notes quote.
"""
import numpy as np
import pytest

import cholla_amd as chl
from helpers import rel_l1
from oracle import oracle as orc

G = 5.0 / 3.0


def kh_run(n, steps, scale, mode, seed=7):
    ng = 4
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_kelvin_helmholtz(n, ng, G, seed=seed)
    orc.lib().orc_set_flux_rounding_model(scale, mode)
    try:
        orc.run(U, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, steps)
    finally:
        orc.lib().orc_set_flux_rounding_model(0.0, 0)
    return U.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]


def l1s(a, b):
    return [rel_l1(a[f], b[f]) for f in range(5)]


def test_model_off_is_the_oracle():
    a = kh_run(16, 4, 0.0, 1)
    b = kh_run(16, 4, 0.0, 0)
    np.testing.assert_array_equal(a, b)
    c = kh_run(16, 4, 2.0, 1)
    assert not np.array_equal(a, c)  # the model is live when on


def test_fp32_flux_rounding_reproduces_the_measured_mixed_error():
    ref = kh_run(32, 20, 0.0, 0)
    l1 = l1s(kh_run(32, 20, 2.0, 1), ref)
    # density, m_x and energy far inside the bar; m_y at the bar; m_z three orders above it
    assert max(l1[0], l1[1], l1[4]) < 1e-6
    assert 5e-6 < l1[2] < 5e-5
    assert 3e-3 < l1[3] < 5e-2


def test_deviation_form_rounding_still_misses_m_z_at_64_cubed():
    ref = kh_run(64, 50, 0.0, 0)
    l1 = l1s(kh_run(64, 50, 2.0, 2), ref)
    assert max(l1[0], l1[1], l1[2], l1[4]) < 1e-6
    assert l1[3] > 2e-5
