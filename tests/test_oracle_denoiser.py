"""The FP64 denoiser oracle against an independent library convolution and against central
finite differences (layer and whole network)."""
import numpy as np
import pytest

import oracle
from scipy.signal import correlate

from oracle import denoiser as dn


def test_conv3x3_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 7, 9, 5))
    w = rng.normal(size=(9, 4, 5))
    b = rng.normal(size=4)
    out = dn.conv3x3(x, w, b)
    for n in range(2):
        for co in range(4):
            k = w[:, co, :].reshape(3, 3, 5)            # [dy+1, dx+1, ci]
            ref = correlate(x[n], k, mode="same") + b[co]
            np.testing.assert_allclose(out[n, :, :, co], ref[:, :, 2] if ref.ndim == 3 else ref, rtol=1e-12, atol=1e-12)


def test_conv3x3_backward_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 5, 6, 3))
    w = rng.normal(size=(9, 2, 3))
    b = rng.normal(size=2)
    g = rng.normal(size=(1, 5, 6, 2))
    gx, gw, gb = dn.conv3x3_backward(x, w, g)
    f = lambda xx, ww, bb: float((dn.conv3x3(xx, ww, bb) * g).sum())
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (0, 2, 3, 1), (0, 4, 5, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps; xm[idx] -= eps
        assert gx[idx] == pytest.approx((f(xp, w, b) - f(xm, w, b)) / (2 * eps), rel=1e-6, abs=1e-8)
    for idx in [(0, 0, 0), (4, 1, 2), (8, 0, 1)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps; wm[idx] -= eps
        assert gw[idx] == pytest.approx((f(x, wp, b) - f(x, wm, b)) / (2 * eps), rel=1e-6, abs=1e-8)
    bp = b.copy(); bp[1] += eps
    bm = b.copy(); bm[1] -= eps
    assert gb[1] == pytest.approx((f(x, w, bp) - f(x, w, bm)) / (2 * eps), rel=1e-6)


def test_pool_upsample_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 4, 6, 3))
    g = rng.normal(size=(1, 2, 3, 3))
    assert (dn.pool(x) * g).sum() == pytest.approx((x * dn.pool_backward(g)).sum())
    y = rng.normal(size=(1, 2, 3, 3))
    G = rng.normal(size=(1, 4, 6, 3))
    assert (dn.upsample(y) * G).sum() == pytest.approx((y * dn.upsample_backward(G)).sum())


def test_display_value_closed_form_and_slope():
    """D(1) is the byte-208 case of Tools/verify_native.py:126 before rounding; the slope
    matches central differences on both sides of the sRGB knee."""
    v, _ = dn.display_value(np.array([1.0]))
    assert int(255 * min(v[0], 1) + 0.5) == 208
    assert oracle.display(1.0, 1.0) == 208
    for L in (1e-4, 2e-3, 0.05, 0.7, 3.0):
        eps = 1e-7 * max(L, 1e-3)
        fd = (dn.display_value(np.array([L + eps]))[0] - dn.display_value(np.array([L - eps]))[0]) / (2 * eps)
        assert dn.display_value(np.array([L]))[1][0] == pytest.approx(fd[0], rel=1e-5)


@pytest.mark.parametrize("display_weight", [0.0, 2.0])
def test_network_gradient_finite_differences(display_weight):
    rng = np.random.default_rng(3)
    params = dn.init_params(4).astype(np.float64)
    raw = np.abs(rng.normal(size=(1, 16, 16, 10)))
    target = np.abs(rng.normal(size=(1, 16, 16, 3))) * 5 + 3     # far from the prediction: stable signs
    loss, grads = dn.loss_and_grads(params, raw, target, display_weight)
    lay, _ = dn.layout()
    eps = 1e-5
    checked = 0
    for name, cin, cout, ci, co, relu, w_off, b_off in lay:
        for k in (w_off + (co * ci) * 4 + 1 * ci + 2, b_off + 1):   # centre tap (co 1, ci 2) and bias 1
            if k >= b_off + min(cout, 2) or (k < b_off and cin <= 2):
                continue
            pp, pm = params.copy(), params.copy()
            pp[k] += eps; pm[k] -= eps
            fd = (dn.loss_and_grads(pp, raw, target, display_weight)[0]
                  - dn.loss_and_grads(pm, raw, target, display_weight)[0]) / (2 * eps)
            assert grads[k] == pytest.approx(fd, rel=1e-4, abs=1e-9), (name, k)
            checked += 1
    assert checked >= 30
