"""The CPU restatement of the surrogate network, pinned by finite
differences (float64, no rounding) and by its parameter count."""
import numpy as np

from oracle import surrogate_ref as R


def test_param_count_matches_library():
    import jmultislicelib_amd as J
    assert R.num_params() == J.lib().jms_nn_num_params() == 63107


def test_gradients_vs_finite_differences():
    rng = np.random.default_rng(0)
    p = np.concatenate([rng.standard_normal(int(np.prod(s))) * (0.3 if name.endswith(".w") else 0.05)
                        for name, s in R.layout()])
    x = rng.standard_normal((2, 8, 32, 32))
    t = rng.standard_normal((2, 3))
    sc = np.array([1.0, 0.5, 2.0])
    _, g = R.loss_and_grads(p, x, t, sc, quantize=False)
    idx = rng.choice(len(p), 40, replace=False)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in R.layout()])
    idx = np.concatenate([idx, offs[:-1]])  # first element of every block
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = R.loss_and_grads(pp, x, t, sc, quantize=False)[0]
        lm = R.loss_and_grads(pm, x, t, sc, quantize=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 0.0], np.float32)
    b = R.bf16(a)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0 + 2 ** -7  # ties to even
    assert abs(b[3] + 3.140625) < 1e-12 and b[4] == 0.0


def test_symmetry_images_form_the_4mm_orbit():
    """The 8 operations are a group (closed under composition, identity
    first) and map a general position to 8 distinct points of the cell."""
    from jmultislicelib_amd import surrogate as S
    ops = [np.array(m, float).reshape(2, 2) for m in S.SYM_4MM]
    assert np.array_equal(ops[0], np.eye(2))
    for A in ops:
        for B in ops:
            assert any(np.array_equal(A @ B, C) for C in ops)
    a = 3.9
    x8, y8 = S.symmetric_images([0.7], [0.31], a)
    pts = {(round(float(x), 5), round(float(y), 5)) for x, y in zip(x8[:, 0], y8[:, 0])}
    assert len(pts) == 8 and all(0 <= x < a and 0 <= y < a for x, y in pts)
    assert abs(x8[0, 0] - 0.7) < 1e-6 and abs(y8[0, 0] - 0.31) < 1e-6
    # every image's orbit is the same set
    for k in range(8):
        xk, yk = S.symmetric_images(x8[k], y8[k], a)
        assert {(round(float(x), 4), round(float(y), 4)) for x, y in zip(xk[:, 0], yk[:, 0])} == \
            {(round(x, 4), round(y, 4)) for x, y in pts}


def test_calibration_recovers_a_smooth_cell_factor():
    """A multiplicative error that is a 4mm-symmetric periodic function of
    the position in the cell (inside the basis) is removed exactly; with no
    error the factors are 1; the basis is invariant under the 8 images."""
    from jmultislicelib_amd import surrogate as S
    rng = np.random.default_rng(3)
    u, v = rng.uniform(0, 1, 400), rng.uniform(0, 1, 400)
    ua, va = rng.uniform(0, 1, 1000), rng.uniform(0, 1, 1000)
    truth = np.exp(rng.standard_normal((1400, 3)) * 0.3)
    err = lambda uu, vv: np.stack([0.02 * np.cos(2 * np.pi * uu) * np.cos(2 * np.pi * 2 * vv) +
                                   0.02 * np.cos(2 * np.pi * 2 * uu) * np.cos(2 * np.pi * vv),
                                   0.01 + 0.0 * uu, -0.03 * np.cos(2 * np.pi * 3 * uu) - 0.03 * np.cos(2 * np.pi * 3 * vv)], 1)
    pred_fit = truth[:400] * np.exp(-err(u, v))
    pred_all = truth[400:] * np.exp(-err(ua, va))
    fac = S.calibrate(pred_fit, truth[:400], u, v, ua, va, kmax=4, ridge=0.0)
    assert np.abs(pred_all * fac / truth[400:] - 1).max() < 1e-9
    one = S.calibrate(truth[:400], truth[:400], u, v, ua, va, kmax=4)
    assert np.abs(one - 1).max() < 1e-12
    x8, y8 = S.symmetric_images(u * 3.9, v * 3.9, 3.9)
    B0 = S._cell_basis(u, v, 5)
    for k in range(8):
        # the images are float32 positions: invariant to their rounding
        assert np.abs(S._cell_basis(x8[k] / 3.9, y8[k] / 3.9, 5) - B0).max() < 2e-5
