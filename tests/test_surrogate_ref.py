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
