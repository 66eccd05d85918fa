"""Surrogate network kernels (include/jms_nn.h) on the GPU:
  * forward outputs and every parameter gradient against the CPU
    restatement evaluated with the same bf16 rounding points (relative L2
    below 2e-3 per parameter block: only the FP32 summation order differs,
    which can flip a bf16 rounding now and then);
  * device input windows against a host evaluation of the same formula;
  * a small training run shows what the trained network gains over the
    method without it: its error on an unseen lattice against a constant
    predictor and against the untrained network, and the surrogate scan
    against the plain multislice it replaces."""
import numpy as np
import pytest

from jmultislicelib_amd import Engine
from jmultislicelib_amd import surrogate as S
from oracle import surrogate_ref as R

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mfma", [True, False])
def test_forward_and_gradients_vs_cpu(mfma):
    """Matrix-core convolutions (bf16 MFMA, the output gradient rounded to
    bf16) and the FMA kernels, each against the restatement with the same
    rounding points."""
    rng = np.random.default_rng(4)
    net = S.Network(max_batch=8, seed=3)
    net.set_kernels(mfma)
    p = net.params().astype(np.float64)
    x = (rng.standard_normal((6, 8, 32, 32)) * 0.5).astype(np.float32)
    xb = R.bf16_bits(x)
    t = rng.uniform(0.05, 0.5, (6, 3))
    sc = np.array([0.3, 0.1, 0.05])
    y = net.forward(xb)
    yr, _ = R.forward(p, R.bf16(x))
    ey = rel_l2(y, yr)
    loss, g = net.gradients(xb, t, sc)
    lr_, gr = R.loss_and_grads(p, R.bf16(x), t, sc, round_dy=mfma)
    print(f"forward rel L2 {ey:.2e}; loss {loss:.6e} vs {lr_:.6e}")
    assert ey < 2e-3 and abs(loss - lr_) <= 2e-3 * abs(lr_)
    o = 0
    for name, shape in R.layout():
        k = int(np.prod(shape))
        e = rel_l2(g[o:o + k], gr[o:o + k])
        print(f"  {name}: rel L2 {e:.2e}")
        assert e < 2e-3, name
        o += k
    net.close()


def test_windows_vs_host():
    rng = np.random.default_rng(5)
    lats = [S.random_lattice(rng) for _ in range(2)]
    cid = np.array([0, 1, 1, 0, 1], np.int32)
    px = rng.uniform(-2, 6, 5).astype(np.float32)
    py = rng.uniform(-2, 6, 5).astype(np.float32)
    smp = S.Samples(lats, cid, px, py, np.zeros((5, 3), np.float32))
    ws = smp.windows()
    xb = ws.windows()
    got = (xb.astype(np.uint32) << 16).view(np.float32)
    maps = [S.potential_maps(l).astype(np.float64) for l in lats]
    bw = S.bin_weights()
    m = S.MAP_M
    j = np.arange(32)
    for s in range(5):
        a = lats[cid[s]].a
        X = px[s] + (j[None, :] - 16) * S.STEP_ANG + 0 * j[:, None]
        Y = py[s] + (j[:, None] - 16) * S.STEP_ANG + 0 * j[None, :]
        u, v = X / a * m, Y / a * m
        u0, v0 = np.floor(u).astype(int) % m, np.floor(v).astype(int) % m
        fu, fv = u - np.floor(u), v - np.floor(v)
        u1, v1 = (u0 + 1) % m, (v0 + 1) % m
        V = [(1 - fv) * ((1 - fu) * mp[v0, u0] + fu * mp[v0, u1]) + fv * ((1 - fu) * mp[v1, u0] + fu * mp[v1, u1])
             for mp in maps[cid[s]]]
        ref = np.stack([(bw[c, 0] * V[0] + bw[c, 1] * V[1]) * S.WIN_SCALE for c in range(8)])
        assert np.abs(got[s] - ref).max() <= 2 ** -7 * np.abs(ref).max(), s
    ws.close()


def test_training_gain_and_audited_scan():
    rng = np.random.default_rng(11)
    kw = dict(n=256, ncell=4)
    train_lats = [S.random_lattice(rng) for _ in range(48)]
    eng = Engine()
    tr = S.simulate(train_lats, 96, seed=1, eng=eng, **kw)
    ens = S.Ensemble(nmembers=2, max_batch=128, seed=7)
    untrained = S.Ensemble(nmembers=2, max_batch=128, seed=7)
    hist = ens.train(tr, epochs=40, batch=128, lr=2e-3)
    untrained.scale = ens.scale
    print("loss first/last epoch:", hist[0], hist[-1])
    assert hist[-1].mean() < 0.5 * hist[0].mean()
    # unseen lattice
    test_lat = S.random_lattice(np.random.default_rng(99))
    te = S.simulate([test_lat], 192, seed=2, eng=eng, **kw)
    tw = te.windows()
    pred, spread = ens.predict(tw)
    pred0, _ = untrained.predict(tw)
    const = np.broadcast_to(tr.t.mean(axis=0), te.t.shape)
    err = lambda p, d: np.mean(np.abs(p[:, d] - te.t[:, d]) / te.t[:, d])
    for d in (1, 2):
        print(f"{S.DET_NAMES[d]}: trained {err(pred, d):.3f}, constant {err(const, d):.3f}, "
              f"untrained {err(pred0, d):.3f}")
        assert err(pred, d) < 0.7 * err(const, d) and err(pred, d) < 0.5 * err(pred0, d)
    # the surrogate scan: re-simulated positions carry the plain values
    eng.setup(S.lattice_setup(test_lat, **kw))
    out, idx, tm = S.surrogate_scan(ens, eng, test_lat, te.px, te.py, audit=0.1, spread_max=0.02, seed=3)
    assert 0 < len(idx) < len(te.px)
    assert np.allclose(out[idx], te.t[idx], rtol=1e-5, atol=1e-9)
    for d in (1, 2):
        assert err(out, d) <= err(pred, d) + 1e-9
    print(f"re-simulated {len(idx)} of {len(te.px)}; timings {tm}")
    tw.close()
    ens.close()
    untrained.close()
    eng.close()


def test_symmetry_images_give_equal_intensities():
    """The perovskite-like stack, the round probe and the annular detectors
    are invariant under 4mm about the A column at the origin, so the plain
    multislice gives the same intensities at the 8 images of a position (the
    premise of training on the images and averaging predictions over them)."""
    lat = S.random_lattice(np.random.default_rng(31))
    eng = Engine()
    eng.setup(S.lattice_setup(lat, n=256, ncell=4))
    x8, y8 = S.symmetric_images([0.83, 1.37, 0.2], [0.29, 0.61, 1.9], lat.a)
    eng.set_batch(24)
    t = S.plain_scan(eng, x8.ravel(), y8.ravel()).reshape(8, 3, 3).astype(np.float64)
    eng.close()
    spread = np.abs(t - t[0]).max(axis=0) / np.abs(t[0])
    print("max relative deviation over the images [position][detector]:\n", spread)
    assert spread.max() < 2e-4
    # and different positions do differ
    assert np.abs(t[0, 0] - t[0, 1]).max() > 1e-2 * np.abs(t[0, 0]).max()
