"""The CPU oracle of the learned scheduler's network (oracle/mlp.py), pinned
without the device: finite differences of the FP64 network, closed forms of
the forward pass and of Adam's first step, BF16 rounding against torch's
own conversion, and the feature map on constant occupancy grids."""
import numpy as np
import pytest

from oracle import mlp


def _data(seed, n=24):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, (n, mlp.FEAT))
    t = rng.normal(0.0, 1.0, n)
    w = rng.uniform(0.5, 1.5, n)
    return X, t, w


def test_parameter_count_matches_the_layout():
    assert mlp.NPARAMS == 160 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 14529


def test_fp64_gradient_matches_finite_differences():
    X, t, w = _data(1)
    p = mlp.init_params(2)
    _, g = mlp.loss_grad(p, X, t, w, emulate=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([
        rng.choice(np.arange(mlp.OFF_W1, mlp.OFF_B1), 120, replace=False),
        np.arange(mlp.OFF_B1, mlp.OFF_W2),
        rng.choice(np.arange(mlp.OFF_W2, mlp.OFF_B2), 120, replace=False),
        np.arange(mlp.OFF_B2, mlp.NPARAMS)])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        lp, _ = mlp.loss_grad(pp, X, t, w, emulate=False)
        lm, _ = mlp.loss_grad(pm, X, t, w, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1.0 + abs(g[i])), (i, fd, g[i])


def test_a_dropped_backward_term_fails_the_difference_check():
    # the check above would catch a missing relu mask on the first layer
    X, t, w = _data(4)
    p = mlp.init_params(5)
    f = mlp.forward(p, X, emulate=False)
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p)
    dy = w * (f["y"] - t) / w.sum()
    dz2 = np.where(f["z2"] > 0, dy[:, None] * w3[None, :], 0.0)
    wrong_gb1 = (dz2 @ W2.T).sum(axis=0)  # no [z1 > 0]
    _, g = mlp.loss_grad(p, X, t, w, emulate=False)
    assert np.max(np.abs(wrong_gb1 - g[mlp.OFF_B1:mlp.OFF_W2])) > 1e-6


def test_forward_closed_form():
    # one path through the network, every weight BF16-exact
    W1 = np.zeros((mlp.FEAT, mlp.HID)); W1[0, 0] = 2.0
    b1 = np.zeros(mlp.HID); b1[0] = -1.0
    W2 = np.zeros((mlp.HID, mlp.HID)); W2[0, 0] = 3.0
    b2 = np.zeros(mlp.HID); b2[0] = -0.5
    w3 = np.zeros(mlp.HID); w3[0] = 0.5
    p = mlp.pack(W1, b1, W2, b2, w3, 0.25)
    x0 = np.array([0.0, 0.25, 0.5, 0.75, 1.0, 2.0])
    X = np.zeros((len(x0), mlp.FEAT)); X[:, 0] = x0
    want = 0.5 * np.maximum(3.0 * np.maximum(2.0 * x0 - 1.0, 0.0) - 0.5, 0.0) + 0.25
    for em in (False, True):
        assert np.array_equal(mlp.forward(p, X, emulate=em)["y"], want)


def test_first_adam_step_is_lr_times_sign_of_gradient():
    X, t, w = _data(6)
    p = mlp.init_params(7)
    lr = 1e-3
    _, g = mlp.loss_grad(p, X, t, w)
    p1, _ = mlp.adam(p, X, t, w, 1, lr)
    np.testing.assert_allclose(p1 - p, -lr * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)


def test_adam_lowers_the_loss():
    X, t, w = _data(8, n=64)
    p = mlp.init_params(9)
    l0, _ = mlp.loss_grad(p, X, t, w)
    p2, _ = mlp.adam(p, X, t, w, 200, 3e-3)
    l2, _ = mlp.loss_grad(p2, X, t, w)
    assert l2 < 0.2 * l0


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    vals = np.concatenate([
        rng.normal(0, 1, 4000).astype(np.float32),
        rng.normal(0, 1e-30, 500).astype(np.float32),
        np.array([0.0, -0.0, 1.0, np.inf, -np.inf, 3.0e38, 1e-45], dtype=np.float32),
        # exact ties: 1 + 2^-8 (even below) and 1 + 3 * 2^-8 (even above)
        np.array([1.00390625, 1.01171875, -1.00390625], dtype=np.float32)])
    ours = mlp.bf16(vals)
    ref = torch.from_numpy(vals).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    assert np.isnan(mlp.bf16(np.float32(np.nan)))


def test_emulated_gradient_close_to_fp64():
    X, t, w = _data(11, n=256)
    p = mlp.init_params(12)
    _, ge = mlp.loss_grad(p, X, t, w, emulate=True)
    _, gf = mlp.loss_grad(p, X, t, w, emulate=False)
    # BF16 operands (8-bit mantissas) move the gradient by a few percent
    assert np.linalg.norm(ge - gf) <= 0.1 * np.linalg.norm(gf)
    assert ge @ gf >= 0.995 * np.linalg.norm(ge) * np.linalg.norm(gf)


def test_features_of_constant_grid():
    count = np.full((6, 5, 7), 7)
    f = mlp.tile_features(count, (7, 5, 6), (3, 2, 1), [0.5, 0.25, 1.0],
                          [0.1, 0.2, 0.3], 16, 16)
    assert np.all(f[:125] == 3.0 / 8.0)  # log2(8) / 8
    assert np.allclose(f[125:128], np.log2(np.array([0.5, 0.25, 1.0]) + 1 / 256) / 8,
                       rtol=0, atol=1e-15)
    assert list(f[128:131]) == [0.1, 0.2, 0.3]
    assert f[131] == 1.0 and f[132] == 0.5
    assert not f[133:].any()


def test_features_follow_dz_dy_dx_order_with_wrap():
    count = np.zeros((3, 3, 3), dtype=int)
    count[0, 1, 2] = 15  # z=0, y=1, x=2
    f = mlp.tile_features(count, (3, 3, 3), (0, 0, 0), [1.0] * 3, [0.5] * 3, 1, 1)
    hit = np.flatnonzero(f[:125])
    # offsets (dz, dy, dx) with (0 + dz) % 3 == 0, dy % 3 == 1, dx % 3 == 2
    want = [(dz + 2) * 25 + (dy + 2) * 5 + (dx + 2)
            for dz in range(-2, 3) for dy in range(-2, 3) for dx in range(-2, 3)
            if dz % 3 == 0 and dy % 3 == 1 and dx % 3 == 2]
    assert hit.tolist() == sorted(want)
    assert np.all(f[hit] == 0.5)
