"""Pins the MLP oracle (oracle/mlp_oracle.py) on the CPU: gradients against finite differences, BF16
rounding against torch's bfloat16 conversion and hand-worked ties, Adam against its closed form, and
the features against a brute-force loop."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _net(nin, hidden, seed):
    rng = np.random.default_rng(seed)
    nw = hidden * nin + hidden * hidden + hidden
    w = rng.normal(size=nw) * 0.4
    b = rng.normal(size=2 * hidden + 1) * 0.1
    return w, b


@pytest.mark.parametrize("nin", [9, 27])
def test_gradients_match_finite_differences(nin):
    hidden = 8
    w, b = _net(nin, hidden, 3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(13, nin))
    t = rng.normal(size=13)
    loss, gw, gb = M.loss_and_grads(x, t, w, b, quantize=False)
    h = 1e-6

    def L(w_, b_):
        return M.loss_and_grads(x, t, w_, b_, quantize=False)[0]

    for k in rng.choice(len(w), 40, replace=False):
        e = np.zeros_like(w)
        e[k] = h
        fd = (L(w + e, b) - L(w - e, b)) / (2 * h)
        assert abs(fd - gw[k]) <= 1e-6 * max(1.0, abs(fd)), k
    for k in range(len(b)):
        e = np.zeros_like(b)
        e[k] = h
        fd = (L(w, b + e) - L(w, b - e)) / (2 * h)
        assert abs(fd - gb[k]) <= 1e-6 * max(1.0, abs(fd)), k
    assert np.isclose(loss, 0.5 * np.mean((M.forward(x, w, b, quantize=False)[0] - t) ** 2))


def test_bf16_rounding():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.uniform(-30, 30, 5000),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1.0, 3.0e38], np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16_round(x), ref)
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert M.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    assert M.bf16_round(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert np.isnan(M.bf16_round(np.float32(np.nan)))


def test_adam_closed_form():
    # first step of Adam moves every parameter by lr * sign(g) (up to eps)
    w = np.array([0.5, -1.0, 2.0])
    b = np.array([0.1])
    gw = np.array([3.0, -0.2, 1e-3])
    gb = np.array([-4.0])
    st = M.adam_state(3, 1)
    w1, b1, st = M.adam_step(w, b, gw, gb, st, lr=0.01)
    assert np.allclose(w1, w - 0.01 * np.sign(gw), atol=1e-7)
    assert np.allclose(b1, b + 0.01, atol=1e-7)
    # a second step with the same gradient moves by lr again (bias-corrected moments equal g, |g|)
    w2, _, _ = M.adam_step(w1, b1, gw, gb, st, lr=0.01)
    assert np.allclose(w2, w - 0.02 * np.sign(gw), atol=1e-6)


@pytest.mark.parametrize("shape", [(5, 4), (4, 3, 5)])
def test_features_bruteforce(shape):
    rng = np.random.default_rng(1)
    n = int(np.prod(shape))
    le = rng.normal(size=n)
    F = M.features(le, shape)
    dim = len(shape)
    nx, ny = shape[0], shape[1]
    nz = shape[2] if dim == 3 else 1
    for e in range(n):
        i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
        t = 0
        for dk in ((-1, 0, 1) if dim == 3 else (0,)):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        assert F[e, t] == le[ii + nx * (jj + ny * kk)] - le[e]
                    else:
                        assert F[e, t] == M.WALL
                    t += 1
    assert np.all(F[:, 13 if dim == 3 else 4] == 0.0)


def test_scaling_clip():
    out = np.array([-10.0, 0.0, 1.0, 10.0])
    assert np.allclose(M.scaling(np.ones(4), out), [1e-4, 1.0, 10.0, 1e4])
