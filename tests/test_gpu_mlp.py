"""The surrogate MLP kernels on the GPU against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import diskfft_amd
    if diskfft_amd.device_count() < 1:
        pytest.fail("no HIP device visible")
    import surrogate
    return surrogate


@pytest.fixture(scope="module")
def mo():
    import mlp_oracle
    return mlp_oracle


def test_mfma_tile_layout(sg, mo):
    print("case mfma probe", flush=True)
    rng = np.random.default_rng(1)
    A = mo.bf16(rng.normal(size=(16, 16)))
    B = mo.bf16(rng.normal(size=(16, 16)))
    D = sg.mfma_probe(A, B)
    assert np.max(np.abs(D - A @ B)) < 1e-5 * np.max(np.abs(A @ B))


def test_forward_matches_oracle(sg, mo):
    print("case mlp forward", flush=True)
    net = sg.MLP(members=3, seed=5)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (200, 4))
    y = net.forward(x)
    for m in range(3):
        w = net.get_params(m).astype(np.float64)
        yo = mo.forward_bf16(w, x)
        y64, _ = mo.forward(w, x)
        # BF16 activations: agreement to a few BF16 ulps with the rounded oracle
        assert np.max(np.abs(y[m] - yo)) / np.max(np.abs(yo)) < 2e-2
        assert np.max(np.abs(y[m] - y64)) / np.max(np.abs(y64)) < 5e-2


def test_gradient_matches_oracle(sg, mo):
    print("case mlp gradient", flush=True)
    net = sg.MLP(members=2, seed=9)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (100, 4))      # not a multiple of 32: ragged last workgroup
    t = rng.normal(0, 1, (100, 3))
    loss, g = net.grad(x, t)
    for m in range(2):
        w = net.get_params(m).astype(np.float64)
        lo, go = mo.loss_grad(w, x, t)
        assert abs(loss[m] - lo) / lo < 2e-2
        cos = float(g[m] @ go / (np.linalg.norm(g[m]) * np.linalg.norm(go)))
        assert cos > 0.999, cos
        assert np.linalg.norm(g[m] - go) / np.linalg.norm(go) < 5e-2
        for k, (o, s) in mo.OFF.items():
            n = int(np.prod(s))
            assert np.linalg.norm(g[m][o:o + n] - go[o:o + n]) <= 0.08 * np.linalg.norm(go[o:o + n]) + 1e-9, k


def test_training_fits_smooth_function(sg):
    print("case mlp training", flush=True)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (2048, 4))
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.cos(x[:, 3] + x[:, 0]),
                  x[:, 1] ** 2 - 0.5 * x[:, 2]], axis=1)
    net = sg.MLP(members=2, seed=11)
    hist = net.train(x, t, epochs=150, batch=256, lr=3e-3)
    assert hist[-1] < 0.02 * hist[0], (hist[0], hist[-1])
    y = net.forward(x)
    assert np.sqrt(np.mean((y[0] - t) ** 2)) < 0.05
