"""HIP kernels of the learned-autogrid network against the float64 CPU
oracle: forward, loss, gradient (backward + weight reduction) and Adam."""
import numpy as np
import pytest

from inca_parallel_amd import capi, synth
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 16)).astype(np.float32)
    t = np.stack([rng.normal(scale=0.3, size=n), rng.integers(0, 4, n),
                  rng.integers(0, 4, n)], axis=1).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    net = capi.MLP(seed=7)
    try:
        x, _ = _data(1000, 1)
        y = net.forward(x)
        ref, _ = M.forward(net.params().astype(np.float64), x.astype(np.float64))
        err = np.abs(y - ref).max() / np.abs(ref).max()
        print("forward max rel err %.2e" % err)
        assert err < 2e-6
    finally:
        net.close()


def test_gradient_matches_oracle():
    net = capi.MLP(seed=8)
    try:
        for n in (1, 37, 700):
            x, t = _data(n, 10 + n)
            loss, g = net.loss_grad(x, t)
            rl, rg = M.loss_grad(net.params().astype(np.float64), x.astype(np.float64),
                                 t.astype(np.float64))
            eg = np.abs(g - rg).max() / np.abs(rg).max()
            print("n=%d loss %.6f (oracle %.6f), grad max rel err %.2e" % (n, loss, rl, eg))
            assert abs(loss - rl) <= 1e-5 * abs(rl)
            assert eg < 1e-4
    finally:
        net.close()


def test_adam_steps_match_oracle():
    net = capi.MLP(seed=9)
    try:
        x, t = _data(300, 5)
        p = net.params().astype(np.float64)
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        for step in range(1, 4):
            _, g = net.loss_grad(x, t)
            net.adam_step(1e-3)
            p, m, v = M.adam(p, g.astype(np.float64), m, v, step, 1e-3)
            assert np.abs(net.params() - p).max() < 1e-6
    finally:
        net.close()


def test_training_reduces_loss():
    net = capi.MLP(seed=3)
    try:
        x, _ = _data(512, 6)
        # learnable targets: a smooth function of the inputs
        t = np.stack([0.3 * np.tanh(x[:, 0] - x[:, 1]), (x[:, 2] > 0) * 1 + (x[:, 3] > 0) * 2,
                      np.clip(np.floor(x[:, 4] + 2), 0, 3)], axis=1).astype(np.float32)
        l0, _ = net.loss_grad(x, t, want_grad=False)
        for _ in range(300):
            net.loss_grad(x, t)
            net.adam_step(3e-3)
        l1, _ = net.loss_grad(x, t, want_grad=False)
        print("loss %.4f -> %.4f" % (l0, l1))
        assert l1 < 0.3 * l0
    finally:
        net.close()


def test_learned_autogrid_uses_predicted_radii():
    s = synth.config_cluster(seed=31, natoms=5, nheavy=3, nocc=3,
                             shells_heavy={0: 2, 1: 1}, shells_h={0: 2})
    p = s.prims()
    net = capi.MLP(seed=4)
    try:
        cent, Ps, sf, nr, na, rad = net.learned_autogrid(p["an"], p["cartes"], p["alpha"],
                                                         p["tmn"], s.ra)
        f = capi.atom_features(p["an"], p["cartes"], p["alpha"], p["tmn"], s.ra)
        y = net.forward(f).astype(np.float64)
        bs = np.array([M.bs_bohr(int(z)) for z in p["an"]])
        assert np.allclose(rad, bs * np.exp(np.clip(y[:, 0], -1.2, 1.2)), rtol=1e-6)
        ref = capi.autogrid(p["an"], p["cartes"], True, False, radius=rad)
        assert np.array_equal(cent, ref[0]) and np.array_equal(Ps, ref[1])
        assert np.array_equal(sf, ref[2])
        assert set(nr) <= set(capi.NRAD_CLASSES) and set(na) <= set(capi.NANG_CLASSES)
        # the origin takes the largest class over all atoms
        assert nr[0] == capi.NRAD_CLASSES[int(np.argmax(y[:, 1:5], axis=1).max())]
    finally:
        net.close()
