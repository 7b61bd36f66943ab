"""Pins of the learned-autogrid network's CPU oracle, and the product's host
feature computation against it (no GPU needed)."""
import math

import numpy as np
import pytest
from scipy.special import erf, logsumexp

from inca_parallel_amd import capi, synth
from oracle import mlp_oracle as M


def _params(seed=0, scale=0.3):
    return np.random.default_rng(seed).normal(scale=scale, size=M.NPARAM)


def test_parameter_count_matches_header():
    assert M.NPARAM == capi.MLP_NPARAM == 9993


def test_gelu_closed_form():
    z = np.linspace(-6, 6, 101)
    assert np.allclose(M.gelu(z), z * 0.5 * (1 + erf(z / np.sqrt(2))), rtol=0, atol=1e-15)
    # derivative by central differences
    h = 1e-6
    assert np.allclose(M.gelu_d(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)


def test_forward_one_atom_by_loops():
    p = _params(1)
    x = np.random.default_rng(2).normal(size=(1, 16))
    W = M.unpack(p)
    a = x[0]
    for li in range(3):
        Wl, bl = W[2 * li], W[2 * li + 1]
        a = [0.5 * z * (1 + math.erf(z / math.sqrt(2)))
             for z in (sum(a[i] * Wl[i][o] for i in range(len(a))) + bl[o]
                       for o in range(Wl.shape[1]))]
    y = [sum(a[i] * W[6][i][o] for i in range(64)) + W[7][o] for o in range(9)]
    assert np.allclose(M.forward(p, x)[0][0], y, rtol=1e-13, atol=1e-13)


def test_loss_against_logsumexp():
    rng = np.random.default_rng(3)
    y = rng.normal(size=(7, 9))
    t = np.stack([rng.normal(size=7), rng.integers(0, 4, 7), rng.integers(0, 4, 7)], axis=1)
    ref = np.mean((y[:, 0] - t[:, 0]) ** 2
                  + logsumexp(y[:, 1:5], axis=1) - y[np.arange(7), 1 + t[:, 1].astype(int)]
                  + logsumexp(y[:, 5:9], axis=1) - y[np.arange(7), 5 + t[:, 2].astype(int)])
    assert abs(M.loss_and_dout(y, t)[0] - ref) < 1e-13


def test_gradient_by_finite_differences():
    rng = np.random.default_rng(4)
    p = _params(5)
    x = rng.normal(size=(6, 16))
    t = np.stack([rng.normal(size=6), rng.integers(0, 4, 6), rng.integers(0, 4, 6)], axis=1)
    _, g = M.loss_grad(p, x, t)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss_grad(p + e, x, t)[0] - M.loss_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_adam_first_step_moves_by_lr():
    g = np.array([0.3, -2.0, 1e-3])
    p1, _, _ = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p1, -0.01 * np.sign(g), rtol=1e-4)


@pytest.mark.parametrize("seed", [3, 11, 40])
def test_atom_features_match_oracle(seed):
    s = synth.config_cluster(seed=seed, natoms=5, nheavy=3, nocc=3,
                             shells_heavy={0: 2, 1: 1, 2: 1}, shells_h={0: 2})
    p = s.prims()
    got = capi.atom_features(p["an"], p["cartes"], p["alpha"], p["tmn"], s.ra)
    ref = M.atom_features(p["an"], p["cartes"], p["alpha"], p["tmn"], s.ra)
    assert got.shape == (5, 16)
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-7)
