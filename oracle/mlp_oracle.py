"""CPU float64 restatement of the learned-autogrid network and its features.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels of
inca_parallel_amd/csrc/mlp.hip and the host features of autogrid.hip.  Only
tests/ may import it; the product never does.

The network has no counterpart in the reference; the feature inputs come
from the reference's molecule description (geninfo: an, cartes, Alpha, TMN,
Ra) and the Bragg-Slater table of get_bs_radius (mod_autogrid.f90:142-185).
Parameter layout (flat): W1 (16x64, [in][out]), b1, W2 (64x64), b2,
W3 (64x64), b3, W4 (64x9), b4.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH, NOUT = 16, 64, 9
SHAPES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
NPARAM = int(sum(np.prod(s) for s in SHAPES))

_ERF = np.vectorize(math.erf)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _ERF(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _ERF(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(p, x):
    """out (n,9) and the cache for backward."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(x, np.float64)
    z1 = x @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    z3 = a2 @ W3 + b3
    a3 = gelu(z3)
    y = a3 @ W4 + b4
    return y, (x, z1, a1, z2, a2, z3, a3)


def loss_and_dout(y, target):
    """mean over atoms of (y0 - t0)^2 + CE(nrad head) + CE(nang head)."""
    n = y.shape[0]
    t = np.asarray(target, np.float64)
    d = np.zeros_like(y)
    e = y[:, 0] - t[:, 0]
    loss = np.sum(e * e)
    d[:, 0] = 2.0 * e / n
    for head in range(2):
        o0 = 1 + 4 * head
        logits = y[:, o0:o0 + 4]
        m = logits.max(axis=1, keepdims=True)
        lse = m[:, 0] + np.log(np.exp(logits - m).sum(axis=1))
        cls = t[:, 1 + head].astype(int)
        loss += np.sum(lse - logits[np.arange(n), cls])
        sm = np.exp(logits - lse[:, None])
        sm[np.arange(n), cls] -= 1.0
        d[:, o0:o0 + 4] = sm / n
    return loss / n, d


def loss_grad(p, x, target):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (x, z1, a1, z2, a2, z3, a3) = forward(p, x)
    loss, d4 = loss_and_dout(y, target)
    d3 = (d4 @ W4.T) * gelu_d(z3)
    d2 = (d3 @ W3.T) * gelu_d(z2)
    d1 = (d2 @ W2.T) * gelu_d(z1)
    g = [x.T @ d1, d1.sum(0), a1.T @ d2, d2.sum(0), a2.T @ d3, d3.sum(0), a3.T @ d4, d4.sum(0)]
    return loss, np.concatenate([v.reshape(-1) for v in g])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    return p - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v


# Bragg-Slater radii in angstrom (mod_autogrid.f90:142-185)
_BS = [0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.38, 1.80, 1.50, 1.25, 1.10,
       1.00, 1.00, 1.00, 1.00, 2.20, 1.80, 1.60, 1.40, 1.35, 1.40, 1.40, 1.40, 1.35, 1.35,
       1.35, 1.35, 1.30, 1.25, 1.15, 1.15, 1.15, 1.00]


def bs_bohr(z):
    return (_BS[z - 1] if 1 <= z <= 36 else 1.0) * 1.8897


def atom_features(an, cartes, alpha, tmn, prim_atom):
    """(natoms,16) descriptors; cartes (natoms,3), tmn (nprim,3), prim_atom 1-based."""
    an = np.asarray(an)
    xyz = np.asarray(cartes, np.float64).reshape(len(an), 3)
    alpha = np.asarray(alpha, np.float64)
    tmn = np.asarray(tmn).reshape(-1, 3)
    ra = np.asarray(prim_atom)
    nat = len(an)
    gmin, gmax = (alpha.min(), alpha.max()) if len(alpha) else (1.0, 1.0)
    f = np.zeros((nat, 16))
    for a in range(nat):
        mine = ra == a + 1
        amin, amax = (alpha[mine].min(), alpha[mine].max()) if mine.any() else (1.0, 1.0)
        lmax = int(tmn[mine].sum(axis=1).max()) if mine.any() else 0
        others = [b for b in range(nat) if b != a]
        d = np.array([np.linalg.norm(xyz[a] - xyz[b]) for b in others])
        z = int(an[a])
        period = 1 if z <= 2 else (2 if z <= 10 else (3 if z <= 18 else 4))
        f[a] = [z / 10.0, float(z == 1), period / 3.0, math.log10(amin) / 3.0,
                math.log10(amax) / 3.0, mine.sum() / 20.0, lmax / 2.0,
                (d.min() if len(d) else 0.0) / 5.0, (d.mean() if len(d) else 0.0) / 5.0,
                np.sum(d < 4.0) / 10.0, nat / 20.0, bs_bohr(z) / 3.0,
                sum(an[b] / dd for b, dd in zip(others, d) if dd > 1e-8) / 10.0,
                math.log10(gmin) / 3.0, math.log10(gmax) / 3.0, 1.0]
    return f
