"""CPU oracle of the learned surrogate (no reference counterpart).

TEST INFRASTRUCTURE ONLY: the checker for the device's feature kernel and
the BF16 MFMA training / inference kernels of cgppd_amd.

The surrogate predicts the inter-molecular energy E_IJ (kcal/mol) of a
molecule pair from 80 features (the canonical order has I < J):
  f[0:32]   Gaussian RBFs of the minimum-image centre distance d,
            exp(-0.5 ((d - 3k) / 3)^2), k = 0..31 (A)
  f[32:35]  unit vector u from centre I to centre J in molecule I's body frame
  f[35:38]  -u in molecule J's body frame
  f[38:58]  residue-type fractions of molecule I (20 bins)
  f[58:78]  residue-type fractions of molecule J
  f[78:80]  zero
A body frame is built from the relative positions of two anchor residues a,
b of the molecule: e1 = a/|a|, e2 = (b - (b.e1) e1)/|...|, e3 = e1 x e2.

MLP: 80 -> 128 -> 128 -> 1, SiLU on both hidden layers.  Precision recipe
(the device follows it step for step): inputs, weights and hidden
activations are rounded to BF16 (round to nearest even) before each matrix
product, products accumulate in FP32, biases and pre-activations are FP32.
Training: loss 0.5 mean((y - t)^2) with t clipped to [-15, 15] kcal/mol, FP32
master weights, Adam (beta1 0.9, beta2 0.999, eps 1e-8).

Pinned by: finite differences of the FP64 version of the same network
(gradients), rotation / translation / permutation invariances of the
features, closed-form RBF values, and the Adam update written out for a
scalar (tests/test_surrogate_oracle.py).
"""
from __future__ import annotations

import numpy as np

N_RBF = 32
RBF_STEP = 3.0
RBF_WIDTH = 3.0
N_FEAT = 80
HIDDEN = 128
TARGET_CLIP = 15.0


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even); returns float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = ((u + r) >> 16).astype(np.uint32) << 16
    nan = np.isnan(a)
    out = out.astype(np.uint32)
    out[nan] = 0x7FC00000
    return out.view(np.float32).reshape(a.shape)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def frame(ra, rb):
    """Body frame (3x3, rows e1 e2 e3) from two anchor relative positions."""
    ra = np.asarray(ra, dtype=np.float64)
    rb = np.asarray(rb, dtype=np.float64)
    n1 = np.linalg.norm(ra)
    if n1 == 0:
        return np.eye(3)
    e1 = ra / n1
    p = rb - np.dot(rb, e1) * e1
    n2 = np.linalg.norm(p)
    if n2 == 0:
        return np.eye(3)
    e2 = p / n2
    return np.stack([e1, e2, np.cross(e1, e2)])


def anchors(rel, mol_off):
    """Anchor residues (a, b) of every molecule: a = farthest from the
    centre, b = largest distance from the line through a (caller indices,
    global).  Molecules of fewer than 3 residues use a = b = first."""
    out = []
    for m in range(len(mol_off) - 1):
        s, e = int(mol_off[m]), int(mol_off[m + 1])
        x = np.asarray(rel[s:e], dtype=np.float64)
        if e - s < 3:
            out.append((s, s))
            continue
        a = int(np.argmax((x * x).sum(1)))
        e1 = x[a] / max(np.linalg.norm(x[a]), 1e-30)
        perp = x - (x @ e1)[:, None] * e1
        b = int(np.argmax((perp * perp).sum(1)))
        out.append((s + a, s + b))
    return np.array(out, dtype=np.int32)


def composition(types, mol_off):
    M = len(mol_off) - 1
    h = np.zeros((M, 20))
    for m in range(M):
        s, e = int(mol_off[m]), int(mol_off[m + 1])
        h[m] = np.bincount(np.asarray(types[s:e]), minlength=20)[:20] / float(e - s)
    return h


def pair_features(ci, cj, fi, fj, hi, hj, box):
    d = np.asarray(cj, dtype=np.float64) - np.asarray(ci, dtype=np.float64)
    d -= box * np.rint(d / box)
    dist = float(np.linalg.norm(d))
    u = d / dist if dist > 0 else np.zeros(3)
    f = np.zeros(N_FEAT)
    mu = RBF_STEP * np.arange(N_RBF)
    f[:N_RBF] = np.exp(-0.5 * ((dist - mu) / RBF_WIDTH) ** 2)
    f[32:35] = fi @ u
    f[35:38] = fj @ (-u)
    f[38:58] = hi
    f[58:78] = hj
    return f


def system_features(centres, rel, anchor, comp, box):
    """Features of every pair I < J of one replica: (M(M-1)/2, 80), pairs in
    row-major (I, J) order.  A molecule whose two anchors coincide (fewer
    than 3 residues) has the identity frame."""
    M = len(centres)
    frames = [np.eye(3) if a == b else frame(rel[a], rel[b]) for a, b in anchor]
    rows = []
    for I in range(M):
        for J in range(I + 1, M):
            rows.append(pair_features(centres[I], centres[J], frames[I], frames[J], comp[I], comp[J], box))
    return np.array(rows).reshape(-1, N_FEAT)


class MLP:
    """Parameters: W1 (80, 128), b1 (128), W2 (128, 128), b2 (128),
    w3 (128), b3 scalar; FP32 masters."""

    def __init__(self, seed=0, params=None):
        if params is not None:
            self.p = {k: np.array(v, dtype=np.float32) for k, v in params.items()}
            return
        rng = np.random.default_rng(seed)
        self.p = {
            "W1": (rng.standard_normal((N_FEAT, HIDDEN)) * np.sqrt(2.0 / N_FEAT)).astype(np.float32),
            "b1": np.zeros(HIDDEN, np.float32),
            "W2": (rng.standard_normal((HIDDEN, HIDDEN)) * np.sqrt(2.0 / HIDDEN)).astype(np.float32),
            "b2": np.zeros(HIDDEN, np.float32),
            "w3": (rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)).astype(np.float32),
            "b3": np.zeros(1, np.float32),
        }

    def forward(self, X, keep=False):
        """BF16 recipe forward; X (B, 80) float32 (rounded to BF16 here)."""
        p = self.p
        x = bf16(X)
        z1 = (x.astype(np.float64) @ bf16(p["W1"]).astype(np.float64)).astype(np.float32) + p["b1"]
        a1 = bf16(silu(z1.astype(np.float64)).astype(np.float32))
        z2 = (a1.astype(np.float64) @ bf16(p["W2"]).astype(np.float64)).astype(np.float32) + p["b2"]
        a2 = bf16(silu(z2.astype(np.float64)).astype(np.float32))
        y = (a2.astype(np.float64) @ bf16(p["w3"]).astype(np.float64)).astype(np.float32) + p["b3"][0]
        if keep:
            return y, (x, z1, a1, z2, a2)
        return y

    def loss_and_grad(self, X, t):
        """0.5 mean((y - clip(t))^2) and its gradient (BF16 recipe)."""
        p = self.p
        B = X.shape[0]
        y, (x, z1, a1, z2, a2) = self.forward(X, keep=True)
        tc = np.clip(t, -TARGET_CLIP, TARGET_CLIP).astype(np.float32)
        dy = ((y - tc) / np.float32(B)).astype(np.float32)
        loss = 0.5 * float(np.mean((y.astype(np.float64) - tc) ** 2))
        g = {}
        g["b3"] = np.array([dy.astype(np.float64).sum()], dtype=np.float32)
        g["w3"] = (a2.astype(np.float64).T @ dy.astype(np.float64)).astype(np.float32)
        w3 = bf16(p["w3"])
        dz2 = (np.outer(dy, w3).astype(np.float32) * silu_grad(z2.astype(np.float64))).astype(np.float32)
        dz2b = bf16(dz2)
        g["b2"] = dz2.astype(np.float64).sum(0).astype(np.float32)
        g["W2"] = (a1.astype(np.float64).T @ dz2b.astype(np.float64)).astype(np.float32)
        da1 = (dz2b.astype(np.float64) @ bf16(p["W2"]).astype(np.float64).T).astype(np.float32)
        dz1 = (da1 * silu_grad(z1.astype(np.float64))).astype(np.float32)
        dz1b = bf16(dz1)
        g["b1"] = dz1.astype(np.float64).sum(0).astype(np.float32)
        g["W1"] = (x.astype(np.float64).T @ dz1b.astype(np.float64)).astype(np.float32)
        return loss, g


class MLP64:
    """The same network entirely in float64 (no BF16 rounding): the
    finite-difference reference for the backward pass."""

    def __init__(self, params):
        self.p = {k: np.array(v, dtype=np.float64) for k, v in params.items()}

    def forward(self, X):
        p = self.p
        a1 = silu(X @ p["W1"] + p["b1"])
        a2 = silu(a1 @ p["W2"] + p["b2"])
        return a2 @ p["w3"] + p["b3"][0]

    def loss(self, X, t):
        y = self.forward(X)
        tc = np.clip(t, -TARGET_CLIP, TARGET_CLIP)
        return 0.5 * np.mean((y - tc) ** 2)

    def grad(self, X, t):
        p = self.p
        B = X.shape[0]
        z1 = X @ p["W1"] + p["b1"]
        a1 = silu(z1)
        z2 = a1 @ p["W2"] + p["b2"]
        a2 = silu(z2)
        y = a2 @ p["w3"] + p["b3"][0]
        dy = (y - np.clip(t, -TARGET_CLIP, TARGET_CLIP)) / B
        g = {"b3": np.array([dy.sum()]), "w3": a2.T @ dy}
        dz2 = np.outer(dy, p["w3"]) * silu_grad(z2)
        g["b2"] = dz2.sum(0)
        g["W2"] = a1.T @ dz2
        dz1 = (dz2 @ p["W2"].T) * silu_grad(z1)
        g["b1"] = dz1.sum(0)
        g["W1"] = X.T @ dz1
        return g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in FP32 (Kingma & Ba 2015, algorithm 1); step >= 1."""
    for k in p:
        m[k] = (np.float32(b1) * m[k] + np.float32(1 - b1) * g[k]).astype(np.float32)
        v[k] = (np.float32(b2) * v[k] + np.float32(1 - b2) * g[k] * g[k]).astype(np.float32)
        mh = m[k] / np.float32(1 - b1 ** step)
        vh = v[k] / np.float32(1 - b2 ** step)
        p[k] = (p[k] - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))).astype(np.float32)
