"""Pins the surrogate oracle (oracle/surrogate.py): BF16 rounding, the
network's gradient against central finite differences of the FP64 network,
the BF16-recipe backward against the FP64 gradient, feature invariances,
closed-form RBFs, and one Adam step in closed form."""
import numpy as np
import pytest

from oracle import surrogate as S


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.4e38], dtype=np.float32)
    got = S.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie -> even mantissa
    assert got[2] == np.float32(1 + 2 ** -6)  # tie -> even (up)
    assert got[3] == -2.5
    assert np.isinf(got[4])                   # above the largest BF16 (3.39e38)
    # every output is a BF16 value and within half a BF16 ulp
    r = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    b = S.bf16(r)
    assert (b.view(np.uint32) & 0xFFFF == 0).all()
    assert (np.abs(b - r) <= np.abs(r) * 2.0 ** -8 + 1e-30).all()


def rand_batch(n=48, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, S.N_FEAT)).astype(np.float32)
    X[:, 78:] = 0
    t = rng.uniform(-20, 20, size=n).astype(np.float32)
    return X, t


def test_fp64_gradient_matches_finite_differences():
    X, t = rand_batch(16)
    net = S.MLP64(S.MLP(seed=3).p)
    g = net.grad(X.astype(np.float64), t.astype(np.float64))
    rng = np.random.default_rng(7)
    for k in net.p:
        flat = net.p[k].reshape(-1)
        for i in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[i]
            h = 1e-6 * max(1.0, abs(old))
            flat[i] = old + h
            lp = net.loss(X, t)
            flat[i] = old - h
            lm = net.loss(X, t)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k].reshape(-1)[i]) <= 1e-6 * max(1.0, abs(fd)), (k, i, fd, g[k].reshape(-1)[i])


def test_bf16_recipe_backward_is_the_gradient():
    """The BF16-recipe gradient agrees with the FP64 gradient to BF16
    accuracy: no dropped term, sign or transposition."""
    X, t = rand_batch(64)
    mlp = S.MLP(seed=5)
    loss, g = mlp.loss_and_grad(X, t)
    net = S.MLP64(mlp.p)
    g64 = net.grad(S.bf16(X).astype(np.float64), t.astype(np.float64))
    assert abs(loss - net.loss(S.bf16(X).astype(np.float64), t)) <= 2e-2 * abs(loss)
    for k in g:
        err = np.linalg.norm(g[k] - g64[k]) / max(np.linalg.norm(g64[k]), 1e-30)
        assert err < 3e-2, (k, err)


def test_rbf_closed_form():
    f = S.pair_features([0, 0, 0], [9.0, 0, 0], np.eye(3), np.eye(3), np.zeros(20), np.zeros(20), 1000.0)
    assert f[3] == pytest.approx(1.0)
    assert f[2] == pytest.approx(np.exp(-0.5))
    assert f[4] == pytest.approx(np.exp(-0.5))
    np.testing.assert_allclose(f[32:35], [1, 0, 0])
    np.testing.assert_allclose(f[35:38], [-1, 0, 0])
    # minimum image: 990 A apart in a 1000 A box is 10 A
    g = S.pair_features([0, 0, 0], [990.0, 0, 0], np.eye(3), np.eye(3), np.zeros(20), np.zeros(20), 1000.0)
    assert g[0] == pytest.approx(np.exp(-0.5 * (10 / 3) ** 2))


def small_system():
    from cgppd_amd import synth
    return synth.make_system("feat", [30, 45, 12, 2, 60], 200.0, 1, 300.0, 300.0, 13)


def test_features_are_rigid_motion_invariant():
    spec = small_system()
    anchor = S.anchors(spec.rel, spec.mol_off)
    comp = S.composition(spec.types, spec.mol_off)
    f0 = S.system_features(spec.centers.astype(np.float64), spec.rel.astype(np.float64), anchor, comp, 1e9)
    rng = np.random.default_rng(2)
    q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    shift = rng.uniform(-50, 50, 3)
    f1 = S.system_features(spec.centers @ q.T + shift, spec.rel @ q.T, anchor, comp, 1e9)
    # distances and compositions are invariant for every pair; the frame
    # components too unless a molecule has fewer than 3 residues (it has no
    # body frame and uses the fixed lab frame)
    np.testing.assert_allclose(f1[:, :32], f0[:, :32], atol=1e-9)
    np.testing.assert_allclose(f1[:, 38:], f0[:, 38:], atol=1e-12)
    framed = [a != b for a, b in anchor]
    row = 0
    for I in range(spec.n_mol):
        for J in range(I + 1, spec.n_mol):
            if framed[I] and framed[J]:
                np.testing.assert_allclose(f1[row, 32:38], f0[row, 32:38], atol=1e-9)
            row += 1


def test_features_of_swapped_molecules():
    """f(J, I) is f(I, J) with the two frame blocks and composition blocks
    exchanged."""
    spec = small_system()
    anchor = S.anchors(spec.rel, spec.mol_off)
    comp = S.composition(spec.types, spec.mol_off)
    frames = [np.eye(3) if a == b else S.frame(spec.rel[a], spec.rel[b]) for a, b in anchor]
    fij = S.pair_features(spec.centers[0], spec.centers[2], frames[0], frames[2], comp[0], comp[2], spec.box)
    fji = S.pair_features(spec.centers[2], spec.centers[0], frames[2], frames[0], comp[2], comp[0], spec.box)
    np.testing.assert_allclose(fij[:32], fji[:32])
    np.testing.assert_allclose(fij[32:35], fji[35:38], atol=1e-12)
    np.testing.assert_allclose(fij[38:58], fji[58:78])


def test_adam_first_step_closed_form():
    p = {"w": np.array([1.0, -2.0, 0.5], np.float32)}
    g = {"w": np.array([0.3, -4.0, 1e-3], np.float32)}
    m = {"w": np.zeros(3, np.float32)}
    v = {"w": np.zeros(3, np.float32)}
    S.adam_step(p, g, m, v, 1, 0.01)
    # m_hat = g, v_hat = g^2: p -= lr g / (|g| + eps)
    want = np.array([1.0, -2.0, 0.5]) - 0.01 * g["w"] / (np.abs(g["w"]) + 1e-8)
    np.testing.assert_allclose(p["w"], want, rtol=1e-6)
