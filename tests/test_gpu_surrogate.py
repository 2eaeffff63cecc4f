"""The surrogate's device kernels against the oracle (oracle/surrogate.py):
BF16 MFMA inference and gradients, one Adam step on a Philox minibatch,
the pair features, and delayed-acceptance MC.

Tolerances: the device and the oracle round the same values to BF16 at the
same points; FP32 accumulation order and expf differ, so a hidden
activation can land one BF16 step apart now and then -- outputs and
gradients are compared by norm (relative 2e-3) and element-wise with a
BF16-scale bound.
"""
import numpy as np
import pytest

import oracle
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def engine_with_system(factory, spec, seed=7):
    e = factory()
    e.load_spec(spec, seed=seed)
    e.mc_prepare()
    return e


def small_spec(n_rep=4):
    from cgppd_amd import synth
    return synth.make_system("sur", [22, 75, 40, 130, 18, 2, 60], 150.0, n_rep, 288.0, 450.0, 31, 2.0, 0.2)


def rand_rows(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, S.N_FEAT), np.float32)
    X[:, :32] = rng.uniform(0, 1, (n, 32))
    X[:, 32:38] = rng.uniform(-1, 1, (n, 6))
    X[:, 38:78] = rng.uniform(0, 0.2, (n, 40))
    t = rng.uniform(-20, 20, n).astype(np.float32)
    return X, t


def rel_norm(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-30))


def test_predict_matches_oracle(gpu_engine_factory):
    e = engine_with_system(gpu_engine_factory, small_spec())
    mlp = S.MLP(seed=11)
    mlp.p["b1"][:] = np.random.default_rng(1).normal(0, 0.1, 128)
    mlp.p["b3"][:] = 0.7
    e.surrogate_set_params(mlp.p)
    X, _ = rand_rows(203, 3)
    y = e.surrogate_predict(X)
    want = mlp.forward(X)
    print("predict rel err", rel_norm(y, want), "max abs", np.abs(y - want).max())
    assert rel_norm(y, want) < 2e-3
    assert np.abs(y - want).max() < 2e-2 * max(1.0, np.abs(want).max())
    e.close()


def test_gradient_matches_oracle(gpu_engine_factory):
    e = engine_with_system(gpu_engine_factory, small_spec())
    mlp = S.MLP(seed=12)
    e.surrogate_set_params(mlp.p)
    X, t = rand_rows(150, 4)
    loss, g = e.surrogate_gradient(X, t)
    want_loss, want = mlp.loss_and_grad(X, t)
    assert abs(loss - want_loss) <= 1e-3 * abs(want_loss)
    for k in want:
        err = rel_norm(g[k], want[k])
        print(k, "grad rel err", err)
        assert err < 2e-3, k
    # parameters unchanged
    for k, v in e.surrogate_params().items():
        np.testing.assert_array_equal(v, mlp.p[k])
    e.close()


def test_adam_step_matches_oracle(gpu_engine_factory):
    e = engine_with_system(gpu_engine_factory, small_spec())
    mlp = S.MLP(seed=13)
    e.surrogate_set_params(mlp.p)
    X, t = rand_rows(500, 5)
    e.surrogate_add_samples(X, t)
    seed, batch, lr = 99, 64, 1e-3
    e.surrogate_train(1, batch=batch, lr=lr, seed=seed)
    got = e.surrogate_params()
    # the same minibatch: Philox(step 1, sample i, tag "SG") picks rows
    idx = [(oracle.philox([1, 0, i, 0x53470000], [seed & 0xFFFFFFFF, seed >> 32])[0] * len(X)) >> 32
           for i in range(batch)]
    _, g = mlp.loss_and_grad(X[idx], t[idx])
    p = {k: v.copy() for k, v in mlp.p.items()}
    m = {k: np.zeros_like(v) for k, v in p.items()}
    v_ = {k: np.zeros_like(v) for k, v in p.items()}
    S.adam_step(p, g, m, v_, 1, lr)
    for k in p:
        big = np.abs(g[k]) > 1e-2 * np.abs(g[k]).max()   # sign(g) is well defined there
        step_dev = got[k] - mlp.p[k]
        step_ora = p[k] - mlp.p[k]
        assert np.abs(step_dev[big] - step_ora[big]).max() <= 1e-3 * lr, k
        assert np.abs(step_dev).max() <= 1.0001 * lr
    e.close()


def test_features_match_oracle(gpu_engine_factory):
    spec = small_spec()
    e = engine_with_system(gpu_engine_factory, spec)
    e.mc_steps(50)
    anchor = S.anchors(spec.rel, spec.mol_off)
    comp = S.composition(spec.types, spec.mol_off)
    for r in range(spec.n_rep):
        pos, rel, cen = e.coordinates(r)
        want = S.system_features(cen.astype(np.float64), rel.astype(np.float64), anchor, comp, spec.box)
        got = e.surrogate_features(r)
        print("features max abs err", np.abs(got - want).max())
        np.testing.assert_allclose(got, want, atol=2e-5)
    e.close()


def test_collect_and_train_reduce_loss(gpu_engine_factory):
    spec = small_spec(8)
    e = engine_with_system(gpu_engine_factory, spec)
    e.surrogate_init(seed=3)
    for _ in range(10):
        e.mc_steps(40)
        e.surrogate_collect()
    P = spec.n_mol * (spec.n_mol - 1) // 2
    # only pairs within the surrogate's range are kept
    assert 0 < e.surrogate_sample_count() <= 10 * spec.n_rep * P
    first = e.surrogate_train(20, batch=256, lr=2e-3, seed=1)
    later = e.surrogate_train(400, batch=256, lr=2e-3, seed=2)
    last = e.surrogate_train(20, batch=256, lr=2e-3, seed=3)
    print("loss", first, later, last)
    assert last < 0.5 * first
    e.close()


def test_delayed_acceptance_with_zero_surrogate_is_plain_metropolis(gpu_engine_factory):
    """A surrogate that is 0 everywhere has dS = 0: stage one always passes
    and stage two is the plain Metropolis test, so the trajectory is the
    oracle's."""
    spec = small_spec()
    e = engine_with_system(gpu_engine_factory, spec, seed=21)
    zero = {k: np.zeros_like(v) for k, v in S.MLP(seed=0).p.items()}
    e.surrogate_set_params(zero)
    e.set_delayed_acceptance(True)
    st = oracle.Remc(spec, seed=21)
    for _ in range(3):
        e.mc_steps(60)
        st.mc_steps(60)
        s = e.stats()
        np.testing.assert_array_equal(s["counters"], st.counters)
        for r in range(spec.n_rep):
            np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
    assert e.stage1_rejections().sum() == 0
    e.close()


def test_delayed_acceptance_keeps_exact_energies(gpu_engine_factory):
    """With a trained surrogate and large moves, stage one rejects some
    proposals; accepted states keep exact running energies."""
    from cgppd_amd import synth
    spec = small_spec(8)
    spec.translate_step, spec.rotate_step = 8.0, 0.8
    e = engine_with_system(gpu_engine_factory, spec, seed=5)
    e.surrogate_init(seed=4)
    for _ in range(8):
        e.mc_steps(50)
        e.surrogate_collect()
    e.surrogate_train(300, batch=256, lr=2e-3, seed=9)
    e.set_delayed_acceptance(True)
    e.run(4, 100, use_graph=True)
    s = e.stats()
    rej = e.stage1_rejections()
    print("stage-one rejections", rej, "counters", s["counters"].sum(0))
    assert rej.sum() > 0
    # every step is counted once: accepted, rejected at stage two, or
    # rejected at stage one (exchanges move counters between replicas)
    assert s["counters"].sum() + rej.sum() == s["steps"].sum()
    _, _, full = e.energy_all()
    for r in range(spec.n_rep):
        assert abs(s["energy"][r] - full[r]) <= 1e-6 * max(1.0, abs(full[r])), r
    e.close()


def test_trained_surrogate_gain_over_the_method_without_it(gpu_engine_factory):
    """What the trained network gains over delayed acceptance without it (an
    all-zero network, which passes every proposal to the exact test): from
    the same state and the same Philox moves, stage one rejects a share of
    the proposals before their exact delta-E rows, so the exact rows (and
    residue pairs) evaluated drop by that share, and the proposals that
    reach stage two are accepted more often than plain proposals.  The
    chain stays exact (running totals equal full evaluations).  Whether the
    saving pays for stage one's own cost is measured by bench.py's
    delayed-acceptance leg, not asserted here."""
    spec = small_spec(8)
    spec.translate_step, spec.rotate_step = 8.0, 0.8
    e = engine_with_system(gpu_engine_factory, spec, seed=5)
    e.surrogate_init(seed=4)
    for _ in range(10):
        e.mc_steps(50)
        e.surrogate_collect()
    e.surrogate_train(400, batch=256, lr=2e-3, seed=9)
    trained = e.surrogate_params()
    start = [e.coordinates(r) for r in range(spec.n_rep)]
    st0 = e.stats()
    ts, rs = spec.step_sizes()
    e.close()
    zero = {k: np.zeros_like(v) for k, v in trained.items()}

    def leg(params):
        f = gpu_engine_factory()
        f.load_spec(spec, seed=5)
        f.set_replicas(np.array([c for _, _, c in start]), np.array([rel for _, rel, _ in start]), st0["temperature"],
                       ts, rs, st0["label"], seed=77)
        f.mc_prepare()
        f.surrogate_set_params(params)
        f.set_delayed_acceptance(True)
        p0 = f.pair_evaluations()
        f.run(4, 100, use_graph=True)
        s = f.stats()
        out = {"rejected": int(f.stage1_rejections().sum()), "pairs": int(f.pair_evaluations() - p0),
               "accepted": int(s["counters"][:, :2].sum()), "steps": int(s["steps"].sum())}
        _, _, full = f.energy_all()
        for r in range(spec.n_rep):
            assert abs(s["energy"][r] - full[r]) <= 1e-6 * max(1.0, abs(full[r])), r
        f.close()
        return out

    with_net, without = leg(trained), leg(zero)
    saved = 1.0 - with_net["pairs"] / without["pairs"]
    reach = with_net["steps"] - with_net["rejected"]
    print(f"trained network: {with_net['rejected']} of {with_net['steps']} proposals rejected at stage one, "
          f"exact residue pairs {with_net['pairs']} against {without['pairs']} without it ({saved:.1%} saved); "
          f"stage-two acceptance {with_net['accepted'] / reach:.3f} against plain {without['accepted'] / without['steps']:.3f}",
          flush=True)
    assert without["rejected"] == 0
    assert with_net["rejected"] > 0.05 * with_net["steps"]
    assert with_net["pairs"] < 0.95 * without["pairs"]
    assert with_net["accepted"] / reach > without["accepted"] / without["steps"]
