"""The benchmark configurations at their full sizes on the device.

- 256 chains of 20-2000 residues (about 106k), 64 replicas: after MC steps,
  1005 molecule-pair entries of the device energy matrices, sampled over
  four replicas and including every entry of the largest molecule's row in
  one of them, against the FP64 oracle of that pair's coordinates (1e-5
  relative); running totals against a full evaluation.
- 40 chains (about 6.2k residues), 32 replicas, with the learned surrogate
  driving delayed acceptance: a briefly trained network over replayed
  hipGraph intervals keeps the running totals equal to full evaluations,
  and an all-zero network reproduces the plain trajectory bit for bit.
- The linked-domain molecule with flexible linkers at its configured 64
  replicas: the oracle's trajectory step by step, coordinates bit for bit
  (a replica whose decision parts from the oracle's on a borderline
  uniform is shown to be borderline and left out afterwards).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

REL = 1e-5


def pair_fp64(spec, pos, I, J):
    """FP64 oracle energy (kcal/mol) of molecules I and J on their own."""
    a, b = int(spec.mol_off[I]), int(spec.mol_off[I + 1])
    c, d = int(spec.mol_off[J]), int(spec.mol_off[J + 1])
    off = np.array([0, b - a, b - a + d - c], np.int32)
    ty = np.concatenate([spec.types[a:b], spec.types[c:d]])
    p = np.concatenate([pos[a:b], pos[c:d]])
    return oracle.total(*oracle.energy_fp64(off, ty, spec.ff.charge, spec.ff.sigma, spec.ff.eps, p, spec.box))


@pytest.fixture(scope="module")
def crowded256():
    from cgppd_amd import synth
    return synth.config(4)


def test_crowded256_sampled_pairs_against_fp64(gpu_engine_factory, crowded256):
    spec = crowded256
    e = gpu_engine_factory()
    e.load_spec(spec, seed=41)
    e.mc_prepare()
    e.run(2, 100, use_graph=True)
    s = e.stats()
    print(f"{spec.name}: {spec.n_res} residues, {spec.n_rep} replicas, counters {s['counters'].sum(0)}", flush=True)
    M = spec.n_mol
    big = int(np.argmax(spec.lengths))
    rng = np.random.default_rng(5)
    checked, worst = 0, 0.0
    for r in (0, 21, 42, 63):
        mat = e.energy_matrix(r)
        pos = e.coordinates(r)[0]
        if r == 21:
            pairs = [(big, J) for J in range(M) if J != big]  # the largest molecule's whole row
            pairs += [tuple(rng.choice(M, 2, replace=False)) for _ in range(250 - len(pairs))]
        else:
            pairs = [tuple(rng.choice(M, 2, replace=False)) for _ in range(250)]
        for I, J in pairs:
            want = pair_fp64(spec, pos, I, J)
            got = mat[I, J]
            assert mat[J, I] == got
            err = abs(got - want)
            assert err <= REL * abs(want) + 1e-7, (r, I, J, got, want)
            if abs(want) > 1e-3:
                worst = max(worst, err / abs(want))
            checked += 1
    print(f"{checked} sampled molecule pairs, worst relative error {worst:.2e} (entries above 1e-3 kcal/mol)",
          flush=True)
    assert checked == 750 + M - 1
    _, _, full = e.energy_all()
    np.testing.assert_allclose(s["energy"], full, rtol=REL, atol=1e-6)
    e.close()


def train_briefly(e, rounds=6, steps=200):
    """On-device samples from plain MC, then a short Adam run."""
    e.surrogate_init(seed=4)
    for _ in range(rounds):
        e.mc_steps(50)
        e.surrogate_collect()
    return e.surrogate_train(steps, batch=1024, lr=2e-3, seed=9)


def test_crowded40_delayed_acceptance_keeps_exact_totals(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.config(1)
    e = gpu_engine_factory()
    e.load_spec(spec, seed=13)
    e.mc_prepare()
    loss = train_briefly(e)
    print(f"{spec.name}: {e.surrogate_sample_count()} training pairs, mean loss {loss:.3f}", flush=True)
    e.set_delayed_acceptance(True)
    e.run(3, 200, use_graph=True)
    s = e.stats()
    rej = e.stage1_rejections()
    print(f"stage-one rejections {rej.sum()}, counters {s['counters'].sum(0)}", flush=True)
    # every step is decided once: at stage two (or plainly), or at stage one
    assert s["counters"].sum() + rej.sum() == s["steps"].sum()
    _, _, full = e.energy_all()
    for r in range(spec.n_rep):
        assert abs(s["energy"][r] - full[r]) <= 1e-6 * max(1.0, abs(full[r])), r
    e.close()


def test_crowded40_zero_surrogate_is_plain_metropolis(gpu_engine_factory):
    """A surrogate that is 0 everywhere accepts every proposal at stage one
    and leaves stage two as the plain Metropolis test on the same uniform."""
    from cgppd_amd import synth
    spec = synth.config(1)
    plain = gpu_engine_factory()
    plain.load_spec(spec, seed=21)
    # one MC step per launch pair, as delayed acceptance runs (windows sum
    # the delta-E in another FP64 order)
    plain.set_window(1)
    plain.mc_prepare()
    da = gpu_engine_factory()
    da.load_spec(spec, seed=21)
    da.mc_prepare()
    da.surrogate_set_params(np.zeros(27009, np.float32))
    da.set_delayed_acceptance(True)
    plain.run(2, 250, use_graph=True)
    da.run(2, 250, use_graph=True)
    a, b = plain.stats(), da.stats()
    np.testing.assert_array_equal(a["counters"], b["counters"])
    np.testing.assert_array_equal(a["temperature"], b["temperature"])
    assert da.stage1_rejections().sum() == 0
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(plain.coordinates(r)[0], da.coordinates(r)[0])
    np.testing.assert_allclose(a["energy"], b["energy"], rtol=1e-12, atol=1e-12)
    plain.close()
    da.close()


def lockstep(e, st, spec, seed, n_steps, rel_tol=REL):
    """One MC step at a time on the device and in the oracle.  The two sides
    sum the same energies in FP32 and in the reference's arithmetic, so a
    decision can differ where the Metropolis uniform lies within the energy
    tolerance of the Boltzmann factor.  Such a replica is checked to be
    borderline at the step where it parted and is not compared after it.
    Returns the replicas that parted."""
    M = spec.n_mol
    mol_flexible = [int(spec.molecule_is_flexible(m)) for m in range(M)]
    parted = np.zeros(spec.n_rep, bool)
    s = e.stats()
    for k in range(n_steps):
        e_before, o_before = s["energy"].copy(), st.energy.copy()
        a_before = s["counters"][:, :2].sum(1)
        oa_before = st.counters[:, :2].sum(1).copy()
        step = st.step.copy()
        e.mc_steps(1)
        st.mc_steps(1)
        s = e.stats()
        acc_d = s["counters"][:, :2].sum(1) - a_before
        acc_o = st.counters[:, :2].sum(1) - oa_before
        for r in np.flatnonzero(~parted & (acc_d != acc_o)):
            # the accepting side knows the move's energy change
            d = s["energy"][r] - e_before[r] if acc_d[r] else st.energy[r] - o_before[r]
            u = oracle.proposal_flex(seed, int(r), int(step[r]), mol_flexible)[3]
            beta = 4184.0 / (8.3144621 * st.temperature[r])
            p = np.exp(-beta * d)
            window = beta * p * rel_tol * max(1.0, abs(e_before[r]))
            print(f"replica {r} parts at step {step[r]}: dE {d:.6g}, u {u:.9f}, exp(-dE/kT) {p:.9f}, "
                  f"window {window:.2e}", flush=True)
            assert d >= -rel_tol * max(1.0, abs(e_before[r]))
            assert abs(u - min(1.0, p)) <= window
            parted[r] = True
        keep = ~parted
        np.testing.assert_array_equal(s["steps"], st.step)
        np.testing.assert_allclose(s["energy"][keep], st.energy[keep], rtol=rel_tol, atol=1e-6)
    for r in np.flatnonzero(~parted):
        np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
    return np.flatnonzero(parted)


def test_linked4x76_at_64_replicas_matches_oracle(gpu_engine_factory):
    """Flex, local-translate, crankshaft and rigid moves of the 322-residue
    molecule at the configured 64 replicas, against the oracle's steps."""
    from cgppd_amd import synth
    spec = synth.config(2)
    assert spec.n_rep == 64
    e = gpu_engine_factory()
    e.load_spec(spec, seed=23)
    e.mc_prepare()
    st = oracle.Remc(spec, seed=23)
    parted = lockstep(e, st, spec, 23, 80)
    s = e.stats()
    print(f"80 steps: gpu {s['counters'].sum(0)} oracle {st.counters.sum(0)}, parted replicas {list(parted)}",
          flush=True)
    # borderline decisions are rare: a systematic difference would part most
    assert len(parted) <= 2
    e.close()
