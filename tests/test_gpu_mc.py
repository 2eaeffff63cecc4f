"""On-device MC and replica exchange against the CPU oracle.

Both sides draw every proposal from the same Philox stream and apply the
move with the same float/double operation order, so with identical
accept/reject decisions the coordinates agree bit for bit; energies agree
to the FP32-vs-FP64 pair-term bar.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def small_spec(n_rep=8):
    from cgppd_amd import synth
    return synth.make_system("mcsmall", [22, 75, 40, 130, 18], 110.0, n_rep, 288.0, 450.0, 31, 2.0, 0.2)


def gpu_and_oracle(factory, spec, seed):
    e = factory()
    e.load_spec(spec, seed=seed)
    e.mc_prepare()
    st = oracle.Remc(spec, seed=seed)
    return e, st


def test_prepare_matrix_and_totals(gpu_engine_factory):
    spec = small_spec()
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 5)
    s = e.stats()
    np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-9)
    for r in range(spec.n_rep):
        np.testing.assert_allclose(e.energy_matrix(r), st.emat[r], rtol=1e-5, atol=1e-9)
    e.close()


def test_mc_trajectory_matches_oracle(gpu_engine_factory):
    spec = small_spec()
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 123)
    for chunk in range(4):
        e.mc_steps(75)
        st.mc_steps(75)
        s = e.stats()
        print(f"after {75 * (chunk + 1)} steps: counters gpu {s['counters'].sum(0)} oracle {st.counters.sum(0)}")
        np.testing.assert_array_equal(s["steps"], st.step)
        np.testing.assert_array_equal(s["counters"], st.counters)
        np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
        for r in range(spec.n_rep):
            pos, rel, cen = e.coordinates(r)
            np.testing.assert_array_equal(pos, st.pos[r])
            np.testing.assert_array_equal(rel, st.rel[r])
            np.testing.assert_array_equal(cen, st.center[r])
    e.close()


def test_running_total_matches_full_recompute(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.config(0)
    e = gpu_engine_factory()
    e.load_spec(spec, seed=9)
    e.mc_prepare()
    e.mc_steps(1000)
    s = e.stats()
    _, _, full = e.energy_all()
    for r in range(spec.n_rep):
        pos, _, _ = e.coordinates(r)
        flj, fdh = oracle.energy_fp64(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps, pos,
                                      spec.box)
        want = oracle.total(flj, fdh)
        assert abs(s["energy"][r] - want) <= 1e-5 * abs(want) + 1e-6, r
        assert abs(full[r] - want) <= 1e-5 * abs(want) + 1e-9, r
    assert s["counters"].sum() == 1000 * spec.n_rep
    e.close()


def test_exchange_rounds_match_oracle(gpu_engine_factory):
    spec = small_spec()
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 77)
    for k in range(6):
        e.run(1, 25, use_graph=False)
        st.mc_steps(25)
        st.exchange()
    s = e.stats()
    np.testing.assert_array_equal(s["counters"], st.counters)
    np.testing.assert_array_equal(s["label"], st.label)
    np.testing.assert_array_equal(s["temperature"], st.temperature)
    np.testing.assert_array_equal(s["exchange"], st.exchange_stats)
    np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
    e.close()


def test_graph_replay_equals_direct_launches(gpu_engine_factory):
    spec = small_spec()
    a, _ = gpu_and_oracle(gpu_engine_factory, spec, 5)
    b = gpu_engine_factory()
    b.load_spec(spec, seed=5)
    b.mc_prepare()
    a.run(4, 30, use_graph=False)
    for _ in range(4):
        b.run(1, 30, use_graph=True)   # replay the captured interval, eager reads between replays
        b.synchronize()
        b.stats()
    sa, sb = a.stats(), b.stats()
    for k in ("energy", "temperature", "label", "counters", "steps", "exchange"):
        np.testing.assert_array_equal(sa[k], sb[k])
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(a.coordinates(r)[0], b.coordinates(r)[0])
    a.close()
    b.close()


def test_config1_mc_steps_keep_totals_consistent(gpu_engine_factory):
    """Full-size 40-chain system: after MC steps the running totals equal a
    fresh full evaluation (size-independent consistency check)."""
    from cgppd_amd import synth
    spec = synth.config(1)
    e = gpu_engine_factory()
    e.load_spec(spec, seed=3)
    e.mc_prepare()
    e.run(2, 100, use_graph=True)
    s = e.stats()
    _, _, full = e.energy_all()
    np.testing.assert_allclose(s["energy"], full, rtol=1e-5, atol=1e-6)
    assert (s["steps"] == 200).all()
    assert s["counters"].sum() == 200 * spec.n_rep
    e.close()


@pytest.mark.parametrize("layout", ["ilanes", "jlanes", "jlanes4"])
def test_pair_layouts_match_oracle(gpu_engine_factory, layout):
    """Both pair-kernel layouts give the oracle's trajectory; j-lanes also
    on a system whose molecules are shorter than a wave (many molecules per
    j-block) and with the per-pair minimum image forced."""
    from cgppd_amd import synth
    for spec in (small_spec(), synth.make_system("tiny", [3, 1, 9, 2, 70, 5, 1, 12], 60.0, 6, 288.0, 450.0, 7, 2.0,
                                                 0.2)):
        for shift in (True, False):
            print(f"layout {layout} spec {spec.name} image_shift {shift}", flush=True)
            e, st = gpu_and_oracle(gpu_engine_factory, spec, 41)
            e.set_pair_layout(layout)
            e.set_image_shift(shift)
            e.mc_steps(120)
            st.mc_steps(120)
            s = e.stats()
            # rotations of the one-residue molecules leave the energy
            # unchanged: the oracle's delta is exactly 0 there and the FP32
            # row's is a rounding residue of either sign, so only the split
            # of accepted moves between "dE < 0" and "Metropolis" may differ
            c, o = s["counters"], st.counters
            np.testing.assert_array_equal(c[:, 0] + c[:, 1], o[:, 0] + o[:, 1])
            np.testing.assert_array_equal(c[:, 2], o[:, 2])
            if spec.name != "tiny":
                np.testing.assert_array_equal(c, o)
            np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
            for r in range(spec.n_rep):
                np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
            e.close()


@pytest.mark.parametrize("layout", ["ilanes", "jlanes", "jlanes4"])
def test_tight_box_image_paths(gpu_engine_factory, layout):
    """A 46 A box holding a 120-residue molecule of bounding radius 21 A
    (above 0.45 L, so its rows take the per-pair rint image) and three small
    ones, whose rows need the per-axis minimum image on most visits: the
    oracle's trajectory."""
    from cgppd_amd import synth
    spec = synth.make_system("tightmc", [120, 8, 5, 3], 46.0, 6, 288.0, 450.0, 11, 2.0, 0.2)
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 5)
    e.set_pair_layout(layout)
    for chunk in range(2):
        e.mc_steps(40)
        st.mc_steps(40)
        s = e.stats()
        c, o = s["counters"], st.counters
        print(f"{layout} after {40 * (chunk + 1)} steps: gpu {c.sum(0)} oracle {o.sum(0)}", flush=True)
        np.testing.assert_array_equal(c[:, 0] + c[:, 1], o[:, 0] + o[:, 1])
        np.testing.assert_array_equal(c[:, 2], o[:, 2])
        np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-6)
        for r in range(spec.n_rep):
            np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
    e.close()


@pytest.mark.parametrize("n_rep", [64, 65, 130])
def test_row_kernel_replica_counts(gpu_engine_factory, n_rep):
    """The row kernel keeps the proposal headers of up to 64 replicas in LDS
    and reads them from memory beyond; its unit-offset scan takes 64
    replicas per pass.  Both sides of each boundary follow the oracle."""
    spec = small_spec(n_rep)
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 23)
    for layout in ("jlanes", "jlanes4"):
        print(f"n_rep {n_rep} layout {layout}", flush=True)
        e.set_pair_layout(layout)
        e.mc_steps(20)
        st.mc_steps(20)
        s = e.stats()
        np.testing.assert_array_equal(s["counters"], st.counters)
        np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
        for r in (0, 63, n_rep - 1):
            np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
    e.close()


def test_pair_layouts_agree_on_config1(gpu_engine_factory):
    """40-chain system: the two layouts make the same decisions and their
    running totals agree far below the parity bar."""
    from cgppd_amd import synth
    spec = synth.config(1)
    out = {}
    for layout in ("ilanes", "jlanes", "jlanes4"):
        e = gpu_engine_factory()
        e.load_spec(spec, seed=12)
        e.mc_prepare()
        e.set_pair_layout(layout)
        e.mc_steps(300)
        s = e.stats()
        out[layout] = (s, e.coordinates(5)[0])
        e.close()
    sa, pa = out["ilanes"]
    for other in ("jlanes", "jlanes4"):
        sb, pb = out[other]
        # molecules far from all others have |dE| near the FP32 rounding
        # level: the decision agrees, its "dE < 0" / "Metropolis" label may not
        ca, cb = sa["counters"], sb["counters"]
        np.testing.assert_array_equal(ca[:, 0] + ca[:, 1], cb[:, 0] + cb[:, 1])
        np.testing.assert_array_equal(ca[:, 2], cb[:, 2])
        np.testing.assert_array_equal(pa, pb)
        # the layouts use different FP32 forms of the pair term and different
        # summation orders: totals of ~0.1 kcal/mol agree to a few 1e-8
        np.testing.assert_allclose(sa["energy"], sb["energy"], rtol=2e-6, atol=1e-7)


def test_jlanes_rows_are_deterministic(gpu_engine_factory):
    """The j-lane kernel sums row partials in fixed point: two runs of the
    same configuration agree bit for bit."""
    from cgppd_amd import synth
    spec = synth.config(1)
    res = []
    for _ in range(2):
        e = gpu_engine_factory()
        e.load_spec(spec, seed=4)
        e.mc_prepare()
        e.set_pair_layout("jlanes")
        e.mc_steps(200)
        res.append(e.stats()["energy"].copy())
        e.close()
    np.testing.assert_array_equal(res[0], res[1])


def test_autotune_keeps_state(gpu_engine_factory):
    """Tuning times the candidates on the pending proposals without changing
    the MC state: the tuned engine follows the untuned engine's trajectory."""
    spec = small_spec()
    a, st = gpu_and_oracle(gpu_engine_factory, spec, 19)
    best = a.autotune(launches=3)
    print("autotune", best, flush=True)
    assert best["layout"] in ("ilanes", "jlanes", "jlanes4")
    assert 2 <= best["wg_per_cu"] <= 8 and best["ms"] > 0
    a.mc_steps(150)
    st.mc_steps(150)
    s = a.stats()
    np.testing.assert_array_equal(s["counters"], st.counters)
    np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(a.coordinates(r)[0], st.pos[r])
    a.close()



@pytest.mark.parametrize("groups,layout", [(2, "jlanes"), (3, "ilanes"), (4, "jlanes4")])
def test_replica_groups_match_oracle(gpu_engine_factory, groups, layout):
    """Replica groups on streams of their own: each replica keeps the
    oracle's trajectory (moves keyed by the replica's index in the
    context), exchanges see every group's energies, and a replayed graph
    (the groups' streams forked and joined inside it) equals direct launches."""
    spec = small_spec(n_rep=9)
    e, st = gpu_and_oracle(gpu_engine_factory, spec, 101)
    e.set_pair_layout(layout)
    e.set_replica_groups(groups)
    assert e.replica_groups() == groups
    b = gpu_engine_factory()
    b.load_spec(spec, seed=101)
    b.mc_prepare()
    b.set_pair_layout(layout)
    b.set_replica_groups(groups)
    for k in range(4):
        e.run(1, 40, use_graph=False)
        b.run(1, 40, use_graph=True)
        st.mc_steps(40)
        st.exchange()
        s = e.stats()
        print(f"{groups} groups, interval {k}: gpu {s['counters'].sum(0)} oracle {st.counters.sum(0)}", flush=True)
        np.testing.assert_array_equal(s["counters"], st.counters)
        np.testing.assert_array_equal(s["temperature"], st.temperature)
        np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-8)
    sb = b.stats()
    for key in ("energy", "temperature", "label", "counters", "steps", "exchange"):
        np.testing.assert_array_equal(s[key], sb[key])
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
        np.testing.assert_array_equal(b.coordinates(r)[0], st.pos[r])
    e.close()
    b.close()
