"""Speculative windows of MC steps (cgppd_set_window) against the oracle.

A window evaluates the delta-E rows of consecutive steps that move distinct
molecules in one launch, against the window's starting state, plus the
energies between the window's proposals; the decisions are then taken in
step order.  The chain must be the one-step-per-launch chain: the same
Philox moves and the same decisions, so the coordinates agree bit for bit
with the oracle's, and the running totals agree to the FP32-vs-FP64 bar.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def spec_of(lengths, n_rep, box=110.0, seed=31):
    from cgppd_amd import synth
    return synth.make_system("win", lengths, box, n_rep, 288.0, 450.0, seed, 2.0, 0.2)


def window_engine(factory, spec, seed, slots, layout="jlanes"):
    e = factory()
    e.load_spec(spec, seed=seed)
    e.set_pair_layout(layout)
    e.set_window(slots)
    e.mc_prepare()
    assert e.window() == slots
    return e


def assert_matches(e, st, spec, rtol=1e-5):
    """Accept/reject agree step for step (the "dE < 0" / "Metropolis" label of
    a decision on a molecule far from all others, |dE| at the FP32 rounding
    level, may differ: the delta-E is summed in another FP64 order)."""
    s = e.stats()
    np.testing.assert_array_equal(s["steps"], st.step)
    c, o = s["counters"], st.counters
    np.testing.assert_array_equal(c[:, 0] + c[:, 1], o[:, 0] + o[:, 1])
    np.testing.assert_array_equal(c[:, 2], o[:, 2])
    np.testing.assert_allclose(s["energy"], st.energy, rtol=rtol, atol=1e-8)
    for r in range(spec.n_rep):
        pos, rel, cen = e.coordinates(r)
        np.testing.assert_array_equal(pos, st.pos[r])
        np.testing.assert_array_equal(rel, st.rel[r])
        np.testing.assert_array_equal(cen, st.center[r])


SYSTEMS = {
    # five molecules: repeats end about a third of the windows of 4
    "five": ([22, 75, 40, 130, 18], 8),
    # one molecule: every window holds one step
    "one": ([60], 4),
    # two molecules: windows of one or two steps
    "two": ([76, 76], 6),
    # twelve ragged molecules with one-residue ones; 40 replicas put the
    # row kernel's proposal headers outside LDS for 4 and 8 slots
    "twelve": ([1, 33, 140, 7, 1, 64, 90, 12, 250, 3, 45, 28], 40),
}


@pytest.mark.parametrize("slots", [2, 4, 8])
@pytest.mark.parametrize("system", list(SYSTEMS))
def test_window_trajectory_matches_oracle(gpu_engine_factory, system, slots):
    lengths, n_rep = SYSTEMS[system]
    spec = spec_of(lengths, n_rep)
    e = window_engine(gpu_engine_factory, spec, 77, slots)
    st = oracle.Remc(spec, seed=77)
    done = 0
    for n in (60, 1, 0, 39, 7):
        print(f"{system} W={slots}: steps {done}..{done + n}", flush=True)
        e.mc_steps(n)
        st.mc_steps(n)
        done += n
        assert_matches(e, st, spec)
    _, _, full = e.energy_all()
    np.testing.assert_allclose(e.stats()["energy"], full, rtol=1e-5, atol=1e-8)
    e.close()


def test_window_jlanes4_and_many_replicas(gpu_engine_factory):
    """Four residues per lane, and 130 replicas (520 window slots: the
    headers in HBM and a multi-pass unit scan)."""
    spec = spec_of([22, 75, 40, 130, 18, 9, 51], 130)
    e = window_engine(gpu_engine_factory, spec, 5, 4, layout="jlanes4")
    st = oracle.Remc(spec, seed=5)
    for n in (30, 13):
        print(f"130 replicas: {n} steps", flush=True)
        e.mc_steps(n)
        st.mc_steps(n)
        assert_matches(e, st, spec)
    e.close()


def test_window_graph_with_exchanges(gpu_engine_factory):
    """Replayed intervals (the run's start and the window launches in the
    graph, extra windows and the exchange launched after it) equal direct
    launches and the oracle's exchange rounds."""
    spec = spec_of([22, 75, 40, 130, 18], 9)
    a = window_engine(gpu_engine_factory, spec, 5, 4)
    b = window_engine(gpu_engine_factory, spec, 5, 4)
    st = oracle.Remc(spec, seed=5)
    for k in range(4):
        print(f"interval {k}", flush=True)
        a.run(1, 50, use_graph=True)
        b.run(1, 50, use_graph=False)
        st.mc_steps(50)
        st.exchange()
        sa, sb = a.stats(), b.stats()
        for key in ("steps", "counters", "temperature", "label", "exchange"):
            np.testing.assert_array_equal(sa[key], sb[key])
        np.testing.assert_array_equal(sa["energy"], sb["energy"])
        np.testing.assert_array_equal(sa["temperature"], st.temperature)
        assert_matches(a, st, spec)
    a.run(3, 50, use_graph=True)
    b.run(3, 50, use_graph=False)
    np.testing.assert_array_equal(a.stats()["energy"], b.stats()["energy"])
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(a.coordinates(r)[0], b.coordinates(r)[0])
    a.close()
    b.close()


def test_window_equals_one_step_per_launch_crowded40(gpu_engine_factory):
    """The 40-chain box (32 replicas): windows of 2, 4 and 8 against one MC
    step per launch pair over 600 steps with exchanges (same decisions and
    coordinates, totals to the FP32 rounding of the pair terms), and the
    oracle over the first 12 steps."""
    from cgppd_amd import synth
    spec = synth.config(1)
    ref = window_engine(gpu_engine_factory, spec, 9, 1)
    st = oracle.Remc(spec, seed=9)
    engines = {w: window_engine(gpu_engine_factory, spec, 9, w) for w in (2, 4, 8)}
    ref.mc_steps(12)
    st.mc_steps(12)
    for w, e in engines.items():
        e.mc_steps(12)
        assert_matches(e, st, spec)
    ref.run(3, 196, use_graph=True)
    s0 = ref.stats()
    for w, e in engines.items():
        print(f"crowded40 W={w}", flush=True)
        e.run(3, 196, use_graph=True)
        s = e.stats()
        np.testing.assert_array_equal(s["steps"], s0["steps"])
        c, o = s["counters"], s0["counters"]
        np.testing.assert_array_equal(c[:, 2], o[:, 2])
        np.testing.assert_array_equal(s["temperature"], s0["temperature"])
        for r in range(spec.n_rep):
            np.testing.assert_array_equal(e.coordinates(r)[0], ref.coordinates(r)[0])
        # the entries between two accepted proposals of a window are their
        # own FP32 sum, not part of a row: FP32 rounding of the pair terms
        np.testing.assert_allclose(s["energy"], s0["energy"], rtol=1e-6, atol=1e-7)
        _, _, full = e.energy_all()
        np.testing.assert_allclose(s["energy"], full, rtol=1e-5, atol=1e-8)
        e.close()
    ref.close()


def test_window_layout_switches(gpu_engine_factory):
    """Windows, one step per launch pair, other window sizes and the
    resident kernel, switched between calls, continue one chain."""
    spec = spec_of([22, 75, 40, 130, 18], 6)
    e = window_engine(gpu_engine_factory, spec, 3, 4)
    st = oracle.Remc(spec, seed=3)
    plan = [(4, "jlanes", 25), (1, "jlanes", 9), (8, "jlanes4", 31), ("resident", None, 17), (2, "jlanes", 20)]
    for slots, layout, n in plan:
        print(f"switch to {slots} {layout}: {n} steps", flush=True)
        if slots == "resident":
            e.set_pair_layout("resident")
        else:
            e.set_pair_layout(layout)
            e.set_window(slots)
            assert e.window() == slots
        e.mc_steps(n)
        st.mc_steps(n)
        assert_matches(e, st, spec)
    e.close()


def test_window_autotune_leaves_state(gpu_engine_factory):
    """The autotuner times window launches on the laid-out window and
    leaves the chain unchanged."""
    spec = spec_of([22, 75, 40, 130, 18, 64, 80], 8, box=130.0)
    e = window_engine(gpu_engine_factory, spec, 21, 4)
    st = oracle.Remc(spec, seed=21)
    e.mc_steps(20)
    st.mc_steps(20)
    tuned = e.autotune(launches=3)
    print(f"tuned {tuned}", flush=True)
    e.mc_steps(40)
    st.mc_steps(40)
    assert_matches(e, st, spec)
    e.close()


def test_window_off_where_not_served(gpu_engine_factory):
    """Replica groups and the i-lane layout run one step per launch pair
    (flexible systems always run the i-lane layout)."""
    spec = spec_of([22, 75, 40, 130, 18], 4)
    e = window_engine(gpu_engine_factory, spec, 1, 4)
    e.set_pair_layout("ilanes")
    assert e.window() == 1
    e.set_pair_layout("jlanes")
    assert e.window() == 4
    e.set_replica_groups(2)
    assert e.window() == 1
    e.set_replica_groups(1)
    assert e.window() == 4
    from cgppd_amd.engine import CgppdError
    with pytest.raises(CgppdError):
        e.set_window(3)
    e.close()
