"""Resident MC steps (CGPPD_LAYOUT_RESIDENT): one workgroup per replica
holds the replica in LDS and runs many MC steps per launch.  Against the
oracle's trajectories (the same Philox moves, so equal decisions give
bit-identical coordinates), with exchanges, graph replays and replica
counts on both sides of the grid's waves; systems outside the kernel's
scope fall back to the pair and finish kernels.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def small_spec(n_rep=8):
    from cgppd_amd import synth
    return synth.make_system("mcsmall", [22, 75, 40, 130, 18], 110.0, n_rep, 288.0, 450.0, 31, 2.0, 0.2)


def resident_engine(factory, spec, seed):
    e = factory()
    e.load_spec(spec, seed=seed)
    e.set_pair_layout("resident")
    e.mc_prepare()
    assert e.pair_layout() == "resident"
    return e


def assert_matches(e, st, spec, rtol=1e-5, labels=True):
    """labels=False: molecules far from all others have |dE| near the FP32
    rounding level, so a decision's "dE < 0" / "Metropolis" label may differ
    while the decision itself agrees."""
    s = e.stats()
    np.testing.assert_array_equal(s["steps"], st.step)
    if labels:
        np.testing.assert_array_equal(s["counters"], st.counters)
    else:
        c, o = s["counters"], st.counters
        np.testing.assert_array_equal(c[:, 0] + c[:, 1], o[:, 0] + o[:, 1])
        np.testing.assert_array_equal(c[:, 2], o[:, 2])
    np.testing.assert_allclose(s["energy"], st.energy, rtol=rtol, atol=1e-8)
    for r in range(spec.n_rep):
        pos, rel, cen = e.coordinates(r)
        np.testing.assert_array_equal(pos, st.pos[r])
        np.testing.assert_array_equal(rel, st.rel[r])
        np.testing.assert_array_equal(cen, st.center[r])


@pytest.mark.parametrize("n_rep", [1, 8, 65, 130])
def test_resident_trajectory_matches_oracle(gpu_engine_factory, n_rep):
    spec = small_spec(n_rep)
    e = resident_engine(gpu_engine_factory, spec, 77)
    st = oracle.Remc(spec, seed=77)
    done = 0
    for n in (60, 1, 0, 39):
        print(f"n_rep {n_rep}: steps {done}..{done + n}", flush=True)
        e.mc_steps(n)
        st.mc_steps(n)
        done += n
        assert_matches(e, st, spec)
    # the running totals are the matrix sums of the coordinates' energies
    _, _, full = e.energy_all()
    np.testing.assert_allclose(e.stats()["energy"], full, rtol=1e-5, atol=1e-8)
    e.close()


def test_resident_graph_with_exchanges_matches_oracle(gpu_engine_factory):
    """Replayed intervals (MC steps and an exchange) equal direct launches
    and the oracle's exchange rounds; the pair and finish kernels continue
    from the resident kernel's state."""
    spec = small_spec(9)
    a = resident_engine(gpu_engine_factory, spec, 5)
    b = resident_engine(gpu_engine_factory, spec, 5)
    st = oracle.Remc(spec, seed=5)
    for k in range(3):
        a.run(1, 50, use_graph=True)
        b.run(1, 50, use_graph=False)
        st.mc_steps(50)
        st.exchange()
        s = a.stats()
        print(f"interval {k}: gpu {s['counters'].sum(0)} oracle {st.counters.sum(0)}", flush=True)
        np.testing.assert_array_equal(s["temperature"], st.temperature)
        assert_matches(a, st, spec)
    sb = b.stats()
    for key in ("energy", "temperature", "label", "counters", "steps", "exchange"):
        np.testing.assert_array_equal(a.stats()[key], sb[key])
    # switch to the j-lane kernels mid-run: they pick up the pending proposal
    a.set_pair_layout("jlanes")
    a.mc_steps(40)
    st.mc_steps(40)
    assert_matches(a, st, spec)
    a.close()
    b.close()


def test_resident_pair_2x76_matches_oracle(gpu_engine_factory):
    """The two-chain system of the reference's published benchmark shape
    (2 x 76 residues, 20 replicas)."""
    from cgppd_amd import synth
    spec = synth.config(0)
    e = resident_engine(gpu_engine_factory, spec, 9)
    st = oracle.Remc(spec, seed=9)
    for _ in range(2):
        e.mc_steps(150)
        st.mc_steps(150)
        assert_matches(e, st, spec)
    e.close()


def test_resident_one_residue_molecules_and_single_molecule(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_system("ones", [1, 1, 3, 1, 40, 1, 2], 60.0, 4, 288.0, 450.0, 12, 2.0, 0.2)
    e = resident_engine(gpu_engine_factory, spec, 3)
    st = oracle.Remc(spec, seed=3)
    e.mc_steps(120)
    st.mc_steps(120)
    assert_matches(e, st, spec, labels=False)
    e.close()
    one = synth.make_system("one", [60], 80.0, 3, 288.0, 450.0, 5, 2.0, 0.2)
    e = resident_engine(gpu_engine_factory, one, 5)
    st = oracle.Remc(one, seed=5)
    e.mc_steps(30)
    st.mc_steps(30)
    assert_matches(e, st, one, labels=False)
    e.close()


def test_resident_falls_back_outside_its_scope(gpu_engine_factory):
    """Flexible molecules, repulsive crowding or replicas too large for LDS
    keep the pair and finish kernels."""
    from cgppd_amd import synth
    flex = synth.make_linked_system("resflex", 2, 30, 6, 0.2, 120.0, 2, 288.0, 450.0, 41)
    e = gpu_engine_factory()
    e.load_spec(flex, seed=41)
    e.set_pair_layout("resident")
    e.mc_prepare()
    assert e.pair_layout() == "ilanes"
    e.mc_steps(5)
    e.close()
    big = synth.make_system("big", [300] * 5, 200.0, 2, 288.0, 450.0, 8, 2.0, 0.2)
    e = gpu_engine_factory()
    e.load_spec(big, seed=8)
    e.set_pair_layout("resident")
    e.mc_prepare()
    assert e.pair_layout() == "jlanes"
    st = oracle.Remc(big, seed=8)
    e.mc_steps(10)
    st.mc_steps(10)
    assert_matches(e, st, big)
    e.close()


def test_resident_rows_of_more_than_64_molecules(gpu_engine_factory):
    """90 molecules: the matrix row of the moved molecule spans several
    entries per lane, read before any wave applies the step."""
    from cgppd_amd import synth
    spec = synth.make_system("many", [3, 5, 2, 8] * 22 + [4, 6], 90.0, 6, 288.0, 450.0, 17, 2.0, 0.2)
    assert spec.n_mol > 64
    e = resident_engine(gpu_engine_factory, spec, 17)
    st = oracle.Remc(spec, seed=17)
    for _ in range(2):
        e.mc_steps(60)
        st.mc_steps(60)
        assert_matches(e, st, spec, labels=False)
    _, _, full = e.energy_all()
    np.testing.assert_allclose(e.stats()["energy"], full, rtol=1e-5, atol=1e-8)
    e.close()


def test_resident_timing_keeps_the_resident_kernel(gpu_engine_factory):
    """Timed steps still run the resident kernel (the oracle's trajectory)
    and report its span per MC step."""
    spec = small_spec(4)
    e = resident_engine(gpu_engine_factory, spec, 21)
    st = oracle.Remc(spec, seed=21)
    e.set_timing(True)
    e.mc_steps(70)
    ms, launches = e.last_kernel_ms()
    e.set_timing(False)
    st.mc_steps(70)
    assert e.pair_layout() == "resident"
    assert launches == 1 and 0.0 < ms < 1.0
    assert_matches(e, st, spec)
    e.close()
