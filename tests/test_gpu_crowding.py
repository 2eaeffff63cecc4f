"""Repulsive crowding (the reference's REPULSIVE_CROWDING, src/Replica.cpp:
416-483, src/CudaFunctions.cu:492-500) on the device against the oracle:
full energies, and MC trajectories in every pair-kernel layout."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def crowded_spec(n_rep=6):
    """Two chains, then eight crowders of 1-3 residues in a 60 A box (41
    crowder pairs within 15 A)."""
    from cgppd_amd import synth
    return synth.make_system("crowdmc", [40, 25, 1, 1, 3, 1, 2, 1, 1, 1], 60.0, n_rep, 288.0, 450.0, 14, 2.0, 0.2), 2


@pytest.fixture
def crowding():
    yield oracle.set_repulsive_crowding
    oracle.set_repulsive_crowding(None)


def crowded_engine(factory, spec, n_nc, seed):
    e = factory()
    e.load_spec(spec, seed=seed, n_noncrowder=n_nc)
    e.set_repulsive_crowding(True)
    e.mc_prepare()
    return e


def fp64_total(spec, pos):
    flj, fdh = oracle.energy_fp64(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps, pos,
                                  spec.box)
    return oracle.total(flj, fdh)


def test_crowded_energy_matches_oracle(gpu_engine_factory, crowding):
    spec, n_nc = crowded_spec()
    e = gpu_engine_factory()
    e.load_spec(spec, seed=3, n_noncrowder=n_nc)
    e.mc_prepare()
    plain = e.stats()["energy"].copy()
    e.close()
    crowding(int(spec.mol_off[n_nc]))
    e = crowded_engine(gpu_engine_factory, spec, n_nc, 3)
    s = e.stats()
    _, _, full = e.energy_all()
    for r in range(spec.n_rep):
        want = fp64_total(spec, e.coordinates(r)[0])
        assert abs(s["energy"][r] - want) <= 1e-5 * abs(want) + 1e-6, r
        assert abs(full[r] - want) <= 1e-5 * abs(want) + 1e-6, r
    assert np.all(np.abs(s["energy"] - plain) > 1e-3)
    e.close()


@pytest.mark.parametrize("layout", ["ilanes", "jlanes", "jlanes4"])
def test_crowded_mc_matches_oracle(gpu_engine_factory, crowding, layout):
    """Crowder moves (every pair a crowder pair), chain moves against blocks
    holding chains and crowders: the oracle's trajectory."""
    spec, n_nc = crowded_spec()
    crowding(int(spec.mol_off[n_nc]))
    e = crowded_engine(gpu_engine_factory, spec, n_nc, 5)
    st = oracle.Remc(spec, seed=5)
    e.set_pair_layout(layout)
    for chunk in range(2):
        e.mc_steps(60)
        st.mc_steps(60)
        s = e.stats()
        c, o = s["counters"], st.counters
        print(f"{layout} after {60 * (chunk + 1)} steps: gpu {c.sum(0)} oracle {o.sum(0)}", flush=True)
        np.testing.assert_array_equal(c[:, 0] + c[:, 1], o[:, 0] + o[:, 1])
        np.testing.assert_array_equal(c[:, 2], o[:, 2])
        np.testing.assert_allclose(s["energy"], st.energy, rtol=1e-5, atol=1e-6)
        for r in range(spec.n_rep):
            np.testing.assert_array_equal(e.coordinates(r)[0], st.pos[r])
    e.close()


@pytest.mark.parametrize("layout", ["ilanes", "jlanes"])
def test_crowded_blocks_keep_totals_consistent(gpu_engine_factory, crowding, layout):
    """Four chains and 250 crowders: j-blocks of chains only, of crowders
    only and one of both.  After MC steps the running totals equal a fresh
    full evaluation and the FP64 oracle's energy of the coordinates."""
    from cgppd_amd import synth
    spec = synth.make_system("crowd250", [150, 120, 90, 60] + [1] * 200 + [3] * 50, 120.0, 4, 288.0, 450.0, 7, 2.0,
                             0.2)
    n_nc = 4
    crowding(int(spec.mol_off[n_nc]))
    e = crowded_engine(gpu_engine_factory, spec, n_nc, 9)
    e.set_pair_layout(layout)
    e.mc_steps(300)
    s = e.stats()
    _, _, full = e.energy_all()
    np.testing.assert_allclose(s["energy"], full, rtol=1e-5, atol=1e-6)
    for r in range(spec.n_rep):
        want = fp64_total(spec, e.coordinates(r)[0])
        assert abs(full[r] - want) <= 1e-5 * abs(want) + 1e-6, r
    assert s["counters"].sum() == 300 * spec.n_rep
    e.close()


def test_flexible_crowder_rejected_in_either_order(gpu_engine_factory):
    """Repulsive crowding treats crowders as rigid bodies: a crowder made
    flexible is refused whether its topology or the switch comes first; a
    flexible non-crowder is accepted."""
    from cgppd_amd import engine, synth
    spec = synth.make_linked_system("crowdflex", 2, 30, 6, 0.2, 120.0, 2, 288.0, 450.0, 41, extra_rigid=(25, 40))
    m0 = slice(int(spec.mol_off[0]), int(spec.mol_off[1]))
    topo = (spec.chain_code[m0], spec.chain_index[m0], False, spec.segments[0])

    def fresh(n_nc):
        e = gpu_engine_factory()
        e.set_force_field(spec.ff.eps, spec.ff.sigma, spec.ff.charge)
        e.set_box(spec.box)
        e.set_system(spec.mol_off, spec.types, n_nc)
        e.set_torsion_table(spec.torsion_v, spec.torsion_delta)
        return e

    # molecule 0 (flexible) is a crowder when n_noncrowder = 0
    e = fresh(0)
    e.set_topology(0, *topo)
    with pytest.raises(engine.CgppdError, match=r"\(-1\).*rigid crowders"):
        e.set_repulsive_crowding(True)
    e.close()
    e = fresh(0)
    e.set_repulsive_crowding(True)
    with pytest.raises(engine.CgppdError, match=r"\(-1\).*rigid crowders"):
        e.set_topology(0, *topo)
    e.close()
    # a flexible non-crowder with rigid crowders is fine in both orders
    for first_topology in (True, False):
        e = fresh(1)
        if first_topology:
            e.set_topology(0, *topo)
        e.set_repulsive_crowding(True)
        if not first_topology:
            e.set_topology(0, *topo)
        e.close()


def test_crowding_switch_invalidates_mc_state(gpu_engine_factory, crowding):
    """Toggling the switch after cgppd_mc_prepare would mix two energy
    models in the rows and the matrix: MC calls then fail until the matrix
    is rebuilt, after which the totals follow the new model."""
    from cgppd_amd import engine
    spec, n_nc = crowded_spec()
    e = gpu_engine_factory()
    e.load_spec(spec, seed=4, n_noncrowder=n_nc)
    e.mc_prepare()
    e.mc_steps(5)
    e.set_repulsive_crowding(True)
    with pytest.raises(engine.CgppdError, match=r"\(-4\)"):
        e.mc_steps(1)
    with pytest.raises(engine.CgppdError, match=r"\(-4\)"):
        e.run(1, 10)
    crowding(int(spec.mol_off[n_nc]))
    e.mc_prepare()
    e.mc_steps(20)
    s = e.stats()
    for r in range(spec.n_rep):
        want = fp64_total(spec, e.coordinates(r)[0])
        assert abs(s["energy"][r] - want) <= 1e-5 * abs(want) + 1e-6, r
    e.close()
