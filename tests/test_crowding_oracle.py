"""The oracle's repulsive-crowding branch (the reference's REPULSIVE_CROWDING,
src/Replica.cpp:416-483, crowderPairPotential at :416-419): closed forms on
two residues, pairs without a crowder unchanged, and the MC oracle's running
total against a fresh evaluation."""
import numpy as np
import pytest

import oracle


@pytest.fixture
def crowding():
    yield oracle.set_repulsive_crowding
    oracle.set_repulsive_crowding(None)


def force_field():
    from cgppd_amd import synth
    return synth.make_system("ff", [2, 2], 100.0, 1, 300.0, 300.0, 1, 2.0, 0.2).ff


def two_residues(r):
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]], dtype=np.float32)
    return np.array([0, 1, 2], dtype=np.int32), pos


@pytest.mark.parametrize("r", [4.0, 9.5, 14.9, 15.1, 30.0])
@pytest.mark.parametrize("first", [0, 1])
def test_two_residue_closed_form(crowding, r, first):
    ff = force_field()
    charged = np.nonzero(ff.charge)[0][:2]
    types = charged.astype(np.int32)
    mol_off, pos = two_residues(r)
    args = (mol_off, types, ff.charge, ff.sigma, ff.eps, pos, 100.0)
    off_ref = oracle.energy_ref(*args)
    crowding(first)
    lj, dh = oracle.energy_ref(*args)
    flj, fdh = oracle.energy_fp64(*args)
    rf = np.float32(pos[1, 0] - pos[0, 0])
    assert dh == 0.0 and fdh == 0.0
    if r < 15.0:
        np.testing.assert_allclose(lj, float((np.float32(6.0) / rf) ** np.float32(12.0)), rtol=2e-6)
        np.testing.assert_allclose(flj, (6.0 / float(rf)) ** 12, rtol=1e-12)
        assert off_ref[1] != 0.0 and abs(off_ref[0] - lj) > 1e-6 * abs(lj)
    else:
        assert lj == 0.0 and flj == 0.0


def test_pairs_without_crowder_unchanged(crowding):
    ff = force_field()
    # A = residues 0-1, B = residue 2, crowder C = residue 3, on a line
    mol_off = np.array([0, 2, 3, 4], dtype=np.int32)
    types = np.array([1, 4, 8, 3], dtype=np.int32)
    pos = np.array([[0, 0, 0], [3.8, 0, 0], [9.0, 0, 0], [-7.0, 0, 0]], dtype=np.float32)
    args = (ff.charge, ff.sigma, ff.eps, pos, 50.0)
    ab_off = oracle.energy_pair_ref(mol_off, 0, 1, types, *args)
    total_off = oracle.energy_fp64(mol_off, types, *args)
    crowding(3)
    assert oracle.energy_pair_ref(mol_off, 0, 1, types, *args) == ab_off
    ac = oracle.energy_pair_ref(mol_off, 0, 2, types, *args)
    bc = oracle.energy_pair_ref(mol_off, 1, 2, types, *args)
    assert ac[1] == 0.0 and bc[1] == 0.0
    np.testing.assert_allclose(ac[0], (6 / 7.0) ** 12 + (6 / 10.8) ** 12, rtol=2e-6)
    assert bc[0] == 0.0  # 16 A apart
    flj, fdh = oracle.energy_fp64(mol_off, types, *args)
    # (FP64 against the reference's float distances and powf)
    np.testing.assert_allclose(flj, ab_off[0] + ac[0], rtol=1e-5)
    np.testing.assert_allclose(fdh, ab_off[1], rtol=1e-5)
    # a first crowder residue past the last residue: no crowders, no change
    crowding(4)
    assert oracle.energy_fp64(mol_off, types, *args) == total_off


def test_mc_oracle_running_total(crowding):
    from cgppd_amd import synth
    spec = synth.make_system("crowdmc", [40, 25, 1, 1, 3, 1, 2, 1, 1, 1], 60.0, 4, 288.0, 450.0, 14, 2.0, 0.2)
    plain = oracle.Remc(spec, seed=5)
    crowding(int(spec.mol_off[2]))
    st = oracle.Remc(spec, seed=5)
    assert np.all(np.abs(st.energy - plain.energy) > 1e-3)
    st.mc_steps(300)
    for r in range(spec.n_rep):
        flj, fdh = oracle.energy_fp64(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps,
                                      st.pos[r], spec.box)
        want = oracle.total(flj, fdh)
        assert abs(st.energy[r] - want) <= 1e-6 * abs(want) + 1e-7, r
    assert st.counters.sum() == 300 * spec.n_rep
