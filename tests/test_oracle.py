"""Pins the CPU oracle (oracle/) before anything is compared against it.

Checks: the reference's own known-answer tests (ten conformations, Kim 2008
comparison), bit-exact agreement with oracle/_ref (built from the reference
sources) on committed golden vectors, Random123 known-answer vectors for
Philox, closed forms for two-residue systems, rigid-body invariants, and the
Metropolis / exchange acceptance laws.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_SYSTEMS, KIM_RESULTS, REFERENCE_CONFORMATIONS, golden_system


def test_constants_match_reference_build(golden):
    rt, dhc, r0 = oracle.constants()
    assert np.float32(rt) == golden["consts"][0]
    assert np.float32(dhc) == golden["consts"][1]
    assert np.float32(r0) == golden["consts"][2]


def test_reference_conformations_exact(conformations, mj_ff):
    """tests/testReferenceConformations.cpp:83-98: almost_equal (1e-5 raw)."""
    rt, dhc, _ = oracle.constants()
    for (off, types, pos), (elj, edh) in zip(conformations, REFERENCE_CONFORMATIONS):
        lj, dh = oracle.energy_ref(off, types, mj_ff.charge, mj_ff.sigma, mj_ff.eps, pos, 1000.0)
        assert abs(lj - elj / np.float32(rt)) < 1e-5
        assert abs(dh - edh / np.float32(dhc)) < 1e-5
        # E(a, b) of the two molecules is the same sum
        plj, pdh = oracle.energy_pair_ref(off, 0, 1, types, mj_ff.charge, mj_ff.sigma, mj_ff.eps, pos, 1000.0)
        assert plj == lj and pdh == dh


def test_kim_comparison(conformations, mj_ff):
    """tests/testReferenceConformations.cpp:109-189.  The expected mean/max
    relative errors there were produced with the conversion fudge factors of
    :113-114 (LJ x 0.6/RT_to_kcalmol, DH x 1.00465), which the current source
    comments out; with them every figure is reproduced to the printed digits,
    without them the LJ errors sit at the 1.4 % RT-conversion offset."""
    rt, dhc, _ = oracle.constants()
    lj_fudge, dh_fudge = 0.6 / np.float32(rt), 1.00465
    e = {"lj": [], "dh": [], "tot": []}
    for (off, types, pos), (klj, kdh) in zip(conformations, KIM_RESULTS):
        lj, dh = oracle.energy_ref(off, types, mj_ff.charge, mj_ff.sigma, mj_ff.eps, pos, 1000.0)
        LJ, DH = lj * np.float32(rt) * lj_fudge, dh * np.float32(dhc) * dh_fudge
        e["lj"].append(abs(klj - LJ) / abs(klj))
        e["dh"].append(abs(kdh - DH) / abs(kdh))
        e["tot"].append(abs((klj + kdh) - (LJ + DH)) / abs(klj + kdh))
    approx = lambda v: pytest.approx(v, abs=5e-7)  # noqa: E731  the expected values carry 6 decimals
    assert np.mean(e["lj"]) == approx(0.001505)
    assert np.mean(e["dh"]) == approx(0.000078)
    assert np.mean(e["tot"]) == approx(0.001073)
    assert max(e["lj"]) == approx(0.002866)
    assert max(e["dh"]) == approx(0.000250)
    assert max(e["tot"]) == approx(0.002202)


def test_conformations_match_reference_build(golden, conformations, mj_ff):
    for i, (off, types, pos) in enumerate(conformations, start=1):
        lj, dh = oracle.energy_ref(off, types, mj_ff.charge, mj_ff.sigma, mj_ff.eps, pos, 1000.0)
        ref = golden[f"conf{i}_energy"]
        assert lj == ref[0] and dh == ref[1]
        assert oracle.total(lj, dh) == ref[2]


@pytest.mark.parametrize("name", GOLDEN_SYSTEMS)
def test_synthetic_systems_bit_exact(golden, name):
    s = golden_system(golden, name)
    lj, dh = oracle.energy_ref(s["mol_off"], s["types"], s["charge"], s["sigma"], s["eps"], s["pos"], s["box"])
    assert lj == s["energy"][0]
    assert dh == s["energy"][1]
    mat = oracle.energy_matrix_ref(s["mol_off"], s["types"], s["charge"], s["sigma"], s["eps"], s["pos"], s["box"])
    np.testing.assert_array_equal(mat, s["matrix"])


@pytest.mark.parametrize("name", GOLDEN_SYSTEMS)
def test_fp64_oracle_agrees(golden, name):
    s = golden_system(golden, name)
    lj, dh = oracle.energy_fp64(s["mol_off"], s["types"], s["charge"], s["sigma"], s["eps"], s["pos"], s["box"])
    assert lj == pytest.approx(s["energy"][0], rel=2e-6)
    assert dh == pytest.approx(s["energy"][1], rel=2e-6, abs=1e-12)


def _two_residue(ti, tj, r, eps, sigma, charge, box=0.0):
    off = np.array([0, 1, 2], dtype=np.int32)
    types = np.array([ti, tj], dtype=np.int32)
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]], dtype=np.float32)
    return oracle.energy_fp64(off, types, charge, sigma, eps, pos, box), oracle.energy_ref(
        off, types, charge, sigma, eps, pos, box)


def test_two_residue_closed_form():
    """Kim & Hummer 2008 eqs. 2-6: attractive and repulsive branches, the
    repulsive switch point, Debye-Hueckel."""
    rng = np.random.default_rng(0)
    eps = rng.uniform(-7, -1, (20, 20)).astype(np.float32)
    eps = np.triu(eps) + np.triu(eps, 1).T
    eps[2, 3] = eps[3, 2] = 1.5  # E_ij > 0: repulsive pair
    sigma = rng.uniform(4.6, 7.4, 20).astype(np.float32)
    charge = np.zeros(20, dtype=np.float32)
    charge[1], charge[3] = 1.0, -1.0
    lam, e0 = float(np.float32(0.159)), float(np.float32(-2.27))
    for ti, tj in [(0, 5), (1, 3), (2, 3)]:
        E = lam * (float(eps[ti, tj]) - e0)
        s = 0.5 * (float(sigma[ti]) + float(sigma[tj]))
        for r in (0.8 * s, s, 2 ** (1 / 6) * s * 0.999, 2 ** (1 / 6) * s * 1.001, 1.7 * s, 12.0, 40.0):
            r = float(np.float32(r))
            t = (s / r) ** 6
            lj = -4 * E * t * (t - 1)
            if E > 0 and r < 2 ** (1 / 6) * s:
                lj = -lj + 2 * E
            dh = float(charge[ti]) * float(charge[tj]) * math.exp(-r / 10.0) / r
            (l64, d64), (lref, dref) = _two_residue(ti, tj, r, eps, sigma, charge)
            assert l64 == pytest.approx(lj, rel=1e-12, abs=1e-15)
            assert d64 == pytest.approx(dh, rel=1e-12, abs=1e-15)
            # at r = sigma the term cancels to 0: compare on the scale 4|E|
            assert lref == pytest.approx(lj, rel=1e-5, abs=4e-6 * abs(E))
            assert dref == pytest.approx(dh, rel=1e-5, abs=1e-9)


def test_repulsive_branch_is_continuous():
    eps = np.full((20, 20), 3.0, dtype=np.float32)
    sigma = np.full(20, 5.0, dtype=np.float32)
    charge = np.zeros(20, dtype=np.float32)
    rc = 2 ** (1 / 6) * 5.0
    (a, _), _ = _two_residue(0, 0, rc * (1 - 1e-6), eps, sigma, charge)
    (b, _), _ = _two_residue(0, 0, rc * (1 + 1e-6), eps, sigma, charge)
    assert a == pytest.approx(b, rel=1e-4)


def test_minimum_image_invariance():
    """Translating one molecule by a box vector leaves every energy unchanged."""
    from cgppd_amd import synth
    spec = synth.make_system("t", [30, 40, 25], 120.0, 1, 300.0, 300.0, 3)
    pos = spec.positions()
    base = oracle.energy_fp64(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps, pos, spec.box)
    moved = pos.copy()
    moved[30:70] += np.float32(spec.box) * np.array([1, -2, 0], dtype=np.float32)
    shifted = oracle.energy_fp64(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps, moved,
                                 spec.box)
    assert shifted[0] == pytest.approx(base[0], rel=1e-5)
    assert shifted[1] == pytest.approx(base[1], rel=1e-5)


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_rotation_matches_reference_quaternion(golden):
    """Quaternion(angle, axis).normalize().rotateVector(v) of the reference
    build, bit for bit."""
    for row, want in zip(golden["rot_in"], golden["rot_out"]):
        angle = float(row[0])
        _, rel, _ = oracle.apply_move(1, row[1:4], 0.0, math.cos(angle / 2.0), math.sin(angle / 2.0), 1000.0,
                                      [0, 0, 0], np.float32(row[4:7]))
        np.testing.assert_array_equal(rel[0], want)


def test_translation_normalise_matches_reference(golden):
    """Vector3f::normalize of the reference build, bit for bit (a translation
    by step 1 from the origin is the normalised vector)."""
    for v, want in zip(golden["nrm_in"], golden["nrm_out"]):
        c, _, _ = oracle.apply_move(0, v.astype(np.float64), 1.0, 1.0, 0.0, 1e30, [0, 0, 0],
                                    np.zeros((1, 3), np.float32))
        np.testing.assert_array_equal(c, want)


def test_rigid_moves_preserve_shape():
    rng = np.random.default_rng(1)
    rel = rng.uniform(-15, 15, (50, 3)).astype(np.float32)
    d0 = np.linalg.norm(rel[:, None] - rel[None], axis=2)
    c_in = np.array([10.0, -20.0, 30.0], dtype=np.float32)
    for rot in (0, 1):
        axis = rng.uniform(-0.5, 0.5, 3)
        c, r2, p2 = oracle.apply_move(rot, axis, 2.0, math.cos(0.1), math.sin(0.1), 150.0, c_in, rel)
        d1 = np.linalg.norm(p2[:, None] - p2[None], axis=2)
        np.testing.assert_allclose(d1, d0, atol=2e-4)
        if rot:
            np.testing.assert_array_equal(c, c_in)  # rotation keeps the centre
        else:
            assert np.linalg.norm(c - c_in) == pytest.approx(2.0, rel=1e-6)
            np.testing.assert_array_equal(r2, rel)


def test_translation_wraps_with_fmod():
    """src/Molecule.cpp:234-241: fmod keeps the sign of the centre."""
    c, _, _ = oracle.apply_move(0, [0.5, 0.0, 0.0], 3.0, 1.0, 0.0, 100.0, [99.0, -99.0, 0.0],
                                np.zeros((1, 3), np.float32))
    assert c[0] == pytest.approx(2.0)
    assert c[1] == pytest.approx(-99.0)


def test_proposal_streams_are_uniform():
    seed, n_mol = 12345, 7
    mols, rots, us = [], [], []
    for step in range(4000):
        m, rot, axis, u = oracle.proposal(seed, 3, step, n_mol)
        mols.append(m)
        rots.append(rot)
        us.append(u)
        assert all(-0.5 <= a <= 0.5 for a in axis)
    counts = np.bincount(mols, minlength=n_mol)
    assert counts.min() > 4000 / n_mol * 0.8
    assert 0.45 < np.mean(rots) < 0.55
    assert 0.47 < np.mean(us) < 0.53


def _tiny_remc(n_rep=4, seed=9):
    from cgppd_amd import synth
    spec = synth.make_system("tiny", [12, 20, 16], 60.0, n_rep, 288.0, 450.0, 21, 2.0, 0.2)
    return spec, oracle.Remc(spec, seed=seed)


def test_delta_e_bookkeeping_matches_full_recompute():
    spec, st = _tiny_remc()
    st.mc_steps(200)
    for r in range(spec.n_rep):
        lj, dh = oracle.energy_ref(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps, st.pos[r],
                                   spec.box)
        assert st.energy[r] == pytest.approx(oracle.total(lj, dh), rel=1e-9, abs=1e-9)
        mat = oracle.energy_matrix_ref(spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps,
                                       st.pos[r], spec.box)
        np.testing.assert_allclose(st.emat[r], mat, rtol=1e-12, atol=1e-12)
        # positions stay rel + centre
        for m in range(spec.n_mol):
            a, b = spec.mol_off[m], spec.mol_off[m + 1]
            np.testing.assert_array_equal(st.pos[r, a:b], st.rel[r, a:b] + st.center[r, m])
    assert st.counters.sum() == 200 * spec.n_rep
    assert (st.step == 200).all()


def test_reference_loop_matches_delta_e_loop_statistically():
    """The CPU baseline (full recompute, float delta) and the delta-E loop use
    the same proposals, so their first decisions coincide."""
    spec, a = _tiny_remc(n_rep=2, seed=4)
    _, b = _tiny_remc(n_rep=2, seed=4)
    a.mc_steps(50)
    b.reference_run(50, 2)
    np.testing.assert_array_equal(a.counters, b.counters)
    np.testing.assert_allclose(a.energy, b.energy, rtol=1e-6, atol=1e-6)


def test_exchange_acceptance_law():
    """src/Simulation.cpp:780-782: P(swap) = min(1, exp(delta))."""
    from cgppd_amd import synth
    spec = synth.make_system("x", [10, 10], 80.0, 2, 300.0, 400.0, 5)
    st = oracle.Remc(spec, seed=77)
    T = st.temperature.copy()
    st.energy[:] = [-1.0, -2.5]
    delta = (1 / T[0] - 1 / T[1]) * (st.energy[0] - st.energy[1]) * float(np.float32(4184.0) / np.float32(8.3144621))
    p = min(1.0, math.exp(delta))
    n, swaps = 4000, 0
    for k in range(n):
        st.temperature[:] = T
        st.rep_at_rank[:] = [0, 1]
        st.exchange_round[0] = 2 * k  # even rounds test pair (0, 1)
        st.exchange()
        swaps += st.temperature[0] != T[0]
    assert abs(swaps / n - p) < 4 * math.sqrt(p * (1 - p) / n) + 1e-3


def test_exchange_swaps_temperature_not_coordinates():
    spec, st = _tiny_remc(n_rep=4)
    pos0 = st.pos.copy()
    st.energy[:] = [5.0, -5.0, 5.0, -5.0]  # favourable swaps for pairs (0,1), (2,3)
    st.exchange()
    np.testing.assert_array_equal(st.pos, pos0)
    assert sorted(st.temperature) == sorted(spec.temperatures())
    assert st.exchange_stats[0] == 2
    assert list(st.rank_of[st.rep_at_rank]) == [0, 1, 2, 3]


def test_lean_sampler_energies_match_fp64(golden):
    """orc_energy_fast (the lean FP64 pair term of the long-run sampler)
    equals the FP64 oracle on the golden systems, wrapped ones included."""
    from cgppd_amd import formats, synth
    for name in GOLDEN_SYSTEMS:
        g = golden_system(golden, name)
        M = len(g["mol_off"]) - 1
        spec = synth.make_system("lean", np.diff(g["mol_off"]), g["box"], 1, 300.0, 300.0, 3)
        spec.types, spec.mol_off = g["types"].astype(np.int32), g["mol_off"].astype(np.int32)
        spec.ff = formats.ForceField([formats.AminoAcid(a.name, a.sname, a.symbol, float(sg), float(q))
                                      for a, sg, q in zip(spec.ff.acids, g["sigma"], g["charge"])],
                                     g["eps"].astype(np.float32))
        spec.centers = np.zeros((M, 3), np.float32)
        spec.rel = g["pos"].astype(np.float32)
        st = oracle.Remc(spec, seed=1)
        want = oracle.total(*oracle.energy_fp64(g["mol_off"], g["types"], g["charge"], g["sigma"], g["eps"],
                                                g["pos"], g["box"]))
        assert st.energy_fast(0) == pytest.approx(want, rel=1e-12, abs=1e-14), name


def test_lean_sampler_takes_the_oracle_decisions():
    """The lean sampler is the Markov chain of orc_remc_mc_steps: on the
    same streams it takes the same decisions and reaches the same
    coordinates; its running totals equal its own fresh evaluations."""
    spec, a = _tiny_remc(n_rep=4, seed=31)
    _, b = _tiny_remc(n_rep=4, seed=31)
    b.refresh_fast()
    a.mc_steps(300)
    b.mc_steps_fast(300, 3)
    np.testing.assert_array_equal(a.counters, b.counters)
    np.testing.assert_array_equal(a.pos, b.pos)
    # (reference-semantics float distances against FP64 ones)
    np.testing.assert_allclose(a.energy, b.energy, rtol=1e-5, atol=1e-6)
    for r in range(spec.n_rep):
        assert b.energy[r] == pytest.approx(b.energy_fast(r), rel=1e-9, abs=1e-12)
