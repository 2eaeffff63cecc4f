"""Flexible linkers on the device against the CPU oracle.

Energies: EonDevice (inter-molecular pairs plus the GPU kernel's
intra-molecular pairs) and internal_molecule_E(false) through the C ABI
against the oracle's reference-semantics restatement (bit-exact with the
reference build, tests/test_flex_oracle.py) and the FP64 oracle, 1e-5
relative; the reference's own expectation for conf1 with a flexible tail
(tests/testReplica.cpp:84-86) with its Catch tolerance.  MC: flex, local
translate and crankshaft proposals drawn from the same Philox streams as
the oracle, coordinates bit for bit.
"""
import numpy as np
import pytest

import oracle
from conftest import FLEX_ENERGY_CASES, flex_case

pytestmark = pytest.mark.gpu

REL = 1e-5


def engine_for_case(factory, case):
    e = factory()
    e.set_force_field(case["eps"], case["sigma"], case["charge"])
    e.set_box(case["box"])
    e.set_system(case["mol_off"], case["types"])
    e.set_flexible_spec(case)
    M = len(case["mol_off"]) - 1
    e.set_replicas(np.zeros((1, M, 3), np.float32), case["pos"][None], seed=1)
    e.upload_positions(0, case["pos"])
    return e


def oracle_parts(case, fp64=False):
    fl = oracle.Flex(case, case["torsion_v"], case["torsion_delta"])
    fn = oracle.energy_fp64 if fp64 else oracle.energy_ref
    lj, dh = fn(case["mol_off"], case["types"], case["charge"], case["sigma"], case["eps"], case["pos"], case["box"])
    gpu_lj, gpu_dh = lj, dh
    internal = np.zeros(5)
    total = oracle.total(lj, dh)
    for m in range(len(case["mol_off"]) - 1):
        args = (fl, m, case["mol_off"], case["types"], case["charge"], case["sigma"], case["eps"], case["pos"],
                case["box"])
        a = oracle.internal(*args, 2, fp64)
        b = oracle.internal(*args, 1, fp64)
        c = oracle.internal(*args, 0, fp64)
        gpu_lj += a[0]
        gpu_dh += a[1]
        internal[0] += b[0]
        internal[1] += b[1]
        internal[2] += b[2]
        internal[3] += np.log(b[3])
        internal[4] += b[4]
        total += oracle.internal_total(c)
    rt, dhc, _ = oracle.constants()
    kcal = np.array([internal[0] * float(np.float32(rt)), internal[1] * float(np.float32(dhc)),
                     internal[2] * 0.5 * 378, internal[3] * -10.0, internal[4]])
    return gpu_lj, gpu_dh, kcal, total


@pytest.mark.parametrize("name", FLEX_ENERGY_CASES)
def test_flexible_energies(gpu_engine_factory, flex_golden, name):
    case = flex_case(flex_golden, name)
    e = engine_for_case(gpu_engine_factory, case)
    lj, dh, tot = e.energy(0)
    internal = e.energy_internal(0)
    want_lj, want_dh, want_int, want_total = oracle_parts(case)
    _, _, _, fp64_total = oracle_parts(case, fp64=True)
    print(f"{name}: gpu lj {lj:.10g} dh {dh:.10g} internal {internal}; oracle {want_lj:.10g} {want_dh:.10g} {want_int}")
    assert abs(lj - want_lj) <= REL * abs(want_lj) + 1e-7
    assert abs(dh - want_dh) <= REL * abs(want_dh) + 1e-7
    scale = max(1.0, np.abs(want_int).max())
    np.testing.assert_allclose(internal, want_int, rtol=REL, atol=1e-7 * scale)
    got_total = tot + internal.sum()
    assert abs(got_total - want_total) <= REL * abs(want_total) + 1e-7
    assert abs(got_total - fp64_total) <= REL * abs(fp64_total) + 1e-7
    # the asynchronous path and energy_all report the same EonDevice total
    e.energy_async(0)
    assert abs(e.energy_wait(0)[2] - tot) <= 1e-12 * max(1.0, abs(tot))
    _, _, tots, mat = e.energy_all(matrix=True)
    assert abs(tots[0] - tot) <= 1e-9 * max(1.0, abs(tot))
    # the matrix's upper triangle with the diagonal (flexible molecules'
    # intra-molecular pairs) sums to the same total
    assert abs(np.triu(mat[0]).sum() - tot) <= 1e-9 * max(1.0, abs(tot))
    e.close()


def test_reference_flexible_replica_total(gpu_engine_factory, flex_golden):
    """tests/testReplica.cpp:84: EonDevice() + internal_molecule_E(false).total()
    == Approx(expected flexible total)."""
    case = flex_case(flex_golden, "conf1_flex")
    e = engine_for_case(gpu_engine_factory, case)
    tot = e.energy(0)[2] + e.energy_internal(0).sum()
    want = -0.068205 - 0.240383 + 0.180426 + 1.555213 + 5.420859
    eps = 100 * np.finfo(np.float32).eps
    assert abs(tot - want) < eps * (1 + max(abs(tot), abs(want))) + 5e-6, (tot, want)
    # EonDeviceNC counts inter-molecular pairs only: the rigid value
    nc = e.energy_nc(0)[2]
    assert abs(nc - (-0.079673 - 0.212202)) < 2e-6
    e.close()


def mc_pair(factory, spec, seed):
    e = factory()
    e.load_spec(spec, seed=seed)
    e.mc_prepare()
    st = oracle.Remc(spec, seed=seed)
    return e, st


def compare_states(e, st, spec, split_counters=True):
    """split_counters=False compares accepted and rejected totals only: a
    rotation of a molecule that is alone leaves its energy unchanged in
    exact arithmetic, so whether the rounding noise in delta counts as
    'accept' (delta < 0) or 'acceptA' (Boltzmann) differs between the FP32
    device sums and the oracle -- both accept it."""
    s = e.stats()
    np.testing.assert_array_equal(s["steps"], st.step)
    if split_counters:
        np.testing.assert_array_equal(s["counters"], st.counters)
    else:
        np.testing.assert_array_equal(s["counters"][:, :2].sum(1), st.counters[:, :2].sum(1))
        np.testing.assert_array_equal(s["counters"][:, 2], st.counters[:, 2])
    np.testing.assert_allclose(s["energy"], st.energy, rtol=REL, atol=1e-6)
    for r in range(spec.n_rep):
        pos, rel, cen = e.coordinates(r)
        np.testing.assert_array_equal(pos, st.pos[r])
        np.testing.assert_array_equal(rel, st.rel[r])
        np.testing.assert_array_equal(cen, st.center[r])


def test_prepare_with_flexible_molecule(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_linked_system("mcflex", 2, 30, 6, 0.2, 120.0, 4, 288.0, 500.0, 41, extra_rigid=(25, 40))
    e, st = mc_pair(gpu_engine_factory, spec, 5)
    s = e.stats()
    np.testing.assert_allclose(s["energy"], st.energy, rtol=REL, atol=1e-6)
    for r in range(spec.n_rep):
        np.testing.assert_allclose(e.energy_matrix(r), st.emat[r], rtol=REL, atol=1e-6)
    e.close()


def test_flexible_mc_trajectory_matches_oracle(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_linked_system("mcflex", 2, 30, 6, 0.2, 120.0, 4, 288.0, 500.0, 41, extra_rigid=(25, 40))
    e, st = mc_pair(gpu_engine_factory, spec, 17)
    for chunk in range(4):
        e.mc_steps(60)
        st.mc_steps(60)
        print(f"after {60 * (chunk + 1)} steps: counters gpu {e.stats()['counters'].sum(0)} oracle {st.counters.sum(0)}")
        compare_states(e, st, spec)
    e.close()


def test_single_flexible_molecule_trajectory(gpu_engine_factory):
    """The linked-domain molecule alone (rotate / flex / local moves)."""
    from cgppd_amd import synth
    spec = synth.config(2)
    spec.n_rep = 4
    e, st = mc_pair(gpu_engine_factory, spec, 23)
    for chunk in range(3):
        e.mc_steps(50)
        st.mc_steps(50)
        compare_states(e, st, spec, split_counters=False)
    e.close()


def test_flexible_running_total_and_graph(gpu_engine_factory):
    from cgppd_amd import synth
    spec = synth.make_linked_system("mcflex", 2, 30, 6, 0.2, 120.0, 4, 288.0, 500.0, 41, extra_rigid=(25, 40))
    a = gpu_engine_factory()
    a.load_spec(spec, seed=3)
    a.mc_prepare()
    b = gpu_engine_factory()
    b.load_spec(spec, seed=3)
    b.mc_prepare()
    a.run(3, 100, use_graph=True)
    b.run(3, 100, use_graph=False)
    sa, sb = a.stats(), b.stats()
    np.testing.assert_array_equal(sa["counters"], sb["counters"])
    for r in range(spec.n_rep):
        np.testing.assert_array_equal(a.coordinates(r)[0], b.coordinates(r)[0])
        tot = a.energy(r)[2] + a.energy_internal(r).sum()
        assert abs(sa["energy"][r] - tot) <= 1e-6 * max(1.0, abs(tot)), (r, sa["energy"][r], tot)
    a.close()
    b.close()


def test_autotune_keeps_flexible_state(gpu_engine_factory):
    """The autotuner's timed launches add into the pending rows, which hold
    the starting internal energy of a proposal that moves one residue: the
    tuned engine must still follow the oracle's trajectory (regression)."""
    from cgppd_amd import synth
    spec = synth.config(2)
    spec.n_rep = 4
    e, st = mc_pair(gpu_engine_factory, spec, 31)
    best = e.autotune(launches=2)
    print("autotune", best, flush=True)
    for chunk in range(2):
        e.mc_steps(60)
        st.mc_steps(60)
        compare_states(e, st, spec, split_counters=False)
    e.close()
