"""The library's topology builder (cgppd_topology_list, host code in
libcgppd_hip.so, no GPU needed) against the oracle's restatement
(oracle/topology.py), itself pinned to the reference build in
tests/test_flex_oracle.py."""
import numpy as np
import pytest

from cgppd_amd import _lib, synth
from conftest import FLEX_CASES, flex_case
from oracle import topology


def molecules(case):
    mol_off = case["mol_off"]
    for m in range(len(mol_off) - 1):
        a, b = int(mol_off[m]), int(mol_off[m + 1])
        yield m, case["chain_code"][a:b], case["chain_index"][a:b], case["all_flexible"][m], case["segments"][m]


def check_molecule(code, index, allf, segs):
    g = topology.Graph([int(c) for c in code], allf, segs, int(max(index)) + 1)
    lib = lambda kind: _lib.topology_list(code, index, allf, segs, kind)  # noqa: E731
    for key, attr in (("bonds", "bonds"), ("angles", "angles"), ("torsions", "torsions"), ("local", "local"),
                      ("crankshaft", "crankshaft"), ("flex", "flex"), ("segment_bonds", "segment_bonds"),
                      ("indirect", "indirect")):
        want = np.array(getattr(g, attr), dtype=np.int32).reshape(lib(key).shape if len(getattr(g, attr)) else -1)
        assert np.array_equal(lib(key).reshape(-1), want.reshape(-1)), key
    dom = np.full(len(code), -1, dtype=np.int32)
    for d, res in enumerate(g.rigid_domains):
        dom[res] = d
    assert np.array_equal(lib("domain_of"), dom)
    sb = np.full(len(code), -1, dtype=np.int32)
    for b, (i, j) in enumerate(g.segment_bonds):
        sb[i] = b
        sb[j] = b
    assert np.array_equal(lib("segbond_of"), sb)
    nb, nb_off, br, br_off = [], [0], [], [0]
    for ri in g.flex:
        for n_ in g.neighbours(ri):
            nb.append(n_)
            br.extend(g.flex_branch(ri, n_))
            br_off.append(len(br))
        nb_off.append(len(nb))
    for key, want in (("flex_nb", nb), ("flex_nb_off", nb_off), ("branch", br), ("branch_off", br_off)):
        assert np.array_equal(lib(key), np.array(want, dtype=np.int32)), key
    assert list(lib("counts")) == [len(g.vertices), g.num_chains]


@pytest.mark.parametrize("name", FLEX_CASES)
def test_library_topology_matches_oracle(flex_golden, name):
    for _m, code, index, allf, segs in molecules(flex_case(flex_golden, name)):
        check_molecule(code, index, allf, segs)


def test_library_topology_linked_config():
    spec = synth.config(2)
    check_molecule(spec.chain_code, spec.chain_index, False, spec.segments[0])


def test_library_topology_edge_cases():
    # all flexible single chain, two chains with a cross-chain segment that
    # closes a loop, a single residue, and a segment on one residue
    check_molecule([65] * 12, [0] * 12, True, [])
    check_molecule([65] * 10 + [66] * 10, [0] * 10 + [1] * 10, False, [[3, 4, 5, 15], [8, 9, 12]])
    check_molecule([65], [0], False, [])
    check_molecule([65] * 5, [0] * 5, False, [[2]])


def test_library_topology_rejects_bad_segments():
    with pytest.raises(RuntimeError):
        _lib.topology_list([65] * 5, [0] * 5, False, [[1, 7]], "bonds")
