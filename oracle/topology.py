"""Topology graph of flexible molecules, restated in Python (oracle).

TEST INFRASTRUCTURE ONLY: the checker for the product's topology builder
(cgppd_amd/csrc/topology.cpp) and the source of the flexible topology the C
oracle's MC uses.

Restates src/Geometry.cpp:
  Graph::add_edge / init            :11-192
  Graph::branch                     :194-222
  Graph::rigid_domain_around        :233-254
  Graph::assign_uids                :256-302
  Molecule::flex's branch choice    src/Molecule.cpp:336-358
  replica-wide UID offsets          src/Replica.cpp:30-38

Pinned by tests/testGeometry.cpp:31-96 (sets for the two-chain topology),
tests/testMolecule.cpp:306-323 and tests/testReplica.cpp:121-175 (counts
and UIDs), and by oracle/_ref/ref_flex_driver (the reference's own
src/Geometry.cpp) on every case of tests/golden/flex_vectors.npz.
"""
from __future__ import annotations

import sys

sys.setrecursionlimit(max(sys.getrecursionlimit(), 20000))


class Graph:
    """Graph::init over one molecule's residues (0-based, caller order)."""

    def __init__(self, chain_code, all_flexible, segments, num_chains):
        n = len(chain_code)
        self.num_chains = num_chains
        self.vertices = set()
        self.edges = set()
        self.adj = {}
        self.flexmap = {}
        for i in range(n - 1):
            if chain_code[i] == chain_code[i + 1]:
                self._add_edge(i, i + 1, bool(all_flexible))
        for seg in segments:
            for a, b in zip(seg[:-1], seg[1:]):
                self._add_edge(int(a), int(b), True)

        bonds, torsions, angles = set(), set(), set()
        local, crank, flexr = set(), set(), set()
        for (j, k) in sorted(self.edges):
            if self.flexmap[(j, k)]:
                bonds.add((j, k))
            nj, nk = self.adj[j], self.adj[k]
            if len(nj) > 1 and len(nk) > 1:
                for i in sorted(nj):
                    if i == k:
                        continue
                    for l in sorted(nk):
                        if l == j:
                            continue
                        if self._flex(i, j) or self._flex(j, k) or self._flex(k, l):
                            torsions.add((i, j, k, l))
        for j in sorted(self.vertices):
            nb = sorted(self.adj[j])
            has_rigid = has_flexible = False
            for i in nb:
                if self._flex(i, j):
                    has_flexible = True
                else:
                    has_rigid = True
                for k in nb:
                    if i < k and (self._flex(i, j) or self._flex(j, k)):
                        angles.add((i, j, k))
            if not has_rigid:
                local.add(j)
                if len(nb) > 1:
                    crank.add(j)
            if has_flexible:
                flexr.add(j)
        self.bonds = sorted(bonds)
        self.angles = sorted(angles)
        self.torsions = sorted(torsions)
        self.local = sorted(local)
        self.crankshaft = sorted(crank)
        self.flex = sorted(flexr)

        todo = set(self.vertices)
        self.rigid_domains = []
        while todo:
            i = min(todo)
            dom = self.rigid_domain_around(i)
            todo -= dom
            if len(dom) > 1:
                self.rigid_domains.append(sorted(dom))

        self.segment_bonds = [(i, j) for (i, j) in self.bonds
                              if chain_code[i] != chain_code[j] or abs(i - j) >= 4]

        indirect = set()
        for (pi, pj) in self.segment_bonds:
            around = sorted(t for t in self.torsions if pi in t and pj in t)
            proposed = set()
            for (ti, tj, tk, tl) in around:
                proposed.add((ti, tl))
                # the reference compares comma expressions: (a, b) == (c, d)
                # evaluates to b == d (src/Geometry.cpp:167-172)
                if tj == pj or ti == pj:
                    proposed.add((ti, tk))
                if tl == pj or tk == pj:
                    proposed.add((tj, tl))
            for (a, b) in sorted(proposed):
                if chain_code[a] != chain_code[b] or abs(a - b) >= 4:
                    indirect.add((a, b) if a < b else (b, a))
        self.indirect = sorted(indirect)

    def _add_edge(self, i, j, flexible):
        self.vertices.update((i, j))
        self.edges.add((i, j) if i < j else (j, i))
        self.adj.setdefault(i, set()).add(j)
        self.adj.setdefault(j, set()).add(i)
        self.flexmap[(i, j)] = flexible
        self.flexmap[(j, i)] = flexible

    def _flex(self, i, j):
        return self.flexmap.get((i, j), False)

    def neighbours(self, i):
        return sorted(self.adj.get(i, ()))

    def branch(self, i, j, visited=frozenset()):
        out = [(j, k) for k in sorted(self.adj.get(j, ())) if k != i and (j, k) not in visited]
        res = {i, j}
        nv = visited | {(i, j)}
        for (a, b) in out:
            res |= self.branch(a, b, nv)
        return res

    def rigid_domain_around(self, i, visited=frozenset()):
        out = [j for j in sorted(self.adj.get(i, ())) if not self._flex(i, j) and (j, i) not in visited]
        dom = {i}
        for j in out:
            dom |= self.rigid_domain_around(j, visited | {(i, j)})
        return dom

    def flex_branch(self, ri, nb):
        """Molecule::flex's set of residues to rotate (src/Molecule.cpp:336-358)."""
        res = set(self.branch(ri, nb))
        dom = self.rigid_domain_around(ri)
        if nb in dom:
            for i in self.neighbours(ri):
                if i != nb and i in dom:
                    res |= self.branch(ri, i)
        return sorted(res)


def molecule_graphs(spec_like):
    """Graphs of every molecule of a system given as a dict or SystemSpec
    with mol_off, chain_code, chain_index, all_flexible, segments."""
    g = _get(spec_like)
    out = []
    for m in range(len(g["mol_off"]) - 1):
        a, b = int(g["mol_off"][m]), int(g["mol_off"][m + 1])
        codes = [int(c) for c in g["chain_code"][a:b]]
        nchains = int(max(g["chain_index"][a:b])) + 1 if b > a else 0
        out.append(Graph(codes, g["all_flexible"][m], g["segments"][m], nchains))
    return out


def _get(s):
    if isinstance(s, dict):
        return s
    return dict(mol_off=s.mol_off, chain_code=s.chain_code, chain_index=s.chain_index, res_seq=s.res_seq,
                all_flexible=s.all_flexible, segments=s.segments)


def assign_uids(spec_like, graphs):
    """(N, 3) chain UID, rigid-domain UID, segment-bond UID per residue
    (src/Geometry.cpp:256-302, offsets from 1 as src/Replica.cpp:32-34)."""
    import numpy as np

    g = _get(spec_like)
    mol_off = g["mol_off"]
    N = int(mol_off[-1])
    uids = np.zeros((N, 3), dtype=np.int32)
    chain_off = dom_off = bond_off = 1
    for m, gr in enumerate(graphs):
        a = int(mol_off[m])
        chain_uid = {}
        for c in range(gr.num_chains):
            chain_uid[c] = chain_off
            chain_off += 1
        dom_uid = []
        for _d in gr.rigid_domains:
            dom_uid.append(dom_off)
            dom_off += 1
        bond_uid = []
        for _b in gr.segment_bonds:
            bond_uid.append(bond_off)
            bond_off += 1
        for d, dom in enumerate(gr.rigid_domains):
            for i in dom:
                uids[a + i, 1] = dom_uid[d]
        for b, (i, j) in enumerate(gr.segment_bonds):
            uids[a + i, 2] = bond_uid[b]
            uids[a + j, 2] = bond_uid[b]
        for i in range(len(gr.vertices)):
            uids[a + i, 0] = chain_uid.get(int(g["chain_index"][a + i]), 0)
    return uids
