// Topology of one molecule for the flexible-linker model: the bonded graph,
// the move sets and the exclusion UIDs.  Host code, built into
// libcgppd_hip.so.
//
// Follows src/Geometry.cpp:34-302 (Graph::init, branch, rigid_domain_around,
// assign_uids) and the branch choice of Molecule::flex
// (src/Molecule.cpp:336-358).  Residue indices are 0-based within the
// molecule, in the caller's order.
#pragma once

#include <array>
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace cgppd_topo {

struct Molecule {
    int n = 0;
    int num_chains = 0;
    int vertex_count = 0;   // residues on some edge (chain UIDs go to the first vertex_count residues)
    bool flexible = false;  // has a segment or is all flexible

    std::vector<std::array<int, 2>> bonds;     // flexible edges, (i < j), sorted
    std::vector<std::array<int, 3>> angles;    // (i, j, k), i < k, sorted
    std::vector<std::array<int, 4>> torsions;  // (i, j, k, l) around edge (j, k), sorted
    std::vector<int> local, crankshaft, flex;  // move sets, ascending
    std::vector<std::vector<int>> domains;     // rigid domains of more than one residue
    std::vector<std::array<int, 2>> segment_bonds;
    std::vector<std::array<int, 2>> indirect;  // (i < j), sorted

    // flex moves: for every flex residue (ascending) and each of its
    // neighbours (ascending), the residues the move rotates (ascending)
    std::vector<int> flex_nb_off;  // flex.size() + 1, into flex_nb
    std::vector<int> flex_nb;
    std::vector<int> branch_off;   // flex_nb.size() + 1, into branch
    std::vector<int> branch;

    // per residue: index of its rigid domain (-1 none), segment bond (-1 none)
    std::vector<int> domain_of, segbond_of;
};

// Builds the topology.  chain_code: the PDB chain identifier of each residue
// (consecutive residues with equal codes are bonded); segments: lists of
// residue indices whose consecutive entries are flexible edges.
// Returns false (with a message) for out-of-range indices.
bool build(int n, const int *chain_code, bool all_flexible, const std::vector<std::vector<int>> &segments,
           int num_chains, Molecule &out, std::string &error);

}  // namespace cgppd_topo
