// ref_flex_driver.cpp -- golden-vector generator for the flexible-linker
// path, built against the reference's own sources (test infrastructure only).
//
// oracle/Makefile compiles this file with -DEnableFlexibleLinkers together
// with src/Geometry.cpp, src/Potential.cpp, src/Residue.cpp,
// src/TorsionalLookupMatrix.cpp, src/AminoAcids.cpp, src/AminoAcid.cpp,
// src/vector3f.cpp and src/Quaternion.cpp where they lie under the reference
// tree, into oracle/_ref/ref_flex_driver.  Nothing from it ships.
//
// Linked unchanged from the reference: Graph::init / branch /
// rigid_domain_around / assign_uids (src/Geometry.cpp), calculate_LJ,
// calculate_DH, calculate_bond, calculate_angle, calculate_torsion and the
// Potential accumulator (src/Potential.cpp), Residue::distance and the
// exclusion predicates (src/Residue.cpp), the torsion table loader
// (src/TorsionalLookupMatrix.cpp).  Restated here because their translation
// units need GSL: the loops of Molecule::E (src/Molecule.cpp:636-699),
// Replica::E (src/Replica.cpp:432-493), the branch selection of
// Molecule::flex (src/Molecule.cpp:331-358) and the GPU kernel's exclusion
// rule (src/CudaFunctions.cu:447-476).
//
// Usage:
//   ref_flex_driver system <AminoAcids> <pair table> <torsion table> < system
//     system = "box L\nnmol M\n", then per molecule
//       "mol <n_res> <all_flexible 0|1> <n_segments>",
//       n_segments lines "seg <k> i1 .. ik" (0-based residue indices),
//       n_res lines "AAA <chain char> <chain index> <resSeq> x y z".
//   Prints the topology of every molecule, the replica-wide UIDs, and the
//   energies (see the keys printed below).
#include "AminoAcids.h"
#include "Geometry.h"
#include "Potential.h"
#include "Residue.h"
#include "TorsionalLookupMatrix.h"
#include "structures.h"
#include "vector3f.h"

#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <vector>

struct Mol {
    std::vector<Residue> res;
    bool all_flexible = false;
    std::vector<segdata> segments;
    int num_chains = 0;
    Graph graph;
    bool is_flexible = false;
};

static void print_set(const char *key, const std::set<int> &s)
{
    std::printf("%s %zu", key, s.size());
    for (int v : s) std::printf(" %d", v);
    std::printf("\n");
}

// Molecule::flex's choice of residues to rotate for (ri, neighbour),
// src/Molecule.cpp:336-358.
static std::set<int> flex_branch(Graph &g, int ri, int neighbour)
{
    const std::vector<int> neighbours = g.neighbours(ri);
    const std::set<int> main_branch = g.branch(ri, neighbour);
    std::set<int> branch(main_branch.begin(), main_branch.end());
    const std::set<int> ri_domain = g.rigid_domain_around(ri);
    if (ri_domain.find(neighbour) != ri_domain.end()) {
        for (int i : neighbours) {
            if (i != neighbour && ri_domain.find(i) != ri_domain.end()) {
                const std::set<int> extra = g.branch(ri, i);
                branch.insert(extra.begin(), extra.end());
            }
        }
    }
    return branch;
}

static void print_potential(const char *key, const Potential &p)
{
    std::printf("%s %.17g %.17g %.17g %.17g %.17g %.17g\n", key, p.LJ, p.DH, p.bond, p.angle, p.torsion, p.total());
}

static int run_system(const char *aa_file, const char *pair_file, const char *torsion_file)
{
    AminoAcids aa;
    aa.init(aa_file, pair_file);
    TorsionalLookupMatrix tor;
    tor.loadData(torsion_file, aa);
    float box = 0.0f;
    int nmol = 0;
    char key[64];
    if (std::scanf("%63s %f", key, &box) != 2 || std::strcmp(key, "box") != 0) return 2;
    if (std::scanf("%63s %d", key, &nmol) != 2 || std::strcmp(key, "nmol") != 0) return 2;
    std::vector<Mol> mols(nmol);
    for (int m = 0; m < nmol; ++m) {
        int n = 0, allf = 0, nseg = 0;
        if (std::scanf("%63s %d %d %d", key, &n, &allf, &nseg) != 4 || std::strcmp(key, "mol") != 0) return 2;
        mols[m].all_flexible = allf != 0;
        for (int s = 0; s < nseg; ++s) {
            int k = 0;
            if (std::scanf("%63s %d", key, &k) != 2 || std::strcmp(key, "seg") != 0) return 2;
            segdata sd;
            for (int t = 0; t < k; ++t) {
                int v;
                if (std::scanf("%d", &v) != 1) return 2;
                sd.residue_indices.push_back(v);
            }
            mols[m].segments.push_back(sd);
        }
        int max_chain = -1;
        for (int i = 0; i < n; ++i) {
            char name[8], chain[4];
            int chain_int, seq;
            float x, y, z;
            if (std::scanf("%7s %3s %d %d %f %f %f", name, chain, &chain_int, &seq, &x, &y, &z) != 7) return 2;
            Residue r;
            r.aminoAcidIndex = aa.getAminoAcidIndex(name);
            r.electrostaticCharge = float(aa.get(r.aminoAcidIndex).electrostaticCharge);
            r.vanderWaalRadius = aa.data[r.aminoAcidIndex].vanderWaalRadius;
            r.position = Vector3f(x, y, z);
            r.chainId = chain[0];
            r.chain_int_id = chain_int;
            r.resSeq = seq;
            r.moleculeId = m;
            mols[m].res.push_back(r);
            if (chain_int > max_chain) max_chain = chain_int;
        }
        mols[m].num_chains = max_chain + 1;
        mols[m].graph.init(mols[m].res, mols[m].all_flexible, mols[m].segments, mols[m].num_chains);
        mols[m].is_flexible = mols[m].all_flexible || !mols[m].segments.empty();
    }
    // replica-wide UIDs, src/Replica.cpp:30-38
    int chain_offset = 1, domain_offset = 1, bond_offset = 1;
    for (int m = 0; m < nmol; ++m)
        mols[m].graph.assign_uids(mols[m].res.data(), chain_offset, domain_offset, bond_offset);

    for (int m = 0; m < nmol; ++m) {
        Graph &g = mols[m].graph;
        std::printf("mol %d flexible %d\n", m, mols[m].is_flexible ? 1 : 0);
        std::printf("bonds %zu", g.bonds.size());
        for (const Bond &b : g.bonds) std::printf(" %d %d", b.i, b.j);
        std::printf("\nangles %zu", g.angles.size());
        for (const Angle &a : g.angles) std::printf(" %d %d %d", a.i, a.j, a.k);
        std::printf("\ntorsions %zu", g.torsions.size());
        for (const Torsion &t : g.torsions) std::printf(" %d %d %d %d", t.i, t.j, t.k, t.l);
        std::printf("\nlocal %zu", g.MC_local_residues.size());
        for (int v : g.MC_local_residues) std::printf(" %d", v);
        std::printf("\ncrankshaft %zu", g.MC_crankshaft_residues.size());
        for (int v : g.MC_crankshaft_residues) std::printf(" %d", v);
        std::printf("\nflex %zu", g.MC_flex_residues.size());
        for (int v : g.MC_flex_residues) std::printf(" %d", v);
        std::printf("\nsegment_bonds %zu", g.segment_bonds.size());
        for (const Pair &p : g.segment_bonds) std::printf(" %d %d", p.i, p.j);
        std::printf("\nindirect %zu", g.indirect_neighbours.size());
        for (const Pair &p : g.indirect_neighbours) std::printf(" %d %d", p.i, p.j);
        std::printf("\ndomains %zu\n", g.rigid_domains.size());
        for (const std::set<int> &d : g.rigid_domains) print_set("domain", d);
        for (int ri : g.MC_flex_residues)
            for (int nb : g.neighbours(ri)) {
                const std::set<int> b = flex_branch(g, ri, nb);
                std::printf("branch %d %d", ri, nb);
                print_set("", b);
            }
        std::printf("uids");
        for (const Residue &r : mols[m].res)
            std::printf(" %d %d %d", r.chain_UID, r.rigid_domain_UID, r.segment_bond_UID);
        std::printf("\n");
    }

    // Replica::E: every inter-molecular residue pair (src/Replica.cpp:432-493)
    Potential inter;
    for (int I = 0; I < nmol; ++I)
        for (int J = I + 1; J < nmol; ++J)
            for (Residue &ri : mols[I].res)
                for (Residue &rj : mols[J].res) {
                    double r(ri.distance(rj, box) + EPS);
                    inter.increment_LJ(calculate_LJ(ri, rj, r, aa));
                    inter.increment_DH(calculate_DH(ri, rj, r));
                }
    print_potential("inter", inter);

    for (int m = 0; m < nmol; ++m) {
        Mol &M = mols[m];
        Residue *R = M.res.data();
        const int n = (int)M.res.size();
        Potential with_nb, without_nb, gpu_rule;
        if (M.is_flexible) {
            // Molecule::E(true): intra-molecular pairs (src/Molecule.cpp:646-668)
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j)
                    if (!R[i].same_rigid_domain_as(R[j]) && !R[i].chain_neighbour_of(R[j]) &&
                        !R[i].bonded_to(R[j]) && !M.graph.indirect_neighbours.count(Pair(i, j))) {
                        double r(R[i].distance(R[j], box) + EPS);
                        with_nb.increment_LJ(calculate_LJ(R[i], R[j], r, aa));
                        with_nb.increment_DH(calculate_DH(R[i], R[j], r));
                    }
            // Molecule::E(false): subtract the indirect neighbours (:669-683)
            for (const Pair &p : M.graph.indirect_neighbours) {
                double r(R[p.i].distance(R[p.j], box) + EPS);
                without_nb.increment_LJ(-calculate_LJ(R[p.i], R[p.j], r, aa));
                without_nb.increment_DH(-calculate_DH(R[p.i], R[p.j], r));
            }
            // bonded terms (:686-696), added to both
            for (const Bond &b : M.graph.bonds) {
                double v = calculate_bond(R, b, box);
                with_nb.increment_bond(v);
                without_nb.increment_bond(v);
            }
            for (const Angle &a : M.graph.angles) {
                double v = calculate_angle(R, a);
                with_nb.increment_angle(v);
                without_nb.increment_angle(v);
            }
            for (const Torsion &t : M.graph.torsions) {
                double v = calculate_torsion(R, t, tor);
                with_nb.increment_torsion(v);
                without_nb.increment_torsion(v);
            }
            // the GPU kernel's intra-molecular pairs: everything except the
            // same rigid domain, the same segment bond and |index difference|
            // < 4 on the same chain (src/CudaFunctions.cu:447-476)
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    const bool same_domain = R[i].rigid_domain_UID && R[i].rigid_domain_UID == R[j].rigid_domain_UID;
                    const bool same_bond = R[i].segment_bond_UID && R[i].segment_bond_UID == R[j].segment_bond_UID;
                    const bool close = R[i].chain_UID == R[j].chain_UID && (j - i) < 4;
                    if (same_domain || same_bond || close) continue;
                    double r(R[i].distance(R[j], box) + EPS);
                    gpu_rule.increment_LJ(calculate_LJ(R[i], R[j], r, aa));
                    gpu_rule.increment_DH(calculate_DH(R[i], R[j], r));
                }
        }
        std::printf("internal_true %d", m);
        print_potential("", with_nb);
        std::printf("internal_false %d", m);
        print_potential("", without_nb);
        std::printf("gpu_intra %d", m);
        print_potential("", gpu_rule);
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 5 && std::strcmp(argv[1], "system") == 0) return run_system(argv[2], argv[3], argv[4]);
    std::fprintf(stderr, "usage: ref_flex_driver system <AminoAcids> <pairs> <torsions> < system\n");
    return 2;
}
