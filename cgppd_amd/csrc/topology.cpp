// Topology builder for flexible molecules (see topology.h).
#include "topology.h"

#include <algorithm>
#include <cstdlib>
#include <string>

namespace cgppd_topo {
namespace {

using Edge = std::pair<int, int>;

// The molecule's bonded graph: vertices are residues that take part in an
// edge; every edge is rigid or flexible (a later add_edge overrides).
struct Graph {
    std::set<int> vertices;
    std::set<Edge> edges;  // (min, max)
    std::map<int, std::set<int>> adj;
    std::map<Edge, bool> flexible;  // both directions

    void add_edge(int i, int j, bool f)
    {
        vertices.insert(i);
        vertices.insert(j);
        edges.insert(i < j ? Edge(i, j) : Edge(j, i));
        adj[i].insert(j);
        adj[j].insert(i);
        flexible[Edge(i, j)] = f;
        flexible[Edge(j, i)] = f;
    }
    bool is_flexible(int i, int j) const
    {
        auto it = flexible.find(Edge(i, j));
        return it != flexible.end() && it->second;
    }
    const std::set<int> &neighbours(int i) const
    {
        static const std::set<int> none;
        auto it = adj.find(i);
        return it == adj.end() ? none : it->second;
    }

    // Residues reached from j when entering it over the directed edge
    // i -> j without reusing a directed edge of the current path, plus i
    // (src/Geometry.cpp:194-222).  `path` holds the directed edges of the
    // path so far and is restored on return.
    void branch(int i, int j, std::set<Edge> &path, std::set<int> &out) const
    {
        std::vector<int> next;
        for (int k : neighbours(j))
            if (k != i && !path.count(Edge(j, k))) next.push_back(k);
        out.insert(i);
        out.insert(j);
        const bool added = path.insert(Edge(i, j)).second;
        for (int k : next) branch(j, k, path, out);
        if (added) path.erase(Edge(i, j));
    }
    std::set<int> branch(int i, int j) const
    {
        std::set<Edge> path;
        std::set<int> out;
        branch(i, j, path, out);
        return out;
    }

    // Residues joined to i by rigid edges (src/Geometry.cpp:233-254).
    void rigid_domain(int i, std::set<Edge> &path, std::set<int> &out) const
    {
        std::vector<int> next;
        for (int j : neighbours(i))
            if (!is_flexible(i, j) && !path.count(Edge(j, i))) next.push_back(j);
        out.insert(i);
        for (int j : next) {
            const bool added = path.insert(Edge(i, j)).second;
            rigid_domain(j, path, out);
            if (added) path.erase(Edge(i, j));
        }
    }
    std::set<int> rigid_domain(int i) const
    {
        std::set<Edge> path;
        std::set<int> out;
        rigid_domain(i, path, out);
        return out;
    }
};

}  // namespace

bool build(int n, const int *chain_code, bool all_flexible, const std::vector<std::vector<int>> &segments,
           int num_chains, Molecule &out, std::string &error)
{
    out = Molecule();
    out.n = n;
    out.num_chains = num_chains;
    out.flexible = all_flexible || !segments.empty();
    for (const auto &seg : segments)
        for (int v : seg)
            if (v < 0 || v >= n) {
                error = "segment residue index " + std::to_string(v) + " outside the molecule";
                return false;
            }

    Graph g;
    for (int i = 0; i + 1 < n; ++i)
        if (chain_code[i] == chain_code[i + 1]) g.add_edge(i, i + 1, all_flexible);
    for (const auto &seg : segments)
        for (size_t k = 0; k + 1 < seg.size(); ++k) g.add_edge(seg[k], seg[k + 1], true);

    std::set<std::array<int, 4>> torsions;
    for (const Edge &e : g.edges) {
        const int j = e.first, k = e.second;
        if (g.is_flexible(j, k)) out.bonds.push_back({j, k});
        const auto &nj = g.neighbours(j), &nk = g.neighbours(k);
        if (nj.size() < 2 || nk.size() < 2) continue;
        for (int i : nj) {
            if (i == k) continue;
            for (int l : nk) {
                if (l == j) continue;
                if (g.is_flexible(i, j) || g.is_flexible(j, k) || g.is_flexible(k, l)) torsions.insert({i, j, k, l});
            }
        }
    }
    out.torsions.assign(torsions.begin(), torsions.end());

    std::set<std::array<int, 3>> angles;
    for (int j : g.vertices) {
        const auto &nb = g.neighbours(j);
        bool rigid_nb = false, flexible_nb = false;
        for (int i : nb) {
            (g.is_flexible(i, j) ? flexible_nb : rigid_nb) = true;
            for (int k : nb)
                if (i < k && (g.is_flexible(i, j) || g.is_flexible(j, k))) angles.insert({i, j, k});
        }
        if (!rigid_nb) {
            out.local.push_back(j);
            if (nb.size() > 1) out.crankshaft.push_back(j);
        }
        if (flexible_nb) out.flex.push_back(j);
    }
    out.angles.assign(angles.begin(), angles.end());

    std::set<int> todo(g.vertices);
    while (!todo.empty()) {
        const std::set<int> dom = g.rigid_domain(*todo.begin());
        for (int v : dom) todo.erase(v);
        if (dom.size() > 1) out.domains.emplace_back(dom.begin(), dom.end());
    }

    for (const auto &b : out.bonds)
        if (chain_code[b[0]] != chain_code[b[1]] || std::abs(b[0] - b[1]) >= 4) out.segment_bonds.push_back(b);

    // Residue pairs next to each other across a segment bond, from the
    // torsions around it (src/Geometry.cpp:153-191).  The reference's
    // conditions compare comma expressions, so "(a, b) == (c, d)" tests
    // b == d; kept as it is.
    std::set<std::array<int, 2>> indirect;
    for (const auto &p : out.segment_bonds) {
        std::set<std::array<int, 2>> proposed;
        for (const auto &t : out.torsions) {
            const bool has_i = t[0] == p[0] || t[1] == p[0] || t[2] == p[0] || t[3] == p[0];
            const bool has_j = t[0] == p[1] || t[1] == p[1] || t[2] == p[1] || t[3] == p[1];
            if (!has_i || !has_j) continue;
            proposed.insert({t[0], t[3]});
            if (t[1] == p[1] || t[0] == p[1]) proposed.insert({t[0], t[2]});
            if (t[3] == p[1] || t[2] == p[1]) proposed.insert({t[1], t[3]});
        }
        for (const auto &q : proposed)
            if (chain_code[q[0]] != chain_code[q[1]] || std::abs(q[0] - q[1]) >= 4)
                indirect.insert({std::min(q[0], q[1]), std::max(q[0], q[1])});
    }
    out.indirect.assign(indirect.begin(), indirect.end());

    out.flex_nb_off.push_back(0);
    out.branch_off.push_back(0);
    for (int ri : out.flex) {
        const std::set<int> dom = g.rigid_domain(ri);
        for (int nb : g.neighbours(ri)) {
            std::set<int> b = g.branch(ri, nb);
            if (dom.count(nb))
                for (int i : g.neighbours(ri))
                    if (i != nb && dom.count(i)) {
                        const std::set<int> extra = g.branch(ri, i);
                        b.insert(extra.begin(), extra.end());
                    }
            out.flex_nb.push_back(nb);
            out.branch.insert(out.branch.end(), b.begin(), b.end());
            out.branch_off.push_back((int)out.branch.size());
        }
        out.flex_nb_off.push_back((int)out.flex_nb.size());
    }

    out.domain_of.assign(n, -1);
    for (size_t d = 0; d < out.domains.size(); ++d)
        for (int v : out.domains[d]) out.domain_of[v] = (int)d;
    out.segbond_of.assign(n, -1);
    for (size_t b = 0; b < out.segment_bonds.size(); ++b) {
        out.segbond_of[out.segment_bonds[b][0]] = (int)b;
        out.segbond_of[out.segment_bonds[b][1]] = (int)b;
    }
    out.vertex_count = (int)g.vertices.size();
    return true;
}

}  // namespace cgppd_topo
