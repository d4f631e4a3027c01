// steppe_amd/csrc/qpgraph_enum.cpp — the bounded qpGraph topology space (host code).
//
// Graphs are parent -> child edge lists over integer node ids: leaves are 0..nleaf-1
// (in the caller's leaf order), the nodes a tree insertion adds are nleaf, nleaf+1, ...,
// and an admixture wiring adds the pair 500000 + 2s, 500000 + 2s + 1 for source edge s.
// The shape hash is a leaf-anchored colour refinement (1-WL) whose final colouring is
// written as text and hashed with 64-bit FNV-1a; the text, the number of rounds and the
// colour codes follow the reference (src/core/qpadm/qpgraph_enumerate.cpp:31-82), so the
// hashes agree bit for bit with its enumerator and with scores keyed by them.
#include "qpgraph_enum.hpp"

#include <algorithm>
#include <atomic>
#include <thread>
#include <unordered_map>
#include <unordered_set>

namespace stp::qgenum {

namespace {

constexpr int kWiringBase = 500000;
constexpr int kExtraRounds = 4;
constexpr std::uint64_t kFnvBasis = 1469598103934665603ULL;
constexpr std::uint64_t kFnvPrime = 1099511628211ULL;

struct Edge {
    int p, c;
};
using Graph = std::vector<Edge>;

// ---- shape hash ---------------------------------------------------------------------
std::uint64_t shape_hash(const Graph& g, int nleaf) {
    std::vector<int> ids;
    ids.reserve(2 * g.size());
    for (const Edge& e : g) {
        ids.push_back(e.p);
        ids.push_back(e.c);
    }
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const int V = static_cast<int>(ids.size());
    auto at = [&](int id) { return static_cast<int>(std::lower_bound(ids.begin(), ids.end(), id) - ids.begin()); };
    std::vector<std::vector<int>> up(V), down(V);
    std::vector<std::pair<int, int>> ge;
    ge.reserve(g.size());
    for (const Edge& e : g) {
        const int a = at(e.p), b = at(e.c);
        down[a].push_back(b);
        up[b].push_back(a);
        ge.emplace_back(a, b);
    }
    std::vector<std::string> col(V), next(V), sig(V), uniq;
    for (int v = 0; v < V; ++v) {
        const int id = ids[v];
        col[v] = (id >= 0 && id < nleaf)
                     ? "L" + std::to_string(id)
                     : "I" + std::to_string(up[v].size()) + "_" + std::to_string(down[v].size());
    }
    std::vector<const std::string*> nb;
    const int rounds = 2 * V + kExtraRounds;
    for (int r = 0; r < rounds; ++r) {
        for (int v = 0; v < V; ++v) {
            std::string& s = sig[v];
            s = col[v];
            s += "|<";
            for (const auto* adj : {&up[v], &down[v]}) {
                nb.clear();
                for (int w : *adj) nb.push_back(&col[w]);
                std::sort(nb.begin(), nb.end(), [](const std::string* x, const std::string* y) { return *x < *y; });
                for (const std::string* x : nb) {
                    s += *x;
                    s += ',';
                }
                s += adj == &up[v] ? ">|<" : ">";
            }
        }
        uniq = sig;
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        for (int v = 0; v < V; ++v)
            next[v] = std::to_string(std::lower_bound(uniq.begin(), uniq.end(), sig[v]) - uniq.begin());
        if (next == col) break;   // a fixed point: every later round reproduces it
        col.swap(next);
    }
    std::vector<std::string> nodes(col), edges;
    std::sort(nodes.begin(), nodes.end());
    edges.reserve(ge.size());
    for (const auto& [a, b] : ge) edges.push_back(col[a] + ">" + col[b]);
    std::sort(edges.begin(), edges.end());
    std::uint64_t h = kFnvBasis;
    auto feed = [&](const std::string& s) {
        for (unsigned char ch : s) {
            h ^= ch;
            h *= kFnvPrime;
        }
    };
    feed("N:");
    for (const auto& s : nodes) {
        feed(s);
        feed(";");
    }
    feed("E:");
    for (const auto& s : edges) {
        feed(s);
        feed(";");
    }
    return h;
}

// ---- names <-> ids --------------------------------------------------------------------
NamedEdges named(const Graph& g, const std::vector<std::string>& leaves) {
    const int nleaf = static_cast<int>(leaves.size());
    auto name = [&](int id) { return id >= 0 && id < nleaf ? leaves[id] : "N" + std::to_string(id); };
    NamedEdges out;
    out.reserve(g.size());
    for (const Edge& e : g) out.emplace_back(name(e.p), name(e.c));
    return out;
}

// Leaves get their index in the leaf list; other nodes the next free ids in order of
// first appearance over the edges (parent, then child).
Graph to_ids(const NamedEdges& edges, const std::vector<std::string>& leaves) {
    std::unordered_map<std::string, int> id;
    for (int i = 0; i < static_cast<int>(leaves.size()); ++i) id.emplace(leaves[i], i);
    int next = static_cast<int>(leaves.size());
    auto get = [&](const std::string& s) {
        const auto it = id.find(s);
        if (it != id.end()) return it->second;
        id.emplace(s, next);
        return next++;
    };
    Graph g;
    g.reserve(edges.size());
    for (const auto& [a, b] : edges) {
        const int p = get(a);
        g.push_back({p, get(b)});
    }
    return g;
}

// ---- trees -------------------------------------------------------------------------------
// Leaf k joins every edge of every tree on leaves 0..k-1 through a new node; a pseudo
// edge above the root lets the new leaf also join above it, and is dropped at the end.
std::vector<Graph> all_trees(int nleaf) {
    std::vector<Graph> cur{Graph{{-1, 0}}};
    for (int k = 1; k < nleaf; ++k) {
        const int w = nleaf + k - 1;
        std::vector<Graph> grown;
        grown.reserve(cur.size() * (cur.empty() ? 0 : cur[0].size()));
        for (const Graph& t : cur)
            for (std::size_t i = 0; i < t.size(); ++i) {
                Graph n;
                n.reserve(t.size() + 2);
                for (std::size_t j = 0; j < t.size(); ++j)
                    if (j != i) n.push_back(t[j]);
                n.push_back({t[i].p, w});
                n.push_back({w, t[i].c});
                n.push_back({w, k});
                grown.push_back(std::move(n));
            }
        cur.swap(grown);
    }
    for (Graph& t : cur) t.erase(std::remove_if(t.begin(), t.end(), [](const Edge& e) { return e.p == -1; }), t.end());
    return cur;
}

bool has_descendant(const std::unordered_map<int, std::vector<int>>& down, int from, int target) {
    std::vector<int> stack{from};
    std::unordered_set<int> seen;
    while (!stack.empty()) {
        const int u = stack.back();
        stack.pop_back();
        const auto it = down.find(u);
        if (it == down.end()) continue;
        for (int v : it->second) {
            if (v == target) return true;
            if (seen.insert(v).second) stack.push_back(v);
        }
    }
    return false;
}

// Every one-admixture wiring of a tree, in order (source edge s, destination edge d over
// the tree without its outgroup edge): split s by x and d by a, add x -> a.  Returns the
// graphs and their hashes; the caller keeps the first of each hash.
void wirings(const Graph& tree, int nleaf, std::vector<Graph>& gs, std::vector<std::uint64_t>& hs) {
    std::unordered_set<int> has_parent, all;
    for (const Edge& e : tree) {
        has_parent.insert(e.c);
        all.insert(e.p);
        all.insert(e.c);
    }
    int root = -1;
    for (const Edge& e : tree)
        if (!has_parent.count(e.p)) {
            root = e.p;
            break;
        }
    int outgroup = -1, n_leaf_children = 0;
    for (const Edge& e : tree)
        if (e.p == root && e.c >= 0 && e.c < nleaf) {
            outgroup = e.c;
            ++n_leaf_children;
        }
    if (n_leaf_children != 1) outgroup = -1;
    Graph base;
    for (const Edge& e : tree)
        if (!(outgroup >= 0 && e.p == root && e.c == outgroup)) base.push_back(e);
    std::unordered_map<int, std::vector<int>> down;
    for (const Edge& e : tree) down[e.p].push_back(e.c);
    for (std::size_t si = 0; si < base.size(); ++si) {
        const int x = kWiringBase + 2 * static_cast<int>(si), a = x + 1;
        for (std::size_t di = 0; di < base.size(); ++di) {
            if (si == di) continue;
            const Edge s = base[si], d = base[di];
            if (s.p == d.p || s.c == d.c || s.c == d.p || d.c == s.p) continue;
            if (has_descendant(down, d.c, s.p)) continue;   // x -> a would close a cycle
            Graph g;
            g.reserve(tree.size() + 3);
            for (const Edge& e : tree)
                if (!((e.p == s.p && e.c == s.c) || (e.p == d.p && e.c == d.c))) g.push_back(e);
            g.push_back({s.p, x});
            g.push_back({x, s.c});
            g.push_back({d.p, a});
            g.push_back({a, d.c});
            g.push_back({x, a});
            hs.push_back(shape_hash(g, nleaf));
            gs.push_back(std::move(g));
        }
    }
}

void keep_new_wirings(const Graph& tree, int nleaf, const std::vector<std::string>& leaves,
                      std::unordered_set<std::uint64_t>& seen, std::vector<Candidate>& out) {
    std::vector<Graph> gs;
    std::vector<std::uint64_t> hs;
    wirings(tree, nleaf, gs, hs);
    for (std::size_t i = 0; i < gs.size(); ++i)
        if (seen.insert(hs[i]).second) out.push_back(Candidate{named(gs[i], leaves), 1, 0, hs[i]});
}

// The tree an admixture graph was wired onto, recovered from its admixture node a.
Graph base_tree(const Graph& g) {
    std::unordered_map<int, int> indeg;
    std::vector<int> order;
    for (const Edge& e : g) {
        if (!indeg.count(e.p)) order.push_back(e.p), indeg[e.p] += 0;
        if (!indeg.count(e.c)) order.push_back(e.c), indeg[e.c] += 0;
        ++indeg[e.c];
    }
    int a = -1;
    for (int u : order)
        if (indeg[u] == 2) {
            a = u;
            break;
        }
    if (a < 0) return g;
    int a_child = -1;
    std::vector<int> a_par;
    for (const Edge& e : g) {
        if (e.p == a) a_child = e.c;
        if (e.c == a) a_par.push_back(e.p);
    }
    auto other_child = [&](int p) {
        for (const Edge& e : g)
            if (e.p == p && e.c != a) return true;
        return false;
    };
    int x = a_par[0], dest_from = a_par.size() > 1 ? a_par[1] : -1;
    if (a_par.size() == 2 && !other_child(a_par[0])) std::swap(x, dest_from);
    int source_to = -1, source_from = -1;
    for (const Edge& e : g) {
        if (e.p == x && e.c != a) source_to = e.c;
        if (e.c == x) source_from = e.p;
    }
    Graph t;
    for (const Edge& e : g)
        if (e.p != a && e.c != a && e.p != x && e.c != x) t.push_back(e);
    if (source_from >= 0 && source_to >= 0) t.push_back({source_from, source_to});
    if (dest_from >= 0 && a_child >= 0) t.push_back({dest_from, a_child});
    return t;
}

// Nearest-neighbour interchanges around each internal edge u -> v: v's child k swaps
// places with v's sibling.
void nni(const Graph& tree, int nleaf, const std::vector<std::string>& leaves, std::unordered_set<std::uint64_t>& seen,
         std::vector<Candidate>& out) {
    const int E = static_cast<int>(tree.size());
    auto children = [&](int u) {
        std::vector<int> ch;
        for (const Edge& e : tree)
            if (e.p == u) ch.push_back(e.c);
        return ch;
    };
    auto find = [&](int p, int c) {
        for (int j = 0; j < E; ++j)
            if (tree[j].p == p && tree[j].c == c) return j;
        return -1;
    };
    for (int ei = 0; ei < E; ++ei) {
        const int u = tree[ei].p, v = tree[ei].c;
        if (v < nleaf) continue;
        const std::vector<int> uch = children(u), vch = children(v);
        if (uch.size() != 2 || vch.size() != 2) continue;
        const int sibling = uch[0] == v ? uch[1] : uch[0];
        const int j_sib = find(u, sibling);
        if (j_sib < 0) continue;
        for (int k = 0; k < 2; ++k) {
            const int j_child = find(v, vch[k]);
            if (j_child < 0) continue;
            Graph g = tree;
            g[j_sib] = {u, vch[k]};
            g[j_child] = {v, sibling};
            const std::uint64_t h = shape_hash(g, nleaf);
            if (seen.insert(h).second) out.push_back(Candidate{named(g, leaves), 0, 0, h});
        }
    }
}

}  // namespace

std::vector<Candidate> enumerate_space(const std::vector<std::string>& leaves, int max_nadmix) {
    const int nleaf = static_cast<int>(leaves.size());
    std::vector<Candidate> out;
    if (nleaf < 1) return out;
    const std::vector<Graph> trees = all_trees(nleaf);
    out.reserve(trees.size());
    for (std::size_t i = 0; i < trees.size(); ++i)
        out.push_back(Candidate{named(trees[i], leaves), 0, static_cast<int>(i), shape_hash(trees[i], nleaf)});
    if (max_nadmix < 1) return out;
    // wirings of each tree in parallel; the first graph of each shape, in tree order, is kept
    const std::size_t T = trees.size();
    std::vector<std::vector<Graph>> gs(T);
    std::vector<std::vector<std::uint64_t>> hs(T);
    std::atomic<std::size_t> next{0};
    auto work = [&] {
        for (std::size_t t = next++; t < T; t = next++) wirings(trees[t], nleaf, gs[t], hs[t]);
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t n_threads = std::min<std::size_t>({16, hw, (T + 7) / 8});
    std::vector<std::thread> pool;
    for (std::size_t i = 1; i < n_threads; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    std::unordered_set<std::uint64_t> seen;
    int id = 0;
    for (std::size_t t = 0; t < T; ++t) {
        for (std::size_t i = 0; i < gs[t].size(); ++i)
            if (seen.insert(hs[t][i]).second) out.push_back(Candidate{named(gs[t][i], leaves), 1, id++, hs[t][i]});
        std::vector<Graph>().swap(gs[t]);
    }
    return out;
}

std::uint64_t graph_hash(const NamedEdges& edges) {
    std::unordered_set<std::string> parents, all;
    for (const auto& [a, b] : edges) {
        parents.insert(a);
        all.insert(a);
        all.insert(b);
    }
    std::vector<std::string> leaf_names;
    for (const auto& n : all)
        if (!parents.count(n)) leaf_names.push_back(n);
    std::sort(leaf_names.begin(), leaf_names.end());
    return shape_hash(to_ids(edges, leaf_names), static_cast<int>(leaf_names.size()));
}

std::vector<Candidate> neighbors(const Candidate& current, const std::vector<std::string>& leaves, int max_nadmix) {
    const int nleaf = static_cast<int>(leaves.size());
    std::vector<Candidate> out;
    std::unordered_set<std::uint64_t> seen{current.hash};
    const Graph g = to_ids(current.edges, leaves);
    const Graph base = current.nadmix == 0 ? g : base_tree(g);
    nni(base, nleaf, leaves, seen, out);
    if (max_nadmix >= 1) {
        keep_new_wirings(base, nleaf, leaves, seen, out);
        std::unordered_set<std::uint64_t> nni_seen{shape_hash(base, nleaf)};
        std::vector<Candidate> moved;
        nni(base, nleaf, leaves, nni_seen, moved);
        for (const Candidate& c : moved) keep_new_wirings(to_ids(c.edges, leaves), nleaf, leaves, seen, out);
    }
    return out;
}

}  // namespace stp::qgenum
