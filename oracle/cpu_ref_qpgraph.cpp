// oracle/cpu_ref_qpgraph.cpp — CPU restatement of the reference's qpGraph fit.
//
// TEST INFRASTRUCTURE ONLY (see cpu_ref.cpp): the parity checker for the HIP qpGraph
// score / adjoint kernels and their optimiser.  Restated from the reference's
// semantics (paths relative to the reference repository root):
//
//   graph model            src/core/qpadm/qpgraph_model.cpp:41-243 (parse_qpgraph)
//   leaf weights           src/core/qpadm/qpgraph_model.hpp:73-110 (fill_pwts_centered)
//   pair products          src/core/qpadm/qpgraph_objective.hpp:24-38 (build_ppwts_2d)
//   NNLS                   src/core/qpadm/qpgraph_objective.hpp:41-100 (Lawson-Hanson active set)
//   edge lengths           src/core/qpadm/qpgraph_objective.hpp:103-160 (opt_edge_lengths)
//   score                  src/core/qpadm/qpgraph_objective.hpp:163-190
//   f3 basis               src/device/cpu/cpu_backend.cpp:2170-2217 (assemble_f3_triples)
//   fit driver             src/core/qpadm/qpgraph_fit.cpp:45-118
//   multistart optimiser   src/device/cpu/cpu_backend.cpp:2936-3045 (qpgraph_fit_fleet),
//                          constants src/core/qpadm/qpgraph_opt_constants.hpp:18-40

#include <cstdint>
#include <functional>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "cref_common.hpp"

namespace {

using namespace cref;

struct Graph {
    int npop = 0, ne = 0, nadmix = 0, npair = 0, npath = 0, base = 0, nedge_total = 0;
    std::vector<std::string> leaves;
    std::vector<int> leaf_to_f2;
    Vec pwts0;                                  // (e, leaf) at e + ne*leaf
    std::vector<int> pe_edge, pe_leaf, pe_path;  // variable (edge, leaf) cells and their paths
    std::vector<int> pae_path, pae_id;          // admixture edges on each path (id 2j+1 / 2j+2)
    std::vector<int> cmb1, cmb2;                // f3 basis: centred columns a <= b
    std::vector<std::string> admix_from, admix_to, edge_from, edge_to;
    std::string error;
    int col_to_leaf(int c) const {
        int seen = 0;
        for (int li = 0; li < npop; ++li) {
            if (li == base) continue;
            if (seen == c) return li;
            ++seen;
        }
        return -1;
    }
};

Graph parse(const std::vector<std::pair<std::string, std::string>>& edges, const std::vector<std::string>& names) {
    Graph g;
    if (edges.empty()) { g.error = "empty edge list"; return g; }
    std::vector<std::string> node;
    std::unordered_map<std::string, int> id;
    auto get = [&](const std::string& s) {
        auto it = id.find(s);
        if (it != id.end()) return it->second;
        id.emplace(s, static_cast<int>(node.size()));
        node.push_back(s);
        return static_cast<int>(node.size()) - 1;
    };
    const int E = static_cast<int>(edges.size());
    std::vector<int> par(E), chi(E);
    for (int e = 0; e < E; ++e) {
        if (edges[e].first.empty() || edges[e].second.empty()) { g.error = "empty endpoint"; return g; }
        par[e] = get(edges[e].first);
        chi[e] = get(edges[e].second);
    }
    const int V = static_cast<int>(node.size());
    g.nedge_total = E;
    std::vector<int> indeg(V, 0), outdeg(V, 0);
    std::vector<std::vector<int>> out(V), in(V);
    for (int e = 0; e < E; ++e) {
        ++outdeg[par[e]];
        ++indeg[chi[e]];
        out[par[e]].push_back(e);
        in[chi[e]].push_back(e);
    }
    int root = -1;
    for (int v = 0; v < V; ++v) {
        if (indeg[v] == 0) {
            if (root != -1) { g.error = "multiple roots"; return g; }
            root = v;
        }
        if (indeg[v] > 2) { g.error = "in-degree > 2"; return g; }
    }
    if (root < 0) { g.error = "no root"; return g; }
    std::vector<int> leaf_nodes;
    for (int v = 0; v < V; ++v)
        if (outdeg[v] == 0) leaf_nodes.push_back(v);
    g.npop = static_cast<int>(leaf_nodes.size());
    g.npair = g.npop * (g.npop - 1) / 2;
    std::vector<int> node_leaf(V, -1);
    for (int li = 0; li < g.npop; ++li) node_leaf[leaf_nodes[li]] = li;
    std::unordered_map<std::string, int> f2idx;
    for (int i = 0; i < static_cast<int>(names.size()); ++i) f2idx.emplace(names[i], i);
    g.leaves.resize(g.npop);
    g.leaf_to_f2.assign(g.npop, -1);
    for (int li = 0; li < g.npop; ++li) {
        g.leaves[li] = node[leaf_nodes[li]];
        auto it = f2idx.find(g.leaves[li]);
        if (it == f2idx.end()) { g.error = "leaf not in the f2 population set"; return g; }
        g.leaf_to_f2[li] = it->second;
    }
    g.base = 0;
    std::vector<int> admix_nodes;
    for (int v = 0; v < V; ++v)
        if (indeg[v] == 2) admix_nodes.push_back(v);
    g.nadmix = static_cast<int>(admix_nodes.size());
    std::vector<int> aid(E, -1);
    for (int j = 0; j < g.nadmix; ++j) {
        const int e0 = in[admix_nodes[j]][0], e1 = in[admix_nodes[j]][1];
        aid[e0] = 2 * j + 1;
        aid[e1] = 2 * j + 2;
        g.admix_from.push_back(node[par[e0]]);
        g.admix_to.push_back(node[chi[e0]]);
    }
    std::vector<int> norm(E, -1);
    for (int e = 0; e < E; ++e)
        if (aid[e] == -1) {
            norm[e] = g.ne++;
            g.edge_from.push_back(node[par[e]]);
            g.edge_to.push_back(node[chi[e]]);
        }
    // root-to-leaf paths in depth-first order over each node's out-edges
    std::vector<std::vector<int>> paths;
    std::vector<int> pleaf, stack;
    std::vector<char> on(V, 0);
    bool cycle = false;
    std::function<void(int)> dfs = [&](int v) {
        if (cycle) return;
        if (outdeg[v] == 0) {
            paths.push_back(stack);
            pleaf.push_back(node_leaf[v]);
            return;
        }
        on[v] = 1;
        for (int e : out[v]) {
            if (on[chi[e]]) { cycle = true; break; }
            stack.push_back(e);
            dfs(chi[e]);
            stack.pop_back();
            if (cycle) break;
        }
        on[v] = 0;
    };
    dfs(root);
    if (cycle) { g.error = "cycle"; return g; }
    g.npath = static_cast<int>(paths.size());
    std::vector<int> pathcount(g.npop, 0);
    for (int l : pleaf) ++pathcount[l];
    g.pwts0.assign(static_cast<std::size_t>(g.ne) * g.npop, 0.0);
    for (int pi = 0; pi < g.npath; ++pi) {
        const double inc = 1.0 / static_cast<double>(pathcount[pleaf[pi]]);
        for (int e : paths[pi])
            if (norm[e] >= 0) g.pwts0[norm[e] + static_cast<std::size_t>(g.ne) * pleaf[pi]] += inc;
    }
    for (int pi = 0; pi < g.npath; ++pi)
        for (int e : paths[pi])
            if (aid[e] > 0) {
                g.pae_path.push_back(pi);
                g.pae_id.push_back(aid[e]);
            }
    std::map<std::pair<int, int>, int> cnt;
    for (int pi = 0; pi < g.npath; ++pi)
        for (int e : paths[pi])
            if (norm[e] >= 0) ++cnt[{norm[e], pleaf[pi]}];
    for (int pi = 0; pi < g.npath; ++pi)
        for (int e : paths[pi])
            if (norm[e] >= 0 && cnt[{norm[e], pleaf[pi]}] < pathcount[pleaf[pi]]) {
                g.pe_edge.push_back(norm[e]);
                g.pe_leaf.push_back(pleaf[pi]);
                g.pe_path.push_back(pi);
            }
    for (int a = 0; a < g.npop - 1; ++a)
        for (int b = a; b < g.npop - 1; ++b) {
            g.cmb1.push_back(a);
            g.cmb2.push_back(b);
        }
    return g;
}

// Centred leaf weights (ne x (npop-1)) for mixture weights theta.
Vec pwts_centred(const Graph& g, const double* theta) {
    Vec path_w(g.npath, 1.0);
    for (std::size_t t = 0; t < g.pae_path.size(); ++t) {
        const int id = g.pae_id[t];
        const double w = theta[(id - 1) / 2];
        path_w[g.pae_path[t]] *= (id % 2 == 1) ? w : (1.0 - w);
    }
    Vec pw = g.pwts0;
    const std::size_t n = g.pe_edge.size();
    for (std::size_t a = 0; a < n; ++a) {
        const int e = g.pe_edge[a], l = g.pe_leaf[a];
        bool first = true;
        for (std::size_t b = 0; b < a; ++b)
            if (g.pe_edge[b] == e && g.pe_leaf[b] == l) { first = false; break; }
        if (!first) continue;
        double s = 0.0;
        for (std::size_t b = a; b < n; ++b)
            if (g.pe_edge[b] == e && g.pe_leaf[b] == l) s += path_w[g.pe_path[b]];
        pw[e + static_cast<std::size_t>(g.ne) * l] = s;
    }
    Vec pc(static_cast<std::size_t>(g.ne) * (g.npop - 1), 0.0);
    int col = 0;
    for (int l = 0; l < g.npop; ++l) {
        if (l == g.base) continue;
        for (int e = 0; e < g.ne; ++e)
            pc[e + static_cast<std::size_t>(g.ne) * col] =
                pw[e + static_cast<std::size_t>(g.ne) * l] - pw[e + static_cast<std::size_t>(g.ne) * g.base];
        ++col;
    }
    return pc;
}

// ppwts[k + np*e] = pc[e, cmb1[k]] * pc[e, cmb2[k]]
Vec pair_products(const Graph& g, const Vec& pc) {
    Vec pp(static_cast<std::size_t>(g.npair) * g.ne, 0.0);
    for (int k = 0; k < g.npair; ++k)
        for (int e = 0; e < g.ne; ++e)
            pp[k + static_cast<std::size_t>(g.npair) * e] =
                pc[e + static_cast<std::size_t>(g.ne) * g.cmb1[k]] * pc[e + static_cast<std::size_t>(g.ne) * g.cmb2[k]];
    return pp;
}

// Lawson-Hanson active set on A (n x n, column-major) and q; tolerance 1e-12.
bool nnls(const Vec& A, const Vec& q, int n, Vec& bl, std::vector<char>* passive_out = nullptr) {
    bl.assign(n, 0.0);
    std::vector<char> passive(n, 0);
    const double eps = 1e-12;
    const int max_iter = 3 * n + 30;
    for (int outer = 0; outer < max_iter; ++outer) {
        Vec w(n, 0.0);
        for (int i = 0; i < n; ++i) {
            double s = q[i];
            for (int j = 0; j < n; ++j) s -= A[i + static_cast<std::size_t>(n) * j] * bl[j];
            w[i] = s;
        }
        int t = -1;
        double best = eps;
        for (int i = 0; i < n; ++i)
            if (!passive[i] && w[i] > best) { best = w[i]; t = i; }
        if (t < 0) break;
        passive[t] = 1;
        for (int inner = 0; inner < max_iter; ++inner) {
            std::vector<int> P;
            for (int i = 0; i < n; ++i)
                if (passive[i]) P.push_back(i);
            const int p = static_cast<int>(P.size());
            Vec Ap(static_cast<std::size_t>(p) * p), qp(p), z;
            for (int a = 0; a < p; ++a) {
                qp[a] = q[P[a]];
                for (int b = 0; b < p; ++b) Ap[a + static_cast<std::size_t>(p) * b] = A[P[a] + static_cast<std::size_t>(n) * P[b]];
            }
            if (!lin_solve(Ap, p, qp, z)) return false;
            bool all_pos = true;
            for (int a = 0; a < p; ++a)
                if (z[a] <= eps) { all_pos = false; break; }
            if (all_pos) {
                for (int i = 0; i < n; ++i) bl[i] = 0.0;
                for (int a = 0; a < p; ++a) bl[P[a]] = z[a];
                break;
            }
            double alpha = 1.0;
            for (int a = 0; a < p; ++a)
                if (z[a] <= eps) {
                    const double denom = bl[P[a]] - z[a];
                    if (denom > eps) {
                        const double r = bl[P[a]] / denom;
                        if (r < alpha) alpha = r;
                    }
                }
            for (int a = 0; a < p; ++a) bl[P[a]] += alpha * (z[a] - bl[P[a]]);
            for (int i = 0; i < n; ++i)
                if (passive[i] && bl[i] <= eps) {
                    passive[i] = 0;
                    bl[i] = 0.0;
                }
        }
    }
    if (passive_out) *passive_out = passive;
    return true;
}

// Ridge-regularised, diagonally scaled GLS for the edge lengths (constrained: NNLS).
bool edge_lengths(const Vec& pp, const Vec& qinv, const Vec& f, int np, int ne, double fudge, bool constrained,
                  Vec& bl) {
    Vec W(static_cast<std::size_t>(np) * ne, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int r = 0; r < np; ++r) {
            double acc = 0.0;
            for (int c = 0; c < np; ++c) acc += qinv[r + static_cast<std::size_t>(np) * c] * pp[c + static_cast<std::size_t>(np) * e];
            W[r + static_cast<std::size_t>(np) * e] = acc;
        }
    Vec cc(static_cast<std::size_t>(ne) * ne, 0.0), q(ne, 0.0);
    for (int e1 = 0; e1 < ne; ++e1) {
        for (int e2 = 0; e2 < ne; ++e2) {
            double acc = 0.0;
            for (int r = 0; r < np; ++r) acc += pp[r + static_cast<std::size_t>(np) * e1] * W[r + static_cast<std::size_t>(np) * e2];
            cc[e1 + static_cast<std::size_t>(ne) * e2] = acc;
        }
        double rr = 0.0;
        for (int r = 0; r < np; ++r) {
            double qf = 0.0;
            for (int c = 0; c < np; ++c) qf += qinv[r + static_cast<std::size_t>(np) * c] * f[c];
            rr += pp[r + static_cast<std::size_t>(np) * e1] * qf;
        }
        q[e1] = rr;
    }
    double trm = 0.0;
    for (int e = 0; e < ne; ++e) trm += cc[e + static_cast<std::size_t>(ne) * e];
    const double ridge = fudge * trm / static_cast<double>(ne);
    for (int e = 0; e < ne; ++e) cc[e + static_cast<std::size_t>(ne) * e] += ridge;
    Vec sc(ne), ccs(static_cast<std::size_t>(ne) * ne), q1(ne);
    for (int e = 0; e < ne; ++e) sc[e] = std::sqrt(cc[e + static_cast<std::size_t>(ne) * e]);
    for (int e1 = 0; e1 < ne; ++e1) {
        q1[e1] = q[e1] / sc[e1];
        for (int e2 = 0; e2 < ne; ++e2)
            ccs[e1 + static_cast<std::size_t>(ne) * e2] = cc[e1 + static_cast<std::size_t>(ne) * e2] / (sc[e1] * sc[e2]);
    }
    Vec bls;
    if (constrained) {
        if (!nnls(ccs, q1, ne, bls)) return false;
    } else if (!lin_solve(ccs, ne, q1, bls)) {
        return false;
    }
    bl.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e) bl[e] = bls[e] / sc[e];
    return true;
}

double score(const Graph& g, const double* theta, const Vec& f, const Vec& qinv, double fudge, bool constrained,
             Vec* bl_out = nullptr, Vec* fit_out = nullptr) {
    const Vec pc = pwts_centred(g, theta);
    const Vec pp = pair_products(g, pc);
    Vec bl;
    if (!edge_lengths(pp, qinv, f, g.npair, g.ne, fudge, constrained, bl)) return std::numeric_limits<double>::infinity();
    const int np = g.npair;
    Vec res(np);
    if (fit_out) fit_out->assign(np, 0.0);
    for (int k = 0; k < np; ++k) {
        double fit = 0.0;
        for (int e = 0; e < g.ne; ++e) fit += pp[k + static_cast<std::size_t>(np) * e] * bl[e];
        res[k] = f[k] - fit;
        if (fit_out) (*fit_out)[k] = fit;
    }
    double s = 0.0;
    for (int a = 0; a < np; ++a) {
        double row = 0.0;
        for (int b = 0; b < np; ++b) row += qinv[a + static_cast<std::size_t>(np) * b] * res[b];
        s += res[a] * row;
    }
    if (bl_out) *bl_out = bl;
    return s;
}

// f3 basis X[k][b] = (f2(c,a) + f2(c,b) - f2(a,b)) / 2 over surviving blocks, with LOO/total.
F4 f3_basis(const Graph& g, const double* f2, const double* vpair, const int* bsz, int P, int nb) {
    F4 x;
    x.nl = g.npair;
    x.nr = 1;
    const std::vector<int> surv = survivors(vpair, P, nb);
    x.nb = static_cast<int>(surv.size());
    for (int s : surv) x.bsize.push_back(bsz[s]);
    const int m = g.npair;
    x.x_blocks.assign(static_cast<std::size_t>(m) * x.nb, 0.0);
    const std::size_t slab = static_cast<std::size_t>(P) * P;
    auto F = [&](int i, int j, int b) { return f2[i + static_cast<std::size_t>(P) * j + slab * b]; };
    const int pc = g.leaf_to_f2[g.base];
    for (int k = 0; k < m; ++k) {
        const int pa = g.leaf_to_f2[g.col_to_leaf(g.cmb1[k])];
        const int pb = g.leaf_to_f2[g.col_to_leaf(g.cmb2[k])];
        for (int s = 0; s < x.nb; ++s) {
            const int b = surv[s];
            x.x_blocks[k + static_cast<std::size_t>(m) * s] = 0.5 * (F(pc, pa, b) + F(pc, pb, b) - F(pa, pb, b));
        }
    }
    if (x.nb > 0) loo_and_total(x);
    return x;
}

namespace opt {
constexpr unsigned long long kInstMul = 0x100000001B3ULL;
constexpr unsigned long long kDimMul = 0x9E3779B97F4A7C15ULL;
constexpr unsigned long long kSeedInc = 0xD1B54A32D192ED03ULL;
constexpr unsigned long long kMix1 = 0xBF58476D1CE4E5B9ULL;
constexpr unsigned long long kMix2 = 0x94D049BB133111EBULL;
constexpr unsigned long long kMantMask = 0xFFFFFFFFFFFFFULL;
constexpr double kMantDiv = 4503599627370496.0;
constexpr double kFdStep = 1e-4, kCurvGuard = 1e-30, kCurvHalf = 0.5, kCurvThresh = 1e-8;
constexpr double kGradStepScale = 0.5, kTrustClamp = 0.5, kBacktrackHalf = 0.5;
constexpr int kMaxBacktrack = 8;
constexpr double kTolDxScale = 1e-2, kTolDsScale = 1e-3;
}  // namespace opt

double init_theta(unsigned inst, int dim) {
    unsigned long long z = static_cast<unsigned long long>(inst) * opt::kInstMul +
                           static_cast<unsigned long long>(dim) * opt::kDimMul + opt::kSeedInc;
    z = (z ^ (z >> 30)) * opt::kMix1;
    z = (z ^ (z >> 27)) * opt::kMix2;
    z = z ^ (z >> 31);
    return static_cast<double>(z & opt::kMantMask) / opt::kMantDiv;
}

struct Fleet {
    double score = 0.0, spread = 0.0;
    Vec theta, lo, hi, bl, fit;
    bool ok = true;
};

// Splitmix multistart; per restart, coordinate projected Newton steps from central
// differences with backtracking, until the largest coordinate and score changes fall
// below the tolerances.
Fleet fleet(const Graph& g, const Vec& f, const Vec& qinv, double fudge, bool constrained, int numstart, int maxit,
            double tol) {
    Fleet out;
    const int D = g.nadmix;
    auto clamp01 = [](double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); };
    auto sco = [&](const Vec& th) { return score(g, th.data(), f, qinv, fudge, constrained); };
    if (D == 0) {
        out.score = score(g, nullptr, f, qinv, fudge, constrained, &out.bl, &out.fit);
        out.ok = std::isfinite(out.score);
        return out;
    }
    double best = std::numeric_limits<double>::infinity(), smin = best, smax = -best;
    Vec best_th(D, 0.0), thmin(D, 1.0), thmax(D, 0.0);
    const double h = opt::kFdStep;
    for (int inst = 0; inst < numstart; ++inst) {
        Vec th(D);
        for (int d = 0; d < D; ++d) th[d] = clamp01(init_theta(static_cast<unsigned>(inst), d));
        double s = sco(th);
        for (int it = 0; it < maxit; ++it) {
            double max_dx = 0.0, max_ds = 0.0;
            for (int d = 0; d < D; ++d) {
                const double w = th[d];
                const double wp = std::min(1.0, w + h), wm = std::max(0.0, w - h);
                Vec thp = th, thm = th;
                thp[d] = wp;
                thm[d] = wm;
                const double sp = sco(thp), sm = sco(thm);
                const double dwp = wp - w, dwm = w - wm;
                double gr, curv;
                if (dwp > 0.0 && dwm > 0.0) {
                    gr = (sp - sm) / (dwp + dwm);
                    curv = (sp - 2.0 * s + sm) / (opt::kCurvHalf * (dwp + dwm) * (dwp + dwm) + opt::kCurvGuard);
                } else if (dwp > 0.0) {
                    gr = (sp - s) / dwp;
                    curv = 1.0;
                } else {
                    gr = (s - sm) / dwm;
                    curv = 1.0;
                }
                double step = (curv > opt::kCurvThresh) ? (gr / curv) : (gr * opt::kGradStepScale);
                if (step > opt::kTrustClamp) step = opt::kTrustClamp;
                if (step < -opt::kTrustClamp) step = -opt::kTrustClamp;
                double wn = clamp01(w - step);
                Vec thn = th;
                thn[d] = wn;
                double sn = sco(thn);
                int bt = 0;
                while (sn > s && bt < opt::kMaxBacktrack) {
                    wn = opt::kBacktrackHalf * (wn + w);
                    thn[d] = wn;
                    sn = sco(thn);
                    ++bt;
                }
                const double dx = std::fabs(wn - w), ds = std::fabs(sn - s);
                if (sn <= s) {
                    th[d] = wn;
                    s = sn;
                }
                if (dx > max_dx) max_dx = dx;
                if (ds > max_ds) max_ds = ds;
            }
            if (max_dx < tol * opt::kTolDxScale && max_ds < tol * opt::kTolDsScale) break;
        }
        if (s < smin) smin = s;
        if (s > smax) smax = s;
        for (int d = 0; d < D; ++d) {
            if (th[d] < thmin[d]) thmin[d] = th[d];
            if (th[d] > thmax[d]) thmax[d] = th[d];
        }
        if (s < best) {
            best = s;
            best_th = th;
        }
    }
    out.score = best;
    out.spread = (std::isfinite(smax) && std::isfinite(smin)) ? (smax - smin) : 0.0;
    out.theta = best_th;
    out.lo = thmin;
    out.hi = thmax;
    const double fin = score(g, best_th.data(), f, qinv, fudge, constrained, &out.bl, &out.fit);
    out.ok = std::isfinite(fin);
    return out;
}

std::vector<std::pair<std::string, std::string>> edge_list(int n, const char* const* from, const char* const* to) {
    std::vector<std::pair<std::string, std::string>> e;
    for (int i = 0; i < n; ++i) e.emplace_back(from[i], to[i]);
    return e;
}

std::vector<std::string> name_list(int n, const char* const* names) {
    std::vector<std::string> v;
    for (int i = 0; i < n; ++i) v.emplace_back(names[i]);
    return v;
}

template <class T>
void copy_out(T* dst, const std::vector<T>& v) {
    if (dst)
        for (std::size_t i = 0; i < v.size(); ++i) dst[i] = v[i];
}

}  // namespace

extern "C" {

// Model dimensions: dims = [npop, nedge_norm, nadmix, npair, npath, base_leaf, n_pe, n_pae].
// Returns 0, or -1 when the graph is invalid.
int cref_qpgraph_model(int n_edges, const char* const* from, const char* const* to, int n_names,
                       const char* const* names, std::int32_t* dims) {
    const Graph g = parse(edge_list(n_edges, from, to), name_list(n_names, names));
    if (!g.error.empty()) return -1;
    const std::int32_t d[8] = {g.npop, g.ne, g.nadmix, g.npair, g.npath, g.base,
                               static_cast<std::int32_t>(g.pe_edge.size()), static_cast<std::int32_t>(g.pae_path.size())};
    for (int i = 0; i < 8; ++i) dims[i] = d[i];
    return 0;
}

// The model tables (sizes from cref_qpgraph_model): pwts0 (ne*npop), pe_edge/pe_leaf/
// pe_path (n_pe), pae_path/pae_id (n_pae), cmb1/cmb2 (npair), leaf_to_f2 (npop).
int cref_qpgraph_tables(int n_edges, const char* const* from, const char* const* to, int n_names,
                        const char* const* names, double* pwts0, std::int32_t* pe_edge, std::int32_t* pe_leaf,
                        std::int32_t* pe_path, std::int32_t* pae_path, std::int32_t* pae_id, std::int32_t* cmb1,
                        std::int32_t* cmb2, std::int32_t* leaf_to_f2) {
    const Graph g = parse(edge_list(n_edges, from, to), name_list(n_names, names));
    if (!g.error.empty()) return -1;
    copy_out(pwts0, g.pwts0);
    auto ci = [](std::int32_t* d, const std::vector<int>& v) {
        if (d)
            for (std::size_t i = 0; i < v.size(); ++i) d[i] = v[i];
    };
    ci(pe_edge, g.pe_edge);
    ci(pe_leaf, g.pe_leaf);
    ci(pe_path, g.pe_path);
    ci(pae_path, g.pae_path);
    ci(pae_id, g.pae_id);
    ci(cmb1, g.cmb1);
    ci(cmb2, g.cmb2);
    ci(leaf_to_f2, g.leaf_to_f2);
    return 0;
}

// f3 basis and its jackknife covariance: f_obs (npair), Q and Qinv (npair x npair,
// column-major) with the diag_f3 ridge; returns 0, -1 (graph), -2 (no blocks), -3 (non-SPD).
int cref_qpgraph_basis(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, int n_edges,
                       const char* const* from, const char* const* to, int n_names, const char* const* names,
                       double diag_f3, double* f_obs, double* Q, double* qinv) {
    const Graph g = parse(edge_list(n_edges, from, to), name_list(n_names, names));
    if (!g.error.empty()) return -1;
    std::vector<int> bs(bsz, bsz + nb);
    const F4 x = f3_basis(g, f2, vpair, bs.data(), P, nb);
    if (g.npair <= 0 || x.nb <= 0) return -2;
    const Cov c = jackknife_cov(x, diag_f3);
    copy_out(f_obs, x.x_total);
    copy_out(Q, c.Q);
    copy_out(qinv, c.Qinv);
    return c.ok ? 0 : -3;
}

// The score at mixture weights theta (nadmix values; may be NULL when nadmix = 0);
// bl (nedge_norm) and fit (npair) are optional outputs.
double cref_qpgraph_score(int n_edges, const char* const* from, const char* const* to, int n_names,
                          const char* const* names, const double* theta, const double* f_obs, const double* qinv,
                          double fudge, int constrained, double* bl, double* fit) {
    const Graph g = parse(edge_list(n_edges, from, to), name_list(n_names, names));
    if (!g.error.empty()) return std::numeric_limits<double>::quiet_NaN();
    const Vec f(f_obs, f_obs + g.npair);
    const Vec qi(qinv, qinv + static_cast<std::size_t>(g.npair) * g.npair);
    Vec blv, fitv;
    const double s = score(g, theta, f, qi, fudge, constrained != 0, &blv, &fitv);
    copy_out(bl, blv);
    copy_out(fit, fitv);
    return s;
}

// The whole single-graph fit (run_qpgraph with the CPU optimiser).  Outputs: scal =
// [score, restart_spread, worst_residual_z, worst_pair_index, status], theta / lo / hi
// (nadmix), bl (nedge_norm), fit (npair).  status: 0 ok, 3 non-SPD, 6 invalid graph.
int cref_qpgraph_fit(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, int n_edges,
                     const char* const* from, const char* const* to, int n_names, const char* const* names,
                     double fudge, double diag_f3, int numstart, int constrained, int maxit, double tol, double* scal,
                     double* theta, double* lo, double* hi, double* bl, double* fit) {
    for (int i = 0; i < 5; ++i) scal[i] = 0.0;
    const Graph g = parse(edge_list(n_edges, from, to), name_list(n_names, names));
    if (!g.error.empty()) { scal[4] = 6; return 0; }
    std::vector<int> bs(bsz, bsz + nb);
    const F4 x = f3_basis(g, f2, vpair, bs.data(), P, nb);
    if (g.npair <= 0 || x.nb <= 0) { scal[4] = 3; return 0; }
    const Cov c = jackknife_cov(x, diag_f3);
    if (!c.ok) { scal[4] = 3; return 0; }
    const Fleet fl = fleet(g, x.x_total, c.Qinv, fudge, constrained != 0, numstart, maxit, tol);
    if (!fl.ok) { scal[4] = 3; return 0; }
    scal[0] = fl.score;
    scal[1] = fl.spread;
    copy_out(theta, fl.theta);
    copy_out(lo, fl.lo);
    copy_out(hi, fl.hi);
    copy_out(bl, fl.bl);
    copy_out(fit, fl.fit);
    double worst = 0.0;
    int wk = -1;
    for (int k = 0; k < g.npair; ++k) {
        const double qkk = c.Q[k + static_cast<std::size_t>(g.npair) * k];
        const double se = qkk > 0.0 ? std::sqrt(qkk) : std::nan("");
        const double z = (x.x_total[k] - fl.fit[k]) / se;
        if (std::isfinite(z) && std::fabs(z) > std::fabs(worst)) {
            worst = z;
            wk = k;
        }
    }
    scal[2] = worst;
    scal[3] = wk;
    return 0;
}

}  // extern "C"
