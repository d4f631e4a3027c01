// steppe_amd/csrc/qpgraph_host.cpp — qpGraph host side: graph parsing, the device
// topology arena, and the resident graph sets the score / adjoint kernels evaluate.
//
// Parsing follows the reference's parse_qpgraph (src/core/qpadm/qpgraph_model.cpp:41-243):
// nodes are registered in first-seen order over the edge list; the root is the one node
// without parents; leaves are the nodes without children, in node order; admixture
// nodes have two parents and their first-listed incoming edge carries theta_j (id 2j+1),
// the second 1 - theta_j (id 2j+2); the remaining edges are the normal edges in input
// order; root-to-leaf paths are enumerated depth-first over each node's out-edges in
// input order; pwts0 adds 1/pathcount(leaf) per path through a normal edge; a
// (edge, leaf) cell crossed by fewer than all of the leaf's paths is variable and is the
// sum of its paths' weights; f3 basis pairs are the centred columns a <= b.
#include <algorithm>
#include <cmath>
#include <limits>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <unordered_map>

#include "engine.hpp"
#include "qpgraph.hpp"

namespace stp {

int QgTopology::col_to_leaf(int c) const { return c < base ? c : c + 1; }

QgTopology parse_qpgraph(const std::vector<std::pair<std::string, std::string>>& edges,
                         const std::vector<std::string>& names, const std::string& base_leaf) {
    STP_REQUIRE(!edges.empty(), "qpgraph: empty edge list");
    QgTopology t;
    std::vector<std::string> node;
    std::unordered_map<std::string, int> index;
    auto node_id = [&](const std::string& s) {
        const auto it = index.find(s);
        if (it != index.end()) return it->second;
        index.emplace(s, static_cast<int>(node.size()));
        node.push_back(s);
        return static_cast<int>(node.size()) - 1;
    };
    const int E = static_cast<int>(edges.size());
    std::vector<int> from(E), to(E);
    for (int e = 0; e < E; ++e) {
        STP_REQUIRE(!edges[e].first.empty() && !edges[e].second.empty(), "qpgraph: an edge has an empty endpoint");
        from[e] = node_id(edges[e].first);
        to[e] = node_id(edges[e].second);
    }
    const int V = static_cast<int>(node.size());
    std::vector<std::vector<int>> out(V), in(V);
    for (int e = 0; e < E; ++e) {
        out[from[e]].push_back(e);
        in[to[e]].push_back(e);
    }
    int root = -1;
    for (int v = 0; v < V; ++v) {
        if (in[v].empty()) {
            STP_REQUIRE(root == -1, "qpgraph: multiple roots (graph is not singly-rooted)");
            root = v;
        }
        STP_REQUIRE(in[v].size() <= 2, "qpgraph: node '" + node[v] + "' has in-degree > 2 (not a valid admixture node)");
    }
    STP_REQUIRE(root != -1, "qpgraph: no root (every node has a parent — graph is cyclic)");
    std::vector<int> leaf_of(V, -1);
    for (int v = 0; v < V; ++v)
        if (out[v].empty()) {
            leaf_of[v] = t.npop++;
            t.leaves.push_back(node[v]);
        }
    STP_REQUIRE(t.npop > 0, "qpgraph: no leaves");
    t.npair = t.npop * (t.npop - 1) / 2;
    std::unordered_map<std::string, int> name_index;
    for (int i = 0; i < static_cast<int>(names.size()); ++i) name_index.emplace(names[i], i);
    for (const auto& l : t.leaves) {
        const auto it = name_index.find(l);
        STP_REQUIRE(it != name_index.end(), "qpgraph: leaf '" + l + "' is not in the f2 population set");
        t.leaf_to_f2.push_back(it->second);
    }
    t.base = 0;
    if (!base_leaf.empty()) {
        const auto it = std::find(t.leaves.begin(), t.leaves.end(), base_leaf);
        STP_REQUIRE(it != t.leaves.end(), "qpgraph: f3basepop '" + base_leaf + "' is not a leaf");
        t.base = static_cast<int>(it - t.leaves.begin());
    }
    // admixture edges and normal edges
    std::vector<int> admix_id(E, 0), normal_id(E, -1);
    for (int v = 0; v < V; ++v)
        if (in[v].size() == 2) {
            const int j = t.nadmix++;
            admix_id[in[v][0]] = 2 * j + 1;
            admix_id[in[v][1]] = 2 * j + 2;
            t.admix_from.push_back(node[from[in[v][0]]]);
            t.admix_to.push_back(node[to[in[v][0]]]);
        }
    for (int e = 0; e < E; ++e)
        if (admix_id[e] == 0) {
            normal_id[e] = t.ne++;
            t.edge_from.push_back(node[from[e]]);
            t.edge_to.push_back(node[to[e]]);
        }
    // root-to-leaf paths, depth-first
    std::vector<std::vector<int>> paths;
    std::vector<int> path_leaf, trail;
    std::vector<char> on_trail(V, 0);
    bool cycle = false;
    std::function<void(int)> walk = [&](int v) {
        if (cycle) return;
        if (out[v].empty()) {
            paths.push_back(trail);
            path_leaf.push_back(leaf_of[v]);
            return;
        }
        on_trail[v] = 1;
        for (int e : out[v]) {
            if (on_trail[to[e]]) {
                cycle = true;
                break;
            }
            trail.push_back(e);
            walk(to[e]);
            trail.pop_back();
            if (cycle) break;
        }
        on_trail[v] = 0;
    };
    walk(root);
    STP_REQUIRE(!cycle, "qpgraph: cycle reachable from root (graph is not a DAG)");
    t.npath = static_cast<int>(paths.size());
    std::vector<int> per_leaf(t.npop, 0);
    for (int l : path_leaf) ++per_leaf[l];
    t.pwts0.assign(static_cast<std::size_t>(t.ne) * t.npop, 0.0);
    for (int p = 0; p < t.npath; ++p) {
        const double inc = 1.0 / static_cast<double>(per_leaf[path_leaf[p]]);
        for (int e : paths[p])
            if (normal_id[e] >= 0) t.pwts0[normal_id[e] + static_cast<std::size_t>(t.ne) * path_leaf[p]] += inc;
    }
    t.pae_off.push_back(0);
    for (int p = 0; p < t.npath; ++p) {
        for (int e : paths[p])
            if (admix_id[e] > 0) t.pae_id.push_back(admix_id[e]);
        t.pae_off.push_back(static_cast<std::int32_t>(t.pae_id.size()));
    }
    // variable cells: crossed by fewer than all of the leaf's paths; each is the sum
    // of its paths' weights in path order
    std::map<std::pair<int, int>, int> crossings;
    for (int p = 0; p < t.npath; ++p)
        for (int e : paths[p])
            if (normal_id[e] >= 0) ++crossings[{normal_id[e], path_leaf[p]}];
    std::map<std::pair<int, int>, int> group_of;
    std::vector<std::vector<int>> group_paths;
    for (int p = 0; p < t.npath; ++p)
        for (int e : paths[p]) {
            const int ne_ = normal_id[e];
            if (ne_ < 0) continue;
            const std::pair<int, int> cell{ne_, path_leaf[p]};
            if (crossings[cell] >= per_leaf[path_leaf[p]]) continue;
            auto it = group_of.find(cell);
            if (it == group_of.end()) {
                it = group_of.emplace(cell, static_cast<int>(group_paths.size())).first;
                group_paths.emplace_back();
                t.grp_edge.push_back(ne_);
                t.grp_leaf.push_back(path_leaf[p]);
            }
            group_paths[it->second].push_back(p);
        }
    t.grp_off.push_back(0);
    for (const auto& g : group_paths) {
        t.grp_paths.insert(t.grp_paths.end(), g.begin(), g.end());
        t.grp_off.push_back(static_cast<std::int32_t>(t.grp_paths.size()));
    }
    for (int a = 0; a < t.npop - 1; ++a)
        for (int b = a; b < t.npop - 1; ++b) {
            t.cmb1.push_back(a);
            t.cmb2.push_back(b);
        }
    return t;
}

// ---------------------------------------------------------------------------------------
namespace {

int ne_bucket(int ne) { return ne <= 32 ? 32 : 64; }

}  // namespace

void HipBackend::qpgraph_upload_topologies(QgSet& s, const std::vector<const QgTopology*>& topos) {
    const int n = static_cast<int>(topos.size());
    STP_REQUIRE(n > 0, "qpgraph: no graphs");
    QgSet* set = &s;
    set->n = n;
    std::vector<QgTopoDev> hdr(n);
    std::vector<std::int32_t> ia;
    std::vector<double> da;
    auto put_i = [&](const std::vector<std::int32_t>& v) {
        const auto o = static_cast<std::int32_t>(ia.size());
        ia.insert(ia.end(), v.begin(), v.end());
        return o;
    };
    for (int g = 0; g < n; ++g) {
        const QgTopology& t = *topos[g];
        STP_REQUIRE(t.ne >= 1 && t.ne <= 64, "qpgraph: graphs need 1..64 normal edges on the device");
        STP_REQUIRE(t.nadmix <= 16, "qpgraph: at most 16 admixture events on the device");
        STP_REQUIRE(t.npair >= 1 && t.npair <= 256, "qpgraph: graphs need 2..23 leaves on the device");
        QgTopoDev h{};
        h.npop = t.npop;
        h.ne = t.ne;
        h.nadmix = t.nadmix;
        h.npair = t.npair;
        h.npath = t.npath;
        h.base = t.base;
        h.ngrp = static_cast<std::int32_t>(t.grp_edge.size());
        h.o_grp_edge = put_i(t.grp_edge);
        h.o_grp_leaf = put_i(t.grp_leaf);
        h.o_grp_off = put_i(t.grp_off);
        h.o_grp_paths = put_i(t.grp_paths);
        h.o_pae_off = put_i(t.pae_off);
        h.o_pae_id = put_i(t.pae_id);
        h.o_cmb1 = put_i(t.cmb1);
        h.o_cmb2 = put_i(t.cmb2);
        h.o_pwts0 = static_cast<std::int32_t>(da.size());
        da.insert(da.end(), t.pwts0.begin(), t.pwts0.end());
        hdr[g] = h;
        set->np_max = std::max(set->np_max, t.npair);
        set->ne_max = std::max(set->ne_max, t.ne);
        set->npop_max = std::max(set->npop_max, t.npop);
        set->npath_max = std::max(set->npath_max, t.npath);
        set->dmax = std::max(set->dmax, t.nadmix);
        set->topo.push_back(t);
        set->slot.push_back(g);
    }
    set->d_topo.alloc(n);
    set->d_iarena.alloc(std::max<std::size_t>(ia.size(), 1));
    set->d_darena.alloc(std::max<std::size_t>(da.size(), 1));
    STP_HIP(hipMemcpyAsync(set->d_topo.get(), hdr.data(), n * sizeof(QgTopoDev), hipMemcpyHostToDevice, stream_));
    if (!ia.empty())
        STP_HIP(hipMemcpyAsync(set->d_iarena.get(), ia.data(), ia.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                               stream_));
    if (!da.empty())
        STP_HIP(hipMemcpyAsync(set->d_darena.get(), da.data(), da.size() * sizeof(double), hipMemcpyHostToDevice,
                               stream_));
    sync();  // the host staging vectors end with this function
}

QgSet* HipBackend::qpgraph_set_create(const std::vector<const QgTopology*>& topos, const F2Blocks* const* f2,
                                      double diag_f3) {
    activate();
    auto set = std::make_unique<QgSet>();
    qpgraph_upload_topologies(*set, topos);
    const int n = set->n;
    // ---- f3 basis and covariance per graph, on the device ----
    const int NP = set->np_max;
    std::vector<std::int32_t> job_topo(n), job_P(n), job_nb(n), l2f(static_cast<std::size_t>(n) * set->npop_max, 0);
    std::vector<const double*> pf2(n), pvp(n);
    std::vector<const std::int32_t*> pbs(n);
    int nb_max = 0;
    for (int g = 0; g < n; ++g) {
        STP_REQUIRE(f2[g] != nullptr, "qpgraph: null f2 handle");
        const F2Blocks& h = *f2[g];
        STP_REQUIRE(h.tier == kTierResident,
                    "qpgraph: the f2 tensor is held in host memory or on disk; fits need it resident in HBM");
        job_topo[g] = g;
        job_P[g] = h.P;
        job_nb[g] = h.n_block;
        nb_max = std::max(nb_max, h.n_block);
        pf2[g] = h.f2.get();
        pvp[g] = h.vpair.get();
        for (int l = 0; l < topos[g]->npop; ++l) {
            STP_REQUIRE(topos[g]->leaf_to_f2[l] < h.P, "qpgraph: a leaf index is outside its f2 tensor");
            l2f[static_cast<std::size_t>(g) * set->npop_max + l] = topos[g]->leaf_to_f2[l];
        }
    }
    // per-job block sizes live in one device array
    std::vector<std::int32_t> bs_all;
    std::vector<std::size_t> bs_off(n);
    for (int g = 0; g < n; ++g) {
        bs_off[g] = bs_all.size();
        bs_all.insert(bs_all.end(), f2[g]->block_sizes.begin(), f2[g]->block_sizes.end());
    }
    DevBuf<std::int32_t> d_bs(std::max<std::size_t>(bs_all.size(), 1)), d_jobs(4 * static_cast<std::size_t>(n) + 1);
    DevBuf<std::int32_t> d_l2f(std::max<std::size_t>(l2f.size(), 1));
    STP_HIP(hipMemcpyAsync(d_bs.get(), bs_all.data(), bs_all.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                           stream_));
    for (int g = 0; g < n; ++g) pbs[g] = d_bs.get() + bs_off[g];
    std::vector<std::int32_t> jobs;
    jobs.insert(jobs.end(), job_topo.begin(), job_topo.end());
    jobs.insert(jobs.end(), job_P.begin(), job_P.end());
    jobs.insert(jobs.end(), job_nb.begin(), job_nb.end());
    STP_HIP(hipMemcpyAsync(d_jobs.get(), jobs.data(), jobs.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                           stream_));
    STP_HIP(hipMemcpyAsync(d_l2f.get(), l2f.data(), l2f.size() * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_));
    DevBuf<const double*> d_pf2(n), d_pvp(n);
    DevBuf<const std::int32_t*> d_pbs(n);
    STP_HIP(hipMemcpyAsync(d_pf2.get(), pf2.data(), n * sizeof(void*), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(d_pvp.get(), pvp.data(), n * sizeof(void*), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(d_pbs.get(), pbs.data(), n * sizeof(void*), hipMemcpyHostToDevice, stream_));
    set->fobs.alloc(static_cast<std::size_t>(n) * NP);
    set->qdiag.alloc(static_cast<std::size_t>(n) * NP);
    set->qinv.alloc(static_cast<std::size_t>(n) * NP * NP);
    set->status.alloc(n);
    QgBasisArgs b{};
    b.topo = set->d_topo.get();
    b.iarena = set->d_iarena.get();
    b.n_jobs = n;
    b.job_topo = d_jobs.get();
    b.job_P = d_jobs.get() + n;
    b.job_nb = d_jobs.get() + 2 * n;
    b.job_leaf_to_f2 = d_l2f.get();
    b.f2 = d_pf2.get();
    b.vpair = d_pvp.get();
    b.bsz = d_pbs.get();
    b.diag_f3 = diag_f3;
    b.fobs = set->fobs.get();
    b.qdiag = set->qdiag.get();
    b.qinv = set->qinv.get();
    b.status = set->status.get();
    b.np_max = NP;
    b.nb_max = std::max(nb_max, 1);
    b.npop_max = set->npop_max;
    const int n_wg = std::min(n, 2 * n_cu_);
    b.scratch_stride = 2 * static_cast<std::int64_t>(NP) * b.nb_max;
    DevBuf<double> scratch(static_cast<std::size_t>(b.scratch_stride) * n_wg);
    b.scratch = scratch.get();
    b.counter = counter_.get() + 48;
    STP_REQUIRE(qg_basis_lds_bytes(b) <= 150 * 1024, "qpgraph: f3 basis too large for the device covariance");
    launch_zero_u32(b.counter, 1, stream_);
    launch_qg_basis(b, n_wg, stream_);
    std::vector<std::int32_t> st(n);
    STP_HIP(hipMemcpyAsync(st.data(), set->status.get(), n * sizeof(std::int32_t), hipMemcpyDeviceToHost, stream_));
    sync();
    set->host_status = st;
    // instance offsets of each graph's data
    return set.release();
}

QgSet* HipBackend::qpgraph_set_from_basis(const std::vector<const QgTopology*>& topos, const double* f_obs,
                                           const double* qinv) {
    activate();
    auto set = std::make_unique<QgSet>();
    qpgraph_upload_topologies(*set, topos);
    const int n = set->n, NP = set->np_max;
    std::vector<double> fo(static_cast<std::size_t>(n) * NP, 0.0), qi(static_cast<std::size_t>(n) * NP * NP, 0.0);
    std::size_t of = 0, oq = 0;
    for (int g = 0; g < n; ++g) {
        const std::size_t m = static_cast<std::size_t>(topos[g]->npair);
        std::copy(f_obs + of, f_obs + of + m, fo.begin() + static_cast<std::ptrdiff_t>(g) * NP);
        std::copy(qinv + oq, qinv + oq + m * m, qi.begin() + static_cast<std::ptrdiff_t>(g) * NP * NP);
        of += m;
        oq += m * m;
    }
    set->fobs.alloc(fo.size());
    set->qdiag.alloc(fo.size());
    set->qinv.alloc(qi.size());
    set->status.alloc(n);
    STP_HIP(hipMemcpyAsync(set->fobs.get(), fo.data(), fo.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(set->qinv.get(), qi.data(), qi.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    std::vector<double> zeros(fo.size(), 0.0);
    STP_HIP(hipMemcpyAsync(set->qdiag.get(), zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice,
                           stream_));
    sync();
    set->host_status.assign(n, 0);
    return set.release();
}

QgSet* HipBackend::qpgraph_set_shared_basis(const std::vector<const QgTopology*>& topos, const double* f_obs,
                                            const double* qinv, const double* qdiag) {
    activate();
    const int m = topos.empty() ? 0 : topos[0]->npair;
    for (const QgTopology* t : topos)
        STP_REQUIRE(t->npair == m, "qpgraph: graphs sharing one f3 basis need the same number of basis pairs");
    auto set = std::make_unique<QgSet>();
    qpgraph_upload_topologies(*set, topos);
    const int n = set->n;
    set->slot.assign(n, 0);
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    set->fobs.alloc(m);
    set->qdiag.alloc(m);
    set->qinv.alloc(mm);
    set->status.alloc(n);
    STP_HIP(hipMemcpyAsync(set->fobs.get(), f_obs, m * sizeof(double), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(set->qinv.get(), qinv, mm * sizeof(double), hipMemcpyHostToDevice, stream_));
    std::vector<double> zeros(m, 0.0);
    STP_HIP(hipMemcpyAsync(set->qdiag.get(), qdiag ? qdiag : zeros.data(), m * sizeof(double),
                           hipMemcpyHostToDevice, stream_));
    sync();
    set->host_status.assign(n, 0);
    return set.release();
}

void HipBackend::qpgraph_set_eval(QgSet& s, int n_inst, const std::int32_t* inst_graph, const double* theta,
                                  const double* bl_in, int mode, double fudge, int constrained, double* score,
                                  double* grad_theta, double* grad_bl, double* bl_out, double* fit_out) {
    activate();
    STP_REQUIRE(mode >= 0 && mode <= 2, "qpgraph: mode must be 0, 1 or 2");
    if (n_inst <= 0) return;
    const int NP = s.np_max, D = std::max(s.dmax, 1), NE = s.ne_max;
    std::vector<std::int64_t> offs(2 * static_cast<std::size_t>(n_inst));
    int T = 32;
    for (int i = 0; i < n_inst; ++i) {
        const int g = inst_graph[i];
        STP_REQUIRE(g >= 0 && g < s.n, "qpgraph: instance graph index out of range");
        STP_REQUIRE(s.host_status[g] == 0, "qpgraph: graph " + std::to_string(g) + " has no valid covariance");
        offs[i] = static_cast<std::int64_t>(s.slot[g]) * NP;
        offs[n_inst + i] = static_cast<std::int64_t>(s.slot[g]) * NP * NP;
        T = std::max(T, ne_bucket(s.topo[g].ne));
    }
    if (mode == 2) STP_REQUIRE(bl_in != nullptr, "qpgraph: the joint mode needs edge lengths");
    // staging: [inst_graph | offsets | theta | bl_in] up, [score | grads | bl] down
    s.i32.ensure(static_cast<std::size_t>(n_inst));
    s.i64.ensure(2 * static_cast<std::size_t>(n_inst));
    s.in.ensure(static_cast<std::size_t>(n_inst) * (D + NE));
    s.out.ensure(static_cast<std::size_t>(n_inst) * (1 + D + 2 * NE + NP));
    STP_HIP(hipMemcpyAsync(s.i32.get(), inst_graph, n_inst * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(s.i64.get(), offs.data(), offs.size() * sizeof(std::int64_t), hipMemcpyHostToDevice, stream_));
    double* d_theta = s.in.get();
    double* d_blin = s.in.get() + static_cast<std::size_t>(n_inst) * D;
    if (theta)
        STP_HIP(hipMemcpyAsync(d_theta, theta, static_cast<std::size_t>(n_inst) * D * sizeof(double),
                               hipMemcpyHostToDevice, stream_));
    if (mode == 2)
        STP_HIP(hipMemcpyAsync(d_blin, bl_in, static_cast<std::size_t>(n_inst) * NE * sizeof(double),
                               hipMemcpyHostToDevice, stream_));
    double* d_score = s.out.get();
    double* d_gth = d_score + n_inst;
    double* d_gbl = d_gth + static_cast<std::size_t>(n_inst) * D;
    double* d_bl = d_gbl + static_cast<std::size_t>(n_inst) * NE;
    double* d_fit = d_bl + static_cast<std::size_t>(n_inst) * NE;
    QgEvalArgs a{};
    a.topo = s.d_topo.get();
    a.iarena = s.d_iarena.get();
    a.darena = s.d_darena.get();
    a.n_inst = n_inst;
    a.inst_topo = s.i32.get();
    a.inst_fobs = s.i64.get();
    a.inst_qinv = s.i64.get() + n_inst;
    a.fobs = s.fobs.get();
    a.qinv = s.qinv.get();
    a.theta = d_theta;
    a.bl_in = d_blin;
    a.dmax = D;
    a.emax = NE;
    a.mode = mode;
    a.fudge = fudge;
    a.constrained = constrained;
    a.score = d_score;
    a.grad_theta = d_gth;
    a.grad_bl = d_gbl;
    a.bl_out = d_bl;
    a.fit_out = fit_out ? d_fit : nullptr;
    a.counter = counter_.get() + 49;
    a.np_max = NP;
    a.ne_max = NE;
    a.npop_max = s.npop_max;
    a.npath_max = std::max(s.npath_max, 1);
    STP_REQUIRE(qg_eval_lds_bytes(a) <= 150 * 1024, "qpgraph: graphs too large for the device score");
    launch_zero_u32(a.counter, 1, stream_);
    mark_begin(kQpgraph);
    launch_qg_eval(T, a, std::min(n_inst, qg_blocks_per_cu(T, false, qg_eval_lds_bytes(a)) * n_cu_), stream_);
    mark_end(kQpgraph);
    STP_HIP(hipMemcpyAsync(score, d_score, n_inst * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (grad_theta && mode != 0)
        STP_HIP(hipMemcpyAsync(grad_theta, d_gth, static_cast<std::size_t>(n_inst) * D * sizeof(double),
                               hipMemcpyDeviceToHost, stream_));
    if (grad_bl && mode == 2)
        STP_HIP(hipMemcpyAsync(grad_bl, d_gbl, static_cast<std::size_t>(n_inst) * NE * sizeof(double),
                               hipMemcpyDeviceToHost, stream_));
    if (bl_out)
        STP_HIP(hipMemcpyAsync(bl_out, d_bl, static_cast<std::size_t>(n_inst) * NE * sizeof(double),
                               hipMemcpyDeviceToHost, stream_));
    if (fit_out)
        STP_HIP(hipMemcpyAsync(fit_out, d_fit, static_cast<std::size_t>(n_inst) * NP * sizeof(double),
                               hipMemcpyDeviceToHost, stream_));
    sync();
}

namespace {

// The reference's multistart point of restart inst, coordinate dim: a splitmix64
// finaliser of a per-(inst, dim) counter, top 52 bits as a fraction
// (src/core/qpadm/qpgraph_opt_constants.hpp:18-26, cpu_backend.cpp:2966-2973).
double multistart_theta(unsigned inst, int dim) {
    std::uint64_t z = static_cast<std::uint64_t>(inst) * 0x100000001B3ull +
                      static_cast<std::uint64_t>(dim) * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double x = static_cast<double>(z & 0xFFFFFFFFFFFFFull) / 4503599627370496.0;
    return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
}

}  // namespace

void HipBackend::qpgraph_set_fit(QgSet& s, int n_inst, const std::int32_t* inst_graph, const double* theta0,
                                 const QgFitOptions& o, double* theta_out, double* score_out,
                                 std::int32_t* iters_out) {
    activate();
    if (n_inst <= 0) return;
    STP_REQUIRE(o.maxit >= 0, "qpgraph: maxit must be >= 0");
    const int NP = s.np_max, D = std::max(s.dmax, 1), NE = s.ne_max;
    std::vector<std::int64_t> offs(2 * static_cast<std::size_t>(n_inst));
    int T = 32;
    for (int i = 0; i < n_inst; ++i) {
        const int g = inst_graph[i];
        STP_REQUIRE(g >= 0 && g < s.n, "qpgraph: instance graph index out of range");
        STP_REQUIRE(s.host_status[g] == 0, "qpgraph: graph " + std::to_string(g) + " has no valid covariance");
        offs[i] = static_cast<std::int64_t>(s.slot[g]) * NP;
        offs[n_inst + i] = static_cast<std::int64_t>(s.slot[g]) * NP * NP;
        T = std::max(T, ne_bucket(s.topo[g].ne));
    }
    s.i32.ensure(2 * static_cast<std::size_t>(n_inst));
    s.i64.ensure(2 * static_cast<std::size_t>(n_inst));
    s.in.ensure(static_cast<std::size_t>(n_inst) * D);
    s.out.ensure(static_cast<std::size_t>(n_inst) * (D + 1));
    STP_HIP(hipMemcpyAsync(s.i32.get(), inst_graph, n_inst * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(s.i64.get(), offs.data(), offs.size() * sizeof(std::int64_t), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(s.in.get(), theta0, static_cast<std::size_t>(n_inst) * D * sizeof(double),
                           hipMemcpyHostToDevice, stream_));
    QgEvalArgs a{};
    a.topo = s.d_topo.get();
    a.iarena = s.d_iarena.get();
    a.darena = s.d_darena.get();
    a.n_inst = n_inst;
    a.inst_topo = s.i32.get();
    a.inst_fobs = s.i64.get();
    a.inst_qinv = s.i64.get() + n_inst;
    a.fobs = s.fobs.get();
    a.qinv = s.qinv.get();
    a.dmax = D;
    a.emax = NE;
    a.mode = 1;
    a.fudge = o.fudge;
    a.constrained = o.constrained;
    a.counter = counter_.get() + 50;
    a.np_max = NP;
    a.ne_max = NE;
    a.npop_max = s.npop_max;
    a.npath_max = std::max(s.npath_max, 1);
    STP_REQUIRE(qg_eval_lds_bytes(a) <= 150 * 1024, "qpgraph: graphs too large for the device score");
    QgFitArgs fa{};
    fa.theta0 = s.in.get();
    fa.theta_out = s.out.get();
    fa.score_out = s.out.get() + static_cast<std::size_t>(n_inst) * D;
    fa.iters_out = s.i32.get() + n_inst;
    fa.maxit = o.maxit;
    fa.gtol = o.gtol;
    fa.xtol = o.xtol;
    fa.ftol = o.ftol;
    launch_zero_u32(a.counter, 1, stream_);
    mark_begin(kQpgraph);
    launch_qg_fit(T, a, fa, std::min(n_inst, qg_blocks_per_cu(T, true, qg_eval_lds_bytes(a)) * n_cu_), stream_);
    mark_end(kQpgraph);
    if (theta_out)
        STP_HIP(hipMemcpyAsync(theta_out, fa.theta_out, static_cast<std::size_t>(n_inst) * D * sizeof(double),
                               hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipMemcpyAsync(score_out, fa.score_out, n_inst * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (iters_out)
        STP_HIP(hipMemcpyAsync(iters_out, fa.iters_out, n_inst * sizeof(std::int32_t), hipMemcpyDeviceToHost,
                               stream_));
    sync();
}

void HipBackend::qpgraph_set_fit_fleet(QgSet& s, int numstart, const QgFitOptions& o, QgFleetOut& out) {
    activate();
    STP_REQUIRE(numstart >= 1, "qpgraph: numstart must be >= 1");
    const int n = s.n, D = std::max(s.dmax, 1), NE = s.ne_max, NP = s.np_max;
    out.score.assign(n, std::numeric_limits<double>::infinity());
    out.restart_spread.assign(n, 0.0);
    out.status.assign(n, 3);
    out.iterations.assign(n, 0);
    out.theta.assign(static_cast<std::size_t>(n) * D, 0.0);
    out.theta_lo.assign(static_cast<std::size_t>(n) * D, 0.0);
    out.theta_hi.assign(static_cast<std::size_t>(n) * D, 0.0);
    out.edge_length.assign(static_cast<std::size_t>(n) * NE, 0.0);
    out.f3_fit.assign(static_cast<std::size_t>(n) * NP, 0.0);
    // Grid seeds: graphs with one or two admixture weights are also scored on a coarse
    // grid over [0, 1] that includes both bounds (kGrid1 points, or kGrid2 x kGrid2) with
    // their gradients.  Seeds are the grid's local minima (no higher than their grid
    // neighbours) and the lower end of each grid interval over which a partial
    // derivative changes sign from - to + (a minimum between two grid points); the
    // kGridSeeds lowest become more L-BFGS starts after the reference's.  A narrow basin
    // that none of the reference's start points falls in is then still found, and so is a
    // minimum at a bound behind a barrier (the score can drop at exactly theta = 0 or 1,
    // where a path's weight vanishes).
    constexpr int kGrid1 = 33, kGrid2 = 9, kGridSeeds = 4;
    std::vector<std::int32_t> gi, g0(n, -1);
    std::vector<double> gx;
    for (int g = 0; g < n; ++g) {
        const int d = s.topo[g].nadmix;
        if (s.host_status[g] != 0 || d < 1 || d > 2) continue;
        const int k = d == 1 ? kGrid1 : kGrid2;
        g0[g] = static_cast<std::int32_t>(gi.size());
        for (int a = 0; a < k; ++a)
            for (int b = 0; b < (d == 2 ? k : 1); ++b) {
                gi.push_back(g);
                for (int j = 0; j < D; ++j)
                    gx.push_back(j == 0 ? a / (k - 1.0) : (j == 1 && d == 2 ? b / (k - 1.0) : 0.0));
            }
    }
    std::vector<std::vector<std::pair<double, int>>> grid_best(n);
    if (!gi.empty()) {
        const int ng = static_cast<int>(gi.size());
        std::vector<double> gsc(ng), ggr(static_cast<std::size_t>(ng) * D);
        qpgraph_set_eval(s, ng, gi.data(), gx.data(), nullptr, 1, o.fudge, o.constrained, gsc.data(), ggr.data(),
                         nullptr, nullptr, nullptr);
        for (int g = 0; g < n; ++g) {
            if (g0[g] < 0) continue;
            const int d = s.topo[g].nadmix, k = d == 1 ? kGrid1 : kGrid2, rows = d == 2 ? k : 1;
            auto id = [&](int a, int b) { return g0[g] + a * rows + b; };
            auto at = [&](int a, int b) { return gsc[id(a, b)]; };
            auto grad = [&](int a, int b, int j) { return ggr[static_cast<std::size_t>(id(a, b)) * D + j]; };
            std::vector<char> seed(static_cast<std::size_t>(k) * rows, 0);
            for (int a = 0; a < k; ++a)
                for (int b = 0; b < rows; ++b) {
                    const double v = at(a, b);
                    if (!std::isfinite(v)) continue;
                    bool low = true;
                    if (a > 0 && std::isfinite(at(a - 1, b)) && at(a - 1, b) < v) low = false;
                    if (a + 1 < k && std::isfinite(at(a + 1, b)) && at(a + 1, b) < v) low = false;
                    if (b > 0 && std::isfinite(at(a, b - 1)) && at(a, b - 1) < v) low = false;
                    if (b + 1 < rows && std::isfinite(at(a, b + 1)) && at(a, b + 1) < v) low = false;
                    if (low) seed[a * rows + b] = 1;
                    // a sign change of d/dtheta_j between this point and the next along j
                    for (int j = 0; j < d; ++j) {
                        const int a2 = a + (j == 0), b2 = b + (j == 1);
                        if (a2 >= k || b2 >= rows || !std::isfinite(at(a2, b2))) continue;
                        if (grad(a, b, j) < 0.0 && grad(a2, b2, j) > 0.0)
                            seed[at(a2, b2) < v ? a2 * rows + b2 : a * rows + b] = 1;
                    }
                }
            auto& v = grid_best[g];
            for (int x = 0; x < k * rows; ++x)
                if (seed[x]) v.emplace_back(gsc[g0[g] + x], g0[g] + x);
            std::stable_sort(v.begin(), v.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            if (v.size() > kGridSeeds) v.resize(kGridSeeds);
        }
    }
    // restarts of every fittable graph: the reference's start points, then the grid seeds
    std::vector<std::int32_t> ig, first(n + 1, 0), n_ref(n, 0);
    std::vector<double> x0;
    for (int g = 0; g < n; ++g) {
        first[g] = static_cast<std::int32_t>(ig.size());
        if (s.host_status[g] != 0) continue;
        const int d = s.topo[g].nadmix;
        const int rs = d > 0 ? numstart : 1;
        n_ref[g] = rs;
        for (int r = 0; r < rs; ++r) {
            ig.push_back(g);
            for (int j = 0; j < D; ++j) x0.push_back(j < d ? multistart_theta(static_cast<unsigned>(r), j) : 0.0);
        }
        for (const auto& [sc0, i] : grid_best[g]) {
            ig.push_back(g);
            x0.insert(x0.end(), gx.begin() + static_cast<std::ptrdiff_t>(i) * D,
                      gx.begin() + static_cast<std::ptrdiff_t>(i + 1) * D);
        }
    }
    first[n] = static_cast<std::int32_t>(ig.size());
    const int m = static_cast<int>(ig.size());
    if (m == 0) return;
    std::vector<double> th(static_cast<std::size_t>(m) * D), sc(m);
    std::vector<std::int32_t> iters(m);
    qpgraph_set_fit(s, m, ig.data(), x0.data(), o, th.data(), sc.data(), iters.data());
    // best restart per graph (first of equal scores, over every start); the spread and
    // per-coordinate range are over the reference's restarts and the best one, so that,
    // as in the reference (where the best is always one of its restarts), the reported
    // weights lie inside their range and the best score inside the spread
    std::vector<std::int32_t> best_g;
    std::vector<double> best_th;
    for (int g = 0; g < n; ++g) {
        if (first[g] == first[g + 1]) continue;
        const int d = s.topo[g].nadmix;
        const int ref_end = first[g] + n_ref[g];
        int b = -1;
        double smin = std::numeric_limits<double>::infinity(), smax = -std::numeric_limits<double>::infinity();
        int it = 0;
        for (int i = first[g]; i < first[g + 1]; ++i) {
            if (i < ref_end) {
                smin = std::min(smin, sc[i]);
                smax = std::max(smax, sc[i]);
            }
            it = std::max(it, iters[i]);
            if (b < 0 || sc[i] < sc[b]) b = i;
        }
        out.iterations[g] = it;
        smin = std::min(smin, sc[b]);
        smax = std::max(smax, sc[b]);
        out.restart_spread[g] = (std::isfinite(smin) && std::isfinite(smax)) ? smax - smin : 0.0;
        for (int j = 0; j < d; ++j) {
            double lo = th[static_cast<std::size_t>(b) * D + j], hi = lo;
            for (int i = first[g]; i < ref_end; ++i) {
                lo = std::min(lo, th[static_cast<std::size_t>(i) * D + j]);
                hi = std::max(hi, th[static_cast<std::size_t>(i) * D + j]);
            }
            out.theta[static_cast<std::size_t>(g) * D + j] = th[static_cast<std::size_t>(b) * D + j];
            out.theta_lo[static_cast<std::size_t>(g) * D + j] = lo;
            out.theta_hi[static_cast<std::size_t>(g) * D + j] = hi;
        }
        out.score[g] = sc[b];
        best_g.push_back(g);
        best_th.insert(best_th.end(), th.begin() + static_cast<std::ptrdiff_t>(b) * D,
                       th.begin() + static_cast<std::ptrdiff_t>(b + 1) * D);
    }
    // edge lengths and fitted f3 at each graph's best weights
    const int nb = static_cast<int>(best_g.size());
    std::vector<double> fsc(nb), bl(static_cast<std::size_t>(nb) * NE), fit(static_cast<std::size_t>(nb) * NP);
    qpgraph_set_eval(s, nb, best_g.data(), best_th.data(), nullptr, 0, o.fudge, o.constrained, fsc.data(), nullptr,
                     nullptr, bl.data(), fit.data());
    for (int k = 0; k < nb; ++k) {
        const int g = best_g[k];
        if (!std::isfinite(fsc[k])) continue;
        out.status[g] = 0;
        std::copy(bl.begin() + static_cast<std::ptrdiff_t>(k) * NE, bl.begin() + static_cast<std::ptrdiff_t>(k + 1) * NE,
                  out.edge_length.begin() + static_cast<std::ptrdiff_t>(g) * NE);
        std::copy(fit.begin() + static_cast<std::ptrdiff_t>(k) * NP, fit.begin() + static_cast<std::ptrdiff_t>(k + 1) * NP,
                  out.f3_fit.begin() + static_cast<std::ptrdiff_t>(g) * NP);
    }
}

void HipBackend::qpgraph_set_basis(const QgSet& s, int g, double* f_obs, double* qdiag, double* qinv) {
    activate();
    STP_REQUIRE(g >= 0 && g < s.n, "qpgraph: graph index out of range");
    const int np = s.topo[g].npair, NP = s.np_max;
    const std::size_t k = static_cast<std::size_t>(s.slot[g]);
    if (f_obs)
        STP_HIP(hipMemcpyAsync(f_obs, s.fobs.get() + k * NP, np * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (qdiag)
        STP_HIP(hipMemcpyAsync(qdiag, s.qdiag.get() + k * NP, np * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (qinv)
        STP_HIP(hipMemcpyAsync(qinv, s.qinv.get() + k * NP * NP,
                               static_cast<std::size_t>(np) * np * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
}

}  // namespace stp
