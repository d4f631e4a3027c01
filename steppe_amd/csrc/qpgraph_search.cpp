// steppe_amd/csrc/qpgraph_search.cpp — the qpGraph topology search
// (HipBackend::qpgraph_search): every tree and one-admixture graph on a population set,
// fitted in one device fleet against one shared f3 basis, the global best, and the
// hill-climb that checks a local search reaches it.
//
// Follows the reference's run_search_impl (src/core/qpadm/qpgraph_search.cpp:110-278):
// candidates come from the bounded enumeration (qpgraph_enum.cpp); each is parsed with
// pops[0] as its f3 base and its basis pairs are renumbered to the shared order
// f3(pops[0]; a, b), a <= b over pops[1..] in the given order (qpgraph_search.cpp:51-105),
// so one f_obs and one Qinv, computed once on the device from the resident f2, serve
// every candidate; all restarts of all candidates run in one device L-BFGS fleet.  The
// best is the first lowest finite score; the hill-climb starts from K evenly spaced
// candidates and moves to the best strictly better neighbour until none is better
// (at most 1000 moves), with scores cached by shape hash.  Here the K climbs advance in
// lockstep so that each round's unscored neighbours are fitted in one fleet.
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>
#include <unordered_map>
#include <unordered_set>

#include "engine.hpp"
#include "qpgraph.hpp"
#include "qpgraph_enum.hpp"

namespace stp {

namespace {

constexpr int kMaxClimbSteps = 1000;
constexpr double kRecoveryRtol = 1e-6;
constexpr double kRecoveryAtol = 1e-9;

// f2 indices of the shared basis pairs (pair_a[k], pair_b[k]) and of the base.
struct SharedPairs {
    int base = -1;
    std::vector<int> a, b;
};

SharedPairs shared_pairs(const std::vector<int>& pop_f2) {
    SharedPairs sp;
    sp.base = pop_f2[0];
    const int n = static_cast<int>(pop_f2.size()) - 1;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            sp.a.push_back(pop_f2[1 + i]);
            sp.b.push_back(pop_f2[1 + j]);
        }
    return sp;
}

// Renumbers a topology's basis pairs to the shared order; false when a pair population
// is not one of its centred (non-base) leaves.
bool use_shared_pairs(QgTopology& t, const SharedPairs& sp) {
    std::unordered_map<int, int> col_of;
    for (int c = 0; c < t.npop - 1; ++c) col_of[t.leaf_to_f2[t.col_to_leaf(c)]] = c;
    const std::size_t m = sp.a.size();
    std::vector<std::int32_t> c1(m), c2(m);
    for (std::size_t k = 0; k < m; ++k) {
        const auto ia = col_of.find(sp.a[k]), ib = col_of.find(sp.b[k]);
        if (ia == col_of.end() || ib == col_of.end()) return false;
        c1[k] = std::min(ia->second, ib->second);
        c2[k] = std::max(ia->second, ib->second);
    }
    t.cmb1 = std::move(c1);
    t.cmb2 = std::move(c2);
    t.npair = static_cast<int>(m);
    return true;
}

// Parse + renumber; nullptr when the candidate cannot be fitted (the reference skips it).
std::unique_ptr<QgTopology> prepare(const qgenum::Candidate& c, const std::vector<std::string>& names,
                                    const std::string& base, const SharedPairs& sp) {
    try {
        auto t = std::make_unique<QgTopology>(parse_qpgraph(c.edges, names, base));
        if (!use_shared_pairs(*t, sp)) return nullptr;
        return t;
    } catch (const ArgError&) {
        return nullptr;
    }
}

}  // namespace

QgSearchResult HipBackend::qpgraph_search(const F2Blocks& f2, const std::vector<std::string>& names,
                                          const std::vector<std::string>& pops, const QgSearchOptions& o) {
    activate();
    QgSearchResult r;
    STP_REQUIRE(pops.size() >= 3, "qpgraph-search: need >= 3 population labels");
    STP_REQUIRE(f2.tier == kTierResident,
                "qpgraph-search: the f2 tensor is held in host memory or on disk; fits need it resident in HBM");
    std::unordered_map<std::string, int> f2_index;
    for (int i = 0; i < static_cast<int>(names.size()); ++i) f2_index.emplace(names[i], i);
    std::vector<int> pop_f2;
    std::unordered_set<std::string> distinct;
    for (const auto& p : pops) {
        const auto it = f2_index.find(p);
        STP_REQUIRE(it != f2_index.end(), "qpgraph-search: invalid pop-set ('" + p + "' is not an f2 population)");
        STP_REQUIRE(distinct.insert(p).second, "qpgraph-search: population '" + p + "' is listed twice");
        pop_f2.push_back(it->second);
    }
    const int max_nadmix = std::min(o.max_nadmix, 1);
    r.cands = qgenum::enumerate_space(pops, max_nadmix);
    for (const auto& c : r.cands) (c.nadmix == 0 ? r.n_trees : r.n_admix1) += 1;
    r.n_candidates = static_cast<int>(r.cands.size());
    const SharedPairs sp = shared_pairs(pop_f2);

    // candidates that parse, in enumeration order
    std::vector<std::unique_ptr<QgTopology>> topo;
    std::vector<const QgTopology*> tp;
    for (std::size_t i = 0; i < r.cands.size(); ++i) {
        auto t = prepare(r.cands[i], names, pops[0], sp);
        if (!t) continue;
        r.fitted.push_back(static_cast<int>(i));
        tp.push_back(t.get());
        topo.push_back(std::move(t));
    }
    STP_REQUIRE(!tp.empty(), "qpgraph-search: no candidate graph can be fitted on this pop-set");

    // the shared basis: f3, unridged Q diagonal and Qinv from the resident f2, once
    const int m = static_cast<int>(sp.a.size());
    std::vector<double> fobs(m), qdiag(m), qinv(static_cast<std::size_t>(m) * m);
    {
        const F2Blocks* f2p = &f2;
        std::unique_ptr<QgSet> one(qpgraph_set_create({tp[0]}, &f2p, o.diag_f3));
        if (one->host_status[0] != 0) {
            r.status = 3;   // non_spd_covariance
            return r;
        }
        qpgraph_set_basis(*one, 0, fobs.data(), qdiag.data(), qinv.data());
    }
    std::unique_ptr<QgSet> set(qpgraph_set_shared_basis(tp, fobs.data(), qinv.data(), qdiag.data()));

    const auto t0 = std::chrono::steady_clock::now();
    QgFleetOut fl;
    qpgraph_set_fit_fleet(*set, o.numstart, o.fit, fl);
    const auto t1 = std::chrono::steady_clock::now();
    r.fit_all_wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    r.topologies_per_s = r.fit_all_wall_ms > 0.0 ? tp.size() / (r.fit_all_wall_ms / 1e3) : 0.0;
    r.score = fl.score;
    r.restart_spread = fl.restart_spread;

    // the best and second-best finite scores (first of equal scores wins)
    int best = -1;
    double best_s = std::numeric_limits<double>::infinity(), second_s = best_s;
    for (int g = 0; g < static_cast<int>(tp.size()); ++g) {
        const double s = fl.score[g];
        if (!std::isfinite(s)) continue;
        if (s < best_s) {
            second_s = best_s;
            best_s = s;
            best = g;
        } else if (s < second_s) {
            second_s = s;
        }
    }
    if (best < 0) {
        r.status = 3;
        return r;
    }
    r.best = best;
    r.best_score = best_s;
    r.second_best_score = second_s;
    const int D = std::max(set->dmax, 1), NE = set->ne_max, NP = set->np_max;
    const QgTopology& bt = *tp[best];
    r.best_topo = bt;
    r.best_fit_status = fl.status[best];
    r.theta.assign(fl.theta.begin() + static_cast<std::ptrdiff_t>(best) * D,
                   fl.theta.begin() + static_cast<std::ptrdiff_t>(best) * D + bt.nadmix);
    r.theta_lo.assign(fl.theta_lo.begin() + static_cast<std::ptrdiff_t>(best) * D,
                      fl.theta_lo.begin() + static_cast<std::ptrdiff_t>(best) * D + bt.nadmix);
    r.theta_hi.assign(fl.theta_hi.begin() + static_cast<std::ptrdiff_t>(best) * D,
                      fl.theta_hi.begin() + static_cast<std::ptrdiff_t>(best) * D + bt.nadmix);
    r.edge_length.assign(fl.edge_length.begin() + static_cast<std::ptrdiff_t>(best) * NE,
                         fl.edge_length.begin() + static_cast<std::ptrdiff_t>(best) * NE + bt.ne);
    r.f3_fit.assign(fl.f3_fit.begin() + static_cast<std::ptrdiff_t>(best) * NP,
                    fl.f3_fit.begin() + static_cast<std::ptrdiff_t>(best) * NP + bt.npair);
    r.f_obs = fobs;
    r.qdiag = qdiag;

    if (o.run_heuristic) {
        std::unordered_map<std::uint64_t, double> cache;
        for (std::size_t k = 0; k < r.fitted.size(); ++k) cache.emplace(r.cands[r.fitted[k]].hash, fl.score[k]);
        // Fits the graphs of a round whose shapes have no score yet, in one fleet.
        auto score_missing = [&](const std::vector<const qgenum::Candidate*>& want) {
            std::vector<std::unique_ptr<QgTopology>> ts;
            std::vector<const QgTopology*> ptr;
            std::vector<std::uint64_t> hs;
            std::unordered_set<std::uint64_t> queued;
            for (const qgenum::Candidate* c : want) {
                if (cache.count(c->hash) || !queued.insert(c->hash).second) continue;
                auto t = prepare(*c, names, pops[0], sp);
                if (!t) {
                    cache[c->hash] = std::numeric_limits<double>::infinity();
                    continue;
                }
                ptr.push_back(t.get());
                hs.push_back(c->hash);
                ts.push_back(std::move(t));
            }
            if (ptr.empty()) return;
            std::unique_ptr<QgSet> s(qpgraph_set_shared_basis(ptr, fobs.data(), qinv.data(), qdiag.data()));
            QgFleetOut out;
            qpgraph_set_fit_fleet(*s, o.numstart, o.fit, out);
            for (std::size_t k = 0; k < hs.size(); ++k) cache[hs[k]] = out.score[k];
        };
        const int K = std::max(1, o.heuristic_seeds);
        const std::size_t N = r.cands.size();
        std::vector<qgenum::Candidate> cur(K);
        std::vector<double> cur_s(K);
        std::vector<char> active(K, 1);
        std::vector<const qgenum::Candidate*> want;
        for (int k = 0; k < K; ++k) {
            cur[k] = r.cands[N > 1 ? (static_cast<std::size_t>(k) * N) / static_cast<std::size_t>(K) : 0];
            want.push_back(&cur[k]);
        }
        score_missing(want);
        for (int k = 0; k < K; ++k) cur_s[k] = cache[cur[k].hash];
        for (int step = 0; step < kMaxClimbSteps; ++step) {
            std::vector<std::vector<qgenum::Candidate>> nb(K);
            want.clear();
            bool any = false;
            for (int k = 0; k < K; ++k) {
                if (!active[k]) continue;
                any = true;
                nb[k] = qgenum::neighbors(cur[k], pops, max_nadmix);
                for (const auto& c : nb[k]) want.push_back(&c);
            }
            if (!any) break;
            score_missing(want);
            for (int k = 0; k < K; ++k) {
                if (!active[k]) continue;
                double best_nb = cur_s[k];
                int j_best = -1;
                for (int j = 0; j < static_cast<int>(nb[k].size()); ++j) {
                    const double s = cache[nb[k][j].hash];
                    if (std::isfinite(s) && s < best_nb) {
                        best_nb = s;
                        j_best = j;
                    }
                }
                if (j_best < 0) {
                    active[k] = 0;
                    continue;
                }
                cur[k] = nb[k][j_best];
                cur_s[k] = best_nb;
            }
        }
        std::uint64_t agg_hash = 0;
        double agg_s = std::numeric_limits<double>::infinity();
        for (int k = 0; k < K; ++k) {
            r.seed_hashes.push_back(cur[k].hash);
            if (std::isfinite(cur_s[k]) && cur_s[k] < agg_s) {
                agg_s = cur_s[k];
                agg_hash = cur[k].hash;
            }
        }
        r.heuristic_recovered = agg_hash == r.cands[r.fitted[best]].hash &&
                                std::fabs(agg_s - best_s) <= kRecoveryAtol + kRecoveryRtol * std::fabs(best_s);
    }
    r.status = 0;
    return r;
}

}  // namespace stp
