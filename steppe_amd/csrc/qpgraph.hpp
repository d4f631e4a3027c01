// steppe_amd/csrc/qpgraph.hpp — qpGraph topology tables and the launch interface of the
// device score / adjoint and f3-basis kernels (qpgraph.hip, qpgraph_host.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace stp {

// A parsed admixture graph (the reference's QpGraphModel,
// src/core/qpadm/qpgraph_model.hpp:23-70), with the variable leaf-weight cells grouped:
// cell (grp_edge[g], grp_leaf[g]) of the leaf-weight matrix is the sum of the path
// weights grp_paths[grp_off[g] .. grp_off[g+1]) (in the reference's summation order);
// path p's weight is the product over its admixture edges pae_id[pae_off[p] ..
// pae_off[p+1]) of theta_j (id 2j+1) or 1 - theta_j (id 2j+2).
struct QgTopology {
    int npop = 0, ne = 0, nadmix = 0, npair = 0, npath = 0, base = 0;
    std::vector<std::string> leaves, admix_from, admix_to, edge_from, edge_to;
    std::vector<std::int32_t> leaf_to_f2;
    std::vector<double> pwts0;  // (e, leaf) at e + ne*leaf
    std::vector<std::int32_t> grp_edge, grp_leaf, grp_off, grp_paths;
    std::vector<std::int32_t> pae_off, pae_id;
    std::vector<std::int32_t> cmb1, cmb2;  // f3 basis pairs of centred columns (a <= b)
    int col_to_leaf(int c) const;
};

// Parses a parent->child edge list against population labels; throws ArgError with
// the reference's messages for invalid graphs (qpgraph_model.cpp:41-243).  base_leaf
// names the f3 base population (the reference's f3basepop); empty: the first leaf.
QgTopology parse_qpgraph(const std::vector<std::pair<std::string, std::string>>& edges,
                         const std::vector<std::string>& names, const std::string& base_leaf = "");

// Device header of one topology inside the int32 / double arenas.
struct QgTopoDev {
    std::int32_t npop, ne, nadmix, npair, npath, base, ngrp;
    std::int32_t o_grp_edge, o_grp_leaf, o_grp_off, o_grp_paths, o_pae_off, o_pae_id, o_cmb1, o_cmb2;
    std::int32_t o_pwts0;  // into the double arena
};

struct QgEvalArgs {
    const QgTopoDev* topo;
    const std::int32_t* iarena;
    const double* darena;
    int n_inst = 0;
    const std::int32_t* inst_topo;   // topology of instance i
    const std::int64_t* inst_fobs;   // offset of its f_obs (npair) in fobs
    const std::int64_t* inst_qinv;   // offset of its Qinv (npair x npair, row-major) in qinv
    const double* fobs;
    const double* qinv;
    const double* theta;  // [n_inst][dmax]
    const double* bl_in;  // [n_inst][emax] (joint mode)
    int dmax = 0, emax = 0;
    int mode = 0;         // 0 profiled score, 1 + d score / d theta, 2 joint score + gradients
    double fudge = 1e-4;
    int constrained = 1;
    double* score;        // [n_inst]
    double* grad_theta;   // [n_inst][dmax]
    double* grad_bl;      // [n_inst][emax] (joint mode)
    double* bl_out;       // [n_inst][emax]
    double* fit_out;      // [n_inst][np_max]: P bl, the fitted f3 basis
    unsigned* counter;
    // LDS layout (doubles): sized from the batch maxima
    int np_max = 0, ne_max = 0, npop_max = 0, npath_max = 0;
};

// Device L-BFGS fits of a batch of instances (graph + start point), one per workgroup.
struct QgFitArgs {
    const double* theta0;     // [n_inst][dmax] start points
    double* theta_out;        // [n_inst][dmax] final mixture weights
    double* score_out;        // [n_inst] final score (+inf: singular edge-length system)
    std::int32_t* iters_out;  // [n_inst] L-BFGS iterations taken; -1 when the start failed
    int maxit = 200;
    double gtol = 1e-11, xtol = 1e-12, ftol = 1e-14;
};

struct QgBasisArgs {
    const QgTopoDev* topo;
    const std::int32_t* iarena;
    int n_jobs = 0;
    const std::int32_t* job_topo;
    const std::int32_t* job_leaf_to_f2;  // [n_jobs][npop_max]
    const double* const* f2;             // per job: [P x P x nb] at i + P*j + P*P*b
    const double* const* vpair;          // per job (may hold nulls: all present)
    const std::int32_t* const* bsz;      // per job block sizes
    const std::int32_t* job_P;
    const std::int32_t* job_nb;
    double diag_f3 = 1e-5;
    double* fobs;     // [n_jobs][np_max]
    double* qdiag;    // [n_jobs][np_max] (unridged Q diagonal)
    double* qinv;     // [n_jobs][np_max * np_max], row-major with leading dimension npair
    std::int32_t* status;  // 0 ok, 3 non-SPD covariance / no blocks
    double* scratch;  // per workgroup
    std::int64_t scratch_stride = 0;
    int np_max = 0, nb_max = 0, npop_max = 0;
    unsigned* counter;
};

void launch_qg_eval(int T, const QgEvalArgs& a, int n_workgroups, hipStream_t st);
void launch_qg_fit(int T, const QgEvalArgs& a, const QgFitArgs& fa, int n_workgroups, hipStream_t st);
std::size_t qg_eval_lds_bytes(const QgEvalArgs& a);
// Workgroups of the persistent score (fit = false) or fit kernel of bucket T that fit on
// one CU at once with lds bytes of dynamic LDS (the launches fill every CU to this).
int qg_blocks_per_cu(int T, bool fit, std::size_t lds);
void launch_qg_basis(const QgBasisArgs& a, int n_workgroups, hipStream_t st);
std::size_t qg_basis_lds_bytes(const QgBasisArgs& a);

}  // namespace stp
