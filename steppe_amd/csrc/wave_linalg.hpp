// steppe_amd/csrc/wave_linalg.hpp — wavefront-level building blocks shared by the fit
// kernels (qpadm.hip, qpgraph.hip): lane reads and uniform values, DPP reductions, the
// pivot search, double-double accumulation, the chi-square tail, an LU solve with one
// matrix row per lane in registers, and an LDS LU factorisation / inverse.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <type_traits>
#include <utility>

namespace stp {
namespace wl {

// ---------------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------------
__device__ inline int lane_id() { return static_cast<int>(threadIdx.x & 63); }

__device__ inline double rl(double x, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

__device__ inline int rli(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// f(std::integral_constant<int, I>) for I = 0 .. N-1, as straight-line code.
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (f(std::integral_constant<int, I>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

// A value every lane computed identically, moved to scalar registers so the compiler
// treats branches on it as uniform.
__device__ inline double uni(double x) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

__device__ inline int first_lane(bool pred) {
    const unsigned long long b = __ballot(pred);
    return b ? __builtin_ctzll(b) : -1;
}

__device__ inline int popc(bool pred) { return __builtin_popcountll(__ballot(pred)); }

// Full-wavefront reductions (all 64 lanes active) over DPP row operations and the
// row broadcasts; the result is read from lane 63.
#define STP_DPP_STEP(op, ident, ctrl, rmask) v = op(v, __builtin_amdgcn_update_dpp(ident, v, ctrl, rmask, 0xF, false))
__device__ inline int wave_max_i32(int v) {
    STP_DPP_STEP(max, INT_MIN, 0xB1, 0xF);   // quad_perm [1,0,3,2]
    STP_DPP_STEP(max, INT_MIN, 0x4E, 0xF);   // quad_perm [2,3,0,1]
    STP_DPP_STEP(max, INT_MIN, 0x141, 0xF);  // row_half_mirror
    STP_DPP_STEP(max, INT_MIN, 0x140, 0xF);  // row_mirror
    STP_DPP_STEP(max, INT_MIN, 0x142, 0xA);  // row_bcast:15
    STP_DPP_STEP(max, INT_MIN, 0x143, 0xC);  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ inline int wave_min_i32(int v) {
    STP_DPP_STEP(min, INT_MAX, 0xB1, 0xF);
    STP_DPP_STEP(min, INT_MAX, 0x4E, 0xF);
    STP_DPP_STEP(min, INT_MAX, 0x141, 0xF);
    STP_DPP_STEP(min, INT_MAX, 0x140, 0xF);
    STP_DPP_STEP(min, INT_MAX, 0x142, 0xA);
    STP_DPP_STEP(min, INT_MAX, 0x143, 0xC);
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ inline unsigned umax_(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ inline unsigned wave_max_u32(unsigned v0) {
    int v = static_cast<int>(v0);
#define STP_UMAX(a, b) static_cast<int>(umax_(static_cast<unsigned>(a), static_cast<unsigned>(b)))
    STP_DPP_STEP(STP_UMAX, 0, 0xB1, 0xF);
    STP_DPP_STEP(STP_UMAX, 0, 0x4E, 0xF);
    STP_DPP_STEP(STP_UMAX, 0, 0x141, 0xF);
    STP_DPP_STEP(STP_UMAX, 0, 0x140, 0xF);
    STP_DPP_STEP(STP_UMAX, 0, 0x142, 0xA);
    STP_DPP_STEP(STP_UMAX, 0, 0x143, 0xC);
#undef STP_UMAX
    return static_cast<unsigned>(__builtin_amdgcn_readlane(v, 63));
}
#undef STP_DPP_STEP

// The lane holding the pivot: largest |value| among candidate lanes (val >= 0;
// non-candidates pass -1), ties broken by the smallest logical position, matching
// the reference's scan with a strict '>' (small_linalg.hpp:62-70).
__device__ inline int argmax_pivot(double val, int pos) {
    const long long bits = __double_as_longlong(val);
    const int hi = static_cast<int>(bits >> 32);
    const int mhi = wave_max_i32(hi);
    bool tie = hi == mhi;
    if (popc(tie) > 1) {
        const unsigned lo = tie ? static_cast<unsigned>(bits & 0xffffffffLL) : 0u;
        const unsigned mlo = wave_max_u32(lo);
        tie = tie && lo == mlo;
        if (popc(tie) > 1) {
            const int mp = wave_min_i32(tie ? pos : INT_MAX);
            tie = tie && pos == mp;
        }
    }
    return first_lane(tie);
}

// ---------------------------------------------------------------------------------
// double-double accumulation (stands in for the reference's long double sums)
// ---------------------------------------------------------------------------------
struct dd {
    double hi, lo;
};

__device__ inline dd dd_norm(double s, double e) {
    const double h = s + e;
    return {h, e - (h - s)};
}
__device__ inline dd dd_add(dd a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b - bb);
    return dd_norm(s, e + a.lo);
}
__device__ inline dd dd_add_dd(dd a, dd b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb);
    return dd_norm(s, e + a.lo + b.lo);
}
// a + x*y with the product kept exactly.
__device__ inline dd dd_fma(dd a, double x, double y) {
    const double p = x * y;
    const double pe = fma(x, y, -p);
    dd r = dd_add(a, p);
    return dd_norm(r.hi, r.lo + pe);
}
// a + x*y for a double-double y.
__device__ inline dd dd_fma_dd(dd a, double x, dd y) {
    const double p = x * y.hi;
    const double pe = fma(x, y.hi, -p);
    dd r = dd_add(a, p);
    return dd_norm(r.hi, r.lo + (pe + x * y.lo));
}
// (double)(a / d)
__device__ inline double dd_div_d(dd a, double d) {
    const double q = a.hi / d;
    const double p = q * d;
    const double pe = fma(q, d, -p);
    const double r = ((a.hi - p) - pe) + a.lo;
    return q + r / d;
}
// (double)(a / b)
__device__ inline double dd_div_dd(dd a, dd b) {
    const double q = a.hi / b.hi;
    const double p = q * b.hi;
    const double pe = fma(q, b.hi, -p);
    const double r = (((a.hi - p) - pe) + a.lo) - q * b.lo;
    return q + r / b.hi;
}

// ---------------------------------------------------------------------------------
// chi-square upper tail: regularised incomplete gamma Q(dof/2, x/2), series below
// a + 1 and Lentz's continued fraction above (the reference's pchisq_upper,
// src/core/internal/pchisq.hpp:16-70, follows the same textbook construction).
// ---------------------------------------------------------------------------------
__device__ inline double gamma_pref(double a, double x) { return exp(-x + a * log(x) - lgamma(a)); }

__device__ inline double chisq_upper(double x, int dof) {
    if (dof <= 0) return __builtin_nan("");
    if (x <= 0.0) return 1.0;
    const double a = 0.5 * dof, xx = 0.5 * x;
    if (xx < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int n = 0; n < 1000; ++n) {
            ap += 1.0;
            del *= xx / ap;
            sum += del;
            if (fabs(del) < fabs(sum) * 1e-15) break;
        }
        return 1.0 - sum * gamma_pref(a, xx);
    }
    const double tiny = 1e-300;
    double b = xx + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i <= 1000; ++i) {
        const double an = -static_cast<double>(i) * (static_cast<double>(i) - a);
        b += 2.0;
        d = an * d + b;
        if (fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < 1e-15) break;
    }
    return gamma_pref(a, xx) * h;
}

__device__ inline int dof_of(int nl, int nr, int r) { return (nl - r) * (nr - r); }

// ---------------------------------------------------------------------------------
// LU solve of an n x n system held one row per lane (lane a < n: row[0..n) of row a,
// right-hand side b).  Partial pivoting with the reference's pivot choice; rows keep
// their lanes and carry a logical position.  Returns false for an exactly zero pivot
// (the reference's singular test); otherwise x[a] is returned in lane a.  The pivot
// row reaches the other lanes by lane reads (to scalar registers) in uniform control
// flow, so the elimination needs no memory round trip per column.
// ---------------------------------------------------------------------------------
template <int T>
__device__ __forceinline__ bool lu_solve_rows(double (&row)[T], double b, int n, double& x_out) {
    const int lane = lane_id();
    const bool act = lane < n;
    int pos = act ? lane : INT_MAX;
    bool ok = true;
    // Straight-line code over compile-time column indices keeps row[] in registers.
    static_for<T>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k < n && ok) {
            const double cand = (act && pos >= k) ? fabs(row[k]) : -1.0;
            const int pl = argmax_pivot(cand, pos);
            const double pv = rl(row[k], pl);
            if (pv == 0.0) {
                ok = false;
            } else {
                const int lk = first_lane(pos == k);
                const int ppl = rli(pos, pl);
                if (lane == pl) pos = k;
                else if (lane == lk) pos = ppl;
                const double rp = 1.0 / pv;
                const bool upd = act && pos > k;
                const double f = row[k] * rp;
                if (upd) row[k] = f;
#pragma unroll
                for (int j = k + 1; j < T; ++j) {
                    if (j < n) {
                        const double pj = rl(row[j], pl);  // the pivot row's entry j
                        if (upd) row[j] -= f * pj;
                    }
                }
            }
        }
    });
    if (!ok) return false;
    // forward substitution (unit lower), the right-hand side moves with its row
#pragma unroll
    for (int j = 0; j < T; ++j) {
        if (j < n) {
            const double yj = rl(b, first_lane(pos == j));
            if (act && pos > j) b -= row[j] * yj;
        }
    }
    // back substitution, column by column
    double x = 0.0;
#pragma unroll
    for (int ii = T - 1; ii >= 0; --ii) {
        if (ii < n) {
            const int li = first_lane(pos == ii);
            const double xi = rl(b, li) / rl(row[ii], li);
            if (act && pos < ii) b -= row[ii] * xi;
            if (lane == ii) x = xi;
        }
    }
    x_out = x;
    return true;
}

// ---------------------------------------------------------------------------------
// LU factorisation of the m x m matrix in LDS (row i at lu + i*ld) by one wavefront,
// lane = row; pos[lane] and lane_of[pos] record the pivoting.  Returns false on an
// exactly zero pivot.
// ---------------------------------------------------------------------------------
__device__ inline bool lds_lu(double* lu, int ld, int m, int* pos_of_lane, int* lane_of_pos) {
    const int lane = lane_id();
    const bool act = lane < m;
    int pos = act ? lane : INT_MAX;
    double* my = lu + (act ? lane : 0) * ld;
    for (int k = 0; k < m; ++k) {
        const double cand = (act && pos >= k) ? fabs(my[k]) : -1.0;
        const int pl = argmax_pivot(cand, pos);
        const double pv = uni(lu[pl * ld + k]);
        if (pv == 0.0) return false;
        const int lk = first_lane(pos == k);
        const int ppl = rli(pos, pl);
        if (lane == pl) pos = k;
        else if (lane == lk) pos = ppl;
        const double rp = 1.0 / pv;
        const double* prow = lu + pl * ld;
        if (act && pos > k) {
            const double f = my[k] * rp;
            my[k] = f;
            for (int j = k + 1; j < m; ++j) my[j] -= f * prow[j];
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (act) {
        pos_of_lane[lane] = pos;
        lane_of_pos[pos] = lane;
    }
    return true;
}

// Column c of the inverse from the LDS LU factors: solves LU x = P e_c with the row
// right-hand sides carried by their lanes; x[i] lands in lane i.
__device__ inline double lds_inverse_column(const double* lu, int ld, int m, const int* pos_of_lane,
                                     const int* lane_of_pos, int c) {
    const int lane = lane_id();
    const bool act = lane < m;
    const int pos = act ? pos_of_lane[lane] : INT_MAX;
    const double* my = lu + (act ? lane : 0) * ld;
    double y = (lane == c) ? 1.0 : 0.0;
    for (int j = 0; j < m; ++j) {
        const double yj = rl(y, lane_of_pos[j]);
        if (act && pos > j) y -= my[j] * yj;
    }
    double x = 0.0;
    for (int i = m - 1; i >= 0; --i) {
        const int li = lane_of_pos[i];
        const double xi = rl(y, li) / lu[li * ld + i];
        if (act && pos < i) y -= my[i] * xi;
        if (lane == i) x = xi;
    }
    return x;
}


}  // namespace wl
}  // namespace stp
