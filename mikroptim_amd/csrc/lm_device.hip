// Device side of the batched Levenberg-Marquardt control (levenberg_marquardt,
// src/optim/levenberg_marquardt.cpp:332-572): the per-run step solve, acceptance and convergence
// tests of one LM cycle run on the GPU, one wavefront (64 lanes) per run, so a cycle needs no
// host round trip of the normal equations or the trial points.
//
// Every floating-point operation is the one the host control performs, in the same order:
// products and sums without contraction (fp contract off below), correctly rounded division and
// square root (__ddiv_rn, __dsqrt_rn), std::max / std::min as (a < b) ? b : a / (b < a) ? b : a.
// Given the same normal equations and merits, device and host control take bit-identical
// trajectories. Sums keep the host order (that of the reference's cholesky / cholesky_solve,
// src/optim/levenberg_marquardt.cpp:70-112): the Cholesky factor (right-looking) and the forward
// substitution run lane-parallel over rows, each entry receiving its subtractions in the host's
// order; the back substitution forms each row's products on the lanes and subtracts them in
// ascending order; the step model (lin, quad) runs on lane 0; maxima (order-independent) are
// wave reductions.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "lm_device.hpp"

#pragma clang fp contract(off)

namespace mko_lmdev {
namespace {

constexpr int kLanes = 64;
constexpr int kMaxP = 128;            // two slots per lane
constexpr int kSlots = kMaxP / kLanes; // 2: the selects below assume it
constexpr double kModelDamping = 1e-12;
constexpr double kFdStepFactor = 6.0554544523933395e-06; // 2^(-52/3)

// Lane k's value of v, in every lane (k uniform).
__device__ __forceinline__ double lane_value(double v, int k) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double scale_of(double x, double typical) { return dmax(fabs(x), typical); }

__device__ __forceinline__ bool blocks(const State& S, int k, double x, double dir) {
    if (S.lower[k] == S.upper[k]) return true;
    return (x <= S.lower[k] && dir < 0.0) || (x >= S.upper[k] && dir > 0.0);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Packed lower triangle of an f x f matrix, column-major: element (row i, column j), i >= j.
__device__ __forceinline__ int tri(int j, int i, int f) { return j * f - (j * (j - 1)) / 2 + (i - j); }
// entries of the packed lower triangle of a P x P matrix
__host__ __device__ __forceinline__ int tri_size(int P) { return P * (P + 1) / 2; }

struct SolveScalars {
    double diag;
    int f;
};

// The solve's per-variable arrays, P entries each, in the dynamic LDS after the packed triangle
// and its row / column bytes (sized by P, not by the largest P: more runs' workgroups fit a CU).
struct SolveLds {
    double* rhs;
    double* delta;
    double* step;
    double* aq;
    double* ratio;
    double* col; // column j of the Cholesky factor while rows below it are updated
    int* idx;
    uint8_t* fixed;
    double& diag;
    int& f;
};

__host__ __device__ __forceinline__ size_t solve_arrays_offset(int P) {
    return (size_t(tri_size(P)) * (sizeof(double) + 2) + 7) / 8 * 8;
}

__device__ __forceinline__ SolveLds solve_lds(double* m, int P, SolveScalars& sc) {
    double* v = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(m) +
                                          solve_arrays_offset(P));
    int* idx = reinterpret_cast<int*>(v + 6 * size_t(P));
    return SolveLds{v,          v + P,     v + 2 * P, v + 3 * P,
                    v + 4 * P,  v + 5 * P, idx,       reinterpret_cast<uint8_t*>(idx + P),
                    sc.diag,    sc.f};
}

// solve_step: (A_FF + mu D_FF) delta_F = -g_F on the free variables with the active set; writes
// trial, step_norm, predicted and solved of the run (trial, step_norm and predicted only when
// the factorisation succeeds, like the host).
// a, g: the run's normal equations (row-major a[q * P + p]). Variable k belongs to lane k % 64,
// slot k / 64 (two slots: up to 128 variables); values a lane keeps per variable sit in
// two-element arrays indexed by the slot.
__device__ void solve_step(const State& S, int run, const double* a, const double* g, double mu,
                           SolveLds& L, double* __restrict__ m) {
    const int n = S.P;
    const int lane = threadIdx.x;
    const double* x = S.x + size_t(run) * n;
    const double* d = S.d + size_t(run) * n;
    for (int k = lane; k < n; k += kLanes) L.fixed[k] = blocks(S, k, x[k], -g[k]) ? 1 : 0;
    __syncthreads();
    for (int round = 0; round <= n; ++round) {
        if (lane == 0) {
            int f = 0;
            for (int k = 0; k < n; ++k)
                if (!L.fixed[k]) L.idx[f++] = k;
            L.f = f;
        }
        for (int k = lane; k < n; k += kLanes) L.delta[k] = 0.0;
        __syncthreads();
        const int f = L.f;
        if (f == 0) break;
        for (int j = 0; j < f; ++j) {
            for (int i = j + ((lane - j) % kLanes + kLanes) % kLanes; i < f; i += kLanes) {
                double v = a[size_t(L.idx[j]) * n + L.idx[i]];
                if (i == j) v += mu * d[L.idx[j]];
                m[tri(j, i, f)] = v;
            }
        }
        for (int i = lane; i < f; i += kLanes) L.rhs[i] = -g[L.idx[i]];
        __syncthreads();
        // row and column of every packed entry (the trailing columns' entries are spread over the
        // lanes below)
        {
            uint8_t* rowof = reinterpret_cast<uint8_t*>(m + tri_size(n));
            uint8_t* colof = rowof + tri_size(n);
            for (int c = 0; c < f; ++c)
                for (int i = lane; i < f; i += kLanes)
                    if (i >= c) {
                        const int e = tri(c, i, f);
                        rowof[e] = uint8_t(i);
                        colof[e] = uint8_t(c);
                    }
        }
        __syncthreads();
        // Cholesky, right-looking: after column j is final, every entry (i, k), j < k <= i,
        // receives l(i, j) l(k, j). Each entry receives the same subtractions in the same
        // (ascending j) order as in the host's left-looking loop, so the factor is bitwise the
        // host's. The trailing columns j+1 ... f-1 are one contiguous range of the packed layout,
        // whose entries are spread evenly over the lanes.
        const uint8_t* rowof = reinterpret_cast<const uint8_t*>(m + tri_size(n));
        const uint8_t* colof = rowof + tri_size(n);
        const int e_end = tri(f - 1, f - 1, f) + 1;
        for (int j = 0; j < f; ++j) {
            if (lane == (j & (kLanes - 1))) L.diag = m[tri(j, j, f)];
            __syncthreads();
            const double diag = L.diag;
            if (!(diag > 0.0) || !isfinite(diag)) {
                if (lane == 0) S.solved[run] = 0;
                return;
            }
            const double l = __dsqrt_rn(diag);
            for (int i = lane; i < f; i += kLanes) {
                if (i == j) {
                    m[tri(j, j, f)] = l;
                } else if (i > j) {
                    const double v = __ddiv_rn(m[tri(j, i, f)], l);
                    m[tri(j, i, f)] = v;
                    L.col[i] = v;
                }
            }
            __syncthreads();
            if (j + 1 < f) {
                for (int e = tri(j + 1, j + 1, f) + lane; e < e_end; e += kLanes)
                    m[e] -= L.col[rowof[e]] * L.col[colof[e]];
            }
            __syncthreads();
        }
        // forward substitution: row i subtracts l(i, k) y_k for k = 0, 1, ... as y_k is final;
        // y_k = (row k's sum) / l(k, k) is formed in every lane alike, so no LDS round trip
        {
            double acc[kSlots];
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const int i = lane + s * kLanes;
                acc[s] = i < f ? L.rhs[i] : 0.0;
            }
            for (int k = 0; k < f; ++k) {
                const double yk = __ddiv_rn(lane_value(k < kLanes ? acc[0] : acc[1], k % kLanes), m[tri(k, k, f)]);
#pragma unroll
                for (int s = 0; s < kSlots; ++s) {
                    const int i = lane + s * kLanes;
                    if (i == k) acc[s] = yk;
                    else if (i > k && i < f) acc[s] -= m[tri(k, i, f)] * yk;
                }
            }
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const int i = lane + s * kLanes;
                if (i < f) L.rhs[i] = acc[s];
            }
        }
        // back substitution, x_i = (b_i - sum_{k > i} l(k, i) x_k) / l(i, i) for i = f-1 ... 0:
        // the lane of row k keeps b_k, then x_k, in a register and forms the product l(k, i) x_k;
        // every lane subtracts the products from b_i in ascending k, as the host does, and
        // divides, so x_i is the same in all lanes and the lane of row i keeps it (no LDS round
        // trip between rows)
        {
            __syncthreads();
            double xk[kSlots];
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const int k = lane + s * kLanes;
                xk[s] = k < f ? L.rhs[k] : 0.0;
            }
            for (int i = f - 1; i >= 0; --i) {
                double p[kSlots];
#pragma unroll
                for (int s = 0; s < kSlots; ++s) {
                    const int k = lane + s * kLanes;
                    p[s] = (k > i && k < f) ? m[tri(i, k, f)] * xk[s] : 0.0;
                }
                double acc = lane_value(i < kLanes ? xk[0] : xk[1], i % kLanes);
                for (int k = i + 1; k < f; ++k) acc -= lane_value(k < kLanes ? p[0] : p[1], k % kLanes);
                const double xi = __ddiv_rn(acc, m[tri(i, i, f)]);
#pragma unroll
                for (int s = 0; s < kSlots; ++s)
                    if (lane + s * kLanes == i) xk[s] = xi;
            }
#pragma unroll
            for (int s = 0; s < kSlots; ++s) {
                const int k = lane + s * kLanes;
                if (k < f) L.rhs[k] = xk[s];
            }
        }
        __syncthreads();
        for (int i = lane; i < f; i += kLanes) L.delta[L.idx[i]] = L.rhs[i];
        __syncthreads();
        bool newly_fixed = false;
        for (int i = lane; i < f; i += kLanes) {
            const int k = L.idx[i];
            if (blocks(S, k, x[k], L.delta[k])) {
                L.fixed[k] = 1;
                newly_fixed = true;
            }
        }
        const bool changed = __any(newly_fixed);
        __syncthreads();
        if (!changed) break;
    }
    double* trial = S.trial + size_t(run) * n;
    for (int k = lane; k < n; k += kLanes) {
        const double t = dmin(dmax(x[k] + L.delta[k], S.lower[k]), S.upper[k]);
        trial[k] = t;
        L.step[k] = t - x[k];
        L.ratio[k] = __ddiv_rn(fabs(L.step[k]), scale_of(x[k], S.typical[k]));
    }
    __syncthreads();
    for (int q = lane; q < n; q += kLanes) {
        double aq = 0.0;
        for (int p = 0; p < n; ++p) aq += a[size_t(q) * n + p] * L.step[p];
        L.aq[q] = aq;
    }
    __syncthreads();
    // dmax(0, .) per variable first: a NaN ratio drops out as in the host's sequential std::max,
    // so the butterfly only sees ordered values and every lane ends with the same maximum
    double r = 0.0;
    for (int k = lane; k < n; k += kLanes) r = dmax(r, dmax(0.0, L.ratio[k]));
    const double nrm = wave_max(r);
    if (lane == 0) {
        double lin = 0.0, quad = 0.0;
        for (int q = 0; q < n; ++q) {
            lin += L.step[q] * g[q];
            quad += L.step[q] * L.aq[q];
        }
        S.step_norm[run] = nrm;
        S.predicted[run] = -(2.0 * lin + quad);
        S.solved[run] = 1;
    }
    __syncthreads();
}

// prepare_step: scaling diagonal D (Marquardt with floor) and the scaled gradient norm. The
// per-variable terms are formed lane-parallel; the two maxima are order-independent (finite
// values), so they equal the host's sequential maxima.
__device__ void prepare_step(const State& S, const Options& o, int run, const double* a,
                             const double* g) {
    const int n = S.P;
    const int lane = threadIdx.x;
    const double* x = S.x + size_t(run) * n;
    double smax_l = 0.0;
    for (int k = lane; k < n; k += kLanes) {
        const double sk = scale_of(x[k], S.typical[k]);
        smax_l = dmax(smax_l, dmax(0.0, a[size_t(k) * n + k] * sk * sk));
    }
    const double smax = wave_max(smax_l);
    const double merit = S.merit[run];
    double gmax_l = 0.0;
    for (int k = lane; k < n; k += kLanes) {
        const double sk = scale_of(x[k], S.typical[k]);
        const double akk = a[size_t(k) * n + k];
        const double floor = smax > 0.0 ? __ddiv_rn(o.diag_floor * smax, sk * sk)
                                        : __ddiv_rn(1.0, sk * sk);
        S.d[size_t(run) * n + k] = dmax(akk, floor);
        double ratio = 0.0;
        if (!blocks(S, k, x[k], -g[k])) {
            const double den = __dsqrt_rn(akk * merit);
            if (den > 0.0) ratio = __ddiv_rn(fabs(g[k]), den);
        }
        gmax_l = dmax(gmax_l, dmax(0.0, ratio));
    }
    const double gmax = wave_max(gmax_l);
    if (lane == 0) S.gradient_norm[run] = gmax;
}

__global__ __launch_bounds__(256) void init_kernel(State S, Options o) {
    const int run = blockIdx.x * blockDim.x + threadIdx.x;
    if (run >= S.n) return;
    const int n = S.P;
    for (int k = 0; k < n; ++k) {
        S.trial[size_t(run) * n + k] = S.x[size_t(run) * n + k];
        S.g[size_t(run) * n + k] = 0.0;
        S.d[size_t(run) * n + k] = 0.0;
    }
    S.mu[run] = o.initial_damping;
    S.nu[run] = 2.0;
    S.step_norm[run] = 0.0;
    S.predicted[run] = 0.0;
    S.gradient_norm[run] = 0.0;
    S.model_decrease[run] = __builtin_nan("");
    S.failed[run] = 0;
    S.hard[run] = 0;
    S.status[run] = MKO_LM_RUNNING;
    S.iterations[run] = 0;
    S.accepted[run] = 0;
    S.jacobians[run] = 0;
    S.history_len[run] = 0;
    S.need_jacobian[run] = 1;
    S.rejected[run] = 0;
    S.solved[run] = 0;
    S.evaluate[run] = 0;
}

__device__ __forceinline__ void push_history(const State& S, int run, double merit) {
    const int c = S.history_len[run];
    S.history[size_t(run) * S.H + (c % S.H)] = merit;
    S.history_len[run] = c + 1;
}

__global__ __launch_bounds__(256) void start_kernel(State S, const double* merit,
                                                    const int64_t* failed, const int64_t* hard) {
    const int run = blockIdx.x * blockDim.x + threadIdx.x;
    if (run >= S.n) return;
    const double m = merit[run];
    S.merit[run] = m;
    S.failed[run] = failed[run];
    S.hard[run] = hard[run];
    push_history(S, run, m);
    if (!isfinite(m)) S.status[run] = MKO_LM_NUMERICAL_ERROR;
    else if (m == 0.0) S.status[run] = MKO_LM_CONVERGED_MERIT;
}

// Ordered compaction in one workgroup of 1024 threads (chunks of 1024 runs, block scan).
__global__ __launch_bounds__(1024) void select_kernel(State S, int kind, int max_iterations,
                                                      const int32_t* in, const int32_t* in_count,
                                                      int32_t* list, int32_t* count, int32_t* pos) {
    __shared__ int32_t wave_total[16];
    __shared__ int32_t base_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int total = kind == kListEvaluate ? *in_count : S.n;
    if (tid == 0) base_s = 0;
    __syncthreads();
    for (int start = 0; start < total; start += 1024) {
        const int e = start + tid;
        bool pick = false;
        int run = -1;
        if (e < total) {
            run = kind == kListEvaluate ? in[e] : e;
            const int st = S.status[run];
            if (kind == kListJacobian) {
                pick = st == MKO_LM_RUNNING && S.need_jacobian[run] &&
                       S.iterations[run] < max_iterations;
            } else if (kind == kListCycle) {
                if (st == MKO_LM_RUNNING) {
                    if (S.iterations[run] >= max_iterations) S.status[run] = MKO_LM_MAX_ITERATIONS;
                    else pick = true;
                }
            } else {
                pick = S.evaluate[run] != 0;
            }
        }
        const uint64_t mask = __ballot(pick);
        const int before = __popcll(mask & ((uint64_t(1) << lane) - 1));
        if (lane == 0) wave_total[wave] = __popcll(mask);
        __syncthreads();
        int offset = base_s;
        for (int w = 0; w < wave; ++w) offset += wave_total[w];
        if (pick) list[offset + before] = run;
        if (pos && run >= 0) pos[run] = pick ? offset + before : -1;
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int w = 0; w < 16; ++w) sum += wave_total[w];
            base_s += sum;
        }
        __syncthreads();
    }
    if (tid == 0) *count = base_s;
}

// Stable split of the Jacobian list `in` (*in_count runs) in one workgroup of 1024 threads: the
// runs whose point has no failed ray into `clean` (in order, count *clean_count), the others
// into `other` (in order, count *other_count).
__global__ __launch_bounds__(1024) void split_kernel(State S, const int32_t* in,
                                                     const int32_t* in_count, int32_t* clean,
                                                     int32_t* clean_count, int32_t* other,
                                                     int32_t* other_count) {
    __shared__ int32_t wave_clean[16], wave_other[16];
    __shared__ int32_t base_c, base_o;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int total = *in_count;
    if (tid == 0) base_c = base_o = 0;
    __syncthreads();
    for (int start = 0; start < total; start += 1024) {
        const int e = start + tid;
        const int run = e < total ? in[e] : -1;
        const bool c = run >= 0 && S.failed[run] == 0;
        const bool o = run >= 0 && !c;
        const uint64_t mc = __ballot(c), mo = __ballot(o);
        const uint64_t below = (uint64_t(1) << lane) - 1;
        if (lane == 0) {
            wave_clean[wave] = __popcll(mc);
            wave_other[wave] = __popcll(mo);
        }
        __syncthreads();
        int oc = base_c, oo = base_o;
        for (int w = 0; w < wave; ++w) {
            oc += wave_clean[w];
            oo += wave_other[w];
        }
        if (c) clean[oc + __popcll(mc & below)] = run;
        if (o) other[oo + __popcll(mo & below)] = run;
        __syncthreads();
        if (tid == 0) {
            for (int w = 0; w < 16; ++w) {
                base_c += wave_clean[w];
                base_o += wave_other[w];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        *clean_count = base_c;
        *other_count = base_o;
    }
}

// Number of runs the next Jacobian list holds (kListJacobian's test) whose point has failed rays.
__global__ __launch_bounds__(1024) void count_failing_kernel(State S, int max_iterations,
                                                             int32_t* count) {
    __shared__ int32_t wave_total[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int n = 0;
    for (int run = tid; run < S.n; run += 1024)
        n += S.status[run] == MKO_LM_RUNNING && S.need_jacobian[run] &&
             S.iterations[run] < max_iterations && S.failed[run] > 0;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if (lane == 0) wave_total[wave] = n;
    __syncthreads();
    if (tid == 0) {
        int sum = 0;
        for (int w = 0; w < 16; ++w) sum += wave_total[w];
        *count = sum;
    }
}

__global__ __launch_bounds__(64) void gather_kernel(State S, const int32_t* list, int count,
                                                    const int32_t* dcount, int trial,
                                                    const int32_t* round_index, double* xs,
                                                    int32_t* tidx) {
    const int j = blockIdx.x;
    if (j >= count || (dcount && j >= *dcount)) return;
    const int run = list[j];
    const double* src = (trial ? S.trial : S.x) + size_t(run) * S.P;
    for (int k = threadIdx.x; k < S.P; k += 64) xs[size_t(j) * S.P + k] = src[k];
    if (threadIdx.x == 0 && tidx) tidx[j] = round_index ? round_index[run] : run;
}

// One wavefront per listed run (round_cycle, the runs that received new normal equations).
__global__ __launch_bounds__(kLanes) void prepare_kernel(State S, Options o, const int32_t* list,
                                                         int count, const int32_t* dcount,
                                                         const double* a_c,
                                                         const double* g_c) {
    __shared__ SolveScalars sc;
    __shared__ int finite_s;
    extern __shared__ double m_tri[]; // packed lower triangle, P (P + 1) / 2 doubles, then the
                                      // row and column (uint8) of each entry, then the arrays
    SolveLds L = solve_lds(m_tri, S.P, sc);
    const int j = blockIdx.x;
    if (j >= count || (dcount && j >= *dcount)) return;
    const int run = list[j];
    const int n = S.P, lane = threadIdx.x;
    bool finite = true;
    for (int e = lane; e < n * n; e += kLanes) {
        const double v = a_c[size_t(j) * n * n + e];
        S.a[size_t(run) * n * n + e] = v;
        finite = finite && isfinite(v);
    }
    for (int k = lane; k < n; k += kLanes) {
        const double v = g_c[size_t(j) * n + k];
        S.g[size_t(run) * n + k] = v;
        finite = finite && isfinite(v);
    }
    const bool all_finite = __all(finite);
    if (lane == 0) {
        S.need_jacobian[run] = 0;
        S.jacobians[run] += 1;
        finite_s = all_finite ? 1 : 0;
        if (!all_finite) S.status[run] = MKO_LM_NUMERICAL_ERROR;
    }
    __syncthreads();
    if (!finite_s) return;
    // the compact inputs are read-only in this kernel; the state copies serve later kernels
    const double* a = a_c + size_t(j) * n * n;
    const double* g = g_c + size_t(j) * n;
    prepare_step(S, o, run, a, g);
    __syncthreads();
    solve_step(S, run, a, g, kModelDamping, L, m_tri);
    if (lane == 0) {
        const bool solved = S.solved[run] != 0;
        const double merit = S.merit[run];
        const double predicted = S.predicted[run];
        S.model_decrease[run] = solved ? __ddiv_rn(predicted, merit) : __builtin_nan("");
        if (S.gradient_norm[run] <= o.gtol) S.status[run] = MKO_LM_CONVERGED_GRADIENT;
        else if (solved && predicted >= 0.0 && predicted <= o.ftol * merit)
            S.status[run] = MKO_LM_CONVERGED_MERIT;
        else if (solved && S.step_norm[run] <= o.xtol) S.status[run] = MKO_LM_CONVERGED_STEP;
    }
}

__global__ __launch_bounds__(kLanes) void step_kernel(State S, Options o, const int32_t* list,
                                                      int count, const int32_t* dcount) {
    __shared__ SolveScalars sc;
    extern __shared__ double m_tri[]; // packed lower triangle, P (P + 1) / 2 doubles, then the
                                      // row and column (uint8) of each entry, then the arrays
    SolveLds L = solve_lds(m_tri, S.P, sc);
    const int j = blockIdx.x;
    if (j >= count || (dcount && j >= *dcount)) return;
    const int run = list[j];
    solve_step(S, run, S.a + size_t(run) * S.P * S.P, S.g + size_t(run) * S.P, S.mu[run], L,
               m_tri);
    if (threadIdx.x == 0) {
        const bool solved = S.solved[run] != 0;
        S.evaluate[run] = 0;
        if (solved && S.rejected[run] && S.step_norm[run] <= o.xtol)
            S.status[run] = MKO_LM_CONVERGED_STEP;
        else if (solved)
            S.evaluate[run] = 1;
    }
}

__global__ __launch_bounds__(256) void update_kernel(State S, Options o, const int32_t* list,
                                                     int count, const int32_t* dcount,
                                                     const int32_t* pos, const double* f,
                                                     const int64_t* failed, const int64_t* hard) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count || (dcount && j >= *dcount)) return;
    const int run = list[j];
    if (S.status[run] == MKO_LM_CONVERGED_STEP) return;
    S.iterations[run] += 1;
    bool accept = false;
    const double merit = S.merit[run];
    if (S.solved[run]) {
        const int e = pos[run];
        const double f_new = f[e];
        const int64_t failed_new = failed[e], hard_new = hard[e];
        const double predicted = S.predicted[run];
        const double rho = __ddiv_rn(merit - f_new, predicted);
        accept = isfinite(f_new) && predicted > 0.0 && f_new < merit &&
                 !(o.reject_new_failures && hard_new > S.hard[run]);
        if (accept) {
            const int n = S.P;
            for (int k = 0; k < n; ++k) S.x[size_t(run) * n + k] = S.trial[size_t(run) * n + k];
            S.merit[run] = f_new;
            S.failed[run] = failed_new;
            S.hard[run] = hard_new;
            const double t = 2.0 * rho - 1.0;
            S.mu[run] = S.mu[run] * dmax(1.0 / 3.0, 1.0 - t * t * t);
            S.nu[run] = 2.0;
            S.need_jacobian[run] = 1;
            S.rejected[run] = 0;
            S.accepted[run] += 1;
            if (f_new == 0.0) S.status[run] = MKO_LM_CONVERGED_MERIT;
        }
    }
    if (!accept) {
        S.mu[run] = S.mu[run] * S.nu[run];
        S.nu[run] = S.nu[run] * 2.0;
        S.rejected[run] = 1;
        if (S.mu[run] > o.max_damping) S.status[run] = MKO_LM_DAMPING_LIMIT;
    }
    const double now = S.merit[run];
    push_history(S, run, now);
    if (o.stall_window > 0 && S.status[run] == MKO_LM_RUNNING) {
        const int w = o.stall_window;
        const int size = S.history_len[run];
        if (size > w) {
            const double old = S.history[size_t(run) * S.H + ((size - 1 - w) % S.H)];
            if (old - now <= o.stall_tolerance * old) S.status[run] = MKO_LM_CONVERGED_STALL;
        }
    }
}

__global__ __launch_bounds__(64) void fd_pattern_kernel(int count, int P, const double* xs,
                                                        const double* lower, const double* upper,
                                                        const double* typical, double* plus,
                                                        double* minus, double* step_plus,
                                                        double* step_minus, uint8_t* stencil) {
    const int j = blockIdx.x;
    if (j >= count) return;
    for (int k = threadIdx.x; k < P; k += 64) {
        const size_t at = size_t(j) * P + k;
        const double x = xs[at];
        const double lo = lower[k], hi = upper[k];
        const double h = kFdStepFactor * dmax(fabs(x), typical[k]);
        const double up = x + h, down = x - h;
        double xa, xb;
        uint8_t sten = MKO_STENCIL_CENTRAL;
        if (up <= hi && down >= lo) {
            xa = up;
            xb = down;
        } else {
            const bool towards_upper = hi - x >= x - lo;
            const double far = towards_upper ? x + 2.0 * h : x - 2.0 * h;
            if (towards_upper ? far <= hi : far >= lo) {
                sten = MKO_STENCIL_ONE_SIDED;
                xa = towards_upper ? up : down;
                xb = far;
            } else {
                sten = MKO_STENCIL_TRUNCATED;
                xa = up <= hi ? up : hi;
                xb = down >= lo ? down : lo;
            }
        }
        plus[at] = xa;
        minus[at] = xb;
        step_plus[at] = xa - x;
        step_minus[at] = xb - x;
        stencil[at] = sten;
    }
}

inline hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }
inline size_t tri_bytes(int P) {
    // the triangle with its entries' row and column bytes, then the solve's arrays (6 doubles,
    // an int and a byte per variable)
    return solve_arrays_offset(P) + size_t(P) * (6 * sizeof(double) + sizeof(int) + 1);
}
inline unsigned blocks_of(int n, int per) { return unsigned((n + per - 1) / per); }

} // namespace

void launch_init(const State& s, const Options& o, void* stream) {
    if (s.n > 0) hipLaunchKernelGGL(init_kernel, dim3(blocks_of(s.n, 256)), dim3(256), 0, st(stream), s, o);
}

void launch_start(const State& s, const double* merit, const int64_t* failed, const int64_t* hard,
                  void* stream) {
    if (s.n > 0)
        hipLaunchKernelGGL(start_kernel, dim3(blocks_of(s.n, 256)), dim3(256), 0, st(stream), s,
                           merit, failed, hard);
}

void launch_select(const State& s, int kind, int max_iterations, const int32_t* in,
                   const int32_t* in_count, int32_t* list, int32_t* count, int32_t* pos,
                   void* stream) {
    hipLaunchKernelGGL(select_kernel, dim3(1), dim3(1024), 0, st(stream), s, kind, max_iterations,
                       in, in_count, list, count, pos);
}

void launch_split(const State& s, const int32_t* in, const int32_t* in_count, int32_t* clean,
                  int32_t* clean_count, int32_t* other, int32_t* other_count, void* stream) {
    hipLaunchKernelGGL(split_kernel, dim3(1), dim3(1024), 0, st(stream), s, in, in_count, clean,
                       clean_count, other, other_count);
}

void launch_count_failing(const State& s, int max_iterations, int32_t* count, void* stream) {
    hipLaunchKernelGGL(count_failing_kernel, dim3(1), dim3(1024), 0, st(stream), s,
                       max_iterations, count);
}

void launch_gather(const State& s, const int32_t* list, int count, bool trial,
                   const int32_t* round_index, double* xs, int32_t* tidx, void* stream,
                   const int32_t* dcount) {
    if (count > 0)
        hipLaunchKernelGGL(gather_kernel, dim3(unsigned(count)), dim3(64), 0, st(stream), s, list,
                           count, dcount, trial ? 1 : 0, round_index, xs, tidx);
}

void launch_prepare(const State& s, const Options& o, const int32_t* list, int count,
                    const double* a_c, const double* g_c, void* stream, const int32_t* dcount) {
    if (count > 0)
        hipLaunchKernelGGL(prepare_kernel, dim3(unsigned(count)), dim3(kLanes), tri_bytes(s.P),
                           st(stream), s, o, list, count, dcount, a_c, g_c);
}

void launch_step(const State& s, const Options& o, const int32_t* list, int count, void* stream,
                 const int32_t* dcount) {
    if (count > 0)
        hipLaunchKernelGGL(step_kernel, dim3(unsigned(count)), dim3(kLanes), tri_bytes(s.P),
                           st(stream), s, o, list, count, dcount);
}

void launch_update(const State& s, const Options& o, const int32_t* list, int count,
                   const int32_t* pos, const double* f, const int64_t* failed,
                   const int64_t* hard, void* stream, const int32_t* dcount) {
    if (count > 0)
        hipLaunchKernelGGL(update_kernel, dim3(blocks_of(count, 256)), dim3(256), 0, st(stream), s,
                           o, list, count, dcount, pos, f, failed, hard);
}

void launch_fd_pattern(int count, int P, const double* xs, const double* lower,
                       const double* upper, const double* typical, double* plus, double* minus,
                       double* step_plus, double* step_minus, uint8_t* stencil, void* stream) {
    if (count > 0)
        hipLaunchKernelGGL(fd_pattern_kernel, dim3(unsigned(count)), dim3(64), 0, st(stream), count,
                           P, xs, lower, upper, typical, plus, minus, step_plus, step_minus,
                           stencil);
}

} // namespace mko_lmdev
