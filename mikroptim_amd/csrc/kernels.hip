// gfx950 kernels of the batched forward-and-Jacobian path and their device-level C entry points.
//
// Work decomposition (all kernels): one workgroup of 256 lanes (4 wave64) per (variant, tuple),
// tuple = (field f, wavelength l), t = f * L + l; the tuple's rays sit on lanes, ray k of the
// tuple on lane k mod 256 (iteration k / 256). The surface table of the (variant, wavelength)
// is staged once per workgroup in LDS (curvature, vertex distance, index, mu = n/n'), so the
// per-ray surface loop reads wave-uniform LDS broadcasts. Every reduction over rays (piston,
// merit partials, piston projection of the Jacobian) is a fixed-order tree: lane-local sum in
// iteration order, xor-butterfly inside the wave, waves summed in index order; results depend
// only on the tuple, never on the batch composition or the launch.
//
// Reference counterparts: residuals_kernel src/cuda/residuals.cu:58, merit_kernel +
// merit_finalize_kernel src/cuda/merit.cu:63/207, fd_difference_kernel src/cuda/fd_jacobian.cu:65,
// adjoint_jacobian_kernel src/cuda/adjoint_jacobian.cu:119, gram_kernel
// src/cuda/normal_equations.cu:48.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "error.hpp"
#include "gauss_legendre_table.h"
#include "internal.hpp"
#include "mikroptim_hip.h"
#include "constraints.hpp"
#include "trace_device.hpp"

namespace {

// RESIDUALS / MERIT traces in FP64 with the Newton-refined division and square root (a few ulp)
// instead of the IEEE sequences: 15 % fewer instructions per ray-surface at unchanged OPD
// parity with the oracle (max |dOPD| 4.2e-13 mm on the NA 0.95 objective either way); 0 keeps
// the IEEE operations.
#ifndef MKO_TRACE_FAST_MATH
#define MKO_TRACE_FAST_MATH 1
#endif
constexpr bool kTraceFastMath = MKO_TRACE_FAST_MATH != 0;

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxFields = 32;
constexpr int kMaxWavelengths = 32;
constexpr int kMaxRaysPerLane = 8; // tuples of at most 2048 rays
constexpr double kPi = 3.14159265358979323846264338327950288;

using mkod::Ray;
using mkod::Soft;
using mkod::V3;

// Everything a trace kernel needs, passed by value (< 1.5 KB of kernel arguments).
struct TraceParams {
    const double* curvature;
    const double* thickness;
    const double* index;
    const double* pupil_radius;
    const double* field_angle;
    const double* field_target;
    const double* reference_curvature;
    int V, S, L, F, T;
    int source, ref_kind;
    int rings, spokes, soft, precision;
    double fail_weight;
    double lambda_ref_nm;
    double src_pt[3];
    double src_na;
    uint64_t sphere_mask;
    uint64_t ext_mask; // surfaces whose extrema a kExt trace reduces
    // device count of the variants to trace (NULL: all V): blocks of variants >= *live exit at
    // once, so that a launch sized for V serves a captured LM cycle whose count the device sets
    const int32_t* live;
    int64_t rays, trace_rows, stride;
    double gl_u[MKO_MAX_RINGS], gl_w[MKO_MAX_RINGS];
    double field_weight[kMaxFields];
    double wavelength_weight[kMaxWavelengths];
    uint8_t pupil_mode[kMaxFields];
};

// Adjoint column map: variable index of curvature[s] / thickness[s] (-1 if not a variable).
struct ParamMap {
    int P;
    // 32-bit entries (here and in glass_col): the reverse sweeps index them by the (uniform)
    // surface, which scalar loads from the kernel arguments serve only at dword granularity
    int32_t curv_col[MKO_MAX_SURFACES];
    int32_t thick_col[MKO_MAX_SURFACES];
    // glass variables (AdjointParameterMap::glass / glass_slope, ray_jacobian.hpp:28-38): the
    // variables of (n0, nu1, nu2) of the medium after surface s, and the Buchdahl coordinate
    // w_l of each wavelength (dn/d(n0, nu1, nu2) = 1, w_l, w_l^2)
    int glass;
    int32_t glass_col[MKO_MAX_SURFACES][3];
    double omega[kMaxWavelengths];
};

// ---- tuple geometry ------------------------------------------------------------------------

struct Tuple {
    int v, f, l, n, M;
    int64_t offset;
    double scale, pscale;
};

__device__ __forceinline__ Tuple tuple_of(const TraceParams& P, int block) {
    Tuple t;
    t.v = block / P.T;
    const int tt = block - t.v * P.T;
    t.f = tt / P.L;
    t.l = tt - t.f * P.L;
    int64_t off = 0;
    for (int f = 0; f < t.f; ++f)
        off += int64_t(P.L) * P.rings * (P.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : P.spokes);
    t.M = P.pupil_mode[t.f] == MKO_PUPIL_AXIS ? 1 : P.spokes;
    t.n = P.rings * t.M;
    t.offset = off + int64_t(t.l) * t.n;
    // tuple_weight / residual_scale / penalty_scale (ray_residual.hpp:301-385): no a*b+c,
    // correctly rounded division and square root, hence bitwise equal to the host values.
    const double wf = P.field_weight[t.f];
    const double wl = P.wavelength_weight[t.l];
    double w = 0.0;
    if (wf > 0.0 && wl > 0.0) {
        double sf = 0.0, sl = 0.0;
        for (int i = 0; i < P.F; ++i) sf += P.field_weight[i];
        for (int i = 0; i < P.L; ++i) sl += P.wavelength_weight[i];
        w = __dmul_rn(__ddiv_rn(wf, sf), __ddiv_rn(wl, sl));
    }
    t.scale = w > 0.0 ? __ddiv_rn(__dsqrt_rn(w), __dmul_rn(P.lambda_ref_nm, 1e-6)) : 0.0;
    t.pscale = w > 0.0 ? __dsqrt_rn(__dmul_rn(P.fail_weight, w)) : 0.0;
    return t;
}

struct Sample {
    double rho, sphi, cphi, weight, rootw;
};

// PupilQuadrature::sample (quadrature.hpp:71-92): ring-major index k = i * M + j; ring i gives
// the radius and the weights, spoke j the angle.
__device__ __forceinline__ void ring_of(const TraceParams& P, const Tuple& t, int i, Sample& s) {
    s.rho = __dsqrt_rn(P.gl_u[i]);
    s.weight = __ddiv_rn(P.gl_w[i], double(t.M));
    s.rootw = __dsqrt_rn(s.weight);
}
__device__ __forceinline__ void spoke_of(const TraceParams& P, const Tuple& t, int j, Sample& s) {
    const uint8_t mode = P.pupil_mode[t.f];
    double phi = 0.0;
    if (mode == MKO_PUPIL_HALF)
        phi = __ddiv_rn(__dmul_rn(double(2 * j + 1), kPi), double(2 * t.M));
    else if (mode == MKO_PUPIL_FULL)
        phi = __ddiv_rn(__dmul_rn(double(2 * j + 1), kPi), double(t.M));
    if (mode == MKO_PUPIL_AXIS) {
        s.sphi = 0.0;
        s.cphi = 1.0;
    } else {
        sincos(phi, &s.sphi, &s.cphi);
    }
}
__device__ __forceinline__ Sample sample_of(const TraceParams& P, const Tuple& t, int k) {
    const int i = k / t.M;
    const int j = k - i * t.M;
    Sample s;
    ring_of(P, t, i, s);
    spoke_of(P, t, j, s);
    return s;
}

// The tuple's ring and spoke values in LDS, formed once per workgroup by the functions above (the
// same bits as sample_of): a ray's sample is then two table reads instead of a square root, a
// division and an FP64 sincos per ray. Tuples with more spokes than the table holds compute
// their samples per ray.
constexpr int kTableSpokes = 128;
struct SampleTable {
    double rho[MKO_MAX_RINGS], weight[MKO_MAX_RINGS], rootw[MKO_MAX_RINGS];
    double sphi[kTableSpokes], cphi[kTableSpokes];
};
// The caller's next __syncthreads publishes the table.
__device__ __forceinline__ void stage_samples(const TraceParams& P, const Tuple& t,
                                              SampleTable& T) {
    if (t.M > kTableSpokes) return;
    const int x = threadIdx.x;
    if (x < P.rings) {
        Sample s;
        ring_of(P, t, x, s);
        T.rho[x] = s.rho;
        T.weight[x] = s.weight;
        T.rootw[x] = s.rootw;
    }
    for (int j = x; j < t.M; j += int(blockDim.x)) {
        Sample s;
        spoke_of(P, t, j, s);
        T.sphi[j] = s.sphi;
        T.cphi[j] = s.cphi;
    }
}
__device__ __forceinline__ Sample sample_at(const TraceParams& P, const Tuple& t,
                                            const SampleTable& T, int k) {
    if (t.M > kTableSpokes) return sample_of(P, t, k);
    const int i = k / t.M;
    const int j = k - i * t.M;
    return Sample{T.rho[i], T.sphi[j], T.cphi[j], T.weight[i], T.rootw[i]};
}

// Surface table of (variant, wavelength) in LDS.
struct SurfLds {
    double c[MKO_MAX_SURFACES];   // effective curvature (0 for stop/plane)
    double t[MKO_MAX_SURFACES];   // thickness[s]
    double n[MKO_MAX_SURFACES];   // index[s][l]
    double mu[MKO_MAX_SURFACES];  // n[s-1] / n[s]
    double imu[MKO_MAX_SURFACES]; // n[s] / n[s-1]
};

__device__ __forceinline__ void stage_surfaces(const TraceParams& P, const Tuple& t, SurfLds& L) {
    const int s = threadIdx.x;
    if (s < P.S) {
        const size_t vs = size_t(t.v) * P.S + s;
        const bool sphere = (P.sphere_mask >> s) & 1ull;
        L.c[s] = sphere ? P.curvature[vs] : 0.0;
        L.t[s] = P.thickness[vs];
        L.n[s] = P.index[vs * P.L + t.l];
    }
    __syncthreads();
    if (s >= 1 && s < P.S) {
        L.mu[s] = __ddiv_rn(L.n[s - 1], L.n[s]);
        L.imu[s] = __ddiv_rn(L.n[s], L.n[s - 1]);
    }
    __syncthreads();
}

__device__ __forceinline__ Ray source_ray(const TraceParams& P, const Tuple& t, const Sample& s,
                                          double a, double sin_t, double cos_t, double n0) {
    Ray r;
    if (P.source == MKO_SOURCE_POINT) {
        const double sin_max = P.src_na / n0;
        const double sr = sin_max * s.rho;
        r.p = V3<double>{P.src_pt[0], P.src_pt[1], P.src_pt[2]};
        r.d = V3<double>{sr * s.sphi, sr * s.cphi, __dsqrt_rn((1.0 - sr) * (1.0 + sr))};
        r.opl = 0.0;
        return r;
    }
    const double rad = a * s.rho;
    r.p = V3<double>{rad * s.sphi, rad * s.cphi, 0.0};
    r.d = V3<double>{0.0, sin_t, cos_t};
    r.opl = mkod::dot(r.p, r.d);
    return r;
}

// Record of one ray's states for the reverse sweep (AdjointRecord, adjoint_record.hpp:66).
template <class RR>
struct Record {
    RR p[MKO_MAX_SURFACES][3];
    RR d[MKO_MAX_SURFACES][3];
    uint8_t branch[MKO_MAX_SURFACES];
    uint8_t image_branch;
    int count;
    static constexpr bool kLastOnly = false; // every surface's state
    __device__ __forceinline__ void put(int s, const Ray& r, uint8_t b) {
        p[s][0] = RR(r.p.x);
        p[s][1] = RR(r.p.y);
        p[s][2] = RR(r.p.z);
        d[s][0] = RR(r.d.x);
        d[s][1] = RR(r.d.y);
        d[s][2] = RR(r.d.z);
        branch[s] = b;
        count = s + 1;
    }
};

struct NoRecord {
    static constexpr bool kLastOnly = false;
    template <class RR>
    __device__ __forceinline__ void put(int, const RR&, uint8_t) {}
    uint8_t image_branch;
};

// Last state of a forward trace (the state on surface S-1, which the reference-sphere step starts
// from); the reverse sweep reconstructs the earlier states with mkod::inverse_step. The traces put
// it once, after the last surface (kLastOnly): a ray that fails on the way leaves it unset, and
// the loop carries no copy of the state.
struct LastState {
    V3<double> p, d;
    uint8_t image_branch;
    int count;
    static constexpr bool kLastOnly = true;
    __device__ __forceinline__ void put(int s, const Ray& r, uint8_t) {
        p = r.p;
        d = r.d;
        count = s + 1;
    }
};

// trace_to_reference_sphere (reference_sphere.hpp:183-222) over the LDS table, in the number
// type R of the ray (double, or df64 for MKO_PRECISION_DF64). Returns the final status;
// soft.penalty holds the accumulated penalty.
template <bool kSoft, bool kFast = false, class R, class Rec>
__device__ __forceinline__ uint8_t trace_ray(const TraceParams& P, const SurfLds& L,
                                             mkod::RayT<R>& r, V3<double> M, double kappa,
                                             Soft& soft, Rec& rec) {
    using mkod::num;
    if constexpr (!Rec::kLastOnly) rec.put(0, r, mkod::kBranchRegular);
    for (int s = 1; s < P.S; ++s) {
        uint8_t branch = 0;
        const uint8_t st = mkod::surface_step<kSoft, kFast>(r, num<R>(L.c[s]), num<R>(L.t[s - 1]),
                                                     num<R>(L.n[s - 1]), num<R>(L.mu[s]), soft,
                                                     branch);
        if (st != MKO_STATUS_OK) return st;
        if constexpr (!Rec::kLastOnly) rec.put(s, r, branch);
    }
    if constexpr (Rec::kLastOnly) rec.put(P.S - 1, r, mkod::kBranchRegular);
    if (!(r.d.z > num<R>(0.0))) return MKO_STATUS_BACKWARD;
    uint8_t rb = 0;
    const uint8_t rst = mkod::reference_step<kSoft, kFast>(
        r, V3<R>{num<R>(M.x), num<R>(M.y), num<R>(M.z)}, num<R>(kappa), P.ref_kind,
        num<R>(L.n[P.S - 1]), soft, rb);
    if (!mkod::in_piston(rst)) return rst;
    rec.image_branch = rb;
    return kSoft ? soft.first : uint8_t(MKO_STATUS_OK);
}

// The start ray in the trace's number type (computed in double, then converted).
template <class R>
__device__ __forceinline__ mkod::RayT<R> ray_as(const Ray& r) {
    using mkod::num;
    return mkod::RayT<R>{V3<R>{num<R>(r.p.x), num<R>(r.p.y), num<R>(r.p.z)},
                         V3<R>{num<R>(r.d.x), num<R>(r.d.y), num<R>(r.d.z)}, num<R>(r.opl)};
}

// ---- surface extrema (SurfaceExtrema, include/mikroptim/trace/surface_extrema.hpp:47-110) -----

// height^2 = x^2 + y^2 at the intersection and cos I = d_before . N, N = (-c x, -c y, 1 - c z),
// each product and sum rounded on its own (RaySurfaceRecord::surface).
__device__ __forceinline__ void surface_extremum(double x, double y, double z, double c, double dx,
                                                 double dy, double dz, double& h, double& ci) {
#pragma clang fp contract(off)
    h = x * x + y * y;
    ci = dx * (-c * x) + dy * (-c * y) + dz * (1.0 - c * z);
}

// trace_ray with the surface extrema of the wave's rays: at every surface s in P.ext_mask (0 =
// start state; the other surfaces keep the neutral 0 / 1),
// the largest height^2 and the smallest cos I over the wave's lanes whose ray is `alive` (has
// passed every surface so far) are reduced across the wave; lane s keeps surface s's values in
// eh / ec (S <= 64 = wave width), replacing them (merge = false) or folding them in (merge =
// true). Every lane of the wave calls this with the same P.S (uniform control flow around the
// cross-lane reductions); lanes without a ray pass alive = false.
template <bool kSoft, bool kFast, class R, class Rec>
__device__ __forceinline__ uint8_t trace_ray_ext(const TraceParams& P, const SurfLds& L,
                                                 mkod::RayT<R>& r, V3<double> M, double kappa,
                                                 Soft& soft, bool alive, int lane, double& eh,
                                                 double& ec, bool merge, Rec& rec) {
    using mkod::num;
    using mkod::to_double;
    auto reduce = [&](int s, double h, double c) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double oh = __shfl_xor(h, off, 64);
            const double oc = __shfl_xor(c, off, 64);
            h = oh > h ? oh : h;
            c = oc < c ? oc : c;
        }
        if (lane == s) {
            eh = merge ? (h > eh ? h : eh) : h;
            ec = merge ? (c < ec ? c : ec) : c;
        }
    };
    uint8_t st = MKO_STATUS_OK;
    {
        double h = 0.0, c = 1.0;
        if (alive)
            surface_extremum(to_double(r.p.x), to_double(r.p.y), to_double(r.p.z), 0.0,
                             to_double(r.d.x), to_double(r.d.y), to_double(r.d.z), h, c);
        if (P.ext_mask & 1ull) reduce(0, h, c);
    }
    if constexpr (!Rec::kLastOnly)
        if (alive) rec.put(0, r, mkod::kBranchRegular);
    for (int s = 1; s < P.S; ++s) {
        double h = 0.0, c = 1.0;
        if (alive) {
            const double dx = to_double(r.d.x), dy = to_double(r.d.y), dz = to_double(r.d.z);
            uint8_t branch = 0;
            const uint8_t sst = mkod::surface_step<kSoft, kFast>(
                r, num<R>(L.c[s]), num<R>(L.t[s - 1]), num<R>(L.n[s - 1]), num<R>(L.mu[s]), soft,
                branch);
            if (sst != MKO_STATUS_OK) {
                alive = false;
                st = sst;
            } else {
                if constexpr (!Rec::kLastOnly) rec.put(s, r, branch);
                surface_extremum(to_double(r.p.x), to_double(r.p.y), to_double(r.p.z), L.c[s], dx,
                                 dy, dz, h, c);
            }
        }
        if ((P.ext_mask >> s) & 1ull) reduce(s, h, c); // uniform: P.ext_mask is a kernel argument
    }
    if (!alive) return st;
    if constexpr (Rec::kLastOnly) rec.put(P.S - 1, r, mkod::kBranchRegular);
    if (!(r.d.z > num<R>(0.0))) return MKO_STATUS_BACKWARD;
    uint8_t rb = 0;
    const uint8_t rst = mkod::reference_step<kSoft, kFast>(
        r, V3<R>{num<R>(M.x), num<R>(M.y), num<R>(M.z)}, num<R>(kappa), P.ref_kind,
        num<R>(L.n[P.S - 1]), soft, rb);
    if (!mkod::in_piston(rst)) return rst;
    rec.image_branch = rb;
    return kSoft ? soft.first : uint8_t(MKO_STATUS_OK);
}

// ---- deterministic block reductions -----------------------------------------------------------

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <class T>
__device__ __forceinline__ T block_sum(T v, T* lds) {
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    T total = lds[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) total += lds[w];
    __syncthreads();
    return total;
}

// ---- RESIDUALS / MERIT ----------------------------------------------------------------------

// Mode 0: RESIDUALS (residual rows, penalty rows, OPD, status). Mode 1: MERIT partials per
// (variant, tuple): sum r^2 + sum p^2, failed and hard-failed counts. R: number type of the ray
// trace (double or df64); OPL, piston and residuals are formed in double.
__device__ __forceinline__ int64_t state_index(const mko_internal::RayStates& rs, int v, int c,
                                               int64_t row) {
    const int64_t slot = rs.slot ? rs.slot[v] : int64_t(rs.slot_offset) + v;
    return (int64_t(c) * rs.slots + slot) * rs.rays + row;
}

// kExt: also the surface extrema of the tuple's rays with status Ok, ext_partial[block][S][2]
// (largest height^2, smallest cos I); a wave that holds a ray with another status traces its
// valid rays a second time for them.
template <int kMode, bool kSoft, int kRpl, class R, bool kStates = false, bool kExt = false>
__global__ __launch_bounds__(kBlock) void trace_kernel(TraceParams P, double* residuals,
                                                       double* opd, uint8_t* status,
                                                       double* partial, int* pfailed,
                                                       int* phard,
                                                       mko_internal::RayStates states,
                                                       double* ext_partial) {
    __shared__ SurfLds L;
    __shared__ SampleTable ST;
    __shared__ double red[kWaves];
    __shared__ int ired[kWaves];
    if (P.live && int(blockIdx.x) / P.T >= *P.live) return;
    const Tuple t = tuple_of(P, blockIdx.x);
    stage_samples(P, t, ST);
    stage_surfaces(P, t, L);
    const size_t vf = size_t(t.v) * P.F + t.f;
    const double a = P.pupil_radius[t.v];
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
    const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;

    // when the spoke count divides the workgroup, lane x keeps spoke x mod M in every ray slot i
    // and its ring advances by 256 / M per slot: the samples without a division per ray
    const bool lane_spoke = kBlock % t.M == 0 && t.M <= kTableSpokes;
    const int ring0 = int(threadIdx.x) / t.M, spoke0 = int(threadIdx.x) - ring0 * t.M;
    const int ring_step = kBlock / t.M;
    auto sample_slot = [&](int i, int k) {
        if (!lane_spoke) return sample_at(P, t, ST, k);
        const int ring = ring0 + i * ring_step;
        return Sample{ST.rho[ring], ST.sphi[spoke0], ST.cphi[spoke0], ST.weight[ring],
                      ST.rootw[ring]};
    };
    double opl[kRpl], pen[kRpl];
    uint8_t st[kRpl];
    double sw = 0.0, swo = 0.0;
    const int lane = threadIdx.x & 63;
    double eh = 0.0, ec = 1.0; // surface `lane`'s extrema over this wave's valid rays
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const int k = threadIdx.x + i * kBlock;
        st[i] = MKO_STATUS_UNSUPPORTED;
        opl[i] = 0.0;
        pen[i] = 0.0;
        if constexpr (kExt) {
            constexpr bool kFast = kTraceFastMath && std::is_same_v<R, double>;
            const bool have = k < t.n;
            Sample s{};
            Ray r0{};
            if (have) {
                s = sample_slot(i, k);
                r0 = source_ray(P, t, s, a, sin_t, cos_t, L.n[0]);
            }
            mkod::RayT<R> r = ray_as<R>(r0);
            Soft soft{0.0, MKO_STATUS_OK};
            double pend_h = 0.0, pend_c = 1.0;
            std::conditional_t<kStates, LastState, NoRecord> rec{};
            const uint8_t sti = trace_ray_ext<kSoft, kFast>(P, L, r, M, kappa, soft, have, lane,
                                                            pend_h, pend_c, false, rec);
            const uint64_t have_mask = __ballot(have);
            const uint64_t ok_mask = __ballot(have && sti == MKO_STATUS_OK);
            if (ok_mask == have_mask) {
                if (lane < P.S) {
                    eh = pend_h > eh ? pend_h : eh;
                    ec = pend_c < ec ? pend_c : ec;
                }
            } else {
                // some ray of the wave did not end with status Ok: its earlier surfaces must not
                // count, so the wave's valid rays are traced again and folded in directly
                mkod::RayT<R> rr = ray_as<R>(r0);
                Soft s2{0.0, MKO_STATUS_OK};
                NoRecord nr;
                trace_ray_ext<kSoft, kFast>(P, L, rr, M, kappa, s2, have && sti == MKO_STATUS_OK,
                                            lane, eh, ec, true, nr);
            }
            if (have) {
                st[i] = sti;
                if constexpr (kStates) {
                    const int64_t row = t.offset + k;
                    const bool okr = sti == MKO_STATUS_OK;
                    if (okr) {
                        states.data[state_index(states, t.v, 0, row)] = rec.p.x;
                        states.data[state_index(states, t.v, 1, row)] = rec.p.y;
                        states.data[state_index(states, t.v, 2, row)] = rec.p.z;
                        states.data[state_index(states, t.v, 3, row)] = rec.d.x;
                        states.data[state_index(states, t.v, 4, row)] = rec.d.y;
                        states.data[state_index(states, t.v, 5, row)] = rec.d.z;
                    }
                    states.data[state_index(states, t.v, 6, row)] =
                        okr ? mkod::to_double(r.opl) : __builtin_nan("");
                }
                opl[i] = mkod::in_piston(sti) ? mkod::to_double(r.opl) : 0.0;
                pen[i] = soft.penalty;
                if (mkod::in_piston(sti)) {
                    sw += s.weight;
                    swo += s.weight * opl[i];
                }
            }
            continue;
        }
        if (k < t.n) {
            const Sample s = sample_slot(i, k);
            mkod::RayT<R> r = ray_as<R>(source_ray(P, t, s, a, sin_t, cos_t, L.n[0]));
            Soft soft{0.0, MKO_STATUS_OK};
            constexpr bool kFast = kTraceFastMath && std::is_same_v<R, double>;
            if constexpr (kStates) {
                LastState last;
                last.count = 0;
                last.image_branch = 0;
                st[i] = trace_ray<kSoft, kFast>(P, L, r, M, kappa, soft, last);
                const int64_t row = t.offset + k;
                const bool okr = st[i] == MKO_STATUS_OK;
                if (okr) {
                    states.data[state_index(states, t.v, 0, row)] = last.p.x;
                    states.data[state_index(states, t.v, 1, row)] = last.p.y;
                    states.data[state_index(states, t.v, 2, row)] = last.p.z;
                    states.data[state_index(states, t.v, 3, row)] = last.d.x;
                    states.data[state_index(states, t.v, 4, row)] = last.d.y;
                    states.data[state_index(states, t.v, 5, row)] = last.d.z;
                }
                states.data[state_index(states, t.v, 6, row)] =
                    okr ? mkod::to_double(r.opl) : __builtin_nan("");
            } else {
                NoRecord nr;
                st[i] = trace_ray<kSoft, kFast>(P, L, r, M, kappa, soft, nr);
            }
            // a ray's OPL counts only with an in-piston status; selecting it here leaves the ray
            // dead on the surface loop's failure exits (no copies of it kept for them)
            opl[i] = mkod::in_piston(st[i]) ? mkod::to_double(r.opl) : 0.0;
            pen[i] = soft.penalty;
            if (mkod::in_piston(st[i])) {
                sw += s.weight;
                swo += s.weight * opl[i];
            }
        }
    }
    const double w_ok = block_sum(sw, red);
    const double sum = block_sum(swo, red);
    const bool valid = w_ok > 0.0;
    const double piston = valid ? sum / w_ok : 0.0;
    double r2 = 0.0, p2 = 0.0;
    int nfail = 0, nhard = 0;
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const int k = threadIdx.x + i * kBlock;
        if (k < t.n) {
            const Sample s = sample_slot(i, k);
            const bool ok = valid && mkod::in_piston(st[i]);
            const double e = ok ? opl[i] - piston : 0.0;
            const double r = t.scale * s.rootw * e;
            double pr = 0.0;
            if (kSoft) {
                const double fp = st[i] == MKO_STATUS_OK ? 0.0
                                  : mkod::is_soft(st[i]) ? pen[i]
                                                         : pen[i] + 1.0;
                pr = t.pscale * s.rootw * fp;
            }
            const int64_t row = t.offset + k;
            if (kMode == 0) {
                const size_t base = size_t(t.v) * P.stride;
                residuals[base + row] = r;
                if (kSoft) residuals[base + P.rays + row] = pr;
                if (opd) opd[size_t(t.v) * P.rays + row] = e;
                if (status) status[size_t(t.v) * P.rays + row] = st[i];
            } else {
                r2 += r * r;
                p2 += pr * pr;
                nfail += st[i] != MKO_STATUS_OK;
                nhard += !mkod::in_piston(st[i]);
            }
        }
    }
    if (kMode == 1) {
        const double tr2 = block_sum(r2, red);
        const double tp2 = kSoft ? block_sum(p2, red) : 0.0;
        const int tf = block_sum(nfail, ired);
        const int th = block_sum(nhard, ired);
        if (threadIdx.x == 0) {
            const size_t idx = size_t(t.v) * P.T + size_t(blockIdx.x - t.v * P.T);
            partial[idx] = kSoft ? tr2 + tp2 : tr2;
            pfailed[idx] = tf;
            phard[idx] = th;
        }
    }
    if constexpr (kExt) {
        __shared__ double ext_w[kWaves][MKO_MAX_SURFACES][2];
        const int wave = threadIdx.x >> 6;
        if (lane < P.S) {
            ext_w[wave][lane][0] = eh;
            ext_w[wave][lane][1] = ec;
        }
        __syncthreads();
        if (int(threadIdx.x) < P.S) {
            double h = ext_w[0][threadIdx.x][0], c = ext_w[0][threadIdx.x][1];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) {
                const double hw = ext_w[w][threadIdx.x][0], cw = ext_w[w][threadIdx.x][1];
                h = hw > h ? hw : h;
                c = cw < c ? cw : c;
            }
            double* out = ext_partial + (size_t(blockIdx.x) * P.S + threadIdx.x) * 2;
            out[0] = h;
            out[1] = c;
        }
    }
}

// ---- tuple packs: consecutive tuples of a variant share a workgroup -------------------------

// Surface tables of the pack's tuples (tuple t0 + j in Ls[j]); as stage_surfaces per tuple.
__device__ __forceinline__ void stage_surfaces_pack(const TraceParams& P, int v, int t0, int tpb,
                                                    SurfLds* Ls) {
    const int n = tpb * P.S;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int j = e / P.S, s = e - j * P.S;
        const int tt = t0 + j;
        if (tt >= P.T) continue;
        const int l = tt - (tt / P.L) * P.L;
        const size_t vs = size_t(v) * P.S + s;
        const bool sphere = (P.sphere_mask >> s) & 1ull;
        Ls[j].c[s] = sphere ? P.curvature[vs] : 0.0;
        Ls[j].t[s] = P.thickness[vs];
        Ls[j].n[s] = P.index[vs * P.L + l];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int j = e / P.S, s = e - j * P.S;
        if (t0 + j >= P.T || s < 1) continue;
        Ls[j].mu[s] = __ddiv_rn(Ls[j].n[s - 1], Ls[j].n[s]);
        Ls[j].imu[s] = __ddiv_rn(Ls[j].n[s], Ls[j].n[s - 1]);
    }
    __syncthreads();
}

// Sum over a pack segment of `seg` lanes (a power of two, 32 ... 256): for seg >= 64 the wave
// sums of the segment's waves added in wave order, for seg < 64 the xor butterfly over the
// segment's lanes (offsets seg / 2 ... 1). The same tree as block_sum of one tuple per workgroup,
// whose other lanes (and waves) hold zeros: the steps it has beyond the segment add exact zeros,
// so the sums are equal bit for bit.
template <class T>
__device__ __forceinline__ T seg_sum(T v, int seg, T* lds) {
    if (seg >= 64) {
        v = wave_sum(v);
        const int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) lds[wave] = v;
        __syncthreads();
        const int w0 = (int(threadIdx.x) / seg) * (seg / 64);
        T total = lds[w0];
        for (int w = 1; w < seg / 64; ++w) total += lds[w0 + w];
        __syncthreads();
        return total;
    }
    for (int off = seg / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// RESIDUALS (kMode 0) / MERIT partials (kMode 1) of tuples of at most `seg` rays (seg 32, 64 or
// 128), 256 / seg consecutive tuples of one variant per workgroup (make_tuple_packs,
// include/mikroptim/cuda/tuple_packs.hpp:61, with a pack size fixed per launch): tuple t0 + j
// on lanes [j seg, (j + 1) seg), one ray per lane, its surface table in Ls[j]. Every value is
// the one trace_kernel computes for the tuple alone (FP64 trace, no extrema).
template <int kMode, bool kSoft, bool kStates>
__global__ __launch_bounds__(kBlock) void trace_pack_kernel(TraceParams P, int seg,
                                                            double* residuals, double* opd,
                                                            uint8_t* status, double* partial,
                                                            int* pfailed, int* phard,
                                                            mko_internal::RayStates states) {
    extern __shared__ SurfLds Ls[];
    __shared__ double red[kWaves];
    __shared__ int ired[kWaves];
    const int tpb = kBlock / seg;
    const int bpv = (P.T + tpb - 1) / tpb;
    const int v = int(blockIdx.x) / bpv;
    const int t0 = (int(blockIdx.x) - v * bpv) * tpb;
    if (P.live && v >= *P.live) return;
    stage_surfaces_pack(P, v, t0, tpb, Ls);
    const int j = int(threadIdx.x) / seg, k = int(threadIdx.x) - j * seg;
    const int tt = t0 + j;
    const bool in_pack = tt < P.T;
    const SurfLds& L = Ls[j];
    const Tuple t = tuple_of(P, v * P.T + (in_pack ? tt : 0));
    const size_t vf = size_t(t.v) * P.F + t.f;
    const double a = P.pupil_radius[t.v];
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
    const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;
    const bool have = in_pack && k < t.n;
    double opl = 0.0, pen = 0.0, sw = 0.0, swo = 0.0;
    uint8_t st = MKO_STATUS_UNSUPPORTED;
    Sample smp{};
    if (have) {
        smp = sample_of(P, t, k);
        mkod::RayT<double> r = source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]);
        Soft soft{0.0, MKO_STATUS_OK};
        if constexpr (kStates) {
            LastState last;
            last.count = 0;
            last.image_branch = 0;
            st = trace_ray<kSoft, kTraceFastMath>(P, L, r, M, kappa, soft, last);
            const int64_t row = t.offset + k;
            const bool okr = st == MKO_STATUS_OK;
            if (okr) {
                states.data[state_index(states, t.v, 0, row)] = last.p.x;
                states.data[state_index(states, t.v, 1, row)] = last.p.y;
                states.data[state_index(states, t.v, 2, row)] = last.p.z;
                states.data[state_index(states, t.v, 3, row)] = last.d.x;
                states.data[state_index(states, t.v, 4, row)] = last.d.y;
                states.data[state_index(states, t.v, 5, row)] = last.d.z;
            }
            states.data[state_index(states, t.v, 6, row)] = okr ? r.opl : __builtin_nan("");
        } else {
            NoRecord nr;
            st = trace_ray<kSoft, kTraceFastMath>(P, L, r, M, kappa, soft, nr);
        }
        opl = r.opl;
        pen = soft.penalty;
        if (mkod::in_piston(st)) {
            sw += smp.weight;
            swo += smp.weight * opl;
        }
    }
    const double w_ok = seg_sum(sw, seg, red);
    const double sum = seg_sum(swo, seg, red);
    const bool valid = w_ok > 0.0;
    const double piston = valid ? sum / w_ok : 0.0;
    double r2 = 0.0, p2 = 0.0;
    int nfail = 0, nhard = 0;
    if (have) {
        const bool ok = valid && mkod::in_piston(st);
        const double e = ok ? opl - piston : 0.0;
        const double r = t.scale * smp.rootw * e;
        double pr = 0.0;
        if (kSoft) {
            const double fp = st == MKO_STATUS_OK ? 0.0 : mkod::is_soft(st) ? pen : pen + 1.0;
            pr = t.pscale * smp.rootw * fp;
        }
        const int64_t row = t.offset + k;
        if (kMode == 0) {
            const size_t base = size_t(t.v) * P.stride;
            residuals[base + row] = r;
            if (kSoft) residuals[base + P.rays + row] = pr;
            if (opd) opd[size_t(t.v) * P.rays + row] = e;
            if (status) status[size_t(t.v) * P.rays + row] = st;
        } else {
            r2 = r * r;
            p2 = pr * pr;
            nfail = st != MKO_STATUS_OK;
            nhard = !mkod::in_piston(st);
        }
    }
    if (kMode == 1) {
        const double tr2 = seg_sum(r2, seg, red);
        const double tp2 = kSoft ? seg_sum(p2, seg, red) : 0.0;
        const int tf = seg_sum(nfail, seg, ired);
        const int th = seg_sum(nhard, seg, ired);
        if (k == 0 && in_pack) {
            const size_t idx = size_t(v) * P.T + size_t(tt);
            partial[idx] = kSoft ? tr2 + tp2 : tr2;
            pfailed[idx] = tf;
            phard[idx] = th;
        }
    }
}

// Extrema of every variant from the tuple partials: out[v][s] = {sqrt(max height^2), min cos I}
// over the variant's tuples (extrema_finalize_kernel of the reference, wavefront_block.hpp:414).
__global__ void extrema_finalize_kernel(int V, int T, int S, const double* partial, double* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(V) * S) return;
    const int64_t v = i / S, s = i - v * S;
    double h = 0.0, c = 1.0;
    for (int t = 0; t < T; ++t) {
        const double* p = partial + ((v * T + t) * S + s) * 2;
        h = p[0] > h ? p[0] : h;
        c = p[1] < c ? p[1] : c;
    }
    out[i * 2] = __dsqrt_rn(h);
    out[i * 2 + 1] = c;
}

// Check of the candidate rays after the differences' traces, one thread per run: on every height
// surface the candidates' largest height over the run's per_run points (ex [runs * per_run][S][2])
// moved by at most D from the run's (ex0); the other rays lie below (1 - delta) of the run's
// largest height at its point and would have to move by more than K D to reach the candidates'
// smallest; likewise for the smallest cos I on the incidence surfaces (the others above
// cmin + delta (1 - cmin)). flag[r] = 1 where that margin is missing or the list overflowed.
__global__ void extrema_check_kernel(int runs, int per_run, int S, uint64_t hmask, uint64_t cmask,
                                     double delta, const double* ex, const double* ex0,
                                     const int* others, const int* overflow, double K,
                                     int* flag) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs) return;
    int bad = overflow[r];
    if (!bad && others[r] > 0) {
        for (int s = 0; s < S; ++s) {
            if (!(((hmask | cmask) >> s) & 1ull)) continue;
            const double h0 = ex0[(size_t(r) * S + s) * 2], c0 = ex0[(size_t(r) * S + s) * 2 + 1];
            double dh = 0.0, dc = 0.0, hmin = h0, cmax = c0;
            for (int j = 0; j < per_run; ++j) {
                const double* e = ex + ((size_t(r) * per_run + j) * S + s) * 2;
                dh = fmax(dh, fabs(e[0] - h0));
                dc = fmax(dc, fabs(e[1] - c0));
                hmin = fmin(hmin, e[0]);
                cmax = fmax(cmax, e[1]);
            }
            if (((hmask >> s) & 1ull) && !((1.0 - delta) * h0 + K * dh < hmin)) bad = 1;
            if (((cmask >> s) & 1ull) && !(c0 + delta * (1.0 - c0) - K * dc > cmax)) bad = 1;
        }
    }
    flag[r] = bad;
}

// ---- candidate rays of the constraint rows' central differences ----------------------------

// Per-ray candidates at the runs' points: one 256-lane workgroup per (run, tuple) of the points'
// batch (variant = run), each ray traced as trace_ray_ext traces it. A ray is a candidate where it
// does not end with status Ok (at a difference point it may), or where on a height surface
// (hmask) its height reaches (1 - delta) of the run's largest, or on an incidence surface (cmask)
// its cos I lies within delta (1 - cos) of the run's smallest (ex0 [runs][S][2]). flags [runs *
// rays] (1 = candidate), tcount [runs][T] candidates per tuple, others [runs] (atomic count of the
// non-candidates).
template <bool kSoft>
__global__ __launch_bounds__(kBlock) void ray_candidates_kernel(TraceParams P, uint64_t hmask,
                                                               uint64_t cmask, double delta,
                                                               const double* ex0, uint8_t* flags,
                                                               int* tcount, int* others) {
    __shared__ SurfLds L;
    __shared__ SampleTable ST;
    __shared__ double thr[2][MKO_MAX_SURFACES];
    __shared__ int red[kWaves];
    const Tuple t = tuple_of(P, blockIdx.x);
    if (int(threadIdx.x) < P.S) {
        const double hmax = ex0[(size_t(t.v) * P.S + threadIdx.x) * 2];
        const double cmin = ex0[(size_t(t.v) * P.S + threadIdx.x) * 2 + 1];
        thr[0][threadIdx.x] = (1.0 - delta) * hmax;
        thr[1][threadIdx.x] = cmin + delta * (1.0 - cmin);
    }
    stage_samples(P, t, ST);
    stage_surfaces(P, t, L); // synchronises
    const size_t vf = size_t(t.v) * P.F + t.f;
    const double a = P.pupil_radius[t.v];
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
    const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;
    const uint64_t mask = hmask | cmask;
    int cands = 0, rest = 0;
    for (int k = threadIdx.x; k < t.n; k += kBlock) {
        const Sample smp = sample_at(P, t, ST, k);
        mkod::RayT<double> r = ray_as<double>(source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]));
        Soft soft{0.0, MKO_STATUS_OK};
        bool cand = false;
        auto mark = [&](int sf, double x, double y, double z, double c, double dx, double dy,
                        double dz) {
            if (!((mask >> sf) & 1ull)) return;
            double h2 = 0.0, ci = 1.0;
            surface_extremum(x, y, z, c, dx, dy, dz, h2, ci);
            if (((hmask >> sf) & 1ull) && __dsqrt_rn(h2) >= thr[0][sf]) cand = true;
            if (((cmask >> sf) & 1ull) && ci <= thr[1][sf]) cand = true;
        };
        mark(0, r.p.x, r.p.y, r.p.z, 0.0, r.d.x, r.d.y, r.d.z);
        uint8_t st = MKO_STATUS_OK;
        for (int sf = 1; sf < P.S; ++sf) {
            const double dx = r.d.x, dy = r.d.y, dz = r.d.z;
            uint8_t branch = 0;
            const uint8_t sst = mkod::surface_step<kSoft, kTraceFastMath>(
                r, L.c[sf], L.t[sf - 1], L.n[sf - 1], L.mu[sf], soft, branch);
            if (sst != MKO_STATUS_OK) {
                st = sst;
                break;
            }
            mark(sf, r.p.x, r.p.y, r.p.z, L.c[sf], dx, dy, dz);
        }
        if (st == MKO_STATUS_OK) {
            if (!(r.d.z > 0.0)) {
                st = MKO_STATUS_BACKWARD;
            } else {
                uint8_t rb = 0;
                const uint8_t rst = mkod::reference_step<kSoft, kTraceFastMath>(
                    r, M, kappa, P.ref_kind, L.n[P.S - 1], soft, rb);
                st = !mkod::in_piston(rst) ? rst : kSoft ? soft.first : uint8_t(MKO_STATUS_OK);
            }
        }
        if (st != MKO_STATUS_OK) cand = true;
        flags[size_t(t.v) * P.rays + t.offset + k] = cand ? 1 : 0;
        cands += cand;
        rest += !cand;
    }
    const int tc = block_sum(cands, red);
    const int tr = block_sum(rest, red);
    if (threadIdx.x == 0) {
        tcount[blockIdx.x] = tc;
        if (tr) atomicAdd(others + t.v, tr);
    }
}

// Ordered candidate list of each run (one 1024-thread workgroup per run): the rows with a flag in
// ascending order (so grouped by tuple), list [runs][cap]; tstart [runs][T + 1] the tuples'
// first entries; overflow[run] = 1 where more than cap rays are candidates.
__global__ __launch_bounds__(1024) void candidate_list_kernel(int T, int64_t rays, int cap,
                                                              const uint8_t* flags,
                                                              const int* tcount, int* tstart,
                                                              int* list, int* overflow) {
    __shared__ int32_t wave_total[16];
    __shared__ int32_t base_s;
    const int r = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        int acc = 0;
        for (int t = 0; t < T; ++t) {
            tstart[size_t(r) * (T + 1) + t] = acc;
            acc += tcount[size_t(r) * T + t];
        }
        tstart[size_t(r) * (T + 1) + T] = acc;
        overflow[r] = acc > cap ? 1 : 0;
        base_s = 0;
    }
    __syncthreads();
    if (tstart[size_t(r) * (T + 1) + T] > cap) return;
    const uint8_t* f = flags + size_t(r) * rays;
    for (int64_t start = 0; start < rays; start += 1024) {
        const int64_t e = start + tid;
        const bool pick = e < rays && f[e] != 0;
        const uint64_t m = __ballot(pick);
        const int before = __popcll(m & ((uint64_t(1) << lane) - 1));
        if (lane == 0) wave_total[wave] = __popcll(m);
        __syncthreads();
        int offset = base_s;
        for (int w = 0; w < wave; ++w) offset += wave_total[w];
        if (pick) list[size_t(r) * cap + offset + before] = int(e);
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int w = 0; w < 16; ++w) sum += wave_total[w];
            base_s += sum;
        }
        __syncthreads();
    }
}

// The extrema [V][S][2] of the differences' batch over the candidate rays of each variant's run
// (run = v / per_run): one wavefront per variant, the tuples holding candidates in order, their
// rays traced and reduced exactly as trace_kernel's extrema trace (trace_ray_ext; a wave with a
// ray that does not end Ok traces its valid rays again), then sqrt of the largest height^2 and
// the smallest cos I (extrema_finalize_kernel). Equal to the extrema over every ray wherever the
// extremal ray of each surface is a candidate.
template <bool kSoft>
__global__ __launch_bounds__(64) void candidate_extrema_kernel(TraceParams P, int per_run,
                                                               int cap, const int* tstart,
                                                               const int* list, double* ext) {
    __shared__ SurfLds L;
    const int v = blockIdx.x;
    const int run = v / per_run;
    const int lane = threadIdx.x;
    const int* ts = tstart + size_t(run) * (P.T + 1);
    const int* lst = list + size_t(run) * cap;
    double eh = 0.0, ec = 1.0;
    // an overflowed list was not written: neutral extrema (the check sends the chunk to the
    // whole trace)
    const int ntuples = ts[P.T] > cap ? 0 : P.T;
    for (int tt = 0; tt < ntuples; ++tt) {
        const int lo = ts[tt], hi = ts[tt + 1];
        if (lo == hi) continue; // uniform across the wave
        const Tuple t = tuple_of(P, v * P.T + tt);
        stage_surfaces(P, t, L);
        const size_t vf = size_t(t.v) * P.F + t.f;
        const double a = P.pupil_radius[t.v];
        double sin_t, cos_t;
        sincos(P.field_angle[vf], &sin_t, &cos_t);
        const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
        const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;
        for (int base = lo; base < hi; base += 64) {
            const int e = base + lane;
            const bool have = e < hi;
            Sample smp{};
            Ray r0{};
            if (have) {
                const int k = int(int64_t(lst[e]) - t.offset);
                smp = sample_of(P, t, k);
                r0 = source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]);
            }
            mkod::RayT<double> r = ray_as<double>(r0);
            Soft soft{0.0, MKO_STATUS_OK};
            double pend_h = 0.0, pend_c = 1.0;
            NoRecord rec;
            const uint8_t sti = trace_ray_ext<kSoft, kTraceFastMath>(
                P, L, r, M, kappa, soft, have, lane, pend_h, pend_c, false, rec);
            const uint64_t have_mask = __ballot(have);
            const uint64_t ok_mask = __ballot(have && sti == MKO_STATUS_OK);
            if (ok_mask == have_mask) {
                if (lane < P.S) {
                    eh = pend_h > eh ? pend_h : eh;
                    ec = pend_c < ec ? pend_c : ec;
                }
            } else {
                mkod::RayT<double> rr = ray_as<double>(r0);
                Soft s2{0.0, MKO_STATUS_OK};
                NoRecord nr;
                trace_ray_ext<kSoft, kTraceFastMath>(P, L, rr, M, kappa, s2,
                                                     have && sti == MKO_STATUS_OK, lane, eh, ec,
                                                     true, nr);
            }
        }
        __syncthreads(); // L is restaged for the next tuple
    }
    if (lane < P.S) {
        ext[(size_t(v) * P.S + lane) * 2] = __dsqrt_rn(eh);
        ext[(size_t(v) * P.S + lane) * 2 + 1] = ec;
    }
}

// Constraint rows of every variant (one thread per variant), rows[v * stride + offset + k] for
// the `count` items in row order; extrema [V][S][2] for the items that need them.
__global__ void constraint_rows_kernel(int V, int S, int F, const double* curvature,
                                       const double* thickness, const double* paraxial_index,
                                       const double* pupil_radius, const double* field_angle,
                                       const double* field_target, uint64_t sphere_mask,
                                       const mko_constraint* items, int count, int need_paraxial,
                                       const double* extrema, double* rows, int64_t stride,
                                       int64_t offset, const double* buchdahl,
                                       const double* glass_hull) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const mkoc::VariantRows vr{curvature + size_t(v) * S,
                               thickness + size_t(v) * S,
                               paraxial_index ? paraxial_index + size_t(v) * S : nullptr,
                               sphere_mask,
                               pupil_radius[v],
                               S,
                               field_angle + size_t(v) * F,
                               field_target + size_t(v) * F,
                               F,
                               buchdahl ? buchdahl + size_t(v) * S * 4 : nullptr,
                               glass_hull};
    mkoc::Paraxial p{1.0, 0.0, 0.0, 0.0, false, false, 0.0, 0.0, 0.0};
    if (need_paraxial) p = mkoc::paraxial(vr.c, vr.t, vr.n, S);
    const double* ext = extrema ? extrema + size_t(v) * S * 2 : nullptr;
    double* out = rows + size_t(v) * stride + offset;
    for (int k = 0; k < count; ++k) out[k] = mkoc::row(items[k], vr, p, ext);
}

// Normal equations of the constraint rows added to a run's A [P][P] and g [P] (the fused
// mixed-precision path forms no J): A(p, q) += sum_e Je(p, e) Je(q, e), g(p) += sum_e Je(p, e)
// re(e), one workgroup per run, Je [runs][P][E], re [runs][E].
__global__ void extra_normal_kernel(int P, int E, const double* Je, const double* re, double* a,
                                    double* g) {
    const int run = blockIdx.x;
    const double* J = Je + size_t(run) * P * E;
    const double* r = re + size_t(run) * E;
    for (int i = threadIdx.x; i < P * (P + 1); i += blockDim.x) {
        const int p = i / (P + 1), q = i - p * (P + 1);
        const double* u = J + size_t(p) * E;
        const double* w = q < P ? J + size_t(q) * E : r;
        double acc = 0.0;
        for (int e = 0; e < E; ++e) acc = fma(u[e], w[e], acc);
        if (q < P) a[(size_t(run) * P + p) * P + q] += acc;
        else g[size_t(run) * P + p] += acc;
    }
}

// Sums of squares of the constraint rows per variant (extra_merit_values,
// src/optim/levenberg_marquardt.cpp:659; trace_extra_merit, src/optim/trace_extra_residuals.cpp:96):
// e[v] = (extra ? extra[v] + : ) sum_{k < E} r_k^2, t[v] = sum_{E <= k < E + Et} r_k^2, ascending.
__global__ void rows_merit_kernel(int V, const double* rows, int64_t stride, int64_t offset, int E,
                                  int Et, const double* extra, double* e_sum, double* t_sum) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double* r = rows + size_t(v) * stride + offset;
    double e = 0.0;
    for (int k = 0; k < E; ++k) e = e + r[k] * r[k];
    if (e_sum) e_sum[v] = extra ? extra[v] + e : e;
    double t = 0.0;
    for (int k = E; k < E + Et; ++k) t = t + r[k] * r[k];
    if (t_sum) t_sum[v] = t;
}

// merit[v] = sum over tuples in (f, l) order + extra[v] (merit_finalize_kernel, merit.cu:207)
__global__ void merit_finalize_kernel(int V, int T, const double* partial, const int* pfailed,
                                      const int* phard, const double* extra,
                                      const double* traced, double* merit, int64_t* failed,
                                      int64_t* hard, const int32_t* live) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || (live && v >= *live)) return;
    double s = 0.0;
    int64_t f = 0, h = 0;
    for (int t = 0; t < T; ++t) {
        s += partial[size_t(v) * T + t];
        f += pfailed[size_t(v) * T + t];
        h += phard[size_t(v) * T + t];
    }
    // (ray merit + sum e^2) + sum t^2 (CpuLmBackend::merit, levenberg_marquardt.cpp:866-888)
    double m = extra ? s + extra[v] : s;
    if (traced) m = m + traced[v];
    merit[v] = m;
    failed[v] = f;
    hard[v] = h;
}

// ---- FP32 prefilter (prefilter_kernel, src/cuda/prefilter.cu:87; per ray prefilter_ray,
// include/mikroptim/merit/prefilter_ray.hpp:171-249) ---------------------------------------------
// The only float path: its values select multistart variants and never enter a merit, residual
// or Jacobian. Each ray is traced in float (hard failure) and carries its boundary margin: the
// smallest distance of the tracer's checks from their failure boundary, as a cosine (d_z before a
// transfer, cos I = sqrt(D) at the sphere, cos I' = sqrt(k) after refraction, the reference
// sphere's discriminant and denominator); for a failed ray the size of the violated check.

struct PfLds {
    float c[MKO_MAX_SURFACES], t[MKO_MAX_SURFACES], n[MKO_MAX_SURFACES];
};

__device__ __forceinline__ float pf_min_margin(float acc, float x) {
    if (x < acc) return x;
    return x >= acc ? acc : 0.0f; // NaN: 0 (uncertain)
}
__device__ __forceinline__ float pf_fail_linear(float x) { return -x > 0.0f ? -x : 0.0f; }
__device__ __forceinline__ float pf_fail_root(float q) {
    return -q > 0.0f ? __builtin_sqrtf(-q) : 0.0f;
}

struct PfRay {
    uint8_t status;
    float opl, margin;
};

__device__ __forceinline__ PfRay pf_trace(const PfLds& L, int S, V3<float> p, V3<float> d,
                                          float opl, V3<float> M, int kind, float kappa) {
    using mkod::cross;
    using mkod::dot;
    PfRay out{MKO_STATUS_OK, 0.0f, 0.0f};
    float margin = 1.0f;
    for (int s = 1; s < S; ++s) {
        const float c = L.c[s], nb = L.n[s - 1], na = L.n[s];
        if (!(d.z > 0.0f)) {
            out.status = MKO_STATUS_BACKWARD;
            out.margin = pf_fail_linear(d.z);
            return out;
        }
        const V3<float> db = d;
        p.z -= L.t[s - 1];
        const float s0 = -p.z / d.z;
        p.x += s0 * d.x;
        p.y += s0 * d.y;
        p.z = 0.0f;
        const float H = c * (p.x * p.x + p.y * p.y);
        const float B = d.z - c * (p.x * d.x + p.y * d.y);
        const float D = B * B - c * H;
        const float denom = D < 0.0f ? 0.0f : B + __builtin_sqrtf(D);
        if (D < 0.0f || !(denom > 0.0f)) {
            const float a = pf_fail_root(D), b = pf_fail_linear(B);
            out.status = MKO_STATUS_MISSED;
            out.margin = a > b ? a : b;
            return out;
        }
        const float sh = H / denom;
        p = p + sh * d;
        opl += nb * (s0 + sh);
        const float mu = nb / na;
        const V3<float> N{-c * p.x, -c * p.y, 1.0f - c * p.z};
        const float cosI = dot(d, N);
        const float k = 1.0f - mu * mu * (1.0f - cosI * cosI);
        if (k < 0.0f) {
            out.status = MKO_STATUS_TIR;
            out.margin = pf_fail_root(k);
            return out;
        }
        d = mu * d + (__builtin_sqrtf(k) - mu * cosI) * N;
        margin = pf_min_margin(margin, db.z);
        margin = pf_min_margin(margin, dot(db, N));
        margin = pf_min_margin(margin, dot(d, N));
    }
    if (!(d.z > 0.0f)) {
        out.status = MKO_STATUS_BACKWARD;
        out.margin = pf_fail_linear(d.z);
        return out;
    }
    margin = pf_min_margin(margin, d.z);
    const float n_image = L.n[S - 1];
    if (kind != MKO_REF_LAST_SURFACE) {
        const V3<float> q = p - M;
        const V3<float> ci = cross(q, d);
        const float t = -dot(ci, ci);
        const float disc = 1.0f + kappa * kappa * t;
        const float b = -dot(q, d);
        float step = 0.0f;
        const bool ok = disc >= 0.0f;
        if (ok) step = b - kappa * t / (__builtin_sqrtf(disc) + 1.0f);
        if (!ok || (!(step >= 0.0f) && !(step < 0.0f))) {
            out.status = MKO_STATUS_MISSED;
            out.margin = disc < 0.0f ? pf_fail_root(disc) : 0.0f;
            return out;
        }
        out.opl = opl + n_image * step;
        out.margin = pf_min_margin(margin, __builtin_sqrtf(disc));
        return out;
    }
    const float r2 = dot(M, M);
    const float inv_r = r2 > 0.0f ? 1.0f / __builtin_sqrtf(r2) : 1.0f;
    const V3<float> q = p - M;
    const float bq = -dot(q, d);
    const float g = dot(p, p - 2.0f * M);
    const float disc = bq * bq - g;
    const float sigma = M.z >= 0.0f ? 1.0f : -1.0f;
    const float root = disc > 0.0f ? __builtin_sqrtf(disc) : 0.0f;
    const float den = sigma * (bq + sigma * root) * inv_r;
    bool ok = !(disc < 0.0f);
    float s_ref = 0.0f;
    if (ok) {
        const float rt = __builtin_sqrtf(disc);
        const bool real_image = M.z >= 0.0f;
        const float dn = real_image ? bq + rt : bq - rt;
        ok = real_image ? dn > 0.0f : dn < 0.0f;
        if (ok) s_ref = g / dn;
    }
    if (!ok) {
        const float a = pf_fail_root(disc * inv_r * inv_r);
        const float c = disc >= 0.0f ? pf_fail_linear(den) : 0.0f;
        out.status = MKO_STATUS_MISSED;
        out.margin = a > c ? a : c;
        return out;
    }
    margin = pf_min_margin(margin, root * inv_r);
    margin = pf_min_margin(margin, den);
    out.opl = opl + n_image * s_ref;
    out.margin = margin;
    return out;
}

__device__ __forceinline__ float block_max_f(float v, float* lds) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    float m = lds[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) m = lds[w] > m ? lds[w] : m;
    __syncthreads();
    return m;
}

// One workgroup per (variant, field, wavelength) tuple: the tuple's float surface table in LDS,
// OPL relative to the axial optical length (prefilter_opl_offset), classes 0 certain valid /
// 1 certain failed / 2, 3 uncertain (margin <= eta), the piston over the certain-valid rays and
// the tuple's sum of squared residuals; partials per tuple: merit (as double), largest optical
// length, counts [3] (certain failed, uncertain valid, uncertain failed).
template <int kRpl>
__global__ __launch_bounds__(kBlock) void prefilter_kernel(TraceParams P, float eta,
                                                           double* pmerit, float* plength,
                                                           int* pcounts) {
    __shared__ PfLds L;
    __shared__ float fred[kWaves];
    __shared__ int ired[kWaves];
    const Tuple t = tuple_of(P, blockIdx.x);
    const int tid = threadIdx.x;
    if (tid < P.S) {
        const size_t vs = size_t(t.v) * P.S + tid;
        L.c[tid] = ((P.sphere_mask >> tid) & 1ull) ? float(P.curvature[vs]) : 0.0f;
        L.t[tid] = float(P.thickness[vs]);
        L.n[tid] = float(P.index[vs * P.L + t.l]);
    }
    __syncthreads();
    float offset = 0.0f;
    for (int s = 0; s + 1 < P.S; ++s) offset += L.n[s] * L.t[s];
    const size_t vf = size_t(t.v) * P.F + t.f;
    const float a = float(P.pupil_radius[t.v]);
    const float th = float(P.field_angle[vf]);
    const float sth = sinf(th), cth = cosf(th);
    const V3<float> M{0.0f, float(P.field_target[vf]), L.t[P.S - 1]};
    const float kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? float(P.reference_curvature[vf]) : 0.0f;
    const float pi = float(kPi);
    float opl[kRpl], wts[kRpl];
    uint8_t cls[kRpl];
    float wok = 0.0f, wopl = 0.0f, optical = 0.0f;
    int nf = 0, nuo = 0, nuf = 0;
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const int k = tid + i * kBlock;
        cls[i] = 4;
        opl[i] = 0.0f;
        wts[i] = 0.0f;
        if (k >= t.n) continue;
        const int ri = k / t.M, j = k - ri * t.M;
        const float rho = __builtin_sqrtf(float(P.gl_u[ri]));
        const float w = float(P.gl_w[ri]) / float(t.M);
        const uint8_t mode = P.pupil_mode[t.f];
        const float phi = mode == MKO_PUPIL_HALF   ? float(2 * j + 1) * pi / float(2 * t.M)
                          : mode == MKO_PUPIL_FULL ? float(2 * j + 1) * pi / float(t.M)
                                                   : 0.0f;
        const float sp = sinf(phi), cp = cosf(phi);
        V3<float> p, d;
        float o;
        if (P.source == MKO_SOURCE_POINT) {
            const float sin_max = float(P.src_na) / L.n[0];
            const float sr = sin_max * rho;
            p = V3<float>{float(P.src_pt[0]), float(P.src_pt[1]), float(P.src_pt[2])};
            d = V3<float>{sr * sp, sr * cp, __builtin_sqrtf((1.0f - sr) * (1.0f + sr))};
            o = 0.0f;
        } else {
            const float r = a * rho;
            p = V3<float>{r * sp, r * cp, 0.0f};
            d = V3<float>{0.0f, sth, cth};
            o = mkod::dot(p, d);
        }
        const PfRay rr = pf_trace(L, P.S, p, d, o - offset, M, P.ref_kind, kappa);
        const bool ok = rr.status == MKO_STATUS_OK;
        const bool certain = rr.margin > eta;
        const uint8_t c = uint8_t(ok ? (certain ? 0 : 2) : (certain ? 1 : 3));
        cls[i] = c;
        opl[i] = rr.opl;
        wts[i] = w;
        if (c == 0) {
            wok += w;
            wopl += w * rr.opl;
            const float len = fabsf(offset) + fabsf(rr.opl);
            optical = len > optical ? len : optical;
        }
        nf += c == 1;
        nuo += c == 2;
        nuf += c == 3;
    }
    const float w_ok = block_sum(wok, fred);
    const float sum_o = block_sum(wopl, fred);
    const bool valid = w_ok > 0.0f;
    const float piston = valid ? sum_o / w_ok : 0.0f;
    const float scale = float(t.scale);
    float r2 = 0.0f;
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const float e = (valid && cls[i] == 0) ? opl[i] - piston : 0.0f;
        const float r = scale * __builtin_sqrtf(wts[i]) * e;
        r2 += r * r;
    }
    const float tr2 = block_sum(r2, fred);
    const float tl = block_max_f(optical, fred);
    const int tf = block_sum(nf, ired);
    const int tuo = block_sum(nuo, ired);
    const int tuf = block_sum(nuf, ired);
    if (tid == 0) {
        pmerit[blockIdx.x] = double(tr2);
        plength[blockIdx.x] = tl;
        pcounts[size_t(blockIdx.x) * 3] = tf;
        pcounts[size_t(blockIdx.x) * 3 + 1] = tuo;
        pcounts[size_t(blockIdx.x) * 3 + 2] = tuf;
    }
}

// Per variant: merit = sum of the tuple partials (double, tuple order), optical length = max,
// counts summed (the second stage of the reference's prefilter).
__global__ void prefilter_finalize_kernel(int V, int T, const double* pmerit,
                                          const float* plength, const int* pcounts,
                                          double* merit, double* optical, int64_t* failed,
                                          int64_t* uncertain_ok, int64_t* uncertain_failed) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double m = 0.0;
    float len = 0.0f;
    int64_t f = 0, uo = 0, uf = 0;
    for (int t = 0; t < T; ++t) {
        const size_t i = size_t(v) * T + t;
        m += pmerit[i];
        len = plength[i] > len ? plength[i] : len;
        f += pcounts[i * 3];
        uo += pcounts[i * 3 + 1];
        uf += pcounts[i * 3 + 2];
    }
    merit[v] = m;
    optical[v] = double(len);
    failed[v] = f;
    uncertain_ok[v] = uo;
    uncertain_failed[v] = uf;
}

// ---- FD batch and FD difference -------------------------------------------------------------

// Builds the 2P+1 FD variants of each run from its template variant: variant 0 = x0,
// 2k+1 = x0 + d+_k e_k, 2k+2 = x0 + d-_k e_k (write_fd_variants, variant_generator.cpp:240).
// The template's per-variant arrays are copied, then the variables written.
struct BatchPtrs {
    double* curvature;
    double* thickness;
    double* index;
    double* pupil_radius;
    double* field_angle;
    double* field_target;
    double* reference_curvature;
    double* paraxial_index;
    double* buchdahl;
};
struct ConstBatchPtrs {
    const double* curvature;
    const double* thickness;
    const double* index;
    const double* pupil_radius;
    const double* field_angle;
    const double* field_target;
    const double* reference_curvature;
    const double* paraxial_index;
    const double* buchdahl;
};

// Buchdahl coordinate w of each wavelength and (at [L]) of lambda_ref, computed on the host from
// the batch's wavelengths (valid = 0: not available).
struct OmegaTable {
    double w[kMaxWavelengths + 1];
    int valid;
};

__global__ void point_batch_kernel(ConstBatchPtrs src, BatchPtrs dst, int runs, int per_run,
                                   const int32_t* template_index, int template_per_run, int S,
                                   int L, int F, int P, const int* var_type,
                                   const int* var_surface, const int* var_component,
                                   const double* x0, const double* plus, const double* minus,
                                   const int32_t* live, OmegaTable om) {
    const int vv = blockIdx.x; // destination variant
    const int run = vv / per_run;
    const int j = vv - run * per_run;
    if (run >= runs || (live && run >= *live)) return;
    const int from = template_index ? template_index[run] : template_per_run ? run : 0;
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        dst.curvature[size_t(vv) * S + i] = src.curvature[size_t(from) * S + i];
        dst.thickness[size_t(vv) * S + i] = src.thickness[size_t(from) * S + i];
        if (src.paraxial_index && dst.paraxial_index)
            dst.paraxial_index[size_t(vv) * S + i] = src.paraxial_index[size_t(from) * S + i];
    }
    for (int i = threadIdx.x; i < S * L; i += blockDim.x)
        dst.index[size_t(vv) * S * L + i] = src.index[size_t(from) * S * L + i];
    for (int i = threadIdx.x; i < F; i += blockDim.x) {
        dst.field_angle[size_t(vv) * F + i] = src.field_angle[size_t(from) * F + i];
        dst.field_target[size_t(vv) * F + i] = src.field_target[size_t(from) * F + i];
        if (src.reference_curvature)
            dst.reference_curvature[size_t(vv) * F + i] = src.reference_curvature[size_t(from) * F + i];
    }
    if (threadIdx.x == 0) dst.pupil_radius[vv] = src.pupil_radius[from];
    if (src.buchdahl && dst.buchdahl)
        for (int i = threadIdx.x; i < 4 * S; i += blockDim.x)
            dst.buchdahl[size_t(vv) * 4 * S + i] = src.buchdahl[size_t(from) * 4 * S + i];
    __syncthreads();
    // variables: x0 of the run, component k replaced for FD variants
    const int kk = j == 0 ? -1 : (j - 1) / 2;
    bool glass = false;
    for (int k = 0; k < P; ++k) glass = glass || var_type[k] == MKO_VAR_GLASS;
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
        double x = x0[size_t(run) * P + k];
        if (k == kk) x = (j & 1) ? plus[size_t(run) * P + k] : minus[size_t(run) * P + k];
        const int s = var_surface[k];
        if (var_type[k] == MKO_VAR_CURVATURE)
            dst.curvature[size_t(vv) * S + s] = x;
        else if (var_type[k] == MKO_VAR_GLASS) {
            if (dst.buchdahl && var_component)
                dst.buchdahl[(size_t(vv) * S + s) * 4 + var_component[k]] = x;
        } else
            dst.thickness[size_t(vv) * S + s] = x;
    }
    if (!glass) return;
    __syncthreads();
    // index of the glass-variable surfaces from their Buchdahl coefficients
    // (VariableSet::apply, variable_set.cpp:271-284; buchdahl_index, dispersion.hpp:93);
    // without coefficients or wavelengths the index becomes NaN (the trace then fails loudly)
    for (int k = threadIdx.x; k < P; k += blockDim.x) {
        if (var_type[k] != MKO_VAR_GLASS || !var_component || var_component[k] != 0) continue;
        const int s = var_surface[k];
        const bool ok = om.valid && dst.buchdahl != nullptr;
        const double* c = ok ? dst.buchdahl + (size_t(vv) * S + s) * 4 : nullptr;
        for (int l = 0; l <= L; ++l) {
#pragma clang fp contract(off)
            double n = __builtin_nan("");
            if (ok) {
                const double w = om.w[l];
                n = c[0] + w * (c[1] + w * (c[2] + w * c[3]));
            }
            if (l < L) dst.index[(size_t(vv) * S + s) * L + l] = n;
            else if (dst.paraxial_index) dst.paraxial_index[size_t(vv) * S + s] = n;
        }
    }
}

// J[i][k][r] = fd_derivative (fd_formula.hpp:39) from the FD residuals; block (i, P) copies r0.
// Status changes per (i, k): rays whose Ok-ness differs between x0 and either point.
__global__ __launch_bounds__(kBlock) void fd_difference_kernel(
    const double* res, int64_t stride, const uint8_t* st, int64_t rays, int runs, int P,
    int64_t rows, const double* step_plus, const double* step_minus, const uint8_t* stencil,
    double* J, double* r0, int64_t* changes, int64_t jstride, int64_t joffset) {
    __shared__ int ired[kWaves];
    const int i = blockIdx.x / (P + 1);
    const int k = blockIdx.x - i * (P + 1);
    const int n = 2 * P + 1;
    const size_t v0 = size_t(i) * n;
    const double* r_0 = res + v0 * stride;
    if (k == P) {
        for (int64_t r = threadIdx.x; r < rows; r += kBlock)
            r0[size_t(i) * jstride + joffset + r] = r_0[r];
        return;
    }
    const double* ra = res + (v0 + 2 * k + 1) * stride;
    const double* rb = res + (v0 + 2 * k + 2) * stride;
    const double da = step_plus[size_t(i) * P + k];
    const double db = step_minus[size_t(i) * P + k];
    const uint8_t sten = stencil[size_t(i) * P + k];
    double* col = J + (size_t(i) * P + k) * jstride + joffset;
    for (int64_t r = threadIdx.x; r < rows; r += kBlock) {
        const double a0 = r_0[r], aa = ra[r], ab = rb[r];
        double v;
        if (sten == MKO_STENCIL_ONE_SIDED) {
            const double ga = aa - a0, gb = ab - a0;
            v = (db * db * ga - da * da * gb) / (da * db * (db - da));
        } else {
            v = (aa - ab) / (da - db);
        }
        col[r] = v;
    }
    if (!st) return;
    int c = 0;
    const uint8_t* s0 = st + v0 * rays;
    const uint8_t* sa = st + (v0 + 2 * k + 1) * rays;
    const uint8_t* sb = st + (v0 + 2 * k + 2) * rays;
    for (int64_t r = threadIdx.x; r < rays; r += kBlock) {
        const bool ok0 = s0[r] == MKO_STATUS_OK;
        c += ((sa[r] == MKO_STATUS_OK) != ok0) || ((sb[r] == MKO_STATUS_OK) != ok0);
    }
    const int total = block_sum(c, ired);
    if (threadIdx.x == 0) changes[size_t(i) * P + k] = total;
}

// ---- ADJOINT Jacobian -----------------------------------------------------------------------

// Writes one seed's per-ray parameter adjoints into J rows: value * factor at column p, row.
template <class TJ>
struct RowSink {
    TJ* col0;     // J of the variant
    int64_t rows; // R
    int64_t row;
    double factor;
    __device__ __forceinline__ void write(int p, double value) const {
        if (p >= 0) col0[int64_t(p) * rows + row] = TJ(factor * value);
    }
};

// backward_ray (adjoint_backward.hpp:496-580) over the record, reverse arithmetic in RB.
template <class RB, class RR, class Sink>
__device__ __forceinline__ void backward_ray(const TraceParams& P, const SurfLds& L,
                                             const ParamMap& map, const Record<RR>& rec,
                                             bool reached_image, V3<double> M, double kappa,
                                             RB seed_opl, RB seed_pen, const Sink& sink, int l) {
    const int S = P.S;
    const int count = rec.count;
    mkod::RayAdj<RB> adj{{RB(0), RB(0), RB(0)}, {RB(0), RB(0), RB(0)}};
    uint8_t next_branch = 0;
    RB n_carry = RB(0); // adjoint of index[j] from step j + 1 (or the reference sphere)
    // JacobianRowWriter::index: the glass variables of the medium after surface j
    const auto index_sink = [&](int j, RB value) {
        if (!map.glass) return;
        const double w = map.omega[l];
        const double slope[3] = {1.0, w, w * w};
#pragma unroll
        for (int c = 0; c < 3; ++c) sink.write(map.glass_col[j][c], double(value) * slope[c]);
    };
    for (int s = count - 1; s >= 0; --s) {
        const V3<RB> p{RB(rec.p[s][0]), RB(rec.p[s][1]), RB(rec.p[s][2])};
        const V3<RB> d{RB(rec.d[s][0]), RB(rec.d[s][1]), RB(rec.d[s][2])};
        if (s == count - 1) {
            if (reached_image && count == S) {
                RB nib = RB(0);
                const RB mz = mkod::reference_step_backward<RB>(
                    p, d, V3<RB>{RB(M.x), RB(M.y), RB(M.z)}, RB(kappa), P.ref_kind,
                    RB(L.n[S - 1]), rec.image_branch, seed_opl, seed_pen, adj, nib);
                sink.write(map.thick_col[S - 1], double(mz));
                n_carry = nib;
            }
        } else {
            const int j = s + 1;
            RB tb = RB(0), cb = RB(0), nbb = RB(0), nab = RB(0);
            mkod::refracting_step_backward<RB>(p, d, RB(L.t[s]), RB(L.c[j]), RB(L.n[s]),
                                               RB(L.n[j]), next_branch, seed_opl, seed_pen, adj,
                                               tb, cb, nbb, nab);
            if ((P.sphere_mask >> j) & 1ull) sink.write(map.curv_col[j], double(cb));
            sink.write(map.thick_col[s], double(tb));
            index_sink(j, nab + n_carry);
            n_carry = nbb;
        }
        next_branch = rec.branch[s];
    }
}

// One workgroup per (variant = run, tuple). Raw per-ray gradients go straight into the
// column-major J of the variant (J[v][p][row]); after the tuple's piston reduction the
// workgroup projects its rows in place: J(k,p) = s sqrt(W_k) (g_k - gbar_p) for rays in the
// piston, 0 otherwise (adjoint_jacobian.cpp:462-479). Penalty rows (soft mode) get
// pscale sqrt(W_k) times the penalty gradient. r0 rows and the status are written as by
// RESIDUALS.
template <bool kSoft, int kRpl, class RB, class RR, class TJ>
__global__ __launch_bounds__(kBlock) void adjoint_kernel(TraceParams P, ParamMap map, TJ* J,
                                                         double* r0, uint8_t* status) {
    __shared__ SurfLds L;
    __shared__ SampleTable ST;
    __shared__ double red[kWaves];
    if (P.live && int(blockIdx.x) / P.T >= *P.live) return;
    const Tuple t = tuple_of(P, blockIdx.x);
    stage_samples(P, t, ST);
    stage_surfaces(P, t, L);
    const size_t vf = size_t(t.v) * P.F + t.f;
    const double a = P.pupil_radius[t.v];
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
    const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;
    const int64_t R = P.stride;
    TJ* col0 = J + size_t(t.v) * map.P * R;

    double opl[kRpl];
    uint8_t st[kRpl];
    double pen[kRpl];
    double sw = 0.0, swo = 0.0;
    Record<RR> rec;
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const int k = threadIdx.x + i * kBlock;
        st[i] = MKO_STATUS_UNSUPPORTED;
        opl[i] = 0.0;
        pen[i] = 0.0;
        if (k >= t.n) continue;
        const int64_t row = t.offset + k;
        for (int p = 0; p < map.P; ++p) {
            col0[int64_t(p) * R + row] = TJ(0);
            if (kSoft) col0[int64_t(p) * R + P.rays + row] = TJ(0);
        }
        const Sample s = sample_at(P, t, ST, k);
        Ray r = source_ray(P, t, s, a, sin_t, cos_t, L.n[0]);
        Soft soft{0.0, MKO_STATUS_OK};
        rec.count = 0;
        rec.image_branch = 0;
        st[i] = trace_ray<kSoft>(P, L, r, M, kappa, soft, rec);
        opl[i] = r.opl;
        pen[i] = soft.penalty;
        const bool reached = mkod::in_piston(st[i]);
        if (reached) {
            sw += s.weight;
            swo += s.weight * r.opl;
            backward_ray<RB, RR>(P, L, map, rec, true, M, kappa, RB(1), RB(0),
                                     RowSink<TJ>{col0, R, row, 1.0}, t.l);
        }
        if (kSoft && soft.first != MKO_STATUS_OK) {
            backward_ray<RB, RR>(P, L, map, rec, reached, M, kappa, RB(0), RB(1),
                                     RowSink<TJ>{col0, R, P.rays + row, t.pscale * s.rootw}, t.l);
        }
    }
    const double w_ok = block_sum(sw, red);
    const double sum = block_sum(swo, red);
    const bool valid = w_ok > 0.0;
    const double piston = valid ? sum / w_ok : 0.0;
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
        const int k = threadIdx.x + i * kBlock;
        if (k >= t.n) continue;
        const Sample s = sample_at(P, t, ST, k);
        const bool ok = valid && mkod::in_piston(st[i]);
        const double e = ok ? opl[i] - piston : 0.0;
        const int64_t row = t.offset + k;
        r0[size_t(t.v) * R + row] = t.scale * s.rootw * e;
        if (kSoft) {
            const double fp = st[i] == MKO_STATUS_OK ? 0.0
                              : mkod::is_soft(st[i]) ? pen[i]
                                                     : pen[i] + 1.0;
            r0[size_t(t.v) * R + P.rays + row] = t.pscale * s.rootw * fp;
        }
        if (status) status[size_t(t.v) * P.rays + row] = st[i];
    }
    __threadfence_block();
    __syncthreads();
    // piston projection, one column at a time (sums in FP64, fixed order)
    for (int p = 0; p < map.P; ++p) {
        TJ* col = col0 + int64_t(p) * R + t.offset;
        double part = 0.0;
#pragma unroll
        for (int i = 0; i < kRpl; ++i) {
            const int k = threadIdx.x + i * kBlock;
            if (k < t.n && mkod::in_piston(st[i])) part += sample_at(P, t, ST, k).weight * double(col[k]);
        }
        const double mean = valid ? block_sum(part, red) / w_ok : 0.0;
#pragma unroll
        for (int i = 0; i < kRpl; ++i) {
            const int k = threadIdx.x + i * kBlock;
            if (k >= t.n) continue;
            const Sample s = sample_at(P, t, ST, k);
            const bool ok = valid && mkod::in_piston(st[i]);
            col[k] = ok ? TJ(t.scale * s.rootw * (double(col[k]) - mean)) : TJ(0);
        }
    }
}

// ---- vector-Jacobian product J^T v without J ---------------------------------------------

// Sink of the VJP sweep: lane `lane` adds each column's value into its own slot of acc [P][64]
// (a column is written once per ray and sweep; the lane's rays follow one another).
struct AccSink {
    double* acc;
    int lane;
    __device__ __forceinline__ void write(int p, double value) const {
        if (p >= 0) acc[p * 64 + lane] += value;
    }
};

// J^T v of one (variant = run, tuple), one wave per tuple, the tuple's rays 64 at a time. The
// reverse sweep is linear in its seeds (ray_gradient_rows, ray_jacobian.hpp:93), so each ray
// runs it once with
//   seed_opl = s (v_k sqrt(W_k) - W_k m),  m = sum_j v_j sqrt(W_j) / W_ok (rays in the piston),
//   seed_pen = pscale sqrt(W_k) u_k      (soft mode, rays with a failure),
// v and u being the cotangents of the ray's residual and penalty rows: sum_k v_k J(k, p) with
// J(k, p) = s sqrt(W_k) (g_kp - gbar_p) equals sum_k g_kp s (v_k sqrt(W_k) - W_k m), so neither
// J nor the piston-projected columns are formed. A first pass traces the tuple for W_ok and the
// sum behind m. part[tuple][p]: the tuple's share, lanes summed in a fixed tree.
template <bool kSoft>
__global__ __launch_bounds__(64) void vjp_kernel(TraceParams P, ParamMap map, const double* grad,
                                                 int64_t gstride, double* part) {
    extern __shared__ double acc[]; // [P][64]
    __shared__ SurfLds L;
    const Tuple t = tuple_of(P, blockIdx.x);
    stage_surfaces(P, t, L);
    const int lane = threadIdx.x;
    for (int p = 0; p < map.P; ++p) acc[p * 64 + lane] = 0.0;
    const size_t vf = size_t(t.v) * P.F + t.f;
    const double a = P.pupil_radius[t.v];
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    const V3<double> M{0.0, P.field_target[vf], L.t[P.S - 1]};
    const double kappa = P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0;
    const double* gv = grad + size_t(t.v) * size_t(gstride);
    double sw = 0.0, svw = 0.0;
    for (int k = lane; k < t.n; k += 64) {
        const Sample smp = sample_of(P, t, k);
        Ray r = source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]);
        Soft soft{0.0, MKO_STATUS_OK};
        NoRecord nr;
        const uint8_t st = trace_ray<kSoft>(P, L, r, M, kappa, soft, nr);
        if (mkod::in_piston(st)) {
            sw += smp.weight;
            svw += gv[t.offset + k] * smp.rootw;
        }
    }
    const double W = wave_sum(sw), SV = wave_sum(svw);
    const bool valid = W > 0.0;
    const double m = valid ? SV / W : 0.0;
    Record<double> rec;
    for (int k = lane; k < t.n; k += 64) {
        const Sample smp = sample_of(P, t, k);
        Ray r = source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]);
        Soft soft{0.0, MKO_STATUS_OK};
        rec.count = 0;
        rec.image_branch = 0;
        const uint8_t st = trace_ray<kSoft>(P, L, r, M, kappa, soft, rec);
        const bool reached = mkod::in_piston(st);
        const double beta =
            valid && reached ? t.scale * (gv[t.offset + k] * smp.rootw - smp.weight * m) : 0.0;
        double gamma = 0.0;
        if (kSoft && soft.first != MKO_STATUS_OK)
            gamma = t.pscale * smp.rootw * gv[P.rays + t.offset + k];
        if (beta != 0.0 || gamma != 0.0)
            backward_ray<double, double>(P, L, map, rec, reached, M, kappa, beta, gamma,
                                         AccSink{acc, lane}, t.l);
    }
    for (int p = 0; p < map.P; ++p) {
        const double total = wave_sum(acc[p * 64 + lane]);
        if (lane == 0) part[size_t(blockIdx.x) * map.P + p] = total;
    }
}

// out[v][p] = sum over the run's tuples of part, in (field, wavelength) order.
__global__ void vjp_finalize_kernel(int V, int T, int Pn, const double* part, double* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(V) * Pn) return;
    const int64_t v = i / Pn, p = i - v * Pn;
    double x = 0.0;
    for (int t = 0; t < T; ++t) x += part[(v * T + t) * Pn + p];
    out[i] = x;
}

// ---- ADJOINT fused with the normal equations (hard failure mode) ---------------------------

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(16))) float f32x16;
constexpr int kMaxVariablesFused = 96;
constexpr int kFusedSlots = 4; // S / Tr slots per tuple in the workspace, one per wave (<= 4)
#ifndef MKO_FUSED_WAVES_PER_EU
#define MKO_FUSED_WAVES_PER_EU 3
#endif

// Per-(run, tuple) partial of the normal equations. With a per-tuple shift c (near the tuple's
// weighted mean gradient), o (near its weighted mean OPL) and u_k = sqrt(W_k) (g_k - c) over the
// rays in the piston:
//   C  = sum_k u_k u_k^T                  (P x P, MFMA, float)
//   S  = sum_k W_k (g_k - c)              (P, double)
//   Tr = sum_k W_k (g_k - c) (OPL_k - o)  (P, double)
// the tuple adds s^2 (C - S S^T / W_ok) to J^T J and s^2 (Tr - (piston - o) S) to J^T r0, because
// J(k,p) = s sqrt(W_k) (g_kp - gbar_p), r_k = s sqrt(W_k) (OPL_k - piston) and
// sum_k W_k (OPL_k - piston) = 0 (adjoint_jacobian.cpp:462-479 for J, normal_equations.cpp for
// the products).
struct TupleNormal {
    double w_ok, piston, s2;
};

// Gradient block of one chunk of rays in LDS: P rows of 64 * kNW rays, 4 bytes per ray (a float
// during the sweep, then bf16 hi and lo). Rays sit in groups of 8 (32 bytes); the group's place in
// row p is XOR-swizzled by p, so that the MFMA operand loads (32 rows at one ray group, 16 bytes
// per lane) hit distinct banks. A group's floats and its bf16 values share the same 32 bytes (hi
// in one half, lo in the other, the order alternating with bit 3 of the swizzled slot for the
// same reason), so the 8 lanes of a group -- one wave -- convert in place without a barrier.
template <int kNW>
struct GradLayout {
    static constexpr int kRow = 64 * kNW;
    static constexpr int kRowBytes = 4 * kRow;
    static constexpr int kMask = kRow / 8 >= 16 ? 15 : kRow / 8 - 1;
    __device__ static __forceinline__ int slot(int group, int p) { return group ^ (p & kMask); }
    __device__ static __forceinline__ int float_off(int p, int k) {
        return p * kRowBytes + ((4 * k) ^ ((p & kMask) << 5));
    }
    __device__ static __forceinline__ int hi_off(int p, int s) {
        return p * kRowBytes + 32 * s + (((s >> 3) & 1) << 4);
    }
    __device__ static __forceinline__ int lo_off(int p, int s) {
        return p * kRowBytes + 32 * s + ((((s >> 3) & 1) ^ 1) << 4);
    }
};


__device__ __forceinline__ double first_lane(double x) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// Gradient sink of the fused kernel. Every column is written at most once per ray and, in hard
// mode, at the same sweep step by every lane whose ray reached the image, so the lanes of a wave
// write column p together: the value of the wave's first such lane (pilot) goes to pilot[p] and
// each lane stores its difference to it in float. The float then carries the variation of the
// gradient over the pupil, not its (often much larger) common part.
template <int kNW>
struct SwzSink {
    unsigned char* G; // gradient block
    double* pilot;    // [P] of this wave
    int k;            // ray of the chunk (thread)
    bool leader;      // first lane of the wave with a reached ray
    __device__ __forceinline__ void write(int p, double value) const {
        if (p < 0) return;
        const double c = first_lane(value);
        *reinterpret_cast<float*>(G + GradLayout<kNW>::float_off(p, k)) = float(value - c);
        if (leader) pilot[p] = c;
    }
};

// backward_ray (adjoint_backward.hpp:496-580) for a ray that reached the image in hard mode, with
// the states on surfaces S-2 ... 1 reconstructed from the last state by inverse steps and the
// state on surface 0 regenerated from the source ray (no per-ray record in memory). kGlass: the
// index adjoints as well, into the glass variables' columns of each medium (n_carry as in
// backward_ray; dn/d(n0, nu1, nu2) = 1, w_l, w_l^2).
template <bool kGlass, class Sink>
__device__ __forceinline__ void backward_fused(const TraceParams& P, const SurfLds& L,
                                               const ParamMap& map, const LastState& last,
                                               const Ray& start, V3<double> M, double kappa,
                                               const Sink& sink, int l) {
    const int S = P.S;
    mkod::RayAdj<double> adj{{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double n_carry = 0.0;
    const auto index_sink = [&](int j, double value) {
        const double w = map.omega[l];
        sink.write(map.glass_col[j][0], value);
        sink.write(map.glass_col[j][1], value * w);
        sink.write(map.glass_col[j][2], value * (w * w));
    };
    {
        double nib = 0.0;
        const double mz = mkod::reference_step_backward<double>(last.p, last.d, M, kappa,
                                                                P.ref_kind, L.n[S - 1],
                                                                last.image_branch, 1.0, 0.0, adj,
                                                                nib);
        sink.write(map.thick_col[S - 1], mz);
        n_carry = nib;
    }
    V3<double> q = last.p, d1 = last.d;
    for (int s = S - 2; s >= 1; --s) {
        const int j = s + 1;
        double tb = 0.0, cb = 0.0, nbb = 0.0, nab = 0.0;
        const mkod::RayState<double> st = mkod::fused_step_backward<kGlass>(
            q, d1, L.c[j], L.c[s], L.t[s], L.n[s], L.mu[j], L.imu[j], adj, tb, cb, &nbb, &nab);
        if ((P.sphere_mask >> j) & 1ull) sink.write(map.curv_col[j], cb);
        sink.write(map.thick_col[s], tb);
        if constexpr (kGlass) {
            index_sink(j, nab + n_carry);
            n_carry = nbb;
        }
        q = st.p;
        d1 = st.d;
    }
    // surface 1 from the regenerated source ray (exact state on surface 0)
    {
        double tb = 0.0, cb = 0.0, nbb = 0.0, nab = 0.0;
        mkod::refracting_step_backward<double>(start.p, start.d, L.t[0], L.c[1], L.n[0], L.n[1],
                                               mkod::kBranchRegular, 1.0, 0.0, adj, tb, cb, nbb,
                                               nab);
        if ((P.sphere_mask >> 1) & 1ull) sink.write(map.curv_col[1], cb);
        sink.write(map.thick_col[0], tb);
        if constexpr (kGlass) index_sink(1, nab + n_carry);
    }
}

// Lane k's value of x (k wave-uniform).
__device__ __forceinline__ double lane_value(double x, int k) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), k);
    return __hiloint2double(hi, lo);
}

// Row sums of one wave's 64 rays in the gradient block, transposed: lane p returns
// A[h] = sum_k W_k G[p'][k] and B[h] = sum_k V_k G[p'][k] for row p' = p + 64 h (h < kRows), G the
// stored floats of ray k of the wave, W_k and V_k lane k's values, summed in ray order (one FMA
// chain per row). Each lane reads its row's 64 floats as 16-byte words; rows >= P give 0.
template <int kNW, int kRows>
__device__ __forceinline__ void wave_row_sums(const unsigned char* G, int Pn, int wave, int lane,
                                              double wk, double vk, double (&A)[2],
                                              double (&B)[2]) {
    using Lay = GradLayout<kNW>;
    A[0] = A[1] = B[0] = B[1] = 0.0;
#pragma unroll 1
    for (int j = 0; j < 8; ++j) {
        float gv[kRows][8];
#pragma unroll
        for (int h = 0; h < kRows; ++h) {
            const int p = lane + 64 * h;
            float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f), y = x;
            if (p < Pn) {
                const unsigned char* at =
                    G + size_t(p) * Lay::kRowBytes + 32 * ((8 * wave + j) ^ (p & Lay::kMask));
                x = *reinterpret_cast<const float4*>(at);
                y = *reinterpret_cast<const float4*>(at + 16);
            }
            gv[h][0] = x.x; gv[h][1] = x.y; gv[h][2] = x.z; gv[h][3] = x.w;
            gv[h][4] = y.x; gv[h][5] = y.y; gv[h][6] = y.z; gv[h][7] = y.w;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double w = lane_value(wk, 8 * j + i), v = lane_value(vk, 8 * j + i);
#pragma unroll
            for (int h = 0; h < kRows; ++h) {
                const double g = double(gv[h][i]);
                A[h] = __fma_rn(w, g, A[h]);
                B[h] = __fma_rn(v, g, B[h]);
            }
        }
    }
}

template <int kNW, class T>
__device__ __forceinline__ T block_sum_n(T v, T* lds) {
    v = wave_sum(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    T total = lds[0];
#pragma unroll
    for (int w = 1; w < kNW; ++w) total += lds[w];
    __syncthreads();
    return total;
}

// C tile t of the upper triangle of a tg x tg grid of 32 x 32 tiles, row-major: (I, J), I <= J.
__device__ __forceinline__ void tile_of(int t, int tg, int& I, int& J) {
    I = 0;
    while (t >= tg - I) {
        t -= tg - I;
        ++I;
    }
    J = I + t;
}

// C of one 32 x 32 tile over the chunk's rays: operand lane = (m = lane % 32, kb = lane / 32),
// 8 rays per operand, hi*hi + hi*lo + lo*hi; on the diagonal the B operands are the A operands.
template <int kNW>
__device__ __forceinline__ void tile_products(const unsigned char* G, int Pn, int I, int J,
                                              int lane, f32x16& acc) {
    using Lay = GradLayout<kNW>;
    const int m = lane & 31, kb = lane >> 5;
    const int pa = I * 32 + m, pb = J * 32 + m;
    // lanes of rows >= P (the padding of the tile) load nothing and take zeros
    const bool va = pa < Pn, vb = pb < Pn;
    bf16x8 zero;
#pragma unroll
    for (int e = 0; e < 8; ++e) zero[e] = __builtin_bit_cast(__bf16, uint16_t(0));
#pragma unroll 4
    for (int step = 0; step < Lay::kRow / 16; ++step) {
        const int grp = 2 * step + kb;
        bf16x8 ahi = zero, alo = zero;
        if (va) {
            const int s = Lay::slot(grp, pa);
            ahi = *reinterpret_cast<const bf16x8*>(G + Lay::hi_off(pa, s));
            alo = *reinterpret_cast<const bf16x8*>(G + Lay::lo_off(pa, s));
        }
        bf16x8 bhi = ahi, blo = alo;
        if (I != J) {
            bhi = zero;
            blo = zero;
            if (vb) {
                const int s = Lay::slot(grp, pb);
                bhi = *reinterpret_cast<const bf16x8*>(G + Lay::hi_off(pb, s));
                blo = *reinterpret_cast<const bf16x8*>(G + Lay::lo_off(pb, s));
            }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi, acc, 0, 0, 0);
    }
}

// C tile out: lane holds column n = lane % 32 of rows 8 (e / 4) + 4 (lane / 32) + e % 4.
__device__ __forceinline__ void store_tile(float* C, int Pn, int I, int J, int lane,
                                           const f32x16& acc) {
    const int col = J * 32 + (lane & 31);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int row = I * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        if (row < Pn && col < Pn) C[size_t(row) * Pn + col] = acc[e];
    }
}

// Dynamic LDS of adjoint_normal_kernel<kNW> for P variables: the gradient block, then the
// waves' pilots [kNW][P] and the shift [P] (doubles).
template <int kNW>
constexpr size_t fused_lds_bytes(int Pn) {
    return size_t(Pn) * GradLayout<kNW>::kRowBytes + size_t(kNW + 1) * Pn * sizeof(double);
}

// One workgroup of kNW waves per (run, tuple); the tuple's rays in chunks of 64 kNW, one ray per
// thread. Per chunk: the forward trace in double (or, kStates, the last state and OPL of each ray
// from a MERIT trace of the same point, mko_internal::RayStates), the reverse sweep over
// reconstructed states into the LDS gradient block (pilot-relative floats); S and Tr of each
// wave's rays in FP64 by transposed wave sums (per-lane accumulators, summed over the waves in
// index order at the end); the block rewritten in place as bf16 hi / lo; C on MFMA
// (v_mfma_f32_32x32x16_bf16, 3 products, the upper tiles of the 32 x 32 tile grid spread over
// the waves; kPMax 96: up to 3 x 3 tiles). The shift c is the weight-averaged pilot of the waves
// of the first chunk that holds a reached ray, o its weighted mean OPL. kStream (kNW 1, tuples of
// at most 64 rays, one chunk): the wave's tiles are formed and stored one after another.
template <int kNW, int kPMax, bool kStates, bool kGlass>
__global__ __launch_bounds__(64 * kNW) __attribute__((amdgpu_waves_per_eu(MKO_FUSED_WAVES_PER_EU)))
void adjoint_normal_kernel(TraceParams P, ParamMap map,
                                                                  float* partC, double* partST,
                                                                  TupleNormal* partT,
                                                                  mko_internal::RayStates states) {
    using Lay = GradLayout<kNW>;
    constexpr int kThreads = 64 * kNW;
    constexpr int kTileGrid = (kPMax + 31) / 32;
    constexpr int kTilesAll = kTileGrid * (kTileGrid + 1) / 2;
    constexpr bool kStream = kNW == 1;
    constexpr int kTiles = kStream ? 1 : (kTilesAll + kNW - 1) / kNW;
    constexpr int kGroups = (kPMax + 15) / 16;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ SurfLds L;
    __shared__ double red[kNW];
    __shared__ double wsum_l[kNW], osum_l[kNW];
    if (P.live && int(blockIdx.x) / P.T >= *P.live) return;
    const Tuple t = tuple_of(P, blockIdx.x);
    stage_surfaces(P, t, L);
    const int Pn = map.P;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    unsigned char* G = lds;
    double* pilot = reinterpret_cast<double*>(lds + size_t(Pn) * Lay::kRowBytes);
    double* shift = pilot + kNW * Pn;
    double* own_pilot = pilot + wave * Pn;
    const size_t vf = size_t(t.v) * P.F + t.f;
    // block-uniform values in scalar registers (they stay live across the sweep)
    const double a = first_lane(P.pupil_radius[t.v]);
    double sin_t, cos_t;
    sincos(P.field_angle[vf], &sin_t, &cos_t);
    sin_t = first_lane(sin_t);
    cos_t = first_lane(cos_t);
    const V3<double> M{0.0, first_lane(P.field_target[vf]), first_lane(L.t[P.S - 1])};
    const double kappa =
        first_lane(P.ref_kind != MKO_REF_LAST_SURFACE ? P.reference_curvature[vf] : 0.0);
    const int tg = (Pn + 31) / 32;
    const int ntiles = tg * (tg + 1) / 2;
    const size_t tuple = size_t(blockIdx.x);
    f32x16 acc[kTiles];
#pragma unroll
    for (int i = 0; i < kTiles; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[i][e] = 0.0f;
    // this wave's slot of the tuple's S and Tr in the workspace ([tuple][4][2 P]): each chunk adds
    // its sums (the first one stores), so no accumulator stays live across the sweeps
    double* slot_st = partST + (tuple * kFusedSlots + wave) * 2 * Pn;
    bool accumulated = false;
    double sw = 0.0, swo = 0.0, o_shift = 0.0;
    bool shift_ready = false;
    const int nchunks = (t.n + kThreads - 1) / kThreads;
    // when the tuple's spoke count divides the chunk, a lane's spoke (k mod M) is the same in every
    // chunk: its angle is formed once here and a ray's sample needs only its ring (the same
    // functions as sample_of, hence the same bits)
    const bool fixed_spoke = kThreads % t.M == 0;
    Sample spoke{};
    if (fixed_spoke) spoke_of(P, t, tid % t.M, spoke);
    for (int ch = 0; ch < nchunks; ++ch) {
        const int k = ch * kThreads + tid;
        double st[7];
        if constexpr (kStates) {
            st[6] = __builtin_nan("");
            if (k < t.n) {
                const int64_t row = t.offset + k;
#pragma unroll
                for (int c = 0; c < 7; ++c) st[c] = states.data[state_index(states, t.v, c, row)];
            }
        }
        // the block (all threads, 16-byte words) and this wave's pilots start at zero
        {
            float4* z = reinterpret_cast<float4*>(G);
            const int words = Pn * (Lay::kRowBytes / 16);
            for (int e = tid; e < words; e += kThreads) z[e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        for (int p = lane; p < Pn; p += 64) own_pilot[p] = 0.0;
        __syncthreads();
        double wk = 0.0, opl = 0.0;
        bool ok = false;
        Sample smp{};
        Ray start;
        LastState last;
        if (k < t.n) {
            if (fixed_spoke) {
                smp = spoke;
                ring_of(P, t, k / t.M, smp);
            } else {
                smp = sample_of(P, t, k);
            }
            Ray r = source_ray(P, t, smp, a, sin_t, cos_t, L.n[0]);
            start = r;
            last.count = 0;
            last.image_branch = 0;
            if constexpr (kStates) {
                opl = st[6];
                ok = !__builtin_isnan(opl);
                last.p = V3<double>{st[0], st[1], st[2]};
                last.d = V3<double>{st[3], st[4], st[5]};
                last.count = P.S;
            } else {
                Soft soft{0.0, MKO_STATUS_OK};
                ok = trace_ray<false, true>(P, L, r, M, kappa, soft, last) == MKO_STATUS_OK;
                opl = r.opl;
            }
        }
        const uint64_t okmask = __ballot(ok);
        if (ok) {
            wk = smp.weight;
            sw += wk;
            swo += wk * opl;
            const bool leader = lane == int(__builtin_ctzll(okmask));
            backward_fused<kGlass>(P, L, map, last, start, M, kappa,
                                   SwzSink<kNW>{G, own_pilot, tid, leader}, t.l);
        }
        __syncthreads();
        // this wave's weight sum W_w; the first chunk with a reached ray fixes o = sum W OPL / W
        const double ww = wave_sum(wk);
        double W = 0.0;
        if (!shift_ready) {
            const double wo = wave_sum(wk * opl);
            if (lane == 0) {
                wsum_l[wave] = ww;
                osum_l[wave] = wo;
            }
            __syncthreads();
            double WO = 0.0;
#pragma unroll
            for (int w = 0; w < kNW; ++w) {
                W += wsum_l[w];
                WO += osum_l[w];
            }
            if (!(W > 0.0)) continue; // no reached ray yet: nothing of this chunk counts
            o_shift = WO / W;
        }
        // A_p = sum_k W_k G_pk and B_p = sum_k W_k (OPL_k - o) G_pk over the wave's rays, G the
        // stored floats relative to the wave's pilot: the parts of S and Tr that do not depend on
        // the shift c (S = A + (pilot - c) W_w, Tr = B + (pilot - c) V_w)
        const double ow = ok ? opl - o_shift : 0.0;
        const double vk = wk * ow;
        double A[2], B[2];
        if constexpr (kPMax > 64) {
            if (Pn > 64) wave_row_sums<kNW, 2>(G, Pn, wave, lane, wk, vk, A, B);
            else wave_row_sums<kNW, 1>(G, Pn, wave, lane, wk, vk, A, B);
        } else {
            wave_row_sums<kNW, 1>(G, Pn, wave, lane, wk, vk, A, B);
        }
        const double vw = wave_sum(vk);
        if (!shift_ready) {
            // c0: the weight-averaged pilot of the waves; then c = c0 + S0 / W with S0 the sum of
            // W (g - c0) over the chunk, i.e. the chunk's weighted mean gradient (the waves' S0
            // pass through their slots of the workspace, summed in wave order)
            for (int p = tid; p < Pn; p += kThreads) {
                double c = 0.0;
#pragma unroll
                for (int w = 0; w < kNW; ++w) c += wsum_l[w] * pilot[w * Pn + p];
                shift[p] = c / W;
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int p = lane + 64 * h;
                if (p < Pn) slot_st[p] = A[h] + (own_pilot[p] - shift[p]) * ww;
            }
            __threadfence_block();
            __syncthreads();
            for (int p = tid; p < Pn; p += kThreads) {
                const double* st0 = partST + tuple * kFusedSlots * 2 * Pn + p;
                double s0 = st0[0];
#pragma unroll
                for (int w = 1; w < kNW; ++w) s0 += st0[w * 2 * Pn];
                shift[p] += s0 / W;
            }
            shift_ready = true;
            __syncthreads();
        }
        // S and Tr of this wave's rays into its slot; the wave's deltas pilot - c, as floats
        // over its pilots (read by the bf16 conversion below)
        float* delta = reinterpret_cast<float*>(own_pilot);
        double d[2] = {0.0, 0.0};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p = lane + 64 * h;
            if (p < Pn) {
                d[h] = own_pilot[p] - shift[p];
                double* dst = slot_st + p;
                const double s = A[h] + d[h] * ww, tr = B[h] + d[h] * vw;
                dst[0] = accumulated ? dst[0] + s : s;
                dst[Pn] = accumulated ? dst[Pn] + tr : tr;
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int p = lane + 64 * h;
            if (p < Pn) delta[p] = float(d[h]);
        }
        __syncthreads();
        accumulated = true;
        // u = sqrt(W) (g - c) of the thread's ray, split into bf16 hi + lo in place (the ray's
        // 32-byte group belongs to 8 lanes of this wave: its floats are read before they are
        // overwritten)
        {
            const float swk = __builtin_sqrtf(float(wk));
            const int grp = tid >> 3, i = tid & 7;
            // rows in blocks of 16: for p = p0 + j (p0 a multiple of 16) the swizzle p & kMask is
            // m = j & kMask, so float_off(p, tid) = p RB + (4 tid ^ 32 m), hi_off(p, grp ^ m) + 2 i
            // = p RB + (hbase ^ (32 m + 16 ((m >> 3) & 1))) and lo_off(...) + 2 i = that ^ 16;
            // two rows at a time share the packed bf16 conversions
            const int fbase = 4 * tid;
            const int hbase = 32 * grp + 16 * ((grp >> 3) & 1) + 2 * i;
            for (int p0 = 0; p0 < Pn; p0 += 16) {
                unsigned char* row = G + size_t(p0) * Lay::kRowBytes;
#pragma unroll 2
                for (int j = 0; j < 16; j += 2) {
                    if (p0 + j < Pn) {
                        const bool two = p0 + j + 1 < Pn;
                        const int m0 = j & Lay::kMask, m1 = (j + 1) & Lay::kMask;
                        unsigned char* r0 = row + j * Lay::kRowBytes;
                        unsigned char* r1 = r0 + Lay::kRowBytes;
                        f32x2 u;
                        u.x = swk * (*reinterpret_cast<const float*>(r0 + (fbase ^ (32 * m0))) +
                                     delta[p0 + j]);
                        u.y = two ? swk * (*reinterpret_cast<const float*>(
                                               r1 + (fbase ^ (32 * m1))) +
                                           delta[p0 + j + 1])
                                  : 0.0f;
                        const bf16x2 hi = __builtin_convertvector(u, bf16x2);
                        const f32x2 d{u.x - float(hi[0]), u.y - float(hi[1])};
                        const bf16x2 lo = __builtin_convertvector(d, bf16x2);
                        const int h0 = hbase ^ (32 * m0 + 16 * ((m0 >> 3) & 1));
                        const int h1 = hbase ^ (32 * m1 + 16 * ((m1 >> 3) & 1));
                        *reinterpret_cast<__bf16*>(r0 + h0) = hi[0];
                        *reinterpret_cast<__bf16*>(r0 + (h0 ^ 16)) = lo[0];
                        if (two) {
                            *reinterpret_cast<__bf16*>(r1 + h1) = hi[1];
                            *reinterpret_cast<__bf16*>(r1 + (h1 ^ 16)) = lo[1];
                        }
                    }
                }
            }
        }
        __syncthreads();
        if constexpr (kStream) {
            float* C = partC + tuple * size_t(Pn) * Pn;
            for (int tt = 0; tt < ntiles; ++tt) {
                int I, J;
                tile_of(tt, tg, I, J);
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[0][e] = 0.0f;
                tile_products<kNW>(G, Pn, I, J, lane, acc[0]);
                store_tile(C, Pn, I, J, lane, acc[0]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < kTiles; ++i) {
                const int tt = wave + i * kNW;
                if (tt < ntiles) {
                    int I, J;
                    tile_of(tt, tg, I, J);
                    tile_products<kNW>(G, Pn, I, J, lane, acc[i]);
                }
            }
        }
        __syncthreads();
    }
    if constexpr (kStream) {
        if (!shift_ready) {
            // no reached ray in the tuple: C = 0
            float* C = partC + tuple * size_t(Pn) * Pn;
            for (int e = tid; e < Pn * Pn; e += kThreads) C[e] = 0.0f;
        }
    } else {
        float* C = partC + tuple * size_t(Pn) * Pn;
#pragma unroll
        for (int i = 0; i < kTiles; ++i) {
            const int tt = wave + i * kNW;
            if (tt < ntiles) {
                int I, J;
                tile_of(tt, tg, I, J);
                store_tile(C, Pn, I, J, lane, acc[i]);
            }
        }
    }
    const double w_ok = block_sum_n<kNW>(sw, red);
    const double sum = block_sum_n<kNW>(swo, red);
    // a tuple without a reached ray: its slots hold zeros (the finalize sums kNW slots)
    if (!accumulated)
        for (int e = lane; e < 2 * Pn; e += 64) slot_st[e] = 0.0;
    if (tid == 0) {
        const bool valid = w_ok > 0.0;
        // piston - o, the OPL shift of Tr
        partT[tuple] = TupleNormal{w_ok, valid ? sum / w_ok - o_shift : 0.0, t.scale * t.scale};
    }
}

// a = sum_t s_t^2 (C_t - S_t S_t^T / W_t), g = sum_t s_t^2 (Tr_t - (piston_t - o_t) S_t), tuples
// in (field, wavelength) order; one workgroup per run. Each thread owns up to 19 entries of the
// upper triangle (mirrored on output); the tuples' S, Tr and factors are staged through LDS in
// chunks of 8 tuples.
constexpr int kFinalizeChunk = 8;
// kE entries per thread: 5 up to 49 variables, 10 up to 71, 19 up to 96 (the kernel is
// instantiated per bound, so the accumulators of the smaller counts do not hold registers and
// more workgroups fit a CU).
constexpr int kFinalizeEntries = (kMaxVariablesFused * (kMaxVariablesFused + 1) / 2 + 255) / 256;

template <int kE>
__global__ __launch_bounds__(256) void adjoint_normal_finalize_kernel(
    int T, int Pn, int slots, const float* partC, const double* partST, const TupleNormal* partT,
    double* a, double* g, const int32_t* live) {
    __shared__ double S_l[kFinalizeChunk][2 * kMaxVariablesFused];
    __shared__ double f1[kFinalizeChunk], f2[kFinalizeChunk], pist[kFinalizeChunk];
    const int v = blockIdx.x;
    if (live && v >= *live) return;
    const int tid = threadIdx.x;
    const int nu = Pn * (Pn + 1) / 2;
    int eq[kE], ep[kE];
    double acc[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        acc[i] = 0.0;
        // upper-triangle index u -> (q, p), q <= p, row-major over q
        const int u = tid + i * 256;
        int q = 0, rem = u;
        while (q < Pn && rem >= Pn - q) {
            rem -= Pn - q;
            ++q;
        }
        eq[i] = q;
        ep[i] = q + rem;
    }
    double gacc = 0.0;
    for (int t0 = 0; t0 < T; t0 += kFinalizeChunk) {
        const int nt = min(kFinalizeChunk, T - t0);
        for (int e = tid; e < nt * 2 * Pn; e += 256) {
            // the waves' slots of S and Tr, summed in wave order
            const int tt = e / (2 * Pn), i = e - tt * 2 * Pn;
            const double* st = partST + (size_t(v) * T + t0 + tt) * kFusedSlots * 2 * Pn + i;
            double x = st[0];
            for (int w = 1; w < slots; ++w) x += st[w * 2 * Pn];
            S_l[tt][i] = x;
        }
        if (tid < nt) {
            const TupleNormal tn = partT[size_t(v) * T + t0 + tid];
            const bool valid = tn.w_ok > 0.0;
            f1[tid] = valid ? tn.s2 : 0.0;
            f2[tid] = valid ? tn.s2 / tn.w_ok : 0.0;
            pist[tid] = valid ? tn.piston : 0.0;
        }
        __syncthreads();
        for (int tt = 0; tt < nt; ++tt) {
            const float* C = partC + (size_t(v) * T + t0 + tt) * size_t(Pn) * Pn;
#pragma unroll
            for (int i = 0; i < kE; ++i) {
                if (tid + i * 256 < nu) {
                    const int q = eq[i], p = ep[i];
                    const double c = double(C[size_t(q) * Pn + p]);
                    acc[i] += f1[tt] * c - f2[tt] * (S_l[tt][p] * S_l[tt][q]);
                }
            }
            if (tid < Pn) gacc += f1[tt] * (S_l[tt][Pn + tid] - pist[tt] * S_l[tt][tid]);
        }
        __syncthreads();
    }
    double* A = a + size_t(v) * Pn * Pn;
#pragma unroll
    for (int i = 0; i < kE; ++i) {
        if (tid + i * 256 < nu) {
            A[size_t(eq[i]) * Pn + ep[i]] = acc[i];
            A[size_t(ep[i]) * Pn + eq[i]] = acc[i];
        }
    }
    if (tid < Pn) g[size_t(v) * Pn + tid] = gacc;
}

// ---- Gram: A = [J r0]^T [J r0] per run -----------------------------------------------------

// 64 x 64 output tile per workgroup over the upper tiles of the (P+1) x (P+1) matrix, 256 threads
// each holding a 4 x 4 block of elements: 16 independent FMA chains per thread (the chains' latency
// is what bounds a Gram of long columns). Rows are staged through LDS 32 at a time, row-major
// [row][column], so a thread reads its 4 consecutive columns as two 16-byte words. Each element is
// one FMA chain over the rows in ascending order starting from 0 (gram_dot,
// normal_equations.hpp:19), products of the stored J values: bitwise the host definition.
__global__ __launch_bounds__(256) void gram_kernel(const double* J, const double* r0, int runs, int P,
                                                   int64_t rows, double* a, double* g,
                                                   const int32_t* live) {
    constexpr int kT = 64, kC = 32, kLd = kT + 2; // row stride in doubles (16-byte aligned rows)
    __shared__ __align__(16) double As[kC][kLd];
    __shared__ __align__(16) double Bs[kC][kLd];
    const int T = (P + 1 + kT - 1) / kT;
    const int tiles = T * (T + 1) / 2;
    const int run = blockIdx.x / tiles;
    if (live && run >= *live) return;
    int tile = blockIdx.x - run * tiles;
    int tq = 0;
    while (tile >= T - tq) {
        tile -= T - tq;
        ++tq;
    }
    const int tp = tq + tile; // column tile >= row tile
    const bool diag = tp == tq;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const double* Jr = J + size_t(run) * P * rows;
    const double* rr = r0 + size_t(run) * rows;
    auto column = [&](int col, int64_t row) {
        return col < P ? Jr[size_t(col) * rows + row] : col == P ? rr[row] : 0.0;
    };
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    const double(*Bsrc)[kLd] = diag ? As : Bs;
    for (int64_t base = 0; base < rows; base += kC) {
        const int n = int(rows - base < kC ? rows - base : kC);
        for (int e = threadIdx.x; e < kT * kC; e += 256) {
            const int c = e / kC, rw = e - c * kC;
            const int64_t row = base + rw;
            As[rw][c] = rw < n ? column(tq * kT + c, row) : 0.0;
            if (!diag) Bs[rw][c] = rw < n ? column(tp * kT + c, row) : 0.0;
        }
        __syncthreads();
        for (int rw = 0; rw < n; ++rw) {
            const double2 a01 = *reinterpret_cast<const double2*>(&As[rw][ty * 4]);
            const double2 a23 = *reinterpret_cast<const double2*>(&As[rw][ty * 4 + 2]);
            const double2 b01 = *reinterpret_cast<const double2*>(&Bsrc[rw][tx * 4]);
            const double2 b23 = *reinterpret_cast<const double2*>(&Bsrc[rw][tx * 4 + 2]);
            const double av[4] = {a01.x, a01.y, a23.x, a23.y};
            const double bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __fma_rn(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
    double* A = a + size_t(run) * P * P;
    double* gr = g + size_t(run) * P;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int p = tq * kT + ty * 4 + i;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = tp * kT + tx * 4 + j;
            const double v = acc[i][j];
            if (p < P && q < P) {
                A[size_t(p) * P + q] = v;
                A[size_t(q) * P + p] = v;
            } else if (p < P && q == P) {
                gr[p] = v;
            } else if (p == P && q < P) {
                gr[q] = v;
            }
        }
    }
}

// ---- Gram on FP64 MFMA ----------------------------------------------------------------------

typedef __attribute__((ext_vector_type(4))) double f64x4;

// [J r0]^T [J r0] of one run per workgroup with v_mfma_f64_16x16x4_f64: the run's P + 1 columns
// staged through LDS 32 rows at a time (doubles, row stride 33), the upper 16 x 16 tiles of the
// (P + 1) x (P + 1) result spread over the 4 waves (kTiles per wave), FP64 accumulation in the
// MFMA unit. Operand lane l: A(m = l % 16, k = l / 16), B(k = l / 16, n = l % 16); result lane l,
// element e: row 4 e + l / 16, column l % 16 (measured on MI355X). Not the host's summation
// order (unlike gram_kernel): the results differ from normal_equations_host by rounding.
template <int kTiles>
__global__ __launch_bounds__(256) void gram_mfma_kernel(const double* J, const double* r0, int P,
                                                        int64_t rows, double* a, double* g) {
    constexpr int kC = 32, kLd = kC + 1;
    extern __shared__ double Xs[]; // [16 tg][kLd]
    const int run = blockIdx.x;
    const int Q = P + 1;
    const int tg = (Q + 15) / 16, cols = 16 * tg;
    const int ntiles = tg * (tg + 1) / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    f64x4 acc[kTiles];
    int tI[kTiles], tJ[kTiles];
#pragma unroll
    for (int i = 0; i < kTiles; ++i) {
        acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
        const int tt = wave + 4 * i;
        tI[i] = -1;
        tJ[i] = -1;
        if (tt < ntiles) tile_of(tt, tg, tI[i], tJ[i]);
    }
    const double* Jr = J + size_t(run) * P * rows;
    const double* rr = r0 + size_t(run) * rows;
    const int m = lane & 15, kk = lane >> 4;
    for (int64_t base = 0; base < rows; base += kC) {
        for (int e = tid; e < cols * kC; e += 256) {
            const int c = e / kC, rw = e - c * kC;
            const int64_t row = base + rw;
            double v = 0.0;
            if (row < rows) v = c < P ? Jr[size_t(c) * rows + row] : c == P ? rr[row] : 0.0;
            Xs[c * kLd + rw] = v;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kTiles; ++i) {
            if (tI[i] < 0) continue;
            const double* xa = Xs + (tI[i] * 16 + m) * kLd + kk;
            const double* xb = Xs + (tJ[i] * 16 + m) * kLd + kk;
#pragma unroll
            for (int st = 0; st < kC / 4; ++st)
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[4 * st], xb[4 * st], acc[i], 0, 0,
                                                              0);
        }
        __syncthreads();
    }
    double* A = a + size_t(run) * P * P;
    double* gr = g + size_t(run) * P;
#pragma unroll
    for (int i = 0; i < kTiles; ++i) {
        if (tI[i] < 0) continue;
        const int col = tJ[i] * 16 + m;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = tI[i] * 16 + 4 * e + kk;
            const double v = acc[i][e];
            if (row < P && col < P) {
                A[size_t(row) * P + col] = v;
                A[size_t(col) * P + row] = v;
            } else if (row < P && col == P) {
                gr[row] = v;
            } else if (row == P && col < P) {
                gr[col] = v;
            }
        }
    }
}

// ---- helpers ---------------------------------------------------------------------------------

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The raw device entry points trace ray rows only; constraint rows are the host layer's
// (mko_engine_*), which evaluates them next to the traces.
void no_rows(const mko_residual_settings& s) {
    if (s.extra && s.extra->count > 0)
        throw std::invalid_argument("constraint rows are evaluated by the mko_engine_* entry points");
}

TraceParams make_params(const mko_batch& b, const mko_residual_settings& s) {
    if (b.surfaces < 2 || b.surfaces > MKO_MAX_SURFACES)
        throw std::invalid_argument("surfaces outside [2, 64]");
    if (b.fields < 1 || b.fields > kMaxFields)
        throw std::invalid_argument("fields outside [1, 32]");
    if (b.wavelengths < 1 || b.wavelengths > kMaxWavelengths)
        throw std::invalid_argument("wavelengths outside [1, 32]");
    if (s.rings < 1 || s.rings > MKO_MAX_RINGS) throw std::invalid_argument("rings outside [1, 16]");
    if (s.spokes < 1) throw std::invalid_argument("spokes < 1");
    if (s.failure != MKO_FAILURE_SOFT && s.failure != MKO_FAILURE_HARD)
        throw std::invalid_argument("unknown failure mode");
    if (!(s.fail_weight >= 0.0) || !std::isfinite(s.fail_weight))
        throw std::invalid_argument("fail_weight must be finite and >= 0");
    if (s.precision != MKO_PRECISION_FP64 && s.precision != MKO_PRECISION_DF64)
        throw std::invalid_argument("trace precision must be MKO_PRECISION_FP64 or MKO_PRECISION_DF64");
    if (b.reference_sphere != MKO_REF_LAST_SURFACE && b.reference_curvature == nullptr)
        throw std::invalid_argument("image reference sphere needs reference_curvature");
    if (b.source == MKO_SOURCE_POINT && b.fields != 1)
        throw std::invalid_argument("point source needs exactly one field");
    TraceParams P{};
    P.curvature = b.curvature;
    P.thickness = b.thickness;
    P.index = b.index;
    P.pupil_radius = b.pupil_radius;
    P.field_angle = b.field_angle;
    P.field_target = b.field_target;
    P.reference_curvature = b.reference_curvature;
    P.V = b.variants;
    P.S = b.surfaces;
    P.L = b.wavelengths;
    P.F = b.fields;
    P.T = b.fields * b.wavelengths;
    P.source = b.source;
    P.ref_kind = b.reference_sphere;
    P.rings = s.rings;
    P.spokes = s.spokes;
    P.soft = s.failure == MKO_FAILURE_SOFT;
    P.precision = s.precision;
    P.fail_weight = s.fail_weight;
    P.lambda_ref_nm = b.reference_wavelength_nm;
    for (int i = 0; i < 3; ++i) P.src_pt[i] = b.source_point[i];
    P.src_na = b.source_na;
    P.sphere_mask = 0;
    for (int i = 0; i < b.surfaces; ++i) {
        if (b.type[i] > MKO_SURFACE_SPHERE)
            throw std::invalid_argument("unsupported surface type (ghost plates are not supported)");
        if (b.type[i] == MKO_SURFACE_SPHERE) P.sphere_mask |= 1ull << i;
    }
    int64_t rays = 0;
    int nmax = 0;
    for (int f = 0; f < b.fields; ++f) {
        if (b.pupil_mode[f] > MKO_PUPIL_AXIS) throw std::invalid_argument("unknown pupil mode");
        P.pupil_mode[f] = b.pupil_mode[f];
        P.field_weight[f] = b.field_weight[f];
        const int n = s.rings * (b.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : s.spokes);
        nmax = std::max(nmax, n);
        rays += int64_t(b.wavelengths) * n;
    }
    if (nmax > kBlock * kMaxRaysPerLane)
        throw std::invalid_argument("more than 2048 pupil samples per tuple");
    for (int l = 0; l < b.wavelengths; ++l) P.wavelength_weight[l] = b.wavelength_weight[l];
    const int off = s.rings * (s.rings - 1) / 2;
    for (int i = 0; i < s.rings; ++i) {
        P.gl_u[i] = mko_gl_u[off + i];
        P.gl_w[i] = mko_gl_w[off + i];
    }
    P.rays = rays;
    P.trace_rows = P.soft ? 2 * rays : rays;
    P.stride = P.trace_rows;
    return P;
}

// Launch shapes of the calling thread (set by the host layer for its calls; the device-layer
// entry points run with the defaults).
thread_local mko_internal::Tuning g_tuning{};

// Segment of a tuple pack for this launch (0: one tuple per workgroup): tuples of at most 128
// rays (FP64 traces without extrema) share a workgroup in segments of 32, 64 or 128 lanes, the
// smallest that holds the largest tuple, or the tuning's pack_rays where the tuples fit it
// (pack_rays 256 turns packing off).
int pack_segment(const TraceParams& P, bool packable) {
    if (!packable || P.precision != MKO_PRECISION_FP64) return 0;
    int nmax = 0;
    for (int f = 0; f < P.F; ++f)
        nmax = std::max(nmax, P.rings * (P.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : P.spokes));
    const int want = g_tuning.pack_rays;
    if (want >= kBlock || nmax > 128 || P.T < 2) return 0;
    int seg = nmax <= 32 ? 32 : nmax <= 64 ? 64 : 128;
    if (want == 32 || want == 64 || want == 128) seg = std::max(seg, want);
    return seg;
}

int rays_per_lane(const TraceParams& P) {
    int nmax = 0;
    for (int f = 0; f < P.F; ++f)
        nmax = std::max(nmax, P.rings * (P.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : P.spokes));
    const int r = (nmax + kBlock - 1) / kBlock;
    return r <= 1 ? 1 : r <= 2 ? 2 : r <= 4 ? 4 : 8;
}

template <int kMode, bool kSoft, class R, bool kStates, bool kExt>
void launch_trace_rpl(int rpl, const TraceParams& P, double* res, double* opd, uint8_t* st,
                      double* part, int* pf, int* ph, const mko_internal::RayStates& states,
                      double* ext, hipStream_t stream) {
    const dim3 grid(unsigned(P.V) * unsigned(P.T));
#define MKO_LAUNCH(RPL)                                                                          \
    hipLaunchKernelGGL((trace_kernel<kMode, kSoft, RPL, R, kStates, kExt>), grid, dim3(kBlock),   \
                       0, stream, P, res, opd, st, part, pf, ph, states, ext)
    switch (rpl) {
    case 1: MKO_LAUNCH(1); break;
    case 2: MKO_LAUNCH(2); break;
    case 4: MKO_LAUNCH(4); break;
    default: MKO_LAUNCH(8); break;
    }
#undef MKO_LAUNCH
}

template <int kMode, bool kSoft, class R>
void launch_trace_ext(int rpl, const TraceParams& P, double* res, double* opd, uint8_t* st,
                      double* part, int* pf, int* ph, double* ext, hipStream_t stream) {
    const mko_internal::RayStates none{};
    if (ext) launch_trace_rpl<kMode, kSoft, R, false, true>(rpl, P, res, opd, st, part, pf, ph, none, ext, stream);
    else launch_trace_rpl<kMode, kSoft, R, false, false>(rpl, P, res, opd, st, part, pf, ph, none, nullptr, stream);
}

template <class R>
void launch_trace_r(int mode, int rpl, const TraceParams& P, double* res, double* opd, uint8_t* st,
                    double* part, int* pf, int* ph, double* ext, hipStream_t stream) {
    if (mode == 0) {
        if (P.soft) launch_trace_ext<0, true, R>(rpl, P, res, opd, st, part, pf, ph, ext, stream);
        else launch_trace_ext<0, false, R>(rpl, P, res, opd, st, part, pf, ph, ext, stream);
    } else {
        if (P.soft) launch_trace_ext<1, true, R>(rpl, P, res, opd, st, part, pf, ph, ext, stream);
        else launch_trace_ext<1, false, R>(rpl, P, res, opd, st, part, pf, ph, ext, stream);
    }
}

// RESIDUALS (mode 0) or MERIT partials (mode 1); ext: extrema partials [V*T][S][2] or NULL;
// states: the LM MERIT's final ray states (mode 1, FP64 trace; stored for the rays that end
// with status Ok) or NULL.
void launch_trace(int mode, const TraceParams& P, double* res, double* opd, uint8_t* st,
                  double* part, int* pf, int* ph, double* ext, hipStream_t stream,
                  const mko_internal::RayStates* states = nullptr) {
    if (int64_t(P.V) * P.T > 0x7fffffffLL) throw std::invalid_argument("batch too large for one launch");
    const int rpl = rays_per_lane(P);
    const int seg = pack_segment(P, ext == nullptr);
    if (seg > 0) {
        if (states && (mode != 1 || P.precision != MKO_PRECISION_FP64))
            throw std::invalid_argument("ray states need the MERIT trace and FP64");
        const int tpb = kBlock / seg;
        const dim3 grid(unsigned(P.V) * unsigned((P.T + tpb - 1) / tpb));
        const size_t lds = size_t(tpb) * sizeof(SurfLds);
        const mko_internal::RayStates none{};
        const mko_internal::RayStates& rs = states ? *states : none;
#define MKO_PACK(M, SOFT, ST)                                                                      \
    hipLaunchKernelGGL((trace_pack_kernel<M, SOFT, ST>), grid, dim3(kBlock), lds, stream, P, seg,   \
                       res, opd, st, part, pf, ph, rs)
        if (states) {
            if (P.soft) MKO_PACK(1, true, true);
            else MKO_PACK(1, false, true);
        } else if (mode == 0) {
            if (P.soft) MKO_PACK(0, true, false);
            else MKO_PACK(0, false, false);
        } else {
            if (P.soft) MKO_PACK(1, true, false);
            else MKO_PACK(1, false, false);
        }
#undef MKO_PACK
        check_hip(hipGetLastError(), "trace pack kernel launch");
        return;
    }
    if (states) {
        if (mode != 1 || P.precision != MKO_PRECISION_FP64)
            throw std::invalid_argument("ray states need the MERIT trace and FP64");
        if (P.soft) {
            if (ext) launch_trace_rpl<1, true, double, true, true>(rpl, P, res, opd, st, part, pf, ph, *states, ext, stream);
            else launch_trace_rpl<1, true, double, true, false>(rpl, P, res, opd, st, part, pf, ph, *states, nullptr, stream);
        } else if (ext) {
            launch_trace_rpl<1, false, double, true, true>(rpl, P, res, opd, st, part, pf, ph, *states, ext, stream);
        } else {
            launch_trace_rpl<1, false, double, true, false>(rpl, P, res, opd, st, part, pf, ph, *states, nullptr, stream);
        }
    } else if (P.precision == MKO_PRECISION_DF64) {
        launch_trace_r<mkod::df64>(mode, rpl, P, res, opd, st, part, pf, ph, ext, stream);
    } else {
        launch_trace_r<double>(mode, rpl, P, res, opd, st, part, pf, ph, ext, stream);
    }
    check_hip(hipGetLastError(), "trace kernel launch");
}

// buchdahl_omega (dispersion.hpp:18-89) of nm / 1000 on the host, contract off
double buchdahl_omega_host(double lambda_nm, bool& ok) {
#pragma clang fp contract(off)
    const double um = lambda_nm / 1000.0;
    const double den = 1.0 + 2.5 * (um - 0.546074);
    ok = ok && den > 0.0;
    return (um - 0.546074) / den;
}

OmegaTable omega_table(const mko_batch& b) {
    OmegaTable t{};
    t.valid = 0;
    if (!b.wavelength_nm || b.wavelengths > kMaxWavelengths) return t;
    bool ok = true;
    for (int l = 0; l < b.wavelengths; ++l) t.w[l] = buchdahl_omega_host(b.wavelength_nm[l], ok);
    t.w[b.wavelengths] = buchdahl_omega_host(b.reference_wavelength_nm, ok);
    t.valid = ok ? 1 : 0;
    return t;
}

// Column map of the variables (AdjointParameters, src/optim/adjoint_jacobian.cpp:79-127); glass
// variables need the batch's wavelengths (host) for their slopes, and `allow_glass`.
static_assert(sizeof(TraceParams) + sizeof(ParamMap) < 3584, "kernel arguments within 4 KB");
ParamMap make_map(const mko_variables& vars, const mko_batch& batch, bool allow_glass) {
    const int S = batch.surfaces;
    ParamMap m{};
    m.P = vars.count;
    if (vars.count < 1 || vars.count > 127) throw std::invalid_argument("1 ... 127 variables");
    for (int s = 0; s < MKO_MAX_SURFACES; ++s) {
        m.curv_col[s] = -1;
        m.thick_col[s] = -1;
        for (int c = 0; c < 3; ++c) m.glass_col[s][c] = -1;
    }
    for (int k = 0; k < vars.count; ++k) {
        const int s = vars.batch_surface[k];
        if (s < 0 || s >= S) throw std::invalid_argument("variable surface out of range");
        if (vars.type[k] < MKO_VAR_CURVATURE || vars.type[k] > MKO_VAR_GLASS)
            throw std::invalid_argument("unsupported variable type");
        if (vars.type[k] == MKO_VAR_GLASS) {
            if (!allow_glass)
                throw std::invalid_argument("glass variables need the FP64 adjoint (not the fused "
                                            "normal equations)");
            if (!vars.component || vars.component[k] < 0 || vars.component[k] > 2)
                throw std::invalid_argument("glass variables need components 0 ... 2");
            int32_t& slot = m.glass_col[s][vars.component[k]];
            if (slot >= 0) throw std::invalid_argument("parameter belongs to two variables");
            slot = k;
            m.glass = 1;
        } else {
            int32_t& slot = vars.type[k] == MKO_VAR_CURVATURE ? m.curv_col[s] : m.thick_col[s];
            if (slot >= 0) throw std::invalid_argument("parameter belongs to two variables");
            slot = k;
        }
    }
    if (m.glass) {
        const OmegaTable t = omega_table(batch);
        if (!t.valid)
            throw std::invalid_argument("glass variables need wavelength_nm with a defined Buchdahl "
                                        "coordinate (lambda > 146.074 nm)");
        for (int l = 0; l < batch.wavelengths; ++l) m.omega[l] = t.w[l];
    }
    return m;
}

template <bool kSoft, class RB, class RR, class TJ>
void launch_adjoint_rpl(int rpl, const TraceParams& P, const ParamMap& m, TJ* J, double* r0,
                        uint8_t* st, hipStream_t stream) {
    const dim3 grid(unsigned(P.V) * unsigned(P.T));
#define MKO_LAUNCH(R)                                                                            \
    hipLaunchKernelGGL((adjoint_kernel<kSoft, R, RB, RR, TJ>), grid, dim3(kBlock), 0, stream, P,  \
                       m, J, r0, st)
    switch (rpl) {
    case 1: MKO_LAUNCH(1); break;
    case 2: MKO_LAUNCH(2); break;
    case 4: MKO_LAUNCH(4); break;
    default: MKO_LAUNCH(8); break;
    }
#undef MKO_LAUNCH
}

#define MKO_TRY(...)                                                                               \
    try {                                                                                          \
        __VA_ARGS__;                                                                               \
        return 0;                                                                                  \
    } catch (const std::invalid_argument& e) {                                                     \
        mko_detail::last_error = e.what();                                                                        \
        return MKO_ERR_INVALID;                                                                    \
    } catch (const std::exception& e) {                                                            \
        mko_detail::last_error = e.what();                                                                        \
        return MKO_ERR_DEVICE;                                                                     \
    }

// Check kernel of the df64 arithmetic (df64_check_kernel, src/cuda/df64_check.cu:17): one
// operation per element, for the bitwise comparison with the host restatement.
__global__ void df64_ops_kernel(int op, int64_t n, const float* a, const float* b, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using mkod::df64;
    const df64 x{a[2 * i], a[2 * i + 1]};
    const df64 y = op == 5 ? df64{0.0f, 0.0f} : df64{b[2 * i], b[2 * i + 1]};
    df64 z;
    switch (op) {
    case 0: z = x + y; break;
    case 1: z = x - y; break;
    case 2: z = x * y; break;
    case 3: z = mkod::mul_fp(x, y.hi); break;
    case 4: z = x / y; break;
    default: z = mkod::sqrt(x); break;
    }
    out[2 * i] = z.hi;
    out[2 * i + 1] = z.lo;
}

} // namespace

namespace mko_internal {

int point_batch(const mko_batch* templ, const int32_t* template_index, int template_per_run,
                int runs, int per_run, const mko_variables* vars_device, const double* x0,
                const double* plus, const double* minus, double* curvature, double* thickness,
                double* index, double* pupil_radius, double* field_angle, double* field_target,
                double* reference_curvature, double* paraxial_index, double* buchdahl,
                void* stream, const int32_t* live) {
    MKO_TRY({
        if (runs <= 0) return 0;
        const ConstBatchPtrs src{templ->curvature, templ->thickness, templ->index,
                                 templ->pupil_radius, templ->field_angle, templ->field_target,
                                 templ->reference_curvature, templ->paraxial_index,
                                 templ->buchdahl};
        const BatchPtrs dst{curvature, thickness, index, pupil_radius, field_angle, field_target,
                            reference_curvature, paraxial_index, buchdahl};
        const OmegaTable om = omega_table(*templ);
        hipLaunchKernelGGL(point_batch_kernel, dim3(unsigned(runs) * unsigned(per_run)), dim3(64),
                           0, static_cast<hipStream_t>(stream), src, dst, runs, per_run,
                           template_index, template_per_run, templ->surfaces, templ->wavelengths,
                           templ->fields, vars_device->count, vars_device->type,
                           vars_device->batch_surface, vars_device->component, x0, plus, minus,
                           live, om);
        check_hip(hipGetLastError(), "point batch launch");
    })
}

template <int kNW, int kPMax, bool kStates, bool kGlass>
void launch_fused(const TraceParams& P, const ParamMap& m, float* partC, double* partST,
                  TupleNormal* partT, const RayStates& states, hipStream_t s) {
    const size_t tuples = size_t(P.V) * P.T;
    hipLaunchKernelGGL((adjoint_normal_kernel<kNW, kPMax, kStates, kGlass>), dim3(unsigned(tuples)),
                       dim3(64 * kNW), fused_lds_bytes<kNW>(m.P), s, P, m, partC, partST, partT,
                       states);
}

template <bool kStates, bool kGlass>
void launch_fused_waves(int waves, const TraceParams& P, const ParamMap& m, float* partC,
                        double* partST, TupleNormal* partT, const RayStates& states,
                        hipStream_t s) {
    const bool wide = m.P > 64;
#define MKO_FUSED(NW, PM) launch_fused<NW, PM, kStates, kGlass>(P, m, partC, partST, partT, states, s)
    switch (waves) {
    case 1: wide ? MKO_FUSED(1, 96) : MKO_FUSED(1, 64); break;
    case 2: wide ? MKO_FUSED(2, 96) : MKO_FUSED(2, 64); break;
    default: wide ? MKO_FUSED(4, 96) : MKO_FUSED(4, 64); break;
    }
#undef MKO_FUSED
}

// Waves per workgroup of the fused kernel: 1 for tuples of at most 64 rays (one chunk, tiles
// formed one after another), 2 up to 128, else 4 (chunks of 256 rays); `requested` (1, 2, 4)
// overrides where it fits (1 needs tuples of at most 64 rays).
int fused_waves(const TraceParams& P, int requested) {
    int nmax = 0;
    for (int f = 0; f < P.F; ++f)
        nmax = std::max(nmax, P.rings * (P.pupil_mode[f] == MKO_PUPIL_AXIS ? 1 : P.spokes));
    if (requested == 2 || requested == 4) return requested;
    if (requested == 1 && nmax <= 64) return 1;
    return nmax <= 64 ? 1 : nmax <= 128 ? 2 : 4;
}

int adjoint_normal_states(const mko_batch* batch, const mko_residual_settings* settings,
                          const mko_variables* vars, double* a, double* g, void* workspace,
                          size_t workspace_bytes, const RayStates& states, void* stream,
                          const int32_t* live, int waves) {
    MKO_TRY({
        TraceParams P = make_params(*batch, *settings);
        P.live = live;
        // glass variables (index adjoints in the reverse sweep) on the LM path's stored states
        const ParamMap m = make_map(*vars, *batch, states.data != nullptr);
        if (P.soft) throw std::invalid_argument("fused adjoint normal equations need hard failure mode");
        if (m.P > kMaxVariablesFused)
            throw std::invalid_argument("fused adjoint normal equations: at most 96 variables");
        if (P.V == 0) return 0;
        if (int64_t(P.V) * P.T > 0x7fffffffLL) throw std::invalid_argument("batch too large");
        const size_t need = mko_adjoint_normal_workspace(P.V, P.F, P.L, m.P);
        if (workspace_bytes < need) throw std::invalid_argument("workspace too small");
        const size_t tuples = size_t(P.V) * P.T;
        auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
        char* w = static_cast<char*>(workspace);
        float* partC = reinterpret_cast<float*>(w);
        double* partST = reinterpret_cast<double*>(w + up(tuples * m.P * m.P * sizeof(float)));
        TupleNormal* partT = reinterpret_cast<TupleNormal*>(
            w + up(tuples * m.P * m.P * sizeof(float)) +
            up(tuples * kFusedSlots * 2 * m.P * sizeof(double)));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int nw = fused_waves(P, waves ? waves : g_tuning.fused_waves);
        if (states.data) {
            if (P.precision != MKO_PRECISION_FP64)
                throw std::invalid_argument("stored ray states need the FP64 trace");
            if (states.rays != P.rays) throw std::invalid_argument("ray states of another layout");
            if (m.glass) launch_fused_waves<true, true>(nw, P, m, partC, partST, partT, states, s);
            else launch_fused_waves<true, false>(nw, P, m, partC, partST, partT, states, s);
        } else {
            launch_fused_waves<false, false>(nw, P, m, partC, partST, partT, states, s);
        }
        check_hip(hipGetLastError(), "adjoint normal launch");
        const int nu = m.P * (m.P + 1) / 2;
        const auto finalize = nu <= 5 * 256    ? adjoint_normal_finalize_kernel<5>
                              : nu <= 10 * 256 ? adjoint_normal_finalize_kernel<10>
                                               : adjoint_normal_finalize_kernel<kFinalizeEntries>;
        hipLaunchKernelGGL(finalize, dim3(unsigned(P.V)), dim3(256), 0, s, P.T, m.P, nw, partC,
                           partST, partT, a, g, live);
        check_hip(hipGetLastError(), "adjoint normal finalize launch");
    })
}

int merit_states(const mko_batch* batch, const mko_residual_settings* settings, double* merit,
                 int64_t* failed, int64_t* hard_failed, double* partial, int* partial_counts,
                 const RayStates& states, void* stream, const int32_t* live) {
    MKO_TRY({
        TraceParams P = make_params(*batch, *settings);
        P.live = live;
        if (P.precision != MKO_PRECISION_FP64)
            throw std::invalid_argument("ray states need the FP64 trace");
        if (states.rays != P.rays) throw std::invalid_argument("ray states of another layout");
        if (P.V == 0) return 0;
        const hipStream_t s = static_cast<hipStream_t>(stream);
        int* pf = partial_counts;
        int* ph = partial_counts + size_t(P.V) * P.T;
        launch_trace(1, P, nullptr, nullptr, nullptr, partial, pf, ph, nullptr, s, &states);
        hipLaunchKernelGGL(merit_finalize_kernel, dim3((P.V + 255) / 256), dim3(256), 0, s, P.V,
                           P.T, partial, pf, ph, nullptr, nullptr, merit, failed, hard_failed, live);
        check_hip(hipGetLastError(), "merit finalize launch");
    })
}

int trace_extrema(const mko_batch* batch, const mko_residual_settings* settings, int mode,
                  double* residuals, int64_t stride, double* opd, uint8_t* status,
                  double* partial, int* partial_counts, const RayStates* states,
                  double* ext_partial, double* extrema, uint64_t surfaces, void* stream) {
    MKO_TRY({
        TraceParams P = make_params(*batch, *settings);
        P.ext_mask = surfaces;
        if (mode == 0) {
            if (stride < P.trace_rows) throw std::invalid_argument("stride smaller than trace rows");
            P.stride = stride;
        }
        if (states && states->rays != P.rays) throw std::invalid_argument("ray states of another layout");
        if (P.V == 0) return 0;
        const hipStream_t s = static_cast<hipStream_t>(stream);
        int* pf = mode == 1 ? partial_counts : nullptr;
        int* ph = mode == 1 ? partial_counts + size_t(P.V) * P.T : nullptr;
        launch_trace(mode, P, residuals, opd, status, partial, pf, ph, ext_partial, s, states);
        if (ext_partial) {
            const int64_t n = int64_t(P.V) * P.S;
            hipLaunchKernelGGL(extrema_finalize_kernel, dim3(unsigned((n + 255) / 256)), dim3(256),
                               0, s, P.V, P.T, P.S, ext_partial, extrema);
            check_hip(hipGetLastError(), "extrema finalize launch");
        }
    })
}

int candidate_extrema(const mko_batch* base, const mko_batch* fd,
                      const mko_residual_settings* settings, int per_run, uint64_t hmask,
                      uint64_t cmask, double delta, const double* ex0, uint8_t* flags,
                      int* tcount, int* others, int* tstart, int* list, int cap, int* overflow,
                      double* ext, void* stream) {
    MKO_TRY({
        TraceParams Pb = make_params(*base, *settings);
        TraceParams Pf = make_params(*fd, *settings);
        if (settings->precision != MKO_PRECISION_FP64)
            throw std::invalid_argument("candidate rays trace in FP64");
        if (Pf.V != Pb.V * per_run || Pf.T != Pb.T || Pf.rays != Pb.rays)
            throw std::invalid_argument("candidate rays: batch shapes differ");
        if (Pb.V == 0) return 0;
        Pb.ext_mask = Pf.ext_mask = hmask | cmask;
        const hipStream_t s = static_cast<hipStream_t>(stream);
        check_hip(hipMemsetAsync(others, 0, size_t(Pb.V) * sizeof(int), s), "memset");
        const dim3 gb(unsigned(Pb.V) * unsigned(Pb.T));
        if (Pb.soft)
            hipLaunchKernelGGL(ray_candidates_kernel<true>, gb, dim3(kBlock), 0, s, Pb, hmask,
                               cmask, delta, ex0, flags, tcount, others);
        else
            hipLaunchKernelGGL(ray_candidates_kernel<false>, gb, dim3(kBlock), 0, s, Pb, hmask,
                               cmask, delta, ex0, flags, tcount, others);
        check_hip(hipGetLastError(), "ray candidates launch");
        hipLaunchKernelGGL(candidate_list_kernel, dim3(unsigned(Pb.V)), dim3(1024), 0, s, Pb.T,
                           Pb.rays, cap, flags, tcount, tstart, list, overflow);
        check_hip(hipGetLastError(), "candidate list launch");
        if (Pf.soft)
            hipLaunchKernelGGL(candidate_extrema_kernel<true>, dim3(unsigned(Pf.V)), dim3(64), 0,
                               s, Pf, per_run, cap, tstart, list, ext);
        else
            hipLaunchKernelGGL(candidate_extrema_kernel<false>, dim3(unsigned(Pf.V)), dim3(64), 0,
                               s, Pf, per_run, cap, tstart, list, ext);
        check_hip(hipGetLastError(), "candidate extrema launch");
    })
}

int extrema_check(int runs, int per_run, int S, uint64_t hmask, uint64_t cmask, double delta,
                  const double* ex, const double* ex0, const int* others, const int* overflow,
                  double K, int* flag, void* stream) {
    MKO_TRY({
        if (runs <= 0) return 0;
        hipLaunchKernelGGL(extrema_check_kernel, dim3((runs + 127) / 128), dim3(128), 0,
                           static_cast<hipStream_t>(stream), runs, per_run, S, hmask, cmask,
                           delta, ex, ex0, others, overflow, K, flag);
        check_hip(hipGetLastError(), "extrema check launch");
    })
}

int prefilter_stats(const mko_batch* batch, const mko_residual_settings* settings, double boundary,
                    void* workspace, double* merit, double* optical, int64_t* failed,
                    int64_t* uncertain_ok, int64_t* uncertain_failed, void* stream) {
    MKO_TRY({
        TraceParams P = make_params(*batch, *settings);
        if (!(boundary >= 0.0) || !std::isfinite(boundary))
            throw std::invalid_argument("prefilter boundary margin must be finite and >= 0");
        if (P.V == 0) return 0;
        if (int64_t(P.V) * P.T > 0x7fffffffLL) throw std::invalid_argument("batch too large");
        const size_t tuples = size_t(P.V) * P.T;
        char* w = static_cast<char*>(workspace);
        double* pm = reinterpret_cast<double*>(w);
        float* pl = reinterpret_cast<float*>(w + tuples * 8);
        int* pc = reinterpret_cast<int*>(w + tuples * 12);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const dim3 grid{unsigned(tuples)}, block{unsigned(kBlock)};
        const float eta = float(boundary);
        switch (rays_per_lane(P)) {
        case 1: hipLaunchKernelGGL((prefilter_kernel<1>), grid, block, 0, s, P, eta, pm, pl, pc); break;
        case 2: hipLaunchKernelGGL((prefilter_kernel<2>), grid, block, 0, s, P, eta, pm, pl, pc); break;
        case 4: hipLaunchKernelGGL((prefilter_kernel<4>), grid, block, 0, s, P, eta, pm, pl, pc); break;
        default: hipLaunchKernelGGL((prefilter_kernel<8>), grid, block, 0, s, P, eta, pm, pl, pc); break;
        }
        check_hip(hipGetLastError(), "prefilter launch");
        hipLaunchKernelGGL(prefilter_finalize_kernel, dim3((P.V + 255) / 256), dim3(256), 0, s,
                           P.V, P.T, pm, pl, pc, merit, optical, failed, uncertain_ok,
                           uncertain_failed);
        check_hip(hipGetLastError(), "prefilter finalize launch");
    })
}

void set_thread_tuning(const Tuning& t) { g_tuning = t; }
Tuning thread_tuning() { return g_tuning; }

int merit_finalize(int V, int T, const double* partial, const int* partial_counts,
                   const double* extra, const double* traced, double* merit, int64_t* failed,
                   int64_t* hard_failed, void* stream) {
    MKO_TRY({
        if (V <= 0) return 0;
        hipLaunchKernelGGL(merit_finalize_kernel, dim3((V + 255) / 256), dim3(256), 0,
                           static_cast<hipStream_t>(stream), V, T, partial, partial_counts,
                           partial_counts + size_t(V) * T, extra, traced, merit, failed,
                           hard_failed, nullptr);
        check_hip(hipGetLastError(), "merit finalize launch");
    })
}

int constraint_rows(const mko_batch* batch, const mko_constraint* items, int count,
                    int need_paraxial, const double* extrema, double* rows, int64_t stride,
                    int64_t offset, void* stream, const double* glass_hull) {
    MKO_TRY({
        if (batch->variants <= 0 || count <= 0) return 0;
        if (glass_hull && !batch->buchdahl)
            throw std::invalid_argument("glass-hull rows need the batch's Buchdahl coefficients");
        uint64_t mask = 0;
        for (int i = 0; i < batch->surfaces; ++i)
            if (batch->type[i] == MKO_SURFACE_SPHERE) mask |= 1ull << i;
        if (need_paraxial && !batch->paraxial_index)
            throw std::invalid_argument("paraxial constraint rows need the batch's paraxial_index");
        hipLaunchKernelGGL(constraint_rows_kernel, dim3((batch->variants + 127) / 128), dim3(128),
                           0, static_cast<hipStream_t>(stream), batch->variants, batch->surfaces,
                           batch->fields, batch->curvature, batch->thickness,
                           batch->paraxial_index, batch->pupil_radius, batch->field_angle,
                           batch->field_target, mask, items, count, need_paraxial, extrema, rows,
                           stride, offset, batch->buchdahl, glass_hull);
        check_hip(hipGetLastError(), "constraint rows launch");
    })
}

int extra_normal(int runs, int P, int E, const double* Je, const double* re, double* a, double* g,
                 void* stream) {
    MKO_TRY({
        if (runs <= 0 || E <= 0) return 0;
        hipLaunchKernelGGL(extra_normal_kernel, dim3(unsigned(runs)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), P, E, Je, re, a, g);
        check_hip(hipGetLastError(), "extra normal launch");
    })
}

int rows_merit(int V, const double* rows, int64_t stride, int64_t offset, int E, int Et,
               const double* extra, double* e_sum, double* t_sum, void* stream) {
    MKO_TRY({
        if (V <= 0) return 0;
        hipLaunchKernelGGL(rows_merit_kernel, dim3((V + 255) / 256), dim3(256), 0,
                           static_cast<hipStream_t>(stream), V, rows, stride, offset, E, Et,
                           extra, e_sum, t_sum);
        check_hip(hipGetLastError(), "rows merit launch");
    })
}

} // namespace mko_internal

namespace {

// Tuple partials of mko_merit_device_noalloc: [kScratchTuples] doubles (merit partials), then
// 2 x [kScratchTuples] ints (failed / hard-failed counts), 8 MiB in all. A static device array
// of the library: each GPU the library runs on holds its own copy until the process ends, and
// every call on that GPU shares it.
constexpr size_t kScratchTuples = size_t(1) << 19;
constexpr size_t kScratchBytes = kScratchTuples * (sizeof(double) + 2 * sizeof(int));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device. In host code `scratch` names a host-side
// stand-in, never the device memory: the address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device (each GPU's copy has its own).
unsigned char* scratch_address() {
    constexpr int kMaxDevices = 64;
    static std::mutex mutex;
    static unsigned char* address[kMaxDevices] = {};
    int device = 0;
    check_hip(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(mutex);
    if (!address[device]) {
        void* ptr = nullptr;
        check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
        address[device] = static_cast<unsigned char*>(ptr);
    }
    return address[device];
}

} // namespace

extern "C" {

const char* mko_last_error(void) { return mko_detail::last_error.c_str(); }

int mko_residuals_device(const mko_batch* batch, const mko_residual_settings* settings,
                         double* residuals, int64_t stride, double* opd, uint8_t* status,
                         void* stream) {
    MKO_TRY({
        no_rows(*settings);
        TraceParams P = make_params(*batch, *settings);
        if (stride < P.trace_rows) throw std::invalid_argument("stride smaller than trace rows");
        P.stride = stride;
        if (P.V == 0) return 0;
        launch_trace(0, P, residuals, opd, status, nullptr, nullptr, nullptr, nullptr,
                     static_cast<hipStream_t>(stream));
    })
}

int mko_merit_device(const mko_batch* batch, const mko_residual_settings* settings,
                     const double* extra_merit, double* merit, int64_t* failed,
                     int64_t* hard_failed, double* partial, int* partial_counts, void* stream) {
    MKO_TRY({
        no_rows(*settings);
        const TraceParams P = make_params(*batch, *settings);
        if (P.V == 0) return 0;
        const hipStream_t s = static_cast<hipStream_t>(stream);
        int* pf = partial_counts;
        int* ph = partial_counts + size_t(P.V) * P.T;
        launch_trace(1, P, nullptr, nullptr, nullptr, partial, pf, ph, nullptr, s);
        hipLaunchKernelGGL(merit_finalize_kernel, dim3((P.V + 255) / 256), dim3(256), 0, s, P.V,
                           P.T, partial, pf, ph, extra_merit, nullptr, merit, failed, hard_failed,
                           nullptr);
        check_hip(hipGetLastError(), "merit finalize launch");
    })
}

int mko_merit_device_noalloc(const mko_batch* batch, const mko_residual_settings* settings,
                             const double* extra_merit, double* merit, int64_t* failed,
                             int64_t* hard_failed, void* stream) {
    MKO_TRY({
        no_rows(*settings);
        const TraceParams P = make_params(*batch, *settings);
        if (P.V == 0) return 0;
        const size_t tuples = size_t(P.V) * size_t(P.T);
        if (tuples > kScratchTuples)
            throw std::invalid_argument(
                "mko_merit_device_noalloc: " + std::to_string(tuples) +
                " (variant, field, wavelength) tuples, at most " + std::to_string(kScratchTuples) +
                " per call (split the batch, or use mko_merit_device with a workspace)");
        unsigned char* base = scratch_address();
        double* partial = reinterpret_cast<double*>(base);
        int* counts = reinterpret_cast<int*>(base + kScratchTuples * sizeof(double));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        int* pf = counts;
        int* ph = counts + tuples;
        // every tuple's partial and counts are written by its trace block before the finalize
        // reads them, so the previous call's contents need no clearing
        launch_trace(1, P, nullptr, nullptr, nullptr, partial, pf, ph, nullptr, s);
        hipLaunchKernelGGL(merit_finalize_kernel, dim3((P.V + 255) / 256), dim3(256), 0, s, P.V,
                           P.T, partial, pf, ph, extra_merit, nullptr, merit, failed, hard_failed,
                           nullptr);
        check_hip(hipGetLastError(), "merit finalize launch");
    })
}

int mko_point_batch_device(const mko_batch* templ, int template_per_run, int runs, int per_run,
                           const mko_variables* vars_device, const double* x0,
                           const double* plus, const double* minus, double* curvature,
                           double* thickness, double* index, double* pupil_radius,
                           double* field_angle, double* field_target,
                           double* reference_curvature, double* paraxial_index, void* stream) {
    return mko_internal::point_batch(templ, nullptr, template_per_run, runs, per_run, vars_device,
                                     x0, plus, minus, curvature, thickness, index, pupil_radius,
                                     field_angle, field_target, reference_curvature,
                                     paraxial_index, nullptr, stream);
}

int mko_fd_difference_device(const double* fd_residuals, int64_t stride,
                             const uint8_t* fd_status, int64_t rays, int runs, int variables,
                             int64_t rows, const double* step_plus, const double* step_minus,
                             const uint8_t* stencil, double* jacobian, double* r0,
                             int64_t* status_changes, void* stream) {
    MKO_TRY({
        if (runs <= 0) return 0;
        hipLaunchKernelGGL(fd_difference_kernel, dim3(unsigned(runs) * unsigned(variables + 1)),
                           dim3(kBlock), 0, static_cast<hipStream_t>(stream), fd_residuals,
                           stride, fd_status, rays, runs, variables, rows, step_plus, step_minus,
                           stencil, jacobian, r0, status_changes, rows, int64_t(0));
        check_hip(hipGetLastError(), "fd difference launch");
    })
}

int mko_adjoint_jacobian_device(const mko_batch* batch, const mko_residual_settings* settings,
                                const mko_variables* vars, int precision, double* jacobian,
                                double* r0, uint8_t* status, void* stream) {
    MKO_TRY({
        no_rows(*settings);
        if (precision != MKO_PRECISION_FP64 && precision != MKO_PRECISION_MIXED)
            throw std::invalid_argument("unknown precision");
        return mko_internal::adjoint_jacobian(batch, settings, vars, jacobian, r0, status, 0,
                                              stream);
    })
}

} // extern "C"

namespace mko_internal {

int adjoint_jacobian(const mko_batch* batch, const mko_residual_settings* settings,
                     const mko_variables* vars, double* jacobian, double* r0, uint8_t* status,
                     int64_t stride, void* stream, const int32_t* live) {
    MKO_TRY({
        TraceParams P = make_params(*batch, *settings);
        P.live = live;
        if (stride > 0) {
            if (stride < P.trace_rows) throw std::invalid_argument("stride smaller than trace rows");
            P.stride = stride;
        }
        const int precision = MKO_PRECISION_FP64;
        const ParamMap m = make_map(*vars, *batch, true);
        if (P.V == 0) return 0;
        if (int64_t(P.V) * P.T > 0x7fffffffLL) throw std::invalid_argument("batch too large");
        const int rpl = rays_per_lane(P);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        // MKO_PRECISION_MIXED computes the stored Jacobian in FP64 as well: on this hardware
        // the reduced-precision stage is the MFMA normal equations of the fused LM path.
        if (precision != MKO_PRECISION_FP64 && precision != MKO_PRECISION_MIXED)
            throw std::invalid_argument("unknown precision");
        double* Jd = static_cast<double*>(jacobian);
        if (P.soft) launch_adjoint_rpl<true, double, double, double>(rpl, P, m, Jd, r0, status, s);
        else launch_adjoint_rpl<false, double, double, double>(rpl, P, m, Jd, r0, status, s);
        check_hip(hipGetLastError(), "adjoint launch");
    })
}

int gram_chain(const double* jacobian, const double* r0, int runs, int variables, int64_t rows,
               double* a, double* g, void* stream, const int32_t* live) {
    MKO_TRY({
        if (runs <= 0) return 0;
        const int T = (variables + 1 + 63) / 64;
        const dim3 grid(unsigned(runs) * unsigned(T * (T + 1) / 2));
        hipLaunchKernelGGL(gram_kernel, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                           jacobian, r0, runs, variables, rows, a, g, live);
        check_hip(hipGetLastError(), "gram launch");
    })
}

int fd_difference(const double* rows, int64_t stride, const uint8_t* status, int64_t rays,
                  int runs, int variables, int64_t count, const double* step_plus,
                  const double* step_minus, const uint8_t* stencil, double* jacobian, double* r0,
                  int64_t* status_changes, int64_t jstride, int64_t joffset, void* stream) {
    MKO_TRY({
        if (runs <= 0 || count <= 0) return 0;
        hipLaunchKernelGGL(fd_difference_kernel, dim3(unsigned(runs) * unsigned(variables + 1)),
                           dim3(kBlock), 0, static_cast<hipStream_t>(stream), rows, stride,
                           status, rays, runs, variables, count, step_plus, step_minus, stencil,
                           jacobian, r0, status_changes, jstride, joffset);
        check_hip(hipGetLastError(), "fd difference launch");
    })
}

} // namespace mko_internal

extern "C" {

size_t mko_adjoint_normal_workspace(int variants, int fields, int wavelengths, int variables) {
    const size_t tuples = size_t(variants) * size_t(fields) * size_t(wavelengths);
    const size_t P = size_t(variables);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return up(tuples * P * P * sizeof(float)) + up(tuples * kFusedSlots * 2 * P * sizeof(double)) +
           up(tuples * sizeof(TupleNormal));
}

int mko_adjoint_normal_device(const mko_batch* batch, const mko_residual_settings* settings,
                              const mko_variables* vars, double* a, double* g, void* workspace,
                              size_t workspace_bytes, void* stream) {
    if (settings && settings->extra && settings->extra->count > 0) {
        mko_detail::last_error = "constraint rows are evaluated by the mko_engine_* entry points";
        return MKO_ERR_INVALID;
    }
    return mko_internal::adjoint_normal_states(batch, settings, vars, a, g, workspace,
                                               workspace_bytes, mko_internal::RayStates{}, stream);
}

int mko_df64_ops_device(int op, int64_t n, const float* a, const float* b, float* out,
                        void* stream) {
    MKO_TRY({
        if (op < 0 || op > 5) throw std::invalid_argument("df64 op outside [0, 5]");
        if (n < 0) throw std::invalid_argument("negative count");
        if (n == 0) return 0;
        hipLaunchKernelGGL(df64_ops_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), op, n, a, b, out);
        check_hip(hipGetLastError(), "df64 ops launch");
    })
}

int mko_gram_mfma_device(const double* jacobian, const double* r0, int runs, int variables,
                         int64_t rows, double* a, double* g, void* stream) {
    MKO_TRY({
        if (variables < 1 || variables > 127) throw std::invalid_argument("1 ... 127 variables");
        if (rows < 0) throw std::invalid_argument("negative row count");
        if (runs <= 0) return 0;
        const int Q = variables + 1, tg = (Q + 15) / 16;
        const size_t lds = size_t(16 * tg) * 33 * sizeof(double);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (tg <= 4)
            hipLaunchKernelGGL(gram_mfma_kernel<3>, dim3(unsigned(runs)), dim3(256), lds, s,
                               jacobian, r0, variables, rows, a, g);
        else
            hipLaunchKernelGGL(gram_mfma_kernel<9>, dim3(unsigned(runs)), dim3(256), lds, s,
                               jacobian, r0, variables, rows, a, g);
        check_hip(hipGetLastError(), "gram mfma launch");
    })
}

size_t mko_vjp_workspace(int variants, int fields, int wavelengths, int variables) {
    return ((size_t(variants) * size_t(fields) * size_t(wavelengths) * size_t(variables) *
             sizeof(double)) + 255) & ~size_t(255);
}

int mko_vjp_device(const mko_batch* batch, const mko_residual_settings* settings,
                   const mko_variables* vars, const double* grad, int64_t grad_stride,
                   double* out, void* workspace, size_t workspace_bytes, void* stream) {
    MKO_TRY({
        no_rows(*settings);
        if (settings->precision != MKO_PRECISION_FP64)
            throw std::invalid_argument("the VJP traces in FP64 (precision MKO_PRECISION_FP64)");
        const TraceParams P = make_params(*batch, *settings);
        const ParamMap m = make_map(*vars, *batch, true);
        if (grad_stride < P.trace_rows) throw std::invalid_argument("grad_stride smaller than trace rows");
        if (P.V == 0) return 0;
        if (int64_t(P.V) * P.T > 0x7fffffffLL) throw std::invalid_argument("batch too large");
        if (workspace_bytes < mko_vjp_workspace(P.V, P.F, P.L, m.P))
            throw std::invalid_argument("workspace too small");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        double* part = static_cast<double*>(workspace);
        const dim3 grid(unsigned(P.V) * unsigned(P.T));
        const size_t lds = size_t(m.P) * 64 * sizeof(double);
        if (P.soft)
            hipLaunchKernelGGL(vjp_kernel<true>, grid, dim3(64), lds, s, P, m, grad, grad_stride,
                               part);
        else
            hipLaunchKernelGGL(vjp_kernel<false>, grid, dim3(64), lds, s, P, m, grad, grad_stride,
                               part);
        check_hip(hipGetLastError(), "vjp launch");
        const int64_t n = int64_t(P.V) * m.P;
        hipLaunchKernelGGL(vjp_finalize_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                           P.V, P.T, m.P, part, out);
        check_hip(hipGetLastError(), "vjp finalize launch");
    })
}

int mko_gram_device(const double* jacobian, const double* r0, int runs, int variables,
                    int64_t rows, double* a, double* g, void* stream) {
    return mko_internal::gram_chain(jacobian, r0, runs, variables, rows, a, g, stream, nullptr);
}

} // extern "C"
