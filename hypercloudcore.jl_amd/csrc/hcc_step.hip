// The explicit time step on the device: stage combine + outflow, gather (gradient + limiter), flux (reconstruction,
// Rusanov, divergence) with the fused MOOD check, the MOOD fixed point on compacted work lists, the CFL estimate,
// and the stepper object of the C ABI.
//
// Layout. Every stage state lives in node records (position then state, padded to 2/4/8 doubles; Layout<D, M>) and
// the limited gradients in gradient records, so a neighbour is read with a few 16-byte loads from one or two cache
// lines. K_E is stored node-major (M doubles per node). One thread per node walks its SELL-64 neighbour slots in
// slot order; the loads of slot k+1 are issued before slot k is computed (index two slots ahead), so the gather
// latency overlaps the arithmetic of the previous slot. Every per-node sum keeps the reference's operation order.
#include "hcc_device.h"
#include "hcc_mlp.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <type_traits>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

using namespace hcc;

// ============================================================================================================
// record access
// ============================================================================================================
template <int NREAD, int STRIDE>
__device__ __forceinline__ void load_rec(const double* __restrict__ base, int64_t i, double (&r)[NREAD]) {
    const double2* p = reinterpret_cast<const double2*>(base + i * STRIDE);
#pragma unroll
    for (int q = 0; q < (NREAD + 1) / 2; ++q) {
        const double2 v = p[q];
        r[2 * q] = v.x;
        if (2 * q + 1 < NREAD) r[2 * q + 1] = v.y;
    }
}
template <int NW, int STRIDE>
__device__ __forceinline__ void store_rec(double* __restrict__ base, int64_t i, const double (&r)[NW]) {
    double2* p = reinterpret_cast<double2*>(base + i * STRIDE);
#pragma unroll
    for (int q = 0; q < STRIDE / 2; ++q) {
        double2 v;
        v.x = (2 * q < NW) ? r[2 * q] : 0.0;
        v.y = (2 * q + 1 < NW) ? r[2 * q + 1] : 0.0;
        p[q] = v;
    }
}

// get_distance (ParticleGrids.jl:209-219) between two record positions and |d|^2, the operations of the search
template <int D>
__device__ __forceinline__ double rec_geom(const GeomArgs& g, const double* pi, const double* pj, double (&dd)[D]) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double x = pj[d] - pi[d];
        // on a non-periodic axis (Lw = invLw = 0) x - 0 * rint(x * 0) equals x + 0.0 for every finite x (a -0 becomes
        // +0 either way); the wrap costs a multiply, a rint and a multiply-subtract more
        if (g.Lw[d] != 0.0) x = x - g.Lw[d] * rint(x * g.invLw[d]);
        else x = x + 0.0;
        dd[d] = x;
        s = (d == 0) ? x * x : s + x * x;
    }
    return s;
}

// ============================================================================================================
// neighbour sources: records gathered from HBM, or the block's halo staged in LDS (SoA by record component)
// ============================================================================================================
template <int D, int M>
struct HbmSrc {
    using L = Layout<D, M>;
    static constexpr bool kPrefetch = true;     // HBM latency: load slot k+1 while slot k computes
    const double* Y;
    const double* G;
    const int8_t* ord;
    const int32_t* nbr;
    __device__ __forceinline__ int32_t idx(int64_t s) const { return nbr[s]; }
    template <int K>
    __device__ __forceinline__ void head(int32_t j, double (&r)[K]) const { load_rec<K, L::PU>(Y, j, r); }
    __device__ __forceinline__ void grad(int32_t j, double (&g)[L::NG]) const { load_rec<L::NG, L::GS>(G, j, g); }
    __device__ __forceinline__ int order(int32_t j) const { return ord[j]; }
};
// A staged halo node in LDS: [position, state | pad to even | gradient record | pad to even] doubles, 16-byte
// aligned so that a neighbour costs a few ds_read_b128 at one base address; the orders follow as bytes.
constexpr int lds_grad_offset(int NR) { return (NR + 1) & ~1; }
constexpr int lds_stride(int NR, int NG) { return (lds_grad_offset(NR) + NG + 1) & ~1; }
// staged record of the CFL estimate: position and wave speeds, padded to an even count of doubles
template <int D>
constexpr int kCflStride = (2 * D + 1) & ~1;
// uint16 offset of the staged index table: after hm records of st doubles and, with orders, hm bytes; 16-byte aligned
__host__ __device__ constexpr size_t lds_index_offset(int64_t hm, int st, bool orders) {
    return ((size_t)hm * st * 8 + (orders ? (size_t)hm : 0) + 15) / 16 * 8;
}

template <int K>
__device__ __forceinline__ void lds_read(const double* p, double (&r)[K]) {
    const double2* q2 = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int q = 0; q < (K + 1) / 2; ++q) {
        const double2 v = q2[q];
        r[2 * q] = v.x;
        if (2 * q + 1 < K) r[2 * q + 1] = v.y;
    }
}

template <int D, int M, bool GRADS, bool LAM = false>
struct LdsSrc {
    using L = Layout<D, M>;
    // the flux loop loads slot k+1 while slot k computes; the gather's loops, with more waves per SIMD, load the
    // next index ahead and the records at the top of each slot
    static constexpr bool kPrefetch = GRADS;
    static constexpr int GO = lds_grad_offset(L::NR);   // gradient record (GRADS)
    static constexpr int ST = GRADS ? lds_stride(L::NR, L::NG) : lds_stride(L::NR, 0);
    // LAM: the wave speeds of halo node l at lam_sh + l LS (LS: D padded to even, 16-byte entries), an array of their
    // own after the records (a record stride of a power of two bytes makes the irregular neighbour reads conflict in
    // the LDS banks)
    static constexpr int LS = (D + 1) & ~1;
    const double* sh;          // halo node l at sh + l * ST
    const int8_t* ord;         // orders of the halo nodes (GRADS)
    const double* lam_sh;      // wave speeds of the halo nodes (LAM)
    const uint16_t* nl;        // the block's local indices in LDS: SELL slot s at nl[s - s0]
    int64_t s0;
    __device__ __forceinline__ int32_t idx(int64_t s) const { return nl[s - s0]; }
    template <int K>
    __device__ __forceinline__ void head(int32_t l, double (&r)[K]) const { lds_read<K>(sh + l * ST, r); }
    __device__ __forceinline__ void grad(int32_t l, double (&g)[L::NG]) const { lds_read<L::NG>(sh + l * ST + GO, g); }
    __device__ __forceinline__ int order(int32_t l) const { return ord[l]; }
    __device__ __forceinline__ void lam(int32_t l, double (&v)[D]) const { lds_read<D>(lam_sh + l * LS, v); }
};

template <class Src>
__device__ __forceinline__ int32_t src_slot(const Src& src, int64_t base, int k, int cnt) {
    const int kk = k < cnt ? k : cnt - 1;                    // clamped: a redundant load past the last slot
    return src.idx(base + (int64_t)kk * kChunk);
}

// The logical workgroup of the XCD-aware order (see node_index).
__device__ __forceinline__ int64_t logical_block(int64_t nblk) { return xcd_block(blockIdx.x, nblk); }

static_assert(kStage == kBlock, "staged passes launch one workgroup per halo block");

// The step's dt of node i: one per instance on an ensemble grid, whose instances fill whole kStage blocks (inst: the
// instance of each block), else dtp[0]
__device__ __forceinline__ double node_dt(const double* dtp, const uint16_t* inst, int64_t i) {
    return inst ? dtp[inst[i / kStage]] : *dtp;
}

// Copies block lb's local-index table into LDS (nl) and returns its first SELL slot. The block's chunks are
// consecutive and each holds 64 x width slots, so the range is 128-byte aligned and a multiple of 64 entries: it is
// copied 8 indices at a time, up to four loads per thread in flight before the first store (a table of neighbourhoods
// up to 32 wide takes one round). The caller synchronises the workgroup before reading it.
__device__ __forceinline__ int64_t stage_index_table(const GeomArgs& g, uint16_t* nl, int64_t lb) {
    const int64_t cpb = kStage / kChunk, nch = (g.N + kChunk - 1) / kChunk;
    const int64_t s0 = g.chunk_off[lb * cpb], s1 = g.chunk_off[(lb + 1) * cpb < nch ? (lb + 1) * cpb : nch];
    const uint4* isrc = reinterpret_cast<const uint4*>(g.nbr_local + s0);
    uint4* idst = reinterpret_cast<uint4*>(nl);
    const int64_t n16 = (s1 - s0) / 8;
    for (int64_t t0 = threadIdx.x; t0 < n16; t0 += 4 * kStage) {
        const int64_t t1 = t0 + kStage, t2 = t0 + 2 * kStage, t3 = t0 + 3 * kStage;
        const uint4 v0 = isrc[t0], v1 = isrc[t1 < n16 ? t1 : t0], v2 = isrc[t2 < n16 ? t2 : t0],
                    v3 = isrc[t3 < n16 ? t3 : t0];   // entries past the table reload t0
        idst[t0] = v0;
        if (t1 < n16) idst[t1] = v1;
        if (t2 < n16) idst[t2] = v2;
        if (t3 < n16) idst[t3] = v3;
    }
    return s0;
}

// Loads the halo of workgroup lb into LDS (records, with GRADS also gradient records and orders, with LAM the wave
// speeds of the CFL estimate, _get_Lambda) and returns the source the neighbour loops read. Coalesced: the halo is
// sorted by node index.
template <int D, int M, bool GRADS, bool LAM = false, int KIND = 0>
__device__ __forceinline__ LdsSrc<D, M, GRADS, LAM> stage_halo(const StepArgs& a, const double* __restrict__ Y,
                                                               double* sh, int64_t lb) {
    using L = Layout<D, M>;
    using S = LdsSrc<D, M, GRADS, LAM>;
    const int32_t h0 = a.g.halo_off[lb], nh = a.g.halo_off[lb + 1] - h0;
    int8_t* ord = reinterpret_cast<int8_t*>(sh + (size_t)a.g.hmax * S::ST);
    double* lam_sh = sh + (size_t)a.g.hmax * S::ST;
    uint16_t* nl = reinterpret_cast<uint16_t*>(sh) + lds_index_offset(a.g.hmax, S::ST + (LAM ? S::LS : 0), GRADS);
    const int64_t s0 = stage_index_table(a.g, nl, lb);
    // two halo nodes per thread per round, both nodes' loads in flight before the first store (a halo of up to
    // 2 kStage nodes takes one round); a second node past the halo reloads the first
    constexpr int U = 2;
    for (int t0 = threadIdx.x; t0 < nh; t0 += U * kStage) {
        int32_t j[U];
        double r[U][L::NR], g[U][GRADS ? L::NG : 1];
        int8_t o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = a.g.halo_idx[h0 + (t0 + u * kStage < nh ? t0 + u * kStage : t0)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            load_rec<L::NR, L::PU>(Y, j[u], r[u]);
            if constexpr (GRADS) {
                load_rec<L::NG, L::GS>(a.G, j[u], g[u]);
                o[u] = a.orders[j[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * kStage;
            if (t >= nh) continue;
            double* dst = sh + t * S::ST;
#pragma unroll
            for (int q = 0; q < L::NR; ++q) dst[q] = r[u][q];
            if constexpr (GRADS) {
#pragma unroll
                for (int q = 0; q < L::NG; ++q) dst[S::GO + q] = g[u][q];
                ord[t] = o[u];
            }
            if constexpr (LAM) {
                double f[M], l[D];
#pragma unroll
                for (int m = 0; m < M; ++m) f[m] = r[u][D + m];
                Pde<D, M, KIND>::maxeig(a.pde, f, l);
#pragma unroll
                for (int d = 0; d < D; ++d) lam_sh[t * S::LS + d] = l[d];
            }
        }
    }
    __syncthreads();
    return S{sh, ord, lam_sh, nl, s0};
}

// ============================================================================================================
// stage combine + outflow (_stage_predictor!, TimeSteppers.jl:400-421; OutflowBC, BoundaryConditions.jl:82-131)
// ============================================================================================================
// A boundary particle with an outflow donor takes the stage value of its interior root (the same combination,
// bit-identical); other boundary particles keep pg.rhos (= U^n during the step). Records keep each node's position.
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_predict(StepArgs a, int64_t nblk, const int32_t* __restrict__ root,
                                                     const double* __restrict__ Un, const double* K0, const double* K1,
                                                     const double* K2, const double* K3, const double* K4, double a0,
                                                     double a1, double a2, double a3, double a4, int nterm, unsigned use,
                                                     double* __restrict__ target) {
    using L = Layout<D, M>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    if (i >= N) return;
    // Only the 16-byte pairs of the record that hold state components are read and written: every state buffer
    // keeps the node positions written at set-up (a pair shared by the last position and the first state component
    // is rewritten with the same position).
    constexpr int P0 = D / 2, P1 = (D + M - 1) / 2, NP = P1 - P0 + 1;
    const bool ghost = a.g.isb[i];
    const int64_t src = ghost ? (int64_t)root[i] : i;
    double r[2 * NP];
    const double2* own = reinterpret_cast<const double2*>(Un + i * L::PU) + P0;
#pragma unroll
    for (int q = 0; q < NP; ++q) { const double2 v = own[q]; r[2 * q] = v.x; r[2 * q + 1] = v.y; }
    if (src >= 0) {
        double sv[2 * NP];
        if (ghost) {
            const double2* sp = reinterpret_cast<const double2*>(Un + src * L::PU) + P0;
#pragma unroll
            for (int q = 0; q < NP; ++q) { const double2 v = sp[q]; sv[2 * q] = v.x; sv[2 * q + 1] = v.y; }
        } else {
#pragma unroll
            for (int q = 0; q < 2 * NP; ++q) sv[q] = r[q];
        }
        // terms unrolled with compile-time indices: the K rows are plain global loads (a runtime-indexed pointer
        // array would be demoted to flat loads, one per component and term, each waited for)
        const double* Ks[kMaxStages] = {K0, K1, K2, K3, K4};
        const double dt = node_dt(a.dtp, a.inst, i);
        const double cs[kMaxStages] = {dt * a0, dt * a1, dt * a2, dt * a3, dt * a4};
        double y[M];
#pragma unroll
        for (int m = 0; m < M; ++m) y[m] = sv[D + m - 2 * P0];
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j) {
            if (j < nterm && ((use >> j) & 1u)) {
                double kv[M];
                if constexpr (M % 2 == 0) load_rec<M, M>(Ks[j], src, kv);
                else {
#pragma unroll
                    for (int m = 0; m < M; ++m) kv[m] = Ks[j][src * M + m];
                }
#pragma unroll
                for (int m = 0; m < M; ++m) y[m] = y[m] + cs[j] * kv[m];
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) r[D + m - 2 * P0] = y[m];
    }
    double2* dst = reinterpret_cast<double2*>(target + i * L::PU) + P0;
#pragma unroll
    for (int q = 0; q < NP; ++q) dst[q] = make_double2(r[2 * q], r[2 * q + 1]);
}

// The stage of a tableau with an implicit part or an implicit source (form_candidate!, TimeSteppers.jl:365-397;
// _stage_predictor!, :400-421; _stage_implicit_step! / _stage_implicit_rhs!, :427-444): Y = U^n + sum_j
// (dt a~_ij) K_E[j] + (dt a_ij) K_I[j] (each term for a non-zero weight); for a stage with the implicit decay source
// the diagonal solve U_out = U_in / (1 + (dt a_ii) rate) and K_I = (U_out - U_in) inv(dt a_ii) if |a_ii| > 1e-14,
// else K_I = -rate Y; K_I = 0 on boundary particles and without an implicit source. Outflow ghosts take their root's
// value (its combination and solve, the same bits); other boundary particles keep pg.rhos.
struct ImexTerms {
    const double* KE[kMaxStages];
    const double* KI[kMaxStages];
    double at[kMaxStages], ai[kMaxStages];   // explicit and implicit weights (a~_ij, a_ij, or b~_j, b_j)
    int nterm;
    double aii;                              // the stage's diagonal implicit weight
    int stage;                               // 1: stage i (K_I[i] formed), 0: the final state
    double* KIout;
};

template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_predict_imex(StepArgs a, int64_t nblk, const int32_t* __restrict__ root,
                                                          const double* __restrict__ Un, ImexTerms t,
                                                          double* __restrict__ target) {
    using L = Layout<D, M>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    if (i >= N) return;
    constexpr int P0 = D / 2, P1 = (D + M - 1) / 2, NP = P1 - P0 + 1;
    const bool ghost = a.g.isb[i];
    const int64_t src = ghost ? (int64_t)root[i] : i;
    double r[2 * NP], ki[M];
#pragma unroll
    for (int m = 0; m < M; ++m) ki[m] = 0.0;
    const double2* own = reinterpret_cast<const double2*>(Un + i * L::PU) + P0;
#pragma unroll
    for (int q = 0; q < NP; ++q) { const double2 v = own[q]; r[2 * q] = v.x; r[2 * q + 1] = v.y; }
    if (src >= 0) {
        double y[M];
        if (ghost) {
            double sv[L::NR];
            load_rec<L::NR, L::PU>(Un, src, sv);
#pragma unroll
            for (int m = 0; m < M; ++m) y[m] = sv[D + m];
        } else {
#pragma unroll
            for (int m = 0; m < M; ++m) y[m] = r[D + m - 2 * P0];
        }
        const double dt = node_dt(a.dtp, a.inst, i);
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j) {
            if (j < t.nterm) {
                if (t.at[j] != 0.0) {
                    const double c = dt * t.at[j];
#pragma unroll
                    for (int m = 0; m < M; ++m) y[m] = y[m] + c * t.KE[j][src * M + m];
                }
                if (t.ai[j] != 0.0) {
                    const double c = dt * t.ai[j];
#pragma unroll
                    for (int m = 0; m < M; ++m) y[m] = y[m] + c * t.KI[j][src * M + m];
                }
            }
        }
        if (t.stage && a.src_mode == HCC_SRC_DECAY_IMPLICIT) {
            const double lam = a.src_rate[src];
            if (fabs(t.aii) > 1e-14) {
                const double den = 1.0 + (dt * t.aii) * lam, inv = 1.0 / (dt * t.aii);
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const double uo = y[m] / den;
                    ki[m] = (uo - y[m]) * inv;
                    y[m] = uo;
                }
            } else {
#pragma unroll
                for (int m = 0; m < M; ++m) ki[m] = (-lam) * y[m];
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) r[D + m - 2 * P0] = y[m];
    }
    double2* dst = reinterpret_cast<double2*>(target + i * L::PU) + P0;
#pragma unroll
    for (int q = 0; q < NP; ++q) dst[q] = make_double2(r[2 * q], r[2 * q + 1]);
    if (t.stage)
#pragma unroll
        for (int m = 0; m < M; ++m) t.KIout[i * M + m] = ghost ? 0.0 : ki[m];
}

// ============================================================================================================
// gather pass: update_content! (MUSCL.jl:351-384), cached reconstruction fit (Interpolators.jl:585-602),
// Venkatakrishnan / Barth-Jespersen limiter (Limiter.jl:61-143)
// ============================================================================================================
// EXT: also store the local extrema of Y over {i} and its neighbours (the limiter's bounds, which are MOODu1's DMP
// bounds too, _dmp_fails MOOD.jl:175-184) for the checks of the flux pass of the same stage. The neighbour loop then
// runs for every interior particle; one without a gradient fit discards the fit sums.
// CFL: also the node's CFL step of get_time_step (the sums of k_cfl over the same slots, with the wave speeds staged
// in LDS), into *dtc (infinity where k_cfl skips the node).
// The 16 indicator features of a 2-D Euler node from its gather (build-defined): per component m with the scale s_m
// (|rho|, |E|, |rho| (|u| + c) for the momenta) the neighbourhood range (max - min) / s_m, the skew
// (max + min - 2 U_m) / s_m and the limited gradient norm h |grad U_m| / s_m; the Mach number |u| / c, p / E, the
// smallest limiter factor and the density's. Stored as bf16.
template <int D, int M, int KIND>
constexpr bool kFeat = KIND == HCC_PDE_EULER && D == 2 && M == 4;

template <int D, int M>
__device__ __forceinline__ void store_features(const StepArgs& a, int64_t i, const double* u, const double* umax,
                                               const double* umin, const double* g, const double* phi) {
    static_assert(M == 4 && D == 2, "2-D Euler features");
    const double rho = u[0], E = u[3];
    const double ux = u[1] / rho, uy = u[2] / rho;
    const double p = (a.pde.gamma - 1.0) * (E - 0.5 * (u[1] * ux + u[2] * uy));
    const double c = p > 0.0 && rho > 0.0 ? sqrt(a.pde.gamma * p / rho) : 0.0;
    const double speed = sqrt(ux * ux + uy * uy);
    const double tiny = 1e-300;
    const double sc[4] = {fmax(fabs(rho), tiny), fmax(fabs(rho) * (speed + c), tiny), fmax(fabs(rho) * (speed + c), tiny),
                          fmax(fabs(E), tiny)};
    const double h = 1.0 / a.g.invL;
    float f[16];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f[m] = (float)((umax[m] - umin[m]) / sc[m]);
        f[4 + m] = (float)((umax[m] + umin[m] - 2.0 * u[m]) / sc[m]);
        f[8 + m] = (float)(h * sqrt(g[m] * g[m] + g[M + m] * g[M + m]) / sc[m]);
    }
    f[12] = (float)(c > 0.0 ? speed / c : 10.0);
    f[13] = (float)(p / fmax(fabs(E), tiny));
    f[14] = (float)fmin(fmin(phi[0], phi[1]), fmin(phi[2], phi[3]));
    f[15] = (float)phi[0];
    __bf16* dst = reinterpret_cast<__bf16*>(a.feat) + i * 16;
#pragma unroll
    for (int q = 0; q < 16; ++q) dst[q] = (__bf16)(isfinite(f[q]) ? fminf(fmaxf(f[q], -1e4f), 1e4f) : 1e4f);
}

template <int D, int M, int KIND, bool EXT, bool CFL = false, bool CUT = false, bool FEAT = false, class Src>
__device__ __forceinline__ void gather_node(const StepArgs& a, const double* __restrict__ Y, const Src& src, int64_t i,
                                            double* dtc = nullptr, bool* cand = nullptr) {
    using L = Layout<D, M>;
    constexpr int NR = L::NR;
    const int64_t N = a.g.N;
    const int cnt = a.g.cnt[i];
    int order = a.orders[i];
    if (order > 1 && cnt < D) { order = 1; a.orders[i] = 1; }   // starved neighbourhood (basis_count(D, 1) = D)
    double g[L::NG];
#pragma unroll
    for (int q = 0; q < L::NG; ++q) g[q] = 0.0;
    const bool fit = order == 2 && a.okg[i];
    const int64_t base = sell_base(a.g, i);
    const double invL = a.g.invL;
    double ri[NR], umax[M], umin[M], b[D][M];
    constexpr int BD = D + 1;
    double X[CFL ? D : 1][BD], li[D], sum_c = 0.0;
    double phi_f[M];   // the limiter factors (features)
#pragma unroll
    for (int m = 0; m < M; ++m) phi_f[m] = 1.0;
    bool cnb_all = false;   // every neighbour's state equals U_i (known only where the neighbour loop ran)
    if (fit || (EXT && a.ext_on) || (CFL && cnt > 0) || FEAT) {
        load_rec<NR, L::PU>(Y, i, ri);
        if constexpr (CFL) {
            double f[M];
#pragma unroll
            for (int m = 0; m < M; ++m) f[m] = ri[D + m];
            Pde<D, M, KIND>::maxeig(a.pde, f, li);
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < BD; ++c) X[r][c] = a.invd[(r * BD + c) * N + i];
        }
        bool nanf[M];   // a NaN among the extrema's operands (hw_max / hw_min)
        cnb_all = true;
#pragma unroll
        for (int m = 0; m < M; ++m) { umax[m] = ri[D + m]; umin[m] = ri[D + m]; nanf[m] = ri[D + m] != ri[D + m]; }
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) b[q][m] = 0.0;
        // a node without neighbours reads no slot (src_slot would index slot -1)
        int32_t j1 = cnt > 0 ? src_slot(src, base, Src::kPrefetch ? 1 : 0, cnt) : 0;
        double rn[NR];
        if constexpr (Src::kPrefetch)
            if (cnt > 0) src.template head<NR>(src_slot(src, base, 0, cnt), rn);
        static_assert(!(CFL && Src::kPrefetch), "the fused CFL estimate reads the staged wave speeds");
        for (int k = 0; k < cnt; ++k) {
            double rj[NR];
            int32_t jcur = 0;
            if constexpr (Src::kPrefetch) {
#pragma unroll
                for (int q = 0; q < NR; ++q) rj[q] = rn[q];
                const int32_t jn = j1;
                j1 = src_slot(src, base, k + 2, cnt);
                src.template head<NR>(jn, rn);
            } else {   // the index of slot k+1 (from HBM/L2) loads while slot k computes; records come from LDS
                jcur = j1;
                j1 = src_slot(src, base, k + 1, cnt);
                src.template head<NR>(jcur, rj);
            }
            // Flat slot: a neighbour whose state equals U_i (==, so +0 and -0 count as equal) has df = +-0 in every
            // component, w (p df) = +-0 for the finite positive weight and finite basis, and b + (+-0) = b (the
            // accumulators start at +0 and never hold -0). Its weight and fit update are skipped: the same bits.
            bool flat = true;
#pragma unroll
            for (int m = 0; m < M; ++m) flat = flat && (rj[D + m] == ri[D + m]);
            cnb_all = cnb_all && flat;
            if (CFL || !flat) {
                double dd[D];
                const double d2 = rec_geom<D>(a.g, ri, rj, dd);
                const double w = CUT ? slot_weight(a.g, d2) : mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
                double p[D];
#pragma unroll
                for (int d = 0; d < D; ++d) p[d] = dd[d] * invL;
                if constexpr (CFL) {   // _cfl_weights_body!: sum_d |c_k,d| max(lambda_d(U_i), lambda_d(U_j))
                    double lj[D], v[BD];
                    src.lam(jcur, lj);
                    v[0] = w * 1.0;
#pragma unroll
                    for (int d = 0; d < D; ++d) v[1 + d] = w * p[d];
                    double dv = 0.0;
#pragma unroll
                    for (int r = 0; r < D; ++r) {
                        double acc = X[r][0] * v[0];
#pragma unroll
                        for (int c = 1; c < BD; ++c) acc = acc + X[r][c] * v[c];
                        const double term = fabs(acc * invL) * jmax(li[r], lj[r]);
                        dv = (r == 0) ? term : dv + term;
                    }
                    sum_c = sum_c + dv;
                }
                if (!flat) {
#pragma unroll
                    for (int m = 0; m < M; ++m) {
                        const double df = rj[D + m] - ri[D + m];
#pragma unroll
                        for (int q = 0; q < D; ++q) b[q][m] = b[q][m] + w * (p[q] * df);
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double fj = rj[D + m];
                umax[m] = hw_max(umax[m], fj);
                umin[m] = hw_min(umin[m], fj);
                nanf[m] = nanf[m] | (fj != fj);
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m)   // jmax / jmin over the same operands: the canonical NaN once one was seen
            if (nanf[m]) { umax[m] = __builtin_nan(""); umin[m] = __builtin_nan(""); }
        if (EXT && a.ext_on) {
            double e[2 * M];
#pragma unroll
            for (int m = 0; m < M; ++m) { e[m] = umax[m]; e[M + m] = umin[m]; }
            store_rec<2 * M, 2 * M>(a.ext, i, e);
        }
        if constexpr (CFL) {
            if (sum_c > kFloatMin) *dtc = (1.0 / (double)D) / sum_c;
        }
    }
    if (fit) {
        double inv[D][D];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) inv[r][c] = a.invg[(r * D + c) * N + i];
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double c = inv[q][0] * b[0][m];
#pragma unroll
                for (int r = 1; r < D; ++r) c = c + inv[q][r] * b[r][m];
                g[q * M + m] = c * invL;
            }
        // A zero gradient (every coefficient +-0) reconstructs +-0 at every interface, which no tolerance
        // (16 eps max(...) >= 0) passes: phi stays 1 and g phi = g, so the limiter's slot loop is skipped.
        bool gzero = true;
#pragma unroll
        for (int q = 0; q < L::NG; ++q) gzero = gzero && (g[q] == 0.0);
        if (a.limiter != HCC_LIM_NONE && !gzero) {
            // per component, for either bound (local max when the reconstruction rises, min when it falls): the
            // bound minus U_i and the tolerance 16 eps max(|U_i|, |bound|), the same operations hoisted out of the loop
            double* phi = phi_f;
            double dup[M], ddn[M], tup[M], tdn[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                phi[m] = 1.0;
                const double ui = ri[D + m];
                dup[m] = umax[m] - ui;
                ddn[m] = umin[m] - ui;
                tup[m] = 16 * kEps * jmax(fabs(ui), fabs(umax[m]));
                tdn[m] = 16 * kEps * jmax(fabs(ui), fabs(umin[m]));
            }
            // walks the slots in order and hands each one's interface increments rec[m] = sum_q g[q, m] (d_q / 2)
            // and whether the tolerance limits them
            auto walk = [&](auto&& fn) {
                int32_t k1 = src_slot(src, base, Src::kPrefetch ? 1 : 0, cnt);
                double pn[D];
                if constexpr (Src::kPrefetch) src.template head<D>(src_slot(src, base, 0, cnt), pn);
                for (int k = 0; k < cnt; ++k) {
                    double pj[D];
                    if constexpr (Src::kPrefetch) {
#pragma unroll
                        for (int d = 0; d < D; ++d) pj[d] = pn[d];
                        const int32_t jn = k1;
                        k1 = src_slot(src, base, k + 2, cnt);
                        src.template head<D>(jn, pn);
                    } else {
                        const int32_t j = k1;
                        k1 = src_slot(src, base, k + 1, cnt);
                        src.template head<D>(j, pj);
                    }
                    double dd[D];
                    rec_geom<D>(a.g, ri, pj, dd);
                    double pint[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) pint[d] = 0.5 * dd[d];
                    double rec[M];
                    bool lim[M];
#pragma unroll
                    for (int m = 0; m < M; ++m) {
                        double v = 0.0;
#pragma unroll
                        for (int q = 0; q < D; ++q) v = v + g[q * M + m] * pint[q];
                        rec[m] = v;
                        lim[m] = fabs(v) > (v > 0.0 ? tup[m] : tdn[m]);
                    }
                    fn(rec, lim);
                }
            };
            // phi from the extreme increments of each side (phi_scan_finish): no quotient per slot
            const bool bj = a.limiter == HCC_LIM_BJ;
            PhiScan sc[M];
#pragma unroll
            for (int m = 0; m < M; ++m) phi_scan_init(sc[m]);
            walk([&](const double (&rec)[M], const bool (&lim)[M]) {
#pragma unroll
                for (int m = 0; m < M; ++m)
                    if (lim[m]) phi_scan_add(sc[m], rec[m]);
            });
            bool decided = true;
#pragma unroll
            for (int m = 0; m < M; ++m) decided = phi_scan_finish(sc[m], dup[m], ddn[m], bj, phi[m]) && decided;
            if (!decided) {
                // a near tie of ratios: the reference's fold over every slot. The components are evaluated without
                // branches (a skipped slot's quotient is computed and discarded), so that the division chains of a
                // slot sit in one basic block and interleave.
#pragma unroll
                for (int m = 0; m < M; ++m) phi[m] = 1.0;
                walk([&](const double (&rec)[M], const bool (&lim)[M]) {
                    bool any = false;
#pragma unroll
                    for (int m = 0; m < M; ++m) any = any || lim[m];
                    if (any) {
#pragma unroll
                        for (int m = 0; m < M; ++m) {
                            const double r = (rec[m] > 0.0 ? dup[m] : ddn[m]) / rec[m];
                            const double pk = bj ? bj_phi(r) : vk_phi(r);
                            phi[m] = lim[m] ? jmin(phi[m], pk) : phi[m];
                        }
                    }
                });
            }
#pragma unroll
            for (int q = 0; q < D; ++q)
#pragma unroll
                for (int m = 0; m < M; ++m) g[q * M + m] = g[q * M + m] * phi[m];
        }
    }
    store_rec<L::NG, L::GS>(a.G, i, g);
    if (a.cnb) {   // the flux pass's flat blocks (k_block_flat): a constant neighbourhood and finite speeds and 2 F
        bool ok = cnb_all;
        if (ok) {
            double f[M], F[D][M], lam[D];
#pragma unroll
            for (int m = 0; m < M; ++m) f[m] = ri[D + m];
            Pde<D, M, KIND>::flux_eig(a.pde, f, F, lam);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ok = ok && isfinite(lam[d]);
#pragma unroll
                for (int m = 0; m < M; ++m) ok = ok && isfinite(F[d][m] + F[d][m]);
            }
        }
        a.cnb[i] = ok ? 1 : 0;
    }
    if constexpr (FEAT && kFeat<D, M, KIND>) {   // ri and the extrema were formed above (FEAT enters that block)
        double u[M];
#pragma unroll
        for (int m = 0; m < M; ++m) u[m] = ri[D + m];
        store_features<D, M>(a, i, u, umax, umin, g, phi_f);
        bool varies = false;
#pragma unroll
        for (int m = 0; m < M; ++m) varies = varies || umax[m] > umin[m];
        if (cand) *cand = order == 2 && varies;
    }
}

// Appends the wave's indicator candidates to a.cand (one atomic per wave); every lane of the wave calls it.
__device__ __forceinline__ void append_cand(const StepArgs& a, bool c, int64_t i) {
    const unsigned long long ball = __ballot(c);
    if (!ball || !a.ind_list) return;
    const unsigned lane = threadIdx.x & 63;
    const unsigned leader = (unsigned)(__ffsll((unsigned long long)ball) - 1);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(a.ind_n, (unsigned long long)__popcll(ball));
    base = __shfl(base, (int)leader);
    if (c) {
        const unsigned long long below = ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        a.ind_list[base + __popcll(below)] = (int32_t)i;
    }
}

// Zero-K byte, above the stage bits (bit s: K_s exactly +-0): bit 5, this step at every stage so far the stage state
// stayed (shortcut of fused_predict: every term zero, no zero component) and MOOD's check passed; bit 6, the last
// step's bit 5 (taken at this step's first stage-1 pass); bit 7, every K of the last step exactly +-0 (likewise).
constexpr unsigned kZStableNow = 0x20u, kZStableLast = 0x40u, kZLastStep = 0x80u;

// gather_node for a particle of a flat gather block (k_block_gflat): every neighbour's state equals U_i, so the fit
// accumulators stay +0 and the gradient is the inverse times zero (g = 0 without a fit; the zeros' signs from zfit),
// which no slot limits. The extrema are not stored again: the particle had a constant neighbourhood (cnb) in the state
// of the full gather that last stored its cnb and its extrema, and this state repeats that state's values (every K
// since then is +-0), so the stored extrema are U_i's value in both halves, value-equal to what the neighbour loop
// would fold (only signs of zeros could differ, which the DMP comparisons cannot see). The starved-order rule and the
// gradient store are those of gather_node.
template <int D, int M>
// store: the gradient record is written. A flat gather block of stage s >= 2 (every halo particle had a constant
// neighbourhood at stage s - 1) may leave it: the stage s - 1 gather wrote this particle the same fit of zero
// accumulators at the same order (its slots were all flat), and a MOOD drop in between (2 -> 1) wrote the zero record of
// order 1 (k_mark_gather, k_mark_dense), which is what this gather would store; the indicator never lowers such a
// particle (no candidate). At stage 1 the orders were reset to the maximum, so the record is stored.
__device__ __forceinline__ void gather_node_flat(const StepArgs& a, int64_t i, bool store = true,
                                                 bool keep_stable = false) {
    using L = Layout<D, M>;
    const int64_t N = a.g.N;
    // keep_stable: the stage-1 flat gather of a chained step (the last step's K's all zero on the halo and its last
    // gather's constant neighbourhoods) without the learned indicator, which lowers orders without a gather. A
    // particle that stayed through the whole last step (every K zero, every check passed: never dropped) kept order 2
    // or the starved order 1 there, and each of its gathers wrote, or as a flat gather of a later stage left, the zero
    // fit at that order; the reset order is the same, so is the record, and it is not stored.
    if (store && keep_stable && a.kz) {
        const unsigned old = a.kz[i];
        if ((old & a.kzmask) == a.kzmask && (old & kZStableNow)) store = false;
    }
    const unsigned zf = a.zfit ? a.zfit[i] : 0u;
    const bool starved = a.zfit ? (zf & kZStarved) != 0u : a.g.cnt[i] < D;
    int order = a.orders[i];
    if (order > 1 && starved) { order = 1; a.orders[i] = 1; }   // starved neighbourhood (basis_count(D, 1) = D)
    double g[L::NG];
#pragma unroll
    for (int q = 0; q < L::NG; ++q) g[q] = 0.0;
    if (order == 2 && (zf & kZGrad)) {   // the rows' zeros as the fit below forms them (k_zero_fits)
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) g[q * M + m] = ((zf >> (kZSign0 + q)) & 1u) ? -0.0 : 0.0;
    } else if (order == 2 && a.okg[i]) {
        const double invL = a.g.invL;
#pragma unroll
        for (int q = 0; q < D; ++q) {
            double inv[D];
#pragma unroll
            for (int r = 0; r < D; ++r) inv[r] = a.invg[(q * D + r) * N + i];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double c = inv[0] * 0.0;
#pragma unroll
                for (int r = 1; r < D; ++r) c = c + inv[r] * 0.0;
                g[q * M + m] = c * invL;
            }
        }
    }
    if (store) store_rec<L::NG, L::GS>(a.G, i, g);
}

// The gather pass over all interior particles (full) or the `changed` mask (TimeSteppers.jl:329-334).
// reset (first stage of a step): orders back to the maximum order and the step's MOOD flags and counters cleared
// (TimeSteppers.jl:531-540), each thread for its own node before the gather reads that node's order.
__device__ __forceinline__ void step_reset(const StepArgs& a, int64_t i) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counters[1] = 0; a.counters[2] = 0x7FF0000000000000ull; a.counters[4] = 0;
    }
    if (i < a.g.N) { a.orders[i] = 2; a.trig[i] = 0; }
}

// The CFL steps of a workgroup's nodes: their minimum into bmin[lb] (one store per workgroup, no contended atomics on
// a single address); every thread of the workgroup calls it.
__device__ __forceinline__ void block_min(const StepArgs& a, double dt, int64_t lb) {
    __shared__ double wmin[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) m = fmin(m, wmin[w]);
        if (lb < (a.g.N + kBlock - 1) / kBlock) a.bmin[lb] = m;
    }
}

// CUT: the stage geometry of a moving grid (slot weights 0 beyond R); FEAT: the full pass stores the indicator
// features
template <int D, int M, int KIND, bool CUT, bool FEAT = false>
__global__ __launch_bounds__(kBlock) void k_gather(StepArgs a, const double* __restrict__ Y, int full, int reset,
                                                    int64_t nblk) {
    if (full && blockIdx.x == 0 && threadIdx.x == 0) { a.counters[0] = 0; a.counters[5] = 0; a.counters[8] = 0; }   // lists 0 of the flux pass's fixed point
    const int64_t i = node_index(nblk);
    if (reset) step_reset(a, i);
    bool cand = false;
    if (i < a.g.N && !a.g.isb[i] && (full || a.changed[i])) {
        if (full) gather_node<D, M, KIND, true, false, CUT, FEAT>(a, Y, HbmSrc<D, M>{Y, a.G, a.orders, a.g.nbr}, i, nullptr, &cand);
        else gather_node<D, M, KIND, false, false, CUT>(a, Y, HbmSrc<D, M>{Y, a.G, a.orders, a.g.nbr}, i);
    }
    if (FEAT && full) append_cand(a, cand, i);
}

// The full gather pass with the workgroup's halo staged in LDS (2-D).
// CFL: the CFL estimate of the state being gathered (stage 1 of a step whose state is U) fused in: per-workgroup
// minima into bmin, reduced by k_cfl_final.
template <int D, int M, int KIND, bool CFL, bool FEAT = false>
__global__ __launch_bounds__(kBlock) void k_gather_staged(StepArgs a, const double* __restrict__ Y, int reset,
                                                           int64_t nblk) {
    extern __shared__ double sh[];
    if (blockIdx.x == 0 && threadIdx.x == 0) { a.counters[0] = 0; a.counters[5] = 0; a.counters[8] = 0; }   // lists 0 of the flux pass's fixed point
    const int64_t pos = logical_block(nblk);
    if (pos >= nblk) return;
    // launch order: the stage's flat-block order for the gathers of stages 2 .. s; otherwise that of the last flux
    // pass, whose non-flat blocks are those of a state one stage earlier (a guess at the long workgroups: any order
    // gives the same results)
    const int32_t* ord = a.gflat ? a.gorder : a.border;
    const int64_t lb = ord ? ord[pos] : pos;
    const int64_t i = lb * kBlock + threadIdx.x;
    if (reset) step_reset(a, i);
    // a flat gather block (uniform per workgroup): no halo, no slot loops; with CFL its minimum in bmin is kept. With
    // FEAT (the indicator on, no collection) its particles are no candidates (every neighbourhood is constant), so
    // no feature of theirs is read and none is stored.
    if (a.gflat && a.gflat[lb]) {   // without CFL: stages 2 .. s (the stage-1 gather of a chain runs with CFL)
        if (i < a.g.N && !a.g.isb[i]) gather_node_flat<D, M>(a, i, CFL, CFL && !FEAT);
        return;
    }
    const auto src = stage_halo<D, M, false, CFL, KIND>(a, Y, sh, lb);
    double dtc = INFINITY;
    bool cand = false;
    if (i < a.g.N && !a.g.isb[i]) gather_node<D, M, KIND, true, CFL, false, FEAT>(a, Y, src, i, &dtc, &cand);
    if (FEAT) append_cand(a, cand, i);
    if (CFL) block_min(a, dtc, lb);
}

// The launch order of the staged passes' blocks from the gather's flat flags (gflat, 0 or 1): the non-flat blocks,
// whose workgroups stage a halo and run slot loops, first, then the flat ones, each group in index order. One list a
// step serves its gathers and flux passes (a block whose gather is not flat is, but for a few, one whose flux pass is
// not flat either). The hardware
// hands out workgroups in launch order as CUs free up, so the long workgroups start first and the short flat ones
// fill the gaps at the end (longest first), instead of a long one starting late and finishing alone; the XCD-aware
// mapping then applies to positions in this list. One workgroup: each thread counts a contiguous range of flags
// (16 per load), a scan of the counts (shuffles in each wave, then the four wave totals) places every block. Results
// do not depend on the order (each block writes its own particles).
constexpr uint32_t kOrderEvery = 4;   // uncaptured steps between new launch orders
__global__ __launch_bounds__(kBlock) void k_block_order(const uint8_t* __restrict__ flat, int32_t* __restrict__ order,
                                                        int64_t nblk) {
    __shared__ int32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t per = (((nblk + kBlock - 1) / kBlock) + 15) & ~(int64_t)15;   // a multiple of 16 flags a thread
    const int64_t b0 = min(nblk, (int64_t)threadIdx.x * per), b1 = min(nblk, b0 + per);
    // 16 flags per aligned 16-byte load (b0 is a multiple of 16; flags past nblk read as flat and are dropped)
    auto chunk = [&](int64_t b) {
        uint4 v = make_uint4(0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u);
        if (b + 16 <= nblk) {
            v = *reinterpret_cast<const uint4*>(flat + b);
        } else {
            uint32_t w[4] = {0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u};
            for (int k = 0; b + k < nblk; ++k) {
                w[k >> 2] &= ~(0xffu << (8 * (k & 3)));
                w[k >> 2] |= (uint32_t)flat[b + k] << (8 * (k & 3));
            }
            v = make_uint4(w[0], w[1], w[2], w[3]);
        }
        return v;
    };
    int32_t c = 0;
    for (int64_t b = b0; b < b1; b += 16) {
        const uint4 v = chunk(b);
        c += 16 - __popc(v.x) - __popc(v.y) - __popc(v.z) - __popc(v.w);
    }
    int32_t x = c;   // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        before += w < wave ? wsum[w] : 0;
        total += wsum[w];
    }
    int64_t nf = before + x - c;      // non-flat blocks before this range
    int64_t fl = total + b0 - nf;     // flat blocks go after all non-flat ones
    for (int64_t b = b0; b < b1; b += 16) {
        const uint4 v = chunk(b);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (b + k >= b1) break;
            if ((w[k >> 2] >> (8 * (k & 3))) & 0xffu) order[fl++] = (int32_t)(b + k);
            else order[nf++] = (int32_t)(b + k);
        }
    }
}

// Whether block lb of the next flux pass is flat: every particle of its halo (its own particles and their
// neighbours) is a boundary particle or has cnb from the full gather just run. Then for every particle i of the block
// and every neighbour j: U_j == U_i (i's neighbourhood is constant), g_i = g_j = 0 (a constant neighbourhood fits an
// exactly zero gradient; boundary particles hold zero gradients), and i's speeds and 2 F_i are finite: every slot is
// flat (flat_slot), the fit accumulators stay +0, and the pass skips the block's halo and its slot loops.
__global__ __launch_bounds__(kBlock) void k_block_flat(GeomArgs g, const uint8_t* __restrict__ cnb,
                                                       uint8_t* __restrict__ bflat, int64_t nblk) {
    const int64_t lb = blockIdx.x;
    if (lb >= nblk) return;
    const int32_t h0 = g.halo_off[lb], nh = g.halo_off[lb + 1] - h0;
    int ok = 1;
    for (int t = threadIdx.x; t < nh; t += kBlock) {
        const int32_t j = g.halo_idx[h0 + t];
        ok &= (g.isb[j] | cnb[j]) ? 1 : 0;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) bflat[lb] = (uint8_t)ok;
}

// Whether the full gather of stage i (1-based, i >= 2) of block lb is flat: every particle h of its halo has K_1 ..
// K_{i-1} exactly +-0 (mask: bits 0 .. i-2 of kz) and, unless it is a boundary particle, a constant neighbourhood in
// Y_{i-1} (cnb of the stage i-1 gather). Then Y_i(h) = U^n(h) + sum (+-0) = Y_{i-1}(h) in value for every halo particle,
// so every particle of the block sees a constant neighbourhood in Y_i: the fit accumulators stay +0, the gradient is
// the inverse times zero, no slot is limited, and the extrema are U_i's value.
__global__ __launch_bounds__(kBlock) void k_block_gflat(GeomArgs g, const uint8_t* __restrict__ cnb,
                                                        const uint8_t* __restrict__ kz, unsigned mask,
                                                        uint8_t* __restrict__ gflat, int64_t nblk) {
    const int64_t lb = blockIdx.x;
    if (lb >= nblk) return;
    const int32_t h0 = g.halo_off[lb], nh = g.halo_off[lb + 1] - h0;
    int ok = 1;
    for (int t = threadIdx.x; t < nh; t += kBlock) {
        const int32_t j = g.halo_idx[h0 + t];
        ok &= ((kz[j] & mask) == mask && (g.isb[j] | cnb[j])) ? 1 : 0;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) gflat[lb] = (uint8_t)ok;
}

// What each particle's fits give for all-zero accumulators (geometry only; order-2 flat paths): the divergence
// fit's K before sources, -(2 (acc invL)) with acc = 0 + sum_d (sum_c invd * 0) (-(2.0 * 0.0) without a cached
// fit), and each gradient row (sum_r invg * 0) invL, formed with the operations of fit_divergence and gather_node.
// kZDiv when K is -0 in every component; kZGrad (with each row's sign bit) when every row is a zero: a non-finite
// inverse or invL leaves the bit clear and the flat paths form the fit. kZStarved: fewer than D neighbours (the
// starved-order rule).
template <int D>
__global__ __launch_bounds__(kBlock) void k_zero_fits(const double* __restrict__ invg, const uint8_t* __restrict__ okg,
                                                       const double* __restrict__ invd, const uint8_t* __restrict__ okd,
                                                       const int32_t* __restrict__ cnt, double invL,
                                                       uint8_t* __restrict__ zfit, int64_t N) {
    constexpr int BD = D + 1;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    unsigned z = 0;
    double kd = -(2.0 * 0.0);
    if (okd[i]) {
        double acc = 0.0;
        for (int d = 0; d < D; ++d) {
            double r = invd[(d * BD + 0) * N + i] * 0.0;
            for (int c = 1; c < BD; ++c) r = r + invd[(d * BD + c) * N + i] * 0.0;
            acc = acc + r;
        }
        kd = -(2.0 * (acc * invL));
    }
    if (kd == 0.0 && signbit(kd)) z |= kZDiv;
    bool rows = true;
    for (int q = 0; q < D; ++q) {
        double c = 0.0;
        if (okg[i]) {
            c = invg[(q * D + 0) * N + i] * 0.0;
            for (int r = 1; r < D; ++r) c = c + invg[(q * D + r) * N + i] * 0.0;
            c = c * invL;
        }
        rows = rows && c == 0.0;
        if (signbit(c)) z |= 1u << (kZSign0 + q);
    }
    if (rows) z |= kZGrad;
    if (cnt[i] < D) z |= kZStarved;
    zfit[i] = (uint8_t)z;
}

// The identity order (no block known to be flat yet)
__global__ __launch_bounds__(kBlock) void k_iota(int32_t* __restrict__ order, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) order[i] = (int32_t)i;
}

// The stage predictor fused into the flux passes (form_candidate!, TimeSteppers.jl:365-397; _stage_predictor!,
// :400-421): Y_i(n) = U^n(n) + sum_j (dt a_ij) K_j(n) needs only node n's data, and its last term is the K this pass
// has just computed for n. A recompute of K (MOOD) recomputes Y_i(n) with it. Outflow ghosts take their interior
// root's value: the thread of root n writes the ghosts listed for it (gidx[goff[n] .. goff[n+1])); ghosts without a
// root keep pg.rhos, which never changes, in every state buffer. count: the pass keeps the accepted-inadmissible
// count of the checkpoint (_count_inadmissible, TimeSteppers.jl:481-490) up to date.
struct PredArgs {
    const double* Un;
    const double* K[kMaxStages];
    double at[kMaxStages];     // tableau entries a_ij (b_j for the final state); the coefficient is dt * a_ij
    int nterm;                 // terms; the last one is the K of this pass
    unsigned use;              // terms with a non-zero tableau entry
    double* target;
    const int32_t* goff;
    const int32_t* gidx;
    const double* dtp;
    const uint16_t* inst;      // ensemble grid: the instance of each kStage block (node_dt)
    int on, count;
    // first: the first full flux pass of its stage in a step that continues an unbroken chain of steps (every state
    // buffer holds what the last step left there, and the zero-K bits describe its K's); kmask: the bits of all stages
    int first;
    unsigned kmask;
    int pass1;   // the first full flux pass of its stage (0: a MOOD recompute)
};


// Every term of the stage predictor exactly +-0 (the zero-K bits kzv of the used terms, this pass's K included) and
// the stage state fi = Y_s(n) of the pass free of zeros and finite: Y_s(n) is U^n(n) plus +-0 terms, so its non-zero
// components are U^n's, and so is the new stage state; it is fi itself (layouts whose written pairs hold the state
// only: no position or padding to fill).
template <int D, int M>
__device__ __forceinline__ bool stage_stays(const PredArgs& p, int64_t n, unsigned kzv, const double (&fi)[M]) {
    constexpr int P0 = D / 2, P1 = (D + M - 1) / 2, NP = P1 - P0 + 1, O = D - 2 * P0;
    if constexpr (O == 0 && 2 * NP == M) {
        const double dt = node_dt(p.dtp, p.inst, n);
        bool same = true;
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j)
            if (j < p.nterm && ((p.use >> j) & 1u)) same = same && ((kzv >> j) & 1u) && isfinite(dt * p.at[j]);
#pragma unroll
        for (int m = 0; m < M; ++m) same = same && fi[m] != 0.0 && isfinite(fi[m]);
        return same;
    }
    return false;
}

// same: stage_stays held (the new stage state is fi, and U^n's record is not read)
template <int D, int M>
__device__ __forceinline__ void fused_predict(const PredArgs& p, int64_t n, const double (&Kn)[M],
                                              uint8_t* kz = nullptr, uint8_t kzv = 0, const double* fi = nullptr,
                                              bool same = false) {
    using L = Layout<D, M>;
    constexpr int P0 = D / 2, P1 = (D + M - 1) / 2, NP = P1 - P0 + 1, O = D - 2 * P0;
    double r[2 * NP];
    bool skip_own = false;
    const double dt = node_dt(p.dtp, p.inst, n);
    if (same) {
#pragma unroll
        for (int m = 0; m < M; ++m) r[m] = fi[m];
        // and the last step's K's all exactly zero too: the target, written with U^{n-1}'s values in the last step
        // (or skipped there by this same rule), holds U^{n-1}(n) = U^n(n), these very bits, so nothing is stored
        // (a store of part of a record costs the memory a read-modify-write of the 64 B)
        if (p.first && (kzv & kZLastStep)) {
            skip_own = true;
        }
    } else {
        const double2* own = reinterpret_cast<const double2*>(p.Un + n * L::PU) + P0;
#pragma unroll
        for (int q = 0; q < NP; ++q) { const double2 v = own[q]; r[2 * q] = v.x; r[2 * q + 1] = v.y; }
    }
#pragma unroll
    for (int j = 0; j < kMaxStages && !same; ++j) {
        if (j < p.nterm && ((p.use >> j) & 1u)) {
            const double cj = dt * p.at[j];
            // K_j(n) exactly +-0 in every component (zero-K bit j): the term adds cj (+-0) = +-0, which leaves a
            // component that is not a zero with its bits (x + (+-0) = x), so where no component is a zero the term and
            // the load of its K row are skipped; a zero component takes the term (the sign of the sum depends on it)
            if (kz && j < p.nterm - 1 && ((kzv >> j) & 1u) && isfinite(cj)) {
                bool nz = true;
#pragma unroll
                for (int m = 0; m < M; ++m) nz = nz && (r[O + m] != 0.0);
                if (nz) continue;
            }
            double kv[M];
            if (j == p.nterm - 1) {
#pragma unroll
                for (int m = 0; m < M; ++m) kv[m] = Kn[m];
            } else if constexpr (M % 2 == 0) {
                load_rec<M, M>(p.K[j], n, kv);
            } else {
#pragma unroll
                for (int m = 0; m < M; ++m) kv[m] = p.K[j][n * M + m];
            }
#pragma unroll
            for (int m = 0; m < M; ++m) r[O + m] = r[O + m] + cj * kv[m];
        }
    }
    if (!skip_own) {
        double2* dst = reinterpret_cast<double2*>(p.target + n * L::PU) + P0;
#pragma unroll
        for (int q = 0; q < NP; ++q) dst[q] = make_double2(r[2 * q], r[2 * q + 1]);
    }
    for (int32_t e = p.goff[n], e1 = p.goff[n + 1]; e < e1; ++e) {
        const int64_t g = p.gidx[e];
        if (kz) kz[g] = kzv;   // a ghost's stage states are its root's
        if constexpr (O == 0) {
            double2* gd = reinterpret_cast<double2*>(p.target + g * L::PU) + P0;
#pragma unroll
            for (int q = 0; q < NP; ++q) gd[q] = make_double2(r[2 * q], r[2 * q + 1]);
        } else {                                   // the first pair holds the ghost's own last coordinate
#pragma unroll
            for (int m = 0; m < M; ++m) p.target[g * L::PU + D + m] = r[O + m];
        }
    }
}

// Adds the wavefront's change of the accepted-inadmissible count (delta per lane in -1..1) to counters[1]; one
// atomic per wavefront that has a change (rare).
__device__ __forceinline__ void add_inadmissible(const StepArgs& a, int delta) {
    const unsigned long long up = __ballot(delta > 0), dn = __ballot(delta < 0);
    if ((up | dn) && (threadIdx.x & 63) == (unsigned)(__ffsll((unsigned long long)__ballot(1)) - 1))
        atomicAdd(&a.counters[1], (unsigned long long)((long long)__popcll(up) - (long long)__popcll(dn)));
}
// The same count from per-lane sums a grid-strided kernel kept over all its items: one atomic per wavefront at its
// end (called by every lane of the wavefront). A wavefront that took an atomic per item queued one contended atomic
// on this single counter per item group, which serialised the list kernels when most particles changed.
__device__ __forceinline__ void add_inadmissible_sum(const StepArgs& a, int d) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
    if (d != 0 && (threadIdx.x & 63) == 0) atomicAdd(&a.counters[1], (unsigned long long)(long long)d);
}

// ============================================================================================================
// flux pass: MUSCL functor + _flux_loop (MUSCL.jl:390-508), Rusanov (FluxFunctions.jl:16-76), sort_flux
// (Utils.jl:95-103), cached divergence fit (apply_weights_divergence, Interpolators.jl:610-626), K_E = -2 div
// (TimeSteppers.jl:354).
//
// With CHECK it also runs MOOD's candidate check on the node it just updated (_mood_flag!, MOOD.jl:289-295): the
// forward Euler substep W = Y + h K is node-local (form_substep!, TimeSteppers.jl:501-518), and the DMP bounds of
// MOODu1 (_dmp_fails, MOOD.jl:175-184) are the extrema of Y over the same neighbour loop. The nodes checked are
// exactly those whose K is (re)computed (all interior nodes, or the dirty ones), as in detect_and_drop!. The
// decrement (_mood_drop!) is applied by k_mark_gather: neighbours read this node's order in this very pass.
//
// FV selects the flux loop: 0 the buffered divergence fit with the geometric sorting (RusanovFlux), 1 the other
// MUSCL variants by runtime flags: the streamed divergence (acc += sum_d c_k,d dF_d with the slot's derivative
// weights c_k, _consume with CachedWeights, MUSCL.jl:431-461) and AlignedRusanovFlux (sorted by sign(c_k) where
// either particle is below the full order, _dissipation_direction, MUSCL.jl:338-339). c_k is recomputed from the
// divergence inverse and the slot geometry with the operations of _weights_body! (MUSCL.jl:211-223), the bits of
// the cached dweights (zero where the node has no cached fit).
// ============================================================================================================
// The node's own data for its flux loop: state record, flux, gradients, orders and (FV = 1) the derivative rows of
// its divergence inverse (zero without a cached fit).
template <int D, int M, int KIND, int FV>
struct FluxNode {
    using L = Layout<D, M>;
    double ri[L::NR], fi[M], gi[L::NG], Fi[D][M];
    double X[FV == 1 ? D : 1][D + 1];
    double w[FV == 1 ? D : 1];   // a moving grid: the particle's grid velocity (the frame of its flux)
    int oi, oi_t;   // interface order source, true order
    bool flat_ok;   // FV = 0: the node's wave speeds are finite and 2 F_i does not overflow (flat_slot)
};

template <int D, int M, int KIND, int FV>
__device__ __forceinline__ void load_flux_node(const StepArgs& a, const double* __restrict__ Y, int64_t i,
                                               FluxNode<D, M, KIND, FV>& nd) {
    using L = Layout<D, M>;
    constexpr int BD = D + 1;
    load_rec<L::NR, L::PU>(Y, i, nd.ri);
    load_rec<L::NG, L::GS>(a.G, i, nd.gi);
#pragma unroll
    for (int m = 0; m < M; ++m) nd.fi[m] = nd.ri[D + m];
    nd.flat_ok = false;
    if constexpr (FV == 0) {
        double lam[D];
        Pde<D, M, KIND>::flux_eig(a.pde, nd.fi, nd.Fi, lam);   // the flux of flux(), and the wave speeds
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            ok = ok && isfinite(lam[d]);
#pragma unroll
            for (int m = 0; m < M; ++m) ok = ok && isfinite(nd.Fi[d][m] + nd.Fi[d][m]);
        }
#pragma unroll
        for (int q = 0; q < L::NG; ++q) ok = ok && (nd.gi[q] == 0.0);
        nd.flat_ok = ok;
    } else {
        Pde<D, M, KIND>::flux(a.pde, nd.fi, nd.Fi);
    }
    nd.oi = a.orders[i];
    nd.oi_t = nd.oi;
    if (FV == 1 && a.packed) { nd.oi_t = nd.oi >> 4; nd.oi &= 15; }
    if constexpr (FV == 1) {
        const bool okd = a.okd[i];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < BD; ++c) nd.X[r][c] = okd ? a.invd[(r * BD + c) * a.g.N + i] : 0.0;
        if (a.ale) {   // F_i - w_i U_i
            Pde<D, M, KIND>::grid_velocity(a.pde, nd.fi, a.ale, a.mover_alpha, nd.w);
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int m = 0; m < M; ++m) nd.Fi[d][m] = nd.Fi[d][m] - nd.w[d] * nd.fi[m];
        }
    }
}

// One slot of the flux loop (MUSCL.jl:473-506): the interface states at the interface orders, Rusanov per axis and the
// flux difference dF = F* - F_i, the fit weight w and basis pt = [1, d invL]; with the streamed divergence (FV = 1)
// also its contraction ctr = sum_d c_k,d dF_d.
template <int D, int M, int KIND, int FV>
__device__ __forceinline__ void slot_flux(const StepArgs& a, const FluxNode<D, M, KIND, FV>& nd,
                                          const double (&rj)[Layout<D, M>::NR], const double (&gj)[Layout<D, M>::NG],
                                          int oj, double (&dF)[D][M], double& w, double (&pt)[D + 1], double (&ctr)[M]) {
    using P = Pde<D, M, KIND>;
    constexpr int BD = D + 1;
    const double invL = a.g.invL;
    double dd[D];
    const double d2 = rec_geom<D>(a.g, nd.ri, rj, dd);
    w = FV == 1 ? slot_weight(a.g, d2) : mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
    int oj_t = oj;
    if (FV == 1 && a.packed) { oj_t = oj >> 4; oj &= 15; }
    int o_i = nd.oi, o_j = oj;
    if (a.strategy == HCC_STRAT_EPD1) { o_i = nd.oi < oj ? nd.oi : oj; o_j = o_i; }
    double pii[D], pjj[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        pii[d] = (o_i == 2) ? 0.5 * dd[d] : 0.0;
        pjj[d] = (o_j == 2) ? -0.5 * dd[d] : 0.0;
    }
    double fij[M], fji[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double si = nd.gi[m] * pii[0];
#pragma unroll
        for (int q = 1; q < D; ++q) si = si + nd.gi[q * M + m] * pii[q];
        double sj = gj[m] * pjj[0];
#pragma unroll
        for (int q = 1; q < D; ++q) sj = sj + gj[q * M + m] * pjj[q];
        const double uj = rj[D + m];
        fij[m] = nd.fi[m] + si;
        fji[m] = uj + sj;
    }
    double Fij[D][M], Fji[D][M], lij[D], lji[D];
    bool ale = false;
    if constexpr (FV == 1) {
        if (a.ale) {   // both interface states in the frame of particle i (ALEFrame(eq, w_i))
            ale = true;
            P::flux_eig_ale(a.pde, fij, nd.w, Fij, lij);
            P::flux_eig_ale(a.pde, fji, nd.w, Fji, lji);
        }
    }
    if (!ale) {
        P::flux_eig(a.pde, fij, Fij, lij);
        P::flux_eig(a.pde, fji, Fji, lji);
    }
    pt[0] = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) pt[1 + d] = dd[d] * invL;
    double cw[D];
    bool alignd = false;
    if constexpr (FV == 1) {
        alignd = a.aligned && (nd.oi_t < 2 || oj_t < 2);
        if (a.streamed || alignd) {
            double v[BD];
#pragma unroll
            for (int c = 0; c < BD; ++c) v[c] = w * pt[c];
#pragma unroll
            for (int r = 0; r < D; ++r) {
                double acc = nd.X[r][0] * v[0];
#pragma unroll
                for (int c = 1; c < BD; ++c) acc = acc + nd.X[r][c] * v[c];
                cw[r] = acc * invL;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        // sort_flux orders the two sides by sign(d). Three exact shortcuts: F_L + F_R is the same sum either way
        // (IEEE addition commutes); max(λ_L, λ_R) is symmetric for the non-negative speeds |u| + c (a NaN stays a
        // NaN); f_R - f_L = ±(f_ji - f_ij) up to the sign of a zero difference. The sign of a zero dissipation only
        // reaches dF = F* - F_i when dF is a zero, and the reference's + nc_jump (+0.0) turns such a zero into +0;
        // the fit accumulators start at +0 and can never hold -0 (x + (-0) = x, exact cancellation gives +0), so a
        // zero dF of either sign leaves them bit for bit unchanged, and the + 0.0 is dropped as well. The side's
        // sign goes on the speed: (-s) x and s (-x) are the same IEEE product.
        const double sd = jmax(lij[d], lji[d]);
        const bool right = alignd ? cw[d] > 0.0 : dd[d] > 0.0;
        const double sds = right ? sd : -sd;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const double x = fji[m] - fij[m];
            const double diss = sds * x;
            const double Fnum = 0.5 * (Fij[d][m] + Fji[d][m] - diss);
            dF[d][m] = Fnum - nd.Fi[d][m];
            if (FV == 1 && a.streamed) ctr[m] = d == 0 ? 0.0 + cw[0] * dF[d][m] : ctr[m] + cw[d] * dF[d][m];
        }
    }
}

// A flat slot of the flux loop (FV = 0) contributes exact zeros to the fit, so flux_node skips it. Flat: both
// gradients are zero (every coefficient +-0), the neighbour's state equals U_i (==), and the node's wave speeds are
// finite with 2 F_i finite (nd.flat_ok). Then both interface states equal U_i up to the signs of zeros (U + sum of
// +-0 terms), so F(f_ij), F(f_ji) and F_i agree in every non-zero value (the Euler flux and speeds only multiply and
// add these values; a zero's sign changes no non-zero result, and |u| + c is the same non-negative value), the
// speeds agree, f_R - f_L = +-0 and s (+-0) = +-0. Each component of the Rusanov flux is 0.5 ((F + F) - (+-0)) = F
// exactly (2 F is exact, and finite by flat_ok), or +-0 where F is a zero, so dF = F* - F_i = +-0, w (pt dF) = +-0
// and b + (+-0) = b: the accumulators start at +0 and never hold -0.
template <int D, int M, int KIND, int FV>
__device__ __forceinline__ bool flat_slot(const FluxNode<D, M, KIND, FV>& nd, const double (&rj)[Layout<D, M>::NR],
                                          const double (&gj)[Layout<D, M>::NG]) {
    bool flat = nd.flat_ok;
#pragma unroll
    for (int m = 0; m < M; ++m) flat = flat && (rj[D + m] == nd.fi[m]);
#pragma unroll
    for (int q = 0; q < Layout<D, M>::NG; ++q) flat = flat && (gj[q] == 0.0);
    return flat;
}

// evaluate_sources of the explicit sources added to K_E (TimeSteppers.jl:354): zero(U) without one, the decay source
// -rate_i U_i with one
template <int M>
__device__ __forceinline__ double explicit_source(const StepArgs& a, int64_t i, const double (&fi)[M], int m) {
    return a.src_mode == HCC_SRC_DECAY_EXPLICIT ? (-a.src_rate[i]) * fi[m] : 0.0;
}

// The divergence of the buffered fit (apply_weights_divergence, Interpolators.jl:610-626) from the accumulators of
// the derivative rows, K_E = -2 div + sources; a node without a cached fit gets a zero divergence.
template <int D, int M>
__device__ __forceinline__ void fit_divergence(const StepArgs& a, int64_t i, const double (&b)[D + 1][D * M],
                                               const double (&fi)[M], double (&Kout)[M]) {
    constexpr int BD = D + 1;
    const int64_t N = a.g.N;
    if (a.okd[i]) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double r = a.invd[(d * BD + 0) * N + i] * b[0][d * M + m];
#pragma unroll
                for (int c = 1; c < BD; ++c) r = r + a.invd[(d * BD + c) * N + i] * b[c][d * M + m];
                acc = acc + r;
            }
            Kout[m] = -(2.0 * (acc * a.g.invL)) + explicit_source<M>(a, i, fi, m);
        }
    } else {
#pragma unroll
        for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * 0.0) + explicit_source<M>(a, i, fi, m);
    }
}

// After K_E of node i: the store, the fused stage predictor and (CHECK) MOOD's candidate check with the decrement
// deferred to k_mark_gather. Returns whether the node's order is to drop. CNB: node i has a constant neighbourhood in
// the checked state (cnb), so its stored extrema equal U_i in value and the DMP bounds are taken from fi (the
// comparisons see values only; a NaN U_i has no constant neighbourhood).
template <int D, int M, int KIND, bool CHECK, bool CNB = false>
__device__ __forceinline__ bool flux_tail(const StepArgs& a, const PredArgs& pred, double* __restrict__ K, int64_t i,
                                          const double (&fi)[M], const double (&Kout)[M], double h) {
    using P = Pde<D, M, KIND>;
#pragma unroll
    for (int m = 0; m < M; ++m) K[i * M + m] = Kout[m];
    unsigned v = 0;       // the zero-K byte (stored after the check, which decides bit 5)
    bool same = false;    // stage_stays
    if (pred.on) {
        if (a.kz) {   // bit s of kz: K_s(i) is +-0 (s: this pass's K, the last term of the predictor)
            bool z = true;
#pragma unroll
            for (int m = 0; m < M; ++m) z = z && Kout[m] == 0.0;
            const int st = pred.nterm - 1;
            const unsigned old = a.kz[i];
            if (st == 0) {   // bits 6 and 7 from the last step's byte before this step's first pass overwrites it
                const unsigned last = pred.first ? ((((old & pred.kmask) == pred.kmask) ? kZLastStep : 0u) |
                                                    ((old & kZStableNow) ? kZStableLast : 0u))
                                                 : (pred.pass1 ? 0u : (old & (kZLastStep | kZStableLast)));
                v = (unsigned)z | last | (pred.pass1 ? kZStableNow : (old & kZStableNow));
            } else {
                v = (old & ~(1u << st)) | ((unsigned)z << st);
            }
            same = stage_stays<D, M>(pred, i, v, fi);
        }
        fused_predict<D, M>(pred, i, Kout, a.kz, (uint8_t)v, fi, same);
    }
    // MOOD's check of the Euler substep
    bool flag = false, adm = true;
    if constexpr (CHECK) {
        double W[M];
#pragma unroll
        for (int m = 0; m < M; ++m) W[m] = fi[m] + h * Kout[m];
        adm = P::admissible(a.pde, W);
        a.inadm[i] = !adm;
        flag = !adm;
        if (!flag) {
            if (a.criterion == HCC_CRIT_ONLY) {
                flag = true;
            } else if (a.criterion == HCC_CRIT_U1) {
                // the DMP bounds: extrema of Y_{i-1} over {i} and its neighbours, stored by this stage's full gather
                // pass (the limiter's local extrema over the same set; the fused check only runs with them stored)
                double mx[M], mn[M];
                if constexpr (CNB) {
#pragma unroll
                    for (int m = 0; m < M; ++m) { mx[m] = fi[m]; mn[m] = fi[m]; }
                } else {
                    double e[2 * M];
                    load_rec<2 * M, 2 * M>(a.ext, i, e);
#pragma unroll
                    for (int m = 0; m < M; ++m) { mx[m] = e[m]; mn[m] = e[M + m]; }
                }
                const double dl = a.delta, d3 = dl * dl * dl;
#pragma unroll
                for (int m = 0; m < M; ++m)
                    if (fabs(mx[m] - mn[m]) >= d3 && (W[m] < mn[m] - dl || W[m] > mx[m] + dl)) flag = true;
            }
        }
    }
    if (pred.on && a.kz) {
        if (!(same && (!CHECK || (adm && !flag)))) v &= ~kZStableNow;
        a.kz[i] = (uint8_t)v;
    }
    if (flag && a.orders[i] > 1) { a.trig[i] = 1; return true; }
    return false;
}

template <int D, int M, int KIND, bool CHECK, int FV, class Src>
__device__ __forceinline__ bool flux_node(const StepArgs& a, const PredArgs& pred, const double* __restrict__ Y,
                                          const Src& src, double* __restrict__ K, int64_t i, double h) {
    using L = Layout<D, M>;
    constexpr int NR = L::NR, NG = L::NG, BD = D + 1;
    const int cnt = a.g.cnt[i];
    FluxNode<D, M, KIND, FV> nd;
    double Kout[M];
    if (cnt == 0) {
        load_rec<NR, L::PU>(Y, i, nd.ri);
#pragma unroll
        for (int m = 0; m < M; ++m) nd.fi[m] = nd.ri[D + m];
#pragma unroll
        for (int m = 0; m < M; ++m) Kout[m] = -0.0 + explicit_source<M>(a, i, nd.fi, m);
    } else {
        const int64_t base = sell_base(a.g, i);
        load_flux_node<D, M, KIND, FV>(a, Y, i, nd);
        double b[BD][D * M], sacc[M];
#pragma unroll
        for (int c = 0; c < BD; ++c)
#pragma unroll
            for (int q = 0; q < D * M; ++q) b[c][q] = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) sacc[m] = 0.0;
        auto slot = [&](const double (&rj)[NR], const double (&gj)[NG], int oj) {
            if constexpr (FV == 0)
                if (flat_slot<D, M, KIND, FV>(nd, rj, gj)) return;
            double dF[D][M], w, pt[BD], ctr[M];
            slot_flux<D, M, KIND, FV>(a, nd, rj, gj, oj, dF, w, pt, ctr);
            if (FV == 1 && a.streamed) {
#pragma unroll
                for (int m = 0; m < M; ++m) sacc[m] = sacc[m] + ctr[m];
            } else {
#pragma unroll
                for (int d = 0; d < D; ++d)
#pragma unroll
                    for (int m = 0; m < M; ++m)
#pragma unroll
                        for (int c = 0; c < BD; ++c) b[c][d * M + m] = b[c][d * M + m] + w * (pt[c] * dF[d][m]);
            }
        };
        if constexpr (Src::kPrefetch) {
            // slot k+1's data loads while slot k computes (index two slots ahead); unrolled by two so that the two
            // buffers alternate instead of being copied
            double ra[NR], ga[NG], rb[NR], gb[NG];
            int oa, ob;
            int32_t jA = src_slot(src, base, 0, cnt), jB = src_slot(src, base, 1, cnt);
            src.template head<NR>(jA, ra);
            src.grad(jA, ga);
            oa = src.order(jA);
            for (int k = 0; k < cnt; k += 2) {
                jA = src_slot(src, base, k + 2, cnt);
                src.template head<NR>(jB, rb);
                src.grad(jB, gb);
                ob = src.order(jB);
                slot(ra, ga, oa);
                if (k + 1 >= cnt) break;
                jB = src_slot(src, base, k + 3, cnt);
                src.template head<NR>(jA, ra);
                src.grad(jA, ga);
                oa = src.order(jA);
                slot(rb, gb, ob);
            }
        } else {
            for (int k = 0; k < cnt; ++k) {
                const int32_t j = src.idx(base + (int64_t)k * kChunk);
                double rj[NR], gj[NG];
                src.template head<NR>(j, rj);
                src.grad(j, gj);
                slot(rj, gj, src.order(j));
            }
        }
        if (FV == 1 && a.streamed) {
            const bool okd = a.okd[i];
#pragma unroll
            for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * (okd ? sacc[m] : 0.0)) + explicit_source<M>(a, i, nd.fi, m);
        } else {
            fit_divergence<D, M>(a, i, b, nd.fi, Kout);
        }
    }
    if constexpr (FV == 1) {   // a moving grid: K_X[p] = w (_rhs_flux!, TimeSteppers.jl:346-349), also without neighbours
        if (a.ale) {
            double w[D];
            Pde<D, M, KIND>::grid_velocity(a.pde, nd.fi, a.ale, a.mover_alpha, w);
#pragma unroll
            for (int d = 0; d < D; ++d) a.KX[i * D + d] = w[d];
        }
    }
    return flux_tail<D, M, KIND, CHECK>(a, pred, K, i, nd.fi, Kout, h);
}

// flux_node for a particle of a flat block (k_block_flat): every slot is flat, so the fit accumulators stay +0 and
// K_E is the divergence of that zero fit (or -0 + sources without neighbours), exactly as flux_node's loop leaves it;
// the store, the predictor and the MOOD check follow as there.
template <int D, int M, int KIND, bool CHECK>
__device__ __forceinline__ bool flux_node_flat(const StepArgs& a, const PredArgs& pred, const double* __restrict__ Y,
                                               double* __restrict__ K, int64_t i, double h) {
    using L = Layout<D, M>;
    constexpr int P0 = D / 2, P1 = (D + M - 1) / 2, NP = P1 - P0 + 1, O = D - 2 * P0;
    // A particle that stayed through the whole last step (every K exactly zero, its state free of zeros and finite,
    // every check passed: bits 7 and 6, or at stage 1 the last step's stage bits and bit 5) and through this step's
    // earlier stages (their bits and bit 5), in the first pass of a chain, without a source and without outflow
    // ghosts. Its stage state is U^n = U^{n-1} and the last step's stage state, bit for bit; its K here is the zero fit,
    // +0 in every component; the check sees last step's Euler substep (the state plus h (+0)) and passes again; the
    // predictor's target already holds the state (bit 7). The pass stores the K row, the byte and the check's
    // inadmissible flag as the general path would, and reads nothing of the particle's record.
    if constexpr (O == 0 && 2 * NP == M) {
        if (a.kz && pred.on && pred.first && a.src_mode == HCC_SRC_NONE) {
            const unsigned old = a.kz[i];
            const int st = pred.nterm - 1;
            const unsigned low = (1u << st) - 1u;
            const bool last = st == 0 ? ((old & pred.kmask) == pred.kmask && (old & kZStableNow))
                                      : ((old & kZLastStep) && (old & kZStableLast));
            const bool now = st == 0 || ((old & low) == low && (old & kZStableNow));
            if (last && now && pred.goff[i + 1] == pred.goff[i]) {
#pragma unroll
                for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
                a.kz[i] = (uint8_t)(st == 0 ? (1u | kZStableNow | kZStableLast | kZLastStep) : (old | (1u << st)));
                if constexpr (CHECK) a.inadm[i] = 0;
                return false;
            }
        }
    }
    double ri[L::NR], fi[M], Kout[M];
    load_rec<L::NR, L::PU>(Y, i, ri);
#pragma unroll
    for (int m = 0; m < M; ++m) fi[m] = ri[D + m];
    // the fit of the zero accumulators is -0 before sources (kZDiv) or there is no fit: no inverse is read
    if ((a.zfit && (a.zfit[i] & kZDiv)) || a.g.cnt[i] == 0) {
#pragma unroll
        for (int m = 0; m < M; ++m) Kout[m] = -0.0 + explicit_source<M>(a, i, fi, m);
    } else {
        double b[D + 1][D * M];
#pragma unroll
        for (int c = 0; c < D + 1; ++c)
#pragma unroll
            for (int q = 0; q < D * M; ++q) b[c][q] = 0.0;
        fit_divergence<D, M>(a, i, b, fi, Kout);
    }
    return flux_tail<D, M, KIND, CHECK, true>(a, pred, K, i, fi, Kout, h);
}

// The lists of the fused fixed point alternate between two copies, so that no kernel resets a count another
// workgroup of the same kernel still reads or fills. Iteration k (k = 0, 1, ...) has parity p = k % 2:
// k_mark_gather reads changed list p and fills work list p (and resets changed and work counts 1 - p), k_flux_list
// reads work list p and fills changed list 1 - p. The full flux pass fills changed list 0 (compact, appended); the
// full gather before it resets changed and work counts 0. k_flux_list writes work item t's entry at index t (the node,
// or -1 if its order stays), so it needs no atomics. Changed list p: length changed_count(p), any node at all
// changed_any(p), entries clist / clist2; work list: one buffer (dlist), length work_count(p).
__device__ __forceinline__ unsigned long long* changed_count(const StepArgs& a, int p) {
    return p ? &a.counters[6] : &a.counters[0];
}
__device__ __forceinline__ int32_t* changed_list(const StepArgs& a, int p) { return p ? a.clist2 : a.clist; }
__device__ __forceinline__ unsigned long long* work_count(const StepArgs& a, int p) {
    return p ? &a.counters[7] : &a.counters[5];
}
// whether changed list p holds a change (its entries can be -1: a recomputed node whose order stays)
__device__ __forceinline__ unsigned long long* changed_any(const StepArgs& a, int p) {
    return p ? &a.counters[9] : &a.counters[8];
}
// Sets changed_any(p) if any lane of the wavefront saw a change over all its items (every lane calls it): one store
// a wavefront instead of one per item group to the same word.
__device__ __forceinline__ void set_changed_any(const StepArgs& a, bool any, int p) {
    if (__ballot(any) && (threadIdx.x & 63) == 0) *changed_any(a, p) = 1;
}

// Appends the flagged particles of a wavefront to changed list p (one atomic per wavefront) and counts them.
__device__ __forceinline__ void append_changed(const StepArgs& a, bool ch, int64_t i, int p = 0) {
    const unsigned long long ball = __ballot(ch);
    if (!ball) return;
    const unsigned lane = threadIdx.x & 63;
    const unsigned leader = (unsigned)(__ffsll((unsigned long long)ball) - 1);
    unsigned long long base = 0;
    if (lane == leader) {
        base = atomicAdd(changed_count(a, p), (unsigned long long)__popcll(ball));
        *changed_any(a, p) = 1;
    }
    base = __shfl(base, (int)leader);
    if (ch) {
        const unsigned long long below = ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        changed_list(a, p)[base + __popcll(below)] = (int32_t)i;
    }
}

template <int D, int M, int KIND, bool CHECK, int FV>
__global__ __launch_bounds__(kBlock, D < 3 ? 2 : 1) void k_flux(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                  double* __restrict__ K, int full, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    int bad = 0;
    if (i < N) {
        // MOOD's Euler substep (form_substep!, TimeSteppers.jl:501-518)
        const double h = CHECK ? a.hratio * node_dt(a.dtp, a.inst, i) : 0.0;
        if (a.g.isb[i]) {
            if (full)
#pragma unroll
                for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
        } else if (full || a.dirty[i]) {
            ch = flux_node<D, M, KIND, CHECK, FV>(a, pred, Y, HbmSrc<D, M>{Y, a.G, a.orders, a.g.nbr}, K, i, h);
            if (CHECK) bad = a.inadm[i];
        }
    }
    if (CHECK) {
        append_changed(a, ch, i);
        if (pred.count) add_inadmissible(a, bad);
    }
}

// The full flux pass with the workgroup's halo (records, gradients, orders) staged in LDS (2-D).
template <int D, int M, int KIND, bool CHECK, int FV>
__global__ __launch_bounds__(kBlock, D < 3 ? 2 : 1) void k_flux_staged(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                         double* __restrict__ K, int64_t nblk) {
    extern __shared__ double sh[];
    const int64_t pos = logical_block(nblk);
    if (pos >= nblk) return;
    const int64_t lb = a.border ? a.border[pos] : pos;
    const double h = CHECK ? a.hratio * node_dt(a.dtp, a.inst, lb * kBlock) : 0.0;
    // a flat block (FV = 0) needs neither its halo nor its slot loops (uniform per workgroup: no barrier is skipped
    // by some threads only)
    const bool bflat = FV == 0 && a.bflat && a.bflat[lb];
    LdsSrc<D, M, true> src{};
    if (!bflat) src = stage_halo<D, M, true>(a, Y, sh, lb);
    const int64_t i = lb * kBlock + threadIdx.x;
    const int64_t N = a.g.N;
    bool ch = false;
    int bad = 0;
    if (i < N) {
        if (a.g.isb[i]) {
#pragma unroll
            for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
        } else {
            if (bflat) ch = flux_node_flat<D, M, KIND, CHECK>(a, pred, Y, K, i, h);
            else ch = flux_node<D, M, KIND, CHECK, FV>(a, pred, Y, src, K, i, h);
            if (CHECK) bad = a.inadm[i];
        }
    }
    if (CHECK) {
        append_changed(a, ch, i);
        if (pred.count) add_inadmissible(a, bad);
    }
}

// The unfused fixed point's recompute list: the dirty interior particles into dlist, their count into work count 0
// (cleared before), one atomic per wavefront; the order of the list does not matter (each entry writes its own K).
__global__ __launch_bounds__(kBlock) void k_compact_dirty(StepArgs a, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const bool take = i < a.g.N && a.dirty[i] && !a.g.isb[i];
    const unsigned long long ball = __ballot(take);
    if (!ball) return;
    const unsigned lane = threadIdx.x & 63;
    const unsigned leader = (unsigned)(__ffsll((unsigned long long)ball) - 1);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(work_count(a, 0), (unsigned long long)__popcll(ball));
    base = __shfl(base, (int)leader);
    if (take) a.dlist[base + __popcll(ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))))] = (int32_t)i;
}
__global__ void k_zero_u64(unsigned long long* p) { if (threadIdx.x == 0) *p = 0; }
// A captured fixed point runs a fixed number of iterations; the iteration after them would start from changed list q.
// If that list holds a node the fixed point has not converged: flag it (the host then re-runs the step uncaptured).
__global__ void k_fixed_point_overflow(const unsigned long long* __restrict__ counters, int q,
                                       unsigned long long* __restrict__ flag) {
    if (threadIdx.x != 0) return;
    const unsigned long long any = q ? counters[9] : counters[8], n = q ? counters[6] : counters[0];
    if (any && n) *flag = 1ull;
}

// The recompute kernels of the fixed point run one node per group of kSplit lanes of a 64-lane workgroup: each lane
// takes every kSplit-th slot of the node (loads and per-slot arithmetic in parallel) and writes what the accumulation
// needs to LDS; then one lane per accumulator adds the slot terms in slot order. Every sum keeps the reference's order,
// so the results are those of flux_node / gather_node bit for bit; the chain of one node shrinks from cnt slots to
// cnt / kSplit slots of work plus cnt additions.
// lanes per node: at least one per accumulator of the fit (D M)
template <int D, int M>
constexpr int kSplit = D * M <= 8 ? 8 : 16;

// The flux pass (with the MOOD check) of the fixed-point iteration with parity p over its work list (dlist, length
// read on the device): the dirty particles (the changed ones and their neighbours), each listed once. Changes go to
// changed list 1 - p. FP = false: a recompute of the unfused fixed point over the list k_compact_dirty made (work
// count 0), K only, as k_flux with full = 0 (a recomputed particle's K_X, its own grid velocity, does not change).
template <int D, int M, int KIND, int FV, bool FP = true, int G = kSplit<D, M>>
__global__ __launch_bounds__(64) void k_flux_list(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                  double* __restrict__ K, int p) {
    using L = Layout<D, M>;
    constexpr int NR = L::NR, NG = L::NG, BD = D + 1, NP = D * M, GPW = 64 / G;
    constexpr int SR = NP + 1 + D;   // per slot: dF (axis major), w, d invL; streamed: the contraction (M values)
    static_assert(NP <= G && M <= SR, "one accumulator owner per lane");
    __shared__ double rows[GPW][G][SR];
    __shared__ double part[GPW][NP];
    const int64_t n = (int64_t)*work_count(a, p);
    const int64_t N = a.g.N;
    const int grp = threadIdx.x / G, gl = threadIdx.x % G;
    if (FP && blockIdx.x == 0 && threadIdx.x == 0) *changed_count(a, 1 - p) = (unsigned long long)n;
    const bool streamed = FV == 1 && a.streamed;
    int dsum = 0;         // this lane's inadmissible delta over its items
    bool anych = false;   // this lane listed a change
    for (int64_t t0 = (int64_t)blockIdx.x * GPW; t0 < n; t0 += (int64_t)gridDim.x * GPW) {
        const int64_t t = t0 + grp;
        int64_t i = 0;
        int take = 0;
        if (t < n) {
            i = a.dlist[t];
            if (gl == 0) {
                if (FP) a.dmark[i] = 0u;   // each dirty node is listed once (k_mark_gather); cleared for the next
                take = !a.g.isb[i];
            }
        }
        take = __shfl(take, grp * G);
        const int cnt = take ? a.g.cnt[i] : 0;
        int mc = cnt;   // slot count of the workgroup's largest node: the chunk loop is uniform
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
        FluxNode<D, M, KIND, FV> nd;
        if (take) load_flux_node<D, M, KIND, FV>(a, Y, i, nd);
        const int64_t base = take ? sell_base(a.g, i) : 0;
        double b[BD], sacc = 0.0;   // owner lane gl: accumulators of pair gl = d M + m (streamed: component gl)
#pragma unroll
        for (int c = 0; c < BD; ++c) b[c] = 0.0;
        // the owner lane's derivative row of the divergence inverse, loaded before the slots (used after them)
        const bool own = take && !streamed && gl < NP;
        double xr[BD];
#pragma unroll
        for (int c = 0; c < BD; ++c) xr[c] = own ? a.invd[((gl / M) * BD + c) * N + i] : 0.0;
        // The lane's slots k = k0 + gl, one per chunk, software-pipelined: the records of the next chunk's slot are
        // loaded before this chunk's arithmetic, its neighbour index one chunk earlier still, so that a chunk waits
        // on no load of its own (a short list's kernel is a chain of dependent loads, not arithmetic).
        int32_t jn = (take && gl < cnt) ? a.g.nbr[base + (int64_t)gl * kChunk] : 0;   // index of the next slot
        double rn[NR], gn[NG];   // records of the next slot
        int on = 0;
        if (take && gl < cnt) {
            load_rec<NR, L::PU>(Y, jn, rn);
            load_rec<NG, L::GS>(a.G, jn, gn);
            on = a.orders[jn];
        }
        jn = (take && gl + G < cnt) ? a.g.nbr[base + (int64_t)(gl + G) * kChunk] : 0;
        for (int k0 = 0; k0 < mc; k0 += G) {
            const int k = k0 + gl;
            double rj[NR], gj[NG];
#pragma unroll
            for (int q = 0; q < NR; ++q) rj[q] = rn[q];
#pragma unroll
            for (int q = 0; q < NG; ++q) gj[q] = gn[q];
            const int oj = on;
            if (take && k + G < cnt) {
                load_rec<NR, L::PU>(Y, jn, rn);
                load_rec<NG, L::GS>(a.G, jn, gn);
                on = a.orders[jn];
            }
            jn = (take && k + 2 * G < cnt) ? a.g.nbr[base + (int64_t)(k + 2 * G) * kChunk] : 0;
            if (take && k < cnt) {
                double dF[D][M], w, pt[BD], ctr[M];
                slot_flux<D, M, KIND, FV>(a, nd, rj, gj, oj, dF, w, pt, ctr);
                double* row = rows[grp][gl];
                if (streamed) {
#pragma unroll
                    for (int m = 0; m < M; ++m) row[m] = ctr[m];
                } else {
#pragma unroll
                    for (int d = 0; d < D; ++d)
#pragma unroll
                        for (int m = 0; m < M; ++m) row[d * M + m] = dF[d][m];
                    row[NP] = w;
#pragma unroll
                    for (int d = 0; d < D; ++d) row[NP + 1 + d] = pt[1 + d];
                }
            }
            __syncthreads();
            if (take) {
                const int ns = min(G, cnt - k0);
                if (streamed) {
                    if (gl < M)
                        for (int q = 0; q < ns; ++q) sacc = sacc + rows[grp][q][gl];
                } else if (gl < NP) {
                    for (int q = 0; q < ns; ++q) {
                        const double* row = rows[grp][q];
                        const double dFv = row[gl], w = row[NP];
                        double pt[BD];
                        pt[0] = 1.0;
#pragma unroll
                        for (int d = 0; d < D; ++d) pt[1 + d] = row[NP + 1 + d];
#pragma unroll
                        for (int c = 0; c < BD; ++c) b[c] = b[c] + w * (pt[c] * dFv);
                    }
                }
            }
            __syncthreads();
        }
        // the derivative row of each pair (apply_weights_divergence), then K_E and the check by the group's first lane
        if (take && streamed && gl < M) part[grp][gl] = sacc;
        if (own && a.okd[i]) {
            double r = xr[0] * b[0];
#pragma unroll
            for (int c = 1; c < BD; ++c) r = r + xr[c] * b[c];
            part[grp][gl] = r;
        }
        __syncthreads();
        bool ch = false;
        int delta = 0;
        if (take && gl == 0) {
            double Kout[M];
            const bool okd = a.okd[i];
            if (cnt == 0) {
#pragma unroll
                for (int m = 0; m < M; ++m) Kout[m] = -0.0 + explicit_source<M>(a, i, nd.fi, m);
            } else if (streamed) {
#pragma unroll
                for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * (okd ? part[grp][m] : 0.0)) + explicit_source<M>(a, i, nd.fi, m);
            } else if (okd) {
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    double acc = 0.0;
#pragma unroll
                    for (int d = 0; d < D; ++d) acc = acc + part[grp][d * M + m];
                    Kout[m] = -(2.0 * (acc * a.g.invL)) + explicit_source<M>(a, i, nd.fi, m);
                }
            } else {
#pragma unroll
                for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * 0.0) + explicit_source<M>(a, i, nd.fi, m);
            }
            if constexpr (FP) {
                const int before = a.inadm[i];
                ch = flux_tail<D, M, KIND, true>(a, pred, K, i, nd.fi, Kout, a.hratio * node_dt(a.dtp, a.inst, i));
                delta = (int)a.inadm[i] - before;
            } else {
                flux_tail<D, M, KIND, false>(a, pred, K, i, nd.fi, Kout, 0.0);
            }
        }
        __syncthreads();
        if constexpr (!FP) continue;
        // work item t's entry of changed list 1 - p: the node if its order is to drop, else -1 (no atomics)
        if (t < n && gl == 0) changed_list(a, 1 - p)[t] = ch ? (int32_t)i : -1;
        anych = anych || ch;
        dsum += delta;
    }
    if constexpr (FP) {
        set_changed_any(a, anych, 1 - p);
        if (pred.count) add_inadmissible_sum(a, dsum);
    }
}

// gather_node (EXT off) of particle i by its group of kSplit lanes (all lanes of the workgroup call it; act: the group
// has a particle to gather). The fit's accumulators are summed in slot order by one owner lane each; the extrema and
// the limiter's phi are minima / maxima (IEEE minimum / maximum: exact, so any order), reduced over the group.
template <int D, int M, int KIND, int G = kSplit<D, M>>
__device__ __forceinline__ void gather_split(const StepArgs& a, const double* __restrict__ Y, int64_t i, bool act,
                                             int cnt, int order, int gl, double (*rows)[1 + D + M], double* bsh) {
    using L = Layout<D, M>;
    constexpr int NR = L::NR, NG = L::NG, NP = D * M;
    static_assert(NP <= G, "one accumulator owner per lane");
    const int64_t N = a.g.N;
    const double invL = a.g.invL;
    if (act && order > 1 && cnt < D) {   // starved neighbourhood (basis_count(D, 1) = D)
        order = 1;
        if (gl == 0) a.orders[i] = 1;
    }
    const bool fit = act && order == 2 && a.okg[i];
    double ri[NR], umax[M], umin[M], g[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) g[q] = 0.0;
    if (fit) load_rec<NR, L::PU>(Y, i, ri);
#pragma unroll
    for (int m = 0; m < M; ++m) { umax[m] = fit ? ri[D + m] : 0.0; umin[m] = umax[m]; }
    int mc = fit ? cnt : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
    const int64_t base = fit ? sell_base(a.g, i) : 0;
    double b = 0.0;   // owner lane gl: b[q][m] of pair gl = q M + m
    for (int k0 = 0; k0 < mc; k0 += G) {
        const int k = k0 + gl;
        if (fit && k < cnt) {
            double rj[NR], dd[D];
            load_rec<NR, L::PU>(Y, a.g.nbr[base + (int64_t)k * kChunk], rj);
            const double d2 = rec_geom<D>(a.g, ri, rj, dd);
            double* row = rows[gl];
            row[0] = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
#pragma unroll
            for (int d = 0; d < D; ++d) row[1 + d] = dd[d] * invL;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double fj = rj[D + m];
                row[1 + D + m] = fj - ri[D + m];
                umax[m] = jmax(umax[m], fj);
                umin[m] = jmin(umin[m], fj);
            }
        }
        __syncthreads();
        if (fit && gl < NP) {
            const int q = gl / M, m = gl % M;
            const int ns = min(G, cnt - k0);
            for (int s = 0; s < ns; ++s) b = b + rows[s][0] * (rows[s][1 + q] * rows[s][1 + D + m]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int m = 0; m < M; ++m) {
            umax[m] = jmax(umax[m], __shfl_xor(umax[m], o));
            umin[m] = jmin(umin[m], __shfl_xor(umin[m], o));
        }
    if (fit && gl < NP) bsh[gl] = b;
    __syncthreads();
    if (fit) {
        double inv[D][D];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) inv[r][c] = a.invg[(r * D + c) * N + i];
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double c = inv[q][0] * bsh[0 * M + m];
#pragma unroll
                for (int r = 1; r < D; ++r) c = c + inv[q][r] * bsh[r * M + m];
                g[q * M + m] = c * invL;
            }
    }
    if (fit && a.limiter != HCC_LIM_NONE) {
        double phi[M], dup[M], ddn[M], tup[M], tdn[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            phi[m] = 1.0;
            const double ui = ri[D + m];
            dup[m] = umax[m] - ui;
            ddn[m] = umin[m] - ui;
            tup[m] = 16 * kEps * jmax(fabs(ui), fabs(umax[m]));
            tdn[m] = 16 * kEps * jmax(fabs(ui), fabs(umin[m]));
        }
        for (int k = gl; k < cnt; k += G) {
            double pj[D], dd[D], pint[D];
            load_rec<D, L::PU>(Y, a.g.nbr[base + (int64_t)k * kChunk], pj);
            rec_geom<D>(a.g, ri, pj, dd);
#pragma unroll
            for (int d = 0; d < D; ++d) pint[d] = 0.5 * dd[d];
            double rec[M];
            bool lim[M], any = false;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < D; ++q) v = v + g[q * M + m] * pint[q];
                rec[m] = v;
                lim[m] = fabs(v) > (v > 0.0 ? tup[m] : tdn[m]);
                any = any || lim[m];
            }
            if (any) {
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const double r = (rec[m] > 0.0 ? dup[m] : ddn[m]) / rec[m];
                    const double pk = a.limiter == HCC_LIM_BJ ? bj_phi(r) : vk_phi(r);
                    phi[m] = lim[m] ? jmin(phi[m], pk) : phi[m];
                }
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int m = 0; m < M; ++m) phi[m] = jmin(phi[m], __shfl_xor(phi[m], o));
#pragma unroll
        for (int q = 0; q < D; ++q)
#pragma unroll
            for (int m = 0; m < M; ++m) g[q * M + m] = g[q * M + m] * phi[m];
    }
    if (act && gl == 0) store_rec<NG, L::GS>(a.G, i, g);
    __syncthreads();
}

// _mood_drop! (MOOD.jl:299-309), one hop of expand_mask! (MOOD.jl:323-351) and the gather of the changed list
// (length counters[0], read on the device), one particle per group of kSplit lanes: the particle drops one order, it
// and its neighbours are marked dirty and the newly marked ones appended to the recompute's work list; then its
// gradient is recomputed at the new order (the gather reads only the particle's own order, which
// the group's first lane just wrote and hands to the group, and stage states).
// The neighbour relation is symmetric, so marking the neighbours of changed particles gives the pull form's set.
// Block 0 also resets the dirty-list length, counts the fixed-point iterations that changed an order and publishes
// the count to the host (pub, pinned host memory) with a system-scope store.
template <int D, int M, int KIND, int G = kSplit<D, M>>
__global__ __launch_bounds__(64) void k_mark_gather(StepArgs a, const double* __restrict__ Y, unsigned long long* pub,
                                                    int p) {
    constexpr int GPW = 64 / G;
    constexpr int kListCap = 512;   // newly marked nodes collected before they are appended
    __shared__ double rows[GPW][G][1 + D + M];
    __shared__ double bsh[GPW][D * M];
    __shared__ int32_t lbuf[kListCap];
    // changed list p: n entries, a node or -1 (none if no entry is a node)
    const int64_t n = *changed_any(a, p) ? (int64_t)*changed_count(a, p) : 0;
    const int32_t* cl = changed_list(a, p);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *changed_count(a, 1 - p) = 0;
        *work_count(a, 1 - p) = 0;
        *changed_any(a, 1 - p) = 0;
        if (n > 0) a.counters[4] += 1;
        __hip_atomic_store(pub, (unsigned long long)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    const int lane = threadIdx.x, grp = lane / G, gl = lane % G;
    // the particle and its neighbours are marked dirty; the lane whose atomicExch sets a node's mark lists the node
    // (once per iteration): collected in LDS over all the workgroup's items, then appended to the work list with one
    // atomicAdd whenever the buffer fills and once at the end (an append per item group queued one contended atomic
    // on the list length per group, which serialised the kernel when most particles changed)
    int nl = 0;   // entries in lbuf (uniform)
    auto flush = [&]() {
        __syncthreads();
        if (nl > 0) {
            unsigned long long wb = 0;
            if (lane == 0) wb = atomicAdd(work_count(a, p), (unsigned long long)nl);
            wb = __shfl(wb, 0);
            for (int e = lane; e < nl; e += 64) a.dlist[wb + e] = lbuf[e];
        }
        __syncthreads();
        nl = 0;
    };
    for (int64_t t0 = (int64_t)blockIdx.x * GPW; t0 < n; t0 += (int64_t)gridDim.x * GPW) {
        const int64_t t = t0 + grp;
        int32_t i = 0;
        int cnt = -1, order = 0;   // cnt -1: no particle for this group
        if (t < n && cl[t] >= 0) {
            i = cl[t];
            cnt = a.g.cnt[i];
            if (gl == 0) {
                order = a.orders[i] - 1;
                a.orders[i] = (int8_t)order;
            }
        }
        order = __shfl(order, grp * G);
        int mc = cnt;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
        const int64_t base = cnt >= 0 ? sell_base(a.g, i) : 0;
        for (int k0 = -1; k0 < mc; k0 += G) {
            if (nl > kListCap - 64) flush();   // a step adds at most 64 entries
            const int k = k0 + gl;
            int32_t j = 0;
            bool first = false;
            if (k < cnt) {
                j = k < 0 ? i : a.g.nbr[base + (int64_t)k * kChunk];
                first = atomicExch(&a.dmark[j], 1u) == 0u;
            }
            const unsigned long long ball = __ballot(first);
            const unsigned long long below = ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
            if (first) lbuf[nl + __popcll(below)] = j;
            nl += __popcll(ball);
        }
        gather_split<D, M, KIND, G>(a, Y, i, cnt >= 0 && !a.g.isb[i], cnt, order, gl, rows[grp], bsh[grp]);
    }
    flush();
}

// The mark pass of a dense fixed-point iteration with parity p (MUSCL order 2; k_flux_marked follows), one lane per
// entry of changed list p: the particle drops one order, to 1 (the only order below 2), so its new gradient is the
// zero record gather_split stores for a particle without a fit; it and its neighbours are marked for the flux pass
// with plain stores (a mark set twice is the same mark; nothing is listed). The bookkeeping is k_mark_gather's.
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_mark_dense(StepArgs a, unsigned long long* pub, int p) {
    using L = Layout<D, M>;
    const int64_t n = *changed_any(a, p) ? (int64_t)*changed_count(a, p) : 0;
    const int32_t* cl = changed_list(a, p);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *changed_count(a, 1 - p) = 0;
        *work_count(a, 1 - p) = 0;
        *changed_any(a, 1 - p) = 0;
        if (n > 0) a.counters[4] += 1;
        __hip_atomic_store(pub, (unsigned long long)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
        const int32_t i = cl[t];
        if (i < 0) continue;
        a.orders[i] = (int8_t)(a.orders[i] - 1);
        if (!a.g.isb[i]) {
            double g[L::NG];
#pragma unroll
            for (int q = 0; q < L::NG; ++q) g[q] = 0.0;
            store_rec<L::NG, L::GS>(a.G, i, g);
        }
        a.dmark[i] = 1u;
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        for (int k = 0; k < cnt; ++k) a.dmark[a.g.nbr[base + (int64_t)k * kChunk]] = 1u;
    }
}

// append_changed with one atomic per workgroup of kBlock threads (every thread calls it): the wavefronts' counts are
// added in LDS, thread 0 takes the workgroup's range of changed list p, each wavefront writes its part in lane order
__device__ __forceinline__ void append_changed_block(const StepArgs& a, bool ch, int64_t i, int p) {
    constexpr int NW = kBlock / 64;
    __shared__ unsigned long long wbase[NW + 1];
    const unsigned long long ball = __ballot(ch);
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    if (lane == 0) wbase[w] = (unsigned long long)__popcll(ball);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int q = 0; q < NW; ++q) { const unsigned long long c = wbase[q]; wbase[q] = tot; tot += c; }
        wbase[NW] = tot ? atomicAdd(changed_count(a, p), tot) : 0ull;
        if (tot) *changed_any(a, p) = 1;
    }
    __syncthreads();
    if (ch) {
        const unsigned long long below = ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        changed_list(a, p)[wbase[NW] + wbase[w] + (unsigned long long)__popcll(below)] = (int32_t)i;
    }
}

// The flux pass of a dense fixed-point iteration with parity p (after k_mark_dense): the staged full
// pass restricted to the marked particles (dmark, cleared here), each recomputed and re-checked as k_flux_list does,
// the changed ones appended to changed list 1 - p (compact). A workgroup without a marked particle ends before
// staging. The set recomputed and the values written are those of the list form, so its bits are too; it is the
// faster form when most particles are marked (one staged halo per block instead of kSplit lanes per particle reading
// through the caches, and no list appends).
template <int D, int M, int KIND, int FV>
__global__ __launch_bounds__(kBlock, D < 3 ? 2 : 1) void k_flux_marked(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                         double* __restrict__ K, int p, int64_t nblk) {
    extern __shared__ double sh[];
    const int64_t pos = logical_block(nblk);
    if (pos >= nblk) return;
    const int64_t lb = a.border ? a.border[pos] : pos;
    const int64_t i = lb * kBlock + threadIdx.x;
    const bool mk = i < a.g.N && a.dmark[i] != 0u;
    if (!__syncthreads_or(mk)) return;   // uniform per workgroup
    if (mk) a.dmark[i] = 0u;
    const LdsSrc<D, M, true> src = stage_halo<D, M, true>(a, Y, sh, lb);
    bool ch = false;
    int delta = 0;
    if (mk && !a.g.isb[i]) {
        const int before = a.inadm[i];
        ch = flux_node<D, M, KIND, true, FV>(a, pred, Y, src, K, i, a.hratio * node_dt(a.dtp, a.inst, i));
        delta = (int)a.inadm[i] - before;
    }
    append_changed_block(a, ch, i, 1 - p);
    if (pred.count) add_inadmissible(a, delta);
}

// ============================================================================================================
// the unfused MOOD path (a stored candidate, or the strategies and criteria the flux pass does not check):
// _mood_flag! + _mood_drop! (MOOD.jl:289-309) on all interior particles or the dirty mask, the forcing pass of
// StrictEPD0 (_mood_force!, MOOD.jl:311-320), expand_mask! (MOOD.jl:323-351)
// ============================================================================================================
// The candidate is the stored one (cand), or with Kp the Euler substep W = Y + h K formed on the fly (form_substep!,
// TimeSteppers.jl:501-518). changed = triggered after the decrement (a flagged particle at order 1 is un-flagged).
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_mood_flag(StepArgs a, const double* __restrict__ Yp, const double* __restrict__ cand,
                                                       const double* __restrict__ Kp, const double* __restrict__ KIp,
                                                       int full, int64_t nblk) {
    using P = Pde<D, M, KIND>;
    using L = Layout<D, M>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    if (i < N && !a.g.isb[i] && (full || a.dirty[i])) {
        double r[L::NR], W[M];
        if (Kp) {
            const double h = a.hratio * node_dt(a.dtp, a.inst, i);
            load_rec<L::NR, L::PU>(Yp, i, r);
#pragma unroll
            for (int m = 0; m < M; ++m)   // with implicit sources the substep is Y + h (K + K_I)
                W[m] = KIp ? r[D + m] + h * (Kp[i * M + m] + KIp[i * M + m]) : r[D + m] + h * Kp[i * M + m];
        } else {
            load_rec<L::NR, L::PU>(cand, i, r);
#pragma unroll
            for (int m = 0; m < M; ++m) W[m] = r[D + m];
        }
        const bool adm = P::admissible(a.pde, W);
        a.inadm[i] = !adm;
        bool flag = !adm;
        if (!flag) {
            if (a.criterion == HCC_CRIT_ONLY) {
                flag = true;
            } else if (a.criterion == HCC_CRIT_U1 || a.criterion == HCC_CRIT_U2) {
                // findLocalExtrema over {i} and its neighbours in Y_prev; _dmp_fails (MOOD.jl:175-184), and for MOODu2
                // (_u2_fails, MOOD.jl:215-270, MUSCL order 2) the components below the plateau are exempt
                double mn[M], mx[M];
                if (a.ext_on) {   // the extrema the stage's full gather stored (the same operands, the same bits)
                    double e[2 * M];
                    load_rec<2 * M, 2 * M>(a.ext, i, e);
#pragma unroll
                    for (int m = 0; m < M; ++m) { mx[m] = e[m]; mn[m] = e[M + m]; }
                } else {
                    load_rec<L::NR, L::PU>(Yp, i, r);
#pragma unroll
                    for (int m = 0; m < M; ++m) { mn[m] = r[D + m]; mx[m] = r[D + m]; }
                    const int cnt = a.g.cnt[i];
                    const int64_t base = sell_base(a.g, i);
                    for (int k = 0; k < cnt; ++k) {
                        load_rec<L::NR, L::PU>(Yp, a.g.nbr[base + (int64_t)k * kChunk], r);
#pragma unroll
                        for (int m = 0; m < M; ++m) { mn[m] = jmin(mn[m], r[D + m]); mx[m] = jmax(mx[m], r[D + m]); }
                    }
                }
                const double dl = a.delta, d3 = dl * dl * dl;
                const bool u2 = a.criterion == HCC_CRIT_U2 && a.range;
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const bool viol = fabs(mx[m] - mn[m]) >= d3 && (W[m] < mn[m] - dl || W[m] > mx[m] + dl);
                    const bool plateau = u2 && mx[m] - mn[m] < a.plateau * a.range[m];
                    if (viol && !plateau) flag = true;
                }
            }
        }
        if (flag) {
            const int o = a.orders[i];
            if (o > 1) { a.orders[i] = (int8_t)(o - 1); a.trig[i] = 1; ch = true; }
        }
    }
    if (i < N) a.changed[i] = ch;
    const unsigned long long ball = __ballot(ch);
    if (ball && ((threadIdx.x & 63) == (unsigned)(__ffsll((unsigned long long)__ballot(1)) - 1)))
        atomicAdd(&a.counters[0], (unsigned long long)__popcll(ball));
}

// forced_order of StrictEPD0 (Halo{1}, MOOD.jl:49-64) with _mood_force! (:311-320): every interior particle pulls the
// minimum order of its triggered neighbours (changed) from the snapshot taken after the decrement, so nothing
// cascades. dirty receives the new changed set (triggered or forced); the caller copies it back into changed.
__global__ __launch_bounds__(kBlock) void k_force(StepArgs a, const int8_t* __restrict__ snap, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= a.g.N) return;
    uint8_t ch = a.changed[i];
    if (!a.g.isb[i]) {
        const int own = snap[i];
        int o = own;
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        for (int k = 0; k < cnt; ++k) {
            const int32_t p = a.g.nbr[base + (int64_t)k * kChunk];
            if (a.changed[p]) o = o < snap[p] ? o : snap[p];
        }
        if (o < own) { a.orders[i] = (int8_t)o; ch = 1; }
    }
    a.dirty[i] = ch;
}

// One hop of expand_mask! (_expand_body!, MOOD.jl:323-351): dst = src or any neighbour in src, in push form on a
// cleared dst: each particle of src marks itself and its neighbours (the neighbour relation is symmetric, so this is
// the pull form's set; concurrent marks write the same 1). Only the few particles of src read their lists.
__global__ __launch_bounds__(kBlock) void k_expand(StepArgs a, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                    int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= a.g.N || !src[i]) return;
    dst[i] = 1;
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    for (int k = 0; k < cnt; ++k) dst[a.g.nbr[base + (int64_t)k * kChunk]] = 1;
}
static hipError_t launch_expand(const StepArgs& a, const uint8_t* src, uint8_t* dst, int64_t N, int64_t nb,
                                hipStream_t str) {
    hipError_t e = hipMemsetAsync(dst, 0, (size_t)std::max<int64_t>(N, 1), str);
    if (e != hipSuccess) return e;
    k_expand<<<grid_for(nb), kBlock, 0, str>>>(a, src, dst, nb);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_copy_u8(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// _nb_min_order (MOOD.jl:66-74) of every particle: EPD2's interface order of the edge (i, j) is the minimum of the
// two (_get_interface_orders, MOOD.jl:82-85), so the flux pass reads these as orders with the edge-minimum rule.
// pack: AlignedRusanovFlux also needs the true orders (whether a pair is below the full order): out = min | order << 4.
// only: the particles whose minimum can have changed since the last pass (the rest keep theirs), or null for all.
__global__ __launch_bounds__(kBlock) void k_nbmin(StepArgs a, int8_t* __restrict__ out, int pack,
                                                   const uint8_t* __restrict__ only, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= a.g.N || (only && !only[i])) return;
    int o = a.orders[i];
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    for (int k = 0; k < cnt; ++k) {
        const int oj = a.orders[a.g.nbr[base + (int64_t)k * kChunk]];
        o = o < oj ? o : oj;
    }
    out[i] = (int8_t)(pack ? o | (a.orders[i] << 4) : o);
}

// _global_range (MOOD.jl:206-213) of the state components of the records Y[0 .. N): each workgroup reduces its nodes
// with Julia's min / max (IEEE minimum / maximum: exact, so the order of the reduction does not matter) into
// part[b * 2M ..]: M minima, M maxima; k_range_final reduces the parts and writes range = max - min.
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_range_part(const double* __restrict__ Y, int64_t N, double* __restrict__ part) {
    using L = Layout<D, M>;
    __shared__ double sl[kBlock / 64][2 * M];
    double lo[M], hi[M];
#pragma unroll
    for (int m = 0; m < M; ++m) { lo[m] = INFINITY; hi[m] = -INFINITY; }
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < N) {
        double r[L::NR];
        load_rec<L::NR, L::PU>(Y, i, r);
#pragma unroll
        for (int m = 0; m < M; ++m) { lo[m] = r[D + m]; hi[m] = r[D + m]; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int m = 0; m < M; ++m) { lo[m] = jmin(lo[m], __shfl_xor(lo[m], o)); hi[m] = jmax(hi[m], __shfl_xor(hi[m], o)); }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) { sl[w][m] = lo[m]; sl[w][M + m] = hi[m]; }
    __syncthreads();
    if (threadIdx.x < 2 * M) {
        const int c = threadIdx.x;
        double v = sl[0][c];
        for (int q = 1; q < kBlock / 64; ++q) v = c < M ? jmin(v, sl[q][c]) : jmax(v, sl[q][c]);
        part[(int64_t)blockIdx.x * 2 * M + c] = v;
    }
}
// one workgroup of 1024 threads: strided over the parts, then across the workgroup (any order gives the same bits)
template <int M>
__global__ __launch_bounds__(1024) void k_range_final(const double* __restrict__ part, int64_t nparts, double* __restrict__ range) {
    __shared__ double sl[16][2 * M];
    double lo[M], hi[M];
#pragma unroll
    for (int m = 0; m < M; ++m) { lo[m] = INFINITY; hi[m] = -INFINITY; }
    for (int64_t b = threadIdx.x; b < nparts; b += 1024)
#pragma unroll
        for (int m = 0; m < M; ++m) { lo[m] = jmin(lo[m], part[b * 2 * M + m]); hi[m] = jmax(hi[m], part[b * 2 * M + M + m]); }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int m = 0; m < M; ++m) { lo[m] = jmin(lo[m], __shfl_xor(lo[m], o)); hi[m] = jmax(hi[m], __shfl_xor(hi[m], o)); }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int m = 0; m < M; ++m) { sl[threadIdx.x >> 6][m] = lo[m]; sl[threadIdx.x >> 6][M + m] = hi[m]; }
    __syncthreads();
    if (threadIdx.x < M) {
        const int m = threadIdx.x;
        double a = sl[0][m], b = sl[0][M + m];
        for (int q = 1; q < 16; ++q) { a = jmin(a, sl[q][m]); b = jmax(b, sl[q][M + m]); }
        range[m] = b - a;
    }
}

// counters[0] into the host's pinned slot pub
__global__ __launch_bounds__(64) void k_publish(const unsigned long long* __restrict__ counters, unsigned long long* pub) {
    if (threadIdx.x == 0) __hip_atomic_store(pub, counters[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The step's statistics into a ring slot in coherent pinned host memory: counters[0..5] and the dt bits at [7] (one
// small kernel instead of two copies through the DMA engine)
__global__ __launch_bounds__(64) void k_stats_out(const unsigned long long* __restrict__ counters,
                                                  const double* __restrict__ dtp, unsigned long long* out) {
    const int t = threadIdx.x;
    if (t < 6) __hip_atomic_store(out + t, counters[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (t == 7) {
        unsigned long long bits;
        const double dt = *dtp;
        __builtin_memcpy(&bits, &dt, sizeof(bits));
        __hip_atomic_store(out + 7, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// _count_inadmissible (TimeSteppers.jl:481-490): interior candidates accepted while not admissible
__global__ __launch_bounds__(kBlock) void k_count_inadm(StepArgs a, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const bool bad = i < a.g.N && !a.g.isb[i] && a.inadm[i];
    const unsigned long long ball = __ballot(bad);
    if (ball && ((threadIdx.x & 63) == (unsigned)(__ffsll((unsigned long long)__ballot(1)) - 1)))
        atomicAdd(&a.counters[1], (unsigned long long)__popcll(ball));
}

// ============================================================================================================
// CFL estimate: get_time_step with cached weights (ParticleGrids.jl:386-450)
// ============================================================================================================
// max_eigenvalue of every particle once per estimate (_get_Lambda, ParticleGrids.jl:294)
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_lambda(StepArgs a, const double* __restrict__ U, int64_t nblk) {
    using P = Pde<D, M, KIND>;
    using L = Layout<D, M>;
    const int64_t i = node_index(nblk);
    if (i >= a.g.N) return;
    double r[L::NR], f[M], l[D];
    load_rec<L::NR, L::PU>(U, i, r);
#pragma unroll
    for (int m = 0; m < M; ++m) f[m] = r[D + m];
    P::maxeig(a.pde, f, l);
#pragma unroll
    for (int d = 0; d < D; ++d) a.lam[i * D + d] = l[d];
}

// _cfl_weights_body! (ParticleGrids.jl:439-450) and the minimum over particles (a workgroup minimum per block,
// reduced by k_cfl_final). The per-slot derivative weights are recomputed from the divergence inverse and the slot
// geometry with the operations of _weights_body! (MUSCL.jl:211-223): the same bits as the cached dweights, without
// streaming them. STAGED: the block's halo positions and wave speeds (computed while staging) sit in LDS.
template <int D, int M, int KIND, bool STAGED>
__global__ __launch_bounds__(kBlock) void k_cfl(StepArgs a, const double* __restrict__ U, int64_t nblk) {
    using L = Layout<D, M>;
    constexpr int BD = D + 1;
    extern __shared__ double sh[];
    const int64_t N = a.g.N;
    constexpr int CST = kCflStride<D>;
    int64_t i, s0 = 0;
    const uint16_t* nl = nullptr;
    if (STAGED) {
        const int64_t lb = logical_block(nblk);
        if (lb >= nblk) return;
        const int32_t h0 = a.g.halo_off[lb], nh = a.g.halo_off[lb + 1] - h0;
        uint16_t* nlw = reinterpret_cast<uint16_t*>(sh) + lds_index_offset(a.g.hmax, CST, false);
        s0 = stage_index_table(a.g, nlw, lb);
        nl = nlw;
        for (int t = threadIdx.x; t < nh; t += kStage) {   // positions and wave speeds (_get_Lambda) of the halo
            const int32_t j = a.g.halo_idx[h0 + t];
            double r[L::NR], f[M], l[D];
            load_rec<L::NR, L::PU>(U, j, r);
#pragma unroll
            for (int m = 0; m < M; ++m) f[m] = r[D + m];
            Pde<D, M, KIND>::maxeig(a.pde, f, l);
#pragma unroll
            for (int d = 0; d < D; ++d) { sh[t * CST + d] = r[d]; sh[t * CST + D + d] = l[d]; }
        }
        __syncthreads();
        i = lb * kStage + threadIdx.x;
    } else {
        i = node_index(nblk);
    }
    double dt = INFINITY;
    if (i < N && !a.g.isb[i] && a.g.cnt[i] > 0) {
        double li[D], X[D][BD], pi[D];
        if (STAGED) {
            double r[L::NR], f[M];
            load_rec<L::NR, L::PU>(U, i, r);
#pragma unroll
            for (int m = 0; m < M; ++m) f[m] = r[D + m];
            Pde<D, M, KIND>::maxeig(a.pde, f, li);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) li[d] = a.lam[i * D + d];
        }
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < BD; ++c) X[r][c] = a.invd[(r * BD + c) * N + i];
        load_rec<D, L::PU>(U, i, pi);
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        const double invL = a.g.invL;
        auto nidx = [&](int k) -> int32_t {
            const int kk = k < cnt ? k : cnt - 1;
            return STAGED ? (int32_t)nl[base + (int64_t)kk * kChunk - s0] : a.g.nbr[base + (int64_t)kk * kChunk];
        };
        auto fetch = [&](int32_t j, double (&p)[D], double (&l)[D]) {
            if (STAGED) {
                double v[2 * D];
                lds_read<2 * D>(sh + j * CST, v);
#pragma unroll
                for (int d = 0; d < D; ++d) { p[d] = v[d]; l[d] = v[D + d]; }
            } else {
                load_rec<D, L::PU>(U, j, p);
#pragma unroll
                for (int d = 0; d < D; ++d) l[d] = a.lam[(int64_t)j * D + d];
            }
        };
        double sum_c = 0.0;
        int32_t j1 = nidx(1);
        double pn[D], ln[D];
        fetch(nidx(0), pn, ln);
        for (int k = 0; k < cnt; ++k) {   // slot k+1 loaded (index k+2) while slot k computes
            double pj[D], lj[D];
#pragma unroll
            for (int d = 0; d < D; ++d) { pj[d] = pn[d]; lj[d] = ln[d]; }
            const int32_t jn = j1;
            j1 = nidx(k + 2);
            fetch(jn, pn, ln);
            double dd[D];
            const double d2 = rec_geom<D>(a.g, pi, pj, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
            double v[BD];
            v[0] = w * 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) v[1 + d] = w * (dd[d] * invL);
            double dv = 0.0;
#pragma unroll
            for (int r = 0; r < D; ++r) {
                double acc = X[r][0] * v[0];
#pragma unroll
                for (int c = 1; c < BD; ++c) acc = acc + X[r][c] * v[c];
                const double term = fabs(acc * invL) * jmax(li[r], lj[r]);
                dv = (r == 0) ? term : dv + term;
            }
            sum_c = sum_c + dv;
        }
        if (sum_c > kFloatMin) dt = (1.0 / (double)D) / sum_c;
    }
    block_min(a, dt, logical_block(nblk));
}

// dt = min over the workgroup minima (ParticleGrids.jl:386-415 reduces with minimum); +inf if no node contributes
// With dtp: also the step's dt = min(cfl * dt_cfl, cap) (the driver loop's min(CFL * get_time_step, t_end - t),
// test/time_integration.jl:107-109, with Julia's min), kept on the device for the step that follows.
__global__ __launch_bounds__(1024) void k_cfl_final(const double* __restrict__ bmin, int64_t nblk, unsigned long long* counters,
                                                    double* dtp, double cfl, double cap) {
    __shared__ double wmin[16];
    double dt = INFINITY;
    for (int64_t b = threadIdx.x; b < nblk; b += 1024) dt = fmin(dt, bmin[b]);
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < 16; ++w) m = fmin(m, wmin[w]);
        counters[2] = (unsigned long long)__double_as_longlong(m);
        if (dtp) *dtp = jmin(cfl * m, cap);
    }
}

__global__ void k_set_dt(double* dtp, double dt, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) dtp[k] = dt;
}

// The CFL steps of an ensemble's instances: instance k (workgroup k) reduces the minima of its blocks
// [blk[k], blk[k + 1]); dtp[k] = min(cfl dt_k, cap), counters[2] the smallest dt_k (bits of a non-negative double
// order as integers)
__global__ __launch_bounds__(256) void k_cfl_final_multi(const double* __restrict__ bmin, const int64_t* __restrict__ blk,
                                                         unsigned long long* counters, double* dtp, double cfl, double cap) {
    __shared__ double wmin[4];
    const int k = blockIdx.x;
    double dt = INFINITY;
    for (int64_t b = blk[k] + threadIdx.x; b < blk[k + 1]; b += 256) dt = fmin(dt, bmin[b]);
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wmin[0];
        for (int w = 1; w < 4; ++w) m = fmin(m, wmin[w]);
        atomicMin(&counters[2], (unsigned long long)__double_as_longlong(m));
        if (dtp) dtp[k] = jmin(cfl * m, cap);
    }
}

// ============================================================================================================
// moving grids: stage positions, and the positions between the records and the grid
// ============================================================================================================
// The position part of _stage_predictor! (TimeSteppers.jl:405-425): X = x^n + sum_j (dt a~_ij) K_X[j], a term for
// each non-zero entry a~_ij (b~_j for the final state), boundary particles exactly at x^n; written into the position
// components of the target state's records (the stage geometry its passes read). x^n: the records of U^n.
struct PosTerms {
    const double* KX[kMaxStages];
    double at[kMaxStages];
    int nterm;
    unsigned use;
};
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_stage_pos(StepArgs a, int64_t nblk, const double* __restrict__ Un,
                                                       PosTerms t, double* __restrict__ target) {
    using L = Layout<D, M>;
    const int64_t n = node_index(nblk);
    if (n >= a.g.N) return;
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = Un[n * L::PU + d];
    if (!a.g.isb[n]) {
        const double dt = node_dt(a.dtp, a.inst, n);
#pragma unroll
        for (int j = 0; j < kMaxStages; ++j)
            if (j < t.nterm && ((t.use >> j) & 1u)) {
                const double c = dt * t.at[j];
#pragma unroll
                for (int d = 0; d < D; ++d) x[d] = x[d] + c * t.KX[j][n * D + d];
            }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) target[n * L::PU + d] = x[d];
}
// to_grid: the records' positions into the grid's position columns (set_positions!, GridMovers.jl:129-133), else the
// grid's (wrapped) positions into the records
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_rec_pos(int64_t N, double* __restrict__ rec, double* __restrict__ px,
                                                     double* __restrict__ py, double* __restrict__ pz, int to_grid) {
    using L = Layout<D, M>;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double* col[3] = {px, py, pz};
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (to_grid) col[d][i] = rec[i * L::PU + d];
        else rec[i * L::PU + d] = col[d][i];
    }
}

// ============================================================================================================
// packing, fills, counters
// ============================================================================================================
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_pack(int64_t N, const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, const double* __restrict__ U_aos,
                                                  double* __restrict__ rec) {
    using L = Layout<D, M>;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double r[L::NR];
    r[0] = px[i];
    if (D > 1) r[1] = py[i];
    if (D > 2) r[2] = pz[i];
#pragma unroll
    for (int m = 0; m < M; ++m) r[D + m] = U_aos[i * M + m];
    store_rec<L::NR, L::PU>(rec, i, r);
}
template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_unpack(int64_t N, const double* __restrict__ rec, double* __restrict__ U_aos) {
    using L = Layout<D, M>;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double r[L::NR];
    load_rec<L::NR, L::PU>(rec, i, r);
#pragma unroll
    for (int m = 0; m < M; ++m) U_aos[i * M + m] = r[D + m];
}
__global__ __launch_bounds__(kBlock) void k_fill_f64(double* p, double v, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ __launch_bounds__(kBlock) void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}
// mode 0: the changed count; 1: everything of a new step / CFL estimate
__global__ void k_reset_counters(unsigned long long* c, int mode) {
    if (threadIdx.x == 0) {
        c[0] = 0;
        if (mode == 1) { c[1] = 0; c[2] = 0x7FF0000000000000ull; c[4] = 0; }
    }
}

// total_mass / interior_mass (src/Grid/Volumes.jl:103-117): sum_i V_i U_i as mapreduce with init = zero(State) does
// it, a left fold in index order; thread m folds component m, the workgroup stages the products in LDS.
template <int D, int M>
__global__ __launch_bounds__(256) void k_mass(const double* __restrict__ U, const double* __restrict__ V,
                                              const uint8_t* __restrict__ isb, int64_t N, int interior_only,
                                              double* __restrict__ out) {
    using L = Layout<D, M>;
    constexpr int C = 1024;
    __shared__ double buf[M][C];
    double acc = 0.0;
    for (int64_t c0 = 0; c0 < N; c0 += C) {
        for (int t = threadIdx.x; t < C; t += 256) {
            const int64_t i = c0 + t;
            if (i < N) {
                double r[L::NR];
                load_rec<L::NR, L::PU>(U, i, r);
                const bool skip = interior_only && isb[i];
#pragma unroll
                for (int m = 0; m < M; ++m) buf[m][t] = skip ? 0.0 : V[i] * r[D + m];
            }
        }
        __syncthreads();
        if (threadIdx.x < M) {
            const int m = threadIdx.x, n = (int)(N - c0 < C ? N - c0 : C);
            for (int t = 0; t < n; ++t) acc = acc + buf[m][t];
        }
        __syncthreads();
    }
    if (threadIdx.x < M) out[threadIdx.x] = acc;
}

// ============================================================================================================
// the stepper
// ============================================================================================================
#include "hcc_muscl3.h"

struct hcc_stepper {
    hcc_ctx* ctx = nullptr;
    hcc_grid* grid = nullptr;
    hcc_stepper_desc desc{};
    int D = 0, M = 0, kind = 0, PU = 0, GS = 0;
    int order = 2;             // MUSCL max_order: 2, or 3 (the kernels of hcc_muscl3.h)
    int fv = 0;                // flux loop variant: 0 buffered fit + RusanovFlux, 1 streamed and/or AlignedRusanovFlux
    int64_t N = 0;
    // geometry caches
    double* invg = nullptr; uint8_t* okg = nullptr;
    double* invd = nullptr; uint8_t* okd = nullptr;
    double* dw = nullptr;         // derivative weights per slot, formed only when the geometry is downloaded
    int64_t dw_cap = 0;           // slots dw holds
    int32_t* root = nullptr;
    int32_t *goff = nullptr, *gidx = nullptr;   // outflow ghosts of each interior root (CSR over nodes)
    bool fresh = true;         // U was (re)loaded: the first stage state is formed by the predictor once
    // states (records) and right-hand sides (node-major)
    double* U = nullptr;       // pg.rhos (U^n during a step)
    double* cand = nullptr;    // rho_cand
    double* Y[kMaxStages] = {};
    double* K[kMaxStages] = {};
    double* KI[kMaxStages] = {};   // K_I of the stages (a tableau with an implicit part or an implicit source)
    double* rate = nullptr;        // decay rates of the source
    int src_mode = HCC_SRC_NONE;
    bool imex = false;             // stages formed by k_predict_imex
    bool moving = false;           // a grid mover: stage positions, stage geometry, a search after each step
    // learned troubled-node indicator (hcc_set_indicator) and the collection of its training data
    bool feat_on = false;          // the full gathers store the indicator features (2-D Euler)
    const hcc_mlp* ind = nullptr;  // the network, nullptr: off
    float ind_thr = 0.0f;          // logit threshold
    uint16_t* feat = nullptr;      // N x 16 bf16 features of the last full gather
    float* ind_z = nullptr;        // logits
    unsigned long long* ind_cnt = nullptr;   // particles the indicator lowered in the last step
    int32_t* ind_list = nullptr;   // the indicator's candidates of the last full gather (N) and their count
    unsigned long long* ind_n = nullptr;
    bool collect = false;
    uint16_t* cfeat = nullptr;     // stages x N x 16: each stage's features (last step)
    int8_t* cord = nullptr;        // 2 x stages x N: orders before each stage's flux pass, after its checkpoint
    double* KX[kMaxStages] = {};   // K_X of the stages (moving): the grid velocity w per interior particle, D per node
    double* G = nullptr;
    double* scratch = nullptr; // N * M doubles for up/downloads
    int8_t* orders = nullptr;
    uint8_t *changed = nullptr, *dirty = nullptr, *trig = nullptr, *inadm = nullptr;
    unsigned* dmark = nullptr;
    int32_t *clist = nullptr, *clist2 = nullptr, *dlist = nullptr;
    double* lam = nullptr;
    double* bmin = nullptr;    // per-workgroup minima of the CFL estimate
    double* ext = nullptr;     // neighbourhood extrema of the stage state (fused MOODu1 checks)
    int8_t* snap = nullptr;    // StrictEPD0: orders after the decrement (the forcing reads them)
    int8_t* nbmin = nullptr;   // EPD2: minimum order over each neighbourhood
    uint8_t* mask2 = nullptr;  // EPD2: the first hop of the dirty mask's expansion
    double* rpart = nullptr;   // MOODu2: per-workgroup extrema of the previous stage ...
    double* range = nullptr;   // ... and its global range per component
    double* dtp = nullptr;     // the step's dt on the device
    unsigned long long* counters = nullptr;
    uint8_t* cnb = nullptr;    // flat neighbourhoods of the last full gather (N) and the flux pass's flat blocks
    uint8_t* bflat = nullptr;
    uint8_t* kz = nullptr;     // per particle: which stages' K are exactly zero (N, boundary particles 0xFF) ...
    uint8_t* gflat = nullptr;  // ... and the flat blocks of the gathers of stages 2 .. s
    int32_t* gorder = nullptr; // launch order of the staged gathers and flux passes (k_block_order)
    // Uncaptured steps list a new launch order every kOrderEvery steps: the non-flat region moves by about a spacing
    // a step and a block spans 16 spacings, so an order a few steps old still starts nearly every long workgroup
    // first, and any order gives the same results (every list is a permutation of the blocks)
    bool order_ok = false;      // gorder holds a listed order (set up as the identity; written by an uncaptured step)
    uint32_t order_serial = 0;  // uncaptured steps run
    uint8_t* zfit = nullptr;   // what the fits give for zero accumulators (k_zero_fits, with the moments)
    bool flat_blocks = false;  // the full gathers record cnb and k_block_flat marks the staged flux pass's flat blocks
    // dynamic LDS of the staged full passes (0: the pass reads neighbours from HBM)
    size_t lds_gather = 0, lds_flux = 0, lds_cfl = 0, lds_gcfl = 0;
    // launch choices (hcc_stepper_set_tuning / hcc_stepper_autotune; results are the same bits under every choice):
    // LDS-staged full passes or neighbour reads through the caches, the lanes per particle of the fixed point's
    // kernels (kSplit or twice it), the particles one sweep of their grid covers, and the CFL estimate fused into the
    // stage-1 gather or run as its own pass, and when the first iteration of a fused fixed point runs in the dense form
    // (k_flux_marked): dense > 0 when the same checkpoint of the last step started from at least N / dense changed
    // particles, 0 never, -1 always
    struct Tuning {
        int staged = 1, split = 0, sweep = 65536, fuse_cfl = 1, dense = 64;
    } tune;
    // hcc_step_cfl on a state whose stage 1 is U: the CFL estimate runs inside the stage-1 gather (cfl, cap)
    bool cfl_fused = false;
    double cfl_num = 0, cfl_cap = 0;
    // the last step was an uncaptured fused CFL step with the zero-K bits tracked, and nothing has touched the state
    // since: its kz bits, its last gather's cnb and the CFL block minima describe U, so the next stage-1 gather may
    // keep a block's minimum where the block is flat (StepRunner::run)
    bool cfl_chain = false;
    // statistics of queued steps, read when a getter asks or the ring is full (a new step does not wait)
    struct StepStats {
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        hipEvent_t evp[4 * kMaxStages] = {};   // start/end of each full gather and flux pass
        unsigned long long* h = nullptr;      // coherent pinned copy of the step counters; [7] = dt ...
        unsigned long long* hd = nullptr;     // ... and its device address (k_stats_out writes it)
        int64_t host_iters = 0;
        bool dt_from_device = false;
        bool graph = false;                   // a captured step: no pass events (gather and flux times read 0)
        bool passes = false;                  // the pass events were recorded (hcc_stepper_set_pass_timing)
    };
    static constexpr int kRing = 4;
    StepStats ring[kRing];
    int ring_head = 0, ring_count = 0;
    unsigned long long* h_pub = nullptr;   // count slots 0, 1 and the captured step's overflow flag (slot 2) in
    unsigned long long* d_pub = nullptr;   // coherent pinned host memory, and their device address
    // hipGraph-captured steps (hcc_stepper_set_graph): one executable graph per state buffer that holds U^n (the
    // buffers of U and the candidate swap every step), captured for one (cfl, cap) and one iteration bound per
    // checkpoint; bounds learned from uncaptured steps (the most non-empty fixed-point iterations seen, plus one)
    bool graph_on = false;
    bool graph_pending = false;    // the last queued step was a graph launch whose overflow flag is unread
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const double* gU[2] = {nullptr, nullptr};
    double gcfl = 0, gcap = 0;
    int gbound[kMaxStages + 2] = {};   // iterations captured per checkpoint (index: checkpoint 2 .. stages + 1)
    int gseen[kMaxStages + 2] = {};    // the most non-empty iterations an uncaptured checkpoint needed
    // the count the first fused fixed-point iteration of each checkpoint started from in the last uncaptured step
    // (the dense form's launch choice, tune.dense)
    int64_t fp_first[kMaxStages + 2] = {};
    int64_t g_launches = 0, g_captures = 0, g_reruns = 0;
    double last_ms = 0, last_gather_ms = 0, last_flux_ms = 0;
    int64_t inadmissible = 0, mood_iters = 0;
    // hcc_step returns without waiting for the device; the statistics of its queued steps (ring slots) are read when
    // a getter asks or the ring is full, and summed into the totals
    double last_dt = 0;
    // pass timing: the events around each stage's gather and flux pass are recorded every pass_every-th uncaptured
    // step (0: never); pass_serial counts the uncaptured steps, tot_timed the steps of the totals that were timed
    int pass_every = 1;
    int64_t pass_serial = 0, tot_timed = 0;
    int64_t tot_steps = 0, tot_iters = 0;
    double tot_ms = 0, tot_gather_ms = 0, tot_flux_ms = 0;
    StepArgs args() const;
};

// MOOD checks the Euler substeps inside the flux pass that produces K (flux_node with CHECK) for the strategies whose
// recompute set is one hop (_order_reach 1, no forcing) and the criteria that need no global reduction
static bool mood_fused(const hcc_stepper_desc& d) {
    const bool strat = d.mood_strategy == HCC_STRAT_NONE || d.mood_strategy == HCC_STRAT_EPD0 ||
                       d.mood_strategy == HCC_STRAT_EPD1;
    const bool crit = d.mood_criterion == HCC_CRIT_U1 || d.mood_criterion == HCC_CRIT_PAD ||
                      d.mood_criterion == HCC_CRIT_ONLY;
    return d.substep_ratio > 0.0 && strat && crit;
}

// The step's kernel arguments. The flat gather blocks of stages 2 .. s need the zero-K bits, which the fused
// predictor's passes keep: tracked when those passes form every stage (fused MOOD or none, no IMEX, no moving grid),
// with the staged passes and without the indicator's features (the flat gather does not form them).
static StepArgs step_args(const hcc_stepper* s) {
    StepArgs a = s->args();
    const bool fused_pred = !s->imex && !s->moving && (s->desc.mood_criterion == HCC_CRIT_NONE || mood_fused(s->desc));
    // zero-K bits (the flat gather blocks): not while training data is collected (its copies take every feature)
    if (s->flat_blocks && s->tune.staged && s->lds_gather && !s->collect && fused_pred) a.kz = s->kz;
    a.kzmask = (1u << s->desc.stages) - 1u;
    return a;
}

StepArgs hcc_stepper::args() const {
    StepArgs a{};
    a.g = grid->args();
    a.pde.gamma = desc.gamma;
    for (int d = 0; d < 3; ++d) {
        double mx = 0.0;
        for (int m = 0; m < 8; ++m) {
            a.pde.c[d][m] = desc.adv_speed[d][m];
            if (m < M) mx = std::max(mx, std::fabs(desc.adv_speed[d][m]));
        }
        a.pde.amax[d] = mx;
    }
    a.limiter = desc.limiter; a.mode = desc.limiter_mode; a.strategy = desc.mood_strategy;
    a.criterion = desc.mood_criterion; a.delta = desc.mood_delta;
    a.plateau = desc.mood_plateau;
    a.ratio = desc.mood_ratio;
    {
        double h = grid->geom.desc.dx[0];
        for (int d = 1; d < D; ++d) h = std::min(h, grid->geom.desc.dx[d]);
        a.hmin = h;
    }
    a.streamed = desc.muscl_streamed ? 1 : 0;
    a.src_mode = src_mode;
    a.src_rate = rate;
    a.aligned = desc.flux == HCC_FLUX_ALIGNED_RUSANOV ? 1 : 0;
    a.packed = 0;
    a.range = (desc.mood_criterion == HCC_CRIT_U2 && desc.mood_plateau > 0.0) ? range : nullptr;
    a.invg = invg; a.invd = invd; a.okg = okg; a.okd = okd;
    a.G = G; a.orders = orders; a.changed = changed; a.dirty = dirty; a.trig = trig; a.inadm = inadm;
    a.counters = counters; a.clist = clist; a.clist2 = clist2; a.dlist = dlist; a.dmark = dmark; a.lam = lam;
    a.bmin = bmin;
    a.ext = ext;
    // the full gathers store the neighbourhood extrema of their state for the DMP checks of MOODu1 and MOODu2 (the
    // fused check in the flux pass, or k_mood_flag): the checkpoint's state is the one the stage's gather read
    a.ext_on = (desc.mood_criterion == HCC_CRIT_U1 || desc.mood_criterion == HCC_CRIT_U2) ? 1 : 0;
    a.dtp = dtp;
    // the flux pass launches in the order of its stage's gather (k_block_order after k_block_gflat): the blocks whose
    // gather is not flat are, but for a few, those whose flux pass is not flat either
    if (flat_blocks && tune.staged) { a.cnb = cnb; a.bflat = bflat; a.border = gorder; }
    a.zfit = zfit;
    a.inst = grid->n_inst > 1 ? grid->d_blk_inst : nullptr;
    a.feat = feat;
    a.ind_list = ind_list;
    a.ind_n = ind_n;
    a.hratio = desc.substep_ratio;
    if (moving) {
        a.ale = desc.mover;
        a.mover_alpha = desc.mover_alpha;
        a.KX = KX[0];
        a.g.R_cut = grid->geom.R_sq;   // the stage geometry keeps the step's lists
    }
    return a;
}

// compiled variants: (D, M, kind)
template <int D, int M, int KIND>
struct Variant {
    static hipStream_t st(const hcc_stepper* s) { return s->ctx->stream; }
    static void gather(const hcc_stepper* s, const StepArgs& a, const double* Y, int full, int reset) {
        const int64_t nb = nblocks(s->N);
        if constexpr (kFeat<D, M, KIND>) {
            if (full && s->feat_on) {   // the full pass with the indicator features
                if (s->lds_gather && s->tune.staged)
                    k_gather_staged<D, M, KIND, false, true><<<grid_for(nb), kBlock, s->lds_gather, st(s)>>>(a, Y, reset, nb);
                else if (s->moving)   // the stage geometry of a moving grid: weight 0 beyond R, as the moments
                    k_gather<D, M, KIND, true, true><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Y, full, reset, nb);
                else k_gather<D, M, KIND, false, true><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Y, full, reset, nb);
                block_flat(s, a, full);
                return;
            }
        }
        if (full && s->lds_gather && s->tune.staged)
            k_gather_staged<D, M, KIND, false><<<grid_for(nb), kBlock, s->lds_gather, st(s)>>>(a, Y, reset, nb);
        else if (s->moving) k_gather<D, M, KIND, true><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Y, full, reset, nb);
        else k_gather<D, M, KIND, false><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Y, full, reset, nb);
        block_flat(s, a, full);
    }
    // after a full gather that recorded cnb: the flat blocks of the flux pass that follows
    static void block_flat(const hcc_stepper* s, const StepArgs& a, int full) {
        if (!full || !a.bflat) return;
        const int64_t nb = nblocks(s->N);
        k_block_flat<<<(unsigned)nb, kBlock, 0, st(s)>>>(a.g, a.cnb, a.bflat, nb);
    }
    // the fixed point's kernels: 64 / kSplit nodes per 64-lane workgroup, grid-strided over the device-side count;
    // up to 65536 / GPW workgroups (32 waves a CU for the first, largest lists; a cap of 16384 / GPW measured about
    // 0.7 % slower on the 1 M-node step, the idle workgroups of a short list costing less than the lost parallelism)
    // (tuning: the lanes per particle, kSplit or twice it, and the particles one grid sweep covers)
    template <int G>
    static unsigned split_grid(const hcc_stepper* s) {
        constexpr int GPW = 64 / G;
        return (unsigned)std::min<int64_t>((s->N + GPW - 1) / GPW, s->tune.sweep / GPW);
    }
    template <class F>
    static void with_split(const hcc_stepper* s, F&& f) {
        constexpr int G0 = kSplit<D, M>;
        if (s->tune.split == 2 * G0) f(std::integral_constant<int, 2 * G0>{});
        else f(std::integral_constant<int, G0>{});
    }
    // the full gather of stage 1 with the CFL estimate of its state fused in, then dt = min(cfl * dt_cfl, cap)
    static void gather_cfl(const hcc_stepper* s, const StepArgs& a, const double* Y, int reset, double cfl, double cap) {
        const int64_t nb = nblocks(s->N);
        bool done = false;
        if constexpr (kFeat<D, M, KIND>) {
            if (s->feat_on) {   // with the indicator features
                k_gather_staged<D, M, KIND, true, true><<<grid_for(nb), kBlock, s->lds_gcfl, st(s)>>>(a, Y, reset, nb);
                done = true;
            }
        }
        if (!done) k_gather_staged<D, M, KIND, true><<<grid_for(nb), kBlock, s->lds_gcfl, st(s)>>>(a, Y, reset, nb);
        block_flat(s, a, 1);
        cfl_final(s, s->dtp, cfl, cap);
    }
    static void mark_gather(const hcc_stepper* s, const StepArgs& a, const double* Y, unsigned long long* pub, int p) {
        with_split(s, [&](auto g) {
            constexpr int G = decltype(g)::value;
            k_mark_gather<D, M, KIND, G><<<split_grid<G>(s), 64, 0, st(s)>>>(a, Y, pub, p);
        });
    }
    // a dense fixed-point iteration with parity p (plain marks, then the staged flux pass of the marked particles);
    // false where the flux pass is not staged (the caller runs the list form)
    static bool mood_dense(const hcc_stepper* s, const StepArgs& a, const PredArgs& pr, const double* Y, double* K,
                           unsigned long long* pub, int p) {
        if (!(s->lds_flux && s->tune.staged)) return false;
        const int64_t nb = nblocks(s->N);
        k_mark_dense<D, M><<<(unsigned)std::min<int64_t>(nb, 4096), kBlock, 0, st(s)>>>(a, pub, p);
        if (s->fv) k_flux_marked<D, M, KIND, 1><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, pr, Y, K, p, nb);
        else k_flux_marked<D, M, KIND, 0><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, pr, Y, K, p, nb);
        return true;
    }
    template <int FV>
    // full: 1 every particle, 0 the dirty ones (a pass over all that skips the clean), -1 the dirty ones of an
    // unfused recompute listed first (kSplit lanes a particle)
    static void flux_fv(const hcc_stepper* s, const StepArgs& a, const PredArgs& p, const double* Y, double* K, int full,
                        bool check) {
        const int64_t nb = nblocks(s->N);
        if (full < 0) {
            k_zero_u64<<<1, 64, 0, st(s)>>>(s->counters + 5);
            k_compact_dirty<<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb);
            with_split(s, [&](auto g) {
                constexpr int G = decltype(g)::value;
                k_flux_list<D, M, KIND, FV, false, G><<<split_grid<G>(s), 64, 0, st(s)>>>(a, p, Y, K, 0);
            });
        } else if (full && s->lds_flux && s->tune.staged) {
            if (check) k_flux_staged<D, M, KIND, true, FV><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, p, Y, K, nb);
            else k_flux_staged<D, M, KIND, false, FV><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, p, Y, K, nb);
        } else if (check) k_flux<D, M, KIND, true, FV><<<grid_for(nb), kBlock, 0, st(s)>>>(a, p, Y, K, full, nb);
        else k_flux<D, M, KIND, false, FV><<<grid_for(nb), kBlock, 0, st(s)>>>(a, p, Y, K, full, nb);
    }
    static void flux(const hcc_stepper* s, const StepArgs& a, const PredArgs& p, const double* Y, double* K, int full,
                     bool check) {
        if (s->fv) flux_fv<1>(s, a, p, Y, K, full, check);
        else flux_fv<0>(s, a, p, Y, K, full, check);
    }
    static void flux_list(const hcc_stepper* s, const StepArgs& a, const PredArgs& pr, const double* Y, double* K, int p) {
        with_split(s, [&](auto g) {
            constexpr int G = decltype(g)::value;
            if (s->fv) k_flux_list<D, M, KIND, 1, true, G><<<split_grid<G>(s), 64, 0, st(s)>>>(a, pr, Y, K, p);
            else k_flux_list<D, M, KIND, 0, true, G><<<split_grid<G>(s), 64, 0, st(s)>>>(a, pr, Y, K, p);
        });
    }
    static void flag(const hcc_stepper* s, const StepArgs& a, const double* Yp, const double* cand, const double* Kp,
                     const double* KIp, int full) {
        const int64_t nb = nblocks(s->N);
        k_mood_flag<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Yp, cand, Kp, KIp, full, nb);
    }
    static void predict_imex(const hcc_stepper* s, const StepArgs& a, const ImexTerms& t, double* target) {
        const int64_t nb = nblocks(s->N);
        k_predict_imex<D, M><<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb, s->root, s->U, t, target);
    }
    static void range(const hcc_stepper* s, const double* Y) {
        const int64_t nb = (s->N + kBlock - 1) / kBlock;
        k_range_part<D, M><<<(unsigned)nb, kBlock, 0, st(s)>>>(Y, s->N, s->rpart);
        k_range_final<M><<<1, 1024, 0, st(s)>>>(s->rpart, nb, s->range);
    }
    static void cfl(const hcc_stepper* s, const StepArgs& a, double* dtp, double cfl, double cap) {
        const int64_t nb = nblocks(s->N);
        if (s->lds_cfl && s->tune.staged) {
            k_cfl<D, M, KIND, true><<<grid_for(nb), kBlock, s->lds_cfl, st(s)>>>(a, s->U, nb);
        } else {
            k_lambda<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, s->U, nb);
            k_cfl<D, M, KIND, false><<<grid_for(nb), kBlock, 0, st(s)>>>(a, s->U, nb);
        }
        cfl_final(s, dtp, cfl, cap);
    }
    // the minimum over the workgroup minima: one dt, or one per ensemble instance
    static void cfl_final(const hcc_stepper* s, double* dtp, double cfl, double cap) {
        const hcc_grid* g = s->grid;
        if (g->n_inst > 1)
            k_cfl_final_multi<<<(unsigned)g->n_inst, 256, 0, st(s)>>>(s->bmin, g->d_inst_blk, s->counters, dtp, cfl, cap);
        else k_cfl_final<<<1, 1024, 0, st(s)>>>(s->bmin, nblocks(s->N), s->counters, dtp, cfl, cap);
    }
    static void predict(const hcc_stepper* s, const StepArgs& a, const double* const* Ks, const double* c, int nterm,
                        unsigned use, double* target) {   // c: tableau entries (scaled by the device dt in the kernel)
        const int64_t nb = nblocks(s->N);
        k_predict<D, M><<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb, s->root, s->U, Ks[0], Ks[1], Ks[2], Ks[3], Ks[4], c[0],
                                                             c[1], c[2], c[3], c[4], nterm, use, target);
    }
    static void stage_pos(const hcc_stepper* s, const StepArgs& a, const PosTerms& t, double* target) {
        const int64_t nb = nblocks(s->N);
        k_stage_pos<D, M><<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb, s->U, t, target);
    }
    static void rec_pos(const hcc_stepper* s, double* rec, int to_grid) {
        const hcc_grid* g = s->grid;
        k_rec_pos<D, M><<<(unsigned)((s->N + kBlock - 1) / kBlock), kBlock, 0, st(s)>>>(
            s->N, rec, g->d_pos[0], g->d_pos[1] ? g->d_pos[1] : g->d_pos[0], g->d_pos[2] ? g->d_pos[2] : g->d_pos[0],
            to_grid);
    }
    static void mass(const hcc_stepper* s, int interior_only, double* out_dev) {
        k_mass<D, M><<<1, 256, 0, st(s)>>>(s->U, s->grid->d_vol, s->grid->d_isb, s->N, interior_only, out_dev);
    }
    static void pack(const hcc_stepper* s, const double* U_aos, double* rec) {
        const hcc_grid* g = s->grid;
        k_pack<D, M><<<(unsigned)((s->N + kBlock - 1) / kBlock), kBlock, 0, st(s)>>>(
            s->N, g->d_pos[0], g->d_pos[1] ? g->d_pos[1] : g->d_pos[0], g->d_pos[2] ? g->d_pos[2] : g->d_pos[0], U_aos, rec);
    }
    static void unpack(const hcc_stepper* s, const double* rec, double* U_aos) {
        k_unpack<D, M><<<(unsigned)((s->N + kBlock - 1) / kBlock), kBlock, 0, st(s)>>>(s->N, rec, U_aos);
    }
};

struct VariantOps {
    void (*gather)(const hcc_stepper*, const StepArgs&, const double*, int, int);
    void (*gather_cfl)(const hcc_stepper*, const StepArgs&, const double*, int, double, double);
    void (*mark_gather)(const hcc_stepper*, const StepArgs&, const double*, unsigned long long*, int);
    void (*flux)(const hcc_stepper*, const StepArgs&, const PredArgs&, const double*, double*, int, bool);
    void (*flux_list)(const hcc_stepper*, const StepArgs&, const PredArgs&, const double*, double*, int);
    bool (*mood_dense)(const hcc_stepper*, const StepArgs&, const PredArgs&, const double*, double*,
                       unsigned long long*, int);   // nullptr: the variant has no dense iteration
    void (*flag)(const hcc_stepper*, const StepArgs&, const double*, const double*, const double*, const double*, int);
    void (*predict_imex)(const hcc_stepper*, const StepArgs&, const ImexTerms&, double*);
    void (*range)(const hcc_stepper*, const double*);
    void (*cfl)(const hcc_stepper*, const StepArgs&, double*, double, double);
    void (*predict)(const hcc_stepper*, const StepArgs&, const double* const*, const double*, int, unsigned, double*);
    void (*stage_pos)(const hcc_stepper*, const StepArgs&, const PosTerms&, double*);
    void (*rec_pos)(const hcc_stepper*, double*, int);
    void (*mass)(const hcc_stepper*, int, double*);
    void (*pack)(const hcc_stepper*, const double*, double*);
    void (*unpack)(const hcc_stepper*, const double*, double*);
    int PU, GS;
};
template <int D, int M, int KIND>
static VariantOps make_ops() {
    using V = Variant<D, M, KIND>;
    return VariantOps{&V::gather, &V::gather_cfl, &V::mark_gather, &V::flux, &V::flux_list, &V::mood_dense, &V::flag,
                      &V::predict_imex, &V::range, &V::cfl, &V::predict,
                      &V::stage_pos, &V::rec_pos, &V::mass, &V::pack, &V::unpack, Layout<D, M>::PU, Layout<D, M>::GS};
}
// MUSCL order 3: the passes of hcc_muscl3.h (neighbours through the caches, one lane a particle in the fixed point);
// the predictors, the unfused check, the ranges and the record packing are the order-2 variant's
template <int D, int M, int KIND>
struct Variant3 : Variant<D, M, KIND> {
    using V2 = Variant<D, M, KIND>;
    static hipStream_t st(const hcc_stepper* s) { return s->ctx->stream; }
    static unsigned list_grid(const hcc_stepper* s) {
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>((s->N + 63) / 64, s->tune.sweep / 64));
    }
    static void gather(const hcc_stepper* s, const StepArgs& a, const double* Y, int full, int reset) {
        const int64_t nb = nblocks(s->N);
        if (full && s->lds_gather && s->tune.staged)
            m3::k_gather_staged<D, M, KIND><<<grid_for(nb), kBlock, s->lds_gather, st(s)>>>(a, Y, reset, nb);
        else m3::k_gather<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Y, full, reset, nb);
    }
    static void cfl(const hcc_stepper* s, const StepArgs& a, double* dtp, double cfl, double cap) {
        const int64_t nb = nblocks(s->N);
        k_lambda<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, s->U, nb);
        m3::k_cfl<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, s->U, nb);
        V2::cfl_final(s, dtp, cfl, cap);
    }
    // never fused for order 3 (cfl_fused needs the staged gather); kept for the table: the estimate, then the gather
    static void gather_cfl(const hcc_stepper* s, const StepArgs& a, const double* Y, int reset, double c, double cap) {
        cfl(s, a, s->dtp, c, cap);
        gather(s, a, Y, 1, reset);
    }
    static void mark_gather(const hcc_stepper* s, const StepArgs& a, const double* Y, unsigned long long* pub, int p) {
        m3::k_mark_gather<D, M, KIND><<<list_grid(s), 64, 0, st(s)>>>(a, Y, pub, p);
    }
    // full: 1 every particle, 0 the dirty ones, -1 the dirty ones listed first (k_compact_dirty)
    static void flux(const hcc_stepper* s, const StepArgs& a, const PredArgs& p, const double* Y, double* K, int full,
                     bool check) {
        const int64_t nb = nblocks(s->N);
        if constexpr (m3::kWide3<D, M>) {   // one particle per 16 lanes, the fit's columns spread over the lanes
            const int64_t nq = (s->N + m3::Split3<D, M>::GPW - 1) / m3::Split3<D, M>::GPW;
            if (full < 0) {
                k_zero_u64<<<1, 64, 0, st(s)>>>(s->counters + 5);
                k_compact_dirty<<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb);
                m3::k_flux_wide_list<D, M, KIND, false><<<list_grid(s), 64, 0, st(s)>>>(a, p, Y, K, 0);
            } else if (check) {
                m3::k_flux_wide<D, M, KIND, true><<<grid_for(nq), 64, 0, st(s)>>>(a, p, Y, K, full, nq);
            } else {
                m3::k_flux_wide<D, M, KIND, false><<<grid_for(nq), 64, 0, st(s)>>>(a, p, Y, K, full, nq);
            }
        } else if (full < 0) {
            k_zero_u64<<<1, 64, 0, st(s)>>>(s->counters + 5);
            k_compact_dirty<<<grid_for(nb), kBlock, 0, st(s)>>>(a, nb);
            m3::k_flux_list<D, M, KIND, false><<<list_grid(s), 64, 0, st(s)>>>(a, p, Y, K, 0);
        } else if (full && s->lds_flux && s->tune.staged) {
            if (check) m3::k_flux_staged<D, M, KIND, true><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, p, Y, K, nb);
            else m3::k_flux_staged<D, M, KIND, false><<<grid_for(nb), kBlock, s->lds_flux, st(s)>>>(a, p, Y, K, nb);
        } else if (check) {
            m3::k_flux<D, M, KIND, true><<<grid_for(nb), kBlock, 0, st(s)>>>(a, p, Y, K, full, nb);
        } else {
            m3::k_flux<D, M, KIND, false><<<grid_for(nb), kBlock, 0, st(s)>>>(a, p, Y, K, full, nb);
        }
    }
    static void flux_list(const hcc_stepper* s, const StepArgs& a, const PredArgs& pr, const double* Y, double* K, int p) {
        if constexpr (m3::kWide3<D, M>) m3::k_flux_wide_list<D, M, KIND, true><<<list_grid(s), 64, 0, st(s)>>>(a, pr, Y, K, p);
        else m3::k_flux_list<D, M, KIND, true><<<list_grid(s), 64, 0, st(s)>>>(a, pr, Y, K, p);
    }
    // the unfused check with MOODu2's curvature override
    static void flag(const hcc_stepper* s, const StepArgs& a, const double* Yp, const double* cand, const double* Kp,
                     const double* KIp, int full) {
        const int64_t nb = nblocks(s->N);
        m3::k_mood_flag<D, M, KIND><<<grid_for(nb), kBlock, 0, st(s)>>>(a, Yp, cand, Kp, KIp, full, nb);
    }
};
template <int D, int M, int KIND>
static VariantOps make_ops3() {
    using V = Variant3<D, M, KIND>;
    return VariantOps{&V::gather, &V::gather_cfl, &V::mark_gather, &V::flux, &V::flux_list, nullptr, &V::flag,
                      &V::predict_imex,
                      &V::range, &V::cfl, &V::predict, &V::stage_pos, &V::rec_pos, &V::mass, &V::pack, &V::unpack,
                      Layout<D, M>::PU, m3::L3<D, M>::GS};
}
// LDS bytes of the staged MUSCL3 gather and flux passes (0: the variant is not compiled)
static void lds_sizes3(int D, int M, int kind, int64_t hm, size_t idx_bytes, size_t& gather, size_t& flux) {
    gather = flux = 0;
    auto set = [&](size_t g, size_t f) { gather = g; flux = f; };
    if (kind == HCC_PDE_LINADV && D == 1 && M == 1)
        set(m3::lds_bytes3<1, 1>(hm, idx_bytes, false), m3::lds_bytes3<1, 1>(hm, idx_bytes, true));
    else if (kind == HCC_PDE_LINADV && D == 2 && M == 1)
        set(m3::lds_bytes3<2, 1>(hm, idx_bytes, false), m3::lds_bytes3<2, 1>(hm, idx_bytes, true));
    else if (kind == HCC_PDE_LINADV && D == 2 && M == 2)
        set(m3::lds_bytes3<2, 2>(hm, idx_bytes, false), m3::lds_bytes3<2, 2>(hm, idx_bytes, true));
    else if (kind == HCC_PDE_EULER && D == 2 && M == 4)
        set(m3::lds_bytes3<2, 4>(hm, idx_bytes, false), m3::lds_bytes3<2, 4>(hm, idx_bytes, true));
    else if (kind == HCC_PDE_LINADV && D == 3 && M == 1)
        set(m3::lds_bytes3<3, 1>(hm, idx_bytes, false), m3::lds_bytes3<3, 1>(hm, idx_bytes, true));
    else if (kind == HCC_PDE_EULER && D == 3 && M == 5)
        set(m3::lds_bytes3<3, 5>(hm, idx_bytes, false), m3::lds_bytes3<3, 5>(hm, idx_bytes, true));
}
// MUSCL3 variants: advection (1, 1), (2, 1), (2, 2), (3, 1), 2-D and 3-D Euler
static bool find_ops3(int D, int M, int kind, VariantOps& ops) {
    if (kind == HCC_PDE_LINADV) {
        if (D == 1 && M == 1) { ops = make_ops3<1, 1, HCC_PDE_LINADV>(); return true; }
        if (D == 2 && M == 1) { ops = make_ops3<2, 1, HCC_PDE_LINADV>(); return true; }
        if (D == 2 && M == 2) { ops = make_ops3<2, 2, HCC_PDE_LINADV>(); return true; }
        if (D == 3 && M == 1) { ops = make_ops3<3, 1, HCC_PDE_LINADV>(); return true; }
    } else if (kind == HCC_PDE_EULER) {
        if (D == 2 && M == 4) { ops = make_ops3<2, 4, HCC_PDE_EULER>(); return true; }
        if (D == 3 && M == 5) { ops = make_ops3<3, 5, HCC_PDE_EULER>(); return true; }
    }
    return false;
}
static bool find_ops(int D, int M, int kind, VariantOps& ops) {
    if (kind == HCC_PDE_LINADV) {
        if (D == 1 && M == 1) { ops = make_ops<1, 1, HCC_PDE_LINADV>(); return true; }
        if (D == 2 && M == 1) { ops = make_ops<2, 1, HCC_PDE_LINADV>(); return true; }
        if (D == 2 && M == 2) { ops = make_ops<2, 2, HCC_PDE_LINADV>(); return true; }
        if (D == 3 && M == 1) { ops = make_ops<3, 1, HCC_PDE_LINADV>(); return true; }
    } else if (kind == HCC_PDE_EULER) {
        if (D == 2 && M == 4) { ops = make_ops<2, 4, HCC_PDE_EULER>(); return true; }
        if (D == 3 && M == 5) { ops = make_ops<3, 5, HCC_PDE_EULER>(); return true; }
    }
    return false;
}

static bool find_ops(const hcc_stepper* s, VariantOps& ops) {
    return s->order == 3 ? find_ops3(s->D, s->M, s->kind, ops) : find_ops(s->D, s->M, s->kind, ops);
}

// prepare_geometry! (MUSCL.jl:172-223): the moment inverses and derivative weights of the grid's current positions;
// R_cut: R^2 on a moving grid's stage geometry (drifted pairs weigh 0), +inf on a searched grid
// dw: the derivative weights' buffer (D S_sell), or null: no pass reads them (the streamed and aligned flux loops
// recompute them in registers), so they are formed only for the geometry download.
static void launch_moments(hcc_stepper* s, double R_cut, double* dw = nullptr) {
    hcc_grid* g = s->grid;
    MomentArgs ma{g->args(), s->invg, s->okg, s->invd, s->okd, dw, g->S_sell};
    ma.g.R_cut = R_cut;
    const int64_t nb = nblocks(s->N);
    hipStream_t st = s->ctx->stream;
    if (s->order == 3) {   // the degree-2 reconstruction and the safeguarded degree-3 divergence
        if (s->D == 1) m3::k_moments3<1><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
        else if (s->D == 2) m3::k_moments3<2><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
        else m3::k_moments3<3><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
        return;
    }
    if (s->D == 1) k_moments<1><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
    else if (s->D == 2) k_moments<2><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
    else k_moments<3><<<grid_for(nb), kBlock, 0, st>>>(ma, nb);
    if (s->zfit) {
        const double invL = ma.g.invL;
        if (s->D == 1) k_zero_fits<1><<<(unsigned)nb, kBlock, 0, st>>>(s->invg, s->okg, s->invd, s->okd, ma.g.cnt, invL, s->zfit, s->N);
        else if (s->D == 2) k_zero_fits<2><<<(unsigned)nb, kBlock, 0, st>>>(s->invg, s->okg, s->invd, s->okd, ma.g.cnt, invL, s->zfit, s->N);
        else k_zero_fits<3><<<(unsigned)nb, kBlock, 0, st>>>(s->invg, s->okg, s->invd, s->okd, ma.g.cnt, invL, s->zfit, s->N);
    }
}

extern "C" int hcc_stepper_destroy(hcc_stepper* s) {
    if (!s) return HCC_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (void* p : {(void*)s->invg, (void*)s->okg, (void*)s->invd, (void*)s->okd, (void*)s->dw, (void*)s->root, (void*)s->goff, (void*)s->gidx, (void*)s->U,
                    (void*)s->cand, (void*)s->G, (void*)s->scratch, (void*)s->orders, (void*)s->changed, (void*)s->dirty,
                    (void*)s->trig, (void*)s->inadm, (void*)s->dmark, (void*)s->clist, (void*)s->clist2, (void*)s->dlist,
                    (void*)s->lam, (void*)s->bmin, (void*)s->ext, (void*)s->snap, (void*)s->nbmin, (void*)s->mask2,
                    (void*)s->rpart, (void*)s->range, (void*)s->dtp, (void*)s->counters, (void*)s->feat, (void*)s->ind_z,
                    (void*)s->ind_cnt, (void*)s->cfeat, (void*)s->cord, (void*)s->ind_list, (void*)s->ind_n,
                    (void*)s->cnb, (void*)s->bflat, (void*)s->kz, (void*)s->gflat, (void*)s->zfit,
                    (void*)s->gorder})
        (void)hipFree(p);
    for (int k = 0; k < kMaxStages; ++k) {
        (void)hipFree(s->Y[k]); (void)hipFree(s->K[k]); (void)hipFree(s->KI[k]); (void)hipFree(s->KX[k]);
    }
    (void)hipFree(s->rate);
    for (auto& r : s->ring) {
        if (r.ev0) (void)hipEventDestroy(r.ev0);
        if (r.ev1) (void)hipEventDestroy(r.ev1);
        for (auto& e : r.evp) if (e) (void)hipEventDestroy(e);
        if (r.h) (void)hipHostFree(r.h);
    }
    if (s->h_pub) (void)hipHostFree(s->h_pub);
    for (auto& g : s->gexec) if (g) (void)hipGraphExecDestroy(g);
    delete s;
    return HCC_OK;
}

extern "C" int hcc_stepper_create(hcc_ctx* ctx, hcc_grid* g, const hcc_stepper_desc* d, hcc_stepper** out) {
    if (!ctx || !g || !d || !out) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: NULL argument");
    *out = nullptr;
    const int D = g->geom.D;
    VariantOps ops;
    if (d->muscl_order != 2 && d->muscl_order != 3)
        return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MUSCL orders 2 and 3 are compiled");
    if (!(d->muscl_order == 3 ? find_ops3(D, d->M, d->pde, ops) : find_ops(D, d->M, d->pde, ops)))
        return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: no compiled kernel for this (D, M, PDE, MUSCL order) -- there is no fallback path");
    if (d->muscl_order == 3) {
        if (d->muscl_streamed || d->flux != HCC_FLUX_RUSANOV)
            return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MUSCL3 runs the buffered divergence with RusanovFlux");
        if (d->mover != HCC_MOVER_NONE) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MUSCL3 on a moving grid");
    }
    if (d->stages < 1 || d->stages > kMaxStages) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: 1..5 stages");
    if (d->limiter < HCC_LIM_NONE || d->limiter > HCC_LIM_VK) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: unknown limiter");
    if (d->mood_strategy < HCC_STRAT_NONE || d->mood_strategy > HCC_STRAT_STRICT_EPD0) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MOOD strategy");
    if (d->mood_criterion < HCC_CRIT_NONE || d->mood_criterion > HCC_CRIT_U2) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MOOD criterion");
    if (!(d->mood_plateau >= 0.0) || !std::isfinite(d->mood_plateau)) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: MOODu2 plateau must be finite and >= 0");
    if (d->flux != HCC_FLUX_RUSANOV && d->flux != HCC_FLUX_ALIGNED_RUSANOV) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: numerical flux");
    bool implicit_part = false;
    for (int i = 0; i < d->stages; ++i) {
        implicit_part = implicit_part || d->b_im[i] != 0.0;
        for (int j = 0; j < d->stages; ++j) {
            implicit_part = implicit_part || d->a_im[i][j] != 0.0;
            if (j > i && d->a_im[i][j] != 0.0) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: a_im must be lower triangular");
        }
    }
    if (implicit_part && d->substep_ratio != 0.0)
        return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: a tableau with an implicit part has no SSP coefficient (substep_ratio 0)");
    if (d->n_bc < 0 || d->n_bc > 16) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: at most 16 boundary conditions");
    for (int i = 0; i < d->stages; ++i)
        for (int j = i; j < d->stages; ++j)
            if (d->a_t[i][j] != 0.0) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: a_t must be strictly lower triangular");
    if (d->mover < HCC_MOVER_NONE || d->mover > HCC_MOVER_SCALED_LAGRANGIAN)
        return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: grid mover");
    if (d->mover != HCC_MOVER_NONE) {
        if (g->n_inst > 1) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: a grid mover on an ensemble grid");
        if (!std::isfinite(d->mover_alpha)) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: mover_alpha must be finite");
        if (d->pde == HCC_PDE_LINADV && d->M != 1)
            return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: a moving grid with linear advection needs M = 1");
        for (int b = 0; b < d->n_bc; ++b)
            if (d->bc_kind[b] == HCC_BC_OUTFLOW)
                return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: outflow boundary conditions on a moving grid");
    }
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<hcc_stepper, int (*)(hcc_stepper*)> s(new hcc_stepper(), hcc_stepper_destroy);
    s->ctx = ctx; s->grid = g; s->desc = *d; s->D = D; s->M = d->M; s->kind = d->pde; s->N = g->N;
    s->PU = ops.PU; s->GS = ops.GS;
    s->order = d->muscl_order;
    s->moving = d->mover != HCC_MOVER_NONE;
    s->fv = (d->muscl_streamed || d->flux == HCC_FLUX_ALIGNED_RUSANOV || s->moving) ? 1 : 0;
    s->imex = implicit_part;
    const int64_t N = g->N, M = d->M;
    hipStream_t st = ctx->stream;
    const int BG = s->order == 3 ? m3::bcount(D, 2) : D, BD = s->order == 3 ? m3::bcount(D, 3) + 1 : D + 1;
    HCC_TRY_HIP(ctx, dalloc(&s->invg, (int64_t)BG * BG * N));
    HCC_TRY_HIP(ctx, dalloc(&s->okg, N));
    HCC_TRY_HIP(ctx, dalloc(&s->invd, (int64_t)D * BD * N));
    HCC_TRY_HIP(ctx, dalloc(&s->okd, N));
    HCC_TRY_HIP(ctx, dalloc(&s->root, N));
    HCC_TRY_HIP(ctx, dalloc(&s->U, s->PU * N));
    HCC_TRY_HIP(ctx, dalloc(&s->cand, s->PU * N));
    HCC_TRY_HIP(ctx, dalloc(&s->scratch, M * N));
    for (int k = 0; k < d->stages; ++k) {
        HCC_TRY_HIP(ctx, dalloc(&s->Y[k], s->PU * N));
        HCC_TRY_HIP(ctx, dalloc(&s->K[k], M * N));
        HCC_TRY_HIP(ctx, hipMemsetAsync(s->K[k], 0, sizeof(double) * std::max<int64_t>(M * N, 1), st));
    }
    if (s->moving)
        for (int k = 0; k < d->stages; ++k) {
            HCC_TRY_HIP(ctx, dalloc(&s->KX[k], (int64_t)D * N));
            HCC_TRY_HIP(ctx, hipMemsetAsync(s->KX[k], 0, sizeof(double) * std::max<int64_t>((int64_t)D * N, 1), st));
        }
    if (s->imex)
        for (int k = 0; k < d->stages; ++k) {
            HCC_TRY_HIP(ctx, dalloc(&s->KI[k], M * N));
            HCC_TRY_HIP(ctx, hipMemsetAsync(s->KI[k], 0, sizeof(double) * std::max<int64_t>(M * N, 1), st));
        }
    HCC_TRY_HIP(ctx, dalloc(&s->G, s->GS * N));
    HCC_TRY_HIP(ctx, dalloc(&s->orders, N));
    HCC_TRY_HIP(ctx, dalloc(&s->changed, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dirty, N));
    HCC_TRY_HIP(ctx, dalloc(&s->trig, N));
    HCC_TRY_HIP(ctx, dalloc(&s->inadm, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dmark, N));
    HCC_TRY_HIP(ctx, dalloc(&s->clist, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dlist, N));   // each dirty particle once per fixed-point iteration
    HCC_TRY_HIP(ctx, dalloc(&s->clist2, N));
    HCC_TRY_HIP(ctx, dalloc(&s->lam, (int64_t)D * N));
    HCC_TRY_HIP(ctx, dalloc(&s->bmin, nblocks(N)));
    HCC_TRY_HIP(ctx, dalloc(&s->ext, 2 * M * N));
    HCC_TRY_HIP(ctx, dalloc(&s->snap, N));
    HCC_TRY_HIP(ctx, dalloc(&s->nbmin, N));
    HCC_TRY_HIP(ctx, dalloc(&s->mask2, N));
    HCC_TRY_HIP(ctx, dalloc(&s->rpart, 2 * M * ((N + kBlock - 1) / kBlock)));
    HCC_TRY_HIP(ctx, dalloc(&s->range, M));
    HCC_TRY_HIP(ctx, dalloc(&s->dtp, std::max(2, g->n_inst)));
    HCC_TRY_HIP(ctx, dalloc(&s->counters, 10));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->dmark, 0, sizeof(unsigned) * std::max<int64_t>(N, 1), st));
    for (uint8_t* p : {s->changed, s->dirty, s->trig, s->inadm})
        HCC_TRY_HIP(ctx, hipMemsetAsync(p, 0, std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->orders, s->order, std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->G, 0, sizeof(double) * std::max<int64_t>(s->GS * N, 1), st));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->scratch, 0, sizeof(double) * std::max<int64_t>(M * N, 1), st));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->counters, 0, 10 * sizeof(unsigned long long), st));
    if (g->max_halo > 0 && s->order == 3) {   // the staged MUSCL3 gather and flux passes where the halos fit LDS
        int lds_max = 0;
        HCC_TRY_HIP(ctx, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        size_t lg = 0, lf = 0;
        lds_sizes3(D, (int)M, d->pde, g->max_halo, (size_t)g->max_block_slots * sizeof(uint16_t), lg, lf);
        s->lds_gather = lg <= (size_t)lds_max ? lg : 0;
        s->lds_flux = lf <= (size_t)lds_max ? lf : 0;
    }
    if (g->max_halo > 0 && !s->moving && s->order == 2) {   // LDS staging of the full passes where the block halos fit a workgroup's LDS
        int lds_max = 0;
        HCC_TRY_HIP(ctx, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
        const size_t hm = (size_t)g->max_halo;
        const int NR = D + (int)M, NG = D * (int)M;
        auto fit = [&](size_t bytes) { return bytes <= (size_t)lds_max ? bytes : 0; };
        const size_t idx_bytes = (size_t)g->max_block_slots * sizeof(uint16_t);
        s->lds_gather = fit(2 * lds_index_offset((int64_t)hm, lds_stride(NR, 0), false) + idx_bytes);
        s->lds_flux = fit(2 * lds_index_offset((int64_t)hm, lds_stride(NR, NG), true) + idx_bytes);
        s->lds_cfl = fit(2 * lds_index_offset((int64_t)hm, (2 * D + 1) & ~1, false) + idx_bytes);
        s->lds_gcfl = fit(2 * lds_index_offset((int64_t)hm, lds_stride(NR, 0) + ((D + 1) & ~1), false) + idx_bytes);
    }
    if (s->lds_flux && s->fv == 0 && !s->moving && s->order == 2 && N > 0) {   // flat blocks of the staged flux pass (k_block_flat)
        HCC_TRY_HIP(ctx, dalloc(&s->cnb, N));
        HCC_TRY_HIP(ctx, dalloc(&s->bflat, nblocks(N)));
        HCC_TRY_HIP(ctx, hipMemset(s->cnb, 0, N));
        HCC_TRY_HIP(ctx, hipMemset(s->bflat, 0, nblocks(N)));
        HCC_TRY_HIP(ctx, dalloc(&s->kz, N));
        HCC_TRY_HIP(ctx, dalloc(&s->gflat, nblocks(N)));
        HCC_TRY_HIP(ctx, hipMemset(s->kz, 0xFF, N));   // boundary particles without a root never change
        HCC_TRY_HIP(ctx, hipMemset(s->gflat, 0, nblocks(N)));
        HCC_TRY_HIP(ctx, dalloc(&s->gorder, nblocks(N)));
        HCC_TRY_HIP(ctx, dalloc(&s->zfit, N));   // formed with the moments
        // no block is known to be flat yet: the order starts as the identity
        k_iota<<<(unsigned)nblocks(nblocks(N)), kBlock, 0, st>>>(s->gorder, nblocks(N));
        HCC_TRY_HIP(ctx, hipGetLastError());
        s->flat_blocks = true;
    }
    for (auto& r : s->ring) {
        HCC_TRY_HIP(ctx, hipEventCreate(&r.ev0));
        HCC_TRY_HIP(ctx, hipEventCreate(&r.ev1));
        for (auto& e : r.evp) HCC_TRY_HIP(ctx, hipEventCreate(&e));
        HCC_TRY_HIP(ctx, hipHostMalloc((void**)&r.h, 8 * sizeof(unsigned long long), hipHostMallocCoherent | hipHostMallocMapped));
        HCC_TRY_HIP(ctx, hipHostGetDevicePointer((void**)&r.hd, r.h, 0));
    }
    HCC_TRY_HIP(ctx, hipHostMalloc((void**)&s->h_pub, 4 * sizeof(unsigned long long), hipHostMallocCoherent | hipHostMallocMapped));
    s->h_pub[0] = s->h_pub[1] = s->h_pub[2] = s->h_pub[3] = 0;
    HCC_TRY_HIP(ctx, hipHostGetDevicePointer((void**)&s->d_pub, s->h_pub, 0));
    if (N > 0) {
        ops.pack(s.get(), s->scratch, s->U);      // positions into the state records, zero state
        ops.pack(s.get(), s->scratch, s->cand);   // every state buffer carries the positions (the predictor
        for (int k = 0; k < d->stages; ++k)       // writes state components only)
            ops.pack(s.get(), s->scratch, s->Y[k]);
        launch_moments(s.get(), INFINITY);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    {   // outflow donor roots (host, from the device lists and the wrapped positions)
        std::vector<int64_t> rstart;
        std::vector<int32_t> idx;
        HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
        int rc = download_lists(g, rstart, idx, nullptr);
        if (rc) return rc;
        std::vector<double> pos((size_t)(N * D));
        rc = hcc_grid_get_positions(g, pos.data());
        if (rc) return rc;
        std::vector<int32_t> root;
        outflow_roots(g->geom, N, pos.data(), g->h_isb.data(), g->h_tags.data(), rstart, idx, d->n_bc, d->bc_tag,
                      d->bc_kind, root);
        for (auto& r : root) if (r == -2) r = 0;   // interior entries are never read
        if (N) HCC_TRY_HIP(ctx, hipMemcpy(s->root, root.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
        // the ghosts of each root, for the predictor fused into the flux passes
        std::vector<int32_t> goff((size_t)N + 1, 0), gidx;
        for (int64_t i = 0; i < N; ++i)
            if (g->h_isb[i] && root[i] >= 0) ++goff[(size_t)root[i] + 1];
        for (int64_t i = 0; i < N; ++i) goff[(size_t)i + 1] += goff[(size_t)i];
        gidx.resize((size_t)std::max<int32_t>(goff[(size_t)N], 1));
        {
            std::vector<int32_t> cur(goff.begin(), goff.end() - 1);
            for (int64_t i = 0; i < N; ++i)
                if (g->h_isb[i] && root[i] >= 0) gidx[(size_t)cur[(size_t)root[i]]++] = (int32_t)i;
        }
        HCC_TRY_HIP(ctx, dalloc(&s->goff, N + 1));
        HCC_TRY_HIP(ctx, dalloc(&s->gidx, (int64_t)gidx.size()));
        HCC_TRY_HIP(ctx, hipMemcpy(s->goff, goff.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice));
        HCC_TRY_HIP(ctx, hipMemcpy(s->gidx, gidx.data(), sizeof(int32_t) * gidx.size(), hipMemcpyHostToDevice));
    }
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    *out = s.release();
    return HCC_OK;
}

static int graph_settle(hcc_stepper* s);

extern "C" int hcc_state_upload(hcc_stepper* s, const double* U) {
    if (!s || (!U && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (!s->N) return HCC_OK;
    if (int rc = graph_settle(s)) return rc;   // a pending captured step completes (or is re-run) on the old state
    VariantOps ops;
    find_ops(s, ops);
    HCC_TRY_HIP(ctx, hipMemcpyAsync(s->scratch, U, sizeof(double) * s->N * s->M, hipMemcpyHostToDevice, ctx->stream));
    // every state buffer: ghosts without an outflow root keep pg.rhos in all stages and are never written again
    ops.pack(s, s->scratch, s->U);
    ops.pack(s, s->scratch, s->cand);
    for (int k = 0; k < s->desc.stages; ++k) ops.pack(s, s->scratch, s->Y[k]);
    s->fresh = true;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

extern "C" int hcc_state_download(hcc_stepper* s, double* U) {
    if (!s || (!U && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (!s->N) return HCC_OK;
    if (int rc = graph_settle(s)) return rc;
    VariantOps ops;
    find_ops(s, ops);
    ops.unpack(s, s->U, s->scratch);
    HCC_TRY_HIP(ctx, hipMemcpyAsync(U, s->scratch, sizeof(double) * s->N * s->M, hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

extern "C" int hcc_time_step(hcc_stepper* s, double* dt) {
    if (!s || !dt) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    VariantOps ops;
    find_ops(s, ops);
    k_reset_counters<<<1, 64, 0, ctx->stream>>>(s->counters, 1);
    if (s->N) ops.cfl(s, s->args(), nullptr, 0.0, 0.0);
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, s->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const unsigned long long bits = ctx->h_pinned[2];
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *dt = v;
    return HCC_OK;
}

extern "C" int hcc_time_steps(hcc_stepper* s, double* dts) {
    if (!s || !dts) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    const int n = s->grid->n_inst;
    if (n == 1) return hcc_time_step(s, dts);
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    VariantOps ops;
    find_ops(s, ops);
    k_reset_counters<<<1, 64, 0, ctx->stream>>>(s->counters, 1);
    ops.cfl(s, s->args(), s->dtp, 1.0, INFINITY);   // dtp[k] = min(1 dt_k, inf) = dt_k; the next step sets dtp again
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipMemcpyAsync(dts, s->dtp, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

// host driver of one step (TimeSteppers.jl:531-577)
namespace {
struct StepRunner {
    hcc_stepper* s;
    VariantOps ops;
    StepArgs a;
    hipEvent_t* evp = nullptr;   // the pass events of this step's statistics slot (nullptr: passes not timed)
    int64_t nb = 0;
    int iters = 0;
    // capture: the fused fixed point of checkpoint i runs gbound[i] iterations queued without host decisions, then
    // k_fixed_point_overflow; no pass events. Uncaptured: nonempty[i] counts the iterations that had work.
    bool graph = false;
    bool ordered = false;   // this step has listed its launch order (from the first gather-flag kernel it ran)
    bool reorder = true;    // this step lists a new launch order (captured steps always do)
    bool chain0 = false;    // the step continues an unbroken chain of steps (PredArgs::first of its full flux passes)
    int nonempty[kMaxStages + 2] = {};
    bool substeps() const { return s->desc.substep_ratio > 0.0 && s->desc.mood_criterion != HCC_CRIT_NONE; }
    // MOOD checks the Euler substep inside the flux pass that produces K (flux_node with CHECK); not with an implicit
    // part, whose stages and substeps involve K_I, nor on a moving grid, whose stage positions and geometry are formed
    // between the passes
    bool fused() const { return mood_fused(s->desc) && !s->imex && !s->moving; }
    bool epd2() const { return s->desc.mood_strategy == HCC_STRAT_EPD2; }
    // the arguments of a flux pass: with EPD2 the interface order of an edge is the minimum of the two particles'
    // neighbourhood minima, i.e. the edge-minimum rule of EPD1 applied to nbmin (computed before the pass)
    // kst: the stage (0-based) whose right-hand side the pass computes, for its K_X on a moving grid. only: a
    // recompute of the fixed point refreshes nbmin where it can have changed since the last pass of the same K, the
    // dropped particles and their neighbours (the first hop of the expansion, mask2); null: every particle.
    StepArgs flux_args(int kst, const uint8_t* only = nullptr) {
        StepArgs f = a;
        if (s->moving) f.KX = s->KX[kst];
        if (!epd2()) return f;
        // the aligned flux and MUSCL3 (its divergence degree follows the particle's own order) also need the true
        // orders: packed as min | order << 4
        const int pack = a.aligned || s->order == 3;
        k_nbmin<<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, s->nbmin, pack, only, nb);
        f.orders = s->nbmin;
        f.strategy = HCC_STRAT_EPD1;
        f.packed = pack;
        return f;
    }
    // the predictor runs inside the flux passes unless MOOD checks the stored candidate itself
    bool fuse_pred() const { return !s->imex && !s->moving && (fused() || s->desc.mood_criterion == HCC_CRIT_NONE); }
    bool y1_in_U = false;
    // state of stage k (1-based): U itself once the boundary conditions hold on it (every step after the first since
    // the state was loaded), else the predictor's copy in Y[0]
    const double* stage_state(int k) const { return k == 1 ? (y1_in_U ? s->U : s->Y[0]) : s->Y[k - 1]; }

    // the terms of form_candidate! (TimeSteppers.jl:365-397) for checkpoint i (1-based; i = s + 1 is the final state):
    // the K buffers and tableau entries; the kernels multiply each entry by the device-resident dt (dt * a_ij)
    void terms(int i, const double** Ks, double* at, int& n_prev, unsigned& use, double*& target) const {
        const int st = s->desc.stages;
        const bool fin = i > st;
        n_prev = fin ? st : i - 1;
        use = 0;
        for (int j = 0; j < kMaxStages; ++j) { Ks[j] = s->K[0]; at[j] = 0.0; }
        for (int j = 0; j < n_prev; ++j) {
            Ks[j] = s->K[j];
            at[j] = fin ? s->desc.b_t[j] : s->desc.a_t[i - 1][j];
            if (at[j] != 0.0) use |= 1u << j;
        }
        target = fin ? s->cand : s->Y[i - 1];
    }
    void predict(int i) {
        const double* Ks[kMaxStages];
        double at[kMaxStages];
        int n_prev;
        unsigned use;
        double* target;
        terms(i, Ks, at, n_prev, use, target);
        if (s->imex) {   // the IMEX stage: both parts of the tableau, the diagonal solve and K_I[i]
            const int st = s->desc.stages;
            const bool fin = i > st;
            ImexTerms t{};
            for (int j = 0; j < kMaxStages; ++j) { t.KE[j] = s->K[0]; t.KI[j] = s->KI[0]; }
            for (int j = 0; j < n_prev; ++j) {
                t.KE[j] = s->K[j];
                t.KI[j] = s->KI[j];
                t.at[j] = fin ? s->desc.b_t[j] : s->desc.a_t[i - 1][j];
                t.ai[j] = fin ? s->desc.b_im[j] : s->desc.a_im[i - 1][j];
            }
            t.nterm = n_prev;
            t.stage = fin ? 0 : 1;
            t.aii = fin ? 0.0 : s->desc.a_im[i - 1][i - 1];
            t.KIout = fin ? nullptr : s->KI[i - 1];
            ops.predict_imex(s, a, t, target);
        } else {
            ops.predict(s, a, Ks, at, n_prev, use, target);
        }
        if (s->moving) {   // X_i (x_cand for the final state) into the target's records
            PosTerms t{};
            for (int j = 0; j < kMaxStages; ++j) { t.KX[j] = s->KX[j < n_prev ? j : 0]; t.at[j] = at[j]; }
            t.nterm = n_prev;
            t.use = use;
            ops.stage_pos(s, a, t, target);
        }
    }
    // the orders into slot k of the collection (k < stages: before stage k+1's flux pass; else after the checkpoint of
    // stage k - stages + 1)
    void snap_orders(int k) {
        (void)hipMemcpyAsync(s->cord + (int64_t)k * s->N, s->orders, s->N, hipMemcpyDeviceToDevice, s->ctx->stream);
    }
    // after stage i's full gather (features stored): the collection's copies, and the indicator's decision
    hipError_t after_gather(int i) {
        hipStream_t str = s->ctx->stream;
        if (s->collect) {
            hipError_t e = hipMemcpyAsync(s->cfeat + (int64_t)(i - 1) * s->N * 16, s->feat, sizeof(uint16_t) * 16 * s->N,
                                          hipMemcpyDeviceToDevice, str);
            if (e != hipSuccess) return e;
            snap_orders(i - 1);
        }
        if (s->ind) {   // the network over the candidates only (their count stays on the device)
            // logits, and the decision applied to the orders in the same kernel
            mlp_forward_list(s->ind, s->feat, s->ind_list, s->ind_n, s->N, s->ind_z, s->orders, s->ind_thr, s->ind_cnt);
        }
        return hipGetLastError();
    }
    // _apply_stage_geometry! (TimeSteppers.jl:520-524) with the frozen topology: the stage positions (held by the
    // stage state's records) into the grid, and the geometry caches of the stage geometry
    void apply_stage_geometry(const double* Y) {
        ops.rec_pos(s, const_cast<double*>(Y), 1);
        launch_moments(s, s->grid->geom.R_sq);
    }
    PredArgs pred_args(int i, bool count) const {
        PredArgs p{};
        terms(i, p.K, p.at, p.nterm, p.use, p.target);
        p.Un = s->U;
        p.goff = s->goff;
        p.gidx = s->gidx;
        p.dtp = s->dtp;
        p.inst = a.inst;
        p.on = 1;
        p.count = count ? 1 : 0;
        p.first = 0;
        p.kmask = (1u << s->desc.stages) - 1u;
        p.pass1 = 0;
        return p;
    }
    // the unfused check's count of changed particles (counters[0]), published into pinned slot 0 and polled
    hipError_t read_changed(unsigned long long& n) {
        *(volatile unsigned long long*)&s->h_pub[0] = kPending;
        k_publish<<<1, 64, 0, s->ctx->stream>>>(s->counters, s->d_pub);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return wait_published(0, n);
    }
    // One iteration of the fused fixed point. Every count it needs is read on the device, so iterations queue without
    // a host round trip, and an iteration enqueued after the fixed point is reached does nothing. Its first kernel
    // publishes the count it starts from into pinned slot q, which holds kPending until then (a count is never
    // kPending): the host polls the slot instead of waiting for an event, which would also hold the stream between the
    // two kernels.
    static constexpr unsigned long long kPending = ~0ull;
    hipError_t mood_iteration(const PredArgs& p, const double* Yp, double* Kp, int q, bool dense = false) {
        *(volatile unsigned long long*)&s->h_pub[q] = kPending;   // its last writer has published (waited on)
        if (!(dense && ops.mood_dense && ops.mood_dense(s, a, p, Yp, Kp, s->d_pub + q, q))) {
            ops.mark_gather(s, a, Yp, s->d_pub + q, q);
            ops.flux_list(s, a, p, Yp, Kp, q);
        }
        return hipGetLastError();
    }
    // The first iteration of a fixed point runs in the dense form (k_flux_marked) when the same checkpoint of the last
    // step started from at least N / tune.dense changed particles: then nearly every particle is recomputed, which
    // the staged pass does at a fraction of the list kernels' cost. The host does not know the count before it queues
    // the iteration; the last step's count is a good guess (the flagged set moves little from step to step), and
    // either form gives the same bits.
    bool dense_first(int i) const {
        const int d = s->tune.dense;
        return d < 0 || (d > 0 && s->fp_first[i] * d >= s->N);
    }
    // The count published in slot q. Polls for up to a second; then waits for the stream, so that a failed launch
    // surfaces as its error instead of a hang. The first polls spin (the usual wait is a few to a few tens of
    // microseconds, one fixed-point iteration); after kSpinPolls the thread yields between polls, so that steppers
    // driven from several threads, or other work on the host, share the cores instead of each burning one.
    static constexpr unsigned kSpinPolls = 4096;
    hipError_t wait_published(int q, unsigned long long& n) {
        volatile unsigned long long* v = &s->h_pub[q];
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned spin = 0; (n = *v) == kPending; ++spin) {
            if (spin >= kSpinPolls) std::this_thread::yield();
            if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
                hipError_t e = hipStreamSynchronize(s->ctx->stream);
                if (e != hipSuccess) return e;
                if ((n = *v) == kPending) return hipErrorUnknown;   // the stream drained without the count
                return hipSuccess;
            }
        }
        return hipSuccess;
    }
    // mood_checkpoint! (TimeSteppers.jl:454-477). Fused path: the first (full) check already ran inside stage i-1's
    // flux pass, which also formed the candidate Y_i and counted the accepted inadmissible candidates; every
    // recompute of K_{i-1} on the dirty set re-checks exactly that set and re-forms its Y_i (the check reads only
    // Y_{i-1} and K_{i-1}). Iteration j publishes the count n_{j-1} its predecessor left; while the host waits for it,
    // the rest of iteration j runs, and iteration j + 1 is enqueued only if n_{j-1} > 0 (if n_j turns out 0,
    // iteration j + 1 does nothing).
    hipError_t checkpoint(int i) {
        if (s->desc.mood_criterion == HCC_CRIT_NONE) {   // Y_i formed by the flux pass, or here (IMEX)
            if (!fuse_pred()) predict(i);
            return hipGetLastError();
        }
        const double* Yp = stage_state(i - 1);
        double* Kp = s->K[i - 2];
        hipStream_t str = s->ctx->stream;
        if (fused()) {
            const PredArgs p = pred_args(i, true);
            if (graph) {   // a fixed number of iterations (an iteration after convergence does nothing), then the check
                for (int k = 0; k < s->gbound[i]; ++k) {
                    hipError_t e = mood_iteration(p, Yp, Kp, k & 1);
                    if (e != hipSuccess) return e;
                }
                k_fixed_point_overflow<<<1, 64, 0, str>>>(s->counters, s->gbound[i] & 1, s->d_pub + 2);
                return hipGetLastError();
            }
            hipError_t e = mood_iteration(p, Yp, Kp, 0, dense_first(i));
            if (e != hipSuccess) return e;
            for (int k = 0;; ++k) {
                const int q = k & 1;
                unsigned long long n = 0;
                if ((e = wait_published(q, n)) != hipSuccess) return e;
                if (k == 0) s->fp_first[i] = (int64_t)n;
                if (n == 0) break;   // n_k = 0: iteration k + 1 does nothing
                ++nonempty[i];
                if ((e = mood_iteration(p, Yp, Kp, q ^ 1)) != hipSuccess) return e;
            }
            return hipGetLastError();
        }
        // Unfused: the check reads the stored candidate, or the Euler substep formed from Y_{i-1} and K_{i-1}; the
        // candidate Y_i is formed after the fixed point (it reads only K_1 .. K_{i-1}).
        const PredArgs off{};
        const bool sub = substeps();
        const double* Kchk = sub ? Kp : nullptr;
        if (a.range) ops.range(s, Yp);   // MOODu2's global range of Y_{i-1} (_global_range, MOOD.jl:206-213)
        const int reach = epd2() ? 2 : 1;   // _order_reach (MUSCL_hooks.jl:10-11)
        int full = 1;
        for (;;) {
            if (!sub) predict(i);
            k_reset_counters<<<1, 64, 0, str>>>(s->counters, 0);
            const double* KIchk = (sub && s->src_mode == HCC_SRC_DECAY_IMPLICIT) ? s->KI[i - 2] : nullptr;
            ops.flag(s, a, Yp, i > s->desc.stages ? s->cand : s->Y[i - 1], Kchk, KIchk, full);
            unsigned long long n = 0;
            hipError_t e = read_changed(n);
            if (e != hipSuccess) return e;
            if (n == 0) break;
            ++iters;
            if (s->desc.mood_strategy == HCC_STRAT_STRICT_EPD0) {   // forcing from the post-decrement snapshot
                k_copy_u8<<<(unsigned)nb, kBlock, 0, str>>>(reinterpret_cast<const uint8_t*>(s->orders),
                                                             reinterpret_cast<uint8_t*>(s->snap), s->N);
                k_force<<<grid_for(nb), kBlock, 0, str>>>(a, s->snap, nb);
                k_copy_u8<<<(unsigned)nb, kBlock, 0, str>>>(s->dirty, s->changed, s->N);
            }
            if (reach == 1) {
                if ((e = launch_expand(a, s->changed, s->dirty, s->N, nb, str)) != hipSuccess) return e;
            } else {
                if ((e = launch_expand(a, s->changed, s->mask2, s->N, nb, str)) != hipSuccess) return e;
                if ((e = launch_expand(a, s->mask2, s->dirty, s->N, nb, str)) != hipSuccess) return e;
            }
            ops.gather(s, a, Yp, 0, 0);
            // a small dirty set (the changed particles and their reach-hop neighbourhoods, estimated at 16 a hop) is
            // listed and recomputed with kSplit lanes a particle; a large one (an implicit stage can drop nearly every
            // particle) by the pass over all particles, which skips the clean ones
            const int list = (int64_t)n * 16 * reach <= s->N / 4 ? -1 : 0;
            ops.flux(s, flux_args(i - 2, reach == 2 ? s->mask2 : nullptr), off, Yp, Kp, list, false);
            full = 0;
        }
        if (sub) predict(i);
        k_count_inadm<<<grid_for(nb), kBlock, 0, str>>>(a, nb);
        return hipGetLastError();
    }
    // a new launch order from this stage's flat flags
    void list_order(int64_t nb) {
        k_block_order<<<1, kBlock, 0, s->ctx->stream>>>(s->gflat, s->gorder, nb);
        if (!graph) s->order_ok = true;
    }
    hipError_t run() {
        hcc_ctx* ctx = s->ctx;
        const int st = s->desc.stages;
        hipStream_t str = ctx->stream;
        nb = nblocks(s->N);
        y1_in_U = !s->fresh && !s->imex;
        // this uncaptured fused CFL step continues an unbroken chain (cfl_chain: the last step was one too, with the
        // zero-K bits tracked, and nothing has touched the state since; an upload makes the next step unfused)
        chain0 = s->cfl_chain && s->cfl_fused && a.kz != nullptr && !graph;
        reorder = graph || !s->order_ok || s->order_serial % kOrderEvery == 0;
        if (!graph) ++s->order_serial;
        // The stage-1 gather resets orders, flags and counters. Gradients: every full gather pass writes each interior
        // particle's record; ghosts keep the zeros written at set-up (the reference's gather skips boundary
        // particles). The full gather also clears the changed count for the check of the flux pass.
        if (s->ind) k_zero_u64<<<1, 64, 0, str>>>(s->ind_cnt);   // a kernel, not a memset: capturable
        for (int i = 1; i <= st; ++i) {
            if (i == 1) {
                if (!y1_in_U) predict(1);   // Y_1 = pg.rhos with the boundary conditions applied
            } else {
                hipError_t e = checkpoint(i);
                if (e != hipSuccess) return e;
                if (s->collect) snap_orders(st + i - 2);   // after stage i-1's checkpoint
                if (s->moving) apply_stage_geometry(stage_state(i));
            }
            const double* Y = stage_state(i);
            if (evp) (void)hipEventRecord(evp[4 * (i - 1) + 0], str);
            if (s->feat_on) k_zero_u64<<<1, 64, 0, str>>>(s->ind_n);
            if (i == 1 && s->cfl_fused && a.kz && s->cfl_chain && !graph) {
                // U = Y_s of the last step plus dt sum b_j K_j: on a block whose halo had every K of that step +-0
                // and constant neighbourhoods in Y_s, U equals Y_s (and the last step's U) in value, so the block's
                // gather is flat and its CFL minimum is the one the last stage-1 gather stored
                k_block_gflat<<<(unsigned)nb, kBlock, 0, str>>>(a.g, s->cnb, s->kz, (1u << st) - 1u, s->gflat, nb);
                if (reorder) list_order(nb);
                ordered = true;
                StepArgs ag = a;
                ag.gflat = s->gflat;
                ag.gorder = s->gorder;
                ops.gather_cfl(s, ag, Y, 1, s->cfl_num, s->cfl_cap);
            } else if (i == 1 && s->cfl_fused) ops.gather_cfl(s, a, Y, 1, s->cfl_num, s->cfl_cap);
            else if (i >= 2 && a.kz) {   // the flat blocks of this gather: K_1 .. K_{i-1} zero on their halos
                k_block_gflat<<<(unsigned)nb, kBlock, 0, str>>>(a.g, s->cnb, s->kz, (1u << (i - 1)) - 1u, s->gflat, nb);
                // one launch order a step: the non-flat region moves by about a spacing a step
                if (!ordered && reorder) list_order(nb);
                ordered = true;
                StepArgs ag = a;
                ag.gflat = s->gflat;
                ag.gorder = s->gorder;
                ops.gather(s, ag, Y, 1, 0);
            } else ops.gather(s, a, Y, 1, i == 1);
            if (s->feat_on) {
                hipError_t e = after_gather(i);
                if (e != hipSuccess) return e;
            }
            if (evp) (void)hipEventRecord(evp[4 * (i - 1) + 1], str);   // end of the gather = start of the flux pass
            PredArgs p = fuse_pred() ? pred_args(i + 1, fused()) : PredArgs{};
            p.first = chain0 ? 1 : 0;
            p.pass1 = 1;
            ops.flux(s, flux_args(i - 1), p, Y, s->K[i - 1], 1, fused());
            if (evp) (void)hipEventRecord(evp[4 * (i - 1) + 3], str);
        }
        hipError_t e = checkpoint(st + 1);
        if (e != hipSuccess) return e;
        if (s->collect) snap_orders(2 * st - 1);
        if (!graph) {   // a captured step swaps when its graph is launched
            std::swap(s->U, s->cand);        // pg.rhos <- rho_cand (its ghosts hold their roots' values)
            s->fresh = false;
        }
        return hipGetLastError();
    }
};
}  // namespace

// Waits for the oldest queued step's work and folds its statistics into the step fields and the totals.
static int fold_oldest(hcc_stepper* s) {
    hcc_ctx* ctx = s->ctx;
    auto& r = s->ring[(s->ring_head - s->ring_count + hcc_stepper::kRing) % hcc_stepper::kRing];
    HCC_TRY_HIP(ctx, hipEventSynchronize(r.ev1));
    --s->ring_count;
    std::memcpy(&s->last_dt, r.h + 7, sizeof(double));
    if (r.dt_from_device && !(s->last_dt > 0 && std::isfinite(s->last_dt)))
        return set_err(ctx, HCC_E_ARG, "hcc_step_cfl: the CFL step was not positive and finite; the state is invalid");
    s->inadmissible = (int64_t)r.h[1];
    s->mood_iters = r.host_iters + (int64_t)r.h[4];
    float ms = 0;
    HCC_TRY_HIP(ctx, hipEventElapsedTime(&ms, r.ev0, r.ev1));
    s->last_ms = ms;
    s->last_gather_ms = s->last_flux_ms = 0;
    for (int i = 0; i < s->desc.stages && r.passes; ++i) {
        float gm = 0, fm = 0;
        HCC_TRY_HIP(ctx, hipEventElapsedTime(&gm, r.evp[4 * i + 0], r.evp[4 * i + 1]));
        HCC_TRY_HIP(ctx, hipEventElapsedTime(&fm, r.evp[4 * i + 1], r.evp[4 * i + 3]));
        s->last_gather_ms += gm;
        s->last_flux_ms += fm;
    }
    s->tot_steps += 1;
    s->tot_timed += r.passes ? 1 : 0;
    s->tot_iters += s->mood_iters;
    s->tot_ms += s->last_ms;
    s->tot_gather_ms += s->last_gather_ms;
    s->tot_flux_ms += s->last_flux_ms;
    return HCC_OK;
}

// Folds every queued step (the getters report the last one).
static int finish_pending(hcc_stepper* s) {
    if (int rc = graph_settle(s)) return rc;
    while (s->ring_count > 0)
        if (int rc = fold_oldest(s)) return rc;
    return HCC_OK;
}

// The end of a moving grid's step (TimeSteppers.jl:571-575): the final positions x_cand (in the new state's records)
// into the grid, a full search with the periodic wrap (pg.neighbor(pg)), the wrapped positions back into the state's
// records, the geometry caches of the new lists and ShepardVolumes() (one sweep). Waits for the device (the search
// sizes its lists on the host).
static int moving_step_end(hcc_stepper* s, const VariantOps& ops) {
    hcc_ctx* ctx = s->ctx;
    hcc_grid* g = s->grid;
    ops.rec_pos(s, s->U, 1);
    HCC_TRY_HIP(ctx, hipGetLastError());
    if (int rc = research(g)) return rc;
    ops.rec_pos(s, s->U, 0);
    launch_moments(s, INFINITY);
    HCC_TRY_HIP(ctx, hipGetLastError());
    return grid_volumes(g, 1);
}

// Queues one step whose dt is already on the device (dtp), then the copies its statistics are read from.
static int launch_step(hcc_stepper* s, const char* who, bool dt_from_device) {
    hcc_ctx* ctx = s->ctx;
    if (s->ring_count == hcc_stepper::kRing)
        if (int rc = fold_oldest(s)) return rc;
    auto& slot = s->ring[s->ring_head];
    StepRunner r;
    r.s = s;
    find_ops(s, r.ops);
    r.a = step_args(s);
    slot.passes = s->pass_every > 0 && s->pass_serial++ % s->pass_every == 0;
    r.evp = slot.passes ? slot.evp : nullptr;
    HCC_TRY_HIP(ctx, hipEventRecord(slot.ev0, ctx->stream));
    hipError_t e = r.run();
    s->cfl_chain = s->cfl_fused && r.a.kz != nullptr && e == hipSuccess;
    if (e != hipSuccess) { ctx->err = std::string(who) + ": " + hipGetErrorString(e); return HCC_E_HIP; }
    if (s->moving)
        if (int rc = moving_step_end(s, r.ops)) return rc;
    k_stats_out<<<1, 64, 0, ctx->stream>>>(s->counters, s->dtp, slot.hd);
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipEventRecord(slot.ev1, ctx->stream));
    for (int i = 0; i < kMaxStages + 2; ++i) s->gseen[i] = std::max(s->gseen[i], r.nonempty[i]);
    slot.host_iters = r.iters;
    slot.dt_from_device = dt_from_device;
    slot.graph = false;
    s->ring_head = (s->ring_head + 1) % hcc_stepper::kRing;
    ++s->ring_count;
    return HCC_OK;
}

// ---- hipGraph-captured steps -------------------------------------------------------------------------------------
static void drop_graphs(hcc_stepper* s) {
    for (int k = 0; k < 2; ++k) {
        if (s->gexec[k]) (void)hipGraphExecDestroy(s->gexec[k]);
        s->gexec[k] = nullptr;
        s->gU[k] = nullptr;
    }
}

// A step can be captured when nothing in it waits for the host: the fused fixed point (or no MOOD), no IMEX stage, no
// moving grid (its search sizes lists on the host), no training-data copies, the CFL estimate fused into stage 1.
static bool graph_eligible(const hcc_stepper* s) {
    const bool mood_ok = s->desc.mood_criterion == HCC_CRIT_NONE || mood_fused(s->desc);
    return mood_ok && !s->imex && !s->moving && !s->collect && !s->fresh && s->cfl_fused;
}

// The last step, if it was a graph launch: waits for it and reads its overflow flag. A fixed point that had not
// converged within its captured iterations left later stages computed from an unconverged state; the step is then
// undone and re-run uncaptured (U^n is intact: a step writes only the candidate buffer, which the host swaps in), its
// statistics slot is dropped, and the bounds grow to what the re-run needed.
static int graph_settle(hcc_stepper* s) {
    if (!s->graph_pending) return HCC_OK;
    hcc_ctx* ctx = s->ctx;
    s->graph_pending = false;
    auto& slot = s->ring[(s->ring_head - 1 + hcc_stepper::kRing) % hcc_stepper::kRing];
    HCC_TRY_HIP(ctx, hipEventSynchronize(slot.ev1));
    volatile unsigned long long* flag = &s->h_pub[2];
    if (*flag == 0) return HCC_OK;
    *flag = 0;
    std::swap(s->U, s->cand);
    s->ring_head = (s->ring_head - 1 + hcc_stepper::kRing) % hcc_stepper::kRing;
    --s->ring_count;
    drop_graphs(s);
    ++s->g_reruns;
    s->cfl_fused = true;   // the same step: stage 1 reads U, the CFL estimate fused into its gather
    const int rc = launch_step(s, "hcc_step_cfl", true);
    s->cfl_fused = false;
    return rc;
}

// One step of hcc_step_cfl as a graph launch: captured (once per U buffer, cfl, cap and bounds), then launched between
// the statistics events. The bounds are the iterations uncaptured steps needed plus one.
static int graph_step(hcc_stepper* s) {
    hcc_ctx* ctx = s->ctx;
    const int st = s->desc.stages;
    s->cfl_chain = false;
    bool stale = s->gcfl != s->cfl_num || s->gcap != s->cfl_cap;
    for (int i = 2; i <= st + 1; ++i) stale = stale || s->gbound[i] < s->gseen[i] + 1;
    if (stale) {
        drop_graphs(s);
        for (int i = 2; i <= st + 1; ++i) s->gbound[i] = s->gseen[i] + 1;
        s->gcfl = s->cfl_num;
        s->gcap = s->cfl_cap;
    }
    int k = s->gU[0] == s->U ? 0 : s->gU[1] == s->U ? 1 : -1;
    if (k < 0) {
        k = !s->gexec[0] ? 0 : !s->gexec[1] ? 1 : -1;
        if (k < 0) { drop_graphs(s); k = 0; }
        StepRunner r;
        r.s = s;
        find_ops(s, r.ops);
        r.a = step_args(s);
        r.graph = true;
        HCC_TRY_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        const hipError_t e = r.run();
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return set_err(ctx, HCC_E_HIP, std::string("hcc_step_cfl: capture failed: ") +
                                                hipGetErrorString(e != hipSuccess ? e : e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&s->gexec[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HCC_TRY_HIP(ctx, e3);
        s->gU[k] = s->U;
        ++s->g_captures;
    }
    if (s->ring_count == hcc_stepper::kRing)
        if (int rc = fold_oldest(s)) return rc;
    auto& slot = s->ring[s->ring_head];
    s->h_pub[2] = 0;   // the last graph launch has been settled: no kernel writes the flag now
    HCC_TRY_HIP(ctx, hipEventRecord(slot.ev0, ctx->stream));
    HCC_TRY_HIP(ctx, hipGraphLaunch(s->gexec[k], ctx->stream));
    k_stats_out<<<1, 64, 0, ctx->stream>>>(s->counters, s->dtp, slot.hd);
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipEventRecord(slot.ev1, ctx->stream));
    slot.host_iters = 0;
    slot.dt_from_device = true;
    slot.graph = true;
    slot.passes = false;
    s->ring_head = (s->ring_head + 1) % hcc_stepper::kRing;
    ++s->ring_count;
    std::swap(s->U, s->cand);   // pg.rhos <- rho_cand
    s->fresh = false;
    s->graph_pending = true;
    ++s->g_launches;
    return HCC_OK;
}

extern "C" int hcc_stepper_set_graph(hcc_stepper* s, int32_t on) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    if (!on) drop_graphs(s);
    s->graph_on = on != 0;
    return HCC_OK;
}

extern "C" int hcc_stepper_graph_stats(hcc_stepper* s, int64_t* launches, int64_t* captures, int64_t* reruns,
                                       int32_t* bounds) {
    if (!s) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    if (launches) *launches = s->g_launches;
    if (captures) *captures = s->g_captures;
    if (reruns) *reruns = s->g_reruns;
    if (bounds)
        for (int i = 0; i < s->desc.stages; ++i) bounds[i] = s->gbound[i + 2];
    return HCC_OK;
}

// ---- tuning -------------------------------------------------------------------------------------------------
// knobs: [0] staged full passes (1) or neighbour reads through the caches (0); [1] lanes per particle of the fixed
// point's kernels (kSplit or twice it; 0: kSplit); [2] particles one sweep of their grid covers (16384 or 65536);
// [3] the CFL estimate fused into the stage-1 gather (1) or its own pass (0); [4] the dense form of a fixed point's
// first iteration: when the last step's count at the checkpoint reached N / knobs[4] (> 0), never (0), always (-1)
static int split_base(const hcc_stepper* s) { return s->D * s->M <= 8 ? 8 : 16; }

extern "C" int hcc_stepper_set_tuning(hcc_stepper* s, const int32_t* knobs) {
    if (!s || !knobs) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    const int g0 = split_base(s);
    const int split = knobs[1] == 0 ? g0 : knobs[1];
    if ((knobs[0] != 0 && knobs[0] != 1) || (split != g0 && split != 2 * g0) ||
        (knobs[2] != 16384 && knobs[2] != 65536) || (knobs[3] != 0 && knobs[3] != 1) || knobs[4] < -1)
        return set_err(ctx, HCC_E_ARG, "hcc_stepper_set_tuning: knobs out of range");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    drop_graphs(s);   // captured steps hold the old launch shapes
    s->tune.staged = knobs[0];
    s->tune.split = split;
    s->tune.sweep = knobs[2];
    s->tune.fuse_cfl = knobs[3];
    s->tune.dense = knobs[4];
    return HCC_OK;
}

extern "C" int hcc_stepper_get_tuning(hcc_stepper* s, int32_t* knobs) {
    if (!s || !knobs) return HCC_E_ARG;
    knobs[0] = s->tune.staged;
    knobs[1] = s->tune.split ? s->tune.split : split_base(s);
    knobs[2] = s->tune.sweep;
    knobs[3] = s->tune.fuse_cfl;
    knobs[4] = s->tune.dense;
    return HCC_OK;
}

// Choices already tuned in this process, by workload shape: (D, M, PDE, flux variant, LDS tables present, log2 of
// the node count, mean neighbour count rounded, ensemble or not).
struct TuneKey {
    int D, M, kind, fv, lds, log2n, nnb, ens;
    bool operator<(const TuneKey& o) const {
        return std::tie(D, M, kind, fv, lds, log2n, nnb, ens) < std::tie(o.D, o.M, o.kind, o.fv, o.lds, o.log2n, o.nnb, o.ens);
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, hcc_stepper::Tuning> g_tuned;

static TuneKey tune_key(const hcc_stepper* s) {
    int l = 0;
    while ((int64_t(1) << (l + 1)) <= s->N) ++l;
    const int64_t ni = std::max<int64_t>(1, s->N);
    return TuneKey{s->D, s->M, s->kind, s->fv, s->lds_gather > 0, l, (int)std::lround((double)s->grid->S / (double)ni),
                   s->grid->n_inst > 1};
}

extern "C" int hcc_step_cfl(hcc_stepper* s, double cfl, double dt_cap);

// Times `steps` CFL steps (after one untimed step; wall clock) of each launch choice from the current state, restored before
// every trial and at the end, and keeps the fastest; a choice must beat the default by more than 2 % (the run-to-run
// spread of the step) to replace it. A shape tuned before in this process takes the stored choice without timing
// (*timed = 0). The statistics (hcc_step_totals, the last step's diagnostics) are those from before the call.
extern "C" int hcc_stepper_autotune(hcc_stepper* s, int32_t steps, double cfl, int32_t* knobs, double* ms_best,
                                    double* ms_default, int32_t* timed) {
    if (!s || steps < 1 || !(cfl > 0)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    const TuneKey key = tune_key(s);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            s->tune = it->second;
            drop_graphs(s);
            if (timed) *timed = 0;
            if (ms_best) *ms_best = 0;
            if (ms_default) *ms_default = 0;
            return knobs ? hcc_stepper_get_tuning(s, knobs) : HCC_OK;
        }
    }
    if (timed) *timed = 1;
    // the state and the statistics to restore
    const int64_t nrec = s->N * (int64_t)s->PU;
    double* save = nullptr;
    HCC_TRY_HIP(ctx, dalloc(&save, nrec));
    std::unique_ptr<double, hipError_t (*)(void*)> keep(save, hipFree);
    HCC_TRY_HIP(ctx, hipMemcpyAsync(save, s->U, sizeof(double) * nrec, hipMemcpyDeviceToDevice, ctx->stream));
    const bool fresh0 = s->fresh, graph0 = s->graph_on;
    const int64_t ts0 = s->tot_steps, ti0 = s->tot_iters;
    const double tm0 = s->tot_ms, tg0 = s->tot_gather_ms, tf0 = s->tot_flux_ms;
    const double lm0 = s->last_ms, lg0 = s->last_gather_ms, lf0 = s->last_flux_ms, ldt0 = s->last_dt;
    const int64_t in0 = s->inadmissible, mi0 = s->mood_iters;
    s->graph_on = false;
    auto restore = [&]() -> int {
        if (int rc = finish_pending(s)) return rc;
        HCC_TRY_HIP(ctx, hipMemcpyAsync(s->U, save, sizeof(double) * nrec, hipMemcpyDeviceToDevice, ctx->stream));
        HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        s->fresh = fresh0;
        s->cfl_chain = false;   // the trial steps' zero-K bits and minima do not describe the restored state
        return HCC_OK;
    };
    const int g0 = split_base(s);
    std::vector<hcc_stepper::Tuning> cands;
    const hcc_stepper::Tuning def{};
    for (int staged : {1, 0}) {
        if (staged && !s->lds_gather) continue;
        for (int split : {g0, 2 * g0})
            for (int sweep : {65536, 16384})
                for (int fuse : {1, 0}) {
                    if (!staged && fuse) continue;   // the fused CFL estimate needs the staged gather
                    hcc_stepper::Tuning t;
                    t.staged = staged; t.split = split; t.sweep = sweep; t.fuse_cfl = fuse;
                    cands.push_back(t);
                }
    }
    double best = INFINITY, tdef = INFINITY;
    hcc_stepper::Tuning bt = s->tune;
    for (const auto& t : cands) {
        s->tune = t;
        double ms = INFINITY;
        for (int rep = 0; rep < 2; ++rep) {   // the faster of two runs
            if (int rc = restore()) return rc;
            if (int rc = hcc_step_cfl(s, cfl, INFINITY)) return rc;   // untimed: first launches, stage 1 of a fresh state
            if (int rc = finish_pending(s)) return rc;
            // wall time of the queued steps: every pass of the step counts, also a CFL pass queued ahead of the step's
            // first event, and so do the host's turn-arounds inside the fixed points
            const auto t0 = std::chrono::steady_clock::now();
            for (int k = 0; k < steps; ++k)
                if (int rc = hcc_step_cfl(s, cfl, INFINITY)) return rc;
            if (int rc = finish_pending(s)) return rc;
            const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            ms = std::min(ms, wall / steps);
        }
        const bool is_def = t.staged == (s->lds_gather ? 1 : 0) && (t.split == g0) && t.sweep == def.sweep &&
                            t.fuse_cfl == (s->lds_gather ? 1 : 0);
        if (is_def) tdef = ms;
        if (ms < best) { best = ms; bt = t; }
    }
    // the default unless a choice is clearly faster
    hcc_stepper::Tuning chosen;
    chosen.staged = s->lds_gather ? 1 : 0;
    chosen.split = g0;
    chosen.fuse_cfl = chosen.staged;
    if (best < 0.98 * tdef) chosen = bt;
    if (int rc = restore()) return rc;
    s->tune = chosen;
    s->graph_on = graph0;
    drop_graphs(s);
    s->tot_steps = ts0; s->tot_iters = ti0; s->tot_ms = tm0; s->tot_gather_ms = tg0; s->tot_flux_ms = tf0;
    s->last_ms = lm0; s->last_gather_ms = lg0; s->last_flux_ms = lf0; s->last_dt = ldt0;
    s->inadmissible = in0; s->mood_iters = mi0;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[key] = chosen;
    }
    if (ms_best) *ms_best = best;
    if (ms_default) *ms_default = tdef;
    return knobs ? hcc_stepper_get_tuning(s, knobs) : HCC_OK;
}

extern "C" int hcc_step(hcc_stepper* s, double t, double dt) {
    (void)t;
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    if (!(dt > 0) || !std::isfinite(dt)) return set_err(ctx, HCC_E_ARG, "hcc_step: dt must be positive and finite");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (s->N == 0) return HCC_OK;
    if (int rc = graph_settle(s)) return rc;
    k_set_dt<<<1, 64, 0, ctx->stream>>>(s->dtp, dt, s->grid->n_inst);
    return launch_step(s, "hcc_step", false);
}

extern "C" int hcc_step_cfl(hcc_stepper* s, double cfl, double dt_cap) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    if (!(cfl > 0) || !std::isfinite(cfl) || !(dt_cap > 0))
        return set_err(ctx, HCC_E_ARG, "hcc_step_cfl: cfl must be positive and finite, dt_cap positive");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (s->N == 0) return HCC_OK;
    if (int rc = graph_settle(s)) return rc;
    // stage 1 reads U itself once the boundary conditions hold on it (every step after the first since the state was
    // loaded): the CFL estimate of U then runs inside the stage-1 gather, which stages the same neighbourhoods
    s->cfl_fused = !s->fresh && !s->imex && s->lds_gcfl > 0 && s->tune.staged && s->tune.fuse_cfl;
    if (s->graph_on && graph_eligible(s)) {
        s->cfl_num = cfl;
        s->cfl_cap = dt_cap;
        const int rc = graph_step(s);
        s->cfl_fused = false;
        return rc;
    }
    if (!s->cfl_fused) {
        VariantOps ops;
        find_ops(s, ops);
        ops.cfl(s, s->args(), s->dtp, cfl, dt_cap);
    }
    s->cfl_num = cfl;
    s->cfl_cap = dt_cap;
    const int rc = launch_step(s, "hcc_step_cfl", true);
    s->cfl_fused = false;
    return rc;
}

extern "C" int hcc_last_dt(hcc_stepper* s, double* dt) {
    if (!s || !dt) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    *dt = s->last_dt;
    return HCC_OK;
}

extern "C" int hcc_get_orders(hcc_stepper* s, int32_t* orders) {
    if (!s || (!orders && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int8_t> o((size_t)s->N);
    if (s->N) HCC_TRY_HIP(ctx, hipMemcpy(o.data(), s->orders, s->N, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < s->N; ++i) orders[i] = o[i];
    return HCC_OK;
}

extern "C" int hcc_get_mood_flags(hcc_stepper* s, uint8_t* flags) {
    if (!s || (!flags && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (s->N) HCC_TRY_HIP(ctx, hipMemcpy(flags, s->trig, s->N, hipMemcpyDeviceToHost));
    return HCC_OK;
}

extern "C" int hcc_get_inadmissible(hcc_stepper* s, int64_t* count) {
    if (!s || !count) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    *count = s->inadmissible;
    return HCC_OK;
}

extern "C" int hcc_get_mood_iterations(hcc_stepper* s, int64_t* count) {
    if (!s || !count) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    *count = s->mood_iters;
    return HCC_OK;
}

extern "C" int hcc_get_moments(hcc_stepper* s, double* inv_grad, uint8_t* ok_grad, double* inv_div, uint8_t* ok_div,
                               double* dweights) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t N = s->N;
    const int D = s->D;
    const int BG = s->order == 3 ? m3::bcount(D, 2) : D, BD = s->order == 3 ? m3::bcount(D, 3) + 1 : D + 1;
    if (inv_grad && N) {
        std::vector<double> v((size_t)(BG * BG * N));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->invg, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) for (int q = 0; q < BG * BG; ++q) inv_grad[i * BG * BG + q] = v[(size_t)(q * N + i)];
    }
    if (inv_div && N) {   // rows 1..D of the BD x BD inverse; row 0 and rows past D are not kept on the device (NaN)
        std::vector<double> v((size_t)(D * BD * N));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->invd, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) {
            for (int c = 0; c < BD * BD; ++c) inv_div[i * BD * BD + c] = NAN;
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < BD; ++c) inv_div[i * BD * BD + (1 + r) * BD + c] = v[(size_t)((r * BD + c) * N + i)];
        }
    }
    if (ok_grad && N) HCC_TRY_HIP(ctx, hipMemcpy(ok_grad, s->okg, N, hipMemcpyDeviceToHost));
    if (ok_div && N) HCC_TRY_HIP(ctx, hipMemcpy(ok_div, s->okd, N, hipMemcpyDeviceToHost));
    if (dweights && N) {
        std::vector<int64_t> rstart, sell;
        std::vector<int32_t> idx;
        int rc = download_lists(s->grid, rstart, idx, &sell);
        if (rc) return rc;
        const int64_t SS = s->grid->S_sell;
        // the derivative weights of the current geometry (the step's passes keep no copy): the moments of the
        // grid's positions and lists, as the last geometry refresh formed them, with the weights written too
        if (SS > s->dw_cap) {
            (void)hipFree(s->dw);
            s->dw = nullptr;
            s->dw_cap = 0;
            HCC_TRY_HIP(ctx, dalloc(&s->dw, (int64_t)D * SS));
            s->dw_cap = SS;
        }
        launch_moments(s, INFINITY, s->dw);
        HCC_TRY_HIP(ctx, hipGetLastError());
        HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<double> v((size_t)(D * SS));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->dw, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < sell.size(); ++k)
            for (int d = 0; d < D; ++d) dweights[k * D + d] = v[(size_t)(d * SS + sell[k])];
    }
    return HCC_OK;
}

extern "C" int hcc_rhs(hcc_stepper* s, const double* Y, double* K, double* gradients) {
    if (!s || (s->N > 0 && (!Y || !K))) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    const int64_t N = s->N;
    const int M = s->M, D = s->D;
    if (!N) return HCC_OK;
    s->cfl_chain = false;   // its passes rewrite cnb for Y
    VariantOps ops;
    find_ops(s, ops);
    hipStream_t st = ctx->stream;
    HCC_TRY_HIP(ctx, hipMemcpyAsync(s->scratch, Y, sizeof(double) * N * M, hipMemcpyHostToDevice, st));
    ops.pack(s, s->scratch, s->Y[0]);
    const StepArgs a = s->args();
    k_fill_f64<<<(unsigned)((s->GS * N + kBlock - 1) / kBlock), kBlock, 0, st>>>(s->G, 0.0, (int64_t)s->GS * N);
    k_fill_u8<<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, st>>>((uint8_t*)s->orders, (uint8_t)s->order, N);
    ops.gather(s, a, s->Y[0], 1, 0);
    ops.flux(s, a, PredArgs{}, s->Y[0], s->K[0], 1, false);
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipMemcpyAsync(K, s->K[0], sizeof(double) * N * M, hipMemcpyDeviceToHost, st));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (gradients) {
        std::vector<double> gs((size_t)(s->GS * N));
        HCC_TRY_HIP(ctx, hipMemcpy(gs.data(), s->G, sizeof(double) * gs.size(), hipMemcpyDeviceToHost));
        const int NG = (s->order == 3 ? m3::bcount(D, 2) : D) * M;   // basis terms x components, term major
        for (int64_t i = 0; i < N; ++i)
            for (int q = 0; q < NG; ++q) gradients[i * NG + q] = gs[(size_t)(i * s->GS + q)];
    }
    return HCC_OK;
}

extern "C" int hcc_last_step_timing(hcc_stepper* s, double* ms_total, double* ms_gather, double* ms_flux) {
    if (!s) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    if (ms_total) *ms_total = s->last_ms;
    if (ms_gather) *ms_gather = s->last_gather_ms;
    if (ms_flux) *ms_flux = s->last_flux_ms;
    return HCC_OK;
}

extern "C" int hcc_step_totals(hcc_stepper* s, int reset, int64_t* steps, int64_t* mood_iterations, double* ms_total,
                               double* ms_gather, double* ms_flux) {
    if (!s) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    if (steps) *steps = s->tot_steps;
    if (mood_iterations) *mood_iterations = s->tot_iters;
    if (ms_total) *ms_total = s->tot_ms;
    if (ms_gather) *ms_gather = s->tot_gather_ms;
    if (ms_flux) *ms_flux = s->tot_flux_ms;
    if (reset) { s->tot_steps = s->tot_iters = s->tot_timed = 0; s->tot_ms = s->tot_gather_ms = s->tot_flux_ms = 0; }
    return HCC_OK;
}

extern "C" int hcc_stepper_set_pass_timing(hcc_stepper* s, int32_t every) {
    if (!s || every < 0) return HCC_E_ARG;
    s->pass_every = every;
    s->pass_serial = 0;
    return HCC_OK;
}

extern "C" int hcc_step_timed_passes(hcc_stepper* s, int64_t* steps) {
    if (!s || !steps) return HCC_E_ARG;
    if (int rc = finish_pending(s)) return rc;
    *steps = s->tot_timed;
    return HCC_OK;
}

extern "C" int hcc_total_mass(hcc_stepper* s, int32_t interior_only, double* out) {
    if (!s || !out) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    const int M = s->M;
    if (s->N == 0) {
        for (int m = 0; m < M; ++m) out[m] = 0.0;
        return HCC_OK;
    }
    VariantOps ops;
    find_ops(s, ops);
    ops.mass(s, interior_only ? 1 : 0, s->scratch);   // after the queued steps, on the context stream
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipMemcpyAsync(out, s->scratch, sizeof(double) * M, hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

// The elements of range(0, T, length = n + 1) (test/time_integration.jl:92): Julia's TwicePrecision range gives the
// correctly rounded k T / n; with a 113-bit significand k T is exact and the quotient, rounded once more to 53 bits,
// is correctly rounded (k T / n lies on a 53-bit rounding boundary only if it is one exactly).
static double snapshot_time(int64_t k, int64_t n, double T) {
    if (k == 0) return 0.0;
    if (k == n) return T;
    const __float128 q = ((__float128)k * (__float128)T) / (__float128)n;
    return (double)q;
}

extern "C" int hcc_solve(hcc_stepper* s, double tmax, double dt, int32_t is_cfl, int32_t snapshots,
                         int32_t remove_ghosts, double* t_hist, double* x_hist, double* u_hist, double* mass_hist,
                         int64_t* steps, double* elapsed) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    if (snapshots < 1 || !std::isfinite(tmax) || !(dt > 0) || !std::isfinite(dt))
        return set_err(ctx, HCC_E_ARG, "hcc_solve: snapshots >= 1, tmax finite, dt positive and finite");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = graph_settle(s)) return rc;
    const int64_t N = s->N;
    const int D = s->D, M = s->M;
    const std::vector<uint8_t>& isb = s->grid->h_isb;
    std::vector<int64_t> keep;   // the particles a snapshot holds (remove_ghosts: interior only, in order)
    for (int64_t i = 0; i < N; ++i)
        if (!remove_ghosts || !isb[(size_t)i]) keep.push_back(i);
    const int64_t ns = (int64_t)keep.size();
    std::vector<double> pos((size_t)(N * D)), U((size_t)(N * M));
    if (x_hist && N) {
        if (int rc = hcc_grid_get_positions(s->grid, pos.data())) return rc;
    }
    for (int c = 0; t_hist && c <= snapshots; ++c) t_hist[c] = NAN;   // a snapshot the loop never reaches
    auto save = [&](int c, double t) -> int {   // saveData! (test/time_integration.jl:13-40)
        if (t_hist) t_hist[c] = t;
        if (u_hist && N) {
            if (int rc = hcc_state_download(s, U.data())) return rc;
            for (int64_t q = 0; q < ns; ++q)
                for (int m = 0; m < M; ++m) u_hist[((int64_t)c * ns + q) * M + m] = U[(size_t)(keep[(size_t)q] * M + m)];
        }
        if (x_hist && s->moving && N)   // the positions move with the grid
            if (int rc = hcc_grid_get_positions(s->grid, pos.data())) return rc;
        if (x_hist)
            for (int64_t q = 0; q < ns; ++q)
                for (int d = 0; d < D; ++d) x_hist[((int64_t)c * ns + q) * D + d] = pos[(size_t)(keep[(size_t)q] * D + d)];
        if (mass_hist) {
            if (int rc = hcc_total_mass(s, 0, mass_hist + (int64_t)c * 2 * M)) return rc;
            if (int rc = hcc_total_mass(s, 1, mass_hist + (int64_t)c * 2 * M + M)) return rc;
        }
        return HCC_OK;
    };
    std::vector<double> t_snap((size_t)snapshots + 1);
    for (int k = 0; k <= snapshots; ++k) t_snap[(size_t)k] = snapshot_time(k, snapshots, tmax);
    double t = 0.0;
    int64_t k_step = 0;
    int counter = 0;
    if (int rc = save(counter, t)) return rc;
    ++counter;
    const auto t0 = std::chrono::steady_clock::now();
    while (t < tmax) {
        double dtv = dt;
        if (is_cfl) {
            double c = 0.0;
            if (int rc = hcc_time_step(s, &c)) return rc;
            dtv = dt * c;
        }
        dtv = jmin(dtv, tmax - t);
        if (dtv <= 1e-12) break;
        if (int rc = hcc_step(s, t, dtv)) return rc;
        t += dtv;
        ++k_step;
        while (counter < snapshots && t >= t_snap[(size_t)counter]) {   // 1-based: snap_counter <= snapshots
            if (int rc = save(counter, t)) return rc;
            ++counter;
        }
    }
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (counter <= snapshots)
        if (int rc = save(snapshots, t)) return rc;
    if (int rc = finish_pending(s)) return rc;
    if (steps) *steps = k_step;
    if (elapsed) *elapsed = secs;
    return HCC_OK;
}

extern "C" int hcc_set_source(hcc_stepper* s, int32_t mode, const double* rates) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    if (mode < HCC_SRC_NONE || mode > HCC_SRC_DECAY_IMPLICIT) return set_err(ctx, HCC_E_ARG, "hcc_set_source: unknown mode");
    if (mode != HCC_SRC_NONE && !rates && s->N) return set_err(ctx, HCC_E_ARG, "hcc_set_source: rates required");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    drop_graphs(s);   // captured steps hold the old configuration
    const int64_t N = s->N, M = s->M;
    hcc_stepper_desc& d = s->desc;
    if (mode == HCC_SRC_DECAY_IMPLICIT) {
        bool zero = true;   // _explicit_implicit_part: an explicit tableau applies the implicit source explicitly
        for (int i = 0; i < d.stages; ++i) {
            zero = zero && d.b_im[i] == 0.0;
            for (int j = 0; j < d.stages; ++j) zero = zero && d.a_im[i][j] == 0.0;
        }
        if (zero)
            for (int i = 0; i < d.stages; ++i) {
                d.b_im[i] = d.b_t[i];
                d.c_im[i] = d.c_t[i];
                for (int j = 0; j < d.stages; ++j) d.a_im[i][j] = d.a_t[i][j];
            }
        if (!s->imex) {
            for (int k = 0; k < d.stages; ++k) {
                HCC_TRY_HIP(ctx, dalloc(&s->KI[k], M * N));
                HCC_TRY_HIP(ctx, hipMemsetAsync(s->KI[k], 0, sizeof(double) * std::max<int64_t>(M * N, 1), ctx->stream));
            }
            s->imex = true;
        }
    }
    if (mode != HCC_SRC_NONE && N) {
        if (!s->rate) HCC_TRY_HIP(ctx, dalloc(&s->rate, N));
        HCC_TRY_HIP(ctx, hipMemcpyAsync(s->rate, rates, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    }
    s->src_mode = mode;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

// ============================================================================================================
// learned troubled-node indicator: features, decision before the flux pass, training data
// ============================================================================================================
static int feature_buffers(hcc_stepper* s) {
    hcc_ctx* ctx = s->ctx;
    if (s->D != 2 || s->M != 4 || s->kind != HCC_PDE_EULER || s->order != 2)
        return set_err(ctx, HCC_E_UNSUPPORTED, "the indicator features are defined for 2-D Euler with MUSCL2");
    if (!s->feat) {
        HCC_TRY_HIP(ctx, dalloc(&s->feat, 16 * s->N));
        HCC_TRY_HIP(ctx, hipMemset(s->feat, 0, sizeof(uint16_t) * std::max<int64_t>(16 * s->N, 1)));
        HCC_TRY_HIP(ctx, dalloc(&s->ind_list, s->N));
        HCC_TRY_HIP(ctx, dalloc(&s->ind_n, 1));
        HCC_TRY_HIP(ctx, hipMemset(s->ind_n, 0, sizeof(unsigned long long)));
    }
    return HCC_OK;
}

extern "C" int hcc_set_indicator(hcc_stepper* s, const hcc_mlp* m, float threshold) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    drop_graphs(s);   // captured steps hold the old configuration
    if (m) {
        if (mlp_context(m) != ctx)
            return set_err(ctx, HCC_E_ARG, "hcc_set_indicator: the network belongs to another context (another stream)");
        if (int rc = feature_buffers(s)) return rc;
        if (!s->ind_z) HCC_TRY_HIP(ctx, dalloc(&s->ind_z, s->N));
        if (!s->ind_cnt) {
            HCC_TRY_HIP(ctx, dalloc(&s->ind_cnt, 1));
            HCC_TRY_HIP(ctx, hipMemset(s->ind_cnt, 0, sizeof(unsigned long long)));
        }
    }
    s->ind = m;
    s->ind_thr = threshold;
    s->feat_on = s->ind || s->collect;
    return HCC_OK;
}

extern "C" int hcc_indicator_collect(hcc_stepper* s, int32_t on) {
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    drop_graphs(s);   // captured steps hold the old configuration
    if (on) {
        if (int rc = feature_buffers(s)) return rc;
        const int st = s->desc.stages;
        if (!s->cfeat) {
            HCC_TRY_HIP(ctx, dalloc(&s->cfeat, (int64_t)st * 16 * s->N));
            HCC_TRY_HIP(ctx, dalloc(&s->cord, (int64_t)2 * st * s->N));
            HCC_TRY_HIP(ctx, hipMemset(s->cord, 0, std::max<int64_t>((int64_t)2 * st * s->N, 1)));
        }
    }
    s->collect = on != 0;
    s->feat_on = s->ind || s->collect;
    return HCC_OK;
}

extern "C" int hcc_get_indicator_samples(hcc_stepper* s, int32_t stage, uint16_t* feats, int8_t* pre, int8_t* post) {
    if (!s || !s->cfeat || stage < 0 || stage >= s->desc.stages)
        return set_err(s ? s->ctx : nullptr, HCC_E_ARG, "hcc_get_indicator_samples: no collection or bad stage");
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = finish_pending(s)) return rc;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t N = s->N;
    const int st = s->desc.stages;
    if (feats) HCC_TRY_HIP(ctx, hipMemcpy(feats, s->cfeat + (int64_t)stage * 16 * N, sizeof(uint16_t) * 16 * N, hipMemcpyDeviceToHost));
    if (pre) HCC_TRY_HIP(ctx, hipMemcpy(pre, s->cord + (int64_t)stage * N, N, hipMemcpyDeviceToHost));
    if (post) HCC_TRY_HIP(ctx, hipMemcpy(post, s->cord + (int64_t)(st + stage) * N, N, hipMemcpyDeviceToHost));
    return HCC_OK;
}

extern "C" int hcc_get_indicator_count(hcc_stepper* s, int64_t* count) {
    if (!s || !count) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    *count = 0;
    if (!s->ind_cnt) return HCC_OK;
    if (int rc = finish_pending(s)) return rc;
    unsigned long long c = 0;
    HCC_TRY_HIP(ctx, hipMemcpy(&c, s->ind_cnt, sizeof(c), hipMemcpyDeviceToHost));
    *count = (int64_t)c;
    return HCC_OK;
}
