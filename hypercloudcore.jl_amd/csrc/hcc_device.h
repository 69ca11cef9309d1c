// Device-side building blocks shared by the kernels of libhcc: launch geometry, the per-slot geometry of the
// compute kernels, the PDE variants, error helpers.
#pragma once
#include "hcc_internal.h"
#include <string>

namespace hcc {

// ============================================================================================================
// errors
// ============================================================================================================
extern thread_local std::string g_create_err;

#define HCC_TRY_HIP(ctx, expr)                                                                 \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return HCC_E_HIP;                                                                  \
        }                                                                                      \
    } while (0)

inline int set_err(hcc_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg; else g_create_err = msg;
    return code;
}

template <class T>
inline hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

// ============================================================================================================
// launch geometry: 256-thread workgroups, XCD-aware block order
// ============================================================================================================
constexpr int kBlock = 256;

// The dispatcher hands workgroup b to XCD b % 8, as the k-th (k = b / 8) of that XCD. Remap so that each XCD runs
// runs of kXcdChunk consecutive logical blocks (Morton-ordered neighbours then share that XCD's L2), the runs dealt
// to the XCDs in turn: work that is concentrated in one part of the cloud (a shock, the non-constant region of a
// Riemann problem) spreads over all eight XCDs instead of falling to the one that owns that eighth of the range.
// A bijection of [0, 8 per) onto itself: whole chunks first, then each XCD's remaining k contiguously.
constexpr int64_t kXcdChunk = 16;
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblk) {
    const int64_t per = (nblk + 7) / 8, x = b & 7, k = b >> 3;
    const int64_t full = per / kXcdChunk * kXcdChunk;   // blocks per XCD in whole chunks
    if (k < full) return ((k / kXcdChunk) * 8 + x) * kXcdChunk + k % kXcdChunk;
    return full * 8 + x * (per - full) + (k - full);
}
__device__ __forceinline__ int64_t node_index(int64_t nblk) {
    return xcd_block(blockIdx.x, nblk) * kBlock + threadIdx.x;
}
inline int64_t nblocks(int64_t N) { return (N + kBlock - 1) / kBlock; }
inline unsigned grid_for(int64_t nblk) { return (unsigned)(((nblk + 7) / 8) * 8); }

// ============================================================================================================
// per-slot geometry inside the compute kernels: distance and weight from the positions, exactly as the search
// ============================================================================================================
template <int D>
struct NodePos {
    double x[D];
};
template <int D>
__device__ __forceinline__ NodePos<D> load_pos(const GeomArgs& g, int64_t i) {
    NodePos<D> p;
    p.x[0] = g.px[i];
    if (D > 1) p.x[1] = g.py[i];
    if (D > 2) p.x[2] = g.pz[i];
    return p;
}
template <int D>
__device__ __forceinline__ double slot_geom(const GeomArgs& g, const NodePos<D>& pi, int32_t j, double (&dd)[D]) {
    NodePos<D> pj = load_pos<D>(g, j);
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        double x = pj.x[d] - pi.x[d];
        x = x - g.Lw[d] * rint(x * g.invLw[d]);
        dd[d] = x;
        s = (d == 0) ? x * x : s + x * x;
    }
    return s;
}
__device__ __forceinline__ int64_t sell_base(const GeomArgs& g, int64_t i) { return g.chunk_off[i >> 6] + (i & 63); }
// Running extrema with Julia's max / min (NaN-propagating, -0 < +0) at three VALU operations per element instead of
// the eight of jmax + jmin: the hardware v_max_f64 / v_min_f64 order signed zeros and, given one NaN operand, return
// the other, so a NaN is tracked in a lane mask (x != x) and the extrema are replaced at the end by the canonical NaN
// that jmax / jmin produce (0x7FF8000000000000). Arithmetic only: the operands are registers.
__device__ __forceinline__ double hw_max(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double hw_min(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// the weight of a slot at squared distance d2: wf(d2), or 0 beyond R_cut (weights[k] = d2 <= R_sq ? wf(d2) : 0)
__device__ __forceinline__ double slot_weight(const GeomArgs& g, double d2) {
    const double w = mls_weight(g.alpha, g.inv_range_sq, d2);
    return d2 <= g.R_cut ? w : 0.0;
}

// ============================================================================================================
// PDE kernels variants: KIND 1 = linear advection (diagonal), KIND 2 = Euler (M = D + 2)
// ============================================================================================================
template <int D, int M, int KIND>
struct Pde {
    __device__ static __forceinline__ void flux(const PdeArgs& e, const double (&U)[M], double (&F)[D][M]) {
        if constexpr (KIND == HCC_PDE_EULER) {
            EulerPrim<D> q = euler_prim<D>(e.gamma, U);
            for (int d = 0; d < D; ++d) {
                F[d][0] = U[1 + d];
                for (int k = 0; k < D; ++k) F[d][1 + k] = (k == d) ? U[1 + d] * q.u[k] + q.p : U[1 + d] * q.u[k];
                F[d][D + 1] = (U[D + 1] + q.p) * q.u[d];
            }
        } else {
            for (int d = 0; d < D; ++d) for (int m = 0; m < M; ++m) F[d][m] = e.c[d][m] * U[m];
        }
    }
    __device__ static __forceinline__ void maxeig(const PdeArgs& e, const double (&U)[M], double (&lam)[D]) {
        if constexpr (KIND == HCC_PDE_EULER) {
            EulerPrim<D> q = euler_prim<D>(e.gamma, U);
            double cs = sqrt((e.gamma * q.p) * q.inv);
            for (int d = 0; d < D; ++d) lam[d] = fabs(q.u[d]) + cs;
        } else {
            for (int d = 0; d < D; ++d) lam[d] = e.amax[d];
        }
    }
    // flux and eigenvalues of one state sharing the primitive variables
    __device__ static __forceinline__ void flux_eig(const PdeArgs& e, const double (&U)[M], double (&F)[D][M], double (&lam)[D]) {
        if constexpr (KIND == HCC_PDE_EULER) {
            EulerPrim<D> q = euler_prim<D>(e.gamma, U);
            for (int d = 0; d < D; ++d) {
                F[d][0] = U[1 + d];
                for (int k = 0; k < D; ++k) F[d][1 + k] = (k == d) ? U[1 + d] * q.u[k] + q.p : U[1 + d] * q.u[k];
                F[d][D + 1] = (U[D + 1] + q.p) * q.u[d];
            }
            double cs = sqrt((e.gamma * q.p) * q.inv);
            for (int d = 0; d < D; ++d) lam[d] = fabs(q.u[d]) + cs;
        } else {
            flux(e, U, F);
            maxeig(e, U, lam);
        }
    }
    // grid_velocity (GridMovers.jl:52-63): the fluid velocity (API.jl:104-109: the advection speed, or Euler's m / rho
    // with the operations of the flux), scaled by alpha for ScaledLagrangianMover
    __device__ static __forceinline__ void grid_velocity(const PdeArgs& e, const double (&U)[M], int mover, double alpha,
                                                         double (&w)[D]) {
        if constexpr (KIND == HCC_PDE_EULER) {
            const double inv = 1.0 / U[0];
            for (int d = 0; d < D; ++d) w[d] = U[1 + d] * inv;
        } else {
            for (int d = 0; d < D; ++d) w[d] = e.c[d][0];
        }
        if (mover == HCC_MOVER_SCALED_LAGRANGIAN)
            for (int d = 0; d < D; ++d) w[d] = alpha * w[d];
    }
    // flux and wave speeds of one state in the frame moving with w (ALEFrame, GridMovers.jl:65-90): F - w U and
    // |u_d - w_d| + c for Euler, |c_d - w_d| for scalar advection
    __device__ static __forceinline__ void flux_eig_ale(const PdeArgs& e, const double (&U)[M], const double (&w)[D],
                                                        double (&F)[D][M], double (&lam)[D]) {
        flux(e, U, F);
        for (int d = 0; d < D; ++d)
            for (int m = 0; m < M; ++m) F[d][m] = F[d][m] - w[d] * U[m];
        if constexpr (KIND == HCC_PDE_EULER) {
            EulerPrim<D> q = euler_prim<D>(e.gamma, U);
            double cs = sqrt((e.gamma * q.p) * q.inv);
            for (int d = 0; d < D; ++d) lam[d] = fabs(q.u[d] - w[d]) + cs;
        } else {
            for (int d = 0; d < D; ++d) lam[d] = fabs(e.c[d][0] - w[d]);
        }
    }
    // is_admissible: finite, and for Euler rho > 0 and p > 0
    __device__ static __forceinline__ bool admissible(const PdeArgs& e, const double (&U)[M]) {
        for (int m = 0; m < M; ++m) if (!isfinite(U[m])) return false;
        if constexpr (KIND == HCC_PDE_EULER) {
            EulerPrim<D> q = euler_prim<D>(e.gamma, U);
            return U[0] > 0.0 && q.p > 0.0;
        }
        return true;
    }
};

// ============================================================================================================
// geometry caches: prepare_geometry! (src/Interpolation/MUSCL.jl:172-223, Interpolators.jl:517-577)
// ============================================================================================================
struct MomentArgs {
    GeomArgs g;
    double* invg; uint8_t* okg;
    double* invd; uint8_t* okd;
    double* dw; int64_t S_sell;
};

template <int D>
__global__ __launch_bounds__(kBlock) void k_moments(MomentArgs a, int64_t nblk) {
    constexpr int BG = D, BD = D + 1;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    if (i >= N) return;
    const NodePos<D> pi = load_pos<D>(a.g, i);
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    const double invL = a.g.invL;
    double AG[BG][BG], AD[BD][BD];
    for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) AG[r][c] = 0.0;
    for (int r = 0; r < BD; ++r) for (int c = 0; c < BD; ++c) AD[r][c] = 0.0;
    for (int k = 0; k < cnt; ++k) {
        int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
        double dd[D];
        double d2 = slot_geom<D>(a.g, pi, j, dd);
        double w = slot_weight(a.g, d2);
        double p[BD];
        p[0] = 1.0;
        for (int d = 0; d < D; ++d) p[1 + d] = dd[d] * invL;
        for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) AG[r][c] = AG[r][c] + w * (p[1 + r] * p[1 + c]);
        for (int r = 0; r < BD; ++r) for (int c = 0; c < BD; ++c) AD[r][c] = AD[r][c] + w * (p[r] * p[c]);
    }
    {
        double U[BG][BG], X[BG][BG];
        bool ok = chol_moment<BG>(AG, U);
        if (ok) chol_inverse<BG>(U, X);
        for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) a.invg[(r * BG + c) * N + i] = ok ? X[r][c] : 0.0;
        a.okg[i] = ok;
    }
    double U[BD][BD], X[BD][BD];
    bool ok = chol_moment<BD>(AD, U);
    if (ok) chol_inverse<BD>(U, X);
    for (int r = 0; r < D; ++r) for (int c = 0; c < BD; ++c) a.invd[(r * BD + c) * N + i] = ok ? X[1 + r][c] : 0.0;
    a.okd[i] = ok;
    if (!a.dw) return;
    // dweights: (rows 2..D+1 of inv(N)) * (w p~) * invL (MUSCL.jl:211-223)
    for (int k = 0; k < cnt; ++k) {
        const int64_t s = base + (int64_t)k * kChunk;
        if (!ok) { for (int d = 0; d < D; ++d) a.dw[d * a.S_sell + s] = 0.0; continue; }
        int32_t j = a.g.nbr[s];
        double dd[D];
        double d2 = slot_geom<D>(a.g, pi, j, dd);
        double w = slot_weight(a.g, d2);
        double v[BD];
        v[0] = w * 1.0;
        for (int d = 0; d < D; ++d) v[1 + d] = w * (dd[d] * invL);
        for (int r = 0; r < D; ++r) {
            double acc = X[1 + r][0] * v[0];
            for (int c = 1; c < BD; ++c) acc = acc + X[1 + r][c] * v[c];
            a.dw[r * a.S_sell + s] = acc * invL;
        }
    }
}


}  // namespace hcc
