// MUSCL order 3 on the device (MUSCL(T, D, M, 3, limiter, RusanovFlux()), src/Interpolation/MUSCL.jl): the degree-2
// reconstruction with the graded Taylor basis, the divergence fit of degree 3 with intercept where the stencil carries
// it (the neighbour-count safeguard of _divergence_index lowers it to 2 or 1, zero-padded in the cache), the direct
// Cholesky solves of the particles below the full order, :soft / :hard limiting of the degree-2 terms, and three MOOD
// levels. Included by hcc_step.hip after the order-2 passes, whose record access, predictor, MOOD check and list
// bookkeeping it shares.
//
// Layout: gradient records hold the BG = basis_count(D, 2) coefficients per component, term major ([b * M + m],
// padded to an even stride). The reconstruction inverse is BG x BG per node, the divergence inverse holds its D
// derivative rows (rows 2 .. D + 1 of the padded BD x BD inverse, BD = basis_count(D, 3) + 1), both SoA
// ([q * N + i]). Neighbours are read through the caches (no LDS staging).
#pragma once

namespace m3 {

// basis_count(D, q) = binomial(D + q, q) - 1 (Interpolators.jl:170-177)
__host__ __device__ constexpr int bcount(int D, int q) {
    int n = 1;
    for (int k = 1; k <= q; ++k) n = (n * (D + k)) / k;
    return q <= 0 ? 0 : n - 1;
}

// Exponent of axis `ax` in term b of the graded basis of degree <= 3 (_generate_exponents, Interpolators.jl:138-155):
// each degree's tuples sorted by (-max(t), reverse(t)); 2-D: x, y | x^2, y^2, xy | x^3, y^3, x^2 y, x y^2.
__host__ __device__ constexpr int texp(int D, int b, int ax) {
    if (D == 1) return ax == 0 ? b + 1 : 0;
    if (D == 2) {
        const int t[9][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}, {1, 1}, {3, 0}, {0, 3}, {2, 1}, {1, 2}};
        return ax < 2 ? t[b][ax] : 0;
    }
    const int t[19][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0},
                          {1, 0, 1}, {0, 1, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0},
                          {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {1, 1, 1}};
    return t[b][ax];
}
__host__ __device__ constexpr int tfact(int e) { return e <= 1 ? 1 : e == 2 ? 2 : 6; }
// prod(t!) of term b, and its degree
__host__ __device__ constexpr int tden(int D, int b) {
    int d = 1;
    for (int ax = 0; ax < D; ++ax) d *= tfact(texp(D, b, ax));
    return d;
}
__host__ __device__ constexpr int tdeg(int D, int b) {
    int s = 0;
    for (int ax = 0; ax < D; ++ax) s += texp(D, b, ax);
    return s;
}

// build_basis (Interpolators.jl:194-218) of the first NB terms: 1/prod(t!) (when not 1) times the powers axis by axis,
// d^2 = d*d and d^3 = (d*d)*d, multiplied left to right
template <int D, int NB>
__device__ __forceinline__ void taylor(const double (&d)[D], double (&p)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int den = tden(D, b);
        bool have = den != 1;
        double t = have ? 1.0 / (double)den : 0.0;
#pragma unroll
        for (int ax = 0; ax < D; ++ax) {
            const int e = texp(D, b, ax);
            if (e == 0) continue;
            const double pw = e == 1 ? d[ax] : e == 2 ? d[ax] * d[ax] : (d[ax] * d[ax]) * d[ax];
            t = have ? t * pw : pw;
            have = true;
        }
        p[b] = t;
    }
}
// the scale of term b's coefficient: invL^degree (build_scale_factors, Interpolators.jl:236-247)
__device__ __forceinline__ double tscale(int deg, double invL) {
    return deg == 1 ? invL : deg == 2 ? invL * invL : (invL * invL) * invL;
}

template <int D, int M>
struct L3 {
    static constexpr int NR = D + M, PU = Layout<D, M>::PU;
    static constexpr int BG = bcount(D, 2);          // reconstruction terms
    static constexpr int NG = BG * M;
    static constexpr int GS = (NG + 1) & ~1;         // gradient record stride (16-byte loads)
    static constexpr int BD = bcount(D, 3) + 1;      // divergence fit with intercept at full degree
    static constexpr int B1 = D + 1;                 // the degree-1 fit with intercept
};

// _divergence_index(Val(0), Val(D), p_order, num_nb) (MUSCL.jl:11-32): the table entry (degree + 1)
__device__ __forceinline__ int div_index(int D, int p, int nb) {
    int idx = p <= 2 ? 2 : (((p & 1) && (D < 3 || p <= 3)) ? p + 1 : p);
    const int base = p > 2 ? p : 2;
    if (idx > base && 2 * nb < 3 * (bcount(D, idx - 1) + 1)) idx = base;
    while (idx > 2 && nb < bcount(D, idx - 1) + 1) --idx;
    return idx;
}

// C \ R for NR columns: U \ (U' \ R), forward then back substitution (StaticArrays' triangular solves)
template <int B, int NRH>
__device__ __forceinline__ void chol_solve(const double (&U)[B][B], const double (&R)[B][NRH], double (&X)[B][NRH]) {
#pragma unroll
    for (int j = 0; j < NRH; ++j) {
        double y[B];
#pragma unroll
        for (int i = 0; i < B; ++i) {
            double s = R[i][j];
#pragma unroll
            for (int k = 0; k < i; ++k) s = s - U[k][i] * y[k];
            y[i] = s / U[i][i];
        }
#pragma unroll
        for (int i = B - 1; i >= 0; --i) {
            double s = y[i];
#pragma unroll
            for (int k = i + 1; k < B; ++k) s = s - U[i][k] * X[k][j];
            X[i][j] = s / U[i][i];
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// neighbour sources: records (and gradient records, orders) through the caches, or the workgroup's halo staged in
// LDS by stage_halo3 (the order-2 layout: [position, state | pad | gradient record | pad], orders as bytes after the
// records, then the block's local-index table)
// ------------------------------------------------------------------------------------------------------------
template <int D, int M>
struct HbmSrc3 {
    using L = L3<D, M>;
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
template <int D, int M, bool GRADS>
struct LdsSrc3 {
    using L = L3<D, M>;
    static constexpr int GO = lds_grad_offset(L::NR);
    static constexpr int ST = GRADS ? lds_stride(L::NR, L::NG) : lds_stride(L::NR, 0);
    const double* sh;
    const int8_t* ord;
    const uint16_t* nl;
    int64_t s0;
    __device__ __forceinline__ int32_t idx(int64_t s) const { return nl[s - s0]; }
    template <int K>
    __device__ __forceinline__ void head(int32_t l, double (&r)[K]) const { lds_read<K>(sh + l * ST, r); }
    __device__ __forceinline__ void grad(int32_t l, double (&g)[L::NG]) const { lds_read<L::NG>(sh + l * ST + GO, g); }
    __device__ __forceinline__ int order(int32_t l) const { return ord[l]; }
};
template <int D, int M>
__device__ __forceinline__ HbmSrc3<D, M> hbm_src(const StepArgs& a, const double* Y) {
    return HbmSrc3<D, M>{Y, a.G, a.orders, a.g.nbr};
}

// Bytes of LDS the staged passes need for halos of up to hm nodes and index tables of up to idx_bytes
template <int D, int M>
__host__ __device__ constexpr size_t lds_bytes3(int64_t hm, size_t idx_bytes, bool grads) {
    using L = L3<D, M>;
    return 2 * lds_index_offset(hm, grads ? lds_stride(L::NR, L::NG) : lds_stride(L::NR, 0), grads) + idx_bytes;
}

// Copies the halo of block lb (records; with GRADS gradient records and orders) and its index table into LDS
template <int D, int M, bool GRADS>
__device__ __forceinline__ LdsSrc3<D, M, GRADS> stage_halo3(const StepArgs& a, const double* __restrict__ Y,
                                                            double* sh, int64_t lb) {
    using L = L3<D, M>;
    using S = LdsSrc3<D, M, GRADS>;
    const int32_t h0 = a.g.halo_off[lb], nh = a.g.halo_off[lb + 1] - h0;
    int8_t* ord = reinterpret_cast<int8_t*>(sh + (size_t)a.g.hmax * S::ST);
    uint16_t* nl = reinterpret_cast<uint16_t*>(sh) + lds_index_offset(a.g.hmax, S::ST, GRADS);
    const int64_t s0 = stage_index_table(a.g, nl, lb);
    for (int t = threadIdx.x; t < nh; t += kBlock) {
        const int32_t j = a.g.halo_idx[h0 + t];
        double r[L::NR];
        load_rec<L::NR, L::PU>(Y, j, r);
        double* dst = sh + t * S::ST;
        if constexpr (GRADS) {
            double g[L::NG];
            load_rec<L::NG, L::GS>(a.G, j, g);
            ord[t] = a.orders[j];
#pragma unroll
            for (int q = 0; q < L::NG; ++q) dst[S::GO + q] = g[q];
        }
#pragma unroll
        for (int q = 0; q < L::NR; ++q) dst[q] = r[q];
    }
    __syncthreads();
    return S{sh, ord, nl, s0};
}

// ------------------------------------------------------------------------------------------------------------
// geometry caches (prepare_geometry!, MUSCL.jl:172-223): the degree-2 reconstruction inverse, the divergence inverse
// of the degree each stencil carries at full order (_FullOrderDivDegree) padded to BD x BD, its derivative rows, and
// (dw set) the per-slot derivative weights rows(inv) (w p~) invL with the full-degree intercept basis
// ------------------------------------------------------------------------------------------------------------
template <int D, int B>
__device__ __forceinline__ bool div_inverse(const MomentArgs& a, const NodePos<D>& pi, int cnt, int64_t base, int q,
                                            double (&X)[bcount(D, 3) + 1][bcount(D, 3) + 1]) {
    constexpr int BD = bcount(D, 3) + 1;
    const double invL = a.g.invL;
    double A[B][B];
    for (int r = 0; r < B; ++r) for (int c = 0; c < B; ++c) A[r][c] = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
        double dd[D], ds[D];
        const double d2 = slot_geom<D>(a.g, pi, j, dd);
        const double w = slot_weight(a.g, d2);
        for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
        double p[B];
        p[0] = 1.0;
        double t[B - 1];
        taylor<D, B - 1>(ds, t);
        for (int c = 1; c < B; ++c) p[c] = t[c - 1];
        for (int r = 0; r < B; ++r) for (int c = 0; c < B; ++c) A[r][c] = A[r][c] + w * (p[r] * p[c]);
    }
    (void)q;
    double U[B][B], Xi[B][B];
    if (!chol_moment<B>(A, U)) return false;
    chol_inverse<B>(U, Xi);
    for (int r = 0; r < BD; ++r) for (int c = 0; c < BD; ++c) X[r][c] = (r < B && c < B) ? Xi[r][c] : 0.0;
    return true;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_moments3(MomentArgs a, int64_t nblk) {
    constexpr int BG = bcount(D, 2), BD = bcount(D, 3) + 1;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    if (i >= N) return;
    const NodePos<D> pi = load_pos<D>(a.g, i);
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    const double invL = a.g.invL;
    {
        double AG[BG][BG];
        for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) AG[r][c] = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double dd[D], ds[D], p[BG];
            const double d2 = slot_geom<D>(a.g, pi, j, dd);
            const double w = slot_weight(a.g, d2);
            for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
            taylor<D, BG>(ds, p);
            for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) AG[r][c] = AG[r][c] + w * (p[r] * p[c]);
        }
        double U[BG][BG], X[BG][BG];
        const bool ok = chol_moment<BG>(AG, U);
        if (ok) chol_inverse<BG>(U, X);
        for (int r = 0; r < BG; ++r) for (int c = 0; c < BG; ++c) a.invg[(r * BG + c) * N + i] = ok ? X[r][c] : 0.0;
        a.okg[i] = ok;
    }
    const int q = div_index(D, 3, cnt) - 1;
    double X[BD][BD];
    bool ok;
    if (q == 3) ok = div_inverse<D, BD>(a, pi, cnt, base, q, X);
    else if (q == 2) ok = div_inverse<D, bcount(D, 2) + 1>(a, pi, cnt, base, q, X);
    else ok = div_inverse<D, D + 1>(a, pi, cnt, base, q, X);
    for (int r = 0; r < D; ++r) for (int c = 0; c < BD; ++c) a.invd[(r * BD + c) * N + i] = ok ? X[1 + r][c] : 0.0;
    a.okd[i] = ok;
    if (!a.dw) return;
    for (int k = 0; k < cnt; ++k) {
        const int64_t s = base + (int64_t)k * kChunk;
        if (!ok) { for (int d = 0; d < D; ++d) a.dw[d * a.S_sell + s] = 0.0; continue; }
        const int32_t j = a.g.nbr[s];
        double dd[D], ds[D], t[BD - 1], v[BD];
        const double d2 = slot_geom<D>(a.g, pi, j, dd);
        const double w = slot_weight(a.g, d2);
        for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
        taylor<D, BD - 1>(ds, t);
        v[0] = w * 1.0;
        for (int c = 1; c < BD; ++c) v[c] = w * t[c - 1];
        for (int r = 0; r < D; ++r) {
            double acc = X[1 + r][0] * v[0];
            for (int c = 1; c < BD; ++c) acc = acc + X[1 + r][c] * v[c];
            a.dw[r * a.S_sell + s] = acc * invL;
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// gather: update_content! (MUSCL.jl:351-384) with the starved-order rule, the cached degree-2 fit at order 3
// (apply_weights, Interpolators.jl:585-602), the direct degree-1 solve at order 2 (_mls_solve, :372-400), and
// _limit_slopes (Limiter.jl:120-143) with the degree-2 interface basis; EXT stores the neighbourhood extrema for
// MOODu1's checks
// ------------------------------------------------------------------------------------------------------------
template <int D, int M, int KIND, bool EXT, class Src>
__device__ __forceinline__ void gather_node(const StepArgs& a, const double* __restrict__ Y, const Src& src, int64_t i) {
    using L = L3<D, M>;
    constexpr int NR = L::NR, BG = L::BG;
    const int64_t N = a.g.N;
    const int cnt = a.g.cnt[i];
    int order = a.orders[i];
    const int o0 = order;
    while (order > 1 && cnt < bcount(D, order - 1)) --order;   // starved neighbourhood
    if (order < o0) a.orders[i] = (int8_t)order;
    double g[L::NG];
#pragma unroll
    for (int q = 0; q < L::NG; ++q) g[q] = 0.0;
    const bool cached = order == 3 && a.okg[i];
    const bool direct = order == 2;
    const bool need = cached || direct || (EXT && a.ext_on) || (order > 1 && a.limiter != HCC_LIM_NONE);
    if (!need) { store_rec<L::NG, L::GS>(a.G, i, g); return; }
    const int64_t base = sell_base(a.g, i);
    const double invL = a.g.invL;
    double ri[NR], umax[M], umin[M];
    load_rec<NR, L::PU>(Y, i, ri);
    bool nanf[M];
#pragma unroll
    for (int m = 0; m < M; ++m) { umax[m] = ri[D + m]; umin[m] = ri[D + m]; nanf[m] = ri[D + m] != ri[D + m]; }
    double b[BG][M], A[D][D];
#pragma unroll
    for (int q = 0; q < BG; ++q)
#pragma unroll
        for (int m = 0; m < M; ++m) b[q][m] = 0.0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) A[r][c] = 0.0;
    for (int k = 0; k < cnt; ++k) {
        double rj[NR];
        src.template head<NR>(src.idx(base + (int64_t)k * kChunk), rj);
        if (cached || direct) {
            double dd[D], ds[D];
            const double d2 = rec_geom<D>(a.g, ri, rj, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
#pragma unroll
            for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
            if (cached) {
                double p[BG];
                taylor<D, BG>(ds, p);
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const double df = rj[D + m] - ri[D + m];
#pragma unroll
                    for (int q = 0; q < BG; ++q) b[q][m] = b[q][m] + w * (p[q] * df);
                }
            } else {   // the degree-1 basis is the distance itself
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int c = 0; c < D; ++c) A[r][c] = A[r][c] + w * (ds[r] * ds[c]);
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const double df = rj[D + m] - ri[D + m];
#pragma unroll
                    for (int q = 0; q < D; ++q) b[q][m] = b[q][m] + w * (ds[q] * df);
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
    for (int m = 0; m < M; ++m)
        if (nanf[m]) { umax[m] = __builtin_nan(""); umin[m] = __builtin_nan(""); }
    if (EXT && a.ext_on) {
        double e[2 * M];
#pragma unroll
        for (int m = 0; m < M; ++m) { e[m] = umax[m]; e[M + m] = umin[m]; }
        store_rec<2 * M, 2 * M>(a.ext, i, e);
    }
    if (cached) {
#pragma unroll
        for (int q = 0; q < BG; ++q) {
            double inv[BG];
#pragma unroll
            for (int r = 0; r < BG; ++r) inv[r] = a.invg[(q * BG + r) * N + i];
            const double sc = tscale(tdeg(D, q), invL);
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double c = inv[0] * b[0][m];
#pragma unroll
                for (int r = 1; r < BG; ++r) c = c + inv[r] * b[r][m];
                g[q * M + m] = c * sc;
            }
        }
    } else if (direct) {
        double U[D][D], rhs[D][M], X[D][M];
        if (chol_moment<D>(A, U)) {
#pragma unroll
            for (int q = 0; q < D; ++q)
#pragma unroll
                for (int m = 0; m < M; ++m) rhs[q][m] = b[q][m];
            chol_solve<D, M>(U, rhs, X);
#pragma unroll
            for (int q = 0; q < D; ++q)
#pragma unroll
                for (int m = 0; m < M; ++m) g[q * M + m] = X[q][m] * invL;
        }
    }
    // order 1: zero; order 3 without a factorisation: the direct solve fails the same way, zero
    bool gzero = true;
#pragma unroll
    for (int q = 0; q < L::NG; ++q) gzero = gzero && (g[q] == 0.0);
    if (order > 1 && a.limiter != HCC_LIM_NONE && !gzero) {
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
        // walks the slots in order: each one's increments at the degree-2 interface basis and their limiting
        auto walk = [&](auto&& fn) {
            for (int k = 0; k < cnt; ++k) {
                double pj[D], dd[D], h[D], pint[BG];
                src.template head<D>(src.idx(base + (int64_t)k * kChunk), pj);
                rec_geom<D>(a.g, ri, pj, dd);
#pragma unroll
                for (int d = 0; d < D; ++d) h[d] = 0.5 * dd[d];
                taylor<D, BG>(h, pint);
                double rec[M];
                bool lim[M];
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < BG; ++q) v = v + g[q * M + m] * pint[q];
                    rec[m] = v;
                    lim[m] = fabs(v) > (v > 0.0 ? tup[m] : tdn[m]);
                }
                fn(rec, lim);
            }
        };
        // phi from the extreme increments of each side (phi_scan_finish), the full fold after a near tie
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
#pragma unroll
            for (int m = 0; m < M; ++m) phi[m] = 1.0;
            walk([&](const double (&rec)[M], const bool (&lim)[M]) {
#pragma unroll
                for (int m = 0; m < M; ++m)
                    if (lim[m]) {
                        const double r = (rec[m] > 0.0 ? dup[m] : ddn[m]) / rec[m];
                        phi[m] = jmin(phi[m], bj ? bj_phi(r) : vk_phi(r));
                    }
            });
        }
        const bool hard = a.mode == HCC_MODE_HARD;
#pragma unroll
        for (int q = 0; q < BG; ++q)
            if (hard || q < D)
#pragma unroll
                for (int m = 0; m < M; ++m) g[q * M + m] = g[q * M + m] * phi[m];
    }
    store_rec<L::NG, L::GS>(a.G, i, g);
}

// the step's order reset (TimeSteppers.jl:544-545): the maximum order 3
__device__ __forceinline__ void step_reset3(const StepArgs& a, int64_t i) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counters[1] = 0; a.counters[2] = 0x7FF0000000000000ull; a.counters[4] = 0;
    }
    if (i < a.g.N) { a.orders[i] = 3; a.trig[i] = 0; }
}

template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_gather(StepArgs a, const double* __restrict__ Y, int full, int reset,
                                                    int64_t nblk) {
    if (full && blockIdx.x == 0 && threadIdx.x == 0) { a.counters[0] = 0; a.counters[5] = 0; a.counters[8] = 0; }
    const int64_t i = node_index(nblk);
    if (reset) step_reset3(a, i);
    if (i < a.g.N && !a.g.isb[i] && (full || a.changed[i])) {
        if (full) gather_node<D, M, KIND, true>(a, Y, hbm_src<D, M>(a, Y), i);
        else gather_node<D, M, KIND, false>(a, Y, hbm_src<D, M>(a, Y), i);
    }
}

// the full gather with the block's halo staged in LDS
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_gather_staged(StepArgs a, const double* __restrict__ Y, int reset,
                                                           int64_t nblk) {
    extern __shared__ double sh[];
    if (blockIdx.x == 0 && threadIdx.x == 0) { a.counters[0] = 0; a.counters[5] = 0; a.counters[8] = 0; }
    const int64_t lb = logical_block(nblk);
    if (lb >= nblk) return;
    const int64_t i = lb * kBlock + threadIdx.x;
    if (reset) step_reset3(a, i);
    const auto src = stage_halo3<D, M, false>(a, Y, sh, lb);
    if (i < a.g.N && !a.g.isb[i]) gather_node<D, M, KIND, true>(a, Y, src, i);
}

// ------------------------------------------------------------------------------------------------------------
// flux: the MUSCL functor (MUSCL.jl:390-424) and _flux_loop (:463-508) with the degree-2 interface basis masked to
// each side's interface order (mask_basis, Interpolators.jl:265-284), Rusanov, and the divergence: the cached fit
// (apply_weights_divergence, Interpolators.jl:610-626; the particle's order is 3, or its fit is the cache's degree)
// or the direct degree-1 fit with intercept (compute_dynamic_divergence -> _mls_divergence, :308-346); K_E = -2 div
// ------------------------------------------------------------------------------------------------------------
template <int D, int M, int KIND>
// The particle's own state fi, gradient record gi and flux Fi (flat, axis major) come from registers or, in the wide
// kernels, from the group's LDS copy.
__device__ __forceinline__ void slot3(const StepArgs& a, const double (&ri)[D + M], const double* fi,
                                      const double* gi, const double* Fi, int oi,
                                      const double (&rj)[D + M], const double* __restrict__ gj, int oj,
                                      double (&dd)[D], double& w, double (&dF)[D][M]) {
    using P = Pde<D, M, KIND>;
    constexpr int BG = L3<D, M>::BG;
    const double d2 = rec_geom<D>(a.g, ri, rj, dd);
    w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
    int o_i = oi, o_j = oj;
    if (a.strategy == HCC_STRAT_EPD1) { o_i = oi < oj ? oi : oj; o_j = o_i; }
    double hi[D], hj[D], pii[BG], pjj[BG];
#pragma unroll
    for (int d = 0; d < D; ++d) { hi[d] = 0.5 * dd[d]; hj[d] = -0.5 * dd[d]; }
    taylor<D, BG>(hi, pii);
    taylor<D, BG>(hj, pjj);
    const int ki = bcount(D, o_i - 1), kj = bcount(D, o_j - 1);
#pragma unroll
    for (int q = 0; q < BG; ++q) {
        if (q >= ki) pii[q] = 0.0;
        if (q >= kj) pjj[q] = 0.0;
    }
    double fij[M], fji[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double si = gi[m] * pii[0], sj = gj[m] * pjj[0];
#pragma unroll
        for (int q = 1; q < BG; ++q) { si = si + gi[q * M + m] * pii[q]; sj = sj + gj[q * M + m] * pjj[q]; }
        fij[m] = fi[m] + si;
        fji[m] = rj[D + m] + sj;
    }
    double Fij[D][M], Fji[D][M], lij[D], lji[D];
    P::flux_eig(a.pde, fij, Fij, lij);
    P::flux_eig(a.pde, fji, Fji, lji);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        // the exact shortcuts of the order-2 slot (commuted F_L + F_R, symmetric max of the speeds, the side's sign
        // on the speed, the dropped + 0.0 of the jump)
        const double sd = jmax(lij[d], lji[d]);
        const double sds = dd[d] > 0.0 ? sd : -sd;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const double x = fji[m] - fij[m];
            const double Fnum = 0.5 * (Fij[d][m] + Fji[d][m] - sds * x);
            dF[d][m] = Fnum - Fi[d * M + m];
        }
    }
}

template <int D, int M, int KIND, bool CHECK, class Src>
__device__ __forceinline__ bool flux_node(const StepArgs& a, const PredArgs& pred, const double* __restrict__ Y,
                                          const Src& src, double* __restrict__ K, int64_t i, double h) {
    using L = L3<D, M>;
    constexpr int NR = L::NR, NG = L::NG, BD = L::BD, B1 = L::B1, DM = D * M;
    const int64_t N = a.g.N;
    const int cnt = a.g.cnt[i];
    double ri[NR], fi[M], Kout[M];
    load_rec<NR, L::PU>(Y, i, ri);
#pragma unroll
    for (int m = 0; m < M; ++m) fi[m] = ri[D + m];
    if (cnt == 0) {
#pragma unroll
        for (int m = 0; m < M; ++m) Kout[m] = -0.0 + explicit_source<M>(a, i, fi, m);
        return flux_tail<D, M, KIND, CHECK>(a, pred, K, i, fi, Kout, h);
    }
    double gi[NG], Fi[D][M];
    load_rec<NG, L::GS>(a.G, i, gi);
    Pde<D, M, KIND>::flux(a.pde, fi, Fi);
    // interface order source and the particle's own order (EPD2: the flux pass reads min | order << 4)
    int oi = a.orders[i], oi_t = oi;
    if (a.packed) { oi_t = oi >> 4; oi &= 15; }
    const int64_t base = sell_base(a.g, i);
    const double invL = a.g.invL;
    const int full_idx = div_index(D, 3, cnt);
    const bool okd = a.okd[i];
    const bool cached = okd && div_index(D, oi_t, cnt) == full_idx;
    if (cached || oi_t == 3) {
        // the cached fit (order 3 without a factorisation: the direct solve of the same matrix fails, zero)
        double b[BD][DM];
#pragma unroll
        for (int c = 0; c < BD; ++c)
#pragma unroll
            for (int q = 0; q < DM; ++q) b[c][q] = 0.0;
        if (okd) {
            for (int k = 0; k < cnt; ++k) {
                const int32_t j = src.idx(base + (int64_t)k * kChunk);
                double rj[NR], gj[NG], dd[D], w, dF[D][M];
                src.template head<NR>(j, rj);
                src.grad(j, gj);
                const int oj = src.order(j);
                slot3<D, M, KIND>(a, ri, fi, gi, &Fi[0][0], oi, rj, gj, a.packed ? (oj & 15) : oj, dd, w, dF);
                double ds[D], t[BD - 1];
#pragma unroll
                for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
                taylor<D, BD - 1>(ds, t);
#pragma unroll
                for (int d = 0; d < D; ++d)
#pragma unroll
                    for (int m = 0; m < M; ++m) {
                        b[0][d * M + m] = b[0][d * M + m] + w * (1.0 * dF[d][m]);
#pragma unroll
                        for (int c = 1; c < BD; ++c) b[c][d * M + m] = b[c][d * M + m] + w * (t[c - 1] * dF[d][m]);
                    }
            }
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
                Kout[m] = -(2.0 * (acc * invL)) + explicit_source<M>(a, i, fi, m);
            }
        } else {
#pragma unroll
            for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * 0.0) + explicit_source<M>(a, i, fi, m);
        }
    } else {
        // below the full order: the degree-1 fit with intercept, assembled and factorised here
        double A[B1][B1], b[B1][DM];
#pragma unroll
        for (int r = 0; r < B1; ++r)
#pragma unroll
            for (int c = 0; c < B1; ++c) A[r][c] = 0.0;
#pragma unroll
        for (int c = 0; c < B1; ++c)
#pragma unroll
            for (int q = 0; q < DM; ++q) b[c][q] = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = src.idx(base + (int64_t)k * kChunk);
            double rj[NR], gj[NG], dd[D], w, dF[D][M];
            src.template head<NR>(j, rj);
            src.grad(j, gj);
            const int oj = src.order(j);
            slot3<D, M, KIND>(a, ri, fi, gi, &Fi[0][0], oi, rj, gj, a.packed ? (oj & 15) : oj, dd, w, dF);
            double p[B1];
            p[0] = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) p[1 + d] = dd[d] * invL;
#pragma unroll
            for (int r = 0; r < B1; ++r)
#pragma unroll
                for (int c = 0; c < B1; ++c) A[r][c] = A[r][c] + w * (p[r] * p[c]);
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int m = 0; m < M; ++m)
#pragma unroll
                    for (int c = 0; c < B1; ++c) b[c][d * M + m] = b[c][d * M + m] + w * (p[c] * dF[d][m]);
        }
        double U[B1][B1];
        if (chol_moment<B1>(A, U)) {
            double E[B1][D], X[B1][D];
#pragma unroll
            for (int r = 0; r < B1; ++r)
#pragma unroll
                for (int d = 0; d < D; ++d) E[r][d] = r == 1 + d ? 1.0 : 0.0;
            chol_solve<B1, D>(U, E, X);
#pragma unroll
            for (int m = 0; m < M; ++m) {
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    double r = X[0][d] * b[0][d * M + m];
#pragma unroll
                    for (int c = 1; c < B1; ++c) r = r + X[c][d] * b[c][d * M + m];
                    acc = acc + r;
                }
                Kout[m] = -(2.0 * (acc * invL)) + explicit_source<M>(a, i, fi, m);
            }
        } else {
#pragma unroll
            for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * 0.0) + explicit_source<M>(a, i, fi, m);
        }
    }
    return flux_tail<D, M, KIND, CHECK>(a, pred, K, i, fi, Kout, h);
}

template <int D, int M, int KIND, bool CHECK>
__global__ __launch_bounds__(kBlock) void k_flux(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                  double* __restrict__ K, int full, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    int bad = 0;
    if (i < N) {
        const double h = CHECK ? a.hratio * node_dt(a.dtp, a.inst, i) : 0.0;
        if (a.g.isb[i]) {
            if (full)
#pragma unroll
                for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
        } else if (full || a.dirty[i]) {
            ch = flux_node<D, M, KIND, CHECK>(a, pred, Y, hbm_src<D, M>(a, Y), K, i, h);
            if (CHECK) bad = a.inadm[i];
        }
    }
    if (CHECK) {
        append_changed(a, ch, i);
        if (pred.count) add_inadmissible(a, bad);
    }
}

// the full flux pass with the block's halo (records, gradient records, orders) staged in LDS
template <int D, int M, int KIND, bool CHECK>
__global__ __launch_bounds__(kBlock) void k_flux_staged(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                         double* __restrict__ K, int64_t nblk) {
    extern __shared__ double sh[];
    const int64_t lb = logical_block(nblk);
    if (lb >= nblk) return;
    const double h = CHECK ? a.hratio * node_dt(a.dtp, a.inst, lb * kBlock) : 0.0;
    const auto src = stage_halo3<D, M, true>(a, Y, sh, lb);
    const int64_t i = lb * kBlock + threadIdx.x;
    bool ch = false;
    int bad = 0;
    if (i < a.g.N) {
        if (a.g.isb[i]) {
#pragma unroll
            for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
        } else {
            ch = flux_node<D, M, KIND, CHECK>(a, pred, Y, src, K, i, h);
            if (CHECK) bad = a.inadm[i];
        }
    }
    if (CHECK) {
        append_changed(a, ch, i);
        if (pred.count) add_inadmissible(a, bad);
    }
}

// ------------------------------------------------------------------------------------------------------------
// the fused fixed point (one lane a particle): k_mark_gather drops the listed orders, marks the particles and their
// neighbours dirty (listing each once), and recomputes the listed particles' gradients; k_flux_list recomputes K, the
// candidate and the check over the work list. The bookkeeping is that of the order-2 kernels.
// ------------------------------------------------------------------------------------------------------------
template <int D, int M, int KIND>
__global__ __launch_bounds__(64) void k_mark_gather(StepArgs a, const double* __restrict__ Y, unsigned long long* pub,
                                                    int p) {
    constexpr int kListCap = 512;
    __shared__ int32_t lbuf[kListCap];
    const int64_t n = *changed_any(a, p) ? (int64_t)*changed_count(a, p) : 0;
    const int32_t* cl = changed_list(a, p);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *changed_count(a, 1 - p) = 0;
        *work_count(a, 1 - p) = 0;
        *changed_any(a, 1 - p) = 0;
        if (n > 0) a.counters[4] += 1;
        __hip_atomic_store(pub, (unsigned long long)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    const int lane = threadIdx.x;
    int nl = 0;   // entries in lbuf over all the workgroup's items (uniform), appended when it fills and at the end
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
    for (int64_t t0 = (int64_t)blockIdx.x * 64; t0 < n; t0 += (int64_t)gridDim.x * 64) {
        const int64_t t = t0 + lane;
        int32_t i = 0;
        int cnt = -1;
        if (t < n && cl[t] >= 0) {
            i = cl[t];
            cnt = a.g.cnt[i];
            a.orders[i] = (int8_t)(a.orders[i] - 1);
        }
        int mc = cnt;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
        const int64_t base = cnt >= 0 ? sell_base(a.g, i) : 0;
        for (int k = -1; k < mc; ++k) {
            if (nl > kListCap - 64) flush();
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
        if (cnt >= 0 && !a.g.isb[i]) gather_node<D, M, KIND, false>(a, Y, hbm_src<D, M>(a, Y), i);
    }
    flush();
}

template <int D, int M, int KIND, bool FP = true>
__global__ __launch_bounds__(64) void k_flux_list(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                  double* __restrict__ K, int p) {
    const int64_t n = (int64_t)*work_count(a, p);
    if (FP && blockIdx.x == 0 && threadIdx.x == 0) *changed_count(a, 1 - p) = (unsigned long long)n;
    int dsum = 0;         // this lane's inadmissible delta over its items
    bool anych = false;   // this lane listed a change
    for (int64_t t0 = (int64_t)blockIdx.x * 64; t0 < n; t0 += (int64_t)gridDim.x * 64) {
        const int64_t t = t0 + threadIdx.x;
        bool ch = false;
        int delta = 0;
        int64_t i = -1;
        if (t < n) {
            i = a.dlist[t];
            if (FP) a.dmark[i] = 0u;
            if (!a.g.isb[i]) {
                if constexpr (FP) {
                    const int before = a.inadm[i];
                    ch = flux_node<D, M, KIND, true>(a, pred, Y, hbm_src<D, M>(a, Y), K, i,
                                                     a.hratio * node_dt(a.dtp, a.inst, i));
                    delta = (int)a.inadm[i] - before;
                } else {
                    flux_node<D, M, KIND, false>(a, pred, Y, hbm_src<D, M>(a, Y), K, i, 0.0);
                }
            }
        }
        if constexpr (!FP) continue;
        if (t < n) changed_list(a, 1 - p)[t] = ch ? (int32_t)i : -1;
        anych = anych || ch;
        dsum += delta;
    }
    if constexpr (FP) {
        set_changed_any(a, anych, 1 - p);
        if (pred.count) add_inadmissible_sum(a, dsum);
    }
}

// ------------------------------------------------------------------------------------------------------------
// The flux pass of wide fits (3-D Euler: the divergence fit has BD = 20 terms for D M = 15 flux columns, 300
// accumulators, far beyond a lane's registers): one particle per group of G = 16 lanes. In each chunk of G slots, lane
// gl runs slot gl of the chunk (slot3: the interface states, fluxes, Rusanov, dF) and stores w, the divergence basis
// and dF in LDS; then lane q < D M adds the chunk's terms of fit column q in slot order (BD terms for the cached fit,
// the B1 = D + 1 terms of the direct degree-1 fit, whose moment matrix lane G - 1 accumulates). Every sum keeps the
// order of flux_node, so the results are its bits. Workgroups of one wave run GPW = 4 particles; the chunk loop runs
// to the wave's largest count, so every barrier is uniform.
// ------------------------------------------------------------------------------------------------------------
template <int D, int M>
struct Split3 {
    static constexpr int G = 16, GPW = 64 / G;
    static constexpr int BD = L3<D, M>::BD, DM = D * M, B1 = D + 1;
    static constexpr int SR = BD + DM;        // per slot: w, the divergence basis (BD - 1 terms), dF (axis major)
    static constexpr int XS = B1 * D + 1;     // the direct fit's solved derivative columns, and whether it solved
    static constexpr int NS = L3<D, M>::NG + D * M;   // the particle's gradient record and flux
    static_assert(DM < G, "one fit column per lane, and a lane for the moment matrix");
};
// the wide variants (register-resident accumulators of flux_node would spill)
template <int D, int M>
constexpr bool kWide3 = L3<D, M>::BD * D * M > 160;

template <int D, int M, int KIND, bool CHECK>
__device__ __forceinline__ bool flux_group3(const StepArgs& a, const PredArgs& pred, const double* __restrict__ Y,
                                            double* __restrict__ K, int64_t i, bool take, double h,
                                            double (*rows)[Split3<D, M>::SR], double* xs, double* part, double* nsh,
                                            double* amat) {
    using L = L3<D, M>;
    using S = Split3<D, M>;
    constexpr int NR = L::NR, NG = L::NG, BD = L::BD, DM = S::DM, B1 = S::B1, G = S::G;
    const int gl = threadIdx.x % G;
    const int64_t N = a.g.N;
    const double invL = a.g.invL;
    int cnt = 0, mode = 0;   // mode 1: the cached fit over the slots; 2: the direct degree-1 fit; 0: no slot loop
    int oi = 0;
    double ri[NR], fi[M];
#pragma unroll
    for (int q = 0; q < NR; ++q) ri[q] = 0.0;
    int64_t base = 0;
    // the particle's gradient record and flux, shared by the group's lanes: nsh[0 .. NG) and nsh[NG .. NG + D M)
    if (take) {
        cnt = a.g.cnt[i];
        load_rec<NR, L::PU>(Y, i, ri);
        if (cnt > 0) {
            for (int q = gl; q < NG; q += G) nsh[q] = a.G[i * L::GS + q];
            base = sell_base(a.g, i);
            oi = a.orders[i];
            int oi_t = oi;
            if (a.packed) { oi_t = oi >> 4; oi &= 15; }
            const bool okd = a.okd[i];
            const bool cached = okd && div_index(D, oi_t, cnt) == div_index(D, 3, cnt);
            mode = (cached || oi_t == 3) ? (okd ? 1 : 0) : 2;
        }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) fi[m] = ri[D + m];
    if (mode && gl == 0) {
        double Fi[D][M];
        Pde<D, M, KIND>::flux(a.pde, fi, Fi);
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
            for (int m = 0; m < M; ++m) nsh[NG + d * M + m] = Fi[d][m];
    }
    __syncthreads();
    int mc = mode ? cnt : 0;   // the wave's largest slot count: a uniform chunk loop
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mc = max(mc, __shfl_xor(mc, o));
    const int nb = mode == 1 ? BD : B1;
    // the direct fit's moment matrix lives in the group's LDS (amat, B1 x B1, touched by lane G - 1 only): in
    // registers its 32 VGPRs sat beside every lane's slot arithmetic and pushed it into scratch spills
    double b[BD];
#pragma unroll
    for (int c = 0; c < BD; ++c) b[c] = 0.0;
    if (gl == G - 1)
#pragma unroll
        for (int q = 0; q < B1 * B1; ++q) amat[q] = 0.0;
    for (int k0 = 0; k0 < mc; k0 += G) {
        const int k = k0 + gl;
        if (mode && k < cnt) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double rj[NR], dd[D], w, dF[D][M];
            load_rec<NR, L::PU>(Y, j, rj);
            const int oj = a.orders[j];
            // the neighbour's gradient record is read term by term where the reconstruction needs it (a copy of all
            // its terms in registers beside the fit accumulators spilled)
            slot3<D, M, KIND>(a, ri, fi, nsh, nsh + NG, oi, rj, a.G + (int64_t)j * L::GS, a.packed ? (oj & 15) : oj,
                              dd, w, dF);
            double ds[D], t[BD - 1];
#pragma unroll
            for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
            taylor<D, BD - 1>(ds, t);   // its first D terms are ds itself: the direct fit's basis [1, d invL]
            double* row = rows[gl];
            row[0] = w;
#pragma unroll
            for (int c = 1; c < BD; ++c) row[c] = t[c - 1];
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int m = 0; m < M; ++m) row[BD + d * M + m] = dF[d][m];
        }
        __syncthreads();
        if (mode) {
            const int ns = min(G, cnt - k0);
            if (gl < DM) {
                for (int s = 0; s < ns; ++s) {
                    const double* row = rows[s];
                    const double w = row[0], dFv = row[BD + gl];
                    b[0] = b[0] + w * (1.0 * dFv);
#pragma unroll
                    for (int c = 1; c < BD; ++c)
                        if (c < nb) b[c] = b[c] + w * (row[c] * dFv);
                }
            } else if (gl == G - 1 && mode == 2) {
                for (int s = 0; s < ns; ++s) {
                    const double* row = rows[s];
                    const double w = row[0];
                    double pb[B1];
                    pb[0] = 1.0;
#pragma unroll
                    for (int d = 0; d < D; ++d) pb[1 + d] = row[1 + d];
#pragma unroll
                    for (int r = 0; r < B1; ++r)
#pragma unroll
                        for (int c = 0; c < B1; ++c) amat[r * B1 + c] = amat[r * B1 + c] + w * (pb[r] * pb[c]);
                }
            }
        }
        __syncthreads();
    }
    if (mode == 2 && gl == G - 1) {   // _mls_divergence: (C \ E)' with E the derivative rows of the identity
        double A[B1][B1], U[B1][B1];
#pragma unroll
        for (int r = 0; r < B1; ++r)
#pragma unroll
            for (int c = 0; c < B1; ++c) A[r][c] = amat[r * B1 + c];
        const bool ok = chol_moment<B1>(A, U);
        if (ok) {
            double E[B1][D], X[B1][D];
#pragma unroll
            for (int r = 0; r < B1; ++r)
#pragma unroll
                for (int d = 0; d < D; ++d) E[r][d] = r == 1 + d ? 1.0 : 0.0;
            chol_solve<B1, D>(U, E, X);
#pragma unroll
            for (int r = 0; r < B1; ++r)
#pragma unroll
                for (int d = 0; d < D; ++d) xs[r * D + d] = X[r][d];
        }
        xs[B1 * D] = ok ? 1.0 : 0.0;
    }
    __syncthreads();
    if (mode && gl < DM) {   // the derivative row of fit column gl = d M + m
        const int d = gl / M;
        double r = 0.0;
        if (mode == 1) {
            r = a.invd[(d * BD + 0) * N + i] * b[0];
#pragma unroll
            for (int c = 1; c < BD; ++c) r = r + a.invd[(d * BD + c) * N + i] * b[c];
        } else if (xs[B1 * D] != 0.0) {
            r = xs[0 * D + d] * b[0];
#pragma unroll
            for (int c = 1; c < B1; ++c) r = r + xs[c * D + d] * b[c];
        }
        part[gl] = r;
    }
    __syncthreads();
    bool ch = false;
    if (take && gl == 0) {
        double Kout[M];
        const bool fit = mode == 1 || (mode == 2 && xs[B1 * D] != 0.0);
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (cnt == 0) {
                Kout[m] = -0.0 + explicit_source<M>(a, i, fi, m);
            } else if (fit) {
                double acc = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) acc = acc + part[d * M + m];
                Kout[m] = -(2.0 * (acc * invL)) + explicit_source<M>(a, i, fi, m);
            } else {
                Kout[m] = -(2.0 * 0.0) + explicit_source<M>(a, i, fi, m);
            }
        }
        ch = flux_tail<D, M, KIND, CHECK>(a, pred, K, i, fi, Kout, h);
    }
    __syncthreads();   // the group's LDS is reused by the next particle
    return ch;
}

// the full (or dirty-only) flux pass of the wide variants: four particles a workgroup, XCD-aware order of the quads
template <int D, int M, int KIND, bool CHECK>
__global__ __launch_bounds__(64, 2) void k_flux_wide(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                 double* __restrict__ K, int full, int64_t nquad) {
    using S = Split3<D, M>;
    __shared__ double rows[S::GPW][S::G][S::SR];
    __shared__ double xs[S::GPW][S::XS];
    __shared__ double part[S::GPW][S::DM];
    __shared__ double nsh[S::GPW][S::NS];
    __shared__ double amat[S::GPW][S::B1 * S::B1];
    const int64_t lq = xcd_block(blockIdx.x, nquad);
    if (lq >= nquad) return;
    const int grp = threadIdx.x / S::G, gl = threadIdx.x % S::G;
    const int64_t i = lq * S::GPW + grp;
    bool take = false;
    double h = 0.0;
    if (i < a.g.N) {
        if (a.g.isb[i]) {
            if (full && gl == 0)
#pragma unroll
                for (int m = 0; m < M; ++m) K[i * M + m] = 0.0;
        } else {
            take = full || a.dirty[i];
            if (CHECK) h = a.hratio * node_dt(a.dtp, a.inst, i);
        }
    }
    const bool ch = flux_group3<D, M, KIND, CHECK>(a, pred, Y, K, i, take, h, rows[grp], xs[grp], part[grp], nsh[grp],
                                                   amat[grp]);
    if (CHECK) {
        const int bad = (take && gl == 0) ? a.inadm[i] : 0;
        append_changed(a, ch, i);
        if (pred.count) add_inadmissible(a, bad);
    }
}

// the fixed point's flux pass of the wide variants over its work list (the bookkeeping of k_flux_list)
template <int D, int M, int KIND, bool FP = true>
__global__ __launch_bounds__(64, 2) void k_flux_wide_list(StepArgs a, PredArgs pred, const double* __restrict__ Y,
                                                      double* __restrict__ K, int p) {
    using S = Split3<D, M>;
    __shared__ double rows[S::GPW][S::G][S::SR];
    __shared__ double xs[S::GPW][S::XS];
    __shared__ double part[S::GPW][S::DM];
    __shared__ double nsh[S::GPW][S::NS];
    __shared__ double amat[S::GPW][S::B1 * S::B1];
    const int64_t n = (int64_t)*work_count(a, p);
    if (FP && blockIdx.x == 0 && threadIdx.x == 0) *changed_count(a, 1 - p) = (unsigned long long)n;
    const int grp = threadIdx.x / S::G, gl = threadIdx.x % S::G;
    int dsum = 0;         // this lane's inadmissible delta over its items
    bool anych = false;   // this lane listed a change
    for (int64_t t0 = (int64_t)blockIdx.x * S::GPW; t0 < n; t0 += (int64_t)gridDim.x * S::GPW) {
        const int64_t t = t0 + grp;
        int64_t i = 0;
        bool take = false;
        int before = 0;
        if (t < n) {
            i = a.dlist[t];
            take = !a.g.isb[i];
            if (FP && gl == 0) {
                a.dmark[i] = 0u;   // each dirty particle is listed once (k_mark_gather); cleared for the next
                if (take) before = a.inadm[i];
            }
        }
        const double h = FP && take ? a.hratio * node_dt(a.dtp, a.inst, i) : 0.0;
        const bool ch = flux_group3<D, M, KIND, FP>(a, pred, Y, K, i, take, h, rows[grp], xs[grp], part[grp], nsh[grp],
                                                    amat[grp]);
        if constexpr (!FP) continue;
        const int delta = (take && gl == 0) ? (int)a.inadm[i] - before : 0;
        if (t < n && gl == 0) changed_list(a, 1 - p)[t] = ch ? (int32_t)i : -1;
        anych = anych || (ch && gl == 0);
        dsum += delta;
    }
    if constexpr (FP) {
        set_changed_any(a, anych, 1 - p);
        if (pred.count) add_inadmissible_sum(a, dsum);
    }
}

// ------------------------------------------------------------------------------------------------------------
// the unfused check (_mood_flag! + _mood_drop!, MOOD.jl:289-309) with MOODu2's smooth-extremum override
// (_u2_fails, MOOD.jl:223-271): a component that violates the DMP (and is not a plateau) is accepted when, at
// order 3, every axis's curvature term (gradient record entry D + d) over the stencil has one sign
// ((c_min h^2)(c_max h^2) > -d^2) and comparable magnitudes among the neighbours within 1.5 h
// (|c|_min >= ratio |c|_max, ratio <= 0: off), or is negligible (max(|c_min|, |c_max|) h^2 < d); h = minimum(dx).
// The gradient records are those of the last gather of the checked stage's state (the full pass, then the
// recomputed particles), as the reference's cached gradients.
// ------------------------------------------------------------------------------------------------------------
template <int D, int M, int KIND>
__device__ __forceinline__ bool curvature_fails(const StepArgs& a, const double* __restrict__ Yp, int64_t i,
                                                const bool (&viol)[M]) {
    using L = L3<D, M>;
    if (a.orders[i] < 3) return true;
    const double h = a.hmin, h2 = h * h, nr = 1.5 * h, near_sq = nr * nr, dl = a.delta;
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    double pi[D], gi[L::NG];
    load_rec<D, L::PU>(Yp, i, pi);
    load_rec<L::NG, L::GS>(a.G, i, gi);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double mini[M], maxi[M], mina[M], maxa[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            mini[m] = gi[(D + d) * M + m];
            maxi[m] = mini[m];
            mina[m] = fabs(mini[m]);
            maxa[m] = mina[m];
        }
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double pj[D], dd[D], gj[L::NG];
            load_rec<D, L::PU>(Yp, j, pj);
            const bool near = rec_geom<D>(a.g, pi, pj, dd) <= near_sq;
            load_rec<L::NG, L::GS>(a.G, j, gj);
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const double c = gj[(D + d) * M + m];
                mini[m] = jmin(mini[m], c);
                maxi[m] = jmax(maxi[m], c);
                if (near) { mina[m] = jmin(mina[m], fabs(c)); maxa[m] = jmax(maxa[m], fabs(c)); }
            }
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (!viol[m]) continue;
            if (jmax(fabs(mini[m]), fabs(maxi[m])) * h2 < dl) continue;
            const bool same_sign = (mini[m] * h2) * (maxi[m] * h2) > -(dl * dl);
            const bool comparable = a.ratio <= 0.0 || mina[m] >= a.ratio * maxa[m];
            if (!(same_sign && comparable)) return true;
        }
    }
    return false;
}

template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_mood_flag(StepArgs a, const double* __restrict__ Yp,
                                                       const double* __restrict__ cand, const double* __restrict__ Kp,
                                                       const double* __restrict__ KIp, int full, int64_t nblk) {
    using P = Pde<D, M, KIND>;
    using L = L3<D, M>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    if (i < N && !a.g.isb[i] && (full || a.dirty[i])) {
        double r[L::NR], W[M];
        if (Kp) {
            const double h = a.hratio * node_dt(a.dtp, a.inst, i);
            load_rec<L::NR, L::PU>(Yp, i, r);
#pragma unroll
            for (int m = 0; m < M; ++m)
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
                double mn[M], mx[M];
                if (a.ext_on) {   // the extrema the stage's full gather stored
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
                const bool u2 = a.criterion == HCC_CRIT_U2;
                bool viol[M], any = false;
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const bool v = fabs(mx[m] - mn[m]) >= d3 && (W[m] < mn[m] - dl || W[m] > mx[m] + dl);
                    const bool plateau = u2 && a.range && mx[m] - mn[m] < a.plateau * a.range[m];
                    viol[m] = v && !plateau;
                    any = any || viol[m];
                }
                flag = any && (!u2 || curvature_fails<D, M, KIND>(a, Yp, i, viol));
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

// ------------------------------------------------------------------------------------------------------------
// CFL estimate with the cached weights (_cfl_weights_body!, ParticleGrids.jl:439-450): the per-slot derivative
// weights of the fit in use recomputed from its derivative rows with the operations of _weights_body!
// (MUSCL.jl:211-223, the full-degree intercept basis); a.lam holds the wave speeds (k_lambda)
// ------------------------------------------------------------------------------------------------------------
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_cfl(StepArgs a, const double* __restrict__ U, int64_t nblk) {
    using L = L3<D, M>;
    constexpr int BD = L::BD;
    const int64_t N = a.g.N;
    const int64_t i = node_index(nblk);
    double dt = INFINITY;
    if (i < N && !a.g.isb[i] && a.g.cnt[i] > 0) {
        double li[D], pi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) li[d] = a.lam[i * D + d];
        load_rec<D, L::PU>(U, i, pi);
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        const double invL = a.g.invL;
        double sum_c = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double pj[D], dd[D], ds[D], t[BD - 1], v[BD];
            load_rec<D, L::PU>(U, j, pj);
            const double d2 = rec_geom<D>(a.g, pi, pj, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
#pragma unroll
            for (int d = 0; d < D; ++d) ds[d] = dd[d] * invL;
            taylor<D, BD - 1>(ds, t);
            v[0] = w * 1.0;
#pragma unroll
            for (int c = 1; c < BD; ++c) v[c] = w * t[c - 1];
            double dv = 0.0;
#pragma unroll
            for (int r = 0; r < D; ++r) {
                double acc = a.invd[(r * BD + 0) * N + i] * v[0];
#pragma unroll
                for (int c = 1; c < BD; ++c) acc = acc + a.invd[(r * BD + c) * N + i] * v[c];
                const double term = fabs(acc * invL) * jmax(li[r], a.lam[(int64_t)j * D + r]);
                dv = (r == 0) ? term : dv + term;
            }
            sum_c = sum_c + dv;
        }
        if (sum_c > kFloatMin) dt = (1.0 / (double)D) / sum_c;
    }
    block_min(a, dt, logical_block(nblk));
}

}  // namespace m3
