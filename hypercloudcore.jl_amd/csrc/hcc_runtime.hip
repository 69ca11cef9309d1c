// libhcc: gfx950 kernels of the explicit time-step hot path and the C ABI around them.
//
// Layout in HBM (N nodes, Morton-ordered; SoA everywhere):
//   positions   px, py, pz                 N doubles each (periodic axes wrapped)
//   neighbours  SELL-64: the 64 nodes of one wavefront form a chunk; slot k of node i lives at
//               chunk_off[i/64] + 64 k + i%64 (int32), so the 64 lanes of a wave read 256 contiguous bytes
//               per slot. Distances and MLS weights are recomputed from the positions in every pass with
//               the operations of the neighbour search (bit-identical to the stored values of the reference).
//   states      U[m * N + i]; gradients G[(a * M + m) * N + i]; orders int8; masks uint8.
// One thread per node, a sequential loop over its neighbour slots in slot order: every per-node sum is
// formed in the reference's order, so results are bit-identical to the Float64 CPU path.
#include "hcc_internal.h"
#include <hipcub/hipcub.hpp>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <memory>
#include <new>

using namespace hcc;

// ============================================================================================================
// errors
// ============================================================================================================
static thread_local std::string g_create_err;

#define HCC_TRY_HIP(ctx, expr)                                                                 \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return HCC_E_HIP;                                                                  \
        }                                                                                      \
    } while (0)

static int set_err(hcc_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg; else g_create_err = msg;
    return code;
}

template <class T>
static hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

// ============================================================================================================
// launch geometry: 256-thread workgroups, XCD-aware block order
// ============================================================================================================
constexpr int kBlock = 256;

// The dispatcher hands workgroup b to XCD b % 8. Remap so that XCD x runs one contiguous range of logical
// blocks: Morton-ordered neighbours then share that XCD's L2.
__device__ __forceinline__ int64_t node_index(int64_t nblk) {
    const int64_t b = blockIdx.x, per = (nblk + 7) / 8;
    const int64_t lb = (b & 7) * per + (b >> 3);
    return lb * kBlock + threadIdx.x;
}
static inline int64_t nblocks(int64_t N) { return (N + kBlock - 1) / kBlock; }
static inline unsigned grid_for(int64_t nblk) { return (unsigned)(((nblk + 7) / 8) * 8); }

// ============================================================================================================
// context
// ============================================================================================================
extern "C" int hcc_ctx_create(int device, hcc_ctx** out) {
    if (!out) return set_err(nullptr, HCC_E_ARG, "hcc_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return set_err(nullptr, HCC_E_HIP, "hcc_ctx_create: no HIP device available");
    if (device < 0 || device >= n) return set_err(nullptr, HCC_E_ARG, "hcc_ctx_create: device out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return set_err(nullptr, HCC_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    auto* c = new (std::nothrow) hcc_ctx();
    if (!c) return set_err(nullptr, HCC_E_ARG, "hcc_ctx_create: out of host memory");
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_pinned, 8 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess) {
        g_create_err = std::string("hcc_ctx_create: ") + hipGetErrorString(e);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return HCC_E_HIP;
    }
    *out = c;
    return HCC_OK;
}

extern "C" int hcc_ctx_destroy(hcc_ctx* ctx) {
    if (!ctx) return HCC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) { (void)hipStreamSynchronize(ctx->stream); (void)hipStreamDestroy(ctx->stream); }
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    delete ctx;
    return HCC_OK;
}

extern "C" const char* hcc_last_error(const hcc_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

extern "C" int hcc_ctx_synchronize(hcc_ctx* ctx) {
    if (!ctx) return HCC_E_ARG;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

// ============================================================================================================
// host set-up entry points
// ============================================================================================================
static int check_domain(const hcc_domain_desc* d, std::string& why) {
    if (!d) { why = "domain descriptor is NULL"; return HCC_E_ARG; }
    if (d->D < 1 || d->D > 3) { why = "D must be 1, 2 or 3"; return HCC_E_ARG; }
    for (int a = 0; a < d->D; ++a) {
        if (!(d->dx[a] > 0)) { why = "dx must be positive"; return HCC_E_ARG; }
        if (!(d->maxs[a] > d->mins[a])) { why = "maxs must exceed mins"; return HCC_E_ARG; }
    }
    if (!(d->R > 0)) { why = "R must be positive"; return HCC_E_ARG; }
    return HCC_OK;
}

extern "C" int hcc_cloud_generate(const hcc_domain_desc* desc, int64_t cap, double* pos, uint8_t* is_boundary,
                                  int32_t* tags, int64_t* n_out, double* canvas, int64_t* morton_ties) {
    std::string why;
    int rc = check_domain(desc, why);
    if (rc) return set_err(nullptr, rc, "hcc_cloud_generate: " + why);
    if (!n_out) return set_err(nullptr, HCC_E_ARG, "hcc_cloud_generate: n_out is NULL");
    DomainGeom g;
    g.init(*desc);
    HostCloud c;
    generate_cloud(g, c);
    *n_out = c.N;
    if (canvas) for (int a = 0; a < g.D; ++a) { canvas[a] = g.cmin[a]; canvas[g.D + a] = g.cmax[a]; }
    if (morton_ties) *morton_ties = c.morton_ties;
    if (c.N > cap) return set_err(nullptr, HCC_E_CAPACITY, "hcc_cloud_generate: cap smaller than the node count");
    if (!pos || !is_boundary || !tags) return set_err(nullptr, HCC_E_ARG, "hcc_cloud_generate: NULL output");
    std::memcpy(pos, c.pos.data(), c.pos.size() * sizeof(double));
    std::memcpy(is_boundary, c.isb.data(), c.isb.size());
    std::memcpy(tags, c.tags.data(), c.tags.size() * sizeof(int32_t));
    return HCC_OK;
}

extern "C" int hcc_rng_uniform(uint64_t seed, int64_t n, double* out) {
    if (n < 0 || (n > 0 && !out)) return set_err(nullptr, HCC_E_ARG, "hcc_rng_uniform: bad arguments");
    Xoshiro r(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = r.uniform();
    return HCC_OK;
}

// ============================================================================================================
// neighbour search kernels (src/Grid/NeighborLogic.jl)
// ============================================================================================================
struct SearchArgs {
    int64_t N;
    double* p[3];
    int32_t periodic[3];
    double cmin[3], L[3], bin_size[3], Lw[3], invLw[3];
    int64_t dims[3];
    double R_sq, alpha, inv_range_sq;
    int32_t* bin_of;
    int32_t* cell_start;
    int32_t* cursor;
    int32_t* cell_particles;
    const int32_t* window;
    int32_t* cnt;
    const int64_t* chunk_off;
    int32_t* nbr;
};

// _wrap_position! (NeighborLogic.jl:279-286): canvas_min + mod(p - canvas_min, L) on periodic axes
template <int D>
__global__ __launch_bounds__(kBlock) void k_wrap(SearchArgs a, int64_t nblk) {
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    for (int d = 0; d < D; ++d)
        if (a.periodic[d]) a.p[d][i] = a.cmin[d] + jmod(a.p[d][i] - a.cmin[d], a.L[d]);
}

// flat bin of a position: clamp(floor((x - min) / size) + 1, 1, dims), x fastest (NeighborLogic.jl:86-102)
template <int D>
__device__ __forceinline__ int32_t flat_bin(const SearchArgs& a, int64_t i) {
    int64_t lin = 0, stride = 1;
    for (int d = 0; d < D; ++d) {
        int64_t ix = (int64_t)floor((a.p[d][i] - a.cmin[d]) / a.bin_size[d]) + 1;
        ix = ix < 1 ? 1 : (ix > a.dims[d] ? a.dims[d] : ix);
        lin += (ix - 1) * stride;
        stride *= a.dims[d];
    }
    return (int32_t)lin;
}

// _bin_count! (NeighborLogic.jl:139-144)
template <int D>
__global__ __launch_bounds__(kBlock) void k_bin_count(SearchArgs a, int64_t nblk) {
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    int32_t b = flat_bin<D>(a, i);
    a.bin_of[i] = b;
    atomicAdd(&a.cursor[b], 1);
}

// _bin_scatter! (NeighborLogic.jl:146-151); the order the atomics land in is undone by k_bin_sort
__global__ __launch_bounds__(kBlock) void k_bin_scatter(SearchArgs a, int64_t nblk) {
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    int32_t b = a.bin_of[i];
    int32_t k = atomicAdd(&a.cursor[b], 1);
    a.cell_particles[a.cell_start[b] + k] = (int32_t)i;
}

// _bin_sort! (NeighborLogic.jl:154-167): each bin's segment in descending particle index
__global__ __launch_bounds__(kBlock) void k_bin_sort(SearchArgs a, int64_t nbins) {
    int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (b >= nbins) return;
    int32_t lo = a.cell_start[b], hi = a.cell_start[b + 1];
    for (int32_t s = lo + 1; s < hi; ++s) {
        int32_t v = a.cell_particles[s];
        int32_t t = s - 1;
        while (t >= lo && a.cell_particles[t] < v) { a.cell_particles[t + 1] = a.cell_particles[t]; --t; }
        a.cell_particles[t + 1] = v;
    }
}

// get_distance (ParticleGrids.jl:209-219) + sum(abs2, .)
template <int D>
__device__ __forceinline__ double search_dist(const SearchArgs& a, int64_t i, int64_t j, double (&dd)[D]) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        double x = a.p[d][j] - a.p[d][i];
        x = x - a.Lw[d] * rint(x * a.invLw[d]);
        dd[d] = x;
        s = (d == 0) ? x * x : s + x * x;
    }
    return s;
}

// _count_neighbours! (NeighborLogic.jl:288-301)
template <int D>
__global__ __launch_bounds__(kBlock) void k_count(SearchArgs a, int64_t nblk) {
    constexpr int NW = D == 1 ? 3 : D == 2 ? 9 : 27;
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    const int32_t* win = a.window + (int64_t)a.bin_of[i] * NW;
    int32_t c = 0;
    for (int w = 0; w < NW; ++w) {
        int32_t b = win[w];
        for (int32_t s = a.cell_start[b]; s < a.cell_start[b + 1]; ++s) {
            int32_t j = a.cell_particles[s];
            if (j == i) continue;
            double dd[D];
            if (search_dist<D>(a, i, j, dd) <= a.R_sq) ++c;
        }
    }
    a.cnt[i] = c;
}

// widest node of each chunk of 64 (one wave per chunk)
__global__ __launch_bounds__(kBlock) void k_chunk_width(const int32_t* cnt, int64_t N, int64_t nchunks, int64_t* width) {
    int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int64_t c = t / kChunk;
    int32_t v = (t < N) ? cnt[t] : 0;
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0 && c < nchunks) width[c] = (int64_t)v * kChunk;
}

// _write_neighbours! (NeighborLogic.jl:309-329), into the SELL-64 layout
template <int D>
__global__ __launch_bounds__(kBlock) void k_write(SearchArgs a, int64_t nblk) {
    constexpr int NW = D == 1 ? 3 : D == 2 ? 9 : 27;
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    const int32_t* win = a.window + (int64_t)a.bin_of[i] * NW;
    int64_t base = a.chunk_off[i / kChunk] + (i % kChunk);
    int32_t k = 0;
    for (int w = 0; w < NW; ++w) {
        int32_t b = win[w];
        for (int32_t s = a.cell_start[b]; s < a.cell_start[b + 1]; ++s) {
            int32_t j = a.cell_particles[s];
            if (j == i) continue;
            double dd[D];
            if (search_dist<D>(a, i, j, dd) <= a.R_sq) { a.nbr[base + (int64_t)k * kChunk] = j; ++k; }
        }
    }
}

// per-slot weight and distance in SELL order (for the host download of the neighbour data)
template <int D>
__global__ __launch_bounds__(kBlock) void k_slot_geometry(SearchArgs a, int64_t nblk, double* w_out, double* d_out, int64_t S_sell) {
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    int64_t base = a.chunk_off[i / kChunk] + (i % kChunk);
    for (int32_t k = 0; k < a.cnt[i]; ++k) {
        int64_t s = base + (int64_t)k * kChunk;
        double dd[D];
        double d2 = search_dist<D>(a, i, a.nbr[s], dd);
        w_out[s] = mls_weight(a.alpha, a.inv_range_sq, d2);
        for (int d = 0; d < D; ++d) d_out[d * S_sell + s] = dd[d];
    }
}

template <class T>
static hipError_t exclusive_scan(hcc_ctx* ctx, const T* in, T* out, int64_t n) {
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, ctx->stream);
    if (e != hipSuccess) return e;
    void* tmp = nullptr;
    e = hipMalloc(&tmp, bytes ? bytes : 1);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, ctx->stream);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    return e != hipSuccess ? e : e2;
}

hcc::GeomArgs hcc_grid::args() const {
    GeomArgs a{};
    a.N = N;
    a.px = d_pos[0]; a.py = d_pos[1] ? d_pos[1] : d_pos[0]; a.pz = d_pos[2] ? d_pos[2] : d_pos[0];
    a.isb = d_isb; a.cnt = d_cnt; a.chunk_off = d_chunk_off; a.nbr = d_nbr;
    for (int d = 0; d < 3; ++d) { a.Lw[d] = geom.Lw[d]; a.invLw[d] = geom.invLw[d]; }
    a.invL = geom.invL; a.alpha = geom.desc.alpha; a.inv_range_sq = geom.inv_range_sq;
    return a;
}

static SearchArgs search_args(const hcc_grid* g) {
    SearchArgs a{};
    a.N = g->N;
    for (int d = 0; d < 3; ++d) {
        a.p[d] = g->d_pos[d];
        a.periodic[d] = d < g->geom.D ? g->geom.desc.periodic[d] : 0;
        a.cmin[d] = g->geom.cmin[d]; a.L[d] = g->geom.L[d]; a.bin_size[d] = g->bin_size[d];
        a.Lw[d] = g->geom.Lw[d]; a.invLw[d] = g->geom.invLw[d]; a.dims[d] = g->dims[d];
    }
    a.R_sq = g->geom.R_sq; a.alpha = g->geom.desc.alpha; a.inv_range_sq = g->geom.inv_range_sq;
    a.bin_of = g->d_bin_of; a.cell_start = g->d_cell_start; a.cursor = g->d_cell_cursor;
    a.cell_particles = g->d_cell_particles; a.window = g->d_bin_window;
    a.cnt = g->d_cnt; a.chunk_off = g->d_chunk_off; a.nbr = g->d_nbr;
    return a;
}

template <int D>
static int run_search(hcc_grid* g) {
    hcc_ctx* ctx = g->ctx;
    hipStream_t st = ctx->stream;
    const int64_t N = g->N, nblk = nblocks(N);
    SearchArgs a = search_args(g);
    if (N > 0) {
        k_wrap<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_cell_cursor, 0, sizeof(int32_t) * (g->nbins + 1), st));
        k_bin_count<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    // segment starts: exclusive prefix sum of the counts; the dummy bin nbins is empty
    HCC_TRY_HIP(ctx, exclusive_scan<int32_t>(ctx, g->d_cell_cursor, g->d_cell_start, g->nbins + 1));
    int32_t nn = (int32_t)N;
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_cell_start + g->nbins + 1, &nn, sizeof(int32_t), hipMemcpyHostToDevice));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_cell_cursor, 0, sizeof(int32_t) * (g->nbins + 1), st));
    if (N > 0) {
        k_bin_scatter<<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        k_bin_sort<<<(unsigned)((g->nbins + kBlock - 1) / kBlock), kBlock, 0, st>>>(a, g->nbins);
        k_count<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    // chunk widths -> SELL offsets
    g->nchunks = (N + kChunk - 1) / kChunk;
    int64_t* d_width = nullptr;
    HCC_TRY_HIP(ctx, dalloc(&d_width, g->nchunks + 1));
    HCC_TRY_HIP(ctx, hipMemsetAsync(d_width, 0, sizeof(int64_t) * (g->nchunks + 1), st));
    HCC_TRY_HIP(ctx, dalloc(&g->d_chunk_off, g->nchunks + 1));
    if (N > 0) k_chunk_width<<<(unsigned)((g->nchunks * kChunk + kBlock - 1) / kBlock), kBlock, 0, st>>>(g->d_cnt, N, g->nchunks, d_width);
    HCC_TRY_HIP(ctx, exclusive_scan<int64_t>(ctx, d_width, g->d_chunk_off, g->nchunks + 1));
    (void)hipFree(d_width);
    HCC_TRY_HIP(ctx, hipMemcpy(&g->S_sell, g->d_chunk_off + g->nchunks, sizeof(int64_t), hipMemcpyDeviceToHost));
    HCC_TRY_HIP(ctx, dalloc(&g->d_nbr, g->S_sell));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_nbr, 0, sizeof(int32_t) * std::max<int64_t>(g->S_sell, 1), st));
    a = search_args(g);
    if (N > 0) k_write<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
    HCC_TRY_HIP(ctx, hipGetLastError());
    std::vector<int32_t> cnt((size_t)N);
    if (N > 0) HCC_TRY_HIP(ctx, hipMemcpyAsync(cnt.data(), g->d_cnt, sizeof(int32_t) * N, hipMemcpyDeviceToHost, st));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    g->S = 0; g->max_nb = 0;
    for (int64_t i = 0; i < N; ++i) { g->S += cnt[i]; g->max_nb = std::max<int64_t>(g->max_nb, cnt[i]); }
    return HCC_OK;
}

extern "C" int hcc_grid_destroy(hcc_grid* g) {
    if (!g) return HCC_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    for (int d = 0; d < 3; ++d) (void)hipFree(g->d_pos[d]);
    (void)hipFree(g->d_isb); (void)hipFree(g->d_tags); (void)hipFree(g->d_bin_of); (void)hipFree(g->d_cell_start); (void)hipFree(g->d_cell_cursor);
    (void)hipFree(g->d_cell_particles); (void)hipFree(g->d_bin_window); (void)hipFree(g->d_cnt); (void)hipFree(g->d_chunk_off); (void)hipFree(g->d_nbr);
    delete g;
    return HCC_OK;
}

extern "C" int hcc_grid_create(hcc_ctx* ctx, const hcc_domain_desc* desc, int64_t N, const double* pos,
                               const uint8_t* is_boundary, const int32_t* tags, hcc_grid** out) {
    if (!ctx || !out) return set_err(ctx, HCC_E_ARG, "hcc_grid_create: NULL context or output");
    *out = nullptr;
    std::string why;
    int rc = check_domain(desc, why);
    if (rc) return set_err(ctx, rc, "hcc_grid_create: " + why);
    if (N < 0 || N > (int64_t)INT32_MAX - 1 || (N > 0 && (!pos || !is_boundary || !tags)))
        return set_err(ctx, HCC_E_ARG, "hcc_grid_create: bad node arrays");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<hcc_grid, int (*)(hcc_grid*)> g(new hcc_grid(), hcc_grid_destroy);
    g->ctx = ctx;
    g->geom.init(*desc);
    g->N = N;
    const int D = desc->D;
    // cell list geometry (GlobalBins, NeighborLogic.jl:22-40): bins at least R wide over the canvas
    g->nbins = 1;
    for (int d = 0; d < 3; ++d) { g->dims[d] = 1; g->bin_size[d] = 1.0; }
    for (int d = 0; d < D; ++d) {
        double size = g->geom.cmax[d] - g->geom.cmin[d];
        g->dims[d] = std::max<int64_t>(1, (int64_t)std::floor(size / desc->R));
        g->bin_size[d] = size / (double)g->dims[d];
        g->nbins *= g->dims[d];
    }
    if (g->nbins > (int64_t)INT32_MAX / 27) return set_err(ctx, HCC_E_ARG, "hcc_grid_create: too many bins");
    std::vector<int32_t> window;
    bin_window_table(D, g->dims, desc->periodic, window);
    // upload
    std::vector<double> col((size_t)N);
    for (int d = 0; d < D; ++d) {
        HCC_TRY_HIP(ctx, dalloc(&g->d_pos[d], N));
        for (int64_t i = 0; i < N; ++i) col[i] = pos[i * D + d];
        HCC_TRY_HIP(ctx, hipMemcpy(g->d_pos[d], col.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    }
    HCC_TRY_HIP(ctx, dalloc(&g->d_isb, N));
    HCC_TRY_HIP(ctx, dalloc(&g->d_tags, N));
    if (N) {
        HCC_TRY_HIP(ctx, hipMemcpy(g->d_isb, is_boundary, N, hipMemcpyHostToDevice));
        HCC_TRY_HIP(ctx, hipMemcpy(g->d_tags, tags, sizeof(int32_t) * N, hipMemcpyHostToDevice));
    }
    g->h_isb.assign(is_boundary, is_boundary + N);
    g->h_tags.assign(tags, tags + N);
    HCC_TRY_HIP(ctx, dalloc(&g->d_bin_of, N));
    HCC_TRY_HIP(ctx, dalloc(&g->d_cell_start, g->nbins + 2));
    HCC_TRY_HIP(ctx, dalloc(&g->d_cell_cursor, g->nbins + 1));
    HCC_TRY_HIP(ctx, dalloc(&g->d_cell_particles, N));
    HCC_TRY_HIP(ctx, dalloc(&g->d_bin_window, window.size()));
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_bin_window, window.data(), sizeof(int32_t) * window.size(), hipMemcpyHostToDevice));
    HCC_TRY_HIP(ctx, dalloc(&g->d_cnt, N));
    rc = D == 1 ? run_search<1>(g.get()) : D == 2 ? run_search<2>(g.get()) : run_search<3>(g.get());
    if (rc) return rc;
    *out = g.release();
    return HCC_OK;
}

extern "C" int hcc_grid_num_slots(const hcc_grid* g, int64_t* S, int64_t* max_nb) {
    if (!g) return HCC_E_ARG;
    if (S) *S = g->S;
    if (max_nb) *max_nb = g->max_nb;
    return HCC_OK;
}

// host copies of the SELL lists as row-major ranges + indices
static int download_lists(hcc_grid* g, std::vector<int64_t>& rstart, std::vector<int32_t>& idx,
                          std::vector<int64_t>* sell_of_slot) {
    hcc_ctx* ctx = g->ctx;
    const int64_t N = g->N;
    std::vector<int32_t> cnt((size_t)N), nbr((size_t)std::max<int64_t>(g->S_sell, 1));
    std::vector<int64_t> off((size_t)(g->nchunks + 1));
    if (N) {
        HCC_TRY_HIP(ctx, hipMemcpy(cnt.data(), g->d_cnt, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
        HCC_TRY_HIP(ctx, hipMemcpy(off.data(), g->d_chunk_off, sizeof(int64_t) * (g->nchunks + 1), hipMemcpyDeviceToHost));
        if (g->S_sell) HCC_TRY_HIP(ctx, hipMemcpy(nbr.data(), g->d_nbr, sizeof(int32_t) * g->S_sell, hipMemcpyDeviceToHost));
    }
    rstart.assign((size_t)N + 1, 0);
    for (int64_t i = 0; i < N; ++i) rstart[i + 1] = rstart[i] + cnt[i];
    idx.assign((size_t)rstart[N], 0);
    if (sell_of_slot) sell_of_slot->assign((size_t)rstart[N], 0);
    for (int64_t i = 0; i < N; ++i) {
        int64_t base = off[i / kChunk] + (i % kChunk);
        for (int32_t k = 0; k < cnt[i]; ++k) {
            idx[rstart[i] + k] = nbr[base + (int64_t)k * kChunk];
            if (sell_of_slot) (*sell_of_slot)[rstart[i] + k] = base + (int64_t)k * kChunk;
        }
    }
    return HCC_OK;
}

extern "C" int hcc_grid_get_neighbors(hcc_grid* g, int64_t* ranges, int32_t* indices, double* weights, double* distances) {
    if (!g) return HCC_E_ARG;
    hcc_ctx* ctx = g->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<int64_t> rstart, sell;
    std::vector<int32_t> idx;
    int rc = download_lists(g, rstart, idx, &sell);
    if (rc) return rc;
    if (ranges) std::memcpy(ranges, rstart.data(), sizeof(int64_t) * rstart.size());
    if (indices && !idx.empty()) std::memcpy(indices, idx.data(), sizeof(int32_t) * idx.size());
    if ((weights || distances) && g->N > 0) {
        const int D = g->geom.D;
        double *dw = nullptr, *dd = nullptr;
        HCC_TRY_HIP(ctx, dalloc(&dw, g->S_sell));
        HCC_TRY_HIP(ctx, dalloc(&dd, g->S_sell * D));
        SearchArgs a = search_args(g);
        int64_t nblk = nblocks(g->N);
        if (D == 1) k_slot_geometry<1><<<grid_for(nblk), kBlock, 0, ctx->stream>>>(a, nblk, dw, dd, g->S_sell);
        else if (D == 2) k_slot_geometry<2><<<grid_for(nblk), kBlock, 0, ctx->stream>>>(a, nblk, dw, dd, g->S_sell);
        else k_slot_geometry<3><<<grid_for(nblk), kBlock, 0, ctx->stream>>>(a, nblk, dw, dd, g->S_sell);
        std::vector<double> hw((size_t)g->S_sell), hd((size_t)(g->S_sell * D));
        HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        HCC_TRY_HIP(ctx, hipMemcpy(hw.data(), dw, sizeof(double) * g->S_sell, hipMemcpyDeviceToHost));
        HCC_TRY_HIP(ctx, hipMemcpy(hd.data(), dd, sizeof(double) * g->S_sell * D, hipMemcpyDeviceToHost));
        (void)hipFree(dw); (void)hipFree(dd);
        for (size_t k = 0; k < sell.size(); ++k) {
            if (weights) weights[k] = hw[sell[k]];
            if (distances) for (int d = 0; d < D; ++d) distances[k * D + d] = hd[d * g->S_sell + sell[k]];
        }
    }
    return HCC_OK;
}

extern "C" int hcc_grid_get_bins(hcc_grid* g, int32_t* bin_of, int64_t* cell_start, int32_t* cell_particles, int64_t* dims) {
    if (!g) return HCC_E_ARG;
    hcc_ctx* ctx = g->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (bin_of && g->N) HCC_TRY_HIP(ctx, hipMemcpy(bin_of, g->d_bin_of, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
    if (cell_particles && g->N) HCC_TRY_HIP(ctx, hipMemcpy(cell_particles, g->d_cell_particles, sizeof(int32_t) * g->N, hipMemcpyDeviceToHost));
    if (cell_start) {
        std::vector<int32_t> cs((size_t)(g->nbins + 2));
        HCC_TRY_HIP(ctx, hipMemcpy(cs.data(), g->d_cell_start, sizeof(int32_t) * cs.size(), hipMemcpyDeviceToHost));
        for (size_t b = 0; b < cs.size(); ++b) cell_start[b] = cs[b];
    }
    if (dims) { dims[0] = g->dims[0]; dims[1] = g->dims[1]; dims[2] = g->dims[2]; dims[3] = g->nbins; }
    return HCC_OK;
}

extern "C" int hcc_grid_get_positions(hcc_grid* g, double* pos) {
    if (!g || (!pos && g->N)) return HCC_E_ARG;
    hcc_ctx* ctx = g->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    const int D = g->geom.D;
    std::vector<double> col((size_t)g->N);
    for (int d = 0; d < D; ++d) {
        if (g->N) HCC_TRY_HIP(ctx, hipMemcpy(col.data(), g->d_pos[d], sizeof(double) * g->N, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < g->N; ++i) pos[i * D + d] = col[i];
    }
    return HCC_OK;
}

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
        double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
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
    // dweights: (rows 2..D+1 of inv(N)) * (w p~) * invL (MUSCL.jl:211-223)
    for (int k = 0; k < cnt; ++k) {
        const int64_t s = base + (int64_t)k * kChunk;
        if (!ok) { for (int d = 0; d < D; ++d) a.dw[d * a.S_sell + s] = 0.0; continue; }
        int32_t j = a.g.nbr[s];
        double dd[D];
        double d2 = slot_geom<D>(a.g, pi, j, dd);
        double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
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

// ============================================================================================================
// step kernels
// ============================================================================================================

// _stage_predictor! (TimeSteppers.jl:400-421) fused with the outflow condition: a boundary particle with an
// outflow donor takes the stage value of its interior root (the same combination, bit-identical), others keep
// pg.rhos (= U^n during the step).
template <int M>
__global__ __launch_bounds__(kBlock) void k_predict(int64_t N, int64_t nblk, const uint8_t* isb, const int32_t* root,
                                                     const double* Un, const double* K0, const double* K1, const double* K2,
                                                     const double* K3, const double* K4, double c0, double c1, double c2,
                                                     double c3, double c4, int nterm, unsigned use, double* target) {
    const int64_t i = node_index(nblk);
    if (i >= N) return;
    const int64_t src = isb[i] ? (int64_t)root[i] : i;
    if (src < 0) {
        for (int m = 0; m < M; ++m) target[m * N + i] = Un[m * N + i];
        return;
    }
    const double* Ks[5] = {K0, K1, K2, K3, K4};
    const double cs[5] = {c0, c1, c2, c3, c4};
    for (int m = 0; m < M; ++m) {
        double y = Un[m * N + src];
        for (int j = 0; j < nterm; ++j)
            if ((use >> j) & 1u) y = y + cs[j] * Ks[j][m * N + src];
        target[m * N + i] = y;
    }
}

// update_content! (MUSCL.jl:351-384): cached reconstruction fit (Interpolators.jl:585-602) and the
// Venkatakrishnan / Barth-Jespersen limiter (Limiter.jl:61-143) for one interior particle.
template <int D, int M, int KIND>
__device__ __forceinline__ void gather_node(const StepArgs& a, const double* __restrict__ Y, int64_t i) {
    const int64_t N = a.g.N;
    const int cnt = a.g.cnt[i];
    int order = a.orders[i];
    if (order > 1 && cnt < D) { order = 1; a.orders[i] = 1; }   // starved neighbourhood (basis_count(D, 1) = D)
    double g[D][M];
    for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m) g[q][m] = 0.0;
    if (order == 2 && a.okg[i]) {
        const NodePos<D> pi = load_pos<D>(a.g, i);
        const int64_t base = sell_base(a.g, i);
        const double invL = a.g.invL;
        double fi[M], umax[M], umin[M], b[D][M];
        for (int m = 0; m < M; ++m) { fi[m] = Y[m * N + i]; umax[m] = fi[m]; umin[m] = fi[m]; }
        for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m) b[q][m] = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double dd[D];
            const double d2 = slot_geom<D>(a.g, pi, j, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
            double p[D];
            for (int d = 0; d < D; ++d) p[d] = dd[d] * invL;
            for (int m = 0; m < M; ++m) {
                const double fj = Y[m * N + j];
                const double df = fj - fi[m];
                for (int q = 0; q < D; ++q) b[q][m] = b[q][m] + w * (p[q] * df);
                umax[m] = jmax(umax[m], fj);
                umin[m] = jmin(umin[m], fj);
            }
        }
        double inv[D][D];
        for (int r = 0; r < D; ++r) for (int c = 0; c < D; ++c) inv[r][c] = a.invg[(r * D + c) * N + i];
        for (int q = 0; q < D; ++q)
            for (int m = 0; m < M; ++m) {
                double c = inv[q][0] * b[0][m];
                for (int r = 1; r < D; ++r) c = c + inv[q][r] * b[r][m];
                g[q][m] = c * invL;
            }
        if (a.limiter != HCC_LIM_NONE) {
            double phi[M];
            for (int m = 0; m < M; ++m) phi[m] = 1.0;
            for (int k = 0; k < cnt; ++k) {
                const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
                double dd[D];
                slot_geom<D>(a.g, pi, j, dd);
                double pint[D];
                for (int d = 0; d < D; ++d) pint[d] = 0.5 * dd[d];
                for (int m = 0; m < M; ++m) {
                    double rec = 0.0;
                    for (int q = 0; q < D; ++q) rec = rec + g[q][m] * pint[q];
                    const double bound = rec > 0.0 ? umax[m] : umin[m];
                    const double tol = 16 * kEps * jmax(fabs(fi[m]), fabs(bound));
                    if (fabs(rec) > tol) {
                        const double r = (bound - fi[m]) / rec;
                        phi[m] = jmin(phi[m], a.limiter == HCC_LIM_BJ ? bj_phi(r) : vk_phi(r));
                    }
                }
            }
            for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m) g[q][m] = g[q][m] * phi[m];
        }
    }
    for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m) a.grads[(q * M + m) * N + i] = g[q][m];
}

// The gather pass over all interior particles (full) or the `changed` mask (TimeSteppers.jl:329-334).
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_gather(StepArgs a, const double* __restrict__ Y, int full, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= a.g.N) return;
    if (a.g.isb[i] || !(full || a.changed[i])) return;
    gather_node<D, M, KIND>(a, Y, i);
}

// The gather pass over a compacted list of particles whose order changed (the `changed` set of the MOOD fixed point).
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_gather_list(StepArgs a, const double* __restrict__ Y, const int32_t* __restrict__ list,
                                                         int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const int64_t i = list[t];
    if (a.g.isb[i]) return;
    gather_node<D, M, KIND>(a, Y, i);
}

// MUSCL functor + _flux_loop (MUSCL.jl:390-508): per slot the EPD-masked midpoint reconstructions, two physical
// fluxes, the Rusanov flux sorted by the sign of the distance (Utils.jl:95-103, FluxFunctions.jl:16-76), and
// the flux difference accumulated straight into the moments of the cached divergence fit
// (apply_weights_divergence, Interpolators.jl:610-626); K_E = -2 div (TimeSteppers.jl:354).
//
// With CHECK the kernel also runs MOOD's candidate check on the node it just updated (_mood_flag!, MOOD.jl:289-295):
// the forward Euler substep W = Y + h K is node-local (form_substep!, TimeSteppers.jl:501-518) and the DMP bounds of
// MOODu1 (_dmp_fails, MOOD.jl:175-184) are the extrema of Y over the same neighbour loop. The set of nodes it checks
// is exactly the set whose K it recomputes (all interior nodes, or the dirty ones), as in detect_and_drop!. The order
// decrement (_mood_drop!) is deferred to k_expand_apply: neighbours read this node's order in this very kernel.
template <int D, int M, int KIND, bool CHECK>
__device__ __forceinline__ bool flux_node(const StepArgs& a, const double* __restrict__ Y, double* __restrict__ K,
                                          int64_t i, double h) {
    using P = Pde<D, M, KIND>;
    constexpr int BD = D + 1;
    const int64_t N = a.g.N;
    const int cnt = a.g.cnt[i];
    double fi[M];
    for (int m = 0; m < M; ++m) fi[m] = Y[m * N + i];
    double mn[M], mx[M];
    for (int m = 0; m < M; ++m) { mn[m] = fi[m]; mx[m] = fi[m]; }
    double Kout[M];
    if (cnt == 0) {
        for (int m = 0; m < M; ++m) Kout[m] = -0.0 + 0.0;
    } else {
        const NodePos<D> pi = load_pos<D>(a.g, i);
        const int64_t base = sell_base(a.g, i);
        const double invL = a.g.invL;
        double gi[D][M], Fi[D][M];
        for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m) gi[q][m] = a.grads[(q * M + m) * N + i];
        P::flux(a.pde, fi, Fi);
        const int oi = a.orders[i];
        double b[BD][D * M];
        for (int c = 0; c < BD; ++c) for (int q = 0; q < D * M; ++q) b[c][q] = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double dd[D];
            const double d2 = slot_geom<D>(a.g, pi, j, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
            const int oj = a.orders[j];
            int o_i = oi, o_j = oj;
            if (a.strategy == HCC_STRAT_EPD1) { o_i = oi < oj ? oi : oj; o_j = o_i; }
            double pii[D], pjj[D];
            for (int d = 0; d < D; ++d) {
                pii[d] = (o_i == 2) ? 0.5 * dd[d] : 0.0;
                pjj[d] = (o_j == 2) ? -0.5 * dd[d] : 0.0;
            }
            double fij[M], fji[M];
            for (int m = 0; m < M; ++m) {
                double si = gi[0][m] * pii[0];
                for (int q = 1; q < D; ++q) si = si + gi[q][m] * pii[q];
                double sj = a.grads[(0 * M + m) * N + j] * pjj[0];
                for (int q = 1; q < D; ++q) sj = sj + a.grads[(q * M + m) * N + j] * pjj[q];
                const double uj = Y[m * N + j];
                if (CHECK) { mn[m] = jmin(mn[m], uj); mx[m] = jmax(mx[m], uj); }
                fij[m] = fi[m] + si;
                fji[m] = uj + sj;
            }
            double Fij[D][M], Fji[D][M], lij[D], lji[D];
            P::flux_eig(a.pde, fij, Fij, lij);
            P::flux_eig(a.pde, fji, Fji, lji);
            double pt[BD];
            pt[0] = 1.0;
            for (int d = 0; d < D; ++d) pt[1 + d] = dd[d] * invL;
            for (int d = 0; d < D; ++d) {
                // sort_flux orders the two sides by sign(d); F_L + F_R is the same sum either way (IEEE addition
                // commutes), only the dissipation f_R - f_L needs the sorted states
                const bool right = dd[d] > 0.0;
                const double sd = right ? jmax(lij[d], lji[d]) : jmax(lji[d], lij[d]);
                for (int m = 0; m < M; ++m) {
                    const double fL = right ? fij[m] : fji[m], fR = right ? fji[m] : fij[m];
                    const double diss = sd * (fR - fL);
                    const double Fnum = 0.5 * (Fij[d][m] + Fji[d][m] - diss);
                    const double dF = (Fnum - Fi[d][m]) + 0.0;
                    for (int c = 0; c < BD; ++c) b[c][d * M + m] = b[c][d * M + m] + w * (pt[c] * dF);
                }
            }
        }
        if (a.okd[i]) {
            for (int m = 0; m < M; ++m) {
                double acc = 0.0;
                for (int d = 0; d < D; ++d) {
                    double r = a.invd[(d * BD + 0) * N + i] * b[0][d * M + m];
                    for (int c = 1; c < BD; ++c) r = r + a.invd[(d * BD + c) * N + i] * b[c][d * M + m];
                    acc = acc + r;
                }
                Kout[m] = -(2.0 * (acc * invL)) + 0.0;
            }
        } else {
            for (int m = 0; m < M; ++m) Kout[m] = -(2.0 * 0.0) + 0.0;
        }
    }
    for (int m = 0; m < M; ++m) K[m * N + i] = Kout[m];
    if (!CHECK) return false;
    double W[M];
    for (int m = 0; m < M; ++m) W[m] = fi[m] + h * Kout[m];
    const bool adm = P::admissible(a.pde, W);
    a.inadm[i] = !adm;
    bool flag = !adm;
    if (!flag) {
        if (a.criterion == HCC_CRIT_ONLY) {
            flag = true;
        } else if (a.criterion == HCC_CRIT_U1) {
            const double dl = a.delta, d3 = dl * dl * dl;
            for (int m = 0; m < M; ++m)
                if (fabs(mx[m] - mn[m]) >= d3 && (W[m] < mn[m] - dl || W[m] > mx[m] + dl)) { flag = true; break; }
        }
    }
    if (flag && a.orders[i] > 1) { a.trig[i] = 1; return true; }
    return false;
}

// Appends the flagged particles of a wavefront to `clist` (one atomic per wavefront) and counts them.
__device__ __forceinline__ void append_changed(const StepArgs& a, bool ch, int64_t i) {
    const unsigned long long ball = __ballot(ch);
    if (!ball) return;
    const unsigned lane = threadIdx.x & 63;
    const unsigned leader = (unsigned)(__ffsll((unsigned long long)ball) - 1);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(&a.counters[0], (unsigned long long)__popcll(ball));
    base = __shfl(base, (int)leader);
    if (ch) {
        const unsigned long long below = ball & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        a.clist[base + __popcll(below)] = (int32_t)i;
    }
}

template <int D, int M, int KIND, bool CHECK>
__global__ __launch_bounds__(kBlock) void k_flux(StepArgs a, const double* __restrict__ Y, double* __restrict__ K, int full,
                                                  double h, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    if (i < N) {
        if (a.g.isb[i]) {
            if (full) for (int m = 0; m < M; ++m) K[m * N + i] = 0.0;
        } else if (full || a.dirty[i]) {
            ch = flux_node<D, M, KIND, CHECK>(a, Y, K, i, h);
        }
        if (CHECK) a.changed[i] = ch;
    }
    if (CHECK) append_changed(a, ch, i);
}

// The flux pass (with the MOOD check) over the compacted dirty list of the fixed point; clears each entry's mark.
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_flux_list(StepArgs a, const double* __restrict__ Y, double* __restrict__ K,
                                                       double h) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t n = (int64_t)a.counters[3];
    bool ch = false;
    int64_t i = 0;
    if (t < n) {
        i = a.dlist[t];
        a.dmark[i] = 0;
        if (!a.g.isb[i]) ch = flux_node<D, M, KIND, true>(a, Y, K, i, h);
        a.changed[i] = ch;
    }
    append_changed(a, ch, i);
}

// _mood_drop! (MOOD.jl:299-309) and one hop of expand_mask! (MOOD.jl:323-351) for the changed list: the particle
// drops one order, and it and its neighbours join the dirty list once (first marker appends). The neighbour
// relation is symmetric, so marking the neighbours of changed particles gives the pull form's set.
__global__ __launch_bounds__(kBlock) void k_mark(StepArgs a, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const int32_t i = a.clist[t];
    a.orders[i] = (int8_t)(a.orders[i] - 1);
    if (atomicExch(&a.dmark[i], 1u) == 0u) a.dlist[atomicAdd(&a.counters[3], 1ull)] = i;
    const int cnt = a.g.cnt[i];
    const int64_t base = sell_base(a.g, i);
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
        if (atomicExch(&a.dmark[j], 1u) == 0u) a.dlist[atomicAdd(&a.counters[3], 1ull)] = j;
    }
}

// _mood_flag! + _mood_drop! (MOOD.jl:289-309) fused, on the forward Euler substep W = Y + h K
// (form_substep!, TimeSteppers.jl:501-518) or on a stored candidate. The DMP test of MOODu1 is _dmp_fails
// (MOOD.jl:175-184). Every particle writes its `changed` entry (the reference clears the mask first).
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_mood_flag(StepArgs a, const double* __restrict__ Yp, const double* __restrict__ Kp,
                                                       const double* __restrict__ cand, double h, int full, int64_t nblk) {
    using P = Pde<D, M, KIND>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    bool ch = false;
    if (i < N && !a.g.isb[i] && (full || a.dirty[i])) {
        double W[M];
        if (cand) for (int m = 0; m < M; ++m) W[m] = cand[m * N + i];
        else for (int m = 0; m < M; ++m) W[m] = Yp[m * N + i] + h * Kp[m * N + i];
        const bool adm = P::admissible(a.pde, W);
        a.inadm[i] = !adm;
        bool flag = !adm;
        if (!flag) {
            if (a.criterion == HCC_CRIT_ONLY) {
                flag = true;
            } else if (a.criterion == HCC_CRIT_U1) {
                double mn[M], mx[M];
                for (int m = 0; m < M; ++m) { mn[m] = Yp[m * N + i]; mx[m] = mn[m]; }
                const int cnt = a.g.cnt[i];
                const int64_t base = sell_base(a.g, i);
                for (int k = 0; k < cnt; ++k) {
                    const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
                    for (int m = 0; m < M; ++m) {
                        const double v = Yp[m * N + j];
                        mn[m] = jmin(mn[m], v);
                        mx[m] = jmax(mx[m], v);
                    }
                }
                const double dl = a.delta, d3 = dl * dl * dl;
                for (int m = 0; m < M; ++m)
                    if (fabs(mx[m] - mn[m]) >= d3 && (W[m] < mn[m] - dl || W[m] > mx[m] + dl)) { flag = true; break; }
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

// expand_mask! by one hop (MOOD.jl:323-351): dirty = changed plus the neighbours of changed particles. With
// apply_drop the decrement of _mood_drop! (MOOD.jl:299-309) flagged by the fused check is applied here first.
__global__ __launch_bounds__(kBlock) void k_expand(StepArgs a, int apply_drop, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= a.g.N) return;
    uint8_t m = a.changed[i];
    if (apply_drop && m) a.orders[i] = (int8_t)(a.orders[i] - 1);
    if (!m) {
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        for (int k = 0; k < cnt; ++k)
            if (a.changed[a.g.nbr[base + (int64_t)k * kChunk]]) { m = 1; break; }
    }
    a.dirty[i] = m;
}

// _count_inadmissible (TimeSteppers.jl:481-490): interior candidates accepted while not admissible
__global__ __launch_bounds__(kBlock) void k_count_inadm(StepArgs a, int64_t nblk) {
    const int64_t i = node_index(nblk);
    const bool bad = i < a.g.N && !a.g.isb[i] && a.inadm[i];
    const unsigned long long ball = __ballot(bad);
    if (ball && ((threadIdx.x & 63) == (unsigned)(__ffsll((unsigned long long)__ballot(1)) - 1)))
        atomicAdd(&a.counters[1], (unsigned long long)__popcll(ball));
}

// max_eigenvalue of every particle once per CFL estimate (_get_Lambda, ParticleGrids.jl:294)
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_lambda(StepArgs a, const double* __restrict__ U, double* __restrict__ lam, int64_t nblk) {
    using P = Pde<D, M, KIND>;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    if (i >= N) return;
    double f[M], l[D];
    for (int m = 0; m < M; ++m) f[m] = U[m * N + i];
    P::maxeig(a.pde, f, l);
    for (int d = 0; d < D; ++d) lam[d * N + i] = l[d];
}

// _cfl_weights_body! (ParticleGrids.jl:439-450) and the minimum over particles (positive doubles order as their
// bit patterns, so the minimum is an integer atomic). The per-slot derivative weights are recomputed from the
// divergence inverse and the slot geometry with the operations of _weights_body! (MUSCL.jl:211-223): the same
// bits as the cached dweights, without streaming them.
template <int D, int M, int KIND>
__global__ __launch_bounds__(kBlock) void k_cfl(StepArgs a, int64_t nblk) {
    constexpr int BD = D + 1;
    const int64_t i = node_index(nblk);
    const int64_t N = a.g.N;
    double dt = INFINITY;
    if (i < N && !a.g.isb[i] && a.g.cnt[i] > 0) {
        double li[D], X[D][BD];
        for (int d = 0; d < D; ++d) li[d] = a.lam[d * N + i];
        for (int r = 0; r < D; ++r) for (int c = 0; c < BD; ++c) X[r][c] = a.invd[(r * BD + c) * N + i];
        const NodePos<D> pi = load_pos<D>(a.g, i);
        const int cnt = a.g.cnt[i];
        const int64_t base = sell_base(a.g, i);
        const double invL = a.g.invL;
        double sum_c = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.g.nbr[base + (int64_t)k * kChunk];
            double dd[D];
            const double d2 = slot_geom<D>(a.g, pi, j, dd);
            const double w = mls_weight(a.g.alpha, a.g.inv_range_sq, d2);
            double v[BD];
            v[0] = w * 1.0;
            for (int d = 0; d < D; ++d) v[1 + d] = w * (dd[d] * invL);
            double dv = 0.0;
            for (int r = 0; r < D; ++r) {
                double acc = X[r][0] * v[0];
                for (int c = 1; c < BD; ++c) acc = acc + X[r][c] * v[c];
                const double term = fabs(acc * invL) * jmax(li[r], a.lam[r * N + j]);
                dv = (r == 0) ? term : dv + term;
            }
            sum_c = sum_c + dv;
        }
        if (sum_c > kFloatMin) dt = (1.0 / (double)D) / sum_c;
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0 && dt < INFINITY) atomicMin(&a.counters[2], (unsigned long long)__double_as_longlong(dt));
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double* p, double v, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ __launch_bounds__(kBlock) void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
    int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) p[t] = v;
}
// mode 0: the changed count; 1: everything of a new step / CFL estimate; 2: the dirty-list length
__global__ void k_reset_counters(unsigned long long* c, int mode) {
    if (threadIdx.x == 0) {
        if (mode == 2) { c[3] = 0; return; }
        c[0] = 0;
        if (mode == 1) { c[1] = 0; c[2] = 0x7FF0000000000000ull; c[3] = 0; }
    }
}

// ============================================================================================================
// the stepper
// ============================================================================================================
struct hcc_stepper {
    hcc_ctx* ctx = nullptr;
    hcc_grid* grid = nullptr;
    hcc_stepper_desc desc{};
    int D = 0, M = 0, kind = 0;
    int64_t N = 0;
    // caches
    double* invg = nullptr; uint8_t* okg = nullptr;
    double* invd = nullptr; uint8_t* okd = nullptr;
    double* dw = nullptr;
    int32_t* root = nullptr;
    // states
    double* U = nullptr;       // pg.rhos (U^n during a step)
    double* cand = nullptr;    // rho_cand
    double* Y[kMaxStages] = {};
    double* K[kMaxStages] = {};
    double* grads = nullptr;
    int8_t* orders = nullptr;
    uint8_t *changed = nullptr, *dirty = nullptr, *trig = nullptr, *inadm = nullptr;
    unsigned long long* counters = nullptr;
    int32_t *clist = nullptr, *dlist = nullptr;
    unsigned* dmark = nullptr;
    double* lam = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evp[4 * kMaxStages] = {};   // start/end of each full gather and flux pass of a step
    double last_ms = 0, last_rhs_ms = 0, last_gather_ms = 0, last_flux_ms = 0;
    int64_t inadmissible = 0, mood_iters = 0;
    StepArgs args() const;
};

StepArgs hcc_stepper::args() const {
    StepArgs a{};
    a.g = grid->args();
    a.pde.gamma = desc.gamma;
    for (int d = 0; d < 3; ++d) {
        double mx = 0.0;
        for (int m = 0; m < 8; ++m) { a.pde.c[d][m] = desc.adv_speed[d][m]; if (m < M) mx = std::max(mx, std::fabs(desc.adv_speed[d][m])); }
        a.pde.amax[d] = mx;
    }
    a.limiter = desc.limiter; a.mode = desc.limiter_mode; a.strategy = desc.mood_strategy;
    a.criterion = desc.mood_criterion; a.delta = desc.mood_delta;
    a.invg = invg; a.invd = invd; a.okg = okg; a.okd = okd;
    a.grads = grads; a.orders = orders; a.changed = changed; a.dirty = dirty; a.trig = trig; a.inadm = inadm;
    a.counters = counters;
    a.clist = clist; a.dlist = dlist; a.dmark = dmark; a.lam = lam;
    return a;
}

// compiled variants: (D, M, kind)
template <int D, int M, int KIND>
struct Variant {
    static void gather(const hcc_stepper* s, const StepArgs& a, const double* Y, int full) {
        int64_t nb = nblocks(s->N);
        k_gather<D, M, KIND><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, Y, full, nb);
    }
    static void flux(const hcc_stepper* s, const StepArgs& a, const double* Y, double* K, int full) {
        int64_t nb = nblocks(s->N);
        k_flux<D, M, KIND, false><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, Y, K, full, 0.0, nb);
    }
    static void flux_check(const hcc_stepper* s, const StepArgs& a, const double* Y, double* K, int full, double h) {
        int64_t nb = nblocks(s->N);
        k_flux<D, M, KIND, true><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, Y, K, full, h, nb);
    }
    static void flag(const hcc_stepper* s, const StepArgs& a, const double* Yp, const double* Kp, const double* cand, double h, int full) {
        int64_t nb = nblocks(s->N);
        k_mood_flag<D, M, KIND><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, Yp, Kp, cand, h, full, nb);
    }
    static void cfl(const hcc_stepper* s, const StepArgs& a) {
        int64_t nb = nblocks(s->N);
        k_lambda<D, M, KIND><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, s->U, s->lam, nb);
        k_cfl<D, M, KIND><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, nb);
    }
    static void gather_list(const hcc_stepper* s, const StepArgs& a, const double* Y, int64_t n) {
        if (n > 0) k_gather_list<D, M, KIND><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s->ctx->stream>>>(a, Y, s->clist, n);
    }
    static void flux_list(const hcc_stepper* s, const StepArgs& a, const double* Y, double* K, int64_t bound, double h) {
        if (bound > 0) k_flux_list<D, M, KIND><<<(unsigned)((bound + kBlock - 1) / kBlock), kBlock, 0, s->ctx->stream>>>(a, Y, K, h);
    }
    static void predict(const hcc_stepper* s, const double* const* Ks, const double* c, int nterm, unsigned use, double* target) {
        int64_t nb = nblocks(s->N);
        k_predict<M><<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(s->N, nb, s->grid->d_isb, s->root, s->U, Ks[0], Ks[1], Ks[2], Ks[3],
                                                                   Ks[4], c[0], c[1], c[2], c[3], c[4], nterm, use, target);
    }
};

struct VariantOps {
    void (*gather)(const hcc_stepper*, const StepArgs&, const double*, int);
    void (*flux)(const hcc_stepper*, const StepArgs&, const double*, double*, int);
    void (*flux_check)(const hcc_stepper*, const StepArgs&, const double*, double*, int, double);
    void (*flag)(const hcc_stepper*, const StepArgs&, const double*, const double*, const double*, double, int);
    void (*cfl)(const hcc_stepper*, const StepArgs&);
    void (*gather_list)(const hcc_stepper*, const StepArgs&, const double*, int64_t);
    void (*flux_list)(const hcc_stepper*, const StepArgs&, const double*, double*, int64_t, double);
    void (*predict)(const hcc_stepper*, const double* const*, const double*, int, unsigned, double*);
};
template <int D, int M, int KIND>
static VariantOps make_ops() {
    using V = Variant<D, M, KIND>;
    return VariantOps{&V::gather, &V::flux, &V::flux_check, &V::flag, &V::cfl, &V::gather_list, &V::flux_list, &V::predict};
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

extern "C" int hcc_stepper_destroy(hcc_stepper* s) {
    if (!s) return HCC_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(s->invg); (void)hipFree(s->okg); (void)hipFree(s->invd); (void)hipFree(s->okd); (void)hipFree(s->dw); (void)hipFree(s->root);
    (void)hipFree(s->U); (void)hipFree(s->cand);
    for (int k = 0; k < kMaxStages; ++k) { (void)hipFree(s->Y[k]); (void)hipFree(s->K[k]); }
    (void)hipFree(s->grads); (void)hipFree(s->orders); (void)hipFree(s->changed); (void)hipFree(s->dirty); (void)hipFree(s->trig); (void)hipFree(s->inadm);
    (void)hipFree(s->counters);
    (void)hipFree(s->clist); (void)hipFree(s->dlist); (void)hipFree(s->dmark); (void)hipFree(s->lam);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (auto& e : s->evp) if (e) (void)hipEventDestroy(e);
    delete s;
    return HCC_OK;
}

extern "C" int hcc_stepper_create(hcc_ctx* ctx, hcc_grid* g, const hcc_stepper_desc* d, hcc_stepper** out) {
    if (!ctx || !g || !d || !out) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: NULL argument");
    *out = nullptr;
    const int D = g->geom.D;
    VariantOps ops;
    if (!find_ops(D, d->M, d->pde, ops))
        return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: no compiled kernel for this (D, M, PDE) -- there is no fallback path");
    if (d->muscl_order != 2) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MUSCL order 2 is the compiled scheme");
    if (d->stages < 1 || d->stages > kMaxStages) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: 1..5 stages");
    if (d->limiter < HCC_LIM_NONE || d->limiter > HCC_LIM_VK) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: unknown limiter");
    if (d->mood_strategy < HCC_STRAT_NONE || d->mood_strategy > HCC_STRAT_EPD1) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MOOD strategy");
    if (d->mood_criterion < HCC_CRIT_NONE || d->mood_criterion > HCC_CRIT_ONLY) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_stepper_create: MOOD criterion");
    if (d->n_bc < 0 || d->n_bc > 16) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: at most 16 boundary conditions");
    for (int i = 0; i < d->stages; ++i)
        for (int j = i; j < d->stages; ++j)
            if (d->a_t[i][j] != 0.0) return set_err(ctx, HCC_E_ARG, "hcc_stepper_create: a_t must be strictly lower triangular");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<hcc_stepper, int (*)(hcc_stepper*)> s(new hcc_stepper(), hcc_stepper_destroy);
    s->ctx = ctx; s->grid = g; s->desc = *d; s->D = D; s->M = d->M; s->kind = d->pde; s->N = g->N;
    const int64_t N = g->N, M = d->M;
    HCC_TRY_HIP(ctx, dalloc(&s->invg, (int64_t)D * D * N));
    HCC_TRY_HIP(ctx, dalloc(&s->okg, N));
    HCC_TRY_HIP(ctx, dalloc(&s->invd, (int64_t)D * (D + 1) * N));
    HCC_TRY_HIP(ctx, dalloc(&s->okd, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dw, (int64_t)D * g->S_sell));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->dw, 0, sizeof(double) * std::max<int64_t>((int64_t)D * g->S_sell, 1), ctx->stream));
    HCC_TRY_HIP(ctx, dalloc(&s->root, N));
    HCC_TRY_HIP(ctx, dalloc(&s->U, M * N));
    HCC_TRY_HIP(ctx, dalloc(&s->cand, M * N));
    for (int k = 0; k < d->stages; ++k) {
        HCC_TRY_HIP(ctx, dalloc(&s->Y[k], M * N));
        HCC_TRY_HIP(ctx, dalloc(&s->K[k], M * N));
        HCC_TRY_HIP(ctx, hipMemsetAsync(s->K[k], 0, sizeof(double) * std::max<int64_t>(M * N, 1), ctx->stream));
    }
    HCC_TRY_HIP(ctx, dalloc(&s->grads, (int64_t)D * M * N));
    HCC_TRY_HIP(ctx, dalloc(&s->orders, N));
    HCC_TRY_HIP(ctx, dalloc(&s->changed, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dirty, N));
    HCC_TRY_HIP(ctx, dalloc(&s->trig, N));
    HCC_TRY_HIP(ctx, dalloc(&s->inadm, N));
    HCC_TRY_HIP(ctx, dalloc(&s->counters, 8));
    HCC_TRY_HIP(ctx, dalloc(&s->clist, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dlist, N));
    HCC_TRY_HIP(ctx, dalloc(&s->dmark, N));
    HCC_TRY_HIP(ctx, dalloc(&s->lam, (int64_t)D * N));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->dmark, 0, sizeof(unsigned) * std::max<int64_t>(N, 1), ctx->stream));
    for (uint8_t* p : {s->changed, s->dirty, s->trig, s->inadm})
        HCC_TRY_HIP(ctx, hipMemsetAsync(p, 0, std::max<int64_t>(N, 1), ctx->stream));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->orders, 2, std::max<int64_t>(N, 1), ctx->stream));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->U, 0, sizeof(double) * std::max<int64_t>(M * N, 1), ctx->stream));
    HCC_TRY_HIP(ctx, hipMemsetAsync(s->counters, 0, 8 * sizeof(unsigned long long), ctx->stream));
    HCC_TRY_HIP(ctx, hipEventCreate(&s->ev0));
    HCC_TRY_HIP(ctx, hipEventCreate(&s->ev1));
    for (auto& e : s->evp) HCC_TRY_HIP(ctx, hipEventCreate(&e));
    // geometry caches
    if (N > 0) {
        MomentArgs ma{g->args(), s->invg, s->okg, s->invd, s->okd, s->dw, g->S_sell};
        int64_t nb = nblocks(N);
        if (D == 1) k_moments<1><<<grid_for(nb), kBlock, 0, ctx->stream>>>(ma, nb);
        else if (D == 2) k_moments<2><<<grid_for(nb), kBlock, 0, ctx->stream>>>(ma, nb);
        else k_moments<3><<<grid_for(nb), kBlock, 0, ctx->stream>>>(ma, nb);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    // outflow donor roots (host, from the device lists and the wrapped positions)
    {
        std::vector<int64_t> rstart;
        std::vector<int32_t> idx;
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
    }
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = s.release();
    return HCC_OK;
}

static void to_soa(const double* aos, std::vector<double>& soa, int64_t N, int M) {
    soa.resize((size_t)(N * M));
    for (int64_t i = 0; i < N; ++i) for (int m = 0; m < M; ++m) soa[(size_t)(m * N + i)] = aos[i * M + m];
}
static void to_aos(const std::vector<double>& soa, double* aos, int64_t N, int M) {
    for (int64_t i = 0; i < N; ++i) for (int m = 0; m < M; ++m) aos[i * M + m] = soa[(size_t)(m * N + i)];
}

extern "C" int hcc_state_upload(hcc_stepper* s, const double* U) {
    if (!s || (!U && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> soa;
    to_soa(U, soa, s->N, s->M);
    if (s->N) HCC_TRY_HIP(ctx, hipMemcpyAsync(s->U, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

extern "C" int hcc_state_download(hcc_stepper* s, double* U) {
    if (!s || (!U && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> soa((size_t)(s->N * s->M));
    if (s->N) HCC_TRY_HIP(ctx, hipMemcpyAsync(soa.data(), s->U, sizeof(double) * soa.size(), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    to_aos(soa, U, s->N, s->M);
    return HCC_OK;
}

extern "C" int hcc_time_step(hcc_stepper* s, double* dt) {
    if (!s || !dt) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    VariantOps ops;
    find_ops(s->D, s->M, s->kind, ops);
    k_reset_counters<<<1, 64, 0, ctx->stream>>>(s->counters, 1);
    if (s->N) ops.cfl(s, s->args());
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, s->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long bits = ctx->h_pinned[2];
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *dt = v;
    return HCC_OK;
}

// host driver of one step (TimeSteppers.jl:531-577)
namespace {
struct StepRunner {
    hcc_stepper* s;
    VariantOps ops;
    StepArgs a;
    int64_t nb;
    int iters = 0;
    bool substeps() const { return s->desc.substep_ratio > 0.0 && s->desc.mood_criterion != HCC_CRIT_NONE; }

    // form_candidate! (TimeSteppers.jl:365-397) for checkpoint i (1-based; i = s + 1 is the final state)
    void predict(int i, double dt) {
        const int st = s->desc.stages;
        const bool fin = i > st;
        const int n_prev = fin ? st : i - 1;
        const double* Ks[5] = {s->K[0], s->K[0], s->K[0], s->K[0], s->K[0]};
        double c[5] = {0, 0, 0, 0, 0};
        unsigned use = 0;
        for (int j = 0; j < n_prev; ++j) {
            Ks[j] = s->K[j];
            const double at = fin ? s->desc.b_t[j] : s->desc.a_t[i - 1][j];
            c[j] = dt * at;
            if (at != 0.0) use |= 1u << j;
        }
        ops.predict(s, Ks, c, n_prev, use, fin ? s->cand : s->Y[i - 1]);
    }
    hipError_t read_changed(unsigned long long& n) {
        hcc_ctx* ctx = s->ctx;
        hipError_t e = hipMemcpyAsync(ctx->h_pinned, s->counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(ctx->stream);
        n = ctx->h_pinned[0];
        return e;
    }
    // MOOD checks the Euler substep with a criterion that reads neighbour states: the check is fused into the flux
    // pass that produces K (flux_node with CHECK).
    bool fused() const { return substeps(); }

    // mood_checkpoint! (TimeSteppers.jl:454-477). The candidate Y_i is formed once the fixed point is reached:
    // on the substep path the check reads only Y_{i-1} and K_{i-1}, so re-forming it every iteration changes nothing.
    // Fused path: the first (full) check already ran inside stage i-1's flux pass; each recompute of K_{i-1} on the
    // dirty set re-checks exactly that set.
    hipError_t checkpoint(int i, double dt) {
        if (s->desc.mood_criterion == HCC_CRIT_NONE) { predict(i, dt); return hipSuccess; }
        const double* Yp = s->Y[i - 2];
        double* Kp = s->K[i - 2];
        const double h = s->desc.substep_ratio * dt;
        if (fused()) {
            // the changed particles of the last check are listed in clist (counters[0] of them)
            for (;;) {
                unsigned long long n = 0;
                hipError_t e = read_changed(n);
                if (e != hipSuccess) return e;
                if (n == 0) break;
                ++iters;
                hipStream_t str = s->ctx->stream;
                k_reset_counters<<<1, 64, 0, str>>>(s->counters, 2);               // dlist length = 0
                k_mark<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, str>>>(a, (int64_t)n);
                ops.gather_list(s, a, Yp, (int64_t)n);
                k_reset_counters<<<1, 64, 0, str>>>(s->counters, 0);               // clist restarts
                const int64_t bound = std::min<int64_t>(s->N, (int64_t)n * (1 + s->grid->max_nb));
                ops.flux_list(s, a, Yp, Kp, bound, h);
            }
            k_count_inadm<<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, nb);
            predict(i, dt);
            return hipGetLastError();
        }
        int full = 1;
        for (;;) {
            predict(i, dt);
            k_reset_counters<<<1, 64, 0, s->ctx->stream>>>(s->counters, 0);
            ops.flag(s, a, Yp, Kp, i > s->desc.stages ? s->cand : s->Y[i - 1], h, full);
            unsigned long long n = 0;
            hipError_t e = read_changed(n);
            if (e != hipSuccess) return e;
            if (n == 0) break;
            ++iters;
            k_expand<<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, 0, nb);
            ops.gather(s, a, Yp, 0);
            ops.flux(s, a, Yp, Kp, 0);
            full = 0;
        }
        k_count_inadm<<<grid_for(nb), kBlock, 0, s->ctx->stream>>>(a, nb);
        return hipGetLastError();
    }
    hipError_t run(double dt) {
        hcc_ctx* ctx = s->ctx;
        const int64_t N = s->N;
        const int st = s->desc.stages;
        nb = nblocks(N);
        const unsigned fb = (unsigned)((s->D * s->M * N + kBlock - 1) / kBlock);
        k_fill_f64<<<fb, kBlock, 0, ctx->stream>>>(s->grads, 0.0, (int64_t)s->D * s->M * N);        // update_size!(interp, N)
        k_fill_u8<<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>((uint8_t*)s->orders, 2, N);
        k_fill_u8<<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(s->trig, 0, N);
        k_reset_counters<<<1, 64, 0, ctx->stream>>>(s->counters, 1);
        for (int i = 1; i <= st; ++i) {
            if (i == 1) predict(1, dt);
            else { hipError_t e = checkpoint(i, dt); if (e != hipSuccess) return e; }
            hipStream_t str = ctx->stream;
            (void)hipEventRecord(s->evp[4 * (i - 1) + 0], str);
            ops.gather(s, a, s->Y[i - 1], 1);
            (void)hipEventRecord(s->evp[4 * (i - 1) + 1], str);
            if (fused()) k_reset_counters<<<1, 64, 0, str>>>(s->counters, 0);
            (void)hipEventRecord(s->evp[4 * (i - 1) + 2], str);
            if (fused()) ops.flux_check(s, a, s->Y[i - 1], s->K[i - 1], 1, s->desc.substep_ratio * dt);
            else ops.flux(s, a, s->Y[i - 1], s->K[i - 1], 1);
            (void)hipEventRecord(s->evp[4 * (i - 1) + 3], str);
        }
        hipError_t e = checkpoint(st + 1, dt);
        if (e != hipSuccess) return e;
        std::swap(s->U, s->cand);        // pg.rhos <- rho_cand
        return hipGetLastError();
    }
};
}  // namespace

extern "C" int hcc_step(hcc_stepper* s, double t, double dt) {
    (void)t;
    if (!s) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    if (!(dt > 0) || !std::isfinite(dt)) return set_err(ctx, HCC_E_ARG, "hcc_step: dt must be positive and finite");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    if (s->N == 0) return HCC_OK;
    StepRunner r;
    r.s = s;
    find_ops(s->D, s->M, s->kind, r.ops);
    r.a = s->args();
    HCC_TRY_HIP(ctx, hipEventRecord(s->ev0, ctx->stream));
    hipError_t e = r.run(dt);
    // the cand buffer moved: refresh the argument block for the next call (args() is rebuilt per call)
    if (e != hipSuccess) { ctx->err = std::string("hcc_step: ") + hipGetErrorString(e); return HCC_E_HIP; }
    HCC_TRY_HIP(ctx, hipEventRecord(s->ev1, ctx->stream));
    HCC_TRY_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, s->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->inadmissible = (int64_t)ctx->h_pinned[1];
    s->mood_iters = r.iters;
    float ms = 0;
    HCC_TRY_HIP(ctx, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->last_gather_ms = s->last_flux_ms = 0;
    for (int i = 0; i < s->desc.stages; ++i) {
        float g = 0, f = 0;
        HCC_TRY_HIP(ctx, hipEventElapsedTime(&g, s->evp[4 * i + 0], s->evp[4 * i + 1]));
        HCC_TRY_HIP(ctx, hipEventElapsedTime(&f, s->evp[4 * i + 2], s->evp[4 * i + 3]));
        s->last_gather_ms += g;
        s->last_flux_ms += f;
    }
    return HCC_OK;
}

extern "C" int hcc_get_orders(hcc_stepper* s, int32_t* orders) {
    if (!s || (!orders && s->N)) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
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
    *count = s->inadmissible;
    return HCC_OK;
}

extern "C" int hcc_get_mood_iterations(hcc_stepper* s, int64_t* count) {
    if (!s || !count) return HCC_E_ARG;
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
    const int D = s->D, BD = D + 1;
    if (inv_grad && N) {
        std::vector<double> v((size_t)(D * D * N));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->invg, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) for (int q = 0; q < D * D; ++q) inv_grad[i * D * D + q] = v[(size_t)(q * N + i)];
    }
    if (inv_div && N) {   // rows 1..D of the (D+1)^2 inverse; row 0 is not kept on the device (NaN)
        std::vector<double> v((size_t)(D * BD * N));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->invd, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) {
            for (int c = 0; c < BD; ++c) inv_div[i * BD * BD + c] = NAN;
            for (int r = 0; r < D; ++r) for (int c = 0; c < BD; ++c)
                inv_div[i * BD * BD + (1 + r) * BD + c] = v[(size_t)((r * BD + c) * N + i)];
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
        std::vector<double> v((size_t)(D * SS));
        HCC_TRY_HIP(ctx, hipMemcpy(v.data(), s->dw, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < sell.size(); ++k) for (int d = 0; d < D; ++d) dweights[k * D + d] = v[(size_t)(d * SS + sell[k])];
    }
    return HCC_OK;
}

extern "C" int hcc_rhs(hcc_stepper* s, const double* Y, double* K, double* gradients) {
    if (!s || (s->N > 0 && (!Y || !K))) return HCC_E_ARG;
    hcc_ctx* ctx = s->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = s->N;
    const int M = s->M, D = s->D;
    if (!N) return HCC_OK;
    VariantOps ops;
    find_ops(D, M, s->kind, ops);
    std::vector<double> soa;
    to_soa(Y, soa, N, M);
    HCC_TRY_HIP(ctx, hipMemcpyAsync(s->Y[0], soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, ctx->stream));
    StepArgs a = s->args();
    int64_t nb = nblocks(N);
    k_fill_f64<<<(unsigned)((D * M * N + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(s->grads, 0.0, (int64_t)D * M * N);
    k_fill_u8<<<(unsigned)((N + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>((uint8_t*)s->orders, 2, N);
    HCC_TRY_HIP(ctx, hipEventRecord(s->ev0, ctx->stream));
    ops.gather(s, a, s->Y[0], 1);
    ops.flux(s, a, s->Y[0], s->K[0], 1);
    HCC_TRY_HIP(ctx, hipEventRecord(s->ev1, ctx->stream));
    HCC_TRY_HIP(ctx, hipGetLastError());
    (void)nb;
    HCC_TRY_HIP(ctx, hipMemcpyAsync(soa.data(), s->K[0], sizeof(double) * soa.size(), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    to_aos(soa, K, N, M);
    if (gradients) {
        std::vector<double> gs((size_t)(D * M * N));
        HCC_TRY_HIP(ctx, hipMemcpy(gs.data(), s->grads, sizeof(double) * gs.size(), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) for (int q = 0; q < D; ++q) for (int m = 0; m < M; ++m)
            gradients[(i * D + q) * M + m] = gs[(size_t)((q * M + m) * N + i)];
    }
    float ms = 0;
    HCC_TRY_HIP(ctx, hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_rhs_ms = ms;
    return HCC_OK;
}

extern "C" int hcc_last_step_timing(hcc_stepper* s, double* ms_total, double* ms_gather, double* ms_flux) {
    if (!s) return HCC_E_ARG;
    if (ms_total) *ms_total = s->last_ms;
    if (ms_gather) *ms_gather = s->last_gather_ms;
    if (ms_flux) *ms_flux = s->last_flux_ms;
    return HCC_OK;
}
