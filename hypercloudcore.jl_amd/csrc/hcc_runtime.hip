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
#include "hcc_device.h"
#include <rocprim/device/device_scan.hpp>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <memory>
#include <new>

using namespace hcc;

thread_local std::string hcc::g_create_err;

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
    double* sp[3];                      // positions in cell order: sp[d][s] = p[d][cell_particles[s]]
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

// The bins' positions in cell order, so that the candidate loops read each bin's positions contiguously
template <int D>
__global__ __launch_bounds__(kBlock) void k_sort_pos(SearchArgs a) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.N) return;
    const int32_t j = a.cell_particles[s];
    for (int d = 0; d < D; ++d) a.sp[d][s] = a.p[d][j];
}

// get_distance (ParticleGrids.jl:209-219) + sum(abs2, .) from particle i to the candidate at cell-order slot cs
// (the same operands as p[j] - p[i] with j = cell_particles[cs])
template <int D>
__device__ __forceinline__ double search_dist_cs(const SearchArgs& a, const double (&pi)[D], int64_t cs, double (&dd)[D]) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        double x = a.sp[d][cs] - pi[d];
        x = x - a.Lw[d] * rint(x * a.invLw[d]);
        dd[d] = x;
        s = (d == 0) ? x * x : s + x * x;
    }
    return s;
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
    double pi[D];
    for (int d = 0; d < D; ++d) pi[d] = a.p[d][i];
    int32_t c = 0;
    for (int w = 0; w < NW; ++w) {
        int32_t b = win[w];
        for (int32_t s = a.cell_start[b]; s < a.cell_start[b + 1]; ++s) {
            if (a.cell_particles[s] == i) continue;
            double dd[D];
            if (search_dist_cs<D>(a, pi, s, dd) <= a.R_sq) ++c;
        }
    }
    a.cnt[i] = c;
}

// widest node of each chunk of 64 (one wave per chunk, grid-stride); the sum and the maximum of the counts into
// stats[0..1], one atomic pair per workgroup
constexpr int kWidthBlocks = 512;
__global__ __launch_bounds__(kBlock) void k_chunk_width(const int32_t* cnt, int64_t N, int64_t nchunks, int64_t* width,
                                                        unsigned long long* stats) {
    __shared__ unsigned long long red[2][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long sum = 0, mx = 0;
    for (int64_t c = (int64_t)blockIdx.x * (kBlock / 64) + wv; c < nchunks; c += (int64_t)gridDim.x * (kBlock / 64)) {
        const int64_t t = c * kChunk + lane;
        int32_t v = (t < N) ? cnt[t] : 0;
        sum += (unsigned long long)v;
        for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
        if (lane == 0) width[c] = (int64_t)v * kChunk;
        mx = max(mx, (unsigned long long)v);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lane == 0) { red[0][wv] = sum; red[1][wv] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) { sum += red[0][w]; mx = max(mx, red[1][w]); }
        if (sum) atomicAdd(&stats[0], sum);
        if (mx) atomicMax(&stats[1], mx);
    }
}

__global__ void k_set_i32(int32_t* p, int32_t v) { if (threadIdx.x == 0) *p = v; }

// _write_neighbours! (NeighborLogic.jl:309-329), into the SELL-64 layout
template <int D>
__global__ __launch_bounds__(kBlock) void k_write(SearchArgs a, int64_t nblk) {
    constexpr int NW = D == 1 ? 3 : D == 2 ? 9 : 27;
    int64_t i = node_index(nblk);
    if (i >= a.N) return;
    const int32_t* win = a.window + (int64_t)a.bin_of[i] * NW;
    int64_t base = a.chunk_off[i / kChunk] + (i % kChunk);
    double pi[D];
    for (int d = 0; d < D; ++d) pi[d] = a.p[d][i];
    int32_t k = 0;
    for (int w = 0; w < NW; ++w) {
        int32_t b = win[w];
        for (int32_t s = a.cell_start[b]; s < a.cell_start[b + 1]; ++s) {
            int32_t j = a.cell_particles[s];
            if (j == i) continue;
            double dd[D];
            if (search_dist_cs<D>(a, pi, s, dd) <= a.R_sq) { a.nbr[base + (int64_t)k * kChunk] = j; ++k; }
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

// rocPRIM's device-wide exclusive sum of integer counts (exact, so any association gives the same result), queued on
// the context's stream; the temporary storage is the grid's, grown when a scan needs more
template <class T>
static hipError_t exclusive_scan(hcc_grid* g, const T* in, T* out, int64_t n) {
    hipStream_t st = g->ctx->stream;
    size_t bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st);
    if (e != hipSuccess) return e;
    if (bytes > g->scan_tmp_bytes) {
        if (g->d_scan_tmp) {
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            (void)hipFree(g->d_scan_tmp);
            g->d_scan_tmp = nullptr;
            g->scan_tmp_bytes = 0;
        }
        if ((e = hipMalloc(&g->d_scan_tmp, bytes)) != hipSuccess) return e;
        g->scan_tmp_bytes = bytes;
    }
    return rocprim::exclusive_scan(g->d_scan_tmp, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), st);
}

hcc::GeomArgs hcc_grid::args() const {
    GeomArgs a{};
    a.N = N;
    a.px = d_pos[0]; a.py = d_pos[1] ? d_pos[1] : d_pos[0]; a.pz = d_pos[2] ? d_pos[2] : d_pos[0];
    a.isb = d_isb; a.cnt = d_cnt; a.chunk_off = d_chunk_off; a.nbr = d_nbr;
    for (int d = 0; d < 3; ++d) { a.Lw[d] = geom.Lw[d]; a.invLw[d] = geom.invLw[d]; }
    a.invL = geom.invL; a.alpha = geom.desc.alpha; a.inv_range_sq = geom.inv_range_sq;
    a.R_cut = INFINITY;
    a.halo_off = d_halo_off; a.halo_idx = d_halo_idx; a.nbr_local = d_nbr_local; a.hmax = (int)max_halo;
    a.bslots = (int)max_block_slots;
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
    for (int d = 0; d < 3; ++d) a.sp[d] = g->d_spos[d];
    a.cnt = g->d_cnt; a.chunk_off = g->d_chunk_off; a.nbr = g->d_nbr;
    return a;
}

template <int D>
static int run_search(hcc_grid* g) {
    hcc_ctx* ctx = g->ctx;
    hipStream_t st = ctx->stream;
    const int64_t N = g->N, nblk = nblocks(N);
    for (int d = 0; d < D; ++d)
        if (!g->d_spos[d]) HCC_TRY_HIP(ctx, dalloc(&g->d_spos[d], N));
    SearchArgs a = search_args(g);
    if (N > 0) {
        k_wrap<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_cell_cursor, 0, sizeof(int32_t) * (g->nbins + 1), st));
        k_bin_count<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    // segment starts: exclusive prefix sum of the counts; the dummy bin nbins is empty
    HCC_TRY_HIP(ctx, exclusive_scan<int32_t>(g, g->d_cell_cursor, g->d_cell_start, g->nbins + 1));
    k_set_i32<<<1, 64, 0, st>>>(g->d_cell_start + g->nbins + 1, (int32_t)N);
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_cell_cursor, 0, sizeof(int32_t) * (g->nbins + 1), st));
    if (N > 0) {
        k_bin_scatter<<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        k_bin_sort<<<(unsigned)((g->nbins + kBlock - 1) / kBlock), kBlock, 0, st>>>(a, g->nbins);
        k_sort_pos<D><<<(unsigned)nblk, kBlock, 0, st>>>(a);
        k_count<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    // chunk widths -> SELL offsets (N, and with it the chunk count, is fixed for the grid's life)
    g->nchunks = (N + kChunk - 1) / kChunk;
    if (!g->d_width) HCC_TRY_HIP(ctx, dalloc(&g->d_width, g->nchunks + 1));
    if (!g->d_chunk_off) HCC_TRY_HIP(ctx, dalloc(&g->d_chunk_off, g->nchunks + 1));
    if (!g->d_stats) HCC_TRY_HIP(ctx, dalloc(&g->d_stats, 3));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_width, 0, sizeof(int64_t) * (g->nchunks + 1), st));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_stats, 0, sizeof(unsigned long long) * 3, st));
    if (N > 0)
        k_chunk_width<<<(unsigned)std::min<int64_t>(kWidthBlocks, (g->nchunks + kBlock / 64 - 1) / (kBlock / 64)), kBlock, 0, st>>>(
            g->d_cnt, N, g->nchunks, g->d_width, g->d_stats);
    HCC_TRY_HIP(ctx, exclusive_scan<int64_t>(g, g->d_width, g->d_chunk_off, g->nchunks + 1));
    HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_stats + 2, g->d_chunk_off + g->nchunks, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    unsigned long long stats[3];
    HCC_TRY_HIP(ctx, hipMemcpyAsync(stats, g->d_stats, sizeof(stats), hipMemcpyDeviceToHost, st));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    g->S = (int64_t)stats[0]; g->max_nb = (int64_t)stats[1]; g->S_sell = (int64_t)stats[2];
    if (g->S_sell > g->nbr_cap || !g->d_nbr) {   // a research grows the capacity by a quarter at once
        if (g->d_nbr) (void)hipFree(g->d_nbr);
        g->d_nbr = nullptr;
        const int64_t cap = g->nbr_cap ? g->S_sell + g->S_sell / 4 : g->S_sell;
        g->nbr_cap = 0;
        HCC_TRY_HIP(ctx, dalloc(&g->d_nbr, cap));
        g->nbr_cap = cap;
    }
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_nbr, 0, sizeof(int32_t) * std::max<int64_t>(g->S_sell, 1), st));
    a = search_args(g);
    if (N > 0) k_write<D><<<grid_for(nblk), kBlock, 0, st>>>(a, nblk);
    HCC_TRY_HIP(ctx, hipGetLastError());
    return HCC_OK;
}

// ============================================================================================================
// particle volumes: update_volumes!(pg, ShepardVolumes(sweeps)) (src/Grid/Volumes.jl:183-231), which the ParticleGrid
// constructor runs with one sweep (src/Grid/ParticleGrids.jl:275). The slot weights are recomputed from the positions
// with the search's operations (the same bits as the stored weights).
// ============================================================================================================
template <int D>
__global__ __launch_bounds__(kBlock) void k_shepard_sums(GeomArgs g, double* S, double* V, double w0, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= g.N) return;
    const NodePos<D> pi = load_pos<D>(g, i);
    const int cnt = g.cnt[i];
    const int64_t base = sell_base(g, i);
    double sv = w0;                                   // _shepard_sums!: w(0) + sum_k w_k
    for (int k = 0; k < cnt; ++k) {
        double dd[D];
        const double d2 = slot_geom<D>(g, pi, g.nbr[base + (int64_t)k * kChunk], dd);
        sv = sv + mls_weight(g.alpha, g.inv_range_sq, d2);
    }
    S[i] = sv;
    V[i] = 1.0 / sv;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_shepard_sweep(GeomArgs g, const double* __restrict__ S,
                                                          const double* __restrict__ V, double* __restrict__ Vt,
                                                          double w0, int64_t nblk) {
    const int64_t i = node_index(nblk);
    if (i >= g.N) return;
    const NodePos<D> pi = load_pos<D>(g, i);
    const int cnt = g.cnt[i];
    const int64_t base = sell_base(g, i);
    double acc = w0 / S[i] * V[i];                    // _shepard_sweep!
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = g.nbr[base + (int64_t)k * kChunk];
        double dd[D];
        const double d2 = slot_geom<D>(g, pi, j, dd);
        acc = acc + mls_weight(g.alpha, g.inv_range_sq, d2) / S[j] * V[j];
    }
    Vt[i] = acc;
}

__global__ __launch_bounds__(kBlock) void k_clamp_volumes(double* V, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) V[i] = jmax(V[i], 0.0);
}

__global__ __launch_bounds__(kBlock) void k_scale_volumes(double* V, int64_t n, double dv, const double* total) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const double f = dv / *total;                     // V .*= domain_volume(pg) / sum(V)
    if (i < n) V[i] = V[i] * f;
}

// The structure of Julia's sum of an array (mapreduce over IndexLinear, Base reduce.jl): fewer than 16 elements fold
// left from a1 + a2; otherwise ranges [lo, hi] are halved at lo + (hi - lo) / 2 until hi - lo < 1024, each leaf folds
// left from its first two elements, and the halves are added. The plan (built once per node count): the leaves left to
// right, then the internal nodes as (node, left, right) triples grouped by height, so that one workgroup adds a whole
// height at a time. Each leaf is folded by one workgroup from LDS. (Julia's leaf loops are @simd, which LLVM may
// reassociate; these are sequential.)
namespace {
struct PwLevels {
    int32_t n = 0;
    int32_t off[40] = {};     // height h's triples at [off[h], off[h + 1]); heights < 40 for any int64 node count
};
struct PwInner { int64_t left, right; int height; };
// node ids: leaf k -> k; internal node k -> -(k + 1) until the leaf count is known
int64_t pw_build(int64_t lo, int64_t hi, std::vector<int64_t>& leaves, std::vector<PwInner>& inner, int& h) {
    if (hi - lo < 1024) {
        leaves.push_back(lo); leaves.push_back(hi);
        h = 0;
        return (int64_t)leaves.size() / 2 - 1;
    }
    const int64_t mid = lo + ((hi - lo) >> 1);
    int hl = 0, hr = 0;
    const int64_t l = pw_build(lo, mid, leaves, inner, hl);
    const int64_t r = pw_build(mid + 1, hi, leaves, inner, hr);
    h = std::max(hl, hr) + 1;
    inner.push_back(PwInner{l, r, h});
    return -(int64_t)inner.size();
}
}  // namespace

__global__ __launch_bounds__(64) void k_leaf_sums(const double* __restrict__ v, const int64_t* __restrict__ leaves,
                                                  double* __restrict__ node) {
    __shared__ double x[1024];                        // a leaf holds at most 1024 elements
    const int64_t lo = leaves[2 * (int64_t)blockIdx.x], hi = leaves[2 * (int64_t)blockIdx.x + 1];
    const int n = (int)(hi - lo + 1);
    for (int k = threadIdx.x; k < n; k += 64) x[k] = v[lo + k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s = x[0];
    if (n > 1) {
        s = x[0] + x[1];
        for (int k = 2; k < n; ++k) s = s + x[k];
    }
    node[blockIdx.x] = s;
}
__global__ __launch_bounds__(1024) void k_pairwise_tree(double* node, const int64_t* __restrict__ tri, PwLevels lv) {
    for (int h = 0; h < lv.n; ++h) {
        for (int64_t t = lv.off[h] + (int64_t)threadIdx.x; t < lv.off[h + 1]; t += 1024)
            node[tri[3 * t]] = node[tri[3 * t + 1]] + node[tri[3 * t + 2]];
        __syncthreads();
    }
}

// The plan for g->N nodes into g->d_vplan (leaf bounds, then the triples) and g->vlevel, once per grid.
static int volume_plan(hcc_grid* g) {
    if (g->vplan_n == g->N) return HCC_OK;
    hcc_ctx* ctx = g->ctx;
    std::vector<int64_t> leaves;
    std::vector<PwInner> inner;
    int h = 0;
    if (g->N < 16) { leaves.push_back(0); leaves.push_back(g->N - 1); }
    else pw_build(0, g->N - 1, leaves, inner, h);
    const int64_t nleaf = (int64_t)leaves.size() / 2, nin = (int64_t)inner.size();
    if (nleaf + nin > INT32_MAX) return set_err(ctx, HCC_E_ARG, "volumes: too many nodes");
    auto id = [&](int64_t c) { return c >= 0 ? c : nleaf + (-c - 1); };
    const int H = nin ? inner.back().height : 0;   // the root is built last
    std::vector<int64_t> cnt((size_t)H + 2, 0), plan((size_t)(2 * nleaf + 3 * nin));
    for (const PwInner& q : inner) ++cnt[(size_t)q.height];
    g->vlevel.assign((size_t)H + 1, 0);
    for (int k = 1; k <= H; ++k) g->vlevel[(size_t)k] = g->vlevel[(size_t)k - 1] + (int32_t)cnt[(size_t)k];
    std::copy(leaves.begin(), leaves.end(), plan.begin());
    std::vector<int64_t> fill(g->vlevel.begin(), g->vlevel.end());
    for (int64_t k = 0; k < nin; ++k) {
        const PwInner& q = inner[(size_t)k];
        const int64_t t = fill[(size_t)q.height - 1]++;
        int64_t* o = plan.data() + 2 * nleaf + 3 * t;
        o[0] = nleaf + k; o[1] = id(q.left); o[2] = id(q.right);
    }
    (void)hipFree(g->d_vplan); g->d_vplan = nullptr;
    (void)hipFree(g->d_vnode); g->d_vnode = nullptr;
    g->vplan_n = -1;
    HCC_TRY_HIP(ctx, dalloc(&g->d_vplan, plan.size()));
    HCC_TRY_HIP(ctx, dalloc(&g->d_vnode, nleaf + nin));
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_vplan, plan.data(), sizeof(int64_t) * plan.size(), hipMemcpyHostToDevice));
    g->vleaf = nleaf;
    g->vplan_n = g->N;
    return HCC_OK;
}

// The volumes into g->d_vol: sums and 1 / S, `sweeps` sweeps, clamp at zero, and the renormalisation to the domain's
// volume (prod(maxs - mins), Volumes.jl:86) with the pairwise sum of the volumes (its leaves sequential, Julia's
// @simd: that one sum agrees with Julia's to rounding only). Queued on the context's stream.
static int compute_volumes(hcc_grid* g, int sweeps) {
    hcc_ctx* ctx = g->ctx;
    const int64_t N = g->N;
    if (N == 0) return HCC_OK;
    hipStream_t st = ctx->stream;
    if (!g->d_vol) HCC_TRY_HIP(ctx, dalloc(&g->d_vol, N));
    if (!g->d_vS) HCC_TRY_HIP(ctx, dalloc(&g->d_vS, N));
    if (!g->d_vT) HCC_TRY_HIP(ctx, dalloc(&g->d_vT, N));
    if (int rc = volume_plan(g)) return rc;
    const GeomArgs a = g->args();
    const double w0 = mls_weight(a.alpha, a.inv_range_sq, 0.0);
    const int64_t nb = nblocks(N);
    const int D = g->geom.D;
    double* S = g->d_vS;
    if (D == 1) k_shepard_sums<1><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, w0, nb);
    else if (D == 2) k_shepard_sums<2><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, w0, nb);
    else k_shepard_sums<3><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, w0, nb);
    for (int sw = 0; sw < sweeps; ++sw) {
        if (D == 1) k_shepard_sweep<1><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, g->d_vT, w0, nb);
        else if (D == 2) k_shepard_sweep<2><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, g->d_vT, w0, nb);
        else k_shepard_sweep<3><<<grid_for(nb), kBlock, 0, st>>>(a, S, g->d_vol, g->d_vT, w0, nb);
        std::swap(g->d_vol, g->d_vT);
    }
    k_clamp_volumes<<<(unsigned)nb, kBlock, 0, st>>>(g->d_vol, N);
    k_leaf_sums<<<(unsigned)g->vleaf, 64, 0, st>>>(g->d_vol, g->d_vplan, g->d_vnode);
    const int64_t nin = g->vleaf - 1;
    if (nin > 0) {
        PwLevels lv;
        lv.n = (int32_t)g->vlevel.size() - 1;
        for (size_t k = 0; k < g->vlevel.size(); ++k) lv.off[k] = g->vlevel[k];
        k_pairwise_tree<<<1, 1024, 0, st>>>(g->d_vnode, g->d_vplan + 2 * g->vleaf, lv);
    }
    double dv = g->geom.L[0];
    for (int d = 1; d < D; ++d) dv = dv * g->geom.L[d];
    k_scale_volumes<<<(unsigned)nb, kBlock, 0, st>>>(g->d_vol, N, dv, g->d_vnode + (g->vleaf + nin - 1));
    HCC_TRY_HIP(ctx, hipGetLastError());
    return HCC_OK;
}

extern "C" int hcc_grid_volumes(hcc_grid* g, int32_t sweeps, double* volumes) {
    if (!g) return HCC_E_ARG;
    hcc_ctx* ctx = g->ctx;
    if (sweeps >= 0 && g->n_inst > 1)
        return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_grid_volumes: an ensemble keeps its instances' volumes (recompute them on the parts)");
    if (sweeps >= 0) {
        HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
        if (int rc = compute_volumes(g, sweeps)) return rc;
    }
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (volumes && g->N) HCC_TRY_HIP(ctx, hipMemcpy(volumes, g->d_vol, sizeof(double) * g->N, hipMemcpyDeviceToHost));
    return HCC_OK;
}

int hcc::grid_volumes(hcc_grid* g, int sweeps) { return compute_volumes(g, sweeps); }

int hcc::research(hcc_grid* g) {
    hcc_ctx* ctx = g->ctx;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (g->d_halo_off) { (void)hipFree(g->d_halo_off); g->d_halo_off = nullptr; }
    if (g->d_halo_idx) { (void)hipFree(g->d_halo_idx); g->d_halo_idx = nullptr; }
    if (g->d_nbr_local) { (void)hipFree(g->d_nbr_local); g->d_nbr_local = nullptr; }
    g->max_halo = 0; g->max_block_slots = 0;
    const int D = g->geom.D;
    return D == 1 ? run_search<1>(g) : D == 2 ? run_search<2>(g) : run_search<3>(g);
}

// ============================================================================================================
// ensembles: one grid assembled from instance grids, each padded to whole kStage blocks
// ============================================================================================================
__global__ __launch_bounds__(kBlock) void k_copy_offset(int32_t* __restrict__ dst, const int32_t* __restrict__ src,
                                                        int64_t n, int32_t add) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t < n) dst[t] = src[t] + add;
}

extern "C" int hcc_grid_concat(hcc_ctx* ctx, int32_t n, hcc_grid* const* parts, hcc_grid** out, int64_t* offsets) {
    if (!ctx || !out || !parts || n < 1 || n > 65535) return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: bad arguments");
    *out = nullptr;
    const hcc_grid* p0 = parts[0];
    for (int k = 0; k < n; ++k) {
        const hcc_grid* p = parts[k];
        if (!p || p->ctx != ctx) return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: a part is NULL or of another context");
        if (p->n_inst != 1) return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: a part is itself an ensemble");
        if (p->N == 0) return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: an empty part");
        bool same = p->geom.D == p0->geom.D && p->geom.R_sq == p0->geom.R_sq && p->geom.inv_range_sq == p0->geom.inv_range_sq &&
                    p->geom.desc.alpha == p0->geom.desc.alpha && p->geom.invL == p0->geom.invL;
        for (int d = 0; d < 3; ++d) same = same && p->geom.Lw[d] == p0->geom.Lw[d] && p->geom.invLw[d] == p0->geom.invLw[d];
        if (!same)
            return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: the parts differ in D, R, alpha, minimum spacing or wrap lengths");
    }
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    std::unique_ptr<hcc_grid, int (*)(hcc_grid*)> g(new hcc_grid(), hcc_grid_destroy);
    g->ctx = ctx;
    g->geom = p0->geom;
    const int D = p0->geom.D;
    bool halos = D == 2;
    int64_t N = 0, S_sell = 0, nst = 0;
    g->n_inst = n;
    g->inst_off.resize((size_t)n + 1);
    g->inst_n.resize((size_t)n);
    std::vector<int64_t> blk_off((size_t)n + 1), s_off((size_t)n + 1), h_off((size_t)n + 1);
    for (int k = 0; k < n; ++k) {
        const hcc_grid* p = parts[k];
        g->inst_off[k] = N; g->inst_n[k] = p->N; blk_off[k] = nst; s_off[k] = S_sell; h_off[k] = 0;
        const int64_t nb = (p->N + kStage - 1) / kStage;
        N += nb * kStage; nst += nb; S_sell += p->S_sell;
        g->S += p->S; g->max_nb = std::max(g->max_nb, p->max_nb);
        halos = halos && p->max_halo > 0 && p->N > 0;
    }
    g->inst_off[n] = N; blk_off[n] = nst; s_off[n] = S_sell;
    if (N > (int64_t)INT32_MAX - 1) return set_err(ctx, HCC_E_ARG, "hcc_grid_concat: more than 2^31 - 2 nodes");
    g->N = N; g->S_sell = S_sell; g->nchunks = N / kChunk;
    for (int d = 0; d < D; ++d) {
        HCC_TRY_HIP(ctx, dalloc(&g->d_pos[d], N));
        HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_pos[d], 0, sizeof(double) * std::max<int64_t>(N, 1), st));
    }
    HCC_TRY_HIP(ctx, dalloc(&g->d_isb, N));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_isb, 1, std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, dalloc(&g->d_tags, N));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_tags, 0, sizeof(int32_t) * std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, dalloc(&g->d_cnt, N));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_cnt, 0, sizeof(int32_t) * std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, dalloc(&g->d_vol, N));
    HCC_TRY_HIP(ctx, hipMemsetAsync(g->d_vol, 0, sizeof(double) * std::max<int64_t>(N, 1), st));
    HCC_TRY_HIP(ctx, dalloc(&g->d_nbr, S_sell));
    g->nbr_cap = S_sell;
    HCC_TRY_HIP(ctx, dalloc(&g->d_chunk_off, g->nchunks + 1));
    g->h_isb.assign((size_t)N, 1);
    g->h_tags.assign((size_t)N, 0);
    std::vector<int64_t> coff((size_t)g->nchunks + 1);
    std::vector<int32_t> hoff;
    if (halos) { HCC_TRY_HIP(ctx, dalloc(&g->d_nbr_local, S_sell)); hoff.assign((size_t)nst + 1, 0); }
    int64_t H = 0;
    std::vector<int64_t> h_tot((size_t)n + 1, 0);
    for (int k = 0; k < n; ++k) {   // halo sizes first: the index array's length
        const hcc_grid* p = parts[k];
        if (!halos) break;
        int32_t last = 0;
        const int64_t nb = blk_off[k + 1] - blk_off[k];
        HCC_TRY_HIP(ctx, hipMemcpy(&last, p->d_halo_off + nb, sizeof(int32_t), hipMemcpyDeviceToHost));
        h_tot[k] = H;
        H += last;
    }
    if (halos) HCC_TRY_HIP(ctx, dalloc(&g->d_halo_idx, std::max<int64_t>(H, 1)));
    std::vector<uint16_t> blk_inst((size_t)std::max<int64_t>(nst, 1));
    for (int k = 0; k < n; ++k) {
        const hcc_grid* p = parts[k];
        const int64_t o = g->inst_off[k], Nk = p->N;
        for (int64_t b = blk_off[k]; b < blk_off[k + 1]; ++b) blk_inst[(size_t)b] = (uint16_t)k;
        for (int d = 0; d < D; ++d)
            HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_pos[d] + o, p->d_pos[d], sizeof(double) * Nk, hipMemcpyDeviceToDevice, st));
        HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_isb + o, p->d_isb, Nk, hipMemcpyDeviceToDevice, st));
        HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_tags + o, p->d_tags, sizeof(int32_t) * Nk, hipMemcpyDeviceToDevice, st));
        HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_cnt + o, p->d_cnt, sizeof(int32_t) * Nk, hipMemcpyDeviceToDevice, st));
        HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_vol + o, p->d_vol, sizeof(double) * Nk, hipMemcpyDeviceToDevice, st));
        std::copy(p->h_isb.begin(), p->h_isb.end(), g->h_isb.begin() + o);
        std::copy(p->h_tags.begin(), p->h_tags.end(), g->h_tags.begin() + o);
        // SELL: the part's chunks follow at chunk o / 64 (a multiple of 4); its padding chunks hold no slots
        std::vector<int64_t> pc((size_t)p->nchunks + 1);
        HCC_TRY_HIP(ctx, hipMemcpy(pc.data(), p->d_chunk_off, sizeof(int64_t) * pc.size(), hipMemcpyDeviceToHost));
        const int64_t c0 = o / kChunk, nck = (blk_off[k + 1] - blk_off[k]) * (kStage / kChunk);
        for (int64_t c = 0; c <= nck; ++c) coff[(size_t)(c0 + c)] = s_off[k] + pc[(size_t)std::min<int64_t>(c, p->nchunks)];
        if (p->S_sell) {
            k_copy_offset<<<(unsigned)((p->S_sell + kBlock - 1) / kBlock), kBlock, 0, st>>>(g->d_nbr + s_off[k], p->d_nbr,
                                                                                          p->S_sell, (int32_t)o);
            if (halos)
                HCC_TRY_HIP(ctx, hipMemcpyAsync(g->d_nbr_local + s_off[k], p->d_nbr_local, sizeof(uint16_t) * p->S_sell,
                                                hipMemcpyDeviceToDevice, st));
        }
        if (halos) {   // block halos: node ids shifted by the instance offset, offsets by the halos before
            const int64_t nb = blk_off[k + 1] - blk_off[k];
            std::vector<int32_t> ph((size_t)nb + 1);
            HCC_TRY_HIP(ctx, hipMemcpy(ph.data(), p->d_halo_off, sizeof(int32_t) * ph.size(), hipMemcpyDeviceToHost));
            for (int64_t b = 0; b <= nb; ++b) hoff[(size_t)(blk_off[k] + b)] = (int32_t)(h_tot[k] + ph[(size_t)b]);
            if (ph[(size_t)nb])
                k_copy_offset<<<(unsigned)((ph[(size_t)nb] + kBlock - 1) / kBlock), kBlock, 0, st>>>(
                    g->d_halo_idx + h_tot[k], p->d_halo_idx, ph[(size_t)nb], (int32_t)o);
            g->max_halo = std::max(g->max_halo, p->max_halo);
            g->max_block_slots = std::max(g->max_block_slots, p->max_block_slots);
        }
        HCC_TRY_HIP(ctx, hipGetLastError());
    }
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_chunk_off, coff.data(), sizeof(int64_t) * coff.size(), hipMemcpyHostToDevice));
    if (halos) {
        HCC_TRY_HIP(ctx, dalloc(&g->d_halo_off, hoff.size()));
        HCC_TRY_HIP(ctx, hipMemcpy(g->d_halo_off, hoff.data(), sizeof(int32_t) * hoff.size(), hipMemcpyHostToDevice));
    } else {
        g->max_halo = 0;
        g->max_block_slots = 0;
    }
    HCC_TRY_HIP(ctx, dalloc(&g->d_blk_inst, blk_inst.size()));
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_blk_inst, blk_inst.data(), sizeof(uint16_t) * blk_inst.size(), hipMemcpyHostToDevice));
    HCC_TRY_HIP(ctx, dalloc(&g->d_inst_blk, blk_off.size()));
    HCC_TRY_HIP(ctx, hipMemcpy(g->d_inst_blk, blk_off.data(), sizeof(int64_t) * blk_off.size(), hipMemcpyHostToDevice));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (offsets) std::copy(g->inst_off.begin(), g->inst_off.end(), offsets);
    *out = g.release();
    return HCC_OK;
}

extern "C" int hcc_grid_destroy(hcc_grid* g) {
    if (!g) return HCC_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    for (int d = 0; d < 3; ++d) (void)hipFree(g->d_pos[d]);
    (void)hipFree(g->d_isb); (void)hipFree(g->d_tags); (void)hipFree(g->d_bin_of); (void)hipFree(g->d_cell_start); (void)hipFree(g->d_cell_cursor);
    (void)hipFree(g->d_cell_particles); (void)hipFree(g->d_bin_window); (void)hipFree(g->d_cnt); (void)hipFree(g->d_chunk_off); (void)hipFree(g->d_nbr);
    (void)hipFree(g->d_halo_off); (void)hipFree(g->d_halo_idx); (void)hipFree(g->d_nbr_local);
    (void)hipFree(g->d_vol);
    (void)hipFree(g->d_blk_inst); (void)hipFree(g->d_inst_blk);
    (void)hipFree(g->d_scan_tmp); (void)hipFree(g->d_width); (void)hipFree(g->d_stats);
    for (int d = 0; d < 3; ++d) (void)hipFree(g->d_spos[d]);
    (void)hipFree(g->d_vS); (void)hipFree(g->d_vT); (void)hipFree(g->d_vnode); (void)hipFree(g->d_vplan);
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
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (D == 2 && N > 0) {   // LDS staging tables of the 2-D passes
        std::vector<int32_t> cnt((size_t)N), nbr((size_t)std::max<int64_t>(g->S_sell, 1)), hoff, hidx;
        std::vector<int64_t> off((size_t)(g->nchunks + 1));
        std::vector<uint16_t> nl;
        HCC_TRY_HIP(ctx, hipMemcpy(cnt.data(), g->d_cnt, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
        HCC_TRY_HIP(ctx, hipMemcpy(off.data(), g->d_chunk_off, sizeof(int64_t) * (g->nchunks + 1), hipMemcpyDeviceToHost));
        if (g->S_sell) HCC_TRY_HIP(ctx, hipMemcpy(nbr.data(), g->d_nbr, sizeof(int32_t) * g->S_sell, hipMemcpyDeviceToHost));
        if (build_halos(N, kStage, cnt, off, nbr, hoff, hidx, nl, g->max_halo)) {
            HCC_TRY_HIP(ctx, dalloc(&g->d_halo_off, hoff.size()));
            HCC_TRY_HIP(ctx, dalloc(&g->d_halo_idx, hidx.size()));
            HCC_TRY_HIP(ctx, dalloc(&g->d_nbr_local, nl.size()));
            HCC_TRY_HIP(ctx, hipMemcpy(g->d_halo_off, hoff.data(), sizeof(int32_t) * hoff.size(), hipMemcpyHostToDevice));
            HCC_TRY_HIP(ctx, hipMemcpy(g->d_halo_idx, hidx.data(), sizeof(int32_t) * hidx.size(), hipMemcpyHostToDevice));
            HCC_TRY_HIP(ctx, hipMemcpy(g->d_nbr_local, nl.data(), sizeof(uint16_t) * nl.size(), hipMemcpyHostToDevice));
            const int64_t cpb = kStage / kChunk, nst = (N + kStage - 1) / kStage;   // chunks per staged block
            for (int64_t b = 0; b < nst; ++b)
                g->max_block_slots = std::max<int64_t>(g->max_block_slots,
                                                       off[(size_t)std::min<int64_t>((b + 1) * cpb, g->nchunks)] - off[(size_t)(b * cpb)]);
        } else {
            g->max_halo = 0;   // a halo too large for 16-bit local indices: the passes read neighbours from HBM
        }
    }
    rc = compute_volumes(g.get(), 1);   // the constructor's ShepardVolumes() (ParticleGrids.jl:275)
    if (rc) return rc;
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
int hcc::download_lists(hcc_grid* g, std::vector<int64_t>& rstart, std::vector<int32_t>& idx,
                          std::vector<int64_t>* sell_of_slot) {
    hcc_ctx* ctx = g->ctx;
    const int64_t N = g->N;
    std::vector<int32_t> cnt((size_t)N), nbr((size_t)std::max<int64_t>(g->S_sell, 1));
    std::vector<int64_t> off((size_t)(g->nchunks + 1));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
    if (g->n_inst > 1) return set_err(ctx, HCC_E_UNSUPPORTED, "hcc_grid_get_bins: an ensemble has no cell list (query the parts)");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));   // a moving step's search may still be queued
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
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));   // positions a queued moving step still writes
    const int D = g->geom.D;
    std::vector<double> col((size_t)g->N);
    for (int d = 0; d < D; ++d) {
        if (g->N) HCC_TRY_HIP(ctx, hipMemcpy(col.data(), g->d_pos[d], sizeof(double) * g->N, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < g->N; ++i) pos[i * D + d] = col[i];
    }
    return HCC_OK;
}

