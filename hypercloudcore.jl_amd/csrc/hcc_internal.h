// Internal declarations of libhcc (not part of the C ABI).
#pragma once
#include "../../include/hcc.h"
#include "hcc_math.h"
#include <string>
#include <vector>

struct hcc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    unsigned long long* h_pinned = nullptr;   // pinned host mirror of the device counters
};

namespace hcc {

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed);
    uint64_t next();
    double uniform();
};

// Geometry of a rectangular domain with its ghost canvas (ComputationalDomain) and the derived constants
// the kernels use.
struct DomainGeom {
    hcc_domain_desc desc{};
    int D = 0;
    double irf = 0;                    // interp_range_factor
    double cmin[3], cmax[3];           // canvas
    double L[3], Lw[3], invLw[3];      // physical lengths; wrap lengths (0 on non-periodic axes)
    double invL = 1, R_sq = 0, inv_range_sq = 0;
    void init(const hcc_domain_desc& d);
};

struct HostCloud {
    int64_t N = 0, morton_ties = 0;
    std::vector<double> pos;
    std::vector<uint8_t> isb;
    std::vector<int32_t> tags;
};

void generate_cloud(const DomainGeom& g, HostCloud& out);
void bin_window_table(int D, const int64_t* dims, const int32_t* periodic, std::vector<int32_t>& table);
// Per-block halos for LDS staging: sorted unique neighbours (and own nodes) of each block of `block` nodes, and
// every SELL slot's neighbour as a local index into its block's halo. Returns false if a halo exceeds 65535.
bool build_halos(int64_t N, int block, const std::vector<int32_t>& cnt, const std::vector<int64_t>& chunk_off,
                 const std::vector<int32_t>& nbr, std::vector<int32_t>& halo_off, std::vector<int32_t>& halo_idx,
                 std::vector<uint16_t>& nbr_local, int64_t& max_halo);
void outflow_roots(const DomainGeom& g, int64_t N, const double* pos, const uint8_t* isb, const int32_t* tags,
                   const std::vector<int64_t>& rstart, const std::vector<int32_t>& nbr, int n_bc, const int32_t* bc_tag,
                   const int32_t* bc_kind, std::vector<int32_t>& root);

// ---- device-side argument blocks (passed by value to kernels) -------------------------------------------
constexpr int kChunk = 64;             // SELL chunk = one wavefront of nodes

struct GeomArgs {
    int64_t N;
    const double* px;
    const double* py;
    const double* pz;
    const uint8_t* isb;
    const int32_t* cnt;        // neighbours per node
    const int64_t* chunk_off;  // SELL-64: first slot of each chunk of 64 nodes
    const int32_t* nbr;        // SELL-64 neighbour indices: slot k of node i at chunk_off[i/64] + 64 k + i%64
    double Lw[3], invLw[3];
    double invL, alpha, inv_range_sq;
    double R_cut;               // a slot weighs 0 where |d|^2 > R_cut: R^2 on the stage geometry of a moving grid, whose
                                // lists stay those of the step's search (_refresh_slots!, GridMovers.jl:111-120); +inf
                                // else (every listed pair is within R)
    // LDS staging tables (2-D): the nodes of workgroup b (kStage consecutive nodes) read their neighbours from the
    // block's halo, the sorted union of their neighbourhoods (and themselves): halo_idx[halo_off[b] .. halo_off[b+1])
    const int32_t* halo_off;
    const int32_t* halo_idx;
    const uint16_t* nbr_local;  // SELL-64 like nbr: the neighbour's position in the block's halo
    int hmax;                   // largest halo of any block
    int bslots;                 // largest SELL slot range of any block (its chunks are consecutive)
};
constexpr int kStage = 256;     // nodes per staged workgroup

struct PdeArgs {
    double gamma;
    double c[3][8];
    double amax[3];
};

// Record layout of the step kernels. A state record holds a node's position followed by its state, padded to
// 2, 4 or 8 doubles; a gradient record the D*M limited gradient coefficients (basis term major), padded likewise.
// One neighbour then costs a few 16-byte loads from one or two cache lines.
template <int D, int M>
struct Layout {
    static constexpr int NR = D + M;
    static constexpr int PU = NR <= 2 ? 2 : NR <= 4 ? 4 : 8;
    static constexpr int NG = D * M;
    static constexpr int GS = NG <= 2 ? 2 : NG <= 4 ? 4 : NG <= 8 ? 8 : 16;
};

struct StepArgs {
    GeomArgs g;
    PdeArgs pde;
    int limiter, mode, strategy, criterion;
    double delta;
    double plateau;              // MOODu2: a component whose local range is below plateau x range[m] is not checked
    double ratio;                // MOODu2: ratio of the nearest curvature magnitudes (its override at MUSCL order 3)
    double hmin;                 // minimum(pg.meta.dx) (the curvature override's spacing h)
    const double* range;         // MOODu2: global range of the previous stage per component (nullptr: plateau off)
    int streamed;                // divergence accumulated per slot with the derivative weights (flux variant 1)
    int aligned;                 // AlignedRusanovFlux: sort by the weights' signs below the full order (variant 1)
    int packed;                  // the orders array of a flux pass holds EPD2's minimum | true order << 4
    int src_mode;                // HCC_SRC_*: the decay source S_i(U) = -src_rate[i] U, explicit or implicit
    const double* src_rate;
    const double* invg;          // reconstruction inverse moments, D*D per node: [q * N + i]
    const double* invd;          // divergence inverse moments, derivative rows 1..D: [(r * (D+1) + c) * N + i]
    const uint8_t* okg;
    const uint8_t* okd;
    double* G;                   // gradient records
    int8_t* orders;
    uint8_t* changed;
    uint8_t* dirty;
    uint8_t* trig;               // mood_flags of the step
    uint8_t* inadm;              // last check: candidate not admissible
    unsigned long long* counters;// [0] / [6]: lengths of changed lists 0 / 1 (clist / clist2; entries -1 are no node),
                                 // [1] inadmissible accepted, [2] dt bits, [4] fixed-point iterations of the step,
                                 // [5] / [7]: work-list lengths of iterations with parity 0 / 1, [8] / [9]: changed
                                 // list 0 / 1 holds a node
    int32_t* clist;              // particles whose order changed in a check (changed list 0)
    int32_t* clist2;             // changed list 1
    int32_t* dlist;              // the recompute's work list: the dirty particles, each once
    unsigned* dmark;             // dirty particles of the fixed point (changed plus one hop); set by the mark kernel
                                 // (the setter lists the node), cleared by the recompute
    double* lam;                 // per-particle wave speeds for the CFL estimate, D per node
    double* bmin;                // per-workgroup minima of the CFL estimate
    double* ext;                 // local extrema of the stage state over each neighbourhood: M maxima, M minima
    int ext_on;                  // full gather passes store ext, the MOOD checks read it (MOODu1, MOODu2)
    const double* dtp;           // the step's dt, device-resident (set by the host or by the CFL estimate); one per
                                 // instance on an ensemble grid
    const uint16_t* inst;        // ensemble grid: the instance of each kStage block (nullptr: one instance)
    uint16_t* feat;              // indicator features of the full gathers: 16 bf16 per node
    int32_t* ind_list;           // the indicator's candidates of the last full gather: particles at the full order
    unsigned long long* ind_n;   // whose neighbourhood is not constant (MOOD's DMP test can flag no other), and count
    double hratio;               // MOOD substep: h = hratio * dt
    int ale;                     // HCC_MOVER_*: flux in the frame of the particle's grid velocity w (generic flux loop)
    double mover_alpha;          // ScaledLagrangianMover: w = mover_alpha x fluid velocity
    double* KX;                  // K_X of the right-hand side being computed: w per interior particle, D per node
    uint8_t* cnb;                // full gathers: 1 if every neighbour's state equals the particle's (==) and its wave
                                 // speeds and 2 F are finite (nullptr: not recorded)
    uint8_t* bflat;              // per kStage block: every halo particle has cnb (or is a boundary particle), so every
                                 // slot of the block's flux pass is flat (k_block_flat; nullptr: not used)
    uint8_t* kz;                 // bit s: the particle's K of stage s (0-based) is exactly +-0 in every component (the
                                 // fused predictor's passes write it, ghosts copy their root's; nullptr: not tracked)
    unsigned kzmask;             // the stage bits of kz: (1 << stages) - 1
    const uint8_t* gflat;        // per kStage block: the stage's full gather is flat (k_block_gflat; nullptr: not used)
    const int32_t* border;       // the staged flux pass's and stage-1 gather's blocks in launch order: the order of the
                                 // stage's gather (nullptr: the XCD-aware order of the block indices)
    const int32_t* gorder;       // the staged gathers' launch order after k_block_gflat: non-flat blocks first
    const uint8_t* zfit;         // per particle, geometry only (k_zero_fits): what its fits give for all-zero
                                 // accumulators, so that the flat paths need not read the inverses (kZ* bits)
};
// bits of StepArgs::zfit
constexpr unsigned kZDiv = 1u;    // the divergence fit of zero accumulators is -0 in every component (before sources)
constexpr unsigned kZGrad = 2u;   // the gradient fit of zero accumulators is +-0 in every row: bit kZSign0 + q is
constexpr unsigned kZSign0 = 2;   // row q's sign (the same for every component)
constexpr unsigned kZStarved = 32u;   // fewer than D neighbours

}  // namespace hcc

struct hcc_grid {
    hcc_ctx* ctx = nullptr;
    hcc::DomainGeom geom;
    int64_t N = 0;
    double* d_pos[3] = {nullptr, nullptr, nullptr};
    uint8_t* d_isb = nullptr;
    int32_t* d_tags = nullptr;
    // cell list
    int64_t dims[3] = {1, 1, 1}, nbins = 0;
    double bin_size[3] = {0, 0, 0};
    int32_t* d_bin_of = nullptr;
    int32_t* d_cell_start = nullptr;   // nbins + 2
    int32_t* d_cell_cursor = nullptr;  // nbins + 1
    int32_t* d_cell_particles = nullptr;
    int32_t* d_bin_window = nullptr;   // nbins * 3^D
    double* d_spos[3] = {nullptr, nullptr, nullptr};   // positions in cell order (the search's candidate loops)
    // neighbours (SELL-64)
    int32_t* d_cnt = nullptr;
    int64_t nchunks = 0, S = 0, S_sell = 0, max_nb = 0;
    int64_t* d_chunk_off = nullptr;
    int32_t* d_nbr = nullptr;
    int32_t* d_halo_off = nullptr;
    int32_t* d_halo_idx = nullptr;
    uint16_t* d_nbr_local = nullptr;
    int64_t max_halo = 0, max_block_slots = 0;
    double* d_vol = nullptr;           // particle volumes (ShepardVolumes), N
    // ensemble (hcc_grid_concat): instance k's nodes at [inst_off[k], inst_off[k] + inst_n[k]), padded to whole kStage
    // blocks with inert boundary nodes; the instance of each block, and each instance's block range
    int n_inst = 1;
    std::vector<int64_t> inst_off, inst_n;
    uint16_t* d_blk_inst = nullptr;
    int64_t* d_inst_blk = nullptr;     // n_inst + 1 block offsets
    std::vector<int32_t> h_tags;
    std::vector<uint8_t> h_isb;
    // scratch kept across searches and volume updates (a moving grid repeats both every step): the prefix sums'
    // temporary storage, the chunk widths, the list statistics, the neighbour array's capacity, and the volumes'
    // Shepard sums, sweep buffer and pairwise-sum plan
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int64_t* d_width = nullptr;
    unsigned long long* d_stats = nullptr;   // sum and max of the counts, S_sell
    int64_t nbr_cap = 0;
    double *d_vS = nullptr, *d_vT = nullptr, *d_vnode = nullptr;
    int64_t* d_vplan = nullptr;
    int64_t vplan_n = -1, vleaf = 0;
    std::vector<int32_t> vlevel;             // the combine's level offsets into the node triples
    hcc::GeomArgs args() const;
};

namespace hcc {
// pg.neighbor(pg) on the device after the positions moved: periodic wrap, cell lists and SELL lists rebuilt (the LDS
// halo tables are dropped: the passes then read neighbours from HBM)
int research(hcc_grid* g);
// update_volumes!(pg, ShepardVolumes(sweeps)) into g->d_vol
int grid_volumes(hcc_grid* g, int sweeps);
int download_lists(hcc_grid* g, std::vector<int64_t>& rstart, std::vector<int32_t>& idx, std::vector<int64_t>* sell_of_slot);
}
