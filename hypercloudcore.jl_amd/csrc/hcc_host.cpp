// Host set-up of the MI355X path: point-cloud generation, Morton order, the 3^D bin window table and the
// outflow donor map. These run once per grid on the host, as in the reference (docs/src/gpu_roadmap.md:142-147:
// point generation and the Morton sort are host-only there too).
#include "hcc_internal.h"
#include <algorithm>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace hcc {

// ---- seeded stream of the jitter and of the synthetic initial conditions --------------------------------
// xoshiro256** seeded by splitmix64; doubles in [0,1) from the top 53 bits.
Xoshiro::Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        s[i] = z ^ (z >> 31);
    }
}
uint64_t Xoshiro::next() {
    auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
}
double Xoshiro::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

// Julia eps(x): float spacing at |x|
static double spacing(double x) {
    x = std::fabs(x);
    if (!(x >= kFloatMin)) return 4.9406564584124654e-324;
    return std::ldexp(1.0, std::ilogb(x) - 52);
}

// ComputationalDomain (src/Grid/Domains.jl:159-191): ghost layers per non-periodic axis, canvas, wrap lengths
void DomainGeom::init(const hcc_domain_desc& d) {
    desc = d;
    D = d.D;
    double max_dx = d.dx[0];
    for (int a = 1; a < D; ++a) max_dx = std::max(max_dx, d.dx[a]);
    irf = d.R / max_dx;
    for (int a = 0; a < 3; ++a) { cmin[a] = cmax[a] = L[a] = Lw[a] = invLw[a] = 0.0; }
    for (int a = 0; a < D; ++a) {
        int64_t ng = (d.dx[a] == max_dx) ? (int64_t)std::ceil(irf)
                                          : (int64_t)std::ceil(irf * (max_dx / d.dx[a]) * (1 - 4 * kEps));
        cmin[a] = d.periodic[a] ? d.mins[a] : d.mins[a] - (double)ng * d.dx[a];
        cmax[a] = d.periodic[a] ? d.maxs[a] : d.maxs[a] + (double)ng * d.dx[a];
        L[a] = d.maxs[a] - d.mins[a];
        Lw[a] = d.periodic[a] ? L[a] : 0.0;
        invLw[a] = d.periodic[a] ? 1.0 / L[a] : 0.0;
    }
    double mindx = d.dx[0];
    for (int a = 1; a < D; ++a) mindx = std::min(mindx, d.dx[a]);
    invL = 1.0 / mindx;
    R_sq = d.R * d.R;
    inv_range_sq = 1.0 / (d.R * d.R);
}

// face tag of a point outside the box (Domains.jl:73-94)
static int box_tag(const DomainGeom& g, const double* p) {
    const auto& d = g.desc;
    bool inside = true;
    for (int a = 0; a < g.D; ++a) if (p[a] < d.mins[a] || p[a] > d.maxs[a]) inside = false;
    if (inside) return 0;
    int tag = 0; double out = 0.0;
    for (int a = 0; a < g.D; ++a) {
        if (d.periodic[a]) continue;
        double lo = d.mins[a] - p[a], hi = p[a] - d.maxs[a];
        if (lo > out) { tag = 2 * a + 1; out = lo; }
        if (hi > out) { tag = 2 * a + 2; out = hi; }
    }
    if (tag > 0) return tag;
    tag = 1; double best = INFINITY;
    for (int a = 0; a < g.D; ++a) {
        if (d.periodic[a]) continue;
        double lo = std::fabs(p[a] - d.mins[a]), hi = std::fabs(p[a] - d.maxs[a]);
        if (lo < best) { tag = 2 * a + 1; best = lo; }
        if (hi < best) { tag = 2 * a + 2; best = hi; }
    }
    return tag;
}

static uint32_t spread2(uint32_t w) {
    w &= 0x0000ffffu;
    w = (w | (w << 8)) & 0x00FF00FFu; w = (w | (w << 4)) & 0x0F0F0F0Fu;
    w = (w | (w << 2)) & 0x33333333u; w = (w | (w << 1)) & 0x55555555u;
    return w;
}
static uint32_t spread3(uint32_t w) {
    w &= 0x000003ffu;
    w = (w | (w << 16)) & 0xFF0000FFu; w = (w | (w << 8)) & 0x0300F00Fu;
    w = (w | (w << 4)) & 0x030C30C3u; w = (w | (w << 2)) & 0x09249249u;
    return w;
}

// Lattice + jitter + wrap, interior/ghost split, narrow ghost band, tags (Domains.jl:220-347), then the
// Morton permutation (src/Grid/Reordering.jl:55-101; equal keys keep their original order).
void generate_cloud(const DomainGeom& g, HostCloud& out) {
    const auto& d = g.desc;
    const int D = g.D;
    int64_t Ns[3] = {1, 1, 1}, wlo[3] = {-1, -1, -1}, whi[3] = {-1, -1, -1};
    for (int a = 0; a < D; ++a) {
        int64_t n = std::max<int64_t>(1, (int64_t)std::nearbyint((g.cmax[a] - g.cmin[a]) / d.dx[a]));
        Ns[a] = d.periodic[a] ? n : n + 1;
        if (!d.periodic[a]) {
            // _wall_index (Domains.jl:202-207): the lattice point that sits on the wall
            for (int side = 0; side < 2; ++side) {
                double w = side ? d.maxs[a] : d.mins[a];
                double r = (w - g.cmin[a]) / d.dx[a];
                int64_t k = (int64_t)std::nearbyint(r);
                double tol = (2 * (spacing(w) + spacing(g.cmin[a])) + kEps * std::fabs(w - g.cmin[a])) / d.dx[a] +
                             4 * kEps * std::fabs(r);
                int64_t idx = std::fabs(r - (double)k) <= std::min(0.25, tol) ? k : -1;
                (side ? whi : wlo)[a] = idx;
            }
        }
    }
    Xoshiro rng(d.seed);
    std::vector<double> inner, cand;
    const int64_t total = Ns[0] * Ns[1] * Ns[2];
    inner.reserve((size_t)(total * D));
    for (int64_t lin = 0; lin < total; ++lin) {
        int64_t I[3] = {lin % Ns[0], (lin / Ns[0]) % Ns[1], lin / (Ns[0] * Ns[1])};
        double p[3];
        bool inside = true;
        for (int a = 0; a < D; ++a) {
            double x = (I[a] == wlo[a]) ? d.mins[a] : (I[a] == whi[a]) ? d.maxs[a] : g.cmin[a] + (double)I[a] * d.dx[a];
            x += d.jitter[a] * (rng.uniform() * 2 - 1);
            if (d.periodic[a]) {     // _wrap_periodic (Domains.jl:314)
                double y = d.mins[a] + jmod(x - d.mins[a], g.Lw[a]);
                x = (y >= d.mins[a] + g.Lw[a]) ? d.mins[a] : y;
            }
            p[a] = x;
            if (x < d.mins[a] || x > d.maxs[a]) inside = false;
        }
        auto& dst = inside ? inner : cand;
        for (int a = 0; a < D; ++a) dst.push_back(p[a]);
    }
    const int64_t ni = (int64_t)inner.size() / D, nc = (int64_t)cand.size() / D;
    double max_dx = d.dx[0];
    for (int a = 1; a < D; ++a) max_dx = std::max(max_dx, d.dx[a]);
    const double cut = g.irf * max_dx + max_dx, cut_sq = cut * cut;
    std::vector<char> keep((size_t)nc, 0);
    if (ni > 0 && nc > 0) {
        // a ghost survives if some interior point lies within the cutoff; cells of width sqrt(cut_sq)
        const double h = std::sqrt(cut_sq);
        double org[3] = {0, 0, 0};
        for (int a = 0; a < D; ++a) org[a] = inner[a];
        for (int64_t i = 1; i < ni; ++i) for (int a = 0; a < D; ++a) org[a] = jmin(org[a], inner[i * D + a]);
        auto cell = [&](const double* p, int64_t* c) { for (int a = 0; a < D; ++a) c[a] = (int64_t)std::floor((p[a] - org[a]) / h); };
        auto hkey = [&](const int64_t* c) {
            uint64_t k = 0;
            for (int a = 0; a < D; ++a) k = k * 2097152ull + (uint64_t)(c[a] + 1048576);
            return k;
        };
        std::unordered_map<uint64_t, std::vector<int64_t>> cells;
        cells.reserve((size_t)ni / 4 + 16);
        for (int64_t i = 0; i < ni; ++i) { int64_t c[3]; cell(&inner[i * D], c); cells[hkey(c)].push_back(i); }
        const int nwin = D == 1 ? 3 : D == 2 ? 9 : 27;
        for (int64_t q = 0; q < nc; ++q) {
            const double* gp = &cand[q * D];
            int64_t c[3]; cell(gp, c);
            bool hit = false;
            for (int w = 0; w < nwin && !hit; ++w) {
                int64_t cc[3]; int rem = w;
                for (int a = 0; a < D; ++a) { cc[a] = c[a] + rem % 3 - 1; rem /= 3; }
                auto it = cells.find(hkey(cc));
                if (it == cells.end()) continue;
                for (int64_t i : it->second) {
                    double s = 0.0;
                    for (int a = 0; a < D; ++a) { double t = gp[a] - inner[i * D + a]; s = a == 0 ? t * t : s + t * t; }
                    if (s <= cut_sq) { hit = true; break; }
                }
            }
            keep[q] = hit;
        }
    }
    std::vector<double> pos(inner);
    for (int64_t q = 0; q < nc; ++q) if (keep[q]) for (int a = 0; a < D; ++a) pos.push_back(cand[q * D + a]);
    const int64_t N = (int64_t)pos.size() / D;
    std::vector<int32_t> tag((size_t)N);
    for (int64_t i = 0; i < N; ++i) tag[i] = box_tag(g, &pos[i * D]);

    // Morton order over the canvas (16-bit keys per axis in 2-D, 10-bit in 3-D, the coordinate in 1-D)
    std::vector<int64_t> perm((size_t)N);
    std::iota(perm.begin(), perm.end(), 0);
    out.morton_ties = 0;
    if (N > 1) {
        double ext[3];
        for (int a = 0; a < D; ++a) ext[a] = g.cmax[a] - g.cmin[a];
        auto akey = [&](double x, int a, double top) {
            double v = std::floor((x - g.cmin[a]) / ext[a] * top);
            v = std::min(std::max(v, 0.0), top);
            return (uint32_t)v;
        };
        if (D == 1) {
            std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return pos[a] < pos[b]; });
            for (int64_t k = 1; k < N; ++k) out.morton_ties += pos[perm[k]] == pos[perm[k - 1]];
        } else {
            std::vector<uint32_t> key((size_t)N);
            for (int64_t i = 0; i < N; ++i) {
                const double* p = &pos[i * D];
                key[i] = D == 2 ? (spread2(akey(p[0], 0, 65535.0)) | (spread2(akey(p[1], 1, 65535.0)) << 1))
                                : (spread3(akey(p[0], 0, 1023.0)) | (spread3(akey(p[1], 1, 1023.0)) << 1) |
                                   (spread3(akey(p[2], 2, 1023.0)) << 2));
            }
            std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
            for (int64_t k = 1; k < N; ++k) out.morton_ties += key[perm[k]] == key[perm[k - 1]];
        }
    }
    out.N = N;
    out.pos.resize((size_t)(N * D)); out.isb.resize((size_t)N); out.tags.resize((size_t)N);
    for (int64_t k = 0; k < N; ++k) {
        int64_t s = perm[k];
        for (int a = 0; a < D; ++a) out.pos[k * D + a] = pos[s * D + a];
        out.tags[k] = tag[s];
        out.isb[k] = tag[s] != 0;
    }
}

// The 3^D bin window of every bin, periodic axes wrapped, repeats and outside bins -> dummy bin nbins
// (src/Grid/NeighborLogic.jl:43-76). Offsets in CartesianIndices((-1:1, ...)) order (x fastest).
void bin_window_table(int D, const int64_t* dims, const int32_t* periodic, std::vector<int32_t>& table) {
    const int nw = D == 1 ? 3 : D == 2 ? 9 : 27;
    const int64_t nb = dims[0] * dims[1] * dims[2];
    table.assign((size_t)(nb * nw), (int32_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
        int64_t c[3] = {b % dims[0], (b / dims[0]) % dims[1], b / (dims[0] * dims[1])};
        int64_t raw[27];
        for (int w = 0; w < nw; ++w) {
            int rem = w; bool ok = true; int64_t f[3] = {0, 0, 0};
            for (int a = 0; a < D; ++a) {
                int64_t v = c[a] + rem % 3 - 1; rem /= 3;
                if (v < 0 || v >= dims[a]) {
                    if (periodic[a]) v = ((v % dims[a]) + dims[a]) % dims[a];
                    else ok = false;
                }
                f[a] = v;
            }
            raw[w] = ok ? f[0] + dims[0] * (f[1] + dims[1] * f[2]) : nb;
        }
        for (int w = 0; w < nw; ++w) {
            bool rep = false;
            for (int k = 0; k < w; ++k) rep |= raw[k] == raw[w];
            table[(size_t)(b * nw + w)] = (int32_t)(rep ? nb : raw[w]);
        }
    }
}

bool build_halos(int64_t N, int block, const std::vector<int32_t>& cnt, const std::vector<int64_t>& chunk_off,
                 const std::vector<int32_t>& nbr, std::vector<int32_t>& halo_off, std::vector<int32_t>& halo_idx,
                 std::vector<uint16_t>& nbr_local, int64_t& max_halo) {
    const int64_t nb = (N + block - 1) / block;
    halo_off.assign((size_t)nb + 1, 0);
    halo_idx.clear();
    nbr_local.assign(nbr.size(), 0);
    max_halo = 0;
    std::vector<int32_t> h;
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t lo = b * block, hi = std::min<int64_t>(N, lo + block);
        h.clear();
        for (int64_t i = lo; i < hi; ++i) {
            h.push_back((int32_t)i);
            const int64_t base = chunk_off[i / kChunk] + (i % kChunk);
            for (int32_t k = 0; k < cnt[i]; ++k) h.push_back(nbr[(size_t)(base + (int64_t)k * kChunk)]);
        }
        std::sort(h.begin(), h.end());
        h.erase(std::unique(h.begin(), h.end()), h.end());
        if (h.size() > 65535) return false;
        max_halo = std::max<int64_t>(max_halo, (int64_t)h.size());
        for (int64_t i = lo; i < hi; ++i) {
            const int64_t base = chunk_off[i / kChunk] + (i % kChunk);
            for (int32_t k = 0; k < cnt[i]; ++k) {
                const size_t s = (size_t)(base + (int64_t)k * kChunk);
                nbr_local[s] = (uint16_t)(std::lower_bound(h.begin(), h.end(), nbr[s]) - h.begin());
            }
        }
        halo_idx.insert(halo_idx.end(), h.begin(), h.end());
        halo_off[(size_t)b + 1] = (int32_t)halo_idx.size();
    }
    return true;
}

// Outflow donor map. OutflowBC (src/Grid/BoundaryConditions.jl:82-131) copies, in up to five passes, the
// state of the nearest donor resolved in an earlier pass (strict < on |d|^2, first slot wins ties); donors
// are interior particles or same-tag particles resolved before. Which donor a particle takes depends on the
// geometry only, and every copy is exact, so each target ends up holding the state of one interior particle:
// root[i]. root = -1 for boundary particles that keep pg.rhos (orphans, other tags, Dirichlet), -2 for interior.
void outflow_roots(const DomainGeom& g, int64_t N, const double* pos, const uint8_t* isb, const int32_t* tags,
                   const std::vector<int64_t>& rstart, const std::vector<int32_t>& nbr, int n_bc, const int32_t* bc_tag,
                   const int32_t* bc_kind, std::vector<int32_t>& root) {
    const int D = g.D;
    root.assign((size_t)N, -1);
    for (int64_t i = 0; i < N; ++i) if (!isb[i]) root[i] = -2;
    std::vector<int64_t> status((size_t)N);
    for (int b = 0; b < n_bc; ++b) {
        if (bc_kind[b] != HCC_BC_OUTFLOW) continue;
        const int tag = bc_tag[b];
        for (int64_t i = 0; i < N; ++i) status[i] = !isb[i] ? 1 : (tags[i] == tag ? 0 : -1);
        for (int pass = 1; pass <= 5; ++pass) {
            std::vector<std::pair<int64_t, int64_t>> resolved;
            for (int64_t i = 0; i < N; ++i) {
                if (status[i] != 0) continue;
                int64_t best = -1; double bd = INFINITY;
                for (int64_t k = rstart[i]; k < rstart[i + 1]; ++k) {
                    int64_t j = nbr[k];
                    if (0 < status[j] && status[j] <= pass) {
                        double s = 0.0;
                        for (int a = 0; a < D; ++a) {
                            double dx = pos[j * D + a] - pos[i * D + a];
                            dx = dx - g.Lw[a] * std::nearbyint(dx * g.invLw[a]);
                            s = a == 0 ? dx * dx : s + dx * dx;
                        }
                        if (s < bd) { bd = s; best = j; }
                    }
                }
                if (best != -1) resolved.push_back({i, best});
            }
            for (auto& r : resolved) {
                status[r.first] = pass + 1;
                root[r.first] = isb[r.second] ? root[r.second] : (int32_t)r.second;
            }
            bool open = false;
            for (int64_t i = 0; i < N; ++i) if (status[i] == 0) { open = true; break; }
            if (!open) break;
        }
    }
}

}  // namespace hcc
