// hc_internal.h -- private structures shared by the hip_context translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "physicell_hip.h"

struct hc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool kernels_loaded = false;
    std::string last_error;
    // per-kernel event timing
    uint32_t timing_mask = 0;
    struct Pending { int kernel; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double total_ms[HC_K_COUNT] = {};
    int64_t launches[HC_K_COUNT] = {};
    // 0: partitioned sweeps where the line length suits them, 1: serial
    // elimination (the reference's operation order, bit-identical to it in FP64)
    int sweep_mode = 0;
    // substep order: 0 automatic, 1 axis-major (x of every substrate, then
    // y, then z), 2 substrate-major (x, y, z of substrate 0, then 1, ...)
    int substep_order = 0;
    // small scratch for reductions read back by the host (max bin etc.)
    uint32_t* d_scalar = nullptr;
    // x lane order per (line length, segments, element size)
    std::map<uint64_t, int> xmap_cache;
    // device copy of host index lists (division parents, removal moves)
    void* d_list = nullptr;
    size_t d_list_bytes = 0;
    // neighbour queries: two maxima and the scan's partials
    void* d_query = nullptr;
    size_t d_query_bytes = 0;
    // this device's address of the static scratch of hc_substrate_totals_noalloc
    // (set by hc_load_kernels, or by that entry point's first call)
    void* totals_scratch = nullptr;
};

struct hc_buf {
    hc_ctx* ctx = nullptr;
    void* dptr = nullptr;
    void* host = nullptr;
    size_t bytes = 0;
    // capacity (max_cells) a secretion plan or hash workspace was last built for
    uint64_t aux = 0;
};

namespace hc {

int fail(hc_ctx* ctx, int code, const std::string& msg);
int hip_fail(hc_ctx* ctx, hipError_t e, const char* what);

// Event pair around one launch of `kernel` when its bit is set in timing_mask.
struct KernelTimer {
    hc_ctx* ctx;
    int kernel;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(hc_ctx* c, int k);
    ~KernelTimer();
};

// After a launch: report a launch error (returns 0 when fine).
int check_launch(hc_ctx* ctx, const char* name);

inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// ---- device-side parameter blocks (passed by value) ----
struct SubConst {          // per substrate, per axis coefficient c = D*dt/d^2
    double c[3];
    double h;              // decay share per sweep
    double T;              // Dirichlet value
    uint32_t dir;          // Dirichlet flag
    uint32_t _pad;
};

// Thomas table layout: [(s*3 + axis)*2 + {0: c', 1: denominator}] * nmax + i
__host__ __device__ inline size_t table_index(uint32_t s, int axis, int which, uint32_t nmax) {
    return ((size_t)(s * 3 + axis) * 2 + which) * nmax;
}

// Layout of the table buffer: SubConst[ns] (padded to 256 B) then the tables.
__host__ __device__ inline size_t subconst_bytes(uint32_t ns) { return ((sizeof(SubConst) * ns + 255) / 256) * 256; }
inline uint32_t grid_nmax(const hc_grid* g) {
    uint32_t m = g->nx > g->ny ? g->nx : g->ny;
    return m > g->nz ? m : g->nz;
}

// ---- partitioned sweep tables (see sweep_part.hip) ----
// Workgroups are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).
// xcd_block renumbers them so that each XCD works through one contiguous
// range of the work: neighbour bins of hash-sorted agents, and the cache lines
// that adjacent line groups of a sweep share, then stay in one XCD's L2
// instead of being fetched by several.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
    constexpr uint32_t NX = 8;
    const uint32_t x = b % NX, idx = b / NX, per = nb / NX, rem = nb % NX;
    return x * per + (x < rem ? x : rem) + idx;
}
// The same in chunks of C consecutive work items dealt to the XCDs in turn:
// neighbouring items still share one L2 (all but one boundary in C), and
// work lists whose cost varies along the list (longer lines first) stay
// balanced across the XCDs.  A tail shorter than 8*C keeps its order.
template <uint32_t C>
__device__ inline uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
    constexpr uint32_t NX = 8;
    if (b >= nb / (NX * C) * (NX * C)) return b;
    const uint32_t x = b % NX, idx = b / NX;
    return (idx / C) * (NX * C) + x * C + idx % C;
}

// A line's N = n-2 unknowns are split into P contiguous segments, one per
// lane (P = 4: 16 lines per wave, or P = 16: 4 lines per wave; see part_p).
// Per (substrate, axis) block: for
// each segment k the arrays 1/den, c', alpha, beta of length
// part_mt(nmax, P), then the two rows (left / right neighbour value) of the
// inverse interface matrix, each of length 2P.  Blocks have a common stride.
// Segments per line.  y and z: 4 (a wave solves 16 lines), or 8 (8 lines)
// for lines longer than 130, whose 4-segment windows of up to 64 values take
// 161 VGPRs in FP64 (3 waves per SIMD) and 92 in FP32 (5 waves) against 88
// (5 waves) and 64 (8 waves) for 8 segments of up to 32.  x lines are staged through LDS, a wave's lines at a
// time: 4 segments (16 lines per wave) while 64 lines of n values fit the
// staging tile (n <= 130 in FP32, n <= 96 in FP64); longer FP32 lines (up to
// 258) take 8 segments (8 lines per wave, windows of up to 32: half the
// interface work per line of 16 segments; the 200-node x sweep of one FP32
// simulation 57.1 -> 45.0 us), everything else 16 segments (4 lines per
// wave; for FP64 200-node lines 8 segments measured 87 -> 92 us).
__host__ __device__ inline int part_p(int axis, uint32_t n, bool fp32) {
    if (axis != 0) return n > 130u ? 8 : 4;
    if (n <= (fp32 ? 130u : 96u)) return 4;
    return (fp32 && n <= 258u) ? 8 : 16;
}
__host__ __device__ inline uint32_t part_mt(uint32_t nmax, int P) {
    return nmax > 2 ? (nmax - 2 + P - 1) / P : 1u;
}
__host__ __device__ inline size_t part_block_p(uint32_t nmax, int P) {
    return (size_t)P * 4 * part_mt(nmax, P) + (size_t)P * 2 * 2 * P;
}
__host__ __device__ inline size_t part_block(uint32_t nmax) {
    const size_t a = part_block_p(nmax, 4), b = part_block_p(nmax, 16), c = part_block_p(nmax, 8);
    const size_t m = a > b ? a : b;
    return ((m > c ? m : c) + 31) / 32 * 32;
}
__host__ __device__ inline size_t serial_table_bytes(uint32_t ns, uint32_t nmax) {
    return ((subconst_bytes(ns) + (size_t)ns * 3 * 2 * nmax * sizeof(double)) + 255) / 256 * 256;
}
__host__ __device__ inline size_t part_offset(uint32_t ns, uint32_t nmax) { return serial_table_bytes(ns, nmax); }
// FP32 copies of the per-segment arrays (for FP32 fields), after the FP64 blocks
__host__ __device__ inline size_t part_fblock(uint32_t nmax) {
    const size_t a = (size_t)4 * 4 * part_mt(nmax, 4), b = (size_t)16 * 4 * part_mt(nmax, 16);
    const size_t c = (size_t)8 * 4 * part_mt(nmax, 8), m = a > b ? a : b;
    return ((m > c ? m : c) + 63) / 64 * 64;
}
__host__ __device__ inline size_t part_f_offset(uint32_t ns, uint32_t nmax) {
    return part_offset(ns, nmax) + (size_t)ns * 3 * part_block(nmax) * sizeof(double);
}
__host__ __device__ inline size_t table_total_bytes(uint32_t ns, uint32_t nmax) {
    return part_f_offset(ns, nmax) + (size_t)ns * 3 * part_fblock(nmax) * sizeof(float);
}

// ---- packed ensemble descriptors (batch.hip) ----
struct BatchSim {            // one simulation of a packed ensemble
    void* rho;               // its density field (FP64 or FP32, as the launch)
    const uint8_t* tab;      // its Thomas tables
    uint32_t nx, ny, nz, ns, nmax;
    uint32_t last_axis;      // the sweep that re-applies Dirichlet (2 in 3-D, 1 in 2-D)
    uint32_t xmap;           // x lane order (x_lane_order)
    uint32_t wide;           // bit a: lines along axis a need the 32-value register window (else 16)
};
// number of lines a sweep along `axis` solves
__host__ __device__ inline uint64_t sweep_lines(int axis, uint32_t nx, uint32_t ny, uint32_t nz) {
    return axis == 0 ? (uint64_t)ny * nz : (axis == 1 ? (uint64_t)nz * nx : (uint64_t)ny * nx);
}
struct BatchWork {           // one line group of one substrate of one simulation
    uint32_t sim, s;
    uint64_t grp;
};

// Lane order of an LDS-staged x sweep: lane -> (line l, segment k) is either
// k = lane / LPW, l = lane % LPW (0) or k = lane % P, l = lane / P (1).  Both
// give the same results; they differ in how the lanes' reads of the staging
// tile (line l at l*n, segment k at floor(k*(n-2)/P)) fall on the 32 LDS
// banks.  Returns the order with fewer bank cycles for lines of n values:
// every read step is costed as the sum, over groups of lanes served together
// (128 bytes per cycle), of the largest number of distinct words on one bank.
inline int x_lane_order(uint32_t n, int P, size_t elem) {
    const int LPW = 64 / P, wpl = (int)(elem / 4), per = 32 / wpl;
    const uint32_t N = n - 2, MM = (N + P - 1) / P;
    long cost[2] = {0, 0};
    for (int tr = 0; tr < 2; tr++)
        for (uint32_t t = 0; t < MM; t++)
            for (int g0 = 0; g0 < 64; g0 += per) {
                uint32_t words[32][32];
                int cnt[32] = {0};
                int worst = 0;
                for (int lane = g0; lane < g0 + per; lane++) {
                    const uint32_t k = tr ? lane % P : lane / LPW, l = tr ? lane / P : lane % LPW;
                    const uint32_t sk = (k * N) / P, mk = ((k + 1) * N) / P - sk;
                    const uint32_t idx = l * n + sk + 1 + (t < mk ? t : mk - 1);
                    for (int q = 0; q < wpl; q++) {
                        const uint32_t wd = idx * wpl + q, b = wd % 32;
                        bool seen = false;
                        for (int i = 0; i < cnt[b]; i++) seen |= words[b][i] == wd;
                        if (!seen && cnt[b] < 32) words[b][cnt[b]++] = wd;
                        worst = cnt[b] > worst ? cnt[b] : worst;
                    }
                }
                cost[tr] += worst;
            }
    return cost[1] < cost[0] ? 1 : 0;
}

inline int ctx_x_lane_order(hc_ctx* ctx, uint32_t n, int P, size_t elem) {
    const uint64_t key = ((uint64_t)n << 16) | ((uint64_t)P << 4) | (uint64_t)elem;
    auto it = ctx->xmap_cache.find(key);
    if (it != ctx->xmap_cache.end()) return it->second;
    const int v = x_lane_order(n, P, elem);
    ctx->xmap_cache[key] = v;
    return v;
}

// ---- launchers implemented in diffusion.hip / cells.hip / sweep_part.hip ----
// substrates s0 .. s0+scount-1 (scount 0: all from s0); returns 1 when the
// grid does not suit the partitioned form (the caller falls back)
int launch_sweep_part(hc_ctx* ctx, int axis, void* rho, const hc_grid* g, const uint8_t* tables,
                      int fuse_dirichlet, uint32_t s0 = 0, uint32_t scount = 0);
// 1 when every sweep of the grid takes the partitioned form
bool part_fits_all(const hc_grid* g, int sweep_mode);
// P: the segments per line of every member in the work list (part_p); mm:
// their register window, 16, or 0 for each member's own (16 or 32 values,
// BatchSim::wide; P = 4); grid_blocks 0: one workgroup per item, else a
// persistent grid of that many workgroups
int launch_sweep_batch(hc_ctx* ctx, int axis, int P, int mm, bool fp32, const BatchSim* sims, const BatchWork* work,
                       uint32_t nwork, uint32_t nmax_all, uint32_t max_line, unsigned grid_blocks);
// register window of a line of n values in P segments (0: does not fit)
int part_window(uint32_t n, int P);
void build_part_tables_host(const hc_grid* g, const hc_substrate* subs, double* out, float* fout);
void build_tables_host(const hc_grid* g, const hc_substrate* subs, std::vector<uint8_t>& out);
inline size_t field_elem(const hc_grid* g) { return g->field_fp32 ? sizeof(float) : sizeof(double); }
int launch_sweep(hc_ctx* ctx, int axis, void* rho, const hc_grid* g, const uint8_t* tables,
                 int fuse_dirichlet);
int launch_dirichlet(hc_ctx* ctx, void* rho, const hc_grid* g, const uint8_t* tables);

// Hash workspace layout (byte offsets inside the hash buffer).
struct HashLayout {
    size_t counts, offsets, slot, tmp, sorted, rec, pos_of, pos32, rpart, rmax, partial, xoff, xslab, xtflag, total;
};
HashLayout hash_layout(uint32_t n_mvox, uint32_t max_cells);

// Mechanics mirror of one simulation (cells.hip): each agent's mechanics
// state in slots ordered like a recent step's records (two sets: the second
// receives a re-sort), the coming step's bin and place in the bin per slot,
// the step's forces by record position; spos: each slot's place in the
// step's (bin, x slab, index) order of the FP32 mark positions (plain form).
struct MirrorLayout {
    size_t pos[2], prev[2], prop[2], idx[2], key[2], slot[2], kstep[2], rpos[2], xslab[2], vp, spos, total;
};
MirrorLayout mirror_layout(uint32_t max_cells);
// builds the hash from the cell fields and the mirror from it
int mirror_load(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, hc_buf* mirror, const hc_mech* mech, const hc_grid* mgrid,
                uint32_t n, uint32_t max_cells, uint32_t n_mvox);
// one mechanics step (hash -> forces -> AB2) on the mirror; the cell fields
// are not touched
int mirror_step(hc_ctx* ctx, hc_buf* hash, hc_buf* mirror, const hc_mech* mech, const hc_grid* mgrid, double dt,
                uint32_t n, uint32_t max_cells, uint32_t n_mvox);
// puts the slots back in the record order of the last step
int mirror_resort(hc_ctx* ctx, hc_buf* mirror, uint32_t n, uint32_t max_cells);
// position, velocity, previous velocity, pressure and mechanics voxel of
// every agent from the mirror into the cell fields
int mirror_sync(hc_ctx* ctx, hc_buf* cells, hc_buf* mirror, uint32_t n, uint32_t max_cells);

struct PlanLayout {
    size_t header, counts, offsets, slot, tmp, sorted, gflag, gvox, gbegin, coef, partial, total;
};
PlanLayout plan_layout(const hc_grid* g, uint32_t max_cells);

// exclusive scan of n uint32 values in place into out (out may equal in),
// out[n] = total; partial is scratch of scan_partial_bytes(n).  clear_in
// (in's address, out elsewhere): in is left zero (histogram counters for the
// next pass), by the scan's last kernel where it can, else by a fill.
size_t scan_partial_bytes(uint64_t n);
int scan_exclusive(hc_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partial,
                   uint32_t* clear_in = nullptr);
// looks up this device's address of hc_substrate_totals_noalloc's static
// scratch (once per device) and keeps it in ctx->totals_scratch
int totals_scratch_init(hc_ctx* ctx);

}  // namespace hc
