// cells_common.h -- device helpers shared by the single-simulation and the
// packed-ensemble agent kernels.
#pragma once

#include "hc_internal.h"

namespace hc {

// 64-byte hash-sorted record (FP64 forces)
struct __attribute__((aligned(16))) CellRec {
    double x, y, z, r, rep, adh, rel;
    uint32_t idx, key;
};
static_assert(sizeof(CellRec) == 64, "record must be 64 bytes");

struct CellPtrs {
    double* f[HC_NUM_F64_FIELDS];
    uint32_t* u[HC_NUM_U32_FIELDS];
};

__host__ __device__ inline CellPtrs cell_ptrs(void* base, uint32_t max_cells) {
    CellPtrs p;
    for (int f = 0; f < HC_NUM_F64_FIELDS; f++) p.f[f] = (double*)base + (size_t)f * max_cells;
    uint32_t* ub = (uint32_t*)((double*)base + (size_t)HC_NUM_F64_FIELDS * max_cells);
    for (int f = 0; f < HC_NUM_U32_FIELDS; f++) p.u[f] = ub + (size_t)f * max_cells;
    return p;
}

// 32-byte hash-sorted record (FP32 forces, the mixed form): xg = x relative
// to the grid's x_min (the mark pass's coordinate), y and z relative to the
// lower corner of the agent's own bin, x relative to the same corner (the
// force law's coordinates: values below ~60 um keep ~4e-6 um of resolution
// wherever the domain sits), radius, repulsion, adhesion, agent index.
struct __attribute__((aligned(16))) CellRec32 {
    float xg, y, z, x, r, rep, adh;
    uint32_t idx;
};
static_assert(sizeof(CellRec32) == 32, "mixed record must be 32 bytes");

// three packed doubles (24 bytes): the mechanics mirrors' positions and
// previous velocities, so AB2 and the gathers move no padding
struct Dbl3 {
    double x, y, z;
};
static_assert(sizeof(Dbl3) == 24, "packed triple");
__device__ inline Dbl3 dbl3(double x, double y, double z) {
    Dbl3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

// FP32 records of a bin are ordered by x slab, then by agent index: XSLABS
// slabs of voxel / XSLABS along x, by the record's bin-relative x.
// xoff[XSLABS * b + s] is the position of the first record of slab s of bin
// b, so xoff[XSLABS * b] = offsets[b], and the XSLABS entries past the last
// bin hold n.  The members of a row of bins whose x lies in an interval are
// then one contiguous range of records, found with two loads.
constexpr uint32_t XSLABS = 4;
__host__ __device__ inline float xslab_scale(double voxel) { return (float)((double)XSLABS / voxel); }
// bin-relative FP32 coordinate, the same expression wherever it is formed
__host__ __device__ inline float bin_rel32(double p, double origin, uint32_t bin, double voxel) {
    return (float)(p - (origin + bin * voxel));
}
__device__ inline uint32_t xslab_of(float xb, float scale) {
    const float s = floorf(xb * scale);
    return !(s > 0.0f) ? 0u : (s >= (float)(XSLABS - 1) ? XSLABS - 1 : (uint32_t)s);
}

// ---------------------------------------------------------------------------
// wavefront primitives
// ---------------------------------------------------------------------------
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Ballot binning: returns the slot of this lane's key inside its bin.  The
// lanes are grouped by key with ballots (no memory traffic in the loop), at
// most 8 groups; each group's lowest lane adds the group's size to the bin
// counter with one atomic, and lanes left after 8 groups (a wave whose keys
// are mostly distinct) add 1 on their own.  All atomics of the wave are in
// flight together; members take consecutive slots from the returned base.
__device__ inline uint32_t ballot_bin(bool valid, uint32_t key, uint32_t* counts) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(valid);
    int leader_of = lane;
    uint32_t rank = 0, cnt = 1;
    for (int g = 0; g < 8 && pending; g++) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lk = __builtin_amdgcn_readlane(key, leader);
        const uint64_t same = __ballot(valid && key == lk) & pending;
        if ((same >> lane) & 1ull) {
            leader_of = leader;
            rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        }
        if (lane == leader) cnt = (uint32_t)__popcll(same);
        pending &= ~same;
    }
    uint32_t base = 0;
    if (valid && leader_of == lane) base = atomicAdd(&counts[key], cnt);
    base = __shfl(base, leader_of, 64);
    return base + rank;
}

// rank inside the bin = number of bin members with a smaller index.  The
// members are read 8 at a time (indices clamped into the bin) so their loads
// are in flight together instead of one latency per member.
__device__ inline uint32_t stable_rank(const uint32_t* __restrict__ tmp, uint32_t beg, uint32_t end, uint32_t c) {
    uint32_t r = 0;
    for (uint32_t t0 = beg; t0 < end; t0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = tmp[t0 + q < end ? t0 + q : end - 1];
#pragma unroll
        for (int q = 0; q < 8; q++) r += (t0 + q < end && v[q] < c) ? 1u : 0u;
    }
    return r;
}


// Bins of at most BIN_SORT_CAP members are put in ascending order by one
// thread each (k_bin_sort: a 16-input bitonic network in registers, padded
// with ~0u), which writes the bin's sorted list and every member's position;
// the gather then reads that position instead of ranking the member against
// its bin (stable_rank), which it still does for larger bins.
constexpr uint32_t BIN_SORT_CAP = 16;

__device__ inline void bin_sort_one(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ tmp,
                                    uint32_t* __restrict__ sorted, uint32_t* __restrict__ pos_of, uint32_t b) {
    const uint32_t beg = offsets[b], cnt = offsets[b + 1] - beg;
    if (cnt == 0 || cnt > BIN_SORT_CAP) return;
    uint32_t a[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) a[q] = q < cnt ? tmp[beg + q] : ~0u;
#pragma unroll
    for (uint32_t k = 2; k <= BIN_SORT_CAP; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t i = 0; i < BIN_SORT_CAP; i++) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t lo = a[i] < a[l] ? a[i] : a[l], hi = a[i] < a[l] ? a[l] : a[i];
                    const bool up = (i & k) == 0;
                    a[i] = up ? lo : hi;
                    a[l] = up ? hi : lo;
                }
            }
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++)
        if (q < cnt) {
            sorted[beg + q] = a[q];
            pos_of[a[q]] = beg + q;
        }
}

// The largest r of a 256-thread block into part[blockIdx.x]; every thread of
// the block must call it (it holds a barrier).
__device__ inline void block_max_to(float r, float* __restrict__ part) {
    __shared__ float wmax[4];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) r = fmaxf(r, __shfl_xor(r, d, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
}

// Counter-based uniform stream on [0, 1): the SplitMix64 finaliser (Steele,
// Lea & Flood 2014) applied to (seed, step, cell, draw).  Stateless, so graph
// replays and parallel lanes need no generator state.
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double uniform01(uint64_t seed, uint64_t step, uint32_t cell, uint32_t draw) {
    const uint64_t h = mix64(seed ^ mix64(step ^ mix64(((uint64_t)cell << 2) | draw)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// the motility buffer (hc_motility_field_offset)
struct MotPtrs {
    double *persistence, *elapsed, *bias, *vec[3];
    uint32_t *motile, *restrict_2d;
    int32_t *chem_index, *chem_dir;
    double* sensitivity;
    uint64_t* hdr;       // [0] step, [1] seed
};

inline MotPtrs mot_ptrs(void* base, uint32_t max_cells, uint32_t ns) {
    uint8_t* b = (uint8_t*)base;
    MotPtrs m;
    m.hdr = (uint64_t*)b;
    auto off = [&](uint32_t f) { return b + hc_motility_field_offset(f, max_cells, ns); };
    m.persistence = (double*)off(HC_M_PERSISTENCE);
    m.elapsed = (double*)off(HC_M_ELAPSED);
    m.bias = (double*)off(HC_M_BIAS);
    m.vec[0] = (double*)off(HC_M_VEC_X);
    m.vec[1] = (double*)off(HC_M_VEC_Y);
    m.vec[2] = (double*)off(HC_M_VEC_Z);
    m.motile = (uint32_t*)off(HC_M_MOTILE);
    m.restrict_2d = (uint32_t*)off(HC_M_RESTRICT_2D);
    m.chem_index = (int32_t*)off(HC_M_CHEM_INDEX);
    m.chem_dir = (int32_t*)off(HC_M_CHEM_DIR);
    m.sensitivity = (double*)off(HC_M_SENSITIVITY);
    return m;
}

int fill_u32(hc_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v);
// the largest of nb block maxima rpart[] into rmax[0] (one block)
int launch_rmax(hc_ctx* ctx, const float* rpart, uint32_t nb, float* rmax);
// k_bin_sort over bins 0 .. nbins-1 (bin_sort_one)
int launch_bin_sort(hc_ctx* ctx, const uint32_t* offsets, const uint32_t* tmp, uint32_t* sorted, uint32_t* pos_of,
                    uint32_t nbins);
// k_bin_sort32 (bin_sort_slabs) over bins 0 .. nbins-1 and the entries past the last
int launch_bin_sort32(hc_ctx* ctx, const uint32_t* offsets, const uint32_t* tmp, const uint8_t* xslab,
                      uint32_t* sorted, uint32_t* rpos, uint32_t* xoff, uint32_t nbins);
int launch_scatter(hc_ctx* ctx, const uint32_t* key, const uint32_t* slot, const uint32_t* offsets, uint32_t* tmp,
                   uint32_t n);
int launch_stable_sort(hc_ctx* ctx, const uint32_t* key, const uint32_t* offsets, const uint32_t* tmp,
                       uint32_t* sorted, uint32_t n);
int launch_sec_flags(hc_ctx* ctx, const uint32_t* key, const uint32_t* sorted, const uint32_t* offsets, uint32_t* flag,
                     uint32_t n, uint32_t V);

struct PlanHeader {
    uint32_t n_elig, ngroups, has_net, n;
};

// Neighbour rows of an agent in bin (vi, vj, vk): the 3 (2-D) or 9 (3-D)
// rows (dk, dj) of up to 3 bins (di = -1, 0, 1).  The bins of one row are
// consecutive keys, so their members form one contiguous range of the
// hash-sorted order, in the reference's visit order (dk, dj, di) x ascending
// index (shaders/mechanics.metal:151-203).  All bin offsets are loaded up
// front (independent loads).  bin0 is the key of the simulation's bin 0.
// One of the 9 neighbour rows (q = (dk+1)*3 + (dj+1); 3 in 2-D): the members
// of bins (vi-1 .. vi+1, vj+dj, vk+dk) are one contiguous range of the
// hash-sorted order; mid0 / mid1 start the agent's column and the next.
struct RowRange {
    uint32_t beg, end, mid0, mid1;
    int dk, dj;
};
__device__ __attribute__((always_inline)) inline RowRange neighbor_row(const uint32_t* __restrict__ offsets,
                                                                       uint32_t bin0, uint32_t vi, uint32_t vj,
                                                                       uint32_t vk, uint32_t mnx, uint32_t mny,
                                                                       uint32_t mnz, int q) {
    const uint32_t lo = vi > 0 ? vi - 1 : 0, hi = vi + 1 < mnx ? vi + 1 : mnx - 1;
    const bool three_d = mnz > 1;
    RowRange r;
    r.dk = three_d ? q / 3 - 1 : 0;
    r.dj = q % 3 - 1;
    const int nk = (int)vk + r.dk, nj = (int)vj + r.dj;
    const bool ok = (three_d || q < 3) && nk >= 0 && nk < (int)mnz && nj >= 0 && nj < (int)mny;
    const uint32_t rb = bin0 + (ok ? (uint32_t)nk * mny * mnx + (uint32_t)nj * mnx : 0u);
    r.beg = ok ? offsets[rb + lo] : 0u;
    r.end = ok ? offsets[rb + hi + 1] : 0u;
    r.mid0 = ok ? offsets[rb + vi] : 0u;
    r.mid1 = ok ? offsets[rb + vi + 1] : 0u;
    return r;
}

// Pair forces of one agent against the hash-sorted members [beg, end), in
// ascending order (shaders/mechanics.metal:151-203).  Two passes per chunk
// of up to 32 members: a distance test marks the candidates that can exert a
// force (their loads issued 4 at a time), then the force law runs over the
// marked ones in the same order, so the lanes of a wave run the expensive
// part about as often as they have interacting neighbours, not once per
// candidate.
//
// A pair inside the cutoff exerts a force only when dist < R_sum or dist <
// rel_max_adh * R_sum; otherwise the force law adds +0 to the sums, which
// leaves them unchanged (they start at +0, so they are never -0).  The mark
// pass therefore also requires d^2 < L^2 (1 + m), L = R_sum * max(1,
// rel_max_adh), with a margin m far above the rounding of L^2 and of the
// square root (2^-40 in FP64, 2^-16 in FP32): every pair it drops has a
// rounded dist >= L, so dropping it changes no bit of the result.
__device__ __attribute__((always_inline)) inline void force_range64(
    const CellRec* __restrict__ rec, uint32_t beg, uint32_t end, uint32_t t, double px, double py, double pz,
    double radius_i, double rep_i, double adh_i, double rel_max_adh, double cut2, double& fx, double& fy,
    double& fz, double& local_pressure) {
    const double relf = rel_max_adh > 1.0 ? rel_max_adh : 1.0;
    constexpr double margin = 1.0 + 0x1p-40;
    for (uint32_t u0 = beg; u0 < end; u0 += 32) {
        const uint32_t cnt = end - u0 < 32u ? end - u0 : 32u;
        uint32_t mask = 0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += 4) {
            double2 xy[4], zr[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t u = u0 + (j0 + q < cnt ? j0 + q : cnt - 1);
                xy[q] = *reinterpret_cast<const double2*>(&rec[u].x);
                zr[q] = *reinterpret_cast<const double2*>(&rec[u].z);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t j = j0 + q;
                const double ddx = px - xy[q].x;
                const double ddy = py - xy[q].y;
                const double ddz = pz - zr[q].x;
                const double dist_sq = ddx * ddx + ddy * ddy + ddz * ddz;
                const double L = (radius_i + zr[q].y) * relf;
                const bool inside = j < cnt && u0 + j != t && !(dist_sq >= cut2) && !(dist_sq < 1e-12) &&
                                    dist_sq < L * L * margin;
                mask |= inside ? (1u << j) : 0u;
            }
        }
        while (mask) {
            const uint32_t j = __builtin_ctz(mask);
            mask &= mask - 1u;
            const uint32_t u = u0 + j;
            const double2 xy = *reinterpret_cast<const double2*>(&rec[u].x);
            const double2 zr = *reinterpret_cast<const double2*>(&rec[u].z);
            const double ddx = px - xy.x;
            const double ddy = py - xy.y;
            const double ddz = pz - zr.x;
            const double dist_sq = ddx * ddx + ddy * ddy + ddz * ddz;
            const double dist = sqrt(dist_sq);
            const double inv_dist = 1.0 / dist;
            const double nxv = ddx * inv_dist;
            const double nyv = ddy * inv_dist;
            const double nzv = ddz * inv_dist;
            const double R_sum = radius_i + zr.y;
            double force_mag = 0.0;
            if (dist < R_sum) {
                const double overlap = 1.0 - dist / R_sum;
                force_mag += overlap * overlap * sqrt(rep_i * rec[u].rep);
                local_pressure += overlap * overlap / 0.027288820670331;
            }
            const double adhesion_distance = rel_max_adh * R_sum;
            if (dist < adhesion_distance) {
                const double t2 = 1.0 - dist / adhesion_distance;
                force_mag -= t2 * t2 * sqrt(adh_i * rec[u].adh);
            }
            fx += force_mag * nxv;
            fy += force_mag * nyv;
            fz += force_mag * nzv;
        }
    }
}

// FP64 forces with an FP32 mark pass.  pos32 holds each hash-sorted agent's
// position relative to the lower corner of its own bin, and its radius, in
// FP32: 16 bytes per candidate instead of the 32 the FP64 mark pass reads,
// and FP32 arithmetic.  The FP32 test keeps every pair the exact test keeps
// when the bin size and the cutoff are at most 32 um and the agent lies
// within two bins of its bin corner (always, for agents inside the grid):
// every coordinate of a candidate truly within the cutoff is then below
// 128 um, the FP32 differences are within 3e-5 um per axis of the exact
// ones, d^2 within 4e-3 um^2, and the FP32 cutoff and interaction radius
// within 2e-3 um^2; the threshold carries 0.01 um^2 of slack.  The marked
// candidates are then tested exactly in FP64 with the expressions of
// force_range64 before its force law, so the pairs that contribute, their
// order and every bit of the result are those of force_range64.
__device__ __attribute__((always_inline)) inline void force_range64m(
    const CellRec* __restrict__ rec, const float4* __restrict__ pos32, uint32_t beg, uint32_t end, uint32_t t,
    uint32_t mid0, uint32_t mid1, float mx, float my, float mz, float oy, float oz, float voxel, double px, double py,
    double pz, double radius_i, double rep_i, double adh_i, double rel_max_adh, double cut2, double& fx, double& fy,
    double& fz, double& local_pressure) {
    const double relf = rel_max_adh > 1.0 ? rel_max_adh : 1.0;
    constexpr double margin = 1.0 + 0x1p-40;
    const float relf32 = (float)relf, ri32 = (float)radius_i, cut2f = (float)cut2;
    const float oxm = -1.0f * voxel, ox0 = 0.0f * voxel, oxp = 1.0f * voxel;
    constexpr uint32_t FB = 4;
    for (uint32_t u0 = beg; u0 < end; u0 += 32) {
        const uint32_t cnt = end - u0 < 32u ? end - u0 : 32u;
        uint32_t mask = 0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += FB) {
            float4 o[FB];
#pragma unroll
            for (int q = 0; q < (int)FB; q++) o[q] = pos32[u0 + (j0 + q < cnt ? j0 + q : cnt - 1)];
#pragma unroll
            for (int q = 0; q < (int)FB; q++) {
                const uint32_t j = j0 + q, u = u0 + j;
                const float ox = u < mid0 ? oxm : (u < mid1 ? ox0 : oxp);
                const float ddx = mx - o[q].x - ox, ddy = my - o[q].y - oy, ddz = mz - o[q].z - oz;
                const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                const float L = (ri32 + o[q].w) * relf32;
                const float L2 = L * L * (1.0f + 0x1p-16f);
                const float lim = (L2 < cut2f ? L2 : cut2f) + 0.01f;
                const bool inside = j < cnt && u != t && d2 < lim;
                mask |= inside ? (1u << j) : 0u;
            }
        }
        while (mask) {
            const uint32_t j = __builtin_ctz(mask);
            mask &= mask - 1u;
            const uint32_t u = u0 + j;
            const double2 xy = *reinterpret_cast<const double2*>(&rec[u].x);
            const double2 zr = *reinterpret_cast<const double2*>(&rec[u].z);
            const double ddx = px - xy.x;
            const double ddy = py - xy.y;
            const double ddz = pz - zr.x;
            const double dist_sq = ddx * ddx + ddy * ddy + ddz * ddz;
            const double Lx = (radius_i + zr.y) * relf;
            if (dist_sq >= cut2 || dist_sq < 1e-12 || !(dist_sq < Lx * Lx * margin)) continue;
            const double dist = sqrt(dist_sq);
            const double inv_dist = 1.0 / dist;
            const double nxv = ddx * inv_dist;
            const double nyv = ddy * inv_dist;
            const double nzv = ddz * inv_dist;
            const double R_sum = radius_i + zr.y;
            double force_mag = 0.0;
            if (dist < R_sum) {
                const double overlap = 1.0 - dist / R_sum;
                force_mag += overlap * overlap * sqrt(rep_i * rec[u].rep);
                local_pressure += overlap * overlap / 0.027288820670331;
            }
            const double adhesion_distance = rel_max_adh * R_sum;
            if (dist < adhesion_distance) {
                const double t2 = 1.0 - dist / adhesion_distance;
                force_mag -= t2 * t2 * sqrt(adh_i * rec[u].adh);
            }
            fx += force_mag * nxv;
            fy += force_mag * nyv;
            fz += force_mag * nzv;
        }
    }
}

// FP64 forces of one agent with an FP32 mark pass over all its neighbour
// rows and a per-lane list of the marked candidates in LDS.
//
// pos32 holds each hash-sorted agent's x relative to the grid's x_min (valid
// while the grid's x extent is at most MARKX_MAX_EXTENT), y and z relative to
// the lower corner of its own bin, and its radius, in FP32.  A member of row
// (dj, dk) then differs from the agent by (x_a - x_m, y_a - dj*v - y_m,
// z_a - dk*v - z_m) with no per-member column offset.  Every candidate truly
// within the cutoff lies within 4096 um of x_min, so its FP32 x is within
// 1.22e-4 um and x differences within 2.5e-4 um; y and z differences are
// within 3e-5 um (bin-relative values below 64 um); d^2 of a pair within
// 32 um is then within 0.025 um^2 of the exact value, interaction radius and
// cutoff included, and the threshold carries MARKX_SLACK = 0.05 um^2.
//
// A candidate passing the FP32 test is appended to the lane's list (list
// entry k at list[k * 64]); the FP64 law then walks the list in order with
// the exact predicates and expressions of force_range64, so the pairs that
// contribute, their order (rows q = 0..8, ascending sorted position) and
// every bit of the sums are force_range64's.  The agent itself passes the
// FP32 test and is dropped by the exact test (d^2 < 1e-12).  The list is
// walked whenever the next step could overflow it, and at the end, so the
// wave runs the FP64 law about as often as its busiest lane has neighbours,
// not once per chunk of candidates.
// a 32-bit address in LDS (the per-lane mark lists)
using lds_u32 = __attribute__((address_space(3))) uint32_t;
constexpr double MARKX_MAX_EXTENT = 3968.0;   // x extent of the grid; plus 2 bins + cutoff < 4096
constexpr float MARKX_SLACK = 0.05f;
constexpr uint32_t MARKX_STEP = 8;            // candidates per step (their loads in flight together)
constexpr uint32_t MARKX_CAP = 24;            // list entries per lane

// The FP64 law of one listed pair (the expressions of force_range64):
// xy = (x, y), zr = (z, r), ra = (rep, adh) of the candidate's record.
__device__ __attribute__((always_inline)) inline void force_pair64(
    const double2 xy, const double2 zr, const double2 ra, double px, double py, double pz, double radius_i,
    double rep_i, double adh_i, double rel_max_adh, double relf, double cut2, double& fx, double& fy, double& fz,
    double& local_pressure) {
    constexpr double margin = 1.0 + 0x1p-40;
    const double ddx = px - xy.x;
    const double ddy = py - xy.y;
    const double ddz = pz - zr.x;
    const double dist_sq = ddx * ddx + ddy * ddy + ddz * ddz;
    const double Lx = (radius_i + zr.y) * relf;
    if (dist_sq >= cut2 || dist_sq < 1e-12 || !(dist_sq < Lx * Lx * margin)) return;
    const double dist = sqrt(dist_sq);
    const double inv_dist = 1.0 / dist;
    const double nxv = ddx * inv_dist;
    const double nyv = ddy * inv_dist;
    const double nzv = ddz * inv_dist;
    const double R_sum = radius_i + zr.y;
    double force_mag = 0.0;
    if (dist < R_sum) {
        const double overlap = 1.0 - dist / R_sum;
        force_mag += overlap * overlap * sqrt(rep_i * ra.x);
        local_pressure += overlap * overlap / 0.027288820670331;
    }
    const double adhesion_distance = rel_max_adh * R_sum;
    if (dist < adhesion_distance) {
        const double t2 = 1.0 - dist / adhesion_distance;
        force_mag -= t2 * t2 * sqrt(adh_i * ra.y);
    }
    fx += force_mag * nxv;
    fy += force_mag * nyv;
    fz += force_mag * nzv;
}

// The law over a lane's list of marked candidates (record positions), in
// list order, LAW_BATCH entries at a time: the records of a batch (48 of
// their 64 bytes each) are loaded together, then the pairs run in order.
constexpr uint32_t LAW_BATCH = 4;
template <uint32_t BATCH = LAW_BATCH>
__device__ __attribute__((always_inline)) inline void force_list64(
    const CellRec* __restrict__ rec, const uint32_t* __restrict__ list, uint32_t cnt, double px, double py,
    double pz, double radius_i, double rep_i, double adh_i, double rel_max_adh, double cut2, double& fx, double& fy,
    double& fz, double& local_pressure) {
    const double relf = rel_max_adh > 1.0 ? rel_max_adh : 1.0;
    for (uint32_t k0 = 0; k0 < cnt; k0 += BATCH) {
        uint32_t u[BATCH];
        u[0] = list[k0 * 64];
#pragma unroll
        for (uint32_t j = 1; j < BATCH; j++) u[j] = k0 + j < cnt ? list[(k0 + j) * 64] : u[0];
        double2 xy[BATCH], zr[BATCH], ra[BATCH];
#pragma unroll
        for (uint32_t j = 0; j < BATCH; j++) {
            xy[j] = *reinterpret_cast<const double2*>(&rec[u[j]].x);
            zr[j] = *reinterpret_cast<const double2*>(&rec[u[j]].z);
            ra[j] = *reinterpret_cast<const double2*>(&rec[u[j]].rep);
        }
#pragma unroll
        for (uint32_t j = 0; j < BATCH; j++)
            if (k0 + j < cnt)
                force_pair64(xy[j], zr[j], ra[j], px, py, pz, radius_i, rep_i, adh_i, rel_max_adh, relf, cut2, fx, fy,
                             fz, local_pressure);
    }
}

// m: the agent's own pos32 entry; rmax: the largest radius of any agent,
// rounded up to FP32; list: this lane's first entry in LDS.  The FP32 test
// uses the agent's largest interaction radius, (r_i + rmax) * max(1, rel),
// for every candidate: no smaller than the pair's, so no pair is lost, and
// the same for all candidates, so the test is a comparison with one
// threshold.  pos32 is padded by MARKX_STEP entries: a step loads past the
// end of its range and masks the extra entries.
__device__ __attribute__((always_inline)) inline void force_agent64x(
    const CellRec* __restrict__ rec, const float4* __restrict__ pos32, const uint32_t* __restrict__ offsets,
    uint32_t* __restrict__ list, uint32_t bin0, uint32_t vi, uint32_t vj, uint32_t vk, uint32_t mnx, uint32_t mny,
    uint32_t mnz, float4 m, float voxel, float rmax, double px, double py, double pz, double radius_i, double rep_i,
    double adh_i, double rel_max_adh, double cut2, double& fx, double& fy, double& fz, double& local_pressure) {
    const float relf32 = (float)(rel_max_adh > 1.0 ? rel_max_adh : 1.0);
    const float L = ((float)radius_i + rmax) * relf32;
    const float lim = fminf(L * L * (1.0f + 0x1p-16f), (float)cut2) + MARKX_SLACK;
    // w: the next list entry (an LDS address: 32-bit arithmetic); every
    // candidate is written there and w moves on past the marked ones
    lds_u32* const l0 = (lds_u32*)list;
    lds_u32* w = l0;
    lds_u32* const wflush = l0 + (MARKX_CAP - MARKX_STEP) * 64;
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
        const RowRange rr = neighbor_row(offsets, bin0, vi, vj, vk, mnx, mny, mnz, q);
        const float my = m.y - (float)rr.dj * voxel, mz = m.z - (float)rr.dk * voxel;
        for (uint32_t u0 = rr.beg; u0 < rr.end; u0 += MARKX_STEP) {
            const float4* p = pos32 + u0;
            const uint32_t rem = rr.end - u0;
            float4 o[MARKX_STEP];
#pragma unroll
            for (uint32_t j = 0; j < MARKX_STEP; j++) o[j] = p[j];
#pragma unroll
            for (uint32_t j = 0; j < MARKX_STEP; j++) {
                const float ddx = m.x - o[j].x, ddy = my - o[j].y, ddz = mz - o[j].z;
                const float d2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddy, ddy, ddx * ddx));
                *w = u0 + j;
                w += (j < rem && d2 < lim) ? 64 : 0;
            }
            if (w > wflush) {
                force_list64(rec, list, (uint32_t)(w - l0) / 64u, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh,
                             cut2, fx, fy, fz, local_pressure);
                w = l0;
            }
        }
    }
    force_list64(rec, list, (uint32_t)(w - l0) / 64u, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh, cut2, fx,
                 fy, fz, local_pressure);
}

// ---------------------------------------------------------------------------
// FP32 forces (mixed form)
// ---------------------------------------------------------------------------
// record position bits of a mark-list entry (the row code above them); mark
// step; margin of the walked x interval beyond the mark's reach (um)
constexpr uint32_t MARK32_POS_BITS = 27;
constexpr uint32_t MARK32_STEP = 8;           // candidates per step (their loads in flight together)
constexpr float XTRIM_MARGIN = 0.05f;
// The force law of one FP32 pair (shaders/mechanics.metal:164-204 in FP32):
// o = the candidate's record, (ox, oy, oz) its bin corner relative to the
// agent's, (mx, my, mz) the agent's bin-relative position.  Exact
// predicates first; the hardware square root and reciprocal (about 1 ulp)
// instead of the correctly rounded sequences.
__device__ __attribute__((always_inline)) inline void force_pair32(
    const float4 a, const float4 b, float ox, float oy, float oz, float mx, float my, float mz, float radius_i,
    float rep_i, float adh_i, float rel_max_adh, float relf, float cut2, float& fx, float& fy, float& fz,
    float& local_pressure) {
    constexpr float margin = 1.0f + 0x1p-16f;
    // a = (xg, y, z, x), b = (r, rep, adh, idx)
    const float ddx = mx - a.w - ox, ddy = my - a.y - oy, ddz = mz - a.z - oz;
    const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
    const float L = (radius_i + b.x) * relf;
    if (d2 >= cut2 || d2 < 1e-12f || !(d2 < L * L * margin)) return;
    const float dist = __builtin_amdgcn_sqrtf(d2);
    const float inv_dist = __builtin_amdgcn_rcpf(dist);
    const float R_sum = radius_i + b.x;
    float force_mag = 0.f;
    if (dist < R_sum) {
        const float overlap = 1.0f - dist * __builtin_amdgcn_rcpf(R_sum);
        force_mag += overlap * overlap * __builtin_amdgcn_sqrtf(rep_i * b.y);
        local_pressure += overlap * overlap * (1.0f / 0.027288820670331f);
    }
    const float adhesion_distance = rel_max_adh * R_sum;
    if (dist < adhesion_distance) {
        const float t2 = 1.0f - dist * __builtin_amdgcn_rcpf(adhesion_distance);
        force_mag -= t2 * t2 * __builtin_amdgcn_sqrtf(adh_i * b.z);
    }
    fx += force_mag * (ddx * inv_dist);
    fy += force_mag * (ddy * inv_dist);
    fz += force_mag * (ddz * inv_dist);
}

// FP32 forces of one agent over one row of bins (the chunked pass): every
// member of [beg, end) in storage order.  A member's bin corner relative to
// the agent's is (di, dj, dk) * voxel: dj, dk are the row's (oy, oz), and
// di = -1, 0, +1 for members before mid0 (the first record of the agent's
// own column), before mid1 (the first of the next column), and after.  A
// pair outside the interaction range adds +0 to the sums, which leaves them
// unchanged (they start at +0, so they are never -0): the sums equal those
// over the pairs the mark pass lists, in the same order.
__device__ __attribute__((always_inline)) inline void force_range32(
    const CellRec32* __restrict__ rec, uint32_t beg, uint32_t end, uint32_t t, uint32_t mid0, uint32_t mid1, float mx,
    float my, float mz, float oy, float oz, float voxel, float radius_i, float rep_i, float adh_i, float rel_max_adh,
    float cut2, float& fx, float& fy, float& fz, float& local_pressure) {
    const float relf = rel_max_adh > 1.0f ? rel_max_adh : 1.0f;
    for (uint32_t u = beg; u < end; u++) {
        if (u == t) continue;
        const float4 a = reinterpret_cast<const float4*>(rec + u)[0];
        const float4 b = reinterpret_cast<const float4*>(rec + u)[1];
        const float ox = (float)(u < mid0 ? -1 : (u < mid1 ? 0 : 1)) * voxel;
        force_pair32(a, b, ox, oy, oz, mx, my, mz, radius_i, rep_i, adh_i, rel_max_adh, relf, cut2, fx, fy, fz,
                     local_pressure);
    }
}

// The law over a lane's list of marked candidates, in list order (loading
// a batch of records together, as force_list64 does, measured slower
// here).  Entry:
// the record position in the low MARK32_POS_BITS bits and the row code
// (dk + 1) * 3 + (dj + 1) above them.  The candidate's column comes from its
// record: xg - x is its bin corner's distance from x_min, a whole number of
// bins (to within 3e-4 um while the grid's x extent is at most
// MARKX_MAX_EXTENT, far below half a bin).
__device__ __attribute__((always_inline)) inline void force_list32(
    const CellRec32* __restrict__ rec, const uint32_t* __restrict__ list, uint32_t cnt, uint32_t t, int vi, float mx,
    float my, float mz, float voxel, float inv_voxel, float radius_i, float rep_i, float adh_i, float rel_max_adh,
    float cut2, float& fx, float& fy, float& fz, float& local_pressure) {
    const float relf = rel_max_adh > 1.0f ? rel_max_adh : 1.0f;
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t e = list[k * 64];
        const uint32_t u = e & ((1u << MARK32_POS_BITS) - 1u), code = e >> MARK32_POS_BITS;
        if (u == t) continue;
        const float4 a = reinterpret_cast<const float4*>(rec + u)[0];   // xg, y, z, x
        const float4 b = reinterpret_cast<const float4*>(rec + u)[1];   // r, rep, adh, idx
        const int di = (int)rintf((a.x - a.w) * inv_voxel) - vi;
        const float ox = (float)di * voxel;
        const float oy = (float)((int)(code % 3u) - 1) * voxel, oz = (float)((int)(code / 3u) - 1) * voxel;
        force_pair32(a, b, ox, oy, oz, mx, my, mz, radius_i, rep_i, adh_i, rel_max_adh, relf, cut2, fx, fy, fz,
                     local_pressure);
    }
}

// 8-byte mark entries (force_agent64t, and force_agent32x<true>): x - x_min
// in 1/16 um, y and z relative to the bin's lower corner plus 64 um in
// 1/512 um, and a signed 16-bit displacement from the entry's position to
// the member's record position (0 where the records are in the entries'
// order).  ok is false when a member does not fit.
constexpr float MARK16_SLACK = 2.25f;
constexpr uint32_t MARK16_STEP = 8;   // entries per step (their loads in flight together)
constexpr uint32_t LAW16_BATCH = 2;   // records per law batch (fewer registers: 5 waves per SIMD)
__host__ __device__ inline uint2 mark_entry16(double xg, double yb, double zb, int disp, bool& ok) {
    const double qx = rint(xg * 16.0), qy = rint((yb + 64.0) * 512.0), qz = rint((zb + 64.0) * 512.0);
    ok = qx >= 0.0 && qx <= 65535.0 && qy >= 0.0 && qy <= 65535.0 && qz >= 0.0 && qz <= 65535.0 &&
         disp >= -32768 && disp <= 32767;
    if (!ok) return make_uint2(0u, 0u);
    return make_uint2((uint32_t)qx | ((uint32_t)qy << 16), (uint32_t)qz | ((uint32_t)disp << 16));
}

// FP32 forces of one agent with a mark pass over x-trimmed rows and a
// per-lane list of the marked candidates in LDS (the mixed form; the chunked
// pass's result bit for bit).
//
// Mark: one threshold per agent, lim = min(L^2 (1 + 2^-16), cut^2) +
// MARKX_SLACK with L = (r_i + rmax) * max(1, rel) in FP32, rmax the largest
// FP32 radius of any agent: FP32 addition and multiplication are monotonic,
// so lim is no smaller than any pair's bound in force_pair32.  The mark's
// d^2 uses xg (x relative to x_min: within 1.2e-4 um while the grid's x
// extent is at most MARKX_MAX_EXTENT) and the bin-relative y, z; for pairs
// within 32 um it is within 0.025 um^2 of the law's d^2, below the slack, so
// every pair the law keeps is marked.
//
// Rows: a candidate marked has |dx| < sqrt(lim) + 3e-4 um, so only the x
// interval [mx - Lx, mx + Lx], Lx = sqrt(lim) + XTRIM_MARGIN, of each row is
// walked: from the slab holding mx - Lx (in the column left of the agent's
// when the interval starts there) to the slab holding mx + Lx.  Slabs come
// from the same FP32 products and floors as the records' (xslab_of, a
// monotonic function of x), so no record with x inside the interval falls
// outside the range.  The range is the same for all 9 rows and is one
// contiguous run of records per row (two loads).
//
// A marked candidate is listed with its row code, and the law walks the
// list in visit order (rows, ascending record position) with the exact
// predicates, so the pairs that contribute, their order and every bit of the
// sums are the chunked pass's.  Record positions stay below 2^27 (callers
// check).

//
// M16: the mark reads the 8-byte entries in mark[] (the records' order)
// instead of the records' first 16 bytes, with MARK16_SLACK (the entries'
// quantisation bound of force_agent64t plus the FP32 law's own rounding,
// below 0.01 um^2 for pairs within 32 um).
template <bool M16 = false>
__device__ __attribute__((always_inline)) inline void force_agent32x(
    const CellRec32* __restrict__ rec, const uint32_t* __restrict__ xoff, uint32_t* __restrict__ list, uint32_t t,
    uint32_t bin0, uint32_t vi, uint32_t vj, uint32_t vk, uint32_t mnx, uint32_t mny, uint32_t mnz, float mxg,
    float mx, float my, float mz, float voxel, float scale, float inv_voxel, float rmax, float radius_i, float rep_i,
    float adh_i, float rel_max_adh, float cut2, float& fx, float& fy, float& fz, float& local_pressure,
    const uint2* __restrict__ mark = nullptr) {
    const float relf = rel_max_adh > 1.0f ? rel_max_adh : 1.0f;
    const float L = (radius_i + rmax) * relf;
    const float lim = fminf(L * L * (1.0f + 0x1p-16f), cut2) + (M16 ? MARK16_SLACK : MARKX_SLACK);
    const float Lx = sqrtf(lim) + XTRIM_MARGIN;
    // the walked x interval, as slab indices of the row (column vi - 1 .. vi + 1)
    const float lo = mx - Lx, hi = mx + Lx;
    const bool left = lo < 0.0f && vi > 0, right = hi >= voxel && vi + 1 < mnx;
    const uint32_t jlo = XSLABS * (left ? vi - 1 : vi) + xslab_of(left ? lo + voxel : lo, scale);
    const uint32_t jhi = XSLABS * (right ? vi + 1 : vi) + xslab_of(right ? hi - voxel : hi, scale) + 1;
    const bool three_d = mnz > 1;
    // w: the next list entry (an LDS address: 32-bit arithmetic); every
    // candidate is written there and w moves on past the marked ones
    lds_u32* const l0 = (lds_u32*)list;
    lds_u32* w = l0;
    lds_u32* const wflush = l0 + (MARKX_CAP - MARK32_STEP) * 64;
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
        const int dk = three_d ? q / 3 - 1 : 0, dj = q % 3 - 1;
        const int nk = (int)vk + dk, nj = (int)vj + dj;
        const bool ok = (three_d || q < 3) && nk >= 0 && nk < (int)mnz && nj >= 0 && nj < (int)mny;
        const uint32_t rb = XSLABS * (bin0 + (ok ? (uint32_t)nk * mny * mnx + (uint32_t)nj * mnx : 0u));
        const uint32_t beg = ok ? xoff[rb + jlo] : 0u, end = ok ? xoff[rb + jhi] : 0u;
        const float myr = my - (float)dj * voxel, mzr = mz - (float)dk * voxel;
        const uint32_t qc = (uint32_t)((dk + 1) * 3 + (dj + 1)) << MARK32_POS_BITS;
        const float myr64 = myr + 64.0f, mzr64 = mzr + 64.0f;
        for (uint32_t u0 = beg; u0 < end; u0 += MARK32_STEP) {
            // entries u0 + j with the row code above them (u0 + j < 2^27)
            const uint32_t eb = u0 + qc, rem = end - u0;
            float d2[MARK32_STEP];
            if (M16) {
                uint2 o[MARK32_STEP];
#pragma unroll
                for (uint32_t j = 0; j < MARK32_STEP; j++) o[j] = mark[u0 + j];
#pragma unroll
                for (uint32_t j = 0; j < MARK32_STEP; j++) {
                    const float ddx = __builtin_fmaf((float)(o[j].x & 0xffffu), -0.0625f, mxg);
                    const float ddy = __builtin_fmaf((float)(o[j].x >> 16), -0x1p-9f, myr64);
                    const float ddz = __builtin_fmaf((float)(o[j].y & 0xffffu), -0x1p-9f, mzr64);
                    d2[j] = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddy, ddy, ddx * ddx));
                }
            } else {
                float4 o[MARK32_STEP];
#pragma unroll
                for (uint32_t j = 0; j < MARK32_STEP; j++) o[j] = *reinterpret_cast<const float4*>(rec + u0 + j);
#pragma unroll
                for (uint32_t j = 0; j < MARK32_STEP; j++) {
                    const float ddx = mxg - o[j].x, ddy = myr - o[j].y, ddz = mzr - o[j].z;
                    d2[j] = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddy, ddy, ddx * ddx));
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < MARK32_STEP; j++) {
                *w = eb + j;
                w += (j < rem && d2[j] < lim) ? 64 : 0;
            }
            if (w > wflush) {
                force_list32(rec, list, (uint32_t)(w - l0) / 64u, t, (int)vi, mx, my, mz, voxel, inv_voxel,
                             radius_i, rep_i, adh_i, rel_max_adh, cut2, fx, fy, fz, local_pressure);
                w = l0;
            }
        }
    }
    force_list32(rec, list, (uint32_t)(w - l0) / 64u, t, (int)vi, mx, my, mz, voxel, inv_voxel, radius_i, rep_i,
                 adh_i, rel_max_adh, cut2, fx, fy, fz, local_pressure);
}

// FP64 forces of one agent with an x-trimmed mark pass over 8-byte mark
// entries.  The records stay in (bin, index) order; the mark entries hold the
// same members per bin in (x slab, index) order (slab starts in xoff, as for
// the FP32 records), each a uint2 (MarkEntry16):
//   x: x - x_min in 1/16 um (16 bits) | y - the bin's lower y in 1/512 um,
//      plus 64 um (16 bits);
//   y: z as y (16 bits) | record position - mark position (signed 16 bits).
// The gather sets *flag when a member does not fit (then every agent takes
// the exact pass).  Quantisation moves x by at most 1/32 um and y, z by at
// most 1/1024 um, the agent's FP32 coordinates and the FMA differences add
// below 2.5e-4 um, so for candidates within 32 um the mark's d^2 is within
// 2.14 um^2 of the exact one: the threshold carries MARK16_SLACK = 2.25 um^2
// above force_agent64x's bound, and every pair the law keeps is marked.  A
// marked candidate has |dx| < sqrt(lim) + 0.032 um, inside the walked x
// interval (force_agent32x's, with XTRIM_MARGIN = 0.05 um).
//
// A marked candidate sets the bit of its record position relative to a
// 64-record window of the row (the row's first record, then 64 on: a row of
// more than 64 records takes one walk per window), and the set bits go to
// the LDS list in ascending order: the law sees the marked candidates in the
// reference's visit order (rows, then ascending record position) without a
// sort, re-tests each with the exact FP64 predicates, and the sums are
// force_agent64x's bit for bit.

// mxg: the agent's x - x_min; my64, mz64: its y, z relative to its bin's
// lower corner plus 64 um (FP32 of the FP64 values); mxb: its bin-relative x
// (bin_rel32, the slabs' expression)
__device__ __attribute__((always_inline)) inline void force_agent64t(
    const CellRec* __restrict__ rec, const uint2* __restrict__ mark, const uint32_t* __restrict__ xoff,
    uint32_t* __restrict__ list, uint32_t vi, uint32_t vj, uint32_t vk, uint32_t mnx, uint32_t mny, uint32_t mnz,
    float mxg, float mxb, float my64, float mz64, float voxel, float scale, float rmax, double px, double py,
    double pz, double radius_i, double rep_i, double adh_i, double rel_max_adh, double cut2, double& fx, double& fy,
    double& fz, double& local_pressure) {
    const float relf32 = (float)(rel_max_adh > 1.0 ? rel_max_adh : 1.0);
    const float L = ((float)radius_i + rmax) * relf32;
    const float lim = fminf(L * L * (1.0f + 0x1p-16f), (float)cut2) + MARK16_SLACK;
    const float Lx = sqrtf(lim) + XTRIM_MARGIN;
    const float lo = mxb - Lx, hi = mxb + Lx;
    const bool left = lo < 0.0f && vi > 0, right = hi >= voxel && vi + 1 < mnx;
    const uint32_t blo = left ? vi - 1 : vi, bhi = right ? vi + 1 : vi;
    const uint32_t jlo = XSLABS * blo + xslab_of(left ? lo + voxel : lo, scale);
    const uint32_t jhi = XSLABS * bhi + xslab_of(right ? hi - voxel : hi, scale) + 1;
    const bool three_d = mnz > 1;
    lds_u32* const l0 = (lds_u32*)list;
    lds_u32* w = l0;
    lds_u32* const wend = l0 + MARKX_CAP * 64;
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
        const int dk = three_d ? q / 3 - 1 : 0, dj = q % 3 - 1;
        const int nk = (int)vk + dk, nj = (int)vj + dj;
        const bool ok = (three_d || q < 3) && nk >= 0 && nk < (int)mnz && nj >= 0 && nj < (int)mny;
        const uint32_t rb = XSLABS * (ok ? (uint32_t)nk * mny * mnx + (uint32_t)nj * mnx : 0u);
        // walked slab range [sb, se); records of the walked bins [r0, r1)
        const uint32_t sb = ok ? xoff[rb + jlo] : 0u, se = ok ? xoff[rb + jhi] : 0u;
        const uint32_t r0 = ok ? xoff[rb + XSLABS * blo] : 0u, r1 = ok ? xoff[rb + XSLABS * (bhi + 1)] : 0u;
        const float myr = my64 - (float)dj * voxel, mzr = mz64 - (float)dk * voxel;
        for (uint32_t wb = r0; sb < se && wb < r1; wb += 64) {
            uint64_t mask = 0;
            for (uint32_t u0 = sb; u0 < se; u0 += MARK16_STEP) {
                const uint2* p = mark + u0;
                const uint32_t rem = se - u0, base = u0 - wb;
                uint2 o[MARK16_STEP];
#pragma unroll
                for (uint32_t j = 0; j < MARK16_STEP; j++) o[j] = p[j];
#pragma unroll
                for (uint32_t j = 0; j < MARK16_STEP; j++) {
                    const float ddx = __builtin_fmaf((float)(o[j].x & 0xffffu), -0.0625f, mxg);
                    const float ddy = __builtin_fmaf((float)(o[j].x >> 16), -0x1p-9f, myr);
                    const float ddz = __builtin_fmaf((float)(o[j].y & 0xffffu), -0x1p-9f, mzr);
                    const float d2 = __builtin_fmaf(ddz, ddz, __builtin_fmaf(ddy, ddy, ddx * ddx));
                    const uint32_t rp = base + j + (uint32_t)((int32_t)o[j].y >> 16);
                    mask |= (j < rem && d2 < lim && rp < 64u) ? (1ull << (rp & 63u)) : 0ull;
                }
            }
            while (mask) {
                if (w == wend) {
                    force_list64<LAW16_BATCH>(rec, list, MARKX_CAP, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh,
                                              cut2, fx, fy, fz, local_pressure);
                    w = l0;
                }
                *w = wb + (uint32_t)__builtin_ctzll(mask);
                w += 64;
                mask &= mask - 1;
            }
        }
    }
    force_list64<LAW16_BATCH>(rec, list, (uint32_t)(w - l0) / 64u, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh,
                              cut2, fx, fy, fz, local_pressure);
}

// The chunked pass over the 9 (3 in 2-D) neighbour rows of an agent.
__device__ __attribute__((always_inline)) inline void force_rows32(
    const CellRec32* __restrict__ rec, const uint32_t* __restrict__ offsets, uint32_t t, uint32_t bin0, uint32_t vi,
    uint32_t vj, uint32_t vk, uint32_t mnx, uint32_t mny, uint32_t mnz, float mx, float my, float mz, float voxel,
    float radius_i, float rep_i, float adh_i, float rel_max_adh, float cut2, float& fx, float& fy, float& fz,
    float& local_pressure) {
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
        const RowRange rr = neighbor_row(offsets, bin0, vi, vj, vk, mnx, mny, mnz, q);
        force_range32(rec, rr.beg, rr.end, t, rr.mid0, rr.mid1, mx, my, mz, (float)rr.dj * voxel,
                      (float)rr.dk * voxel, voxel, radius_i, rep_i, adh_i, rel_max_adh, cut2, fx, fy, fz,
                      local_pressure);
    }
}

// Bins of at most BIN_SORT_CAP members: the hash's ascending list (as
// bin_sort_one) and each member's FP32 record position by (x slab, index),
// into rpos; every bin: its slab starts into xoff (xoff[XSLABS * b + s]).
// Larger bins leave their members' positions to the gather (slab_rank).
// Thread nbins writes the XSLABS entries past the last bin.
// rec_pos non-null (the plain form's 8-byte mark entries): tmp holds
// (agent << 2 | x slab) (k_scatter_slab; sorting these values sorts by
// agent), rpos receives the entry position by (x slab, index), rec_pos the
// record position by index, and thread nbins clears *xtflag.
__device__ inline void bin_sort_slabs(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ tmp,
                                      const uint8_t* __restrict__ xslab, uint32_t* __restrict__ sorted,
                                      uint32_t* __restrict__ rpos, uint32_t* __restrict__ xoff, uint32_t b,
                                      uint32_t nbins, uint32_t* __restrict__ rec_pos = nullptr,
                                      uint32_t* __restrict__ xtflag = nullptr) {
    if (b == nbins) {
        const uint32_t n = offsets[nbins];
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * nbins) = make_uint4(n, n, n, n);
        if (xtflag) xtflag[0] = 0u;
        return;
    }
    const uint32_t beg = offsets[b], cnt = offsets[b + 1] - beg;
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    if (cnt == 0) {
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) = make_uint4(beg, beg, beg, beg);
        return;
    }
    if (cnt > BIN_SORT_CAP) {
        for (uint32_t p = beg; p < beg + cnt; p++) {
            const uint32_t s = rec_pos ? (tmp[p] & 3u) : xslab[tmp[p]];
            c0 += s == 0 ? 1u : 0u;
            c1 += s == 1 ? 1u : 0u;
            c2 += s == 2 ? 1u : 0u;
        }
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) =
            make_uint4(beg, beg + c0, beg + c0 + c1, beg + c0 + c1 + c2);
        return;
    }
    uint32_t a[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) a[q] = q < cnt ? tmp[beg + q] : ~0u;
#pragma unroll
    for (uint32_t k = 2; k <= BIN_SORT_CAP; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t i = 0; i < BIN_SORT_CAP; i++) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint32_t lo = a[i] < a[l] ? a[i] : a[l], hi = a[i] < a[l] ? a[l] : a[i];
                    const bool up = (i & k) == 0;
                    a[i] = up ? lo : hi;
                    a[l] = up ? hi : lo;
                }
            }
    uint32_t s[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) s[q] = q < cnt ? (rec_pos ? (a[q] & 3u) : xslab[a[q]]) : XSLABS;
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) {
        c0 += s[q] == 0 ? 1u : 0u;
        c1 += s[q] == 1 ? 1u : 0u;
        c2 += s[q] == 2 ? 1u : 0u;
    }
    const uint32_t x1 = beg + c0, x2 = x1 + c1, x3 = x2 + c2;
    *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) = make_uint4(beg, x1, x2, x3);
    // members in ascending index order: each takes the next position of its slab
    uint32_t n0 = beg, n1 = x1, n2 = x2, n3 = x3;
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++)
        if (q < cnt) {
            const uint32_t ag = rec_pos ? a[q] >> 2 : a[q];
            sorted[beg + q] = ag;
            if (rec_pos) rec_pos[ag] = beg + q;
            const uint32_t sq = s[q];
            rpos[ag] = sq == 0 ? n0 : (sq == 1 ? n1 : (sq == 2 ? n2 : n3));
            n0 += sq == 0 ? 1u : 0u;
            n1 += sq == 1 ? 1u : 0u;
            n2 += sq == 2 ? 1u : 0u;
            n3 += sq == 3 ? 1u : 0u;
        }
}

// FP32 record position of member c (slab sc) of a bin of more than
// BIN_SORT_CAP members: the members of a smaller slab, or of the same slab
// and a smaller index, come first.
__device__ inline uint32_t slab_rank(const uint32_t* __restrict__ tmp, const uint8_t* __restrict__ xslab,
                                     uint32_t beg, uint32_t end, uint32_t c, uint32_t sc) {
    uint32_t r = 0;
    for (uint32_t t0 = beg; t0 < end; t0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = tmp[t0 + q < end ? t0 + q : end - 1];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const uint32_t sv = xslab[v[q]];
            r += (t0 + q < end && (sv < sc || (sv == sc && v[q] < c))) ? 1u : 0u;
        }
    }
    return r;
}

}  // namespace hc
