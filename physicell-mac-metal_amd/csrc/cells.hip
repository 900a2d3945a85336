// cells.hip -- agent-side kernels for gfx950 (FP64): voxel keys, secretion,
// the spatial hash as a counting sort, repulsion/adhesion forces and AB2.
//
// Hash (replaces the atomic 64-slot buckets of shaders/mechanics.metal:27-85):
//   k_mech_keys      key = clamp(trunc((p - min)/30)) per agent, written to
//                    field 27; the histogram is built with wave-aggregated
//                    atomics (lanes holding the same key are found with
//                    ballots; one atomic per distinct key per wave) and each
//                    lane gets its slot inside the bin.
//   scan             exclusive prefix sum of the bin counts (wavefront scans).
//   k_scatter        provisional placement: list[offset[key] + slot] = agent.
//   k_stable_gather  rank = number of agents in the same bin with a smaller
//                    index; writes the ascending-index list and a 64-byte
//                    record (position, radius, repulsion, adhesion,
//                    relative adhesion distance, index, key) per agent in
//                    hash-sorted order, so the force kernel reads neighbours
//                    from contiguous memory.
// Forces (shaders/mechanics.metal:91-216) run one agent per lane in
// hash-sorted order and visit the 27 (9 in 2-D) bins in the reference's
// (dk, dj, di) order, agents of a bin in ascending index.
#include <map>
#include <mutex>

#include "cells_common.h"

namespace hc {



// ---------------------------------------------------------------------------
// exclusive scan (3 kernels): 1024 values per block of 256 threads
// ---------------------------------------------------------------------------
static constexpr int SCAN_BS = 256, SCAN_PER = 4, SCAN_TILE = SCAN_BS * SCAN_PER;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[SCAN_BS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < SCAN_BS / 64; k++) {
        if (k < w) off += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_local(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out, uint64_t n,
                                                        uint32_t* __restrict__ partial) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER;
    uint32_t v[SCAN_PER];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_partials(uint32_t* __restrict__ partial, uint32_t nb) {
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += SCAN_BS) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < nb ? partial[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, &total);
        if (i < nb) partial[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_add(uint32_t* __restrict__ out, uint64_t n,
                                                      const uint32_t* __restrict__ partial, uint32_t nb) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER;
    const uint32_t add = partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++)
        if (base + k < n) out[base + k] += add;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partial[nb];
}

// Two-kernel scan for up to SCAN2_MAX_TILES tiles of 4096 values: tile sums,
// then each tile adds the sum of the tiles before it (read by the whole
// block from the partial sums) to its own exclusive scan.  Counts are read
// twice and offsets written once, with no serial pass over the partials.
static constexpr int SCAN2_PER = 16, SCAN2_TILE = SCAN_BS * SCAN2_PER;
static constexpr uint32_t SCAN2_MAX_TILES = 2048;

__device__ inline uint32_t block_sum(uint32_t v) {
    __shared__ uint32_t ws[SCAN_BS / 64];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
#pragma unroll
    for (int k = 0; k < SCAN_BS / 64; k++) t += ws[k];
    __syncthreads();
    return t;
}

__device__ inline void scan2_load(const uint32_t* in, uint64_t base, uint64_t n, uint32_t* v) {
    if (base + SCAN2_PER <= n) {
#pragma unroll
        for (int q = 0; q < SCAN2_PER / 4; q++) {
            const uint4 w = *reinterpret_cast<const uint4*>(in + base + 4 * q);
            v[4 * q] = w.x;
            v[4 * q + 1] = w.y;
            v[4 * q + 2] = w.z;
            v[4 * q + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN2_PER; k++) v[k] = base + k < n ? in[base + k] : 0u;
    }
}

__global__ __launch_bounds__(SCAN_BS) void k_scan2_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                          uint32_t* __restrict__ partial) {
    uint32_t v[SCAN2_PER];
    scan2_load(in, (uint64_t)blockIdx.x * SCAN2_TILE + (uint64_t)threadIdx.x * SCAN2_PER, n, v);
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN2_PER; k++) sum += v[k];
    const uint32_t tot = block_sum(sum);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// in and out may be the same array (in-place scans); clear: null, or in's
// address when out is another array (each thread zeroes what it read)
__global__ __launch_bounds__(SCAN_BS) void k_scan2_down(const uint32_t* in, uint32_t* out,
                                                        uint64_t n, const uint32_t* __restrict__ partial,
                                                        uint32_t nb, uint32_t* clear) {
    uint32_t pre = 0;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += SCAN_BS) pre += partial[i];
    const uint32_t tile_base = block_sum(pre);
    const uint64_t base = (uint64_t)blockIdx.x * SCAN2_TILE + (uint64_t)threadIdx.x * SCAN2_PER;
    uint32_t v[SCAN2_PER];
    scan2_load(in, base, n, v);
    if (clear) {
        if (base + SCAN2_PER <= n) {
#pragma unroll
            for (int q = 0; q < SCAN2_PER / 4; q++)
                *reinterpret_cast<uint4*>(clear + base + 4 * q) = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (int k = 0; k < SCAN2_PER; k++)
                if (base + k < n) clear[base + k] = 0u;
        }
    }
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN2_PER; k++) sum += v[k];
    uint32_t total;
    uint32_t run = tile_base + block_exclusive_scan(sum, &total);
    if (base + SCAN2_PER <= n) {
#pragma unroll
        for (int q = 0; q < SCAN2_PER / 4; q++) {
            uint4 w;
            w.x = run; run += v[4 * q];
            w.y = run; run += v[4 * q + 1];
            w.z = run; run += v[4 * q + 2];
            w.w = run; run += v[4 * q + 3];
            *reinterpret_cast<uint4*>(out + base + 4 * q) = w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < SCAN2_PER; k++) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = tile_base + total;
}

size_t scan_partial_bytes(uint64_t n) {
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(nb + 1) * sizeof(uint32_t);
}

// in and out hold n (+1 for out) values from 16-byte aligned addresses
int scan_exclusive(hc_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partial,
                   uint32_t* clear_in) {
    const uint64_t nb2 = (n + SCAN2_TILE - 1) / SCAN2_TILE;
    if (nb2 > 0 && nb2 <= SCAN2_MAX_TILES && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0) {
        hipLaunchKernelGGL(k_scan2_reduce, dim3((uint32_t)nb2), dim3(SCAN_BS), 0, ctx->stream, in, n, partial);
        hipLaunchKernelGGL(k_scan2_down, dim3((uint32_t)nb2), dim3(SCAN_BS), 0, ctx->stream, in, out, n, partial,
                           (uint32_t)nb2, clear_in);
        return check_launch(ctx, "scan");
    }
    if (clear_in) {
        const int r = scan_exclusive(ctx, in, out, n, partial);
        return r ? r : fill_u32(ctx, clear_in, n, 0u);
    }
    const uint32_t nb = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
    if (nb == 0) {
        hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_BS), 0, ctx->stream, partial, 0u);
        hipLaunchKernelGGL(k_scan_add, dim3(1), dim3(SCAN_BS), 0, ctx->stream, out, (uint64_t)0, partial, 0u);
        return check_launch(ctx, "scan(empty)");
    }
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(SCAN_BS), 0, ctx->stream, in, out, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SCAN_BS), 0, ctx->stream, partial, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_BS), 0, ctx->stream, out, n, partial, nb);
    return check_launch(ctx, "scan");
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int fill_u32(hc_ctx* ctx, uint32_t* p, uint64_t n, uint32_t v) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, p, n, v);
    return check_launch(ctx, "fill");
}

// ---------------------------------------------------------------------------
// cell -> diffusion voxel (src/microenvironment.mm:176-184)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_voxel_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, uint32_t* __restrict__ out,
                                                    uint32_t n, hc_grid g) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    int i = (int)((x[c] - g.x_min) / g.dx);
    int j = (int)((y[c] - g.y_min) / g.dy);
    int k = (int)((z[c] - g.z_min) / g.dz);
    i = (i < 0) ? 0 : (i >= (int)g.nx ? (int)g.nx - 1 : i);
    j = (j < 0) ? 0 : (j >= (int)g.ny ? (int)g.ny - 1 : j);
    k = (k < 0) ? 0 : (k >= (int)g.nz ? (int)g.nz - 1 : k);
    out[c] = (uint32_t)k * g.nx * g.ny + (uint32_t)j * g.nx + (uint32_t)i;
}

// ---------------------------------------------------------------------------
// mechanics hash
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mech_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, uint32_t* __restrict__ key_out,
                                                   uint32_t* __restrict__ counts, uint32_t* __restrict__ slot,
                                                   uint32_t n, double x_min, double y_min, double z_min,
                                                   double voxel, uint32_t mnx, uint32_t mny, uint32_t mnz,
                                                   uint8_t* __restrict__ xslab) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    const bool valid = c < n;
    uint32_t key = 0;
    if (valid) {
        // shaders/mechanics.metal:61-72
        const double px = x[c];
        int vi = (int)((px - x_min) / voxel);
        int vj = (int)((y[c] - y_min) / voxel);
        int vk = (int)((z[c] - z_min) / voxel);
        vi = vi < 0 ? 0 : (vi > (int)mnx - 1 ? (int)mnx - 1 : vi);
        vj = vj < 0 ? 0 : (vj > (int)mny - 1 ? (int)mny - 1 : vj);
        vk = vk < 0 ? 0 : (vk > (int)mnz - 1 ? (int)mnz - 1 : vk);
        key = (uint32_t)vk * mny * mnx + (uint32_t)vj * mnx + (uint32_t)vi;
        key_out[c] = key;
        // FP32 records: the x slab of the record's bin-relative x
        if (xslab) xslab[c] = (uint8_t)xslab_of(bin_rel32(px, x_min, (uint32_t)vi, voxel), xslab_scale(voxel));
    }
    const uint32_t s = ballot_bin(valid, key, counts);
    if (valid) slot[c] = s;
}

__global__ __launch_bounds__(256) void k_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ slot,
                                                 const uint32_t* __restrict__ offsets, uint32_t* __restrict__ tmp,
                                                 uint32_t n) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    tmp[offsets[key[c]] + slot[c]] = c;
}

// k_scatter with each agent's x slab in the entry's low bits (agent << 2 |
// slab; agents below 2^30), for the plain form's 8-byte mark entries
__global__ __launch_bounds__(256) void k_scatter_slab(const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ slot,
                                                      const uint8_t* __restrict__ xslab,
                                                      const uint32_t* __restrict__ offsets, uint32_t* __restrict__ tmp,
                                                      uint32_t n) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    tmp[offsets[key[c]] + slot[c]] = (c << 2) | xslab[c];
}


// the largest of nb block maxima (one block)
__global__ __launch_bounds__(256) void k_rmax(const float* __restrict__ rpart, uint32_t nb, float* __restrict__ rmax) {
    float r = 0.0f;
    for (uint32_t b = threadIdx.x; b < nb; b += 256) r = fmaxf(r, rpart[b]);
    for (int d = 32; d >= 1; d >>= 1) r = fmaxf(r, __shfl_xor(r, d, 64));
    __shared__ float wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) rmax[0] = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
}

__global__ __launch_bounds__(256) void k_bin_sort(const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                                                  uint32_t* __restrict__ pos_of, uint32_t nbins) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b < nbins) bin_sort_one(offsets, tmp, sorted, pos_of, b);
}

// FP32 records: the ascending lists, record positions (rpos) and slab starts
// (xoff) of bins 0 .. nbins-1, and the entries past the last bin
__global__ __launch_bounds__(256) void k_bin_sort32(const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ tmp,
                                                    const uint8_t* __restrict__ xslab, uint32_t* __restrict__ sorted,
                                                    uint32_t* __restrict__ rpos, uint32_t* __restrict__ xoff,
                                                    uint32_t nbins, uint32_t* __restrict__ rec_pos,
                                                    uint32_t* __restrict__ xtflag) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b <= nbins) bin_sort_slabs(offsets, tmp, xslab, sorted, rpos, xoff, b, nbins, rec_pos, xtflag);
}

__global__ __launch_bounds__(256) void k_stable_gather(CellPtrs cp, const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ tmp,
                                                       uint32_t* __restrict__ sorted, const uint32_t* __restrict__ pos_of,
                                                       CellRec* __restrict__ rec, float4* __restrict__ pos32,
                                                       uint32_t n, double x_min, double y_min, double z_min,
                                                       double voxel, uint32_t mnx, uint32_t mny, int x_grid,
                                                       float* __restrict__ rpart, const uint32_t* __restrict__ spos,
                                                       const uint8_t* __restrict__ xslab,
                                                       uint32_t* __restrict__ xtflag) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    // this block's largest radius, rounded up to FP32 (no lane returns before
    // the barrier)
    if (rpart) block_max_to(c < n ? __double2float_ru(cp.f[HC_F_RADIUS][c]) : 0.0f, rpart);
    if (c >= n) return;
    const uint32_t b = key[c];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos;
    if (end - beg <= BIN_SORT_CAP) {
        pos = pos_of[c];                                  // placed by k_bin_sort (k_bin_sort32)
    } else {
        // spos: tmp holds (agent << 2 | slab), ordered as the agents are
        pos = beg + stable_rank(tmp, beg, end, spos ? c << 2 : c);
        sorted[pos] = c;
    }
    if (spos) {
        if (c == 0) {
            double* org = reinterpret_cast<double*>(xtflag + 2);
            org[0] = x_min;
            org[1] = y_min;
            org[2] = z_min;
        }
        // 8-byte mark entry at its (x slab, index) position (force_agent64t)
        uint32_t sp;
        if (end - beg <= BIN_SORT_CAP) {
            sp = spos[c];
        } else {
            const uint32_t sc = xslab[c];
            sp = beg;
            for (uint32_t u = beg; u < end; u++) {
                const uint32_t v = tmp[u], sv = v & 3u, iv = v >> 2;
                sp += (sv < sc || (sv == sc && iv < c)) ? 1u : 0u;
            }
        }
        const uint32_t bj = (b / mnx) % mny, bk = b / (mnx * mny);
        bool fits;
        reinterpret_cast<uint2*>(pos32)[sp] =
            mark_entry16(cp.f[HC_F_POS_X][c] - x_min, cp.f[HC_F_POS_Y][c] - (y_min + bj * voxel),
                         cp.f[HC_F_POS_Z][c] - (z_min + bk * voxel), (int)pos - (int)sp, fits);
        if (!fits) xtflag[0] = 1u;
    } else if (pos32) {
        // x relative to the grid's x_min (force_agent64x) or to the bin corner
        // (force_range64m); y and z relative to the bin corner
        const uint32_t bi = b % mnx, bj = (b / mnx) % mny, bk = b / (mnx * mny);
        pos32[pos] = make_float4((float)(cp.f[HC_F_POS_X][c] - (x_grid ? x_min : x_min + bi * voxel)),
                                 (float)(cp.f[HC_F_POS_Y][c] - (y_min + bj * voxel)),
                                 (float)(cp.f[HC_F_POS_Z][c] - (z_min + bk * voxel)), (float)cp.f[HC_F_RADIUS][c]);
    }
    CellRec r;
    r.x = cp.f[HC_F_POS_X][c];
    r.y = cp.f[HC_F_POS_Y][c];
    r.z = cp.f[HC_F_POS_Z][c];
    r.r = cp.f[HC_F_RADIUS][c];
    r.rep = cp.f[HC_F_REPULSION][c];
    r.adh = cp.f[HC_F_ADHESION][c];
    r.rel = cp.f[HC_F_REL_MAX_ADH][c];
    r.idx = c;
    r.key = b;
    rec[pos] = r;
}

// ---------------------------------------------------------------------------
// forces (shaders/mechanics.metal:91-216)
// ---------------------------------------------------------------------------

// pos32 is null unless the bin size and the cutoff allow the FP32 mark pass
// (force_range64m); agents more than two bins from their bin corner take the
// FP64 mark pass either way.
// x_grid: pos32 holds x relative to the grid's x_min (force_agent64x, with
// its list of marked candidates in LDS), else relative to the bin corner
// (force_range64m).
// vout non-null: velocity and pressure go to vout[t] in hash-sorted order
// (the hash-ordered mirror), else to the cell fields at the agent's index.
__global__ __launch_bounds__(256) void k_forces(const CellRec* __restrict__ rec, const float4* __restrict__ pos32,
                                                const uint32_t* __restrict__ offsets,
                                                double* __restrict__ vx, double* __restrict__ vy,
                                                double* __restrict__ vz, double* __restrict__ pressure,
                                                uint32_t n, uint32_t mnx, uint32_t mny, uint32_t mnz,
                                                double max_dist, float voxel, int x_grid,
                                                const float* __restrict__ rmax, double4* __restrict__ vout) {
    __shared__ uint32_t marks[4 * MARKX_CAP * 64];
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= n) return;
    const CellRec me = rec[t];
    const double px = me.x, py = me.y, pz = me.z;
    const double radius_i = me.r, rep_i = me.rep, adh_i = me.adh, rel_max_adh = me.rel;
    double fx = 0.0, fy = 0.0, fz = 0.0, local_pressure = 0.0;
    const uint32_t vi = me.key % mnx, vj = (me.key / mnx) % mny, vk = me.key / (mnx * mny);
    const double cut2 = max_dist * max_dist;
    const float4 m = pos32 ? pos32[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float near = 2.0f * voxel;
    const bool yz_near = fabsf(m.y) < near && fabsf(m.z) < near;
    if (pos32 && x_grid && yz_near && m.x > -near && m.x < (float)mnx * voxel + near) {
        uint32_t* list = marks + (threadIdx.x >> 6) * (MARKX_CAP * 64) + (threadIdx.x & 63);
        force_agent64x(rec, pos32, offsets, list, 0u, vi, vj, vk, mnx, mny, mnz, m, voxel, rmax[0], px, py, pz,
                       radius_i, rep_i, adh_i, rel_max_adh, cut2, fx, fy, fz, local_pressure);
        if (vout) {
            vout[t] = make_double4(fx, fy, fz, local_pressure);
            return;
        }
        const uint32_t c = me.idx;
        vx[c] = fx;
        vy[c] = fy;
        vz[c] = fz;
        pressure[c] = local_pressure;
        return;
    }
    const bool mark32 = pos32 && !x_grid && fabsf(m.x) < near && yz_near;
#pragma unroll 1
    for (int q = 0; q < 9; q++) {
        const RowRange rr = neighbor_row(offsets, 0u, vi, vj, vk, mnx, mny, mnz, q);
        if (mark32)
            force_range64m(rec, pos32, rr.beg, rr.end, t, rr.mid0, rr.mid1, m.x, m.y, m.z, (float)rr.dj * voxel,
                           (float)rr.dk * voxel, voxel, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh, cut2, fx,
                           fy, fz, local_pressure);
        else
            force_range64(rec, rr.beg, rr.end, t, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh, cut2, fx, fy,
                          fz, local_pressure);
    }
    if (vout) {
        vout[t] = make_double4(fx, fy, fz, local_pressure);
        return;
    }
    const uint32_t c = me.idx;
    vx[c] = fx;
    vy[c] = fy;
    vz[c] = fz;
    pressure[c] = local_pressure;
}

// FP64 forces on the mirror with the x-trimmed mark pass (force_agent64t):
// 8-byte mark entries in (bin, x slab, index) order, xoff the slab starts;
// velocity and pressure to vout[t] in record order.  The agent's own mark
// coordinates come from its FP64 record.  Agents more than two bins from
// their bin corner, and every agent of a step whose gather set xtflag, take
// the exact FP64 pass over the rows.
__global__ __launch_bounds__(256) void k_forces_xt(const CellRec* __restrict__ rec, const uint2* __restrict__ mark,
                                                   const uint32_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ xoff,
                                                   const uint32_t* __restrict__ xtflag, uint32_t n, uint32_t mnx,
                                                   uint32_t mny, uint32_t mnz, double max_dist, double voxel,
                                                   const double* __restrict__ origin,
                                                   const float* __restrict__ rmax, double4* __restrict__ vout,
                                                   double* __restrict__ vx, double* __restrict__ vy,
                                                   double* __restrict__ vz, double* __restrict__ pressure) {
    __shared__ uint32_t marks[4 * MARKX_CAP * 64];
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= n) return;
    const CellRec me = rec[t];
    const double px = me.x, py = me.y, pz = me.z;
    const double radius_i = me.r, rep_i = me.rep, adh_i = me.adh, rel_max_adh = me.rel;
    double fx = 0.0, fy = 0.0, fz = 0.0, local_pressure = 0.0;
    const uint32_t vi = me.key % mnx, vj = (me.key / mnx) % mny, vk = me.key / (mnx * mny);
    const double cut2 = max_dist * max_dist;
    const double x_min = origin[0], y_min = origin[1], z_min = origin[2];
    const double yb = py - (y_min + vj * voxel), zb = pz - (z_min + vk * voxel);
    const float mxg = (float)(px - x_min);
    const float vox = (float)voxel, near = 2.0f * vox;
    if (xtflag[0] == 0u && fabs(yb) < (double)near && fabs(zb) < (double)near && mxg > -near &&
        mxg < (float)mnx * vox + near) {
        uint32_t* list = marks + (threadIdx.x >> 6) * (MARKX_CAP * 64) + (threadIdx.x & 63);
        force_agent64t(rec, mark, xoff, list, vi, vj, vk, mnx, mny, mnz, mxg, bin_rel32(px, x_min, vi, voxel),
                       (float)(yb + 64.0), (float)(zb + 64.0), vox, xslab_scale(voxel), rmax[0], px, py, pz, radius_i,
                       rep_i, adh_i, rel_max_adh, cut2, fx, fy, fz, local_pressure);
    } else {
#pragma unroll 1
        for (int q = 0; q < 9; q++) {
            const RowRange rr = neighbor_row(offsets, 0u, vi, vj, vk, mnx, mny, mnz, q);
            force_range64(rec, rr.beg, rr.end, t, px, py, pz, radius_i, rep_i, adh_i, rel_max_adh, cut2, fx, fy, fz,
                          local_pressure);
        }
    }
    if (vout) {
        vout[t] = make_double4(fx, fy, fz, local_pressure);
        return;
    }
    const uint32_t c = me.idx;
    vx[c] = fx;
    vy[c] = fy;
    vz[c] = fz;
    pressure[c] = local_pressure;
}

__global__ __launch_bounds__(256) void k_neighbors(const CellRec* __restrict__ rec, const uint32_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ counts, uint32_t* __restrict__ idx,
                                                   uint32_t n, uint32_t mnx, uint32_t mny, uint32_t mnz,
                                                   double max_dist, uint32_t cap) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const CellRec me = rec[t];
    const uint32_t vi = me.key % mnx, vj = (me.key / mnx) % mny, vk = me.key / (mnx * mny);
    const int dk_min = (mnz > 1) ? -1 : 0, dk_max = (mnz > 1) ? 1 : 0;
    const double cut2 = max_dist * max_dist;
    uint32_t cnt = 0;
    for (int dk = dk_min; dk <= dk_max; dk++) {
        const int nk = (int)vk + dk;
        if (nk < 0 || nk >= (int)mnz) continue;
        for (int dj = -1; dj <= 1; dj++) {
            const int nj = (int)vj + dj;
            if (nj < 0 || nj >= (int)mny) continue;
            for (int di = -1; di <= 1; di++) {
                const int ni = (int)vi + di;
                if (ni < 0 || ni >= (int)mnx) continue;
                const uint32_t nb = (uint32_t)nk * mny * mnx + (uint32_t)nj * mnx + (uint32_t)ni;
                for (uint32_t u = offsets[nb]; u < offsets[nb + 1]; u++) {
                    if (u == t) continue;
                    const double ddx = me.x - rec[u].x, ddy = me.y - rec[u].y, ddz = me.z - rec[u].z;
                    const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (d2 >= cut2) continue;
                    if (d2 < 1e-12) continue;
                    if (cnt < cap) idx[(uint64_t)me.idx * cap + cnt] = rec[u].idx;
                    cnt++;
                }
            }
        }
    }
    counts[me.idx] = cnt;
}

__global__ __launch_bounds__(256) void k_max_bin(const uint32_t* __restrict__ offsets, uint32_t nb,
                                                 uint32_t* __restrict__ out) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    uint32_t v = b < nb ? offsets[b + 1] - offsets[b] : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(out, v);
}

// ---------------------------------------------------------------------------
// AB2 (shaders/integrate.metal:56-79)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_integrate(CellPtrs cp, uint32_t n, double dt, double x_min,
                                                   double y_min, double z_min, double x_max,
                                                   double y_max, double z_max) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double px = cp.f[HC_F_POS_X][i], py = cp.f[HC_F_POS_Y][i], pz = cp.f[HC_F_POS_Z][i];
    const double vx = cp.f[HC_F_VEL_X][i], vy = cp.f[HC_F_VEL_Y][i], vz = cp.f[HC_F_VEL_Z][i];
    px += dt * (1.5 * vx - 0.5 * cp.f[HC_F_PREV_VX][i]);
    py += dt * (1.5 * vy - 0.5 * cp.f[HC_F_PREV_VY][i]);
    pz += dt * (1.5 * vz - 0.5 * cp.f[HC_F_PREV_VZ][i]);
    px = px < x_min ? x_min : (px > x_max ? x_max : px);
    py = py < y_min ? y_min : (py > y_max ? y_max : py);
    pz = pz < z_min ? z_min : (pz > z_max ? z_max : pz);
    cp.f[HC_F_POS_X][i] = px;
    cp.f[HC_F_POS_Y][i] = py;
    cp.f[HC_F_POS_Z][i] = pz;
    cp.f[HC_F_PREV_VX][i] = vx;
    cp.f[HC_F_PREV_VY][i] = vy;
    cp.f[HC_F_PREV_VZ][i] = vz;
}

// ---------------------------------------------------------------------------
// secretion (src/cell_phenotype.cpp:526-564), grouped by voxel
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_sec_keys(const uint32_t* __restrict__ voxel, const uint32_t* __restrict__ alive,
                                                  uint32_t* __restrict__ key_out, uint32_t* __restrict__ counts,
                                                  uint32_t* __restrict__ slot, uint32_t n, uint32_t V) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    const bool valid = c < n;
    uint32_t key = V;                                   // V = "not secreting" bin
    if (valid) {
        const uint32_t v = voxel[c];
        if (alive[c] && v < V) key = v;                // src/main.mm:360, src/cell_phenotype.cpp:534
        key_out[c] = key;
    }
    const uint32_t s = ballot_bin(valid, key, counts);
    if (valid) slot[c] = s;
}

// group-start flags over hash-sorted positions
__global__ __launch_bounds__(256) void k_sec_flags(const uint32_t* __restrict__ key, const uint32_t* __restrict__ sorted,
                                                   const uint32_t* __restrict__ offsets, uint32_t* __restrict__ flag,
                                                   uint32_t n, uint32_t V) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const uint32_t n_elig = offsets[V];
    uint32_t f = 0;
    if (t < n_elig) f = (t == 0 || key[sorted[t]] != key[sorted[t - 1]]) ? 1u : 0u;
    flag[t] = f;
}

// C: coefficient storage, the field's type (FP32 fields keep FP32 coefficients,
// halving the bytes the per-substep apply reads; they are evaluated in FP64
// and rounded once)
template <typename C>
__global__ __launch_bounds__(256) void k_sec_groups(CellPtrs cp, const double* __restrict__ rates,
                                                    const uint32_t* __restrict__ key, const uint32_t* __restrict__ sorted,
                                                    const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ gid,
                                                    uint32_t* __restrict__ gvox, uint32_t* __restrict__ gbegin,
                                                    C* __restrict__ coef, PlanHeader* __restrict__ hdr,
                                                    uint32_t n, uint32_t max_cells, uint32_t ns, uint32_t V,
                                                    double dt, double V_voxel) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n_elig = offsets[V];
    if (t == 0) {
        hdr->n_elig = n_elig;
        hdr->ngroups = gid[n];
        hdr->n = n;
        gbegin[gid[n]] = n_elig;
    }
    if (t >= n_elig) return;
    const uint32_t c = sorted[t];
    const uint32_t v = key[c];
    if (t == 0 || key[sorted[t - 1]] != v) {
        const uint32_t g = gid[t];
        gvox[g] = v;
        gbegin[g] = t;
    }
    const double* sec = rates;
    const double* upt = rates + (size_t)max_cells * ns;
    const double* sat = rates + (size_t)2 * max_cells * ns;
    const double* net = rates + (size_t)3 * max_cells * ns;
    const double V_cell = cp.f[HC_F_TOTAL_VOL][c];
    const double internal_constant = dt * V_cell / V_voxel;
    bool any_net = false;
    for (uint32_t s = 0; s < ns; s++) {
        const size_t off = (size_t)c * ns + s;
        const double a = dt * net[off] / V_voxel;
        const double temp1 = internal_constant * sec[off] * sat[off];
        const double temp2 = 1.0 + internal_constant * (sec[off] + upt[off]);
        coef[((size_t)s * 3 + 0) * max_cells + t] = (C)a;
        coef[((size_t)s * 3 + 1) * max_cells + t] = (C)temp1;
        coef[((size_t)s * 3 + 2) * max_cells + t] = (C)temp2;
        any_net |= (a != 0.0);
    }
    if (any_net) atomicOr(&hdr->has_net, 1u);
}

template <typename F>
__global__ __launch_bounds__(256) void k_sec_apply(F* __restrict__ rho, const uint32_t* __restrict__ gvox,
                                                   const uint32_t* __restrict__ gbegin,
                                                   const F* __restrict__ coef,
                                                   const PlanHeader* __restrict__ hdr, uint32_t max_cells,
                                                   uint32_t ns, uint64_t V) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= hdr->ngroups) return;
    const bool has_net = hdr->has_net != 0;
    const uint32_t v = gvox[g];
    const uint32_t b = gbegin[g], e = gbegin[g + 1];
    for (uint32_t s = 0; s < ns; s++) {
        const F* ca = coef + ((size_t)s * 3 + 0) * max_cells;
        const F* c1 = coef + ((size_t)s * 3 + 1) * max_cells;
        const F* c2 = coef + ((size_t)s * 3 + 2) * max_cells;
        double r = (double)rho[(uint64_t)s * V + v];
        for (uint32_t t = b; t < e; t++) {
            if (has_net) r += (double)ca[t];
            r = (r + (double)c1[t]) / (double)c2[t];
            if (r < 0.0) r = 0.0;
        }
        rho[(uint64_t)s * V + v] = (F)r;
    }
}

int probe_kernels_cells(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_scan_local, (const void*)k_scan_partials, (const void*)k_scan_add,
                         (const void*)k_scan2_reduce, (const void*)k_scan2_down,
                         (const void*)k_fill_u32, (const void*)k_voxel_keys, (const void*)k_mech_keys,
                         (const void*)k_scatter, (const void*)k_scatter_slab, (const void*)k_stable_gather, (const void*)k_rmax, (const void*)k_forces,
                         (const void*)k_neighbors, (const void*)k_max_bin, (const void*)k_integrate,
                         (const void*)k_sec_keys, (const void*)k_sec_flags, (const void*)k_sec_groups<double>,
                         (const void*)k_sec_groups<float>,
                         (const void*)k_sec_apply<double>, (const void*)k_sec_apply<float>};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: cell kernel");
    }
    return 0;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

HashLayout hash_layout(uint32_t n_mvox, uint32_t max_cells) {
    HashLayout L;
    size_t o = 0;
    L.counts = o;  o = align256(o + ((size_t)n_mvox + 1) * 4);
    L.offsets = o; o = align256(o + ((size_t)n_mvox + 2) * 4);
    L.slot = o;    o = align256(o + (size_t)max_cells * 4);
    L.tmp = o;     o = align256(o + (size_t)max_cells * 4);
    L.sorted = o;  o = align256(o + (size_t)max_cells * 4);
    L.rec = o;     o = align256(o + (size_t)max_cells * sizeof(CellRec));
    L.pos_of = o;  o = align256(o + (size_t)max_cells * 4);
    // pos32 carries MARKX_STEP entries of padding: the mark pass loads whole
    // steps past the end of a range and masks the extra entries
    L.pos32 = o;   o = align256(o + ((size_t)max_cells + MARKX_STEP) * sizeof(float4));
    // largest radius: one partial per gather block, then the maximum
    L.rpart = o;   o = align256(o + (size_t)blocks_for(max_cells ? max_cells : 1, 256) * 4);
    L.rmax = o;    o = align256(o + 4);
    L.partial = o; o = align256(o + scan_partial_bytes((uint64_t)n_mvox + 1));
    // FP32 records: slab starts per bin (and past the last bin), slab per agent
    L.xoff = o;    o = align256(o + ((size_t)n_mvox + 1) * XSLABS * 4);
    L.xslab = o;   o = align256(o + (size_t)max_cells);
    // set by the gather when a member does not fit its 8-byte mark entry;
    // the grid origin (x_min, y_min, z_min) follows at byte 8
    L.xtflag = o;  o = align256(o + 32);
    L.total = o;
    return L;
}

PlanLayout plan_layout(const hc_grid* g, uint32_t max_cells) {
    PlanLayout L;
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    size_t o = 0;
    L.header = o;  o = align256(o + sizeof(PlanHeader));
    L.counts = o;  o = align256(o + (V + 1) * 4);
    L.offsets = o; o = align256(o + (V + 2) * 4);
    L.slot = o;    o = align256(o + (size_t)max_cells * 4);
    L.tmp = o;     o = align256(o + (size_t)max_cells * 4);
    L.sorted = o;  o = align256(o + (size_t)max_cells * 4);
    L.gflag = o;   o = align256(o + ((size_t)max_cells + 1) * 4);   // flags, then group ids (in place)
    L.gvox = o;    o = align256(o + ((size_t)max_cells + 1) * 4);
    L.gbegin = o;  o = align256(o + ((size_t)max_cells + 2) * 4);
    L.coef = o;    o = align256(o + (size_t)3 * g->n_substrates * max_cells * sizeof(double));
    size_t pb = scan_partial_bytes(V + 1);
    const size_t pb2 = scan_partial_bytes(max_cells);
    if (pb2 > pb) pb = pb2;
    L.partial = o; o = align256(o + pb);
    L.total = o;
    return L;
}

}  // namespace hc

using namespace hc;

static int check_cells(hc_ctx* ctx, hc_buf* cells, uint32_t n, uint32_t max_cells, const char* who) {
    if (!ctx || !cells) return fail(ctx, HC_ERR_ARG, std::string(who) + ": null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, std::string(who) + ": n > max_cells");
    if (hc_buf_size(cells) < hc_cells_bytes(max_cells))
        return fail(ctx, HC_ERR_ARG, std::string(who) + ": cell buffer smaller than hc_cells_bytes(max_cells)");
    return 0;
}

extern "C" {

int hc_voxel_keys(hc_ctx* ctx, hc_buf* cells, const hc_grid* g, uint32_t n, uint32_t max_cells) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_voxel_keys");
    if (r) return r;
    if (!g || !g->nx || !g->ny || !g->nz) return fail(ctx, HC_ERR_ARG, "hc_voxel_keys: bad grid");
    if (!n) return 0;
    CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    KernelTimer tm(ctx, HC_K_VOXEL_KEYS);
    hipLaunchKernelGGL(k_voxel_keys, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                       cp.f[HC_F_POS_X], cp.f[HC_F_POS_Y], cp.f[HC_F_POS_Z],
                       cp.u[HC_F_VOXEL_IDX - HC_NUM_F64_FIELDS], n, *g);
    return check_launch(ctx, "k_voxel_keys");
}

size_t hc_secretion_plan_bytes(const hc_grid* g, uint32_t max_cells) {
    return g ? plan_layout(g, max_cells).total : 0;
}


}  // extern "C"

namespace hc {

__global__ __launch_bounds__(256) void k_stable_sort(const uint32_t* __restrict__ key,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ tmp,
                                                     uint32_t* __restrict__ sorted, uint32_t n) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const uint32_t b = key[c];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    sorted[beg + stable_rank(tmp, beg, end, c)] = c;
}

// ---------------------------------------------------------------------------
// mixed form: 32-byte records, FP32 force arithmetic
// ---------------------------------------------------------------------------
// Records (CellRec32) in each bin by x slab, then agent index; positions
// from k_bin_sort32 (rpos, in pos_of) or, in bins above BIN_SORT_CAP, by
// slab_rank here, which also places those bins' ascending lists.  Own
// relative adhesion distance is read from the SoA.
__global__ __launch_bounds__(256) void k_stable_gather32(CellPtrs cp, const uint32_t* __restrict__ key,
                                                         const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ tmp,
                                                         uint32_t* __restrict__ sorted,
                                                         const uint32_t* __restrict__ pos_of,
                                                         const uint8_t* __restrict__ xslab,
                                                         CellRec32* __restrict__ rec, uint32_t n, double x_min,
                                                         double y_min, double z_min, double voxel, uint32_t mnx,
                                                         uint32_t mny, float* __restrict__ rpart) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    // this block's largest FP32 radius (the values the records hold)
    block_max_to(c < n ? (float)cp.f[HC_F_RADIUS][c] : 0.0f, rpart);
    if (c >= n) return;
    const uint32_t b = key[c];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos;
    if (end - beg <= BIN_SORT_CAP) {
        pos = pos_of[c];                                  // placed by k_bin_sort32
    } else {
        sorted[beg + stable_rank(tmp, beg, end, c)] = c;
        pos = beg + slab_rank(tmp, xslab, beg, end, c, xslab[c]);
    }
    const uint32_t bi = b % mnx, bj = (b / mnx) % mny, bk = b / (mnx * mny);
    const double px = cp.f[HC_F_POS_X][c];
    CellRec32 r;
    r.xg = (float)(px - x_min);
    r.y = bin_rel32(cp.f[HC_F_POS_Y][c], y_min, bj, voxel);
    r.z = bin_rel32(cp.f[HC_F_POS_Z][c], z_min, bk, voxel);
    r.x = bin_rel32(px, x_min, bi, voxel);
    r.r = (float)cp.f[HC_F_RADIUS][c];
    r.rep = (float)cp.f[HC_F_REPULSION][c];
    r.adh = (float)cp.f[HC_F_ADHESION][c];
    r.idx = c;
    rec[pos] = r;
}

// LIST: the mark pass over x-trimmed rows with a list of marked candidates
// (force_agent32x; rmax is the largest FP32 radius); else the chunked pass
// (force_rows32), which populations of 2^27 agents or more, grids wider than
// MARKX_MAX_EXTENT in x and forces_fp32 = 2 take.  Two kernels, so that each
// gets the registers its own path needs.
// vout non-null (the mechanics mirror): the records' index field is the
// agent's mirror slot, mkey the slots' bins, rel_adh the slots' properties
// (double4, relative adhesion distance in w) and the forces go to vout in
// record order.
// M16 (the mirror): the mark pass reads the 8-byte entries in mark[],
// unless the gather set *xtflag (a member that does not fit its entry):
// then every agent takes the chunked pass.
template <bool LIST, bool M16 = false>
__global__ __launch_bounds__(256) void k_forces32(const CellRec32* __restrict__ rec, const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ xoff,
                                                  const uint32_t* __restrict__ mkey,
                                                  const double* __restrict__ rel_adh, double* __restrict__ vx,
                                                  double* __restrict__ vy, double* __restrict__ vz,
                                                  double* __restrict__ pressure, uint32_t n, uint32_t mnx, uint32_t mny,
                                                  uint32_t mnz, float max_dist, double voxel,
                                                  const float* __restrict__ rmax, double4* __restrict__ vout,
                                                  const uint2* __restrict__ mark = nullptr,
                                                  const uint32_t* __restrict__ xtflag = nullptr) {
    __shared__ uint32_t marks[LIST ? 4 * MARKX_CAP * 64 : 1];
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= n) return;
    const CellRec32 me = rec[t];
    const uint32_t key = mkey[me.idx];
    const float rel_max_adh = (float)(vout ? rel_adh[4 * (size_t)me.idx + 3] : rel_adh[me.idx]);
    float fx = 0.f, fy = 0.f, fz = 0.f, local_pressure = 0.f;
    const uint32_t vi = key % mnx, vj = (key / mnx) % mny, vk = key / (mnx * mny);
    const float cut2 = max_dist * max_dist, vox = (float)voxel;
    if (LIST && (!M16 || xtflag[0] == 0u)) {
        uint32_t* list = marks + (threadIdx.x >> 6) * (MARKX_CAP * 64) + (threadIdx.x & 63);
        force_agent32x<M16>(rec, xoff, list, t, 0u, vi, vj, vk, mnx, mny, mnz, me.xg, me.x, me.y, me.z, vox,
                            xslab_scale(voxel), (float)(1.0 / voxel), rmax[0], me.r, me.rep, me.adh, rel_max_adh,
                            cut2, fx, fy, fz, local_pressure, mark);
    } else {
        force_rows32(rec, offsets, t, 0u, vi, vj, vk, mnx, mny, mnz, me.x, me.y, me.z, vox, me.r, me.rep, me.adh,
                     rel_max_adh, cut2, fx, fy, fz, local_pressure);
    }
    if (vout) {
        vout[t] = make_double4(fx, fy, fz, local_pressure);
        return;
    }
    const uint32_t c = me.idx;
    vx[c] = fx;
    vy[c] = fy;
    vz[c] = fz;
    pressure[c] = local_pressure;
}

// ---------------------------------------------------------------------------
// compensated substrate totals (Neumaier), fixed combination order
// ---------------------------------------------------------------------------
struct Neu {
    double s, c;
};
__device__ inline void neu_add(Neu& a, double x) {
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
__device__ inline void neu_merge(Neu& a, const Neu& b) {
    neu_add(a, b.s);
    a.c += b.c;
}

static constexpr int TOT_BLOCKS = 256;

template <typename F>
__global__ __launch_bounds__(256) void k_totals(const F* __restrict__ rho, uint64_t V, Neu* __restrict__ partial) {
    const uint32_t s = blockIdx.y;
    const F* r = rho + (uint64_t)s * V;
    Neu a{0.0, 0.0};
    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (uint64_t)gridDim.x * 256)
        neu_add(a, (double)r[v]);
    __shared__ Neu red[256];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) neu_merge(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)s * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_totals_final(const Neu* __restrict__ partial, uint32_t nb, uint32_t ns, double* __restrict__ out) {
    const uint32_t s = threadIdx.x;
    if (s >= ns) return;
    Neu a{0.0, 0.0};
    for (uint32_t b = 0; b < nb; b++) neu_merge(a, partial[(size_t)s * nb + b]);
    out[s] = a.s + a.c;
}

}  // namespace hc

// Scratch of hc_substrate_totals_noalloc: per-block partial sums of up to
// HC_TOTALS_NOALLOC_MAX_SUBSTRATES substrates, then their results.  Each GPU
// holds its own copy from module load until the process ends; every call on
// that GPU shares it (so calls that overlap on one GPU race).  Every partial
// and result a call reads is written by that call first, so nothing needs
// clearing between calls.
static constexpr size_t TOTALS_PARTIAL_BYTES =
    (size_t)HC_TOTALS_NOALLOC_MAX_SUBSTRATES * 256 /* TOT_BLOCKS */ * 16 /* sizeof(Neu) */;
static constexpr size_t TOTALS_SCRATCH_BYTES = TOTALS_PARTIAL_BYTES + HC_TOTALS_NOALLOC_MAX_SUBSTRATES * sizeof(double);
static __device__ __align__(256) unsigned char scratch[TOTALS_SCRATCH_BYTES];

namespace hc {

static_assert(TOT_BLOCKS == 256 && sizeof(Neu) == 16, "TOTALS_PARTIAL_BYTES assumes 256 blocks of 16-byte partials");

// The device address of `scratch` on each GPU, looked up once per device with
// that device current (the host-side name is only a stand-in for the symbol).
int totals_scratch_init(hc_ctx* ctx) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(ctx->device);
    if (it == addr.end()) {
        hipError_t e = hipSetDevice(ctx->device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(ctx, e, "totals scratch: hipGetSymbolAddress");
        it = addr.emplace(ctx->device, p).first;
    }
    ctx->totals_scratch = it->second;
    return 0;
}

int probe_kernels_cells2(hc_ctx* ctx);
int probe_kernels_cells2(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_stable_sort, (const void*)k_stable_gather32, (const void*)k_forces32<true>, (const void*)k_forces32<false>, (const void*)k_forces32<true, true>,
                         (const void*)k_bin_sort, (const void*)k_bin_sort32,
                         (const void*)k_totals<double>, (const void*)k_totals<float>, (const void*)k_totals_final};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: cell kernel (mixed form)");
    }
    return 0;
}

int launch_rmax(hc_ctx* ctx, const float* rpart, uint32_t nb, float* rmax) {
    hipLaunchKernelGGL(k_rmax, dim3(1), dim3(256), 0, ctx->stream, rpart, nb, rmax);
    return check_launch(ctx, "k_rmax");
}

int launch_bin_sort(hc_ctx* ctx, const uint32_t* offsets, const uint32_t* tmp, uint32_t* sorted, uint32_t* pos_of,
                    uint32_t nbins) {
    if (!nbins) return 0;
    hipLaunchKernelGGL(k_bin_sort, dim3(blocks_for(nbins, 256)), dim3(256), 0, ctx->stream, offsets, tmp, sorted,
                       pos_of, nbins);
    return check_launch(ctx, "k_bin_sort");
}

int launch_bin_sort32(hc_ctx* ctx, const uint32_t* offsets, const uint32_t* tmp, const uint8_t* xslab,
                      uint32_t* sorted, uint32_t* rpos, uint32_t* xoff, uint32_t nbins) {
    hipLaunchKernelGGL(k_bin_sort32, dim3(blocks_for(nbins + 1, 256)), dim3(256), 0, ctx->stream, offsets, tmp, xslab,
                       sorted, rpos, xoff, nbins, (uint32_t*)nullptr, (uint32_t*)nullptr);
    return check_launch(ctx, "k_bin_sort32");
}

int launch_scatter(hc_ctx* ctx, const uint32_t* key, const uint32_t* slot, const uint32_t* offsets, uint32_t* tmp,
                   uint32_t n) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, key, slot, offsets, tmp, n);
    return check_launch(ctx, "k_scatter");
}

int launch_stable_sort(hc_ctx* ctx, const uint32_t* key, const uint32_t* offsets, const uint32_t* tmp,
                       uint32_t* sorted, uint32_t n) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_stable_sort, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, key, offsets, tmp, sorted, n);
    return check_launch(ctx, "k_stable_sort");
}

int launch_sec_flags(hc_ctx* ctx, const uint32_t* key, const uint32_t* sorted, const uint32_t* offsets, uint32_t* flag,
                     uint32_t n, uint32_t V) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_sec_flags, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, key, sorted, offsets, flag, n, V);
    return check_launch(ctx, "k_sec_flags");
}

struct HashPtrs {
    uint32_t *counts, *offsets, *slot, *tmp, *sorted, *partial, *pos_of;
    CellRec* rec;
    float4* pos32;
    float *rpart, *rmax;
    uint32_t* xoff;
    uint8_t* xslab;
    uint32_t* xtflag;
};

static HashPtrs hash_ptrs(hc_buf* hash, uint32_t n_mvox, uint32_t max_cells) {
    const HashLayout L = hash_layout(n_mvox, max_cells);
    uint8_t* P = (uint8_t*)hash->dptr;
    HashPtrs h;
    h.counts = (uint32_t*)(P + L.counts);
    h.offsets = (uint32_t*)(P + L.offsets);
    h.slot = (uint32_t*)(P + L.slot);
    h.tmp = (uint32_t*)(P + L.tmp);
    h.sorted = (uint32_t*)(P + L.sorted);
    h.partial = (uint32_t*)(P + L.partial);
    h.rec = (CellRec*)(P + L.rec);
    h.pos_of = (uint32_t*)(P + L.pos_of);
    h.pos32 = (float4*)(P + L.pos32);
    h.xoff = (uint32_t*)(P + L.xoff);
    h.xslab = (uint8_t*)(P + L.xslab);
    h.xtflag = (uint32_t*)(P + L.xtflag);
    h.rpart = (float*)(P + L.rpart);
    h.rmax = (float*)(P + L.rmax);
    return h;
}

}  // namespace hc

extern "C" {

int hc_secretion_plan(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, const hc_grid* g, double dt,
                      uint32_t n, uint32_t max_cells, hc_buf* plan) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_secretion_plan");
    if (r) return r;
    if (!g || !rates || !plan) return fail(ctx, HC_ERR_ARG, "hc_secretion_plan: null argument");
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    if (V >= 0xFFFFFFFFull) return fail(ctx, HC_ERR_CAPACITY, "hc_secretion_plan: grid too large for 32-bit voxel keys");
    const PlanLayout L = plan_layout(g, max_cells);
    if (hc_buf_size(plan) < L.total) return fail(ctx, HC_ERR_ARG, "hc_secretion_plan: plan buffer too small");
    if (hc_buf_size(rates) < hc_rates_bytes(max_cells, g->n_substrates))
        return fail(ctx, HC_ERR_ARG, "hc_secretion_plan: rates buffer too small");
    uint8_t* P = (uint8_t*)plan->dptr;
    uint32_t* counts = (uint32_t*)(P + L.counts);
    uint32_t* offsets = (uint32_t*)(P + L.offsets);
    uint32_t* slot = (uint32_t*)(P + L.slot);
    uint32_t* tmp = (uint32_t*)(P + L.tmp);
    uint32_t* sorted = (uint32_t*)(P + L.sorted);
    uint32_t* gflag = (uint32_t*)(P + L.gflag);
    uint32_t* gvox = (uint32_t*)(P + L.gvox);
    uint32_t* gbegin = (uint32_t*)(P + L.gbegin);
    uint32_t* partial = (uint32_t*)(P + L.partial);
    double* coef = (double*)(P + L.coef);
    PlanHeader* hdr = (PlanHeader*)(P + L.header);
    // per-cell sort keys live in gbegin until the groups are written
    uint32_t* key = gbegin;
    CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    plan->aux = max_cells;
    KernelTimer tm(ctx, HC_K_SEC_PLAN);
    if ((r = fill_u32(ctx, counts, V + 1, 0))) return r;
    if ((r = fill_u32(ctx, (uint32_t*)hdr, sizeof(PlanHeader) / 4, 0))) return r;
    const unsigned nb = blocks_for(n ? n : 1, 256);
    if (n) {
        hipLaunchKernelGGL(k_sec_keys, dim3(nb), dim3(256), 0, ctx->stream,
                           cp.u[HC_F_VOXEL_IDX - HC_NUM_F64_FIELDS], cp.u[HC_F_IS_ALIVE - HC_NUM_F64_FIELDS],
                           key, counts, slot, n, (uint32_t)V);
        if ((r = check_launch(ctx, "k_sec_keys"))) return r;
    }
    if ((r = scan_exclusive(ctx, counts, offsets, V + 1, partial))) return r;
    if (n) {
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, ctx->stream, key, slot, offsets, tmp, n);
        hipLaunchKernelGGL(k_stable_sort, dim3(nb), dim3(256), 0, ctx->stream, key, offsets, tmp, sorted, n);
        hipLaunchKernelGGL(k_sec_flags, dim3(nb), dim3(256), 0, ctx->stream, key, sorted, offsets, gflag, n,
                           (uint32_t)V);
        if ((r = check_launch(ctx, "secretion plan sort"))) return r;
    }
    if ((r = scan_exclusive(ctx, gflag, gflag, n, partial))) return r;
    // keys move out of gbegin before the group table is written: copy them to tmp
    // (tmp is free after the stable sort)
    if (n) {
        hipError_t e = hipMemcpyAsync(tmp, key, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_secretion_plan: key copy");
    }
    if (g->field_fp32)
        hipLaunchKernelGGL(k_sec_groups<float>, dim3(nb), dim3(256), 0, ctx->stream, cp, (const double*)rates->dptr,
                           tmp, sorted, offsets, gflag, gvox, gbegin, (float*)coef, hdr, n, max_cells,
                           g->n_substrates, (uint32_t)V, dt, g->dx * g->dy * g->dz);
    else
        hipLaunchKernelGGL(k_sec_groups<double>, dim3(nb), dim3(256), 0, ctx->stream, cp, (const double*)rates->dptr,
                           tmp, sorted, offsets, gflag, gvox, gbegin, coef, hdr, n, max_cells,
                           g->n_substrates, (uint32_t)V, dt, g->dx * g->dy * g->dz);
    return check_launch(ctx, "k_sec_groups");
}

int hc_secretion_apply(hc_ctx* ctx, hc_buf* density, hc_buf* plan, const hc_grid* g) {
    if (!ctx || !density || !plan || !g) return fail(ctx, HC_ERR_ARG, "hc_secretion_apply: null argument");
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    if (hc_buf_size(density) < V * g->n_substrates * field_elem(g))
        return fail(ctx, HC_ERR_ARG, "hc_secretion_apply: density buffer too small");
    const uint8_t* P = (const uint8_t*)plan->dptr;
    const uint32_t max_cells = (uint32_t)plan->aux;
    if (!max_cells) return fail(ctx, HC_ERR_STATE, "hc_secretion_apply: plan was never built (call hc_secretion_plan)");
    if (plan_layout(g, max_cells).total > hc_buf_size(plan))
        return fail(ctx, HC_ERR_ARG, "hc_secretion_apply: grid does not match the plan");
    const PlanLayout L = plan_layout(g, max_cells);
    if (!g->n_substrates || !max_cells) return 0;
    KernelTimer tm(ctx, HC_K_SEC_APPLY);
    if (g->field_fp32)
        hipLaunchKernelGGL(k_sec_apply<float>, dim3(blocks_for(max_cells, 256)), dim3(256), 0, ctx->stream,
                           (float*)density->dptr, (const uint32_t*)(P + L.gvox), (const uint32_t*)(P + L.gbegin),
                           (const float*)(P + L.coef), (const PlanHeader*)(P + L.header), max_cells,
                           g->n_substrates, V);
    else
        hipLaunchKernelGGL(k_sec_apply<double>, dim3(blocks_for(max_cells, 256)), dim3(256), 0, ctx->stream,
                           (double*)density->dptr, (const uint32_t*)(P + L.gvox), (const uint32_t*)(P + L.gbegin),
                           (const double*)(P + L.coef), (const PlanHeader*)(P + L.header), max_cells,
                           g->n_substrates, V);
    return check_launch(ctx, "k_sec_apply");
}

int hc_secretion(hc_ctx* ctx, hc_buf* density, hc_buf* cells, hc_buf* rates, const hc_grid* g,
                 double dt, uint32_t n, uint32_t max_cells, hc_buf* plan) {
    int r = hc_secretion_plan(ctx, cells, rates, g, dt, n, max_cells, plan);
    if (r) return r;
    return hc_secretion_apply(ctx, density, plan, g);
}

size_t hc_hash_bytes(uint32_t n_mvox, uint32_t max_cells) { return hash_layout(n_mvox, max_cells).total; }

// The FP32 mark pass of the FP64 force kernel is exact for bin sizes and
// cutoffs up to 32 um (the error bound in force_range64m).
static bool mark32_ok(const hc_mech* mech) {
    return mech->mechanics_voxel_size > 0.0 && mech->mechanics_voxel_size <= 32.0 &&
           mech->max_interaction_distance <= 32.0;
}
// ... and its x relative to the grid's x_min (force_agent64x) when the grid's
// x extent keeps every FP32 x of a candidate in range below 4096 um
static int markx_ok(const hc_mech* mech) {
    return mark32_ok(mech) && (double)mech->mech_grid_nx * mech->mechanics_voxel_size <= MARKX_MAX_EXTENT ? 1 : 0;
}

// FP32 forces take the x-trimmed mark pass with a list (force_agent32x) when
// the mark's FP32 x relative to x_min keeps its error bound (markx_ok), the
// record positions fit the list entries, and forces_fp32 is not 2 (the
// chunked pass on request)
static bool list32_ok(const hc_mech* mech, uint32_t n, uint32_t n_mvox) {
    return mech->forces_fp32 == 1 && markx_ok(mech) && mech->mechanics_voxel_size >= 0.5 &&
           n < (1u << MARK32_POS_BITS) && n_mvox < (1u << 30);
}

// FP64 forces on the mirror take the x-trimmed mark pass (force_agent64t)
// under the same conditions, less the list entries' row code
static bool xtrim64_ok(const hc_mech* mech, uint32_t n_mvox) {
    return !mech->forces_fp32 && markx_ok(mech) && mech->mechanics_voxel_size >= 0.5 && n_mvox < (1u << 30);
}

static int check_hash(hc_ctx* ctx, hc_buf* hash, uint32_t n_mvox, uint32_t max_cells, const char* who) {
    if (!hash || n_mvox == 0) return fail(ctx, HC_ERR_ARG, std::string(who) + ": bad hash argument");
    if (hc_buf_size(hash) < hc_hash_bytes(n_mvox, max_cells))
        return fail(ctx, HC_ERR_ARG, std::string(who) + ": hash buffer smaller than hc_hash_bytes");
    return 0;
}

int hc_build_hash(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, const hc_grid* mgrid,
                  uint32_t n, uint32_t max_cells, uint32_t n_mvox) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_build_hash");
    if (r || (r = check_hash(ctx, hash, n_mvox, max_cells, "hc_build_hash"))) return r;
    if (!mech || !mgrid) return fail(ctx, HC_ERR_ARG, "hc_build_hash: null parameters");
    if ((uint64_t)mech->mech_grid_nx * mech->mech_grid_ny * mech->mech_grid_nz != n_mvox)
        return fail(ctx, HC_ERR_ARG, "hc_build_hash: n_mvox != mech grid size");
    HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    const bool xt = xtrim64_ok(mech, n_mvox) && max_cells <= (1u << 30);
    KernelTimer tm(ctx, HC_K_HASH);
    if ((r = fill_u32(ctx, h.counts, n_mvox, 0))) return r;
    const unsigned nb = blocks_for(n ? n : 1, 256);
    if (n) {
        hipLaunchKernelGGL(k_mech_keys, dim3(nb), dim3(256), 0, ctx->stream, cp.f[HC_F_POS_X], cp.f[HC_F_POS_Y],
                           cp.f[HC_F_POS_Z], cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS], h.counts, h.slot, n,
                           mgrid->x_min, mgrid->y_min, mgrid->z_min, mech->mechanics_voxel_size,
                           mech->mech_grid_nx, mech->mech_grid_ny, mech->mech_grid_nz,
                           mech->forces_fp32 || xt ? h.xslab : nullptr);
        if ((r = check_launch(ctx, "k_mech_keys"))) return r;
    }
    if ((r = scan_exclusive(ctx, h.counts, h.offsets, n_mvox, h.partial))) return r;
    if (n) {
        const uint32_t* key = cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS];
        if (xt)
            hipLaunchKernelGGL(k_scatter_slab, dim3(nb), dim3(256), 0, ctx->stream, key, h.slot, h.xslab, h.offsets,
                               h.tmp, n);
        else
            hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, ctx->stream, key, h.slot, h.offsets, h.tmp, n);
        if (mech->forces_fp32) {
            if ((r = launch_bin_sort32(ctx, h.offsets, h.tmp, h.xslab, h.sorted, h.pos_of, h.xoff, n_mvox))) return r;
            hipLaunchKernelGGL(k_stable_gather32, dim3(nb), dim3(256), 0, ctx->stream, cp, key, h.offsets, h.tmp,
                               h.sorted, h.pos_of, h.xslab, reinterpret_cast<CellRec32*>(h.rec), n, mgrid->x_min,
                               mgrid->y_min, mgrid->z_min, mech->mechanics_voxel_size, mech->mech_grid_nx,
                               mech->mech_grid_ny, h.rpart);
        } else if (xt) {
            // the plain form's 8-byte mark entries: the bin sort also places
            // each agent's entry by (x slab, index), into slot (dead after
            // the scatter)
            hipLaunchKernelGGL(k_bin_sort32, dim3(blocks_for(n_mvox + 1, 256)), dim3(256), 0, ctx->stream, h.offsets,
                               h.tmp, h.xslab, h.sorted, h.slot, h.xoff, n_mvox, h.pos_of, h.xtflag);
            hipLaunchKernelGGL(k_stable_gather, dim3(nb), dim3(256), 0, ctx->stream, cp, key, h.offsets, h.tmp,
                               h.sorted, h.pos_of, h.rec, h.pos32, n, mgrid->x_min, mgrid->y_min, mgrid->z_min,
                               mech->mechanics_voxel_size, mech->mech_grid_nx, mech->mech_grid_ny, 1, h.rpart,
                               h.slot, h.xslab, h.xtflag);
        } else {
            if ((r = launch_bin_sort(ctx, h.offsets, h.tmp, h.sorted, h.pos_of, n_mvox))) return r;
            hipLaunchKernelGGL(k_stable_gather, dim3(nb), dim3(256), 0, ctx->stream, cp, key, h.offsets, h.tmp,
                               h.sorted, h.pos_of, h.rec, mark32_ok(mech) ? h.pos32 : nullptr, n, mgrid->x_min, mgrid->y_min,
                               mgrid->z_min, mech->mechanics_voxel_size, mech->mech_grid_nx, mech->mech_grid_ny,
                               markx_ok(mech), markx_ok(mech) ? h.rpart : nullptr, (const uint32_t*)nullptr,
                               (const uint8_t*)nullptr, (uint32_t*)nullptr);
        }
        if (mech->forces_fp32 || markx_ok(mech))
            hipLaunchKernelGGL(k_rmax, dim3(1), dim3(256), 0, ctx->stream, h.rpart, nb, h.rmax);
        if ((r = check_launch(ctx, "hash scatter"))) return r;
    }
    return 0;
}

int hc_forces(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, uint32_t n, uint32_t max_cells,
              uint32_t n_mvox) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_forces");
    if (r || (r = check_hash(ctx, hash, n_mvox, max_cells, "hc_forces"))) return r;
    if (!mech) return fail(ctx, HC_ERR_ARG, "hc_forces: null parameters");
    if (!n) return 0;
    HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    KernelTimer tm(ctx, HC_K_FORCES);
    if (mech->forces_fp32)
        hipLaunchKernelGGL(list32_ok(mech, n, n_mvox) ? k_forces32<true> : k_forces32<false>,
                           dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                           reinterpret_cast<const CellRec32*>(h.rec), h.offsets, h.xoff,
                           cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS], cp.f[HC_F_REL_MAX_ADH],
                           cp.f[HC_F_VEL_X], cp.f[HC_F_VEL_Y], cp.f[HC_F_VEL_Z], cp.f[HC_F_PRESSURE], n,
                           mech->mech_grid_nx, mech->mech_grid_ny, mech->mech_grid_nz,
                           (float)mech->max_interaction_distance, mech->mechanics_voxel_size, h.rmax,
                           (double4*)nullptr, (const uint2*)nullptr, (const uint32_t*)nullptr);
    else if (xtrim64_ok(mech, n_mvox) && max_cells <= (1u << 30))
        hipLaunchKernelGGL(k_forces_xt, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, h.rec,
                           reinterpret_cast<const uint2*>(h.pos32), h.offsets, h.xoff, h.xtflag, n,
                           mech->mech_grid_nx, mech->mech_grid_ny, mech->mech_grid_nz,
                           mech->max_interaction_distance, mech->mechanics_voxel_size,
                           reinterpret_cast<const double*>(h.xtflag + 2), h.rmax, (double4*)nullptr,
                           cp.f[HC_F_VEL_X], cp.f[HC_F_VEL_Y], cp.f[HC_F_VEL_Z], cp.f[HC_F_PRESSURE]);
    else
        hipLaunchKernelGGL(k_forces, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, h.rec,
                           mark32_ok(mech) ? h.pos32 : nullptr, h.offsets, cp.f[HC_F_VEL_X], cp.f[HC_F_VEL_Y],
                           cp.f[HC_F_VEL_Z], cp.f[HC_F_PRESSURE], n, mech->mech_grid_nx, mech->mech_grid_ny,
                           mech->mech_grid_nz, mech->max_interaction_distance, (float)mech->mechanics_voxel_size,
                           markx_ok(mech), h.rmax, (double4*)nullptr);
    return check_launch(ctx, "k_forces");
}

int hc_forces_only_pipeline(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, const hc_grid* mgrid,
                            uint32_t n, uint32_t max_cells, uint32_t n_mvox) {
    int r = hc_build_hash(ctx, cells, hash, mech, mgrid, n, max_cells, n_mvox);
    if (r) return r;
    return hc_forces(ctx, cells, hash, mech, n, max_cells, n_mvox);
}

int hc_integrate(hc_ctx* ctx, hc_buf* cells, const hc_grid* mgrid, double dt, uint32_t n, uint32_t max_cells) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_integrate");
    if (r) return r;
    if (!mgrid) return fail(ctx, HC_ERR_ARG, "hc_integrate: null grid");
    if (!n) return 0;
    CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    // domain bounds [min, min + n*d] of the mech grid (shaders/integrate.metal:63-69)
    const double x_max = mgrid->x_min + (double)mgrid->nx * mgrid->dx;
    const double y_max = mgrid->y_min + (double)mgrid->ny * mgrid->dy;
    const double z_max = mgrid->z_min + (double)mgrid->nz * mgrid->dz;
    KernelTimer tm(ctx, HC_K_INTEGRATE);
    hipLaunchKernelGGL(k_integrate, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, cp, n, dt, mgrid->x_min,
                       mgrid->y_min, mgrid->z_min, x_max, y_max, z_max);
    return check_launch(ctx, "k_integrate");
}

int hc_mechanics_pipeline(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, const hc_grid* mgrid,
                          double dt, uint32_t n, uint32_t max_cells, uint32_t n_mvox) {
    int r = hc_forces_only_pipeline(ctx, cells, hash, mech, mgrid, n, max_cells, n_mvox);
    if (r) return r;
    return hc_integrate(ctx, cells, mgrid, dt, n, max_cells);
}

int hc_hash_download(hc_ctx* ctx, hc_buf* hash, uint32_t n_mvox, uint32_t max_cells, uint32_t n,
                     uint32_t* offsets_out, uint32_t* cells_out) {
    if (!ctx || !hash || !offsets_out || (!cells_out && n)) return fail(ctx, HC_ERR_ARG, "hc_hash_download: null argument");
    int r = check_hash(ctx, hash, n_mvox, max_cells, "hc_hash_download");
    if (r) return r;
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, "hc_hash_download: n > max_cells");
    HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    hipError_t e = hipMemcpyAsync(offsets_out, h.offsets, ((size_t)n_mvox + 1) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && n) e = hipMemcpyAsync(cells_out, h.sorted, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_hash_download");
}

int hc_hash_max_bin(hc_ctx* ctx, hc_buf* hash, uint32_t n_mvox, uint32_t* max_out) {
    if (!ctx || !hash || !max_out) return fail(ctx, HC_ERR_ARG, "hc_hash_max_bin: null argument");
    const HashLayout L = hash_layout(n_mvox, 0);
    const uint32_t* offsets = (const uint32_t*)((uint8_t*)hash->dptr + L.offsets);
    int r = fill_u32(ctx, ctx->d_scalar, 1, 0);
    if (r) return r;
    hipLaunchKernelGGL(k_max_bin, dim3(blocks_for(n_mvox, 256)), dim3(256), 0, ctx->stream, offsets, n_mvox,
                       ctx->d_scalar);
    if ((r = check_launch(ctx, "k_max_bin"))) return r;
    hipError_t e = hipMemcpyAsync(max_out, ctx->d_scalar, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_hash_max_bin");
}

int hc_neighbor_lists(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, uint32_t n,
                      uint32_t max_cells, uint32_t n_mvox, uint32_t cap, hc_buf* counts, hc_buf* idx) {
    int r = check_cells(ctx, cells, n, max_cells, "hc_neighbor_lists");
    if (r || (r = check_hash(ctx, hash, n_mvox, max_cells, "hc_neighbor_lists"))) return r;
    if (!mech || !counts || !idx) return fail(ctx, HC_ERR_ARG, "hc_neighbor_lists: null argument");
    if (mech->forces_fp32)
        return fail(ctx, HC_ERR_STATE, "hc_neighbor_lists: needs the FP64 records (forces_fp32 = 0)");
    if (hc_buf_size(counts) < (size_t)n * 4 || hc_buf_size(idx) < (size_t)n * cap * 4)
        return fail(ctx, HC_ERR_ARG, "hc_neighbor_lists: output buffers too small");
    if (!n) return 0;
    HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    hipLaunchKernelGGL(k_neighbors, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, h.rec, h.offsets,
                       (uint32_t*)counts->dptr, (uint32_t*)idx->dptr, n, mech->mech_grid_nx, mech->mech_grid_ny,
                       mech->mech_grid_nz, mech->max_interaction_distance, cap);
    return check_launch(ctx, "k_neighbors");
}

int hc_substrate_totals(hc_ctx* ctx, hc_buf* density, const hc_grid* g, double* out) {
    if (!ctx || !density || !g || !out) return fail(ctx, HC_ERR_ARG, "hc_substrate_totals: null argument");
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    const uint32_t ns = g->n_substrates;
    if (!ns) return 0;
    if (ns > 1024) return fail(ctx, HC_ERR_ARG, "hc_substrate_totals: more than 1024 substrates");
    if (hc_buf_size(density) < V * ns * field_elem(g)) return fail(ctx, HC_ERR_ARG, "hc_substrate_totals: buffer too small");
    void* mem = nullptr;
    const size_t pbytes = (size_t)ns * TOT_BLOCKS * sizeof(Neu);
    hipError_t e = hipMalloc(&mem, pbytes + ns * sizeof(double));
    if (e != hipSuccess) return hip_fail(ctx, e, "hc_substrate_totals: scratch");
    Neu* part = (Neu*)mem;
    double* dout = (double*)((char*)mem + pbytes);
    if (g->field_fp32)
        hipLaunchKernelGGL(k_totals<float>, dim3(TOT_BLOCKS, ns), dim3(256), 0, ctx->stream, (const float*)density->dptr, V, part);
    else
        hipLaunchKernelGGL(k_totals<double>, dim3(TOT_BLOCKS, ns), dim3(256), 0, ctx->stream, (const double*)density->dptr, V, part);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1024), 0, ctx->stream, part, (uint32_t)TOT_BLOCKS, ns, dout);
    int r = check_launch(ctx, "k_totals");
    if (!r) {
        e = hipMemcpyAsync(out, dout, ns * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) r = hip_fail(ctx, e, "hc_substrate_totals");
    }
    (void)hipStreamSynchronize(ctx->stream);        // cleanup: r holds the first failure
    (void)hipFree(mem);
    return r;
}

int hc_substrate_totals_noalloc(hc_ctx* ctx, hc_buf* density, const hc_grid* g, double* out) {
    if (!ctx || !density || !g || !out) return fail(ctx, HC_ERR_ARG, "hc_substrate_totals_noalloc: null argument");
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    const uint32_t ns = g->n_substrates;
    if (!ns) return 0;
    if (ns > HC_TOTALS_NOALLOC_MAX_SUBSTRATES)
        return fail(ctx, HC_ERR_CAPACITY, "hc_substrate_totals_noalloc: " + std::to_string(ns) +
                                              " substrates exceed the static scratch's maximum of " +
                                              std::to_string(HC_TOTALS_NOALLOC_MAX_SUBSTRATES));
    if (hc_buf_size(density) < V * ns * field_elem(g))
        return fail(ctx, HC_ERR_ARG, "hc_substrate_totals_noalloc: buffer too small");
    if (!ctx->totals_scratch) {
        int r = totals_scratch_init(ctx);
        if (r) return r;
    }
    // the device address of this GPU's copy of `scratch` (never the host-side name)
    Neu* part = (Neu*)ctx->totals_scratch;
    double* dout = (double*)((char*)ctx->totals_scratch + TOTALS_PARTIAL_BYTES);
    if (g->field_fp32)
        hipLaunchKernelGGL(k_totals<float>, dim3(TOT_BLOCKS, ns), dim3(256), 0, ctx->stream, (const float*)density->dptr, V, part);
    else
        hipLaunchKernelGGL(k_totals<double>, dim3(TOT_BLOCKS, ns), dim3(256), 0, ctx->stream, (const double*)density->dptr, V, part);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1024), 0, ctx->stream, part, (uint32_t)TOT_BLOCKS, ns, dout);
    int r = check_launch(ctx, "k_totals (static scratch)");
    if (r) return r;
    hipError_t e = hipMemcpyAsync(out, dout, ns * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_substrate_totals_noalloc");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Mechanics mirror of one simulation (FP64 forces).
//
// The agents' order in the cell SoA (their index) need not follow space, so
// every per-agent pass of the mechanics step in index order scatters: the
// ballot of the key pass finds about one key per lane (one atomic each), the
// gather reads each agent's fields from unrelated cache lines, and the force
// pass stores velocity and pressure at the agent's index.  The mirror keeps
// each agent's mechanics state -- position, previous velocity, radius,
// repulsion, adhesion, relative adhesion distance (FP64), index -- in slots
// ordered like the records of a recent hash build.  Slots keep their order
// from step to step (mirror_resort puts them back in record order as agents
// change bins); each slot's bin and place in the bin for the coming step are
// computed by AB2 itself, with coherent ballots, and each slot knows its
// record position of the step (rpos), through which AB2 reads its forces.
// Per step the gather reads 72 bytes of slot state (positions packed in 24
// bytes) and writes the 64-byte record and the mark entry, AB2 reads and
// writes 48 bytes of state beside its 32-byte forces.
// The cell fields (position, velocity, previous velocity, pressure,
// mechanics voxel) are written back, by agent index, only when something
// else reads them (mirror_sync).  Keys, bins, the in-bin order (ascending
// agent index), the records and every arithmetic step are those of the
// direct path, so the results are the same bit for bit.  The slots' radii
// change only by a reload (mirror_load's hash build), so the largest radius
// that build computes serves every step until the next load; the scan clears
// the bin counts it reads, for AB2's histogram of the coming step.
// ---------------------------------------------------------------------------
namespace hc {

MirrorLayout mirror_layout(uint32_t max_cells) {
    MirrorLayout L;
    size_t o = 0;
    for (int k = 0; k < 2; k++) {
        L.pos[k] = o;   o = align256(o + (size_t)max_cells * sizeof(Dbl3));
        L.prev[k] = o;  o = align256(o + (size_t)max_cells * sizeof(Dbl3));
        L.prop[k] = o;  o = align256(o + (size_t)max_cells * sizeof(double4));
        L.idx[k] = o;   o = align256(o + (size_t)max_cells * 4);
        L.key[k] = o;   o = align256(o + (size_t)max_cells * 4);
        L.slot[k] = o;  o = align256(o + (size_t)max_cells * 4);
        L.kstep[k] = o; o = align256(o + (size_t)max_cells * 4);
        L.rpos[k] = o;  o = align256(o + (size_t)max_cells * 4);
        L.xslab[k] = o; o = align256(o + (size_t)max_cells);
    }
    L.vp = o;      o = align256(o + (size_t)max_cells * sizeof(double4));
    L.spos = o;    o = align256(o + (size_t)max_cells * 4);
    L.total = o;
    return L;
}

// one set of slot arrays (set 1 receives a re-sort)
struct MSet {
    Dbl3 *pos, *prev;
    double4* prop;                  // radius, repulsion, adhesion, relative max adhesion distance
    uint32_t *idx, *key, *slot, *kstep, *rpos;
    uint8_t* xslab;                 // the x slab of the slot's FP32 record (or mark position) in its coming bin
    uint32_t* spos;                 // plain form, x-trimmed marks: the slot's mark position this step
};

// the slots from a hash built on the cell fields, in its record order: state,
// bin and place in the bin (the hash's counts hold the histogram)
__global__ __launch_bounds__(256) void k_m_load(const CellRec* __restrict__ rec, const uint32_t* __restrict__ offsets,
                                                CellPtrs cp, MSet a, uint32_t n, double x_min, double voxel,
                                                uint32_t mnx, int slabs) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const CellRec r = rec[q];
    if (slabs) a.xslab[q] = (uint8_t)xslab_of(bin_rel32(r.x, x_min, r.key % mnx, voxel), xslab_scale(voxel));
    a.pos[q] = dbl3(r.x, r.y, r.z);
    a.prev[q] = dbl3(cp.f[HC_F_PREV_VX][r.idx], cp.f[HC_F_PREV_VY][r.idx], cp.f[HC_F_PREV_VZ][r.idx]);
    a.prop[q] = make_double4(r.r, r.rep, r.adh, r.rel);
    a.idx[q] = r.idx;
    a.key[q] = r.key;
    a.slot[q] = q - offsets[r.key];
    a.kstep[q] = r.key;
    a.rpos[q] = q;
}

// Bins of up to BIN_SORT_CAP members (slots in tmp): the members ordered by
// agent index (a bitonic network over (index, slot) pairs); writes the bin's
// agent list and each slot's record position.
__global__ __launch_bounds__(256) void k_m_bin_sort(const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ tmp, MSet a,
                                                    uint32_t* __restrict__ sorted, uint32_t nbins) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nbins) return;
    const uint32_t beg = offsets[b], cnt = offsets[b + 1] - beg;
    if (cnt == 0 || cnt > BIN_SORT_CAP) return;
    uint64_t v[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) {
        const uint32_t m = q < cnt ? tmp[beg + q] : 0u;
        v[q] = q < cnt ? ((uint64_t)a.idx[m] << 32) | m : ~0ull;
    }
#pragma unroll
    for (uint32_t k = 2; k <= BIN_SORT_CAP; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t i = 0; i < BIN_SORT_CAP; i++) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t lo = v[i] < v[l] ? v[i] : v[l], hi = v[i] < v[l] ? v[l] : v[i];
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++)
        if (q < cnt) {
            sorted[beg + q] = (uint32_t)(v[q] >> 32);
            a.rpos[(uint32_t)v[q]] = beg + q;
        }
}

// This step's records (k_stable_gather's values) and FP32 mark positions
// from the slots; slots of bins above BIN_SORT_CAP are ranked here by agent
// index.  xt: 8-byte mark entries (mark_entry16) in (bin, x slab, index)
// order into pos32's memory (force_agent64t), a member that does not fit
// setting xtflag; else FP32 mark positions in record order with the radius
// in w.
__global__ __launch_bounds__(256) void k_m_gather(MSet a, const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                                                  CellRec* __restrict__ rec, float4* __restrict__ pos32, uint32_t n,
                                                  double x_min, double y_min, double z_min, double voxel,
                                                  uint32_t mnx, uint32_t mny, int x_grid, int xt,
                                                  uint32_t* __restrict__ xtflag, float* __restrict__ rpart) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const double4 pr = p < n ? a.prop[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (rpart) block_max_to(__double2float_ru(pr.x), rpart);
    if (p >= n) return;
    const uint32_t b = a.key[p], id = a.idx[p];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos, sp = 0;
    if (end - beg <= BIN_SORT_CAP) {
        pos = a.rpos[p];                                  // placed by k_m_bin_sort (k_m_bin_sort32<true>)
        if (xt) sp = a.spos[p];
    } else if (xt) {
        const uint32_t sc = a.xslab[p];
        uint32_t ri = 0, rs = 0;
        for (uint32_t u = beg; u < end; u++) {
            const uint32_t o = tmp[u];
            const uint32_t io = a.idx[o], so = a.xslab[o];
            ri += io < id ? 1u : 0u;
            rs += (so < sc || (so == sc && io < id)) ? 1u : 0u;
        }
        pos = beg + ri;
        sp = beg + rs;
        sorted[pos] = id;
        a.rpos[p] = pos;
    } else {
        uint32_t rank = 0;
        for (uint32_t u0 = beg; u0 < end; u0 += 8) {
            uint32_t w[8];
#pragma unroll
            for (int q = 0; q < 8; q++) w[q] = a.idx[tmp[u0 + q < end ? u0 + q : end - 1]];
#pragma unroll
            for (int q = 0; q < 8; q++) rank += (u0 + q < end && w[q] < id) ? 1u : 0u;
        }
        pos = beg + rank;
        sorted[pos] = id;
        a.rpos[p] = pos;
    }
    a.kstep[p] = b;
    const Dbl3 x = a.pos[p];
    if (xt) {
        if (p == 0) {
            double* org = reinterpret_cast<double*>(xtflag + 2);
            org[0] = x_min;
            org[1] = y_min;
            org[2] = z_min;
        }
        const uint32_t bj = (b / mnx) % mny, bk = b / (mnx * mny);
        bool fits;
        reinterpret_cast<uint2*>(pos32)[sp] = mark_entry16(x.x - x_min, x.y - (y_min + bj * voxel),
                                                           x.z - (z_min + bk * voxel), (int)pos - (int)sp, fits);
        if (!fits) xtflag[0] = 1u;
    } else if (pos32) {
        const uint32_t bi = b % mnx, bj = (b / mnx) % mny, bk = b / (mnx * mny);
        pos32[pos] = make_float4((float)(x.x - (x_grid ? x_min : x_min + bi * voxel)),
                                 (float)(x.y - (y_min + bj * voxel)), (float)(x.z - (z_min + bk * voxel)),
                                 (float)pr.x);
    }
    CellRec r;
    r.x = x.x;
    r.y = x.y;
    r.z = x.z;
    r.r = pr.x;
    r.rep = pr.y;
    r.adh = pr.z;
    r.rel = pr.w;
    r.idx = id;
    r.key = b;
    rec[pos] = r;
}

// Mixed form: the slots from an FP32 hash built on the cell fields, in its
// record order (state from the cell fields; bins, places and slabs as the
// hash build computed them)
__global__ __launch_bounds__(256) void k_m_load32(const CellRec32* __restrict__ rec,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint8_t* __restrict__ xslab, CellPtrs cp, MSet a, uint32_t n) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const uint32_t c = rec[q].idx;
    const uint32_t key = cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS][c];
    a.pos[q] = dbl3(cp.f[HC_F_POS_X][c], cp.f[HC_F_POS_Y][c], cp.f[HC_F_POS_Z][c]);
    a.prev[q] = dbl3(cp.f[HC_F_PREV_VX][c], cp.f[HC_F_PREV_VY][c], cp.f[HC_F_PREV_VZ][c]);
    a.prop[q] = make_double4(cp.f[HC_F_RADIUS][c], cp.f[HC_F_REPULSION][c], cp.f[HC_F_ADHESION][c],
                             cp.f[HC_F_REL_MAX_ADH][c]);
    a.idx[q] = c;
    a.key[q] = key;
    a.slot[q] = q - offsets[key];
    a.kstep[q] = key;
    a.rpos[q] = q;
    a.xslab[q] = xslab[c];
}

// Bins of up to BIN_SORT_CAP members (slots in tmp) in ascending agent index
// (the bin's list) and each slot's place by (x slab, index): the mixed
// form's FP32 record position (rpos), or, REC64, the plain form's mark entry
// position (spos) beside its record position by index (rpos); every bin's
// slab starts (xoff) and the entries past the last bin (grid of nbins + 1
// threads; REC64: that thread also clears xtflag).  Larger bins: slab starts
// only (the gather ranks their slots).
template <bool REC64>
__global__ __launch_bounds__(256) void k_m_bin_sort32(const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ tmp, MSet a,
                                                      uint32_t* __restrict__ sorted, uint32_t* __restrict__ xoff,
                                                      uint32_t nbins, uint32_t* __restrict__ xtflag) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b > nbins) return;
    if (b == nbins) {
        const uint32_t n = offsets[nbins];
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * nbins) = make_uint4(n, n, n, n);
        if (xtflag) xtflag[0] = 0u;                      // the gather that follows sets it
        return;
    }
    const uint32_t beg = offsets[b], cnt = offsets[b + 1] - beg;
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    if (cnt == 0 || cnt > BIN_SORT_CAP) {
        for (uint32_t q = beg; q < beg + cnt; q++) {
            const uint32_t sq = a.xslab[tmp[q]];
            c0 += sq == 0 ? 1u : 0u;
            c1 += sq == 1 ? 1u : 0u;
            c2 += sq == 2 ? 1u : 0u;
        }
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) =
            make_uint4(beg, beg + c0, beg + c0 + c1, beg + c0 + c1 + c2);
        return;
    }
    // (index, slab, slot): the slab rides along with the sort (slots below
    // 2^30), so it is loaded with the index instead of after the sort
    uint64_t v[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) {
        const uint32_t m = q < cnt ? tmp[beg + q] : 0u;
        v[q] = q < cnt ? ((uint64_t)a.idx[m] << 32) | ((uint64_t)a.xslab[m] << 30) | m : ~0ull;
    }
#pragma unroll
    for (uint32_t k = 2; k <= BIN_SORT_CAP; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t i = 0; i < BIN_SORT_CAP; i++) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t lo = v[i] < v[l] ? v[i] : v[l], hi = v[i] < v[l] ? v[l] : v[i];
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
    uint32_t sl[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) {
        sl[q] = q < cnt ? ((uint32_t)v[q] >> 30) : XSLABS;
        c0 += sl[q] == 0 ? 1u : 0u;
        c1 += sl[q] == 1 ? 1u : 0u;
        c2 += sl[q] == 2 ? 1u : 0u;
    }
    const uint32_t x1 = beg + c0, x2 = x1 + c1, x3 = x2 + c2;
    *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) = make_uint4(beg, x1, x2, x3);
    uint32_t n0 = beg, n1 = x1, n2 = x2, n3 = x3;
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++)
        if (q < cnt) {
            sorted[beg + q] = (uint32_t)(v[q] >> 32);
            const uint32_t sq = sl[q];
            const uint32_t sp = sq == 0 ? n0 : (sq == 1 ? n1 : (sq == 2 ? n2 : n3));
            const uint32_t m = (uint32_t)v[q] & 0x3fffffffu;
            if (REC64) {
                a.rpos[m] = beg + q;
                a.spos[m] = sp;
            } else {
                a.rpos[m] = sp;
            }
            n0 += sq == 0 ? 1u : 0u;
            n1 += sq == 1 ? 1u : 0u;
            n2 += sq == 2 ? 1u : 0u;
            n3 += sq == 3 ? 1u : 0u;
        }
}

// Mixed form: this step's FP32 records (k_stable_gather32's values) from the
// slots; the records' index field holds the slot.  Slots of bins above
// BIN_SORT_CAP are ranked here.
__global__ __launch_bounds__(256) void k_m_gather32(MSet a, const uint32_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                                                    CellRec32* __restrict__ rec, uint32_t n, double x_min,
                                                    double y_min, double z_min, double voxel, uint32_t mnx,
                                                    uint32_t mny, float* __restrict__ rpart, uint2* __restrict__ mark,
                                                    uint32_t* __restrict__ xtflag) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const double4 pr = p < n ? a.prop[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (rpart) block_max_to((float)pr.x, rpart);
    if (p >= n) return;
    const uint32_t b = a.key[p];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos;
    if (end - beg <= BIN_SORT_CAP) {
        pos = a.rpos[p];                                  // placed by k_m_bin_sort32
    } else {
        const uint32_t id = a.idx[p], sc = a.xslab[p];
        uint32_t ri = 0, rs = 0;
        for (uint32_t u = beg; u < end; u++) {
            const uint32_t o = tmp[u];
            const uint32_t io = a.idx[o], so = a.xslab[o];
            ri += io < id ? 1u : 0u;
            rs += (so < sc || (so == sc && io < id)) ? 1u : 0u;
        }
        sorted[beg + ri] = id;
        pos = beg + rs;
        a.rpos[p] = pos;
    }
    a.kstep[p] = b;
    const Dbl3 x = a.pos[p];
    const uint32_t bi = b % mnx, bj = (b / mnx) % mny, bk = b / (mnx * mny);
    CellRec32 r;
    r.xg = (float)(x.x - x_min);
    r.y = bin_rel32(x.y, y_min, bj, voxel);
    r.z = bin_rel32(x.z, z_min, bk, voxel);
    r.x = bin_rel32(x.x, x_min, bi, voxel);
    r.r = (float)pr.x;
    r.rep = (float)pr.y;
    r.adh = (float)pr.z;
    r.idx = p;
    rec[pos] = r;
    if (mark) {
        bool fits;
        mark[pos] = mark_entry16(x.x - x_min, x.y - (y_min + bj * voxel), x.z - (z_min + bk * voxel), 0, fits);
        if (!fits) xtflag[0] = 1u;
    }
}

// AB2 (shaders/integrate.metal:56-79) on the slots, with this step's forces
// by record position; then the bin (shaders/mechanics.metal:61-72) and the
// place in it of the new position for the coming step (counts must be zero)
__global__ __launch_bounds__(256) void k_m_integrate(MSet a, const double4* __restrict__ vp,
                                                     uint32_t* __restrict__ counts, uint32_t n, double dt,
                                                     double x_min, double y_min, double z_min, double x_max,
                                                     double y_max, double z_max, double mx_min, double my_min,
                                                     double mz_min, double voxel, uint32_t mnx, uint32_t mny,
                                                     uint32_t mnz, int slabs) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < n;
    uint32_t key = 0;
    if (valid) {
        const double4 v = vp[a.rpos[p]];
        const Dbl3 pv = a.prev[p], x = a.pos[p];
        double px = x.x, py = x.y, pz = x.z;
        px += dt * (1.5 * v.x - 0.5 * pv.x);
        py += dt * (1.5 * v.y - 0.5 * pv.y);
        pz += dt * (1.5 * v.z - 0.5 * pv.z);
        px = px < x_min ? x_min : (px > x_max ? x_max : px);
        py = py < y_min ? y_min : (py > y_max ? y_max : py);
        pz = pz < z_min ? z_min : (pz > z_max ? z_max : pz);
        a.pos[p] = dbl3(px, py, pz);
        a.prev[p] = dbl3(v.x, v.y, v.z);
        int vi = (int)((px - mx_min) / voxel);
        int vj = (int)((py - my_min) / voxel);
        int vk = (int)((pz - mz_min) / voxel);
        vi = vi < 0 ? 0 : (vi > (int)mnx - 1 ? (int)mnx - 1 : vi);
        vj = vj < 0 ? 0 : (vj > (int)mny - 1 ? (int)mny - 1 : vj);
        vk = vk < 0 ? 0 : (vk > (int)mnz - 1 ? (int)mnz - 1 : vk);
        key = (uint32_t)vk * mny * mnx + (uint32_t)vj * mnx + (uint32_t)vi;
        a.key[p] = key;
        if (slabs) a.xslab[p] = (uint8_t)xslab_of(bin_rel32(px, mx_min, (uint32_t)vi, voxel), xslab_scale(voxel));
    }
    const uint32_t s = ballot_bin(valid, key, counts);
    if (valid) a.slot[p] = s;
}

// the cell fields from the slots, by agent index
__global__ __launch_bounds__(256) void k_m_sync(MSet a, const double4* __restrict__ vp, CellPtrs cp, uint32_t n) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const uint32_t c = a.idx[p];
    const Dbl3 x = a.pos[p], pv = a.prev[p];
    const double4 v = vp[a.rpos[p]];
    cp.f[HC_F_POS_X][c] = x.x;
    cp.f[HC_F_POS_Y][c] = x.y;
    cp.f[HC_F_POS_Z][c] = x.z;
    cp.f[HC_F_VEL_X][c] = v.x;
    cp.f[HC_F_VEL_Y][c] = v.y;
    cp.f[HC_F_VEL_Z][c] = v.z;
    cp.f[HC_F_PRESSURE][c] = v.w;
    cp.f[HC_F_PREV_VX][c] = pv.x;
    cp.f[HC_F_PREV_VY][c] = pv.y;
    cp.f[HC_F_PREV_VZ][c] = pv.z;
    cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS][c] = a.kstep[p];
}

// re-sort: each slot's state to its record position of the last step (set
// b), then back into set a slot by slot
__global__ __launch_bounds__(256) void k_m_permute(MSet a, MSet b, uint32_t n) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = a.rpos[p];
    b.pos[q] = a.pos[p];
    b.prev[q] = a.prev[p];
    b.prop[q] = a.prop[p];
    b.idx[q] = a.idx[p];
    b.key[q] = a.key[p];
    b.slot[q] = a.slot[p];
    b.kstep[q] = a.kstep[p];
    b.xslab[q] = a.xslab[p];
    b.rpos[q] = q;
}
__global__ __launch_bounds__(256) void k_m_copy(MSet a, MSet b, uint32_t n) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    a.pos[p] = b.pos[p];
    a.prev[p] = b.prev[p];
    a.prop[p] = b.prop[p];
    a.idx[p] = b.idx[p];
    a.key[p] = b.key[p];
    a.slot[p] = b.slot[p];
    a.kstep[p] = b.kstep[p];
    a.xslab[p] = b.xslab[p];
    a.rpos[p] = b.rpos[p];
}

int probe_kernels_mirror(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_m_load, (const void*)k_m_bin_sort, (const void*)k_m_gather,
                         (const void*)k_m_integrate, (const void*)k_m_sync, (const void*)k_m_permute,
                         (const void*)k_m_copy, (const void*)k_m_load32, (const void*)k_m_bin_sort32<false>,
                         (const void*)k_m_bin_sort32<true>, (const void*)k_m_gather32, (const void*)k_forces_xt};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: mechanics mirror kernel");
    }
    return 0;
}

static int mirror_check(hc_ctx* ctx, hc_buf* hash, hc_buf* mirror, const hc_mech* mech, uint32_t n,
                        uint32_t max_cells, uint32_t n_mvox, const char* who) {
    if (!ctx || !hash || !mirror || !mech) return fail(ctx, HC_ERR_ARG, std::string(who) + ": null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, std::string(who) + ": n > max_cells");
    // slots take 30 bits beside the x slab in the bin sort's values
    if (max_cells > (1u << 30)) return fail(ctx, HC_ERR_ARG, std::string(who) + ": max_cells above 2^30");
    if (hc_buf_size(hash) < hc_hash_bytes(n_mvox, max_cells) || hc_buf_size(mirror) < mirror_layout(max_cells).total)
        return fail(ctx, HC_ERR_ARG, std::string(who) + ": buffer too small");
    return 0;
}

static MSet mirror_set(hc_buf* mirror, uint32_t max_cells, int k) {
    const MirrorLayout L = mirror_layout(max_cells);
    uint8_t* P = (uint8_t*)mirror->dptr;
    MSet m;
    m.pos = (Dbl3*)(P + L.pos[k]);
    m.prev = (Dbl3*)(P + L.prev[k]);
    m.prop = (double4*)(P + L.prop[k]);
    m.idx = (uint32_t*)(P + L.idx[k]);
    m.key = (uint32_t*)(P + L.key[k]);
    m.slot = (uint32_t*)(P + L.slot[k]);
    m.kstep = (uint32_t*)(P + L.kstep[k]);
    m.rpos = (uint32_t*)(P + L.rpos[k]);
    m.xslab = (uint8_t*)(P + L.xslab[k]);
    m.spos = (uint32_t*)(P + L.spos);
    return m;
}
static double4* mirror_vp(hc_buf* mirror, uint32_t max_cells) {
    return (double4*)((uint8_t*)mirror->dptr + mirror_layout(max_cells).vp);
}

int mirror_load(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, hc_buf* mirror, const hc_mech* mech, const hc_grid* mgrid,
                uint32_t n, uint32_t max_cells, uint32_t n_mvox) {
    int r = mirror_check(ctx, hash, mirror, mech, n, max_cells, n_mvox, "mirror_load");
    if (r || (r = hc_build_hash(ctx, cells, hash, mech, mgrid, n, max_cells, n_mvox))) return r;
    if (!n) return 0;
    const HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    if (mech->forces_fp32)
        hipLaunchKernelGGL(k_m_load32, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                           reinterpret_cast<const CellRec32*>(h.rec), h.offsets, h.xslab,
                           cell_ptrs(cells->dptr, max_cells), mirror_set(mirror, max_cells, 0), n);
    else
        hipLaunchKernelGGL(k_m_load, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, h.rec, h.offsets,
                           cell_ptrs(cells->dptr, max_cells), mirror_set(mirror, max_cells, 0), n, mgrid->x_min,
                           mech->mechanics_voxel_size, mech->mech_grid_nx, xtrim64_ok(mech, n_mvox) ? 1 : 0);
    return check_launch(ctx, "k_m_load");
}

int mirror_step(hc_ctx* ctx, hc_buf* hash, hc_buf* mirror, const hc_mech* mech, const hc_grid* mgrid, double dt,
                uint32_t n, uint32_t max_cells, uint32_t n_mvox) {
    int r = mirror_check(ctx, hash, mirror, mech, n, max_cells, n_mvox, "mirror_step");
    if (r) return r;
    if (!mgrid || (uint64_t)mech->mech_grid_nx * mech->mech_grid_ny * mech->mech_grid_nz != n_mvox)
        return fail(ctx, HC_ERR_ARG, "mirror_step: n_mvox != mech grid size");
    if (!n) return 0;
    const HashPtrs h = hash_ptrs(hash, n_mvox, max_cells);
    const MSet a = mirror_set(mirror, max_cells, 0);
    double4* vp = mirror_vp(mirror, max_cells);
    const unsigned nb = blocks_for(n, 256);
    const int xg = markx_ok(mech);
    const int fp32 = mech->forces_fp32 ? 1 : 0;
    const int xt = xtrim64_ok(mech, n_mvox) ? 1 : 0;
    const bool list32 = fp32 && list32_ok(mech, n, n_mvox);
    if (fp32) {
        // the mixed form: FP32 records in (bin, x slab, index) order
        {
            KernelTimer tm(ctx, HC_K_HASH);
            if ((r = scan_exclusive(ctx, h.counts, h.offsets, n_mvox, h.partial, h.counts))) return r;
            hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, ctx->stream, a.key, a.slot, h.offsets, h.tmp, n);
            hipLaunchKernelGGL(k_m_bin_sort32<false>, dim3(blocks_for(n_mvox + 1, 256)), dim3(256), 0, ctx->stream,
                               h.offsets, h.tmp, a, h.sorted, h.xoff, n_mvox, list32 ? h.xtflag : nullptr);
            hipLaunchKernelGGL(k_m_gather32, dim3(nb), dim3(256), 0, ctx->stream, a, h.offsets, h.tmp, h.sorted,
                               reinterpret_cast<CellRec32*>(h.rec), n, mgrid->x_min, mgrid->y_min, mgrid->z_min,
                               mech->mechanics_voxel_size, mech->mech_grid_nx, mech->mech_grid_ny, (float*)nullptr,
                               list32 ? reinterpret_cast<uint2*>(h.pos32) : nullptr, h.xtflag);
            if ((r = check_launch(ctx, "mirror hash (mixed)"))) return r;
        }
        KernelTimer tm(ctx, HC_K_FORCES);
        hipLaunchKernelGGL((list32 ? k_forces32<true, true> : k_forces32<false, false>), dim3(nb), dim3(256), 0,
                           ctx->stream, reinterpret_cast<const CellRec32*>(h.rec), h.offsets, h.xoff, a.key,
                           (const double*)a.prop, (double*)nullptr, (double*)nullptr, (double*)nullptr,
                           (double*)nullptr, n, mech->mech_grid_nx, mech->mech_grid_ny, mech->mech_grid_nz,
                           (float)mech->max_interaction_distance, mech->mechanics_voxel_size, h.rmax, vp,
                           reinterpret_cast<const uint2*>(h.pos32), h.xtflag);
        if ((r = check_launch(ctx, "k_forces32 (mirror)"))) return r;
    } else {
        // the slots' bins and places in them come from the last AB2 (or load)
        KernelTimer tm(ctx, HC_K_HASH);
        if ((r = scan_exclusive(ctx, h.counts, h.offsets, n_mvox, h.partial, h.counts))) return r;
        hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, ctx->stream, a.key, a.slot, h.offsets, h.tmp, n);
        if (xt)
            hipLaunchKernelGGL(k_m_bin_sort32<true>, dim3(blocks_for(n_mvox + 1, 256)), dim3(256), 0, ctx->stream,
                               h.offsets, h.tmp, a, h.sorted, h.xoff, n_mvox, h.xtflag);
        else
            hipLaunchKernelGGL(k_m_bin_sort, dim3(blocks_for(n_mvox, 256)), dim3(256), 0, ctx->stream, h.offsets,
                               h.tmp, a, h.sorted, n_mvox);
        hipLaunchKernelGGL(k_m_gather, dim3(nb), dim3(256), 0, ctx->stream, a, h.offsets, h.tmp, h.sorted, h.rec,
                           mark32_ok(mech) ? h.pos32 : nullptr, n, mgrid->x_min, mgrid->y_min, mgrid->z_min,
                           mech->mechanics_voxel_size, mech->mech_grid_nx, mech->mech_grid_ny, xg, xt, h.xtflag,
                           (float*)nullptr);
        if ((r = check_launch(ctx, "mirror hash"))) return r;
    }
    if (!fp32) {
        KernelTimer tm(ctx, HC_K_FORCES);
        if (xt)
            hipLaunchKernelGGL(k_forces_xt, dim3(nb), dim3(256), 0, ctx->stream, h.rec,
                               reinterpret_cast<const uint2*>(h.pos32), h.offsets, h.xoff, h.xtflag, n,
                               mech->mech_grid_nx, mech->mech_grid_ny, mech->mech_grid_nz,
                               mech->max_interaction_distance, mech->mechanics_voxel_size,
                               reinterpret_cast<const double*>(h.xtflag + 2), h.rmax, vp, (double*)nullptr,
                               (double*)nullptr, (double*)nullptr, (double*)nullptr);
        else
            hipLaunchKernelGGL(k_forces, dim3(nb), dim3(256), 0, ctx->stream, h.rec,
                               mark32_ok(mech) ? h.pos32 : nullptr, h.offsets, (double*)nullptr, (double*)nullptr,
                               (double*)nullptr, (double*)nullptr, n, mech->mech_grid_nx, mech->mech_grid_ny,
                               mech->mech_grid_nz, mech->max_interaction_distance,
                               (float)mech->mechanics_voxel_size, xg, h.rmax, vp);
        if ((r = check_launch(ctx, "k_forces (mirror)"))) return r;
    }
    KernelTimer tm(ctx, HC_K_INTEGRATE);
    const double x_max = mgrid->x_min + (double)mgrid->nx * mgrid->dx;
    const double y_max = mgrid->y_min + (double)mgrid->ny * mgrid->dy;
    const double z_max = mgrid->z_min + (double)mgrid->nz * mgrid->dz;
    hipLaunchKernelGGL(k_m_integrate, dim3(nb), dim3(256), 0, ctx->stream, a, (const double4*)vp, h.counts, n, dt,
                       mgrid->x_min, mgrid->y_min, mgrid->z_min, x_max, y_max, z_max, mgrid->x_min, mgrid->y_min,
                       mgrid->z_min, mech->mechanics_voxel_size, mech->mech_grid_nx, mech->mech_grid_ny,
                       mech->mech_grid_nz, fp32 || xt ? 1 : 0);
    return check_launch(ctx, "k_m_integrate");
}

int mirror_resort(hc_ctx* ctx, hc_buf* mirror, uint32_t n, uint32_t max_cells) {
    if (!ctx || !mirror || n > max_cells || hc_buf_size(mirror) < mirror_layout(max_cells).total)
        return fail(ctx, HC_ERR_ARG, "mirror_resort: bad argument");
    if (!n) return 0;
    const MSet a = mirror_set(mirror, max_cells, 0), b = mirror_set(mirror, max_cells, 1);
    hipLaunchKernelGGL(k_m_permute, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a, b, n);
    hipLaunchKernelGGL(k_m_copy, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, a, b, n);
    return check_launch(ctx, "mirror re-sort");
}

int mirror_sync(hc_ctx* ctx, hc_buf* cells, hc_buf* mirror, uint32_t n, uint32_t max_cells) {
    int r = check_cells(ctx, cells, n, max_cells, "mirror_sync");
    if (r) return r;
    if (!mirror || hc_buf_size(mirror) < mirror_layout(max_cells).total)
        return fail(ctx, HC_ERR_ARG, "mirror_sync: mirror buffer too small");
    if (!n) return 0;
    hipLaunchKernelGGL(k_m_sync, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                       mirror_set(mirror, max_cells, 0), (const double4*)mirror_vp(mirror, max_cells),
                       cell_ptrs(cells->dptr, max_cells), n);
    return check_launch(ctx, "k_m_sync");
}

}  // namespace hc
