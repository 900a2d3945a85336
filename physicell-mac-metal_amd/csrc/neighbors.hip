// neighbors.hip -- neighbour queries on the device hash for the phenotype
// side: the contact graph of writeCellNeighborGraph
// (src/output_writer.cpp:457-490, an O(N^2) pair loop in the reference), the
// interaction candidates of standardCellCellInteractions
// (src/cell_phenotype.cpp:858-968, a Moore-neighbourhood scan of the
// mechanics hash) and the attachment candidates of tryFormAllAttachments
// (src/cell_interactions.cpp:411-478, another O(N^2) pair loop).
//
// Both visit the 27 (9 in 2-D) bins around a cell through the hash's
// offsets and its ascending per-bin cell list, and read the current positions
// and radii from the cell SoA in FP64.  Two passes give a CSR result: count
// per cell, exclusive scan, fill.  Contact and attachment lists are sorted
// by neighbour index (the order of the reference's j loops); interaction
// lists keep the visiting order (dk, dj, di, then ascending index).
#include "cells_common.h"

#include <algorithm>
#include <cstring>

namespace hc {

struct QueryArgs {
    CellPtrs cp;
    const uint32_t* bin_off;    // n_mvox + 1
    const uint32_t* sorted;     // hash-sorted cell indices
    uint32_t n, mnx, mny, mnz;
};

// contact: j > i, both alive with volume above 1e-15, |d|^2 <= (max(rel_i,
// rel_j) (r_i + r_j))^2 in 3-D (src/output_writer.cpp:466-486)
// interact: the reference's candidate scan: up to 512 cells j != i of the
// Moore neighbourhood in visiting order, then volume_j >= 1e-15, 2-D
// distance^2 <= (2.5 r_i + r_j)^2 and distance <= 1.25 (r_i + r_j)
// (src/cell_phenotype.cpp:893-967)
// attach: j > i, both alive, sqrt(|d|^2) <= max(rel_i, rel_j) (r_i + r_j)
// with d = p_j - p_i (tryFormAttachment's distance test,
// src/cell_interactions.cpp:429-440); rates, capacity and draws are the host's
template <int KIND, bool FILL>
__global__ __launch_bounds__(256) void k_query(QueryArgs a, uint32_t* __restrict__ counts,
                                               const uint32_t* __restrict__ offsets, uint32_t* __restrict__ idx,
                                               uint32_t* __restrict__ maxes) {
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const bool valid = t < a.n;
    const uint32_t i = valid ? a.sorted[t] : 0u;
    const double* const* f = a.cp.f;
    const uint32_t* alive = a.cp.u[HC_F_IS_ALIVE - HC_NUM_F64_FIELDS];
    if (!FILL && KIND != 1 && maxes) {
        // largest radius and adhesion scale over the wave, then one atomic
        // per wave: the host checks that every contact distance fits in one
        // mechanics voxel (non-negative doubles compare as their bit patterns)
        double mr = valid && f[HC_F_RADIUS][i] > 0.0 ? f[HC_F_RADIUS][i] : 0.0;
        double ml = valid && f[HC_F_REL_MAX_ADH][i] > 0.0 ? f[HC_F_REL_MAX_ADH][i] : 0.0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const double orr = __shfl_xor(mr, d, 64), orl = __shfl_xor(ml, d, 64);
            mr = orr > mr ? orr : mr;
            ml = orl > ml ? orl : ml;
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax((unsigned long long*)maxes, (unsigned long long)__double_as_longlong(mr));
            atomicMax((unsigned long long*)maxes + 1, (unsigned long long)__double_as_longlong(ml));
        }
    }
    if (!valid) return;
    const double xi = f[HC_F_POS_X][i], yi = f[HC_F_POS_Y][i], zi = f[HC_F_POS_Z][i], ri = f[HC_F_RADIUS][i];
    const double reli = f[HC_F_REL_MAX_ADH][i];
    uint32_t cnt = 0;
    const uint32_t base = FILL ? offsets[i] : 0u;
    if ((KIND == 0 && !(alive[i] && f[HC_F_TOTAL_VOL][i] > 1e-15)) || (KIND == 2 && !alive[i])) {
        if (!FILL) counts[i] = 0;
        return;
    }
    const uint32_t key = a.cp.u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS][i];
    const uint32_t vi = key % a.mnx, vj = (key / a.mnx) % a.mny, vk = key / (a.mnx * a.mny);
    const int dk0 = a.mnz > 1 ? -1 : 0, dk1 = a.mnz > 1 ? 1 : 0;
    uint32_t visited = 0;
    for (int dk = dk0; dk <= dk1; dk++) {
        const int nk = (int)vk + dk;
        if (nk < 0 || nk >= (int)a.mnz) continue;
        for (int dj = -1; dj <= 1; dj++) {
            const int nj = (int)vj + dj;
            if (nj < 0 || nj >= (int)a.mny) continue;
            for (int di = -1; di <= 1; di++) {
                const int ni = (int)vi + di;
                if (ni < 0 || ni >= (int)a.mnx) continue;
                const uint32_t b = ((uint32_t)nk * a.mny + (uint32_t)nj) * a.mnx + (uint32_t)ni;
                for (uint32_t u = a.bin_off[b]; u < a.bin_off[b + 1]; u++) {
                    const uint32_t j = a.sorted[u];
                    bool take;
                    if (KIND == 0) {
                        if (j <= i || !alive[j] || !(f[HC_F_TOTAL_VOL][j] > 1e-15)) continue;
                        const double dx = xi - f[HC_F_POS_X][j], dy = yi - f[HC_F_POS_Y][j],
                                     dz = zi - f[HC_F_POS_Z][j];
                        const double relj = f[HC_F_REL_MAX_ADH][j];
                        const double md = (reli > relj ? reli : relj) * (ri + f[HC_F_RADIUS][j]);
                        take = dx * dx + dy * dy + dz * dz <= md * md;
                    } else if (KIND == 2) {
                        if (j <= i || !alive[j]) continue;
                        const double dx = f[HC_F_POS_X][j] - xi, dy = f[HC_F_POS_Y][j] - yi,
                                     dz = f[HC_F_POS_Z][j] - zi;
                        const double relj = f[HC_F_REL_MAX_ADH][j];
                        const double md = (reli < relj ? relj : reli) * (ri + f[HC_F_RADIUS][j]);
                        take = !(sqrt(dx * dx + dy * dy + dz * dz) > md);
                    } else {
                        if (j == i || j >= a.n) continue;
                        if (visited >= 512u) continue;                  // the reference's MAX_NEIGHBORS
                        visited++;
                        if (f[HC_F_TOTAL_VOL][j] < 1e-15) continue;
                        const double dx = f[HC_F_POS_X][j] - xi, dy = f[HC_F_POS_Y][j] - yi;
                        const double d2 = dx * dx + dy * dy, rj = f[HC_F_RADIUS][j];
                        const double max_r = 2.5 * ri + rj;
                        take = !(d2 > max_r * max_r) && !(sqrt(d2) > (ri + rj) * 1.25);
                    }
                    if (take) {
                        if (FILL) idx[base + cnt] = j;
                        cnt++;
                    }
                }
            }
        }
    }
    if (!FILL) {
        counts[i] = cnt;
    } else if (KIND != 1) {
        // ascending neighbour index (insertion sort of this cell's short list)
        for (uint32_t p = 1; p < cnt; p++) {
            const uint32_t v = idx[base + p];
            uint32_t q = p;
            while (q > 0 && idx[base + q - 1] > v) {
                idx[base + q] = idx[base + q - 1];
                q--;
            }
            idx[base + q] = v;
        }
    }
}

int probe_kernels_neighbors(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_query<0, false>, (const void*)k_query<0, true>,
                         (const void*)k_query<1, false>, (const void*)k_query<1, true>,
                         (const void*)k_query<2, false>, (const void*)k_query<2, true>};
    for (const void* fn : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, fn);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: neighbour query kernel");
    }
    return 0;
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_neighbor_query(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, uint32_t n, uint32_t max_cells,
                      uint32_t n_mvox, int kind, hc_buf* offsets, hc_buf* idx, uint64_t* total) {
    if (!ctx || !cells || !hash || !mech || !offsets || !total)
        return fail(ctx, HC_ERR_ARG, "hc_neighbor_query: null argument");
    if (kind != HC_QUERY_CONTACT && kind != HC_QUERY_INTERACT && kind != HC_QUERY_ATTACH)
        return fail(ctx, HC_ERR_ARG, "hc_neighbor_query: unknown kind");
    if (n > max_cells || hc_buf_size(cells) < hc_cells_bytes(max_cells))
        return fail(ctx, HC_ERR_ARG, "hc_neighbor_query: cell buffer");
    if ((uint64_t)mech->mech_grid_nx * mech->mech_grid_ny * mech->mech_grid_nz != n_mvox ||
        hc_buf_size(hash) < hc_hash_bytes(n_mvox, max_cells))
        return fail(ctx, HC_ERR_ARG, "hc_neighbor_query: hash buffer or mech grid");
    if (hc_buf_size(offsets) < ((size_t)n + 1) * 4) return fail(ctx, HC_ERR_ARG, "hc_neighbor_query: offsets buffer");
    const HashLayout L = hash_layout(n_mvox, max_cells);
    QueryArgs a;
    a.cp = cell_ptrs(cells->dptr, max_cells);
    a.bin_off = (const uint32_t*)((uint8_t*)hash->dptr + L.offsets);
    a.sorted = (const uint32_t*)((uint8_t*)hash->dptr + L.sorted);
    a.n = n;
    a.mnx = mech->mech_grid_nx;
    a.mny = mech->mech_grid_ny;
    a.mnz = mech->mech_grid_nz;
    uint32_t* off = (uint32_t*)offsets->dptr;
    // scratch: the scan's partials after the two maxima
    const size_t need = 16 + scan_partial_bytes((uint64_t)n + 1);
    if (ctx->d_query_bytes < need) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_neighbor_query: wait");
        if (ctx->d_query) (void)hipFree(ctx->d_query);
        ctx->d_query = nullptr;
        ctx->d_query_bytes = 0;
        if ((e = hipMalloc(&ctx->d_query, need)) != hipSuccess) return hip_fail(ctx, e, "hc_neighbor_query: scratch");
        ctx->d_query_bytes = need;
    }
    uint32_t* maxes = (uint32_t*)ctx->d_query;
    uint32_t* partial = maxes + 4;
    int r = fill_u32(ctx, maxes, 4, 0);
    if (r) return r;
    const unsigned nb = blocks_for(n ? n : 1, 256);
    if (n) {
        if (kind == HC_QUERY_CONTACT)
            hipLaunchKernelGGL((k_query<0, false>), dim3(nb), dim3(256), 0, ctx->stream, a, off, nullptr, nullptr,
                               maxes);
        else if (kind == HC_QUERY_ATTACH)
            hipLaunchKernelGGL((k_query<2, false>), dim3(nb), dim3(256), 0, ctx->stream, a, off, nullptr, nullptr,
                               maxes);
        else
            hipLaunchKernelGGL((k_query<1, false>), dim3(nb), dim3(256), 0, ctx->stream, a, off, nullptr, nullptr,
                               nullptr);
        if ((r = check_launch(ctx, "k_query count"))) return r;
    }
    if ((r = scan_exclusive(ctx, off, off, n, partial))) return r;
    uint32_t host[5] = {0, 0, 0, 0, 0};
    hipError_t e = hipMemcpyAsync(host, off + n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host + 1, maxes, 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hc_neighbor_query: read back");
    *total = host[0];
    if (kind != HC_QUERY_INTERACT && n) {
        double rmax, relmax;
        std::memcpy(&rmax, host + 1, 8);
        std::memcpy(&relmax, host + 3, 8);
        const double reach = (relmax > 0.0 ? relmax : 0.0) * 2.0 * rmax;
        if (reach > mech->mechanics_voxel_size)
            return fail(ctx, HC_ERR_CAPACITY, "hc_neighbor_query: a contact or attachment distance (max scale x 2 max radius = " +
                                                  std::to_string(reach) + ") exceeds the mechanics voxel");
    }
    if (!idx) return 0;
    if (hc_buf_size(idx) < (size_t)host[0] * 4)
        return fail(ctx, HC_ERR_CAPACITY, "hc_neighbor_query: idx buffer holds fewer than *total entries");
    if (n && host[0]) {
        if (kind == HC_QUERY_CONTACT)
            hipLaunchKernelGGL((k_query<0, true>), dim3(nb), dim3(256), 0, ctx->stream, a, nullptr, off,
                               (uint32_t*)idx->dptr, nullptr);
        else if (kind == HC_QUERY_ATTACH)
            hipLaunchKernelGGL((k_query<2, true>), dim3(nb), dim3(256), 0, ctx->stream, a, nullptr, off,
                               (uint32_t*)idx->dptr, nullptr);
        else
            hipLaunchKernelGGL((k_query<1, true>), dim3(nb), dim3(256), 0, ctx->stream, a, nullptr, off,
                               (uint32_t*)idx->dptr, nullptr);
        return check_launch(ctx, "k_query fill");
    }
    return 0;
}

}  // extern "C"
