// phenotype.hip -- the SoA side of updatePhenotype on the device: the volume
// model and radii (src/cell_phenotype.cpp:52-113), division
// (src/cell_phenotype.cpp:571-720) and removal (src/cell_phenotype.cpp:833-844
// with CellData::removeCell, src/cell_data.cpp:408-542).  The decisions of
// which cells divide and which are removed belong to the cycle and death
// models and come from the host as index lists.
//
// Values are FP64 (the reference keeps volumes and radii in FP32 and the
// compartments in FP64).  Division draws its angle and its oncoprotein noise
// from the counter-based stream (uniform01) keyed by (seed, division
// counter, position in the list, draw); the noise is a Box-Muller normal.
#include "cells_common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace hc {

constexpr double FOUR_THIRDS_PI = 4.188790204786391;       // src/cell_phenotype.cpp:45

struct VolPtrs {
    double* f[HC_NUM_V_FIELDS];
    uint64_t* hdr;       // [0] division counter, [1] seed
};

inline VolPtrs vol_ptrs(void* base, uint32_t max_cells) {
    VolPtrs v;
    v.hdr = (uint64_t*)base;
    for (int f = 0; f < HC_NUM_V_FIELDS; f++)
        v.f[f] = (double*)((uint8_t*)base + hc_volume_field_offset(f, max_cells));
    return v;
}

// updateVolume, one cell per lane
__global__ __launch_bounds__(256) void k_volume(CellPtrs cp, VolPtrs vp, const uint32_t* __restrict__ mask, uint32_t n,
                                                double dt, int use_ratio) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || (mask && !mask[i])) return;
    double total = cp.f[HC_F_TOTAL_VOL][i];
    double V_fluid = vp.f[HC_V_FLUID][i];
    double nuclear_solid = vp.f[HC_V_SOLID_NUC][i];
    double cyto_solid = vp.f[HC_V_SOLID_CYTO][i];
    // fluid (:59-61)
    V_fluid += dt * vp.f[HC_V_FLUID_RATE][i] * (vp.f[HC_V_TARGET_FLUID_FRAC][i] * total - V_fluid);
    if (V_fluid < 0.0) V_fluid = 0.0;
    // split of the fluid (:64-66)
    double nuclear = cp.f[HC_F_NUCLEAR_VOL][i];
    const double nuclear_fluid = (nuclear / (total + 1e-16)) * V_fluid;
    const double cyto_fluid = V_fluid - nuclear_fluid;
    // nuclear solid (:70-72)
    nuclear_solid += dt * vp.f[HC_V_NUC_RATE][i] * (vp.f[HC_V_TARGET_SOLID_NUC][i] - nuclear_solid);
    if (nuclear_solid < 0.0) nuclear_solid = 0.0;
    // cytoplasmic solid (:78-89)
    double target_sc = vp.f[HC_V_TARGET_SOLID_CYTO][i];
    if (use_ratio) {
        target_sc = vp.f[HC_V_TARGET_CN_RATIO][i] * vp.f[HC_V_TARGET_SOLID_NUC][i];
        vp.f[HC_V_TARGET_SOLID_CYTO][i] = target_sc;
    }
    cyto_solid += dt * vp.f[HC_V_CYTO_RATE][i] * (target_sc - cyto_solid);
    if (cyto_solid < 0.0) cyto_solid = 0.0;
    // calcification (:92-93)
    double calc_frac = vp.f[HC_V_CALC_FRAC][i];
    calc_frac += dt * vp.f[HC_V_CALC_RATE][i] * (1.0 - calc_frac);
    // totals and geometry (:96-112)
    nuclear = nuclear_solid + nuclear_fluid;
    const double cytoplasmic = cyto_solid + cyto_fluid;
    total = cytoplasmic + nuclear;
    const double fluid_fraction = V_fluid / (total + 1e-16);
    cp.f[HC_F_TOTAL_VOL][i] = total;
    cp.f[HC_F_NUCLEAR_VOL][i] = nuclear;
    vp.f[HC_V_FLUID][i] = V_fluid;
    vp.f[HC_V_SOLID_NUC][i] = nuclear_solid;
    vp.f[HC_V_SOLID_CYTO][i] = cyto_solid;
    vp.f[HC_V_FLUID_FRAC][i] = fluid_fraction;
    vp.f[HC_V_CALC_FRAC][i] = calc_frac;
    cp.f[HC_F_RADIUS][i] = cbrt(total / FOUR_THIRDS_PI);
    cp.f[HC_F_NUCLEAR_RAD][i] = cbrt(nuclear / FOUR_THIRDS_PI);
}

// divideCell, one division per lane: parent parents[k], daughter n0 + k
__global__ __launch_bounds__(256) void k_divide(CellPtrs cp, double* __restrict__ rates, VolPtrs vp, MotPtrs mp,
                                                int has_mot, const uint32_t* __restrict__ parents, uint32_t count,
                                                uint32_t n0, uint32_t max_cells, uint32_t ns) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const uint32_t p = parents[k], d = n0 + k;
    const uint64_t step = vp.hdr[0], seed = vp.hdr[1];
    // CellData's fresh values for everything divideCell does not set
    // (src/cell_data.cpp:544-640)
    for (int f = 0; f < HC_NUM_F64_FIELDS; f++) cp.f[f][d] = 0.0;
    cp.f[HC_F_ONCOPROTEIN][d] = 1.0;
    // halve the volumes and compartments (:579-600)
    const double vt = cp.f[HC_F_TOTAL_VOL][p] * 0.5, vn = cp.f[HC_F_NUCLEAR_VOL][p] * 0.5;
    cp.f[HC_F_TOTAL_VOL][p] = cp.f[HC_F_TOTAL_VOL][d] = vt;
    cp.f[HC_F_NUCLEAR_VOL][p] = cp.f[HC_F_NUCLEAR_VOL][d] = vn;
    const int halved[] = {HC_V_FLUID, HC_V_SOLID_CYTO, HC_V_SOLID_NUC, HC_V_RUPTURE_VOL, HC_V_TARGET_SOLID_NUC,
                          HC_V_TARGET_SOLID_CYTO};
    for (int f : halved) {
        const double v = vp.f[f][p] * 0.5;
        vp.f[f][p] = vp.f[f][d] = v;
    }
    const int inherited[] = {HC_V_FLUID_FRAC, HC_V_CYTO_RATE, HC_V_NUC_RATE, HC_V_FLUID_RATE, HC_V_CALC_RATE,
                             HC_V_CALC_FRAC, HC_V_TARGET_FLUID_FRAC, HC_V_REL_RUPTURE_VOL};
    for (int f : inherited) vp.f[f][d] = vp.f[f][p];
    vp.f[HC_V_TARGET_CN_RATIO][d] = 3.62;
    // radii of both (:603-610)
    const double r = cbrt(vt / FOUR_THIRDS_PI), nr = cbrt(vn / FOUR_THIRDS_PI);
    cp.f[HC_F_RADIUS][p] = cp.f[HC_F_RADIUS][d] = r;
    cp.f[HC_F_NUCLEAR_RAD][p] = cp.f[HC_F_NUCLEAR_RAD][d] = nr;
    // position: half a radius away at a random angle in the x-y plane (:613-618)
    const double pi = 3.14159265358979323846;
    const double theta = uniform01(seed, step, k, 0) * 2.0 * pi;
    const double offset = r * 0.5;
    cp.f[HC_F_POS_X][d] = cp.f[HC_F_POS_X][p] + offset * cos(theta);
    cp.f[HC_F_POS_Y][d] = cp.f[HC_F_POS_Y][p] + offset * sin(theta);
    cp.f[HC_F_POS_Z][d] = cp.f[HC_F_POS_Z][p];
    // type, mechanics, motility speed and bias (:621-627)
    const int copied[] = {HC_F_REPULSION, HC_F_ADHESION, HC_F_REL_MAX_ADH, HC_F_MOT_SPEED, HC_F_MOT_BIAS};
    for (int f : copied) cp.f[f][d] = cp.f[f][p];
    // oncoprotein noise ~ N(0, 0.02 * level), Box-Muller (:630-635)
    const double onco = cp.f[HC_F_ONCOPROTEIN][p];
    const double u1 = 1.0 - uniform01(seed, step, k, 1), u2 = uniform01(seed, step, k, 2);
    const double noise = 0.02 * onco * (sqrt(-2.0 * log(u1)) * cos(2.0 * pi * u2));
    const double od = onco + noise, op = onco - noise;
    cp.f[HC_F_ONCOPROTEIN][d] = od > 0.0 ? od : 0.0;
    cp.f[HC_F_ONCOPROTEIN][p] = op > 0.0 ? op : 0.0;
    // integer fields: type and voxels inherited, phase 0, alive (:621, :638-639, :713-716)
    uint32_t* const* u = cp.u;
    u[HC_F_CELL_TYPE - HC_NUM_F64_FIELDS][d] = u[HC_F_CELL_TYPE - HC_NUM_F64_FIELDS][p];
    u[HC_F_CUR_PHASE - HC_NUM_F64_FIELDS][p] = 0u;
    u[HC_F_CUR_PHASE - HC_NUM_F64_FIELDS][d] = 0u;
    u[HC_F_IS_ALIVE - HC_NUM_F64_FIELDS][d] = 1u;
    u[HC_F_VOXEL_IDX - HC_NUM_F64_FIELDS][d] = u[HC_F_VOXEL_IDX - HC_NUM_F64_FIELDS][p];
    u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS][d] = u[HC_F_MECH_VOXEL - HC_NUM_F64_FIELDS][p];
    // secretion, uptake, saturation, net export (:699-709)
    const uint64_t R = (uint64_t)max_cells * ns;
    for (int a = 0; a < 4; a++)
        for (uint32_t s = 0; s < ns; s++)
            rates[a * R + (uint64_t)d * ns + s] = rates[a * R + (uint64_t)p * ns + s];
    if (has_mot) {
        mp.persistence[d] = mp.persistence[p];
        mp.elapsed[d] = 0.0;
        mp.bias[d] = mp.bias[p];
        for (int c = 0; c < 3; c++) mp.vec[c][d] = mp.vec[c][p];
        mp.motile[d] = mp.motile[p];
        mp.restrict_2d[d] = mp.restrict_2d[p];
        mp.chem_index[d] = mp.chem_index[p];
        mp.chem_dir[d] = mp.chem_dir[p];
        for (uint32_t s = 0; s < ns; s++)
            mp.sensitivity[(uint64_t)d * ns + s] = mp.sensitivity[(uint64_t)p * ns + s];
    }
}

__global__ void k_divide_advance(uint64_t* hdr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) hdr[0] += 1;
}

// removal: copy cell src[k] into dst[k] (every dst < every src), one pair per lane
__global__ __launch_bounds__(256) void k_remove_copy(CellPtrs cp, double* __restrict__ rates, VolPtrs vp, MotPtrs mp,
                                                     int has_vol, int has_mot, const uint32_t* __restrict__ pairs,
                                                     uint32_t count, uint32_t max_cells, uint32_t ns) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const uint32_t d = pairs[2 * k], s = pairs[2 * k + 1];
    for (int f = 0; f < HC_NUM_F64_FIELDS; f++) cp.f[f][d] = cp.f[f][s];
    for (int f = 0; f < HC_NUM_U32_FIELDS; f++) cp.u[f][d] = cp.u[f][s];
    const uint64_t R = (uint64_t)max_cells * ns;
    for (int a = 0; a < 4; a++)
        for (uint32_t q = 0; q < ns; q++) rates[a * R + (uint64_t)d * ns + q] = rates[a * R + (uint64_t)s * ns + q];
    if (has_vol)
        for (int f = 0; f < HC_NUM_V_FIELDS; f++) vp.f[f][d] = vp.f[f][s];
    if (has_mot) {
        mp.persistence[d] = mp.persistence[s];
        mp.elapsed[d] = mp.elapsed[s];
        mp.bias[d] = mp.bias[s];
        for (int c = 0; c < 3; c++) mp.vec[c][d] = mp.vec[c][s];
        mp.motile[d] = mp.motile[s];
        mp.restrict_2d[d] = mp.restrict_2d[s];
        mp.chem_index[d] = mp.chem_index[s];
        mp.chem_dir[d] = mp.chem_dir[s];
        for (uint32_t q = 0; q < ns; q++)
            mp.sensitivity[(uint64_t)d * ns + q] = mp.sensitivity[(uint64_t)s * ns + q];
    }
}

int probe_kernels_phenotype(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_volume, (const void*)k_divide, (const void*)k_divide_advance,
                         (const void*)k_remove_copy};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: phenotype kernel");
    }
    return 0;
}

// A device copy of a small host list, in the context's list scratch.
static int upload_list(hc_ctx* ctx, const uint32_t* host, size_t count, uint32_t** dev) {
    const size_t bytes = std::max<size_t>(16, count * sizeof(uint32_t));
    if (ctx->d_list_bytes < bytes) {
        if (ctx->d_list) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return hip_fail(ctx, e, "list scratch: wait");
            (void)hipFree(ctx->d_list);
            ctx->d_list = nullptr;
            ctx->d_list_bytes = 0;
        }
        hipError_t e = hipMalloc(&ctx->d_list, bytes);
        if (e != hipSuccess) return hip_fail(ctx, e, "list scratch: hipMalloc");
        ctx->d_list_bytes = bytes;
    }
    if (count) {
        // the previous user of the scratch may still be queued: order the copy
        // on the stream, and wait for it so the host list may be freed
        hipError_t e = hipMemcpyAsync(ctx->d_list, host, count * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "list scratch: upload");
    }
    *dev = (uint32_t*)ctx->d_list;
    return 0;
}

}  // namespace hc

using namespace hc;

extern "C" {

size_t hc_volume_field_offset(uint32_t field, uint32_t max_cells) {
    return 256 + (size_t)field * max_cells * sizeof(double);
}

size_t hc_volume_bytes(uint32_t max_cells) { return hc_volume_field_offset(HC_NUM_V_FIELDS, max_cells); }

int hc_volume_update(hc_ctx* ctx, hc_buf* cells, hc_buf* volume, hc_buf* mask, uint32_t n, uint32_t max_cells,
                     double dt, int use_ratio) {
    if (!ctx || !cells || !volume) return fail(ctx, HC_ERR_ARG, "hc_volume_update: null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, "hc_volume_update: n > max_cells");
    if (hc_buf_size(cells) < hc_cells_bytes(max_cells) || hc_buf_size(volume) < hc_volume_bytes(max_cells))
        return fail(ctx, HC_ERR_ARG, "hc_volume_update: buffer too small");
    if (mask && hc_buf_size(mask) < (size_t)n * sizeof(uint32_t))
        return fail(ctx, HC_ERR_ARG, "hc_volume_update: mask buffer too small");
    if (!n) return 0;
    KernelTimer tm(ctx, HC_K_VOLUME);
    hipLaunchKernelGGL(k_volume, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                       cell_ptrs(cells->dptr, max_cells), vol_ptrs(volume->dptr, max_cells),
                       mask ? (const uint32_t*)mask->dptr : nullptr, n, dt, use_ratio);
    return check_launch(ctx, "k_volume");
}

int hc_divide(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, hc_buf* volume, hc_buf* motility, uint32_t ns,
              const uint32_t* parents, uint32_t count, uint32_t n, uint32_t max_cells, uint32_t* n_out) {
    if (!ctx || !cells || !rates || !volume || !n_out || (count && !parents))
        return fail(ctx, HC_ERR_ARG, "hc_divide: null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, "hc_divide: n > max_cells");
    if (hc_buf_size(cells) < hc_cells_bytes(max_cells) || hc_buf_size(volume) < hc_volume_bytes(max_cells) ||
        hc_buf_size(rates) < hc_rates_bytes(max_cells, ns) ||
        (motility && hc_buf_size(motility) < hc_motility_bytes(max_cells, ns)))
        return fail(ctx, HC_ERR_ARG, "hc_divide: buffer too small");
    // distinct parents below n (each division touches only its parent and its
    // own daughter, so the divisions run in parallel)
    std::vector<uint32_t> chk(parents, parents + count);
    std::sort(chk.begin(), chk.end());
    for (uint32_t i = 0; i < count; i++)
        if (chk[i] >= n || (i && chk[i] == chk[i - 1]))
            return fail(ctx, HC_ERR_ARG, "hc_divide: parents must be distinct cells below n");
    // as CellData::addCell, divisions beyond max_cells do not happen
    const uint32_t done = std::min(count, max_cells - n);
    *n_out = n + done;
    const VolPtrs vp = vol_ptrs(volume->dptr, max_cells);
    if (done) {
        uint32_t* dlist = nullptr;
        int rc = upload_list(ctx, parents, done, &dlist);
        if (rc) return rc;
        MotPtrs mp{};
        if (motility) mp = mot_ptrs(motility->dptr, max_cells, ns);
        KernelTimer tm(ctx, HC_K_DIVIDE);
        hipLaunchKernelGGL(k_divide, dim3(blocks_for(done, 256)), dim3(256), 0, ctx->stream,
                           cell_ptrs(cells->dptr, max_cells), (double*)rates->dptr, vp, mp, motility ? 1 : 0, dlist,
                           done, n, max_cells, ns);
    }
    hipLaunchKernelGGL(k_divide_advance, dim3(1), dim3(64), 0, ctx->stream, vp.hdr);
    return check_launch(ctx, "k_divide");
}

int hc_remove_plan(const uint32_t* indices, uint32_t count, uint32_t n, uint32_t* dst, uint32_t* src,
                   uint32_t* n_pairs) {
    if ((count && !indices) || !n_pairs || count > n) return HC_ERR_ARG;
    std::vector<uint32_t> idx(indices, indices + count);
    std::sort(idx.begin(), idx.end(), [](uint32_t a, uint32_t b) { return a > b; });
    for (uint32_t i = 0; i < count; i++)
        if (idx[i] >= n || (i && idx[i] == idx[i - 1])) return HC_ERR_ARG;
    // Replays the descending removeCell loop on labels: positions at or above
    // m = n - count only ever move into holes below m or further down the
    // tail, and cells below m never move, so each hole below m ends up holding
    // one original cell from [m, n)
    const uint32_t m = n - count;
    std::vector<uint32_t> lab(count);
    for (uint32_t q = 0; q < count; q++) lab[q] = m + q;
    uint32_t cur = n, np = 0;
    for (uint32_t i : idx) {
        const uint32_t last = cur - 1;
        if (i != last) {
            if (i >= m) {
                lab[i - m] = lab[last - m];
            } else {
                if (dst) dst[np] = i;
                if (src) src[np] = lab[last - m];
                np++;
            }
        }
        cur--;
    }
    *n_pairs = np;
    return 0;
}

int hc_remove(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, hc_buf* volume, hc_buf* motility, uint32_t ns,
              const uint32_t* indices, uint32_t count, uint32_t n, uint32_t max_cells, uint32_t* n_out) {
    if (!ctx || !cells || !rates || !n_out || (count && !indices))
        return fail(ctx, HC_ERR_ARG, "hc_remove: null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, "hc_remove: n > max_cells");
    if (hc_buf_size(cells) < hc_cells_bytes(max_cells) || hc_buf_size(rates) < hc_rates_bytes(max_cells, ns) ||
        (volume && hc_buf_size(volume) < hc_volume_bytes(max_cells)) ||
        (motility && hc_buf_size(motility) < hc_motility_bytes(max_cells, ns)))
        return fail(ctx, HC_ERR_ARG, "hc_remove: buffer too small");
    std::vector<uint32_t> dst(count), src(count);
    uint32_t np = 0;
    if (hc_remove_plan(indices, count, n, dst.data(), src.data(), &np))
        return fail(ctx, HC_ERR_ARG, "hc_remove: indices must be distinct cells below n");
    *n_out = n - count;
    if (!np) return 0;
    std::vector<uint32_t> pairs(2 * (size_t)np);
    for (uint32_t k = 0; k < np; k++) {
        pairs[2 * k] = dst[k];
        pairs[2 * k + 1] = src[k];
    }
    uint32_t* dpairs = nullptr;
    int rc = upload_list(ctx, pairs.data(), pairs.size(), &dpairs);
    if (rc) return rc;
    VolPtrs vp{};
    if (volume) vp = vol_ptrs(volume->dptr, max_cells);
    MotPtrs mp{};
    if (motility) mp = mot_ptrs(motility->dptr, max_cells, ns);
    KernelTimer tm(ctx, HC_K_REMOVE);
    hipLaunchKernelGGL(k_remove_copy, dim3(blocks_for(np, 256)), dim3(256), 0, ctx->stream,
                       cell_ptrs(cells->dptr, max_cells), (double*)rates->dptr, vp, mp, volume ? 1 : 0,
                       motility ? 1 : 0, dpairs, np, max_cells, ns);
    return check_launch(ctx, "k_remove_copy");
}

}  // extern "C"
