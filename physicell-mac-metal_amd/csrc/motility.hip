// motility.hip -- motility on the device (src/motility.cpp:179-293).
//
// Between the force step and AB2, every live motile cell adds its motility
// vector to its velocity; the vector is resampled (persistent random walk
// biased along substrate gradients) when the persistence time is below dt or
// a uniform draw is below dt/persistence.  The random stream is
// counter-based: uniform(seed, step, cell, draw) from the SplitMix64
// finaliser (Steele, Lea & Flood 2014), so runs are reproducible, graph
// replays draw fresh numbers (the step counter lives in device memory and is
// advanced by a kernel), and the FP64 oracle draws the same numbers.
//
// Gradients are central differences of the substrate-major field at the
// cell's voxel, indices clamped (src/motility.cpp:45-86; the reference reads
// density[voxel*ns + s], the same thing for one substrate).
#include "motility_common.h"

namespace hc {

template <typename F>
__global__ __launch_bounds__(256) void k_motility(CellPtrs cp, MotPtrs mp, const F* __restrict__ rho, hc_grid g,
                                                  double dt, uint32_t n, int use_sensitivity) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) motility_one<F>(cp, mp, rho, g, dt, i, use_sensitivity);
}

// the step counter advances after every call (one thread; graph-safe)
__global__ void k_motility_advance(uint64_t* hdr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) hdr[0] += 1;
}

int probe_kernels_motility(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_motility<double>, (const void*)k_motility<float>,
                         (const void*)k_motility_advance};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: motility kernel");
    }
    return 0;
}

}  // namespace hc

using namespace hc;

extern "C" {

size_t hc_motility_field_offset(uint32_t field, uint32_t max_cells, uint32_t n_substrates) {
    const size_t m = max_cells, head = 256;
    const size_t d = ((head + 6 * m * 8 + 4 * m * 4) + 7) / 8 * 8;   // start of the sensitivities
    if (field <= HC_M_VEC_Z) return head + field * m * 8;
    if (field <= HC_M_CHEM_DIR) return head + 6 * m * 8 + (field - HC_M_MOTILE) * m * 4;
    (void)n_substrates;
    return d;
}

size_t hc_motility_bytes(uint32_t max_cells, uint32_t n_substrates) {
    return hc_motility_field_offset(HC_M_SENSITIVITY, max_cells, n_substrates) +
           (size_t)max_cells * n_substrates * 8;
}

int hc_motility(hc_ctx* ctx, hc_buf* cells, hc_buf* motility, hc_buf* density, const hc_grid* g, double dt,
                uint32_t n, uint32_t max_cells, int use_sensitivity) {
    if (!ctx || !cells || !motility || !g) return fail(ctx, HC_ERR_ARG, "hc_motility: null argument");
    if (n > max_cells) return fail(ctx, HC_ERR_ARG, "hc_motility: n > max_cells");
    if (hc_buf_size(cells) < hc_cells_bytes(max_cells)) return fail(ctx, HC_ERR_ARG, "hc_motility: cells buffer too small");
    if (hc_buf_size(motility) < hc_motility_bytes(max_cells, g->n_substrates))
        return fail(ctx, HC_ERR_ARG, "hc_motility: motility buffer too small");
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    if (density && hc_buf_size(density) < V * g->n_substrates * field_elem(g))
        return fail(ctx, HC_ERR_ARG, "hc_motility: density buffer too small");
    const CellPtrs cp = cell_ptrs(cells->dptr, max_cells);
    const MotPtrs mp = mot_ptrs(motility->dptr, max_cells, g->n_substrates);
    if (n) {
        const dim3 grid(blocks_for(n, 256));
        KernelTimer tm(ctx, HC_K_MOTILITY);
        if (g->field_fp32)
            hipLaunchKernelGGL(k_motility<float>, grid, dim3(256), 0, ctx->stream, cp, mp,
                               density ? (const float*)density->dptr : nullptr, *g, dt, n, use_sensitivity);
        else
            hipLaunchKernelGGL(k_motility<double>, grid, dim3(256), 0, ctx->stream, cp, mp,
                               density ? (const double*)density->dptr : nullptr, *g, dt, n, use_sensitivity);
    }
    hipLaunchKernelGGL(k_motility_advance, dim3(1), dim3(64), 0, ctx->stream, mp.hdr);
    return check_launch(ctx, "k_motility");
}

}  // extern "C"
