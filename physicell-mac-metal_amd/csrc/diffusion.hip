// diffusion.hip -- LOD Thomas diffusion-decay sweeps for gfx950 (FP64).
//
// Each sweep solves, for every line of every substrate along one axis, the
// per-line system of shaders/diffusion_2d.metal:28-157 (x) / :165-292 (y):
// unknowns u1..u_{n-2}, edge rows with the boundary diagonal 1+c+h and +c*T
// on the right-hand side when Dirichlet is on, Dirichlet set before and
// re-applied after, n<=2 a no-op, n==3 a single interior node.  The z sweep
// is the same system with stride nx*ny.
//
// The elimination factors c'_i and denominators den_i depend only on
// (substrate, axis, i), so they are built once on the host (same operation
// sequence as the reference's in-kernel recurrence) and read from a table;
// each lane keeps only the d' recurrence:
//     d'_1 = rhs_1 / den_1,  d'_i = (rhs_i - a*d'_{i-1}) / den_i   (a = -c)
//     u_{n-2} = d'_{n-2},    u_i = d'_i - c'_i * u_{i+1}
// d' is written in place and read back by the backward pass.
//
// Kernels:
//   k_sweep_lane<AXIS>  one line per lane; lanes walk adjacent x positions, so
//                       every step of the recurrence is one coalesced row of
//                       the plane (y and z sweeps; also any axis with n <= 3).
//   k_sweep_x_lds       x lines are contiguous in memory: a wave stages a
//                       64-line x 16-element chunk through LDS (coalesced
//                       128-byte row segments in, transposed to one line per
//                       lane for the recurrence, and back out).
#include "hc_internal.h"

#include <cmath>

namespace hc {

static constexpr int LANE_BLOCK = 256;
static constexpr int XCH = 16;          // x chunk (elements per line per stage)
static constexpr int XWAVES = 4;        // waves per block in the x kernel

// ---------------------------------------------------------------------------
// one line per lane
// ---------------------------------------------------------------------------
template <int AXIS, int PF, typename F>
__global__ __launch_bounds__(LANE_BLOCK) void k_sweep_lane(
    F* __restrict__ rho, const uint8_t* __restrict__ tables, uint32_t nx, uint32_t ny,
    uint32_t nz, uint32_t nmax, int fuse_dirichlet) {
    const uint32_t s = blockIdx.y;
    const uint64_t L = (uint64_t)blockIdx.x * LANE_BLOCK + threadIdx.x;
    const uint64_t V = (uint64_t)nx * ny * nz;
    uint32_t n;
    uint64_t stride, base, lines;
    bool face = false;
    if (AXIS == 0) {
        n = nx; stride = 1; lines = (uint64_t)ny * nz; base = L * nx;
    } else if (AXIS == 1) {
        n = ny; stride = nx; lines = (uint64_t)nz * nx;
        const uint64_t k = L / nx, i = L % nx;
        base = k * nx * ny + i;
        face = (i == 0 || i == nx - 1);
    } else {
        n = nz; stride = (uint64_t)nx * ny; lines = (uint64_t)ny * nx;
        const uint64_t j = L / nx, i = L % nx;
        base = L;
        face = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
    }
    if (L >= lines) return;

    const SubConst* sc = reinterpret_cast<const SubConst*>(tables);
    const double* tab = reinterpret_cast<const double*>(tables + subconst_bytes(gridDim.y));
    const double c = sc[s].c[AXIS];
    const double T = sc[s].T;
    const bool dir = sc[s].dir != 0;
    F* p = rho + (uint64_t)s * V + base;

    if (fuse_dirichlet && dir && face) {
        // Every voxel of this line is on a face of another axis: the
        // Dirichlet re-apply that follows the substep sets all of it to T.
        for (uint32_t m = 0; m < n; m++) p[(uint64_t)m * stride] = (F)T;
        return;
    }
    if (dir) { p[0] = (F)T; p[(uint64_t)(n - 1) * stride] = (F)T; }
    if (n <= 2) return;
    const double* cpt = tab + table_index(s, AXIS, 0, nmax);
    const double* dnt = tab + table_index(s, AXIS, 1, nmax);
    if (n == 3) {
        double d = (double)p[stride];
        if (dir) { d += c * T; d += c * T; }
        p[stride] = (F)(d / dnt[1]);
        if (dir) { p[0] = (F)T; p[2 * stride] = (F)T; }
        return;
    }
    const double a = -c;
    double dprev;
    {
        double d = (double)p[stride];
        if (dir) d += c * T;
        dprev = d / dnt[1];
        p[stride] = (F)dprev;
    }
    // interior rows 2..n-3, PF-element chunks, next chunk loaded ahead
    const uint32_t iend = n - 2;
    uint32_t i = 2;
    if (i + PF <= iend) {
        double cur[PF];
#pragma unroll
        for (int k = 0; k < PF; k++) cur[k] = p[(uint64_t)(i + k) * stride];
        while (true) {
            const bool more = i + 2 * PF <= iend;
            double nxt[PF];
            if (more) {
#pragma unroll
                for (int k = 0; k < PF; k++) nxt[k] = p[(uint64_t)(i + PF + k) * stride];
            }
#pragma unroll
            for (int k = 0; k < PF; k++) {
                dprev = (cur[k] - a * dprev) / dnt[i + k];
                p[(uint64_t)(i + k) * stride] = (F)dprev;
            }
            i += PF;
            if (!more) break;
#pragma unroll
            for (int k = 0; k < PF; k++) cur[k] = nxt[k];
        }
    }
    for (; i < iend; i++) {
        const double d = (double)p[(uint64_t)i * stride];
        dprev = (d - a * dprev) / dnt[i];
        p[(uint64_t)i * stride] = (F)dprev;
    }
    {
        double d = (double)p[(uint64_t)(n - 2) * stride];
        if (dir) d += c * T;
        dprev = (d - a * dprev) / dnt[n - 2];
        p[(uint64_t)(n - 2) * stride] = (F)dprev;
    }
    // back substitution rows n-3..1
    double u = dprev;
    int j = (int)n - 3;
    if (j - (PF - 1) >= 1) {
        double cur[PF];
#pragma unroll
        for (int k = 0; k < PF; k++) cur[k] = p[(uint64_t)(j - k) * stride];
        while (true) {
            const bool more = j - PF - (PF - 1) >= 1;
            double nxt[PF];
            if (more) {
#pragma unroll
                for (int k = 0; k < PF; k++) nxt[k] = p[(uint64_t)(j - PF - k) * stride];
            }
#pragma unroll
            for (int k = 0; k < PF; k++) {
                u = cur[k] - cpt[j - k] * u;
                p[(uint64_t)(j - k) * stride] = (F)u;
            }
            j -= PF;
            if (!more) break;
#pragma unroll
            for (int k = 0; k < PF; k++) cur[k] = nxt[k];
        }
    }
    for (; j >= 1; j--) {
        u = (double)p[(uint64_t)j * stride] - cpt[j] * u;
        p[(uint64_t)j * stride] = (F)u;
    }
    if (dir) { p[0] = (F)T; p[(uint64_t)(n - 1) * stride] = (F)T; }
}

// ---------------------------------------------------------------------------
// x lines through LDS (n >= 4)
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(64 * XWAVES) void k_sweep_x_lds(
    F* __restrict__ rho, const uint8_t* __restrict__ tables, uint32_t nx,
    uint64_t lines, uint64_t V, uint32_t nmax) {
    __shared__ double tile[XWAVES][XCH][65];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t s = blockIdx.y;
    const uint64_t line0 = ((uint64_t)blockIdx.x * XWAVES + w) * 64;
    const bool valid = line0 + lane < lines;

    const SubConst* sc = reinterpret_cast<const SubConst*>(tables);
    const double* tab = reinterpret_cast<const double*>(tables + subconst_bytes(gridDim.y));
    const double c = sc[s].c[0];
    const double T = sc[s].T;
    const bool dir = sc[s].dir != 0;
    const double a = -c;
    const double* cpt = tab + table_index(s, 0, 0, nmax);
    const double* dnt = tab + table_index(s, 0, 1, nmax);
    F* base = rho + (uint64_t)s * V;
    const uint32_t n = nx;
    double (*t)[65] = tile[w];

    // forward elimination, chunk by chunk
    double dprev = 0.0;
    for (uint32_t c0 = 0; c0 < n; c0 += XCH) {
#pragma unroll
        for (int r = 0; r < XCH; r++) {
            const uint32_t idx = r * 64 + lane;
            const uint32_t l = idx / XCH, e = idx % XCH;
            const uint64_t Lg = line0 + l;
            const uint32_t col = c0 + e;
            t[e][l] = (Lg < lines && col < n) ? (double)base[Lg * nx + col] : 0.0;
        }
        __syncthreads();
        if (valid) {
            for (int e = 0; e < XCH; e++) {
                const uint32_t i = c0 + e;
                if (i >= n) break;
                if (i == 0 || i == n - 1) {
                    if (dir) t[e][lane] = T;
                    continue;
                }
                double d = t[e][lane];
                if (i == 1) {
                    if (dir) d += c * T;
                    dprev = d / dnt[1];
                } else {
                    if (i == n - 2 && dir) d += c * T;
                    dprev = (d - a * dprev) / dnt[i];
                }
                t[e][lane] = dprev;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < XCH; r++) {
            const uint32_t idx = r * 64 + lane;
            const uint32_t l = idx / XCH, e = idx % XCH;
            const uint64_t Lg = line0 + l;
            const uint32_t col = c0 + e;
            if (Lg < lines && col < n) base[Lg * nx + col] = (F)t[e][l];
        }
        __syncthreads();
    }
    // back substitution, last chunk first
    double u = dprev;
    const uint32_t last_c0 = ((n - 1) / XCH) * XCH;
    for (int64_t c0 = last_c0; c0 >= 0; c0 -= XCH) {
        // chunks holding only rows > n-3 need no work
        if ((uint32_t)c0 > n - 3) continue;
#pragma unroll
        for (int r = 0; r < XCH; r++) {
            const uint32_t idx = r * 64 + lane;
            const uint32_t l = idx / XCH, e = idx % XCH;
            const uint64_t Lg = line0 + l;
            const uint32_t col = (uint32_t)c0 + e;
            t[e][l] = (Lg < lines && col < n) ? (double)base[Lg * nx + col] : 0.0;
        }
        __syncthreads();
        if (valid) {
            for (int e = XCH - 1; e >= 0; e--) {
                const uint32_t i = (uint32_t)c0 + e;
                if (i > n - 3 || i < 1) continue;
                u = t[e][lane] - cpt[i] * u;
                t[e][lane] = u;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < XCH; r++) {
            const uint32_t idx = r * 64 + lane;
            const uint32_t l = idx / XCH, e = idx % XCH;
            const uint64_t Lg = line0 + l;
            const uint32_t col = (uint32_t)c0 + e;
            if (Lg < lines && col < n) base[Lg * nx + col] = (F)t[e][l];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// standalone Dirichlet re-apply (all faces)
// ---------------------------------------------------------------------------
template <typename F>
__global__ __launch_bounds__(256) void k_dirichlet(F* __restrict__ rho,
                                                   const uint8_t* __restrict__ tables,
                                                   uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t s = blockIdx.y;
    const SubConst* sc = reinterpret_cast<const SubConst*>(tables);
    if (!sc[s].dir) return;
    const uint64_t V = (uint64_t)nx * ny * nz;
    const uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const uint32_t i = (uint32_t)(v % nx);
    const uint32_t j = (uint32_t)((v / nx) % ny);
    const uint32_t k = (uint32_t)(v / ((uint64_t)nx * ny));
    const bool face = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || (nz > 1 && (k == 0 || k == nz - 1));
    if (face) rho[(uint64_t)s * V + v] = (F)sc[s].T;
}

int probe_kernels_part(hc_ctx* ctx);
int probe_kernels_diffusion(hc_ctx* ctx) {
    if (int r = probe_kernels_part(ctx)) return r;
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_sweep_lane<0, 8, double>, (const void*)k_sweep_lane<1, 8, double>,
                         (const void*)k_sweep_lane<2, 8, double>, (const void*)k_sweep_x_lds<double>,
                         (const void*)k_dirichlet<double>, (const void*)k_sweep_lane<0, 8, float>,
                         (const void*)k_sweep_lane<1, 8, float>, (const void*)k_sweep_lane<2, 8, float>,
                         (const void*)k_sweep_x_lds<float>, (const void*)k_dirichlet<float>};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: diffusion kernel");
    }
    return 0;
}

template <typename F>
static int launch_sweep_serial(hc_ctx* ctx, int axis, F* rho, const hc_grid* g, const uint8_t* tables,
                               int fuse_dirichlet) {
    const uint32_t nx = g->nx, ny = g->ny, nz = g->nz, ns = g->n_substrates;
    const uint32_t nmax = grid_nmax(g);
    const uint64_t V = (uint64_t)nx * ny * nz;
    if (axis == 0 && nx >= 4 && !fuse_dirichlet) {
        KernelTimer tm(ctx, HC_K_SWEEP_X);
        const uint64_t lines = (uint64_t)ny * nz;
        dim3 grid(blocks_for(lines, 64 * XWAVES), ns);
        hipLaunchKernelGGL(k_sweep_x_lds<F>, grid, dim3(64 * XWAVES), 0, ctx->stream, rho, tables, nx, lines, V, nmax);
        return check_launch(ctx, "k_sweep_x_lds");
    }
    const uint32_t n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const uint64_t lines = axis == 0 ? (uint64_t)ny * nz : (axis == 1 ? (uint64_t)nz * nx : (uint64_t)ny * nx);
    dim3 grid(blocks_for(lines, LANE_BLOCK), ns);
    KernelTimer tm(ctx, n <= 3 ? HC_K_SWEEP_SMALL : HC_K_SWEEP_X + axis);
    if (axis == 0)
        hipLaunchKernelGGL((k_sweep_lane<0, 8, F>), grid, dim3(LANE_BLOCK), 0, ctx->stream, rho, tables, nx, ny, nz, nmax,
                           fuse_dirichlet);
    else if (axis == 1)
        hipLaunchKernelGGL((k_sweep_lane<1, 8, F>), grid, dim3(LANE_BLOCK), 0, ctx->stream, rho, tables, nx, ny, nz, nmax,
                           fuse_dirichlet);
    else
        hipLaunchKernelGGL((k_sweep_lane<2, 8, F>), grid, dim3(LANE_BLOCK), 0, ctx->stream, rho, tables, nx, ny, nz, nmax,
                           fuse_dirichlet);
    return check_launch(ctx, "k_sweep_lane");
}

int launch_sweep(hc_ctx* ctx, int axis, void* rho, const hc_grid* g, const uint8_t* tables, int fuse_dirichlet) {
    if (g->n_substrates == 0 || g->nx == 0 || g->ny == 0 || g->nz == 0) return 0;
    if (ctx->sweep_mode != 1) {
        const int r = launch_sweep_part(ctx, axis, rho, g, tables, fuse_dirichlet);
        if (r != 1) return r;
    }
    if (g->field_fp32) return launch_sweep_serial<float>(ctx, axis, (float*)rho, g, tables, fuse_dirichlet);
    return launch_sweep_serial<double>(ctx, axis, (double*)rho, g, tables, fuse_dirichlet);
}

int launch_dirichlet(hc_ctx* ctx, void* rho, const hc_grid* g, const uint8_t* tables) {
    const uint64_t V = (uint64_t)g->nx * g->ny * g->nz;
    if (!V || !g->n_substrates) return 0;
    KernelTimer tm(ctx, HC_K_DIRICHLET);
    dim3 grid(blocks_for(V, 256), g->n_substrates);
    if (g->field_fp32)
        hipLaunchKernelGGL(k_dirichlet<float>, grid, dim3(256), 0, ctx->stream, (float*)rho, tables, g->nx, g->ny, g->nz);
    else
        hipLaunchKernelGGL(k_dirichlet<double>, grid, dim3(256), 0, ctx->stream, (double*)rho, tables, g->nx, g->ny, g->nz);
    return check_launch(ctx, "k_dirichlet");
}

// Host construction of the per-substrate constants and the elimination
// tables.  The operation sequence is the reference's in-kernel recurrence
// (shaders/diffusion_2d.metal:89-142) evaluated once in FP64.
void build_tables_host(const hc_grid* g, const hc_substrate* subs, std::vector<uint8_t>& out) {
    const uint32_t ns = g->n_substrates, nmax = grid_nmax(g);
    out.assign(table_total_bytes(ns, nmax), 0);
    build_part_tables_host(g, subs, reinterpret_cast<double*>(out.data() + part_offset(ns, nmax)),
                           reinterpret_cast<float*>(out.data() + part_f_offset(ns, nmax)));
    SubConst* sc = reinterpret_cast<SubConst*>(out.data());
    double* tab = reinterpret_cast<double*>(out.data() + subconst_bytes(ns));
    const double sp[3] = {g->dx, g->dy, g->dz};
    const uint32_t nn[3] = {g->nx, g->ny, g->nz};
    for (uint32_t s = 0; s < ns; s++) {
        const double D = subs[s].diffusion_coefficient, lam = subs[s].decay_rate;
        // decay share per sweep: 0.5*lambda*dt with two sweeps (2-D,
        // src/microenvironment.mm:91); lambda*dt/3 with three (nz > 1)
        const double h = g->nz > 1 ? lam * g->dt / 3.0 : 0.5 * lam * g->dt;
        sc[s].h = h;
        sc[s].T = subs[s].target_value;
        sc[s].dir = subs[s].dirichlet_enabled ? 1u : 0u;
        for (int ax = 0; ax < 3; ax++) {
            const double c = D * g->dt / (sp[ax] * sp[ax]);
            sc[s].c[ax] = c;
            const uint32_t n = nn[ax];
            double* cp = tab + table_index(s, ax, 0, nmax);
            double* den = tab + table_index(s, ax, 1, nmax);
            const double diag_interior = 1.0 + 2.0 * c + h;
            const double diag_boundary = 1.0 + c + h;
            if (n == 3) { den[1] = diag_interior; continue; }
            if (n < 4) continue;
            den[1] = diag_boundary;
            cp[1] = -c / diag_boundary;
            double c_prev = cp[1];
            for (uint32_t i = 2; i < n - 2; i++) {
                const double denom = diag_interior - (-c) * c_prev;
                den[i] = denom;
                c_prev = -c / denom;
                cp[i] = c_prev;
            }
            den[n - 2] = diag_boundary - (-c) * c_prev;
        }
    }
}

}  // namespace hc

using namespace hc;

extern "C" {

size_t hc_thomas_table_bytes(const hc_grid* g) {
    if (!g) return 0;
    return table_total_bytes(g->n_substrates, grid_nmax(g));
}

int hc_set_substep_order(hc_ctx* ctx, int order) {
    if (!ctx || order < 0 || order > 2) return fail(ctx, HC_ERR_ARG, "hc_set_substep_order: order must be 0, 1 or 2");
    ctx->substep_order = order;
    return 0;
}

int hc_set_sweep_mode(hc_ctx* ctx, int mode) {
    if (!ctx || mode < 0 || mode > 2) return fail(ctx, HC_ERR_ARG, "hc_set_sweep_mode: mode must be 0, 1 or 2");
    ctx->sweep_mode = mode;
    return 0;
}

int hc_build_thomas_tables(hc_ctx* ctx, const hc_grid* g, const hc_substrate* subs, hc_buf* tables) {
    if (!ctx || !g || (!subs && g->n_substrates) || !tables) return fail(ctx, HC_ERR_ARG, "hc_build_thomas_tables: null argument");
    if (hc_buf_size(tables) < hc_thomas_table_bytes(g))
        return fail(ctx, HC_ERR_ARG, "hc_build_thomas_tables: table buffer too small");
    std::vector<uint8_t> host;
    build_tables_host(g, subs, host);
    hipError_t e = hipMemcpyAsync(tables->dptr, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "hc_build_thomas_tables");
}

static int check_grid_args(hc_ctx* ctx, hc_buf* density, const hc_grid* g, hc_buf* tables, const char* who) {
    if (!ctx || !density || !g || !tables) return fail(ctx, HC_ERR_ARG, std::string(who) + ": null argument");
    const uint64_t need = (uint64_t)g->nx * g->ny * g->nz * g->n_substrates * field_elem(g);
    if (hc_buf_size(density) < need) return fail(ctx, HC_ERR_ARG, std::string(who) + ": density buffer too small");
    if (hc_buf_size(tables) < hc_thomas_table_bytes(g))
        return fail(ctx, HC_ERR_ARG, std::string(who) + ": tables too small (build them first)");
    return 0;
}

int hc_diffusion_sweep(hc_ctx* ctx, int axis, hc_buf* density, const hc_grid* g,
                       const hc_substrate* subs, hc_buf* tables, int fuse_dirichlet) {
    (void)subs;  // the per-substrate constants live in the table buffer
    int r = check_grid_args(ctx, density, g, tables, "hc_diffusion_sweep");
    if (r) return r;
    if (axis < 0 || axis > 2) return fail(ctx, HC_ERR_ARG, "hc_diffusion_sweep: axis must be 0, 1 or 2");
    return launch_sweep(ctx, axis, density->dptr, g, (const uint8_t*)tables->dptr, fuse_dirichlet);
}

int hc_dirichlet(hc_ctx* ctx, hc_buf* density, const hc_grid* g, const hc_substrate* subs) {
    if (!ctx || !density || !g || (!subs && g->n_substrates)) return fail(ctx, HC_ERR_ARG, "hc_dirichlet: null argument");
    const uint64_t need = (uint64_t)g->nx * g->ny * g->nz * g->n_substrates * field_elem(g);
    if (hc_buf_size(density) < need) return fail(ctx, HC_ERR_ARG, "hc_dirichlet: density buffer too small");
    // the kernel reads the Dirichlet flags/values from a SubConst block
    std::vector<uint8_t> host(subconst_bytes(g->n_substrates), 0);
    SubConst* sc = reinterpret_cast<SubConst*>(host.data());
    for (uint32_t s = 0; s < g->n_substrates; s++) {
        sc[s].T = subs[s].target_value;
        sc[s].dir = subs[s].dirichlet_enabled ? 1u : 0u;
    }
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, host.size());
    if (e != hipSuccess) return hip_fail(ctx, e, "hc_dirichlet: constants");
    e = hipMemcpyAsync(d, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream);
    int rr = e == hipSuccess ? launch_dirichlet(ctx, density->dptr, g, (const uint8_t*)d)
                             : hip_fail(ctx, e, "hc_dirichlet: upload");
    (void)hipStreamSynchronize(ctx->stream);        // cleanup: rr holds the first failure
    (void)hipFree(d);
    return rr;
}

int hc_diffusion_substep(hc_ctx* ctx, hc_buf* density, const hc_grid* g, const hc_substrate* subs,
                         hc_buf* tables) {
    (void)subs;
    int r = check_grid_args(ctx, density, g, tables, "hc_diffusion_substep");
    if (r) return r;
    void* rho = density->dptr;
    const uint8_t* tb = (const uint8_t*)tables->dptr;
    const bool three_d = g->nz > 1;
    // Substrate-major order (x, y, z of one substrate, then the next) gives
    // the same results -- the sweeps of different substrates are independent
    // -- and could keep one substrate's field in the last-level cache between
    // its sweeps.  Measured on the 200^3 x 4 grid it is slower (four launches
    // per axis of a quarter of the work each: 424 us vs 313 us per substep in
    // FP64), so the default is axis-major.
    if (ctx->substep_order == 2 && part_fits_all(g, ctx->sweep_mode)) {
        for (uint32_t s = 0; s < g->n_substrates; s++) {
            if ((r = launch_sweep_part(ctx, 0, rho, g, tb, 0, s, 1))) return r;
            if ((r = launch_sweep_part(ctx, 1, rho, g, tb, three_d ? 0 : 1, s, 1))) return r;
            if (three_d && (r = launch_sweep_part(ctx, 2, rho, g, tb, 1, s, 1))) return r;
        }
        return 0;
    }
    if ((r = launch_sweep(ctx, 0, rho, g, tb, 0))) return r;
    if ((r = launch_sweep(ctx, 1, rho, g, tb, three_d ? 0 : 1))) return r;
    if (three_d && (r = launch_sweep(ctx, 2, rho, g, tb, 1))) return r;
    return 0;
}

}  // extern "C"
