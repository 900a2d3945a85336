// sweep_part.hip -- partitioned LOD Thomas sweep (FP64, gfx950).
//
// Solves the same per-line system as the serial sweep (edge rows with the
// boundary diagonal 1+c+h and +c*T with Dirichlet; shaders/diffusion_2d.metal
// :89-155) without the forward pass's d' ever leaving the chip:
//
//   * the N = n-2 unknowns of a line are split into 4 contiguous segments,
//     one per lane; a lane keeps its segment in registers (<= MM values);
//   * each lane solves its segment with the neighbouring unknowns set to zero
//     (y = A_k^-1 d_k, elimination factors from a line-independent table);
//   * the true solution is u = y + uL*alpha_k + uR*beta_k, with
//     alpha_k = c A_k^-1 e_first, beta_k = c A_k^-1 e_last (tables), where uL
//     and uR are the last unknown of the segment to the left and the first of
//     the one to the right;
//   * uL and uR follow from the 8 segment end values y of the line (gathered
//     with cross-lane shuffles) through two rows of the inverse of the 8x8
//     interface system, also a table.
//
// Every value is read once and written once (16 bytes per voxel per sweep in
// FP64).  For y and z, 16 lines x 4 segments fill a wave and the lanes of one
// segment walk 16 adjacent x positions, so each load is a full 128-byte row
// segment.  x lines are contiguous: 4 lines x 16 segments fill a wave, which
// copies its 4 lines (4*n contiguous values) through LDS with 16-byte loads.
#include "hc_internal.h"

#include <algorithm>
#include <cmath>

namespace hc {

// Contiguous copy of nval doubles between global memory and an LDS tile by
// one wave.  The source may be only 8-byte aligned: a leading odd element is
// copied alone so the rest moves as 16-byte pairs.  The loads of a lane (at
// most MM/2 + 1 pairs, held in the caller's register window) are all issued
// before the LDS writes.
template <int MM>
__device__ inline void stage_in(const double* __restrict__ src, double* tile, uint32_t nval, uint32_t lane,
                                double (&reg)[MM + 2]) {
    const uint32_t head = ((uintptr_t)src & 15u) ? 1u : 0u;
    if (head && lane == 0 && nval) tile[0] = src[0];
    const uint32_t npair = (nval - (nval ? head : 0)) / 2;
    const double2* s2 = reinterpret_cast<const double2*>(src + head);
    constexpr int R = MM / 2 + 1;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = lane + 64u * r;
        if (e < npair) {
            const double2 q = s2[e];
            reg[2 * r] = q.x;
            reg[2 * r + 1] = q.y;
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = lane + 64u * r;
        if (e < npair) {
            tile[head + 2 * e] = reg[2 * r];
            tile[head + 2 * e + 1] = reg[2 * r + 1];
        }
    }
    const uint32_t done = head + 2 * npair;
    if (done < nval && lane == 0) tile[done] = src[done];
}

template <int MM>
__device__ inline void stage_out(double* __restrict__ dst, const double* tile, uint32_t nval, uint32_t lane) {
    const uint32_t head = ((uintptr_t)dst & 15u) ? 1u : 0u;
    if (head && lane == 0 && nval) dst[0] = tile[0];
    const uint32_t npair = (nval - (nval ? head : 0)) / 2;
    double2* d2 = reinterpret_cast<double2*>(dst + head);
    constexpr int R = MM / 2 + 1;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = lane + 64u * r;
        if (e < npair) d2[e] = make_double2(tile[head + 2 * e], tile[head + 2 * e + 1]);
    }
    const uint32_t done = head + 2 * npair;
    if (done < nval && lane == 0) dst[done] = tile[done];
}

// Each wave owns its LDS tile: lanes of one wave exchange data through it, so
// ordering the wave's own LDS accesses is enough (no workgroup barrier).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int AXIS, int P, int MM, bool XLDS>
__global__ __launch_bounds__(256) void k_sweep_part(
    double* __restrict__ rho, const uint8_t* __restrict__ tables, uint32_t nx, uint32_t ny, uint32_t nz,
    uint32_t nmax, int fuse_dirichlet) {
    constexpr int LPW = 64 / P;                          // lines per wave
    constexpr int WPB = 4;
    extern __shared__ double smem[];
    const uint32_t s = blockIdx.y, ns = gridDim.y;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t k = lane / LPW, l = lane % LPW;
    const uint32_t n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const uint32_t N = n - 2;
    const uint32_t MT = part_mt(nmax, P);
    const uint32_t blk = (uint32_t)part_block(nmax);
    const uint64_t V = (uint64_t)nx * ny * nz;

    // this (substrate, axis) table block -> LDS
    const double* gt = reinterpret_cast<const double*>(tables + part_offset(ns, nmax)) + (size_t)(s * 3 + AXIS) * blk;
    const uint32_t used = (uint32_t)part_block_p(nmax, P);
    for (uint32_t q = threadIdx.x; q < used; q += blockDim.x) smem[q] = gt[q];
    const double* tab = smem;

    const uint64_t lines_total = AXIS == 0 ? (uint64_t)ny * nz : (AXIS == 1 ? (uint64_t)nz * nx : (uint64_t)ny * nx);
    const uint64_t ngroups = (lines_total + (uint64_t)WPB * LPW - 1) / ((uint64_t)WPB * LPW);
    __syncthreads();                                    // tables visible
    // x blocks walk several line groups so their table copy is reused; y/z
    // blocks take one group each
    auto group = [&](uint64_t grp, const double* tabg) {
    const uint64_t Lw0 = (grp * WPB + w) * LPW;     // first line of this wave
    const uint64_t L = Lw0 + l;
    uint64_t stride, base, lines;
    bool face = false;
    if (AXIS == 0) {
        stride = 1; lines = (uint64_t)ny * nz; base = L * nx;
    } else if (AXIS == 1) {
        stride = nx; lines = (uint64_t)nz * nx;
        const uint64_t kk = L / nx, i = L % nx;
        base = kk * nx * ny + i;
        face = (i == 0 || i == nx - 1);
    } else {
        stride = (uint64_t)nx * ny; lines = (uint64_t)ny * nx;
        const uint64_t j = L / nx, i = L % nx;
        base = L;
        face = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
    }
    const bool valid = L < lines;
    const SubConst* sc = reinterpret_cast<const SubConst*>(tables);
    const double c = sc[s].c[AXIS];
    const double T = sc[s].T;
    const bool dir = sc[s].dir != 0;
    const bool fill = fuse_dirichlet && dir && face;
    const uint32_t sk = (k * N) / P;
    const uint32_t mk = ((k + 1) * N) / P - sk;
    double* p = rho + (uint64_t)s * V + base;

    double* tile = smem + blk + (size_t)w * LPW * n;

    double v[MM + 2];
    if (XLDS) {
        // The wave's LPW consecutive x lines are LPW*n contiguous values: copy
        // them with 16-byte loads, all issued before the first LDS write (the
        // register window v[] is free at this point), then read the segments.
        const uint64_t rem = lines > Lw0 ? lines - Lw0 : 0;
        const uint32_t nval = (uint32_t)((rem < (uint64_t)LPW ? rem : (uint64_t)LPW) * n);
        stage_in<MM>(rho + (uint64_t)s * V + Lw0 * nx, tile, nval, lane, v);
        wave_sync();
#pragma unroll
        for (int t = 0; t < MM; t++) v[t] = (t < (int)mk) ? tile[l * n + sk + 1 + t] : 0.0;
    } else {
#pragma unroll
        for (int t = 0; t < MM; t++) v[t] = (t < (int)mk && valid) ? p[(uint64_t)(sk + 1 + t) * stride] : 0.0;
    }

    const double* idn = tabg + (k * 4 + 0) * MT;
    const double* cpl = tabg + (k * 4 + 1) * MT;
    const double* al = tabg + (k * 4 + 2) * MT;
    const double* be = tabg + (k * 4 + 3) * MT;
    const double* gl = tabg + (size_t)P * 4 * MT + (k * 2 + 0) * 2 * P;
    const double* gr = tabg + (size_t)P * 4 * MT + (k * 2 + 1) * 2 * P;

    if (dir) {
        const double ct = c * T;
        if (k == 0) v[0] += ct;
#pragma unroll
        for (int t = 0; t < MM; t++)
            if (k == P - 1 && t == (int)mk - 1) v[t] += ct;
    }
    // segment solve with zero coupling
    double dp = 0.0;
#pragma unroll
    for (int t = 0; t < MM; t++)
        if (t < (int)mk) {
            dp = (v[t] + c * dp) * idn[t];
            v[t] = dp;
        }
    double y = 0.0;
#pragma unroll
    for (int t = MM - 1; t >= 0; t--)
        if (t < (int)mk) {
            y = v[t] - cpl[t] * y;
            v[t] = y;
        }
    const double yS = v[0];
    double yE = 0.0;
#pragma unroll
    for (int t = 0; t < MM; t++)
        if (t == (int)mk - 1) yE = v[t];
    // interface values of the line (its P lanes are l, l+LPW, l+2*LPW, ...)
    double uL = 0.0, uR = 0.0;
#pragma unroll
    for (int j = 0; j < P; j++) {
        const double a = __shfl(yS, j * LPW + l, 64);
        const double b = __shfl(yE, j * LPW + l, 64);
        uL += gl[2 * j] * a + gl[2 * j + 1] * b;
        uR += gr[2 * j] * a + gr[2 * j + 1] * b;
    }
#pragma unroll
    for (int t = 0; t < MM; t++)
        if (t < (int)mk) v[t] = v[t] + al[t] * uL + be[t] * uR;

    if (XLDS) {
#pragma unroll
        for (int t = 0; t < MM; t++)
            if (t < (int)mk) tile[l * n + sk + 1 + t] = v[t];
        if (dir) {
            if (k == 0) tile[l * n] = T;
            if (k == P - 1) tile[l * n + n - 1] = T;
        }
        wave_sync();
        const uint64_t rem = lines > Lw0 ? lines - Lw0 : 0;
        const uint32_t nval = (uint32_t)((rem < (uint64_t)LPW ? rem : (uint64_t)LPW) * n);
        stage_out<MM>(rho + (uint64_t)s * V + Lw0 * nx, tile, nval, lane);
    } else {
        if (valid) {
#pragma unroll
            for (int t = 0; t < MM; t++)
                if (t < (int)mk) p[(uint64_t)(sk + 1 + t) * stride] = fill ? T : v[t];
            if (dir) {
                if (k == 0) p[0] = T;
                if (k == P - 1) p[(uint64_t)(n - 1) * stride] = T;
            }
        }
    }
    if (XLDS) wave_sync();                          // tile reused by the next group
    };
    if (XLDS) {
        for (uint64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
            // keep the table reads inside the loop (hoisting them all into
            // registers would cost more registers than the window has)
            uint32_t toff = 0;
            asm volatile("" : "+v"(toff));
            group(grp, tab + toff);
        }
    } else {
        group(blockIdx.x, tab);
    }
}

int probe_kernels_part(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_sweep_part<0, 16, 16, true>, (const void*)k_sweep_part<0, 16, 32, true>,
                         (const void*)k_sweep_part<0, 16, 16, false>, (const void*)k_sweep_part<0, 16, 32, false>,
                         (const void*)k_sweep_part<1, 4, 16, false>, (const void*)k_sweep_part<1, 4, 32, false>,
                         (const void*)k_sweep_part<1, 4, 64, false>, (const void*)k_sweep_part<2, 4, 16, false>,
                         (const void*)k_sweep_part<2, 4, 32, false>, (const void*)k_sweep_part<2, 4, 64, false>};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: partitioned sweep");
    }
    return 0;
}

template <int AXIS, int P, bool XLDS>
static void launch_part_axis(hc_ctx* ctx, int mm, dim3 grid, dim3 block, size_t shm, double* rho,
                             const uint8_t* tables, const hc_grid* g, uint32_t nmax, int fuse) {
    if (mm == 16)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 16, XLDS>), grid, block, shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse);
    else if (mm == 32)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 32, XLDS>), grid, block, shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse);
    else if (P == 4)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 64, XLDS>), grid, block, shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse);
}

int launch_sweep_part(hc_ctx* ctx, int axis, double* rho, const hc_grid* g, const uint8_t* tables,
                      int fuse_dirichlet) {
    const int P = part_p(axis);
    const uint32_t n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    if (n < 2 + 2 * (uint32_t)P) return 1;                  // too short to partition
    const uint32_t m = (n - 2 + P - 1) / P;                  // longest segment
    const int mm = m <= 16 ? 16 : (m <= 32 ? 32 : (m <= 64 && P == 4 ? 64 : 0));
    if (!mm) return 1;
    const uint32_t nmax = grid_nmax(g);
    const int LPW = 64 / P;
    const bool xlds = axis == 0 && ctx->sweep_mode != 2;
    const size_t shm = (part_block(nmax) + (xlds ? (size_t)4 * LPW * n : 0)) * sizeof(double);
    if (shm > 64 * 1024) return 1;
    const uint64_t lines = axis == 0 ? (uint64_t)g->ny * g->nz
                                     : (axis == 1 ? (uint64_t)g->nz * g->nx : (uint64_t)g->ny * g->nx);
    // x: enough blocks to fill the chip a few times over, each looping over
    // line groups; y/z: one block per group
    const unsigned groups = blocks_for(lines, LPW * 4);
    const unsigned per_sub = (unsigned)std::max<uint64_t>(1, 4096 / std::max<uint32_t>(1, g->n_substrates));
    dim3 grid(xlds ? std::min(groups, per_sub) : groups, g->n_substrates);
    dim3 block(256);
    KernelTimer tm(ctx, HC_K_SWEEP_X + axis);
    if (axis == 0) {
        if (xlds) launch_part_axis<0, 16, true>(ctx, mm, grid, block, shm, rho, tables, g, nmax, fuse_dirichlet);
        else launch_part_axis<0, 16, false>(ctx, mm, grid, block, shm, rho, tables, g, nmax, fuse_dirichlet);
    } else if (axis == 1) {
        launch_part_axis<1, 4, false>(ctx, mm, grid, block, shm, rho, tables, g, nmax, fuse_dirichlet);
    } else {
        launch_part_axis<2, 4, false>(ctx, mm, grid, block, shm, rho, tables, g, nmax, fuse_dirichlet);
    }
    return check_launch(ctx, "k_sweep_part");
}

// Host construction of the partitioned tables (FP64).
static void solve_local(const std::vector<double>& den, const std::vector<double>& cpl, double c,
                        const std::vector<double>& d, std::vector<double>& out) {
    const size_t m = den.size();
    out.assign(m, 0.0);
    double dp = 0.0;
    for (size_t t = 0; t < m; t++) {
        dp = (d[t] + c * dp) / den[t];
        out[t] = dp;
    }
    double y = 0.0;
    for (size_t t = m; t-- > 0;) {
        y = out[t] - cpl[t] * y;
        out[t] = y;
    }
}

static bool invert(std::vector<double>& A, int n, std::vector<double>& inv) {
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int col = 0; col < n; col++) {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
        if (A[(size_t)piv * n + col] == 0.0) return false;
        if (piv != col)
            for (int j = 0; j < n; j++) {
                std::swap(A[(size_t)piv * n + j], A[(size_t)col * n + j]);
                std::swap(inv[(size_t)piv * n + j], inv[(size_t)col * n + j]);
            }
        const double d = A[(size_t)col * n + col];
        for (int j = 0; j < n; j++) { A[(size_t)col * n + j] /= d; inv[(size_t)col * n + j] /= d; }
        for (int r = 0; r < n; r++) {
            if (r == col) continue;
            const double f = A[(size_t)r * n + col];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) {
                A[(size_t)r * n + j] -= f * A[(size_t)col * n + j];
                inv[(size_t)r * n + j] -= f * inv[(size_t)col * n + j];
            }
        }
    }
    return true;
}

void build_part_tables_host(const hc_grid* g, const hc_substrate* subs, double* out) {
    const uint32_t ns = g->n_substrates, nmax = grid_nmax(g);
    const size_t blk = part_block(nmax);
    const double sp[3] = {g->dx, g->dy, g->dz};
    const uint32_t nn[3] = {g->nx, g->ny, g->nz};
    for (uint32_t s = 0; s < ns; s++) {
        const double D = subs[s].diffusion_coefficient, lam = subs[s].decay_rate;
        const double h = g->nz > 1 ? lam * g->dt / 3.0 : 0.5 * lam * g->dt;
        for (int ax = 0; ax < 3; ax++) {
            const int P = part_p(ax);
            const uint32_t MT = part_mt(nmax, P);
            double* B = out + (size_t)(s * 3 + ax) * blk;
            std::fill(B, B + blk, 0.0);
            const uint32_t n = nn[ax];
            if (n < 2 + 2 * P) continue;
            const uint32_t N = n - 2;
            const double c = D * g->dt / (sp[ax] * sp[ax]);
            const double bI = 1.0 + 2.0 * c + h, bB = 1.0 + c + h;
            std::vector<std::vector<double>> al(P), be(P);
            for (int k = 0; k < P; k++) {
                const uint32_t sk = (k * N) / P, mk = ((k + 1) * N) / P - sk;
                std::vector<double> den(mk), cp(mk), cpl(mk);
                for (uint32_t t = 0; t < mk; t++) {
                    const uint32_t r = sk + t;
                    const double b = (r == 0 || r == N - 1) ? bB : bI;
                    den[t] = t == 0 ? b : b - (-c) * cp[t - 1];
                    cp[t] = -c / den[t];
                }
                cpl = cp;
                cpl[mk - 1] = 0.0;
                std::vector<double> e0(mk, 0.0), eL(mk, 0.0);
                e0[0] = c;
                eL[mk - 1] = c;
                solve_local(den, cpl, c, e0, al[k]);
                solve_local(den, cpl, c, eL, be[k]);
                for (uint32_t t = 0; t < mk; t++) {
                    B[(k * 4 + 0) * MT + t] = 1.0 / den[t];
                    B[(k * 4 + 1) * MT + t] = cpl[t];
                    B[(k * 4 + 2) * MT + t] = al[k][t];
                    B[(k * 4 + 3) * MT + t] = be[k][t];
                }
            }
            // interface system over z = (L0, R0, L1, R1, ...): L_k = u(first row of
            // segment k), R_k = u(last row); u = y + R_{k-1} alpha_k + L_{k+1} beta_k
            const int n2 = 2 * P;
            std::vector<double> M((size_t)n2 * n2, 0.0), Mi;
            for (int k = 0; k < P; k++) {
                M[(size_t)(2 * k) * n2 + 2 * k] = 1.0;
                M[(size_t)(2 * k + 1) * n2 + 2 * k + 1] = 1.0;
                if (k > 0) {
                    M[(size_t)(2 * k) * n2 + 2 * k - 1] -= al[k].front();
                    M[(size_t)(2 * k + 1) * n2 + 2 * k - 1] -= al[k].back();
                }
                if (k < P - 1) {
                    M[(size_t)(2 * k) * n2 + 2 * k + 2] -= be[k].front();
                    M[(size_t)(2 * k + 1) * n2 + 2 * k + 2] -= be[k].back();
                }
            }
            if (!invert(M, n2, Mi)) continue;   // never singular: |alpha|,|beta| < 1
            double* G = B + (size_t)P * 4 * MT;
            for (int k = 0; k < P; k++)
                for (int j = 0; j < n2; j++) {
                    G[(k * 2 + 0) * n2 + j] = k > 0 ? Mi[(size_t)(2 * k - 1) * n2 + j] : 0.0;
                    G[(k * 2 + 1) * n2 + j] = k < P - 1 ? Mi[(size_t)(2 * k + 2) * n2 + j] : 0.0;
                }
        }
    }
}

}  // namespace hc
