// sweep_part.hip -- partitioned LOD Thomas sweep (gfx950), FP64 or FP32 fields.
//
// Solves the same per-line system as the serial sweep (edge rows with the
// boundary diagonal 1+c+h and +c*T with Dirichlet; shaders/diffusion_2d.metal
// :89-155) without the forward pass's d' ever leaving the chip:
//
//   * the N = n-2 unknowns of a line are split into P contiguous segments,
//     one per lane; a lane keeps its segment in registers (<= MM values);
//   * each lane solves its segment with the neighbouring unknowns set to zero
//     (y = A_k^-1 d_k, elimination factors from a line-independent table);
//   * the true solution is u = y + uL*alpha_k + uR*beta_k, with
//     alpha_k = c A_k^-1 e_first, beta_k = c A_k^-1 e_last (tables), where uL
//     and uR are the last unknown of the segment to the left and the first of
//     the one to the right;
//   * uL and uR follow from the 2P segment end values y of the line (gathered
//     with cross-lane shuffles) through two rows of the inverse of the 2P x 2P
//     interface system, also a table.
//
// Every value is read once and written once (16 bytes per voxel per sweep in
// FP64, 8 in FP32).  For y and z, 16 lines x 4 segments fill a wave and the
// lanes of one segment walk 16 adjacent x positions, so each load is a full
// row segment.  x lines are contiguous: a wave copies its lines (LPW*n
// contiguous values) through LDS with 16-byte loads; short lines use 16
// lines x 4 segments per wave, long ones 4 lines x 16 segments, so that the
// staging tile fits and the per-line interface work stays small.
//
// FP32 fields (the mixed form): values and the segment recurrence are FP32;
// the elimination tables are computed in FP64 and rounded once; the Dirichlet
// terms c*T and the interface solve (uL, uR) are evaluated in FP64.
//
// The same group solver serves one simulation (k_sweep_part: one workgroup
// per line group) and a packed ensemble (k_sweep_part_batch: a persistent
// grid walking a work list of (simulation, substrate, line group)).
#include "hc_internal.h"

#include <algorithm>
#include <type_traits>
#include <cmath>

namespace hc {

// 16-byte vector of T as a native vector value (one dwordx4 / b128 access)
template <typename T>
struct Vec16T;
template <>
struct Vec16T<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec16T<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <typename T>
using Vec16 = typename Vec16T<T>::type;
// T in the global address space: field accesses through such pointers are
// global loads and stores even where the pointer was read from memory (the
// packed ensemble's descriptors), not passed as a kernel argument
template <typename T>
using global_t = __attribute__((address_space(1))) T;

// Contiguous copy of nval values between global memory and an LDS tile by
// one wave.  A misaligned head and the tail are copied as scalars so the bulk
// moves in 16-byte vectors; all of a lane's loads (at most RV vectors) are
// issued before its first LDS write.  The tile pointer carries the same
// misalignment as the global pointer (it is shifted by lead elements from a
// 16-byte boundary), so the vectors are 16-byte LDS accesses as well.
template <typename T, int RV>
__device__ inline void stage_in(const global_t<T>* __restrict__ src, T* tile, uint32_t nval, uint32_t lane) {
    constexpr uint32_t E = 16 / sizeof(T);
    uint32_t head = (uint32_t)(((16u - ((uintptr_t)src & 15u)) & 15u) / sizeof(T));
    if (head > nval) head = nval;
    if (lane < head) tile[lane] = src[lane];
    const uint32_t nvec = (nval - head) / E;
    const global_t<Vec16<T>>* s2 = reinterpret_cast<const global_t<Vec16<T>>*>(src + head);
    Vec16<T>* t2 = reinterpret_cast<Vec16<T>*>(tile + head);
    if (nvec) {                                           // wave-uniform
        // unconditional loads (lanes past the end re-read the last vector),
        // so that all of them are in flight before the first wait
        Vec16<T> buf[RV];
#pragma unroll
        for (int r = 0; r < RV; r++) {
            const uint32_t e = lane + 64u * r;
            buf[r] = s2[e < nvec ? e : nvec - 1];
        }
        __builtin_amdgcn_sched_barrier(0);               // keep the loads ahead of the LDS writes
#pragma unroll
        for (int r = 0; r < RV; r++) {
            const uint32_t e = lane + 64u * r;
            if (e < nvec) t2[e] = buf[r];
        }
    }
    const uint32_t done = head + E * nvec;
    if (lane < nval - done) tile[done + lane] = src[done + lane];
}

// Ordinary stores: the next sweep (y) reads the values soon after.  Non-
// temporal stores here left FP64 x unchanged and, with 8-segment FP32 x
// lines, slowed the following y sweep from 47 to 78 us (config 2).
template <typename T, int RV>
__device__ inline void stage_out(global_t<T>* __restrict__ dst, const T* tile, uint32_t nval, uint32_t lane) {
    constexpr uint32_t E = 16 / sizeof(T);
    uint32_t head = (uint32_t)(((16u - ((uintptr_t)dst & 15u)) & 15u) / sizeof(T));
    if (head > nval) head = nval;
    if (lane < head) dst[lane] = tile[lane];
    const uint32_t nvec = (nval - head) / E;
    global_t<Vec16<T>>* d2 = reinterpret_cast<global_t<Vec16<T>>*>(dst + head);
    const Vec16<T>* t2 = reinterpret_cast<const Vec16<T>*>(tile + head);
#pragma unroll
    for (int r = 0; r < RV; r++) {
        const uint32_t e = lane + 64u * r;
        if (e < nvec) d2[e] = t2[e];
    }
    const uint32_t done = head + E * nvec;
    if (lane < nval - done) dst[done + lane] = tile[done + lane];
}

// Elements of one wave's x staging tile: LPW lines of n values, one vector
// of lead-in (the tile is shifted to share the global misalignment), rounded
// to whole 16-byte vectors.
template <typename T>
__host__ __device__ inline uint32_t xtile_elems(uint32_t lpw, uint32_t n) {
    constexpr uint32_t E = 16 / sizeof(T);
    return (lpw * n + 2 * E - 1) / E * E;
}


// Each wave owns its LDS tile: lanes of one wave exchange data through it, so
// ordering the wave's own LDS accesses is enough (no workgroup barrier).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS image of one (substrate, axis) table block: interface rows (FP64) then
// the per-segment arrays 1/den, c', alpha, beta (field type).  Each segment's
// part is padded by one element (odd stride), so that the segments of a wave,
// which read the same row or array position at the same time, hit different
// LDS banks.
template <int P, typename T>
struct PartLds {
    static constexpr uint32_t GS = 4 * P + 1;                 // per segment: left row, right row, pad
    static __host__ __device__ uint32_t arr_stride(uint32_t MT) { return 4 * MT + 1; }
    static __host__ __device__ size_t g_count() { return (size_t)P * GS; }
    // plus 4 elements after the arrays: seg_solve / seg_join read table
    // positions in batches of 4, up to 3 past the last segment's end
    static __host__ __device__ size_t bytes(uint32_t nmax) {
        const size_t b = g_count() * sizeof(double) + ((size_t)P * arr_stride(part_mt(nmax, P)) + 4) * sizeof(T);
        return (b + 15) / 16 * 16;
    }
};

template <int P, typename T>
__device__ inline void load_part_tables(const global_t<uint8_t>* __restrict__ tables, uint32_t ns, uint32_t nmax,
                                        uint32_t s, int axis, double* tG, T* tarr) {
    const uint32_t MT = part_mt(nmax, P);
    const uint32_t na = (uint32_t)P * 4 * MT, AS = PartLds<P, T>::arr_stride(MT);
    const global_t<double>* dblk = reinterpret_cast<const global_t<double>*>(tables + part_offset(ns, nmax)) + (size_t)(s * 3 + axis) * part_block(nmax);
    for (uint32_t q = threadIdx.x; q < (uint32_t)(4 * P * P); q += blockDim.x)
        tG[(q / (4 * P)) * PartLds<P, T>::GS + q % (4 * P)] = dblk[na + q];
    if (sizeof(T) == sizeof(double)) {
        for (uint32_t q = threadIdx.x; q < na; q += blockDim.x) tarr[(q / (4 * MT)) * AS + q % (4 * MT)] = (T)dblk[q];
    } else {
        const global_t<float>* fblk = reinterpret_cast<const global_t<float>*>(tables + part_f_offset(ns, nmax)) + (size_t)(s * 3 + axis) * part_fblock(nmax);
        for (uint32_t q = threadIdx.x; q < na; q += blockDim.x) tarr[(q / (4 * MT)) * AS + q % (4 * MT)] = (T)fblk[q];
    }
}

// The part of one (substrate, axis) table block that a sweep of lines of n
// nodes reads (segment positions t < part_mt(n, P) <= MM), held in registers
// between its global loads and its LDS stores.  A workgroup that loads its
// tables issues them first, then its field loads, and stores the tables to
// LDS only when the field loads are in flight: one memory latency before the
// solve instead of two.
template <int P, int MM, typename T, int BS = 256>
struct TableRegs {
    static constexpr int NA = (P * 4 * MM + BS - 1) / BS;
    static constexpr int NG = (4 * P * P + BS - 1) / BS;
    T a[NA];
    double g[NG];

    __device__ inline void load(const global_t<uint8_t>* __restrict__ tables, uint32_t ns, uint32_t nmax, uint32_t s,
                                int axis, uint32_t n) {
        const uint32_t MT = part_mt(nmax, P), mt = part_mt(n, P), na = (uint32_t)P * 4 * MT;
        const global_t<double>* dblk = reinterpret_cast<const global_t<double>*>(tables + part_offset(ns, nmax)) +
                                       (size_t)(s * 3 + axis) * part_block(nmax);
        const global_t<T>* ablk;
        if constexpr (sizeof(T) == sizeof(double))
            ablk = (const global_t<T>*)dblk;
        else
            ablk = reinterpret_cast<const global_t<T>*>(tables + part_f_offset(ns, nmax)) +
                   (size_t)(s * 3 + axis) * part_fblock(nmax);
#pragma unroll
        for (int r = 0; r < NG; r++) {
            const uint32_t q = threadIdx.x + (uint32_t)BS * r;
            g[r] = q < (uint32_t)(4 * P * P) ? dblk[na + q] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < NA; r++) {
            const uint32_t q = threadIdx.x + (uint32_t)BS * r;
            const uint32_t k = q / (4 * MM), j = (q / MM) % 4, t = q % MM;
            a[r] = (q < (uint32_t)(P * 4 * MM) && t < mt) ? ablk[(k * 4 + j) * MT + t] : (T)0;
        }
    }
    __device__ inline void store(uint32_t nmax, uint32_t n, double* tG, T* tarr) const {
        const uint32_t MT = part_mt(nmax, P), mt = part_mt(n, P), AS = PartLds<P, T>::arr_stride(MT);
#pragma unroll
        for (int r = 0; r < NG; r++) {
            const uint32_t q = threadIdx.x + (uint32_t)BS * r;
            if (q < (uint32_t)(4 * P * P)) tG[(q / (4 * P)) * PartLds<P, T>::GS + q % (4 * P)] = g[r];
        }
#pragma unroll
        for (int r = 0; r < NA; r++) {
            const uint32_t q = threadIdx.x + (uint32_t)BS * r;
            const uint32_t k = q / (4 * MM), j = (q / MM) % 4, t = q % MM;
            if (q < (uint32_t)(P * 4 * MM) && t < mt) tarr[k * AS + j * MT + t] = a[r];
        }
    }
};

struct NoTables {
    __device__ inline void operator()() const {}
};

// The recurrences of one lane's segment (mk <= MM unknowns in v[0..mk)),
// written without a branch per element: the table values are read from LDS
// SB at a time (so a batch costs one LDS latency instead of one per element;
// the reads past the segment's end stay inside the table block or the LDS
// that follows it, and their values are not used), positions t >= mk are
// kept out of the chains by selects, and batches at or beyond Mu (the largest
// mk of the wave, uniform) are skipped.  Same operations in the same order as
// the element-by-element form.
//   forward: dp = (v[t] + c*dp) / den[t] (times the reciprocal), v[t] = dp;
//   backward: y = v[t] - c'[t]*y, v[t] = y (c' = 0 at the segment's end);
//   yE = the forward pass's last value, yS = v[0] after the backward pass.
template <int MM, typename T, int SB = 4>
__device__ __attribute__((always_inline)) inline void seg_solve(T (&v)[MM], uint32_t mk, uint32_t Mu, const T* idn,
                                                                const T* cpl, T cc, double& yS, double& yE) {
    T dp = 0;
#pragma unroll
    for (int c = 0; c < MM / SB; c++) {
        if ((uint32_t)(SB * c) < Mu) {
            T d[SB];
#pragma unroll
            for (int q = 0; q < SB; q++) {
                const uint32_t t = SB * c + q;
                d[q] = idn[t];
            }
#pragma unroll
            for (int q = 0; q < SB; q++) {
                const int t = SB * c + q;
                const T nd = (v[t] + cc * dp) * d[q];
                dp = (uint32_t)t < mk ? nd : dp;
                v[t] = nd;                                 // positions t >= mk are never used
            }
        }
    }
    yE = (double)dp;
    T y = 0;
#pragma unroll
    for (int c = MM / SB - 1; c >= 0; c--) {
        if ((uint32_t)(SB * c) < Mu) {
            T d[SB];
#pragma unroll
            for (int q = 0; q < SB; q++) {
                const uint32_t t = SB * c + q;
                d[q] = cpl[t];
            }
#pragma unroll
            for (int q = SB - 1; q >= 0; q--) {
                const int t = SB * c + q;
                const T ny = v[t] - d[q] * y;
                y = (uint32_t)t < mk ? ny : y;
                v[t] = ny;
            }
        }
    }
    yS = (double)v[0];
}

// u = y + uL*alpha + uR*beta over the segment (same batching)
template <int MM, typename T, int SB = 4>
__device__ __attribute__((always_inline)) inline void seg_join(T (&v)[MM], uint32_t mk, uint32_t Mu, const T* al,
                                                               const T* be, T tuL, T tuR) {
#pragma unroll
    for (int c = 0; c < MM / SB; c++) {
        if ((uint32_t)(SB * c) < Mu) {
            T a[SB], b[SB];
#pragma unroll
            for (int q = 0; q < SB; q++) {
                const uint32_t t = SB * c + q;
                a[q] = al[t];
                b[q] = be[t];
            }
#pragma unroll
            for (int q = 0; q < SB; q++) {
                const int t = SB * c + q;
                v[t] = v[t] + a[q] * tuL + b[q] * tuR;
            }
        }
    }
}

// One group of WPB*LPW lines (a workgroup of WPB waves) of one substrate of
// one simulation.  rho points at that substrate's field.
// xmap: lane order of the x staging tile (x_lane_order); y and z use order 0.
// mid() runs once the group's field loads are issued and before the first
// table read (a workgroup that stages its tables stores them and syncs there;
// every thread of the workgroup calls it).
template <int AXIS, int P, int MM, bool XLDS, typename T, typename Mid = NoTables, int WPB = 4>
__device__ __attribute__((always_inline)) inline void sweep_group(T* __restrict__ rho, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t MT,
                                   const T* tarr, const double* tG, double c, double T0, bool dir, int fuse_dirichlet,
                                   uint64_t grp, T* tile_base, uint32_t xmap, const Mid& mid = Mid()) {
    constexpr int LPW = 64 / P;
    constexpr int RV = MM / (16 / (int)sizeof(T)) + 2;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool tr = XLDS && xmap;
    const uint32_t k = tr ? lane % P : lane / LPW, l = tr ? lane / P : lane % LPW;
    const uint32_t n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const uint32_t N = n - 2;
    const uint64_t Lw0 = (grp * WPB + w) * LPW;          // first line of this wave
    const uint64_t L = Lw0 + l;
    uint64_t stride, base, lines;
    bool face = false;
    if (AXIS == 0) {
        stride = 1; lines = (uint64_t)ny * nz; base = L * nx;
    } else if (AXIS == 1) {
        // a substrate's y or z lines number nz*nx or ny*nx < 2^32: 32-bit
        // division (a 64-bit one costs about 60 VALU instructions)
        stride = nx; lines = (uint64_t)nz * nx;
        const uint32_t kk = (uint32_t)L / nx, i = (uint32_t)L - kk * nx;
        base = (uint64_t)kk * nx * ny + i;
        face = (i == 0 || i == nx - 1);
    } else {
        stride = (uint64_t)nx * ny; lines = (uint64_t)ny * nx;
        const uint32_t j = (uint32_t)L / nx, i = (uint32_t)L - j * nx;
        base = L;
        face = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
    }
    const bool valid = L < lines;
    const bool fill = fuse_dirichlet && dir && face;
    const uint32_t sk = (k * N) / P;
    const uint32_t mk = ((k + 1) * N) / P - sk;
    global_t<T>* grho = (global_t<T>*)rho;
    global_t<T>* p = grho + base;
    T* tile = tile_base + (size_t)w * xtile_elems<T>(LPW, n) + (((uintptr_t)(rho + Lw0 * nx) & 15u) / sizeof(T));
    const uint64_t rem = lines > Lw0 ? lines - Lw0 : 0;
    const uint32_t nval = (uint32_t)((rem < (uint64_t)LPW ? rem : (uint64_t)LPW) * n);

    // Dirichlet term c*T on the line's first and last unknown, added in FP64
    const double ct = c * T0;
    T v[MM];
    if (XLDS) {
        // the wave's LPW consecutive x lines are LPW*n contiguous values
        stage_in<T, RV>(grho + Lw0 * nx, tile, nval, lane);
        mid();
        wave_sync();
        if (dir) {
            // in the tile, by the lanes whose segments hold those unknowns
            // (at most one of the two per lane: P >= 4)
            const uint32_t at = k == 0 ? l * n + 1 : l * n + n - 2;
            if (k == 0 || k == P - 1) tile[at] = (T)((double)tile[at] + ct);
            wave_sync();
        }
#pragma unroll
        for (int t = 0; t < MM; t++) v[t] = (t < (int)mk) ? tile[l * n + sk + 1 + t] : (T)0;
    } else {
#pragma unroll
        for (int t = 0; t < MM; t++) v[t] = (t < (int)mk && valid) ? p[(uint64_t)(sk + 1 + t) * stride] : (T)0;
        mid();
    }

    const T* idn = tarr + k * PartLds<P, T>::arr_stride(MT);
    const T* cpl = idn + MT;
    const T* al = idn + 2 * MT;
    const T* be = idn + 3 * MT;
    const double* gl = tG + k * PartLds<P, T>::GS;
    const double* gr = gl + 2 * P;

    if (!XLDS && dir) {
        if (k == 0) v[0] = (T)((double)v[0] + ct);
        // the last segment's length is the same in every lane, so its last
        // unknown is one register position for the whole wave
        const uint32_t jl = __builtin_amdgcn_readfirstlane(N - ((P - 1) * N) / P - 1);
        const bool last = k == P - 1;
#pragma unroll
        for (int t = 0; t < MM; t++)
            if (t == (int)jl) {                            // uniform: one position is taken
                // a scalar branch, not evaluated at every position and selected
                __asm__ volatile("");
                if (last) v[t] = (T)((double)v[t] + ct);
            }
    }
    // segment solve with zero coupling; the segment's last value yE: the
    // backward pass starts there with no right neighbour, so it keeps the
    // forward pass's last value
    const uint32_t Mu = __builtin_amdgcn_readfirstlane((N + P - 1) / P);
    double yS, yE;
    seg_solve<MM, T>(v, mk, Mu, idn, cpl, (T)c, yS, yE);
    // interface values of the line (its P lanes are l, l+LPW, l+2*LPW, ...), FP64
    double uL = 0.0, uR = 0.0;
    if (XLDS && P == 16) {
        // x: every lane publishes its segment's two end values once, in the
        // first KiB of the wave's staging tile; each lane then forms uL and uR
        // as dot products with its two rows of the inverse interface matrix,
        // in two independent FMA chains each, instead of 2P shuffles of FP64
        // values and one chain of 2P dependent adds.  (With 4 segments the
        // shuffles are few and cost fewer registers.)  The tile's interior
        // values are all in registers by now; its lines' end nodes are not
        // rewritten without Dirichlet, so the lanes owning them keep them in
        // registers across the exchange.  A 16-segment tile (n > 96) holds
        // more than 1 KiB.
        const T keep0 = tile[l * n], keep1 = tile[l * n + n - 1];
        double2* ex = reinterpret_cast<double2*>(tile_base + (size_t)w * xtile_elems<T>(LPW, n));
        wave_sync();
        ex[lane] = make_double2(yS, yE);
        wave_sync();
        double aL0 = 0.0, aL1 = 0.0, aR0 = 0.0, aR1 = 0.0;
        // the line's P entries sit STEP apart from its first; a uniform
        // branch per lane order, so every read has an immediate offset
        auto dots = [&](const double2* e, auto step) {
#pragma unroll
            for (int j = 0; j < P; j++) {
                const double2 ab = e[j * decltype(step)::value];
                aL0 = __builtin_fma(gl[2 * j], ab.x, aL0);
                aR0 = __builtin_fma(gr[2 * j], ab.x, aR0);
                aL1 = __builtin_fma(gl[2 * j + 1], ab.y, aL1);
                aR1 = __builtin_fma(gr[2 * j + 1], ab.y, aR1);
                // one step's reads at a time: hoisting all P steps' reads
                // would hold 12*P registers
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        if (tr)
            dots(ex + l * P, std::integral_constant<int, 1>());
        else
            dots(ex + l, std::integral_constant<int, LPW>());
        uL = aL0 + aL1;
        uR = aR0 + aR1;
        wave_sync();
        if (k == 0) tile[l * n] = keep0;
        if (k == P - 1) tile[l * n + n - 1] = keep1;
    } else {
#pragma unroll
        for (int j = 0; j < P; j++) {
            const int src = tr ? (int)l * P + j : j * LPW + (int)l;
            const double a = __shfl(yS, src, 64);
            const double b = __shfl(yE, src, 64);
            uL += gl[2 * j] * a + gl[2 * j + 1] * b;
            uR += gr[2 * j] * a + gr[2 * j + 1] * b;
        }
    }
    seg_join<MM, T>(v, mk, Mu, al, be, (T)uL, (T)uR);

    const T tT = (T)T0;
    if (XLDS) {
#pragma unroll
        for (int t = 0; t < MM; t++)
            if (t < (int)mk) tile[l * n + sk + 1 + t] = v[t];
        if (dir) {
            if (k == 0) tile[l * n] = tT;
            if (k == P - 1) tile[l * n + n - 1] = tT;
        }
        wave_sync();
        stage_out<T, RV>(grho + Lw0 * nx, tile, nval, lane);
        wave_sync();                                      // tile reused by the next group
    } else if (valid) {
#pragma unroll
        for (int t = 0; t < MM; t++)
            if (t < (int)mk) p[(uint64_t)(sk + 1 + t) * stride] = fill ? tT : v[t];
        if (dir) {
            if (k == 0) p[0] = tT;
            if (k == P - 1) p[(uint64_t)(n - 1) * stride] = tT;
        }
    }
}

// ---------------------------------------------------------------------------
// one simulation: one workgroup per (line group, substrate)
// ---------------------------------------------------------------------------
// FP32 y / z sweeps of one simulation with windows of up to 32 values ask
// for 8 waves per SIMD: they fit in 61-63 VGPRs without spills, against
// 66-68 (7 waves) when the compiler chooses (mixed config-2 step -0.7%).
// The ensemble's item kernels measured slower that way (y spills 12 bytes).
template <int AXIS, int MM, typename T>
constexpr int yz32_waves() {
    return (AXIS != 0 && MM <= 32 && sizeof(T) == sizeof(float)) ? 8 : 1;
}

template <int AXIS, int P, int MM, bool XLDS, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(yz32_waves<AXIS, MM, T>()))) void k_sweep_part(T* __restrict__ rho, const uint8_t* __restrict__ tables,
                                                    uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nmax,
                                                    int fuse_dirichlet, uint32_t xmap, uint32_t s0, uint32_t ns) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* tG = reinterpret_cast<double*>(smem_raw);
    T* tarr = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::g_count() * sizeof(double));
    T* tile = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::bytes(nmax));
    const SubConst* sc = reinterpret_cast<const SubConst*>(tables);
    const uint32_t MT = part_mt(nmax, P);
    // one line group per workgroup (substrates s0 .. s0+gridDim.y-1 of ns),
    // in XCD chunks of consecutive groups (xcd_chunk); table loads first,
    // then the field loads, then the tables go to LDS.  (x measured faster
    // this way than with fewer workgroups each looping over groups: FP64
    // config-2 x sweep 108.5 -> 101.7 us.)
    const uint32_t lin = xcd_chunk<16>(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const uint32_t grp = lin % gridDim.x, s = s0 + lin / gridDim.x;
    const double c = sc[s].c[AXIS], T0 = sc[s].T;
    const bool dir = sc[s].dir != 0;
    T* rs = rho + (uint64_t)s * nx * ny * nz;
    const uint32_t n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    TableRegs<P, MM, T> treg;
    treg.load((const global_t<uint8_t>*)tables, ns, nmax, s, AXIS, n);
    __builtin_amdgcn_sched_barrier(0);
    auto mid = [&]() {
        treg.store(nmax, n, tG, tarr);
        __syncthreads();
    };
    sweep_group<AXIS, P, MM, XLDS, T, decltype(mid)>(rs, nx, ny, nz, MT, tarr, tG, c, T0, dir, fuse_dirichlet, grp,
                                                     tile, XLDS ? xmap : 0u, mid);
}

// ---------------------------------------------------------------------------
// packed ensemble.  MM = 0: the list holds members of both register windows
// (16 and 32 values per lane, BatchSim::wide) and each item takes its
// member's window (a uniform branch); one launch per axis then covers them
// all, with no tail between a list of long lines and a list of short ones.
// ---------------------------------------------------------------------------
template <int AXIS, int P, int MM, bool XLDS, typename T>
__device__ __attribute__((always_inline)) inline void batch_group(const BatchSim& sm, uint32_t s, uint64_t grp,
                                                                  const T* tarr, const double* tG, T* tile) {
    const global_t<SubConst>* sc = (const global_t<SubConst>*)sm.tab;
    uint32_t toff = 0;
    asm volatile("" : "+v"(toff));
    T* rs = reinterpret_cast<T*>(sm.rho) + (uint64_t)s * sm.nx * sm.ny * sm.nz;
    sweep_group<AXIS, P, MM, XLDS, T>(rs, sm.nx, sm.ny, sm.nz, part_mt(sm.nmax, P), tarr + toff, tG + toff,
                                      sc[s].c[AXIS], sc[s].T, sc[s].dir != 0, sm.last_axis == (uint32_t)AXIS ? 1 : 0,
                                      grp, tile, sm.xmap);
}

// a persistent grid walks the work list; each workgroup takes a contiguous
// run of items, so consecutive items mostly share their tables
template <int AXIS, int P, int MM, bool XLDS, typename T>
__global__ __launch_bounds__(256) void k_sweep_part_batch(const BatchSim* __restrict__ sims,
                                                          const BatchWork* __restrict__ work, uint32_t nwork,
                                                          uint32_t nmax_all) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* tG = reinterpret_cast<double*>(smem_raw);
    T* tarr = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::g_count() * sizeof(double));
    T* tile = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::bytes(nmax_all));
    const uint32_t per = (nwork + gridDim.x - 1) / gridDim.x;
    const uint32_t w0 = blockIdx.x * per;
    const uint32_t w1 = w0 + per < nwork ? w0 + per : nwork;
    uint32_t cur_sim = 0xFFFFFFFFu, cur_s = 0xFFFFFFFFu;
    for (uint32_t wi = w0; wi < w1; wi++) {
        const BatchWork it = work[wi];
        const BatchSim sm = sims[it.sim];
        if (it.sim != cur_sim || it.s != cur_s) {
            __syncthreads();                              // previous tables no longer read
            load_part_tables<P, T>((const global_t<uint8_t>*)sm.tab, sm.ns, sm.nmax, it.s, AXIS, tG, tarr);
            __syncthreads();
            cur_sim = it.sim;
            cur_s = it.s;
        }
        if constexpr (MM != 0)
            batch_group<AXIS, P, MM, XLDS, T>(sm, it.s, it.grp, tarr, tG, tile);
        else if ((sm.wide >> AXIS) & 1u)
            batch_group<AXIS, P, 32, XLDS, T>(sm, it.s, it.grp, tarr, tG, tile);
        else
            batch_group<AXIS, P, 16, XLDS, T>(sm, it.s, it.grp, tarr, tG, tile);
    }
}

// one workgroup per work item: table loads first, then the field loads, then
// the tables go to LDS
template <int AXIS, int P, int MM, bool XLDS, typename T>
__device__ __attribute__((always_inline)) inline void item_group(const BatchSim& sm, const BatchWork& it, double* tG,
                                                                 T* tarr, T* tile) {
    const uint32_t n = AXIS == 0 ? sm.nx : (AXIS == 1 ? sm.ny : sm.nz);
    TableRegs<P, MM, T> treg;
    treg.load((const global_t<uint8_t>*)sm.tab, sm.ns, sm.nmax, it.s, AXIS, n);
    __builtin_amdgcn_sched_barrier(0);
    const global_t<SubConst>* sc = (const global_t<SubConst>*)sm.tab;
    T* rs = reinterpret_cast<T*>(sm.rho) + (uint64_t)it.s * sm.nx * sm.ny * sm.nz;
    sweep_group<AXIS, P, MM, XLDS, T>(rs, sm.nx, sm.ny, sm.nz, part_mt(sm.nmax, P), tarr, tG, sc[it.s].c[AXIS],
                                      sc[it.s].T, sc[it.s].dir != 0, sm.last_axis == (uint32_t)AXIS ? 1 : 0, it.grp,
                                      tile, sm.xmap, [&]() {
                                          treg.store(sm.nmax, n, tG, tarr);
                                          __syncthreads();
                                      });
}

template <int AXIS, int P, int MM, bool XLDS, typename T>
__global__ __launch_bounds__(256) void k_sweep_part_item(const BatchSim* __restrict__ sims,
                                                         const BatchWork* __restrict__ work, uint32_t nmax_all) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* tG = reinterpret_cast<double*>(smem_raw);
    T* tarr = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::g_count() * sizeof(double));
    T* tile = reinterpret_cast<T*>(smem_raw + PartLds<P, T>::bytes(nmax_all));
    const BatchWork it = work[xcd_chunk<16>(blockIdx.x, gridDim.x)];
    const BatchSim sm = sims[it.sim];
    if constexpr (MM != 0)
        item_group<AXIS, P, MM, XLDS, T>(sm, it, tG, tarr, tile);
    else if ((sm.wide >> AXIS) & 1u)
        item_group<AXIS, P, 32, XLDS, T>(sm, it, tG, tarr, tile);
    else
        item_group<AXIS, P, 16, XLDS, T>(sm, it, tG, tarr, tile);
}

int probe_kernels_part(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {
        (const void*)k_sweep_part<0, 16, 16, true, double>, (const void*)k_sweep_part<0, 16, 32, true, double>,
        (const void*)k_sweep_part<0, 16, 16, false, double>, (const void*)k_sweep_part<0, 16, 32, false, double>,
        (const void*)k_sweep_part<1, 4, 16, false, double>, (const void*)k_sweep_part<1, 4, 32, false, double>,
        (const void*)k_sweep_part<1, 4, 64, false, double>, (const void*)k_sweep_part<2, 4, 16, false, double>,
        (const void*)k_sweep_part<2, 4, 32, false, double>, (const void*)k_sweep_part<2, 4, 64, false, double>,
        (const void*)k_sweep_part<1, 8, 16, false, double>, (const void*)k_sweep_part<1, 8, 32, false, double>,
        (const void*)k_sweep_part<1, 8, 64, false, double>, (const void*)k_sweep_part<2, 8, 16, false, double>,
        (const void*)k_sweep_part<2, 8, 32, false, double>, (const void*)k_sweep_part<2, 8, 64, false, double>,
        (const void*)k_sweep_part<0, 16, 16, true, float>, (const void*)k_sweep_part<0, 16, 32, true, float>,
        (const void*)k_sweep_part<1, 4, 16, false, float>, (const void*)k_sweep_part<1, 4, 32, false, float>,
        (const void*)k_sweep_part<1, 4, 64, false, float>, (const void*)k_sweep_part<2, 4, 16, false, float>,
        (const void*)k_sweep_part<2, 4, 32, false, float>, (const void*)k_sweep_part<2, 4, 64, false, float>,
        (const void*)k_sweep_part<0, 4, 16, true, double>, (const void*)k_sweep_part<0, 4, 32, true, double>,
        (const void*)k_sweep_part<0, 4, 16, false, double>, (const void*)k_sweep_part<0, 4, 32, false, double>,
        (const void*)k_sweep_part<0, 4, 16, true, float>, (const void*)k_sweep_part<0, 4, 32, true, float>,
        (const void*)k_sweep_part<0, 4, 16, false, float>, (const void*)k_sweep_part<0, 4, 32, false, float>,
        (const void*)k_sweep_part<0, 8, 16, true, float>, (const void*)k_sweep_part<0, 8, 32, true, float>,
        (const void*)k_sweep_part<0, 8, 16, false, float>, (const void*)k_sweep_part<0, 8, 32, false, float>,
        (const void*)k_sweep_part_batch<0, 8, 32, true, float>, (const void*)k_sweep_part_item<0, 8, 32, true, float>,
        (const void*)k_sweep_part_batch<0, 16, 16, true, double>, (const void*)k_sweep_part_item<0, 16, 16, true, double>,
        (const void*)k_sweep_part_batch<0, 4, 0, true, double>, (const void*)k_sweep_part_item<0, 4, 0, true, double>,
        (const void*)k_sweep_part_batch<1, 4, 0, false, double>, (const void*)k_sweep_part_item<1, 4, 0, false, double>,
        (const void*)k_sweep_part_batch<2, 4, 0, false, double>, (const void*)k_sweep_part_item<2, 4, 0, false, double>,
        (const void*)k_sweep_part_batch<0, 16, 16, true, float>, (const void*)k_sweep_part_item<0, 16, 16, true, float>,
        (const void*)k_sweep_part_batch<0, 4, 0, true, float>, (const void*)k_sweep_part_item<0, 4, 0, true, float>,
        (const void*)k_sweep_part_batch<1, 4, 0, false, float>, (const void*)k_sweep_part_item<1, 4, 0, false, float>,
        (const void*)k_sweep_part_batch<2, 4, 0, false, float>, (const void*)k_sweep_part_item<2, 4, 0, false, float>};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: partitioned sweep");
    }
    return 0;
}

// window size for a line of n nodes split into P segments (0 = does not fit)
int part_window(uint32_t n, int P) {
    if (n < 2 + (uint32_t)P) return 0;          // every segment holds at least one unknown
    const uint32_t m = (n - 2 + P - 1) / P;
    return m <= 16 ? 16 : (m <= 32 ? 32 : (m <= 64 && (P == 4 || P == 8) ? 64 : 0));
}

template <int AXIS, int P, bool XLDS, typename T>
static void launch_part_axis(hc_ctx* ctx, int mm, dim3 grid, size_t shm, T* rho, const uint8_t* tables,
                             const hc_grid* g, uint32_t nmax, int fuse, uint32_t xmap, uint32_t s0) {
    if (mm == 16)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 16, XLDS, T>), grid, dim3(256), shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse, xmap, s0, g->n_substrates);
    else if (mm == 32)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 32, XLDS, T>), grid, dim3(256), shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse, xmap, s0, g->n_substrates);
    else if constexpr ((P == 4 || P == 8) && AXIS != 0)
        hipLaunchKernelGGL((k_sweep_part<AXIS, P, 64, false, T>), grid, dim3(256), shm, ctx->stream, rho, tables, g->nx, g->ny, g->nz, nmax, fuse, xmap, s0, g->n_substrates);
}

// LDS bytes of a partitioned sweep launch along `axis` (0: the grid does not
// take the partitioned form)
template <typename T>
static size_t part_shm(int axis, const hc_grid* g, int sweep_mode) {
    const uint32_t n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    const int P = part_p(axis, n, g->field_fp32 != 0);
    if (!part_window(n, P)) return 0;
    const uint32_t nmax = grid_nmax(g);
    const bool xlds = axis == 0 && sweep_mode != 2;
    const size_t tab_bytes = P == 16 ? PartLds<16, T>::bytes(nmax) : (P == 8 ? PartLds<8, T>::bytes(nmax) : PartLds<4, T>::bytes(nmax));
    const size_t shm = tab_bytes + (xlds ? (size_t)4 * xtile_elems<T>(64 / P, n) * sizeof(T) : 0);
    return shm > 64 * 1024 ? 0 : shm;
}

bool part_fits_all(const hc_grid* g, int sweep_mode) {
    if (sweep_mode == 1) return false;
    for (int ax = 0; ax < (g->nz > 1 ? 3 : 2); ax++)
        if (!(g->field_fp32 ? part_shm<float>(ax, g, sweep_mode) : part_shm<double>(ax, g, sweep_mode))) return false;
    return true;
}

template <typename T>
static int launch_sweep_part_t(hc_ctx* ctx, int axis, T* rho, const hc_grid* g, const uint8_t* tables,
                               int fuse_dirichlet, uint32_t s0, uint32_t scount) {
    const uint32_t n = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    const int P = part_p(axis, n, g->field_fp32 != 0);
    const int mm = part_window(n, P);
    const size_t shm = part_shm<T>(axis, g, ctx->sweep_mode);
    if (!mm || !shm) return 1;
    const uint32_t nmax = grid_nmax(g);
    const int LPW = 64 / P;
    const bool xlds = axis == 0 && ctx->sweep_mode != 2;
    const uint32_t xmap = xlds ? (uint32_t)ctx_x_lane_order(ctx, n, P, sizeof(T)) : 0u;
    // one workgroup per line group (of 4 waves x LPW lines) and substrate
    const unsigned groups = blocks_for(sweep_lines(axis, g->nx, g->ny, g->nz), LPW * 4);
    dim3 grid(groups, scount);
    KernelTimer tm(ctx, HC_K_SWEEP_X + axis);
    if (axis == 0 && P == 4) {
        if (xlds) launch_part_axis<0, 4, true, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
        else launch_part_axis<0, 4, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
    } else if (axis == 0 && P == 8) {
        if (xlds) launch_part_axis<0, 8, true, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
        else launch_part_axis<0, 8, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
    } else if (axis == 0) {
        if (xlds) launch_part_axis<0, 16, true, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
        else launch_part_axis<0, 16, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
    } else if (axis == 1) {
        if (P == 8) launch_part_axis<1, 8, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
        else launch_part_axis<1, 4, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
    } else {
        if (P == 8) launch_part_axis<2, 8, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
        else launch_part_axis<2, 4, false, T>(ctx, mm, grid, shm, rho, tables, g, nmax, fuse_dirichlet, xmap, s0);
    }
    return check_launch(ctx, "k_sweep_part");
}

int launch_sweep_part(hc_ctx* ctx, int axis, void* rho, const hc_grid* g, const uint8_t* tables, int fuse, uint32_t s0,
                      uint32_t scount) {
    if (scount == 0) scount = g->n_substrates - s0;
    if (g->field_fp32) return launch_sweep_part_t<float>(ctx, axis, (float*)rho, g, tables, fuse, s0, scount);
    return launch_sweep_part_t<double>(ctx, axis, (double*)rho, g, tables, fuse, s0, scount);
}

template <int AX, int PP, int M, bool X, typename T>
static void launch_batch_kernel(hc_ctx* ctx, unsigned grid_blocks, size_t shm, const BatchSim* sims,
                                const BatchWork* work, uint32_t nwork, uint32_t nmax_all) {
    if (grid_blocks == 0)
        hipLaunchKernelGGL((k_sweep_part_item<AX, PP, M, X, T>), dim3(nwork), dim3(256), shm, ctx->stream, sims, work,
                           nmax_all);
    else
        hipLaunchKernelGGL((k_sweep_part_batch<AX, PP, M, X, T>), dim3(std::min<unsigned>(grid_blocks, nwork)),
                           dim3(256), shm, ctx->stream, sims, work, nwork, nmax_all);
}

template <typename T>
constexpr bool fp32_t() {
    return sizeof(T) == sizeof(float);
}

template <typename T>
static int launch_batch_t(hc_ctx* ctx, int axis, int P, int mm, const BatchSim* sims, const BatchWork* work,
                          uint32_t nwork, uint32_t nmax_all, uint32_t max_line, unsigned grid_blocks) {
    if (nwork == 0) return 0;
    if (axis != 0 && P != 4) return fail(ctx, HC_ERR_ARG, "batched sweep: y and z use 4 segments per line");
    // mm: 16 (every member's lines fit 16 values per lane; P = 16), 32
    // (P = 8, FP32 x lines), or 0 (each member's own window, 16 or 32; P = 4)
    const bool ok = (P == 4 && (mm == 0 || mm == 16)) || (P == 16 && mm == 16) || (P == 8 && axis == 0 && fp32_t<T>() && mm == 32);
    if (!ok || part_window(max_line, P) > (mm ? mm : 32))
        return fail(ctx, HC_ERR_CAPACITY, "batched sweep: line length outside the list's register window");
    const int LPW = 64 / P;
    const size_t tab_bytes = P == 16 ? PartLds<16, T>::bytes(nmax_all)
                                     : (P == 8 ? PartLds<8, T>::bytes(nmax_all) : PartLds<4, T>::bytes(nmax_all));
    const size_t shm = tab_bytes + (axis == 0 ? (size_t)4 * xtile_elems<T>(LPW, max_line) * sizeof(T) : 0);
    if (shm > 64 * 1024) return fail(ctx, HC_ERR_CAPACITY, "batched sweep: LDS");
    KernelTimer tm(ctx, HC_K_SWEEP_X + axis);
    if (axis == 0 && P == 16) launch_batch_kernel<0, 16, 16, true, T>(ctx, grid_blocks, shm, sims, work, nwork, nmax_all);
    else if (axis == 0 && P == 8) launch_batch_kernel<0, 8, 32, true, T>(ctx, grid_blocks, shm, sims, work, nwork, nmax_all);
    else if (axis == 0) launch_batch_kernel<0, 4, 0, true, T>(ctx, grid_blocks, shm, sims, work, nwork, nmax_all);
    else if (axis == 1) launch_batch_kernel<1, 4, 0, false, T>(ctx, grid_blocks, shm, sims, work, nwork, nmax_all);
    else launch_batch_kernel<2, 4, 0, false, T>(ctx, grid_blocks, shm, sims, work, nwork, nmax_all);
    return check_launch(ctx, "batched sweep");
}

int launch_sweep_batch(hc_ctx* ctx, int axis, int P, int mm, bool fp32, const BatchSim* sims, const BatchWork* work,
                       uint32_t nwork, uint32_t nmax_all, uint32_t max_line, unsigned grid_blocks) {
    if (fp32) return launch_batch_t<float>(ctx, axis, P, mm, sims, work, nwork, nmax_all, max_line, grid_blocks);
    return launch_batch_t<double>(ctx, axis, P, mm, sims, work, nwork, nmax_all, max_line, grid_blocks);
}

// Host construction of the partitioned tables (FP64, plus an FP32 rounding
// of the per-segment arrays for FP32 fields).
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

void build_part_tables_host(const hc_grid* g, const hc_substrate* subs, double* out, float* fout) {
    const uint32_t ns = g->n_substrates, nmax = grid_nmax(g);
    const size_t blk = part_block(nmax), fblk = part_fblock(nmax);
    const double sp[3] = {g->dx, g->dy, g->dz};
    const uint32_t nn[3] = {g->nx, g->ny, g->nz};
    for (uint32_t s = 0; s < ns; s++) {
        const double D = subs[s].diffusion_coefficient, lam = subs[s].decay_rate;
        const double h = g->nz > 1 ? lam * g->dt / 3.0 : 0.5 * lam * g->dt;
        for (int ax = 0; ax < 3; ax++) {
            const int P = part_p(ax, nn[ax], g->field_fp32 != 0);
            const uint32_t MT = part_mt(nmax, P);
            double* B = out + (size_t)(s * 3 + ax) * blk;
            float* F = fout + (size_t)(s * 3 + ax) * fblk;
            std::fill(B, B + blk, 0.0);
            std::fill(F, F + fblk, 0.0f);
            const uint32_t n = nn[ax];
            if (n < 2 + (uint32_t)P) continue;
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
                    const double vals[4] = {1.0 / den[t], cpl[t], al[k][t], be[k][t]};
                    for (int q = 0; q < 4; q++) {
                        B[(k * 4 + q) * MT + t] = vals[q];
                        F[(k * 4 + q) * MT + t] = (float)vals[q];
                    }
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
