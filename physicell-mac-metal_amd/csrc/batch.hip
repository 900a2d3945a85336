// batch.hip -- packed ensembles: many simulations of uneven grid and
// population size stepped together (gfx950).
//
// Layout.  Each member keeps its own density field, tables, cell SoA and rate
// arrays (so members can join and leave between steps); everything the step
// iterates over is packed:
//   * sweep work lists: one entry per (member, substrate, line group), walked
//     by one persistent grid per axis (k_sweep_part_batch, sweep_part.hip);
//   * agent tiles: one entry per (member, 256 consecutive agents);
//   * the hash: one counting sort over the concatenated bin space of all
//     members (member m's mech voxel v is global bin bin_off[m] + v), so keys,
//     counts, the ascending list and the hash-sorted records are packed arrays
//     over all agents;
//   * the secretion plan: one counting sort over the concatenated voxel space.
// A member's arithmetic does not depend on the other members: its results are
// identical alone and in any mix (and identical to the single-simulation path).
#include "motility_common.h"
#include "host/physicell_host.h"

#include <algorithm>
#include <cstring>
#include <memory>

namespace hc {

struct EnsSim {
    double* soa;            // member cell SoA (hc_cells_bytes(max_cells))
    const double* rates;    // member rate arrays
    void* rho;              // member density field
    const uint8_t* tables;  // member Thomas tables
    uint32_t max_cells, n, ns, nmax;
    uint32_t nx, ny, nz, V;
    double x_min, y_min, z_min, dx, dy, dz, dt_diff;
    double mvoxel, max_dist, dt_mech;
    uint32_t mnx, mny, mnz, mot;   // mot: bit 0 motility on, bit 1 chemotaxis by sensitivities
    uint32_t cell_off, bin_off, vox_off, _pad2;
};

struct AgentTile {
    uint32_t sim, c0;
};

__device__ inline double* sim_field(const EnsSim& m, int f) { return m.soa + (size_t)f * m.max_cells; }
__device__ inline uint32_t* sim_ufield(const EnsSim& m, int f) {
    return reinterpret_cast<uint32_t*>(m.soa + (size_t)HC_NUM_F64_FIELDS * m.max_cells) +
           (size_t)(f - HC_NUM_F64_FIELDS) * m.max_cells;
}

// mech keys (shaders/mechanics.metal:61-72) into the global bin space
__global__ __launch_bounds__(256) void k_b_mech_keys(const EnsSim* __restrict__ sims,
                                                     const AgentTile* __restrict__ tiles,
                                                     uint32_t* __restrict__ counts, uint32_t* __restrict__ kglob,
                                                     uint32_t* __restrict__ slot, uint8_t* __restrict__ xslab) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    const bool valid = c < m.n;
    uint32_t key = 0;
    if (valid) {
        const double* x = sim_field(m, HC_F_POS_X);
        const double* y = sim_field(m, HC_F_POS_Y);
        const double* z = sim_field(m, HC_F_POS_Z);
        const double px = x[c];
        int vi = (int)((px - m.x_min) / m.mvoxel);
        int vj = (int)((y[c] - m.y_min) / m.mvoxel);
        int vk = (int)((z[c] - m.z_min) / m.mvoxel);
        vi = vi < 0 ? 0 : (vi > (int)m.mnx - 1 ? (int)m.mnx - 1 : vi);
        vj = vj < 0 ? 0 : (vj > (int)m.mny - 1 ? (int)m.mny - 1 : vj);
        vk = vk < 0 ? 0 : (vk > (int)m.mnz - 1 ? (int)m.mnz - 1 : vk);
        const uint32_t local = (uint32_t)vk * m.mny * m.mnx + (uint32_t)vj * m.mnx + (uint32_t)vi;
        sim_ufield(m, HC_F_MECH_VOXEL)[c] = local;
        key = m.bin_off + local;
        kglob[m.cell_off + c] = key;
        // FP32 records: the x slab of the record's bin-relative x
        if (xslab)
            xslab[m.cell_off + c] =
                (uint8_t)xslab_of(bin_rel32(px, m.x_min, (uint32_t)vi, m.mvoxel), xslab_scale(m.mvoxel));
    }
    const uint32_t s = ballot_bin(valid, key, counts);
    if (valid) slot[m.cell_off + c] = s;
}

// stable placement + hash-sorted records (FP64 or bin-relative FP32)
// A member whose mechanics grid spans at most MARKX_MAX_EXTENT in x takes the
// FP64 mark pass of force_agent64x (pos32 x relative to its x_min); the
// gather and the force kernel decide this alike.
__device__ inline bool member_markx(const EnsSim& m) {
    return m.mvoxel > 0.0 && m.mvoxel <= 32.0 && m.max_dist <= 32.0 &&
           (double)m.mnx * m.mvoxel <= MARKX_MAX_EXTENT;
}

// FP32: records in each bin by x slab, then index (positions from
// k_bin_sort32, or slab_rank in bins above BIN_SORT_CAP).
template <bool FP32>
__global__ __launch_bounds__(256) void k_b_gather(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                  const uint32_t* __restrict__ kglob, const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ tmp, uint32_t* __restrict__ sorted,
                                                  const uint32_t* __restrict__ pos_of, const uint8_t* __restrict__ xslab,
                                                  void* __restrict__ recs, float4* __restrict__ pos32,
                                                  float* __restrict__ rpart) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];                       // one member per workgroup
    const uint32_t c = tl.c0 + threadIdx.x;
    // the tile's largest radius (FP32 as the records hold it; rounded up for
    // the FP64 mark pass); no lane returns before the barrier
    const double rc = c < m.n ? sim_field(m, HC_F_RADIUS)[c] : 0.0;
    block_max_to(FP32 ? (float)rc : __double2float_ru(rc), rpart);
    if (c >= m.n) return;
    const uint32_t g = m.cell_off + c;
    const uint32_t b = kglob[g];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos;
    if (end - beg <= BIN_SORT_CAP) {
        pos = pos_of[g];                                  // placed by k_bin_sort / k_bin_sort32
    } else if (FP32) {
        sorted[beg + stable_rank(tmp, beg, end, g)] = g;
        pos = beg + slab_rank(tmp, xslab, beg, end, g, xslab[g]);
    } else {
        pos = beg + stable_rank(tmp, beg, end, g);
        sorted[pos] = g;
    }
    const double px = sim_field(m, HC_F_POS_X)[c], py = sim_field(m, HC_F_POS_Y)[c], pz = sim_field(m, HC_F_POS_Z)[c];
    if (FP32) {
        const uint32_t lb = b - m.bin_off;
        const uint32_t bi = lb % m.mnx, bj = (lb / m.mnx) % m.mny, bk = lb / (m.mnx * m.mny);
        CellRec32 r;
        r.xg = (float)(px - m.x_min);
        r.y = bin_rel32(py, m.y_min, bj, m.mvoxel);
        r.z = bin_rel32(pz, m.z_min, bk, m.mvoxel);
        r.x = bin_rel32(px, m.x_min, bi, m.mvoxel);
        r.r = (float)sim_field(m, HC_F_RADIUS)[c];
        r.rep = (float)sim_field(m, HC_F_REPULSION)[c];
        r.adh = (float)sim_field(m, HC_F_ADHESION)[c];
        r.idx = g;
        reinterpret_cast<CellRec32*>(recs)[pos] = r;
    } else {
        if (pos32) {
            // x relative to the member's x_min (force_agent64x) or to the bin
            // corner (force_range64m); y and z relative to the bin corner
            const uint32_t lb = b - m.bin_off;
            const uint32_t bi = lb % m.mnx, bj = (lb / m.mnx) % m.mny, bk = lb / (m.mnx * m.mny);
            pos32[pos] = make_float4((float)(px - (member_markx(m) ? m.x_min : m.x_min + bi * m.mvoxel)),
                                     (float)(py - (m.y_min + bj * m.mvoxel)),
                                     (float)(pz - (m.z_min + bk * m.mvoxel)), (float)sim_field(m, HC_F_RADIUS)[c]);
        }
        CellRec r;
        r.x = px;
        r.y = py;
        r.z = pz;
        r.r = sim_field(m, HC_F_RADIUS)[c];
        r.rep = sim_field(m, HC_F_REPULSION)[c];
        r.adh = sim_field(m, HC_F_ADHESION)[c];
        r.rel = sim_field(m, HC_F_REL_MAX_ADH)[c];
        r.idx = g;
        r.key = b;
        reinterpret_cast<CellRec*>(recs)[pos] = r;
    }
}

// The single simulation's force passes over the member's bins: the mark
// passes with a list of marked candidates (FP32: LIST, force_agent32x; FP64:
// force_agent64x where member_markx holds), else the chunked ones.  rmax: the
// largest FP32 radius over all members.
// A member's agents hold the sorted positions [cell_off, cell_off + n) (its
// bins follow the earlier members' bins), so the agent tiles of a member,
// taken as sorted positions, give each workgroup one member: its descriptor
// is uniform (scalar registers) instead of a per-lane gather.
// FP32 members whose mechanics grid keeps the FP32 mark's error bound
// (member_markx, bins of at least 0.5 um) take the x-trimmed mark pass with a
// list; the others the chunked pass.
// MIRROR (FP32 only): the records' index field is the agent's mirror slot,
// kglob its bin (the mirror's key array), the relative adhesion distance
// comes from the mirror, the mark pass reads the 8-byte entries in mark[]
// (every agent takes the chunked pass when k_bm_gather set *xtflag) and the
// forces go to vrec in record order.
template <bool FP32, bool LIST = true, bool MIRROR = false>
__global__ __launch_bounds__(256) void k_b_forces(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                  const void* __restrict__ recs, const float4* __restrict__ pos32,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ xoff,
                                                  const uint32_t* __restrict__ kglob,
                                                  const float* __restrict__ rmax,
                                                  const float4* __restrict__ mprop, float4* __restrict__ vrec,
                                                  const uint2* __restrict__ mark,
                                                  const uint32_t* __restrict__ xtflag) {
    __shared__ uint32_t marks[LIST ? 4 * MARKX_CAP * 64 : 1];
    uint32_t* list = marks + (LIST ? (threadIdx.x >> 6) * (MARKX_CAP * 64) + (threadIdx.x & 63) : 0);
    const AgentTile tl = tiles[xcd_block(blockIdx.x, gridDim.x)];
    const EnsSim m = sims[tl.sim];
    if (tl.c0 + threadIdx.x >= m.n) return;
    const uint32_t t = m.cell_off + tl.c0 + threadIdx.x;   // sorted position
    uint32_t idx, key;
    if (FP32) {
        const CellRec32 me = reinterpret_cast<const CellRec32*>(recs)[t];
        idx = me.idx;
        key = kglob[idx];
    } else {
        const CellRec me = reinterpret_cast<const CellRec*>(recs)[t];
        idx = me.idx;
        key = me.key;
    }
    const uint32_t c = idx - m.cell_off;
    const uint32_t lk = key - m.bin_off;
    const uint32_t vi = lk % m.mnx, vj = (lk / m.mnx) % m.mny, vk = lk / (m.mnx * m.mny);
    double ofx = 0.0, ofy = 0.0, ofz = 0.0, opr = 0.0;
    if (FP32) {
        const CellRec32 me = reinterpret_cast<const CellRec32*>(recs)[t];
        const CellRec32* rec = reinterpret_cast<const CellRec32*>(recs);
        const float rel = MIRROR ? mprop[idx].w : (float)sim_field(m, HC_F_REL_MAX_ADH)[c];
        const float cut = (float)m.max_dist, vox = (float)m.mvoxel;
        float fx = 0.f, fy = 0.f, fz = 0.f, pr = 0.f;
        if (LIST && member_markx(m) && m.mvoxel >= 0.5 && (!MIRROR || xtflag[0] == 0u))
            force_agent32x<MIRROR>(rec, xoff, list, t, m.bin_off, vi, vj, vk, m.mnx, m.mny, m.mnz, me.xg, me.x, me.y,
                                   me.z, vox, xslab_scale(m.mvoxel), (float)(1.0 / m.mvoxel), rmax[0], me.r, me.rep,
                                   me.adh, rel, cut * cut, fx, fy, fz, pr, mark);
        else
            force_rows32(rec, offsets, t, m.bin_off, vi, vj, vk, m.mnx, m.mny, m.mnz, me.x, me.y, me.z, vox, me.r,
                         me.rep, me.adh, rel, cut * cut, fx, fy, fz, pr);
        if (MIRROR) {
            vrec[t] = make_float4(fx, fy, fz, pr);
            return;
        }
        ofx = fx; ofy = fy; ofz = fz; opr = pr;
    } else {
        const CellRec me = reinterpret_cast<const CellRec*>(recs)[t];
        const CellRec* rec = reinterpret_cast<const CellRec*>(recs);
        const double cut2 = m.max_dist * m.max_dist;
        double fx = 0.0, fy = 0.0, fz = 0.0, pr = 0.0;
        // FP32 mark pass where its error bound holds (force_range64m)
        const float vox = (float)m.mvoxel, near = 2.0f * vox;
        const float4 m32 = pos32 ? pos32[t] : make_float4(0.f, 0.f, 0.f, 0.f);
        const bool xgrid = member_markx(m);
        const bool yz_near = fabsf(m32.y) < near && fabsf(m32.z) < near;
        const bool markx = pos32 && rmax && xgrid && yz_near && m32.x > -near && m32.x < (float)m.mnx * vox + near;
        const bool mark32 = pos32 && !xgrid && m.mvoxel > 0.0 && m.mvoxel <= 32.0 && m.max_dist <= 32.0 &&
                            fabsf(m32.x) < near && yz_near;
        if (markx)
            force_agent64x(rec, pos32, offsets, list, m.bin_off, vi, vj, vk, m.mnx, m.mny, m.mnz, m32, vox, rmax[0],
                           me.x, me.y, me.z, me.r, me.rep, me.adh, me.rel, cut2, fx, fy, fz, pr);
#pragma unroll 1
        for (int q = 0; q < (markx ? 0 : 9); q++) {
            const RowRange rr = neighbor_row(offsets, m.bin_off, vi, vj, vk, m.mnx, m.mny, m.mnz, q);
            if (mark32)
                force_range64m(rec, pos32, rr.beg, rr.end, t, rr.mid0, rr.mid1, m32.x, m32.y, m32.z,
                               (float)rr.dj * vox, (float)rr.dk * vox, vox, me.x, me.y, me.z, me.r, me.rep, me.adh,
                               me.rel, cut2, fx, fy, fz, pr);
            else
                force_range64(rec, rr.beg, rr.end, t, me.x, me.y, me.z, me.r, me.rep, me.adh, me.rel, cut2, fx,
                              fy, fz, pr);
        }
        ofx = fx; ofy = fy; ofz = fz; opr = pr;
    }
    sim_field(m, HC_F_VEL_X)[c] = ofx;
    sim_field(m, HC_F_VEL_Y)[c] = ofy;
    sim_field(m, HC_F_VEL_Z)[c] = ofz;
    sim_field(m, HC_F_PRESSURE)[c] = opr;
}

// AB2 (shaders/integrate.metal:56-79), clamp to the member's mech-grid extent
__global__ __launch_bounds__(256) void k_b_integrate(const EnsSim* __restrict__ sims,
                                                     const AgentTile* __restrict__ tiles) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t i = tl.c0 + threadIdx.x;
    if (i >= m.n) return;
    const double x_max = m.x_min + (double)m.mnx * m.mvoxel;
    const double y_max = m.y_min + (double)m.mny * m.mvoxel;
    const double z_max = m.z_min + (double)m.mnz * m.mvoxel;
    const double dt = m.dt_mech;
    double* X = sim_field(m, HC_F_POS_X);
    double* Y = sim_field(m, HC_F_POS_Y);
    double* Z = sim_field(m, HC_F_POS_Z);
    double* PX = sim_field(m, HC_F_PREV_VX);
    double* PY = sim_field(m, HC_F_PREV_VY);
    double* PZ = sim_field(m, HC_F_PREV_VZ);
    const double vx = sim_field(m, HC_F_VEL_X)[i], vy = sim_field(m, HC_F_VEL_Y)[i], vz = sim_field(m, HC_F_VEL_Z)[i];
    double px = X[i], py = Y[i], pz = Z[i];
    px += dt * (1.5 * vx - 0.5 * PX[i]);
    py += dt * (1.5 * vy - 0.5 * PY[i]);
    pz += dt * (1.5 * vz - 0.5 * PZ[i]);
    px = px < m.x_min ? m.x_min : (px > x_max ? x_max : px);
    py = py < m.y_min ? m.y_min : (py > y_max ? y_max : py);
    pz = pz < m.z_min ? m.z_min : (pz > z_max ? z_max : pz);
    X[i] = px; Y[i] = py; Z[i] = pz;
    PX[i] = vx; PY[i] = vy; PZ[i] = vz;
}

// secretion plan (src/cell_phenotype.cpp:526-564) over the global voxel space
__global__ __launch_bounds__(256) void k_b_sec_keys(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                    uint32_t* __restrict__ counts, uint32_t* __restrict__ kv,
                                                    uint32_t* __restrict__ slot, uint32_t vtot) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    const bool valid = c < m.n;
    uint32_t key = vtot;
    if (valid) {
        const uint32_t v = sim_ufield(m, HC_F_VOXEL_IDX)[c];
        if (sim_ufield(m, HC_F_IS_ALIVE)[c] && v < m.V) key = m.vox_off + v;
        kv[m.cell_off + c] = key;
    }
    const uint32_t s = ballot_bin(valid, key, counts);
    if (valid) slot[m.cell_off + c] = s;
}

// C: coefficient storage, the field's type (FP32 fields keep FP32 coefficients,
// halving the bytes the per-substep apply reads; they are evaluated in FP64
// and rounded once)
template <typename C>
__global__ __launch_bounds__(256) void k_b_sec_groups(const EnsSim* __restrict__ sims,
                                                      const uint32_t* __restrict__ agent_sim,
                                                      const uint32_t* __restrict__ kv, const uint32_t* __restrict__ sorted,
                                                      const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ gid,
                                                      uint32_t* __restrict__ gsim, uint32_t* __restrict__ gvox,
                                                      uint32_t* __restrict__ gbegin, C* __restrict__ coef,
                                                      PlanHeader* __restrict__ hdr, uint32_t ntot, uint32_t vtot,
                                                      uint32_t ns_max) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n_elig = offsets[vtot];
    if (t == 0) {
        hdr->n_elig = n_elig;
        hdr->ngroups = gid[ntot];
        hdr->n = ntot;
        gbegin[gid[ntot]] = n_elig;
    }
    if (t >= n_elig) return;
    const uint32_t g = sorted[t];
    const uint32_t simi = agent_sim[g];
    const EnsSim m = sims[simi];
    const uint32_t c = g - m.cell_off;
    const uint32_t v = kv[g];
    if (t == 0 || kv[sorted[t - 1]] != v) {
        const uint32_t q = gid[t];
        gsim[q] = simi;
        gvox[q] = v - m.vox_off;
        gbegin[q] = t;
    }
    const size_t stride = (size_t)m.max_cells * m.ns;
    const double* sec = m.rates;
    const double* upt = m.rates + stride;
    const double* sat = m.rates + 2 * stride;
    const double* net = m.rates + 3 * stride;
    const double V_voxel = m.dx * m.dy * m.dz;
    const double dt = m.dt_diff;
    const double V_cell = sim_field(m, HC_F_TOTAL_VOL)[c];
    const double internal_constant = dt * V_cell / V_voxel;
    bool any_net = false;
    for (uint32_t s = 0; s < m.ns; s++) {
        const size_t off = (size_t)c * m.ns + s;
        const double a = dt * net[off] / V_voxel;
        const double temp1 = internal_constant * sec[off] * sat[off];
        const double temp2 = 1.0 + internal_constant * (sec[off] + upt[off]);
        coef[((size_t)s * 3 + 0) * ntot + t] = (C)a;
        coef[((size_t)s * 3 + 1) * ntot + t] = (C)temp1;
        coef[((size_t)s * 3 + 2) * ntot + t] = (C)temp2;
        any_net |= (a != 0.0);
    }
    (void)ns_max;
    if (any_net) atomicOr(&hdr->has_net, 1u);
}

template <typename F>
__global__ __launch_bounds__(256) void k_b_sec_apply(const EnsSim* __restrict__ sims, const uint32_t* __restrict__ gsim,
                                                     const uint32_t* __restrict__ gvox, const uint32_t* __restrict__ gbegin,
                                                     const F* __restrict__ coef, const PlanHeader* __restrict__ hdr,
                                                     uint32_t ntot) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= hdr->ngroups) return;
    const bool has_net = hdr->has_net != 0;
    const EnsSim m = sims[gsim[q]];
    F* rho = reinterpret_cast<F*>(m.rho);
    const uint32_t v = gvox[q];
    const uint32_t b = gbegin[q], e = gbegin[q + 1];
    for (uint32_t s = 0; s < m.ns; s++) {
        const F* ca = coef + ((size_t)s * 3 + 0) * ntot;
        const F* c1 = coef + ((size_t)s * 3 + 1) * ntot;
        const F* c2 = coef + ((size_t)s * 3 + 2) * ntot;
        double r = (double)rho[(uint64_t)s * m.V + v];
        for (uint32_t t = b; t < e; t++) {
            if (has_net) r += (double)ca[t];
            r = (r + (double)c1[t]) / (double)c2[t];
            if (r < 0.0) r = 0.0;
        }
        rho[(uint64_t)s * m.V + v] = (F)r;
    }
}

// motility of every member that has it (src/motility.cpp:179-293), between
// the forces and AB2, as the single simulation's k_motility: same per-agent
// update, the member's own field, grid, step counter and seed
template <typename F>
__global__ __launch_bounds__(256) void k_b_motility(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                    const MotPtrs* __restrict__ mots) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    if (!(m.mot & 1u) || c >= m.n) return;
    hc_grid g{};
    g.nx = m.nx; g.ny = m.ny; g.nz = m.nz; g.n_substrates = m.ns;
    g.dx = m.dx; g.dy = m.dy; g.dz = m.dz; g.x_min = m.x_min; g.y_min = m.y_min; g.z_min = m.z_min;
    motility_one<F>(cell_ptrs(m.soa, m.max_cells), mots[tl.sim], reinterpret_cast<const F*>(m.rho), g, m.dt_mech, c,
                    (m.mot >> 1) & 1u);
}

// each member's step counter advances after every motility call
__global__ void k_b_motility_advance(const EnsSim* __restrict__ sims, const MotPtrs* __restrict__ mots,
                                     uint32_t nsims) {
    const uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s < nsims && (sims[s].mot & 1u)) mots[s].hdr[0] += 1;
}

// ---------------------------------------------------------------------------
// Agent mirror of the batched FP32-mixed ensemble
// ---------------------------------------------------------------------------
// Agents in index order are not in spatial order, so every per-agent pass of
// the mechanics step in index order scatters: the key pass's ballots find
// about one bin per lane (one atomic each), the gather writes its records
// and the forces their results at unrelated addresses.  The mirror keeps each
// agent's mechanics state -- FP64 position and previous velocity, FP32
// radius, repulsion, adhesion and relative adhesion distance, index -- in
// slots ordered like the records of a recent step.  Slots keep their order
// from step to step (a periodic re-sort restores it as agents change bins),
// the next step's bin and slot in the bin are computed by AB2 itself, and
// the forces are written in record order; each slot knows its record
// position of the step (rpos).  Records, their order (bin, x slab, index),
// every arithmetic step and so every result are those of the direct path.
// The cell fields (position, velocity, previous velocity, pressure,
// mechanics voxel) are written back by index only when something reads them.
struct BMirror {
    Dbl3* pos;        // x, y, z: the FP64 AB2 accumulator
    Dbl3* prev;       // previous velocity
    float4* prop;     // radius, repulsion, adhesion, relative max adhesion distance (FP32)
    uint32_t* idx;    // global agent index (member cell_off + cell)
    uint32_t* key;    // global bin of the coming step
    uint32_t* slot;   // slot inside that bin
    uint8_t* xslab;   // x slab of the agent's FP32 record in that bin
    uint32_t* rpos;   // record position in the last step
    uint32_t* kstep;  // bin of the last step (the cell's mechanics voxel)
    float4* v;        // forces of the last step by record position: vx, vy, vz, pressure
};

// bin and x slab of a position (k_b_mech_keys's expressions)
__device__ inline uint32_t bm_key(const EnsSim& m, double px, double py, double pz, uint8_t& xs) {
    int vi = (int)((px - m.x_min) / m.mvoxel);
    int vj = (int)((py - m.y_min) / m.mvoxel);
    int vk = (int)((pz - m.z_min) / m.mvoxel);
    vi = vi < 0 ? 0 : (vi > (int)m.mnx - 1 ? (int)m.mnx - 1 : vi);
    vj = vj < 0 ? 0 : (vj > (int)m.mny - 1 ? (int)m.mny - 1 : vj);
    vk = vk < 0 ? 0 : (vk > (int)m.mnz - 1 ? (int)m.mnz - 1 : vk);
    xs = (uint8_t)xslab_of(bin_rel32(px, m.x_min, (uint32_t)vi, m.mvoxel), xslab_scale(m.mvoxel));
    return m.bin_off + (uint32_t)vk * m.mny * m.mnx + (uint32_t)vj * m.mnx + (uint32_t)vi;
}

// the mirror from the cell fields, slot = index; bins and slots of the
// current positions (counts must be zero)
__global__ __launch_bounds__(256) void k_bm_load(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                 BMirror mr, uint32_t* __restrict__ counts) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    const bool valid = c < m.n;
    const uint32_t p = m.cell_off + c;
    uint32_t key = 0;
    uint8_t xs = 0;
    if (valid) {
        const double px = sim_field(m, HC_F_POS_X)[c], py = sim_field(m, HC_F_POS_Y)[c],
                     pz = sim_field(m, HC_F_POS_Z)[c];
        mr.pos[p] = dbl3(px, py, pz);
        mr.prev[p] = dbl3(sim_field(m, HC_F_PREV_VX)[c], sim_field(m, HC_F_PREV_VY)[c],
                          sim_field(m, HC_F_PREV_VZ)[c]);
        mr.prop[p] = make_float4((float)sim_field(m, HC_F_RADIUS)[c], (float)sim_field(m, HC_F_REPULSION)[c],
                                 (float)sim_field(m, HC_F_ADHESION)[c], (float)sim_field(m, HC_F_REL_MAX_ADH)[c]);
        mr.idx[p] = p;
        key = bm_key(m, px, py, pz, xs);
        mr.key[p] = key;
        mr.xslab[p] = xs;
    }
    const uint32_t sl = ballot_bin(valid, key, counts);
    if (valid) mr.slot[p] = sl;
}

// Provisional placement of the mirror's slots in their bins, as one 64-bit
// entry each: agent index << 32 | x slab << 28 | slot (indices and slots
// below 2^27), so that the bin sort and the ranks of large bins need no
// further loads.
__global__ __launch_bounds__(256) void k_bm_scatter(const EnsSim* __restrict__ sims,
                                                    const AgentTile* __restrict__ tiles, BMirror mr,
                                                    const uint32_t* __restrict__ offsets,
                                                    uint64_t* __restrict__ tmp64) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    if (c >= m.n) return;
    const uint32_t p = m.cell_off + c;
    tmp64[offsets[mr.key[p]] + mr.slot[p]] = ((uint64_t)mr.idx[p] << 32) | ((uint64_t)mr.xslab[p] << 28) | p;
}

// Bins of at most BIN_SORT_CAP members: the entries in ascending agent index
// (the bin's list) and each slot's record position by (x slab, index);
// every bin: its slab starts (xoff).  Larger bins: slab starts only.
__global__ __launch_bounds__(256) void k_bm_bin_sort(const uint32_t* __restrict__ offsets,
                                                     const uint64_t* __restrict__ tmp64, BMirror mr,
                                                     uint32_t* __restrict__ sorted, uint32_t* __restrict__ xoff,
                                                     uint32_t nbins, uint32_t* __restrict__ xtflag) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b > nbins) return;
    if (b == nbins) {
        const uint32_t n = offsets[nbins];
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * nbins) = make_uint4(n, n, n, n);
        xtflag[0] = 0u;                                  // k_bm_gather sets it
        return;
    }
    const uint32_t beg = offsets[b], cnt = offsets[b + 1] - beg;
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    if (cnt == 0) {
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) = make_uint4(beg, beg, beg, beg);
        return;
    }
    if (cnt > BIN_SORT_CAP) {
        for (uint32_t q = beg; q < beg + cnt; q++) {
            const uint32_t sq = (uint32_t)(tmp64[q] >> 28) & 3u;
            c0 += sq == 0 ? 1u : 0u;
            c1 += sq == 1 ? 1u : 0u;
            c2 += sq == 2 ? 1u : 0u;
        }
        *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) =
            make_uint4(beg, beg + c0, beg + c0 + c1, beg + c0 + c1 + c2);
        return;
    }
    uint64_t a[BIN_SORT_CAP];
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) a[q] = q < cnt ? tmp64[beg + q] : ~0ull;
#pragma unroll
    for (uint32_t k = 2; k <= BIN_SORT_CAP; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t i = 0; i < BIN_SORT_CAP; i++) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t lo = a[i] < a[l] ? a[i] : a[l], hi = a[i] < a[l] ? a[l] : a[i];
                    const bool up = (i & k) == 0;
                    a[i] = up ? lo : hi;
                    a[l] = up ? hi : lo;
                }
            }
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++) {
        const uint32_t sq = q < cnt ? (uint32_t)(a[q] >> 28) & 3u : XSLABS;
        c0 += sq == 0 ? 1u : 0u;
        c1 += sq == 1 ? 1u : 0u;
        c2 += sq == 2 ? 1u : 0u;
    }
    const uint32_t x1 = beg + c0, x2 = x1 + c1, x3 = x2 + c2;
    *reinterpret_cast<uint4*>(xoff + (size_t)XSLABS * b) = make_uint4(beg, x1, x2, x3);
    uint32_t n0 = beg, n1 = x1, n2 = x2, n3 = x3;
#pragma unroll
    for (uint32_t q = 0; q < BIN_SORT_CAP; q++)
        if (q < cnt) {
            sorted[beg + q] = (uint32_t)(a[q] >> 32);
            const uint32_t sq = (uint32_t)(a[q] >> 28) & 3u;
            mr.rpos[(uint32_t)a[q] & ((1u << 28) - 1u)] = sq == 0 ? n0 : (sq == 1 ? n1 : (sq == 2 ? n2 : n3));
            n0 += sq == 0 ? 1u : 0u;
            n1 += sq == 1 ? 1u : 0u;
            n2 += sq == 2 ? 1u : 0u;
            n3 += sq == 3 ? 1u : 0u;
        }
}

// FP32 records from the mirror (k_b_gather<true>'s values); slots of bins
// above BIN_SORT_CAP are ranked here.  The record's index field holds the slot.
__global__ __launch_bounds__(256) void k_bm_gather(const EnsSim* __restrict__ sims,
                                                   const AgentTile* __restrict__ tiles, BMirror mr,
                                                   const uint32_t* __restrict__ offsets,
                                                   const uint64_t* __restrict__ tmp64, uint32_t* __restrict__ sorted,
                                                   CellRec32* __restrict__ recs, float* __restrict__ rpart,
                                                   uint2* __restrict__ mark, uint32_t* __restrict__ xtflag) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    const uint32_t p = m.cell_off + c;
    const float4 pr = c < m.n ? mr.prop[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (rpart) block_max_to(pr.x, rpart);
    if (c >= m.n) return;
    const uint32_t b = mr.key[p];
    const uint32_t beg = offsets[b], end = offsets[b + 1];
    uint32_t pos;
    if (end - beg <= BIN_SORT_CAP) {
        pos = mr.rpos[p];
    } else {
        const uint32_t id = mr.idx[p], sc = mr.xslab[p];
        uint32_t ri = 0, rs = 0;
        for (uint32_t q = beg; q < end; q++) {
            const uint64_t e = tmp64[q];
            const uint32_t io = (uint32_t)(e >> 32), so = (uint32_t)(e >> 28) & 3u;
            ri += io < id ? 1u : 0u;
            rs += (so < sc || (so == sc && io < id)) ? 1u : 0u;
        }
        sorted[beg + ri] = id;
        pos = beg + rs;
        mr.rpos[p] = pos;
    }
    mr.kstep[p] = b;
    const Dbl3 x = mr.pos[p];
    const uint32_t lb = b - m.bin_off;
    const uint32_t bi = lb % m.mnx, bj = (lb / m.mnx) % m.mny, bk = lb / (m.mnx * m.mny);
    CellRec32 r;
    r.xg = (float)(x.x - m.x_min);
    r.y = bin_rel32(x.y, m.y_min, bj, m.mvoxel);
    r.z = bin_rel32(x.z, m.z_min, bk, m.mvoxel);
    r.x = bin_rel32(x.x, m.x_min, bi, m.mvoxel);
    r.r = pr.x;
    r.rep = pr.y;
    r.adh = pr.z;
    r.idx = p;
    recs[pos] = r;
    bool fits;
    mark[pos] = mark_entry16(x.x - m.x_min, x.y - (m.y_min + bj * m.mvoxel), x.z - (m.z_min + bk * m.mvoxel), 0, fits);
    if (!fits) xtflag[0] = 1u;
}

// AB2 (shaders/integrate.metal:56-79, k_b_integrate's expressions) on the
// mirror, then the bin and slot of the new position for the coming step
// (counts must be zero)
__global__ __launch_bounds__(256) void k_bm_integrate(const EnsSim* __restrict__ sims,
                                                      const AgentTile* __restrict__ tiles, BMirror mr,
                                                      uint32_t* __restrict__ counts) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    const bool valid = c < m.n;
    const uint32_t p = m.cell_off + c;
    uint32_t key = 0;
    uint8_t xs = 0;
    if (valid) {
        const double x_max = m.x_min + (double)m.mnx * m.mvoxel;
        const double y_max = m.y_min + (double)m.mny * m.mvoxel;
        const double z_max = m.z_min + (double)m.mnz * m.mvoxel;
        const double dt = m.dt_mech;
        const float4 f = mr.v[mr.rpos[p]];
        const double vx = (double)f.x, vy = (double)f.y, vz = (double)f.z;
        const Dbl3 pv = mr.prev[p];
        const Dbl3 x = mr.pos[p];
        double px = x.x, py = x.y, pz = x.z;
        px += dt * (1.5 * vx - 0.5 * pv.x);
        py += dt * (1.5 * vy - 0.5 * pv.y);
        pz += dt * (1.5 * vz - 0.5 * pv.z);
        px = px < m.x_min ? m.x_min : (px > x_max ? x_max : px);
        py = py < m.y_min ? m.y_min : (py > y_max ? y_max : py);
        pz = pz < m.z_min ? m.z_min : (pz > z_max ? z_max : pz);
        mr.pos[p] = dbl3(px, py, pz);
        mr.prev[p] = dbl3(vx, vy, vz);
        key = bm_key(m, px, py, pz, xs);
        mr.key[p] = key;
        mr.xslab[p] = xs;
    }
    const uint32_t sl = ballot_bin(valid, key, counts);
    if (valid) mr.slot[p] = sl;
}

// the cell fields from the mirror, by agent index: position, velocity
// (forces of the last step), previous velocity, pressure, mechanics voxel
__global__ __launch_bounds__(256) void k_bm_sync(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                 BMirror mr) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c0 = tl.c0 + threadIdx.x;
    if (c0 >= m.n) return;
    const uint32_t p = m.cell_off + c0;
    const uint32_t c = mr.idx[p] - m.cell_off;
    const Dbl3 x = mr.pos[p], pv = mr.prev[p];
    const float4 f = mr.v[mr.rpos[p]];
    sim_field(m, HC_F_POS_X)[c] = x.x;
    sim_field(m, HC_F_POS_Y)[c] = x.y;
    sim_field(m, HC_F_POS_Z)[c] = x.z;
    sim_field(m, HC_F_VEL_X)[c] = (double)f.x;
    sim_field(m, HC_F_VEL_Y)[c] = (double)f.y;
    sim_field(m, HC_F_VEL_Z)[c] = (double)f.z;
    sim_field(m, HC_F_PRESSURE)[c] = (double)f.w;
    sim_field(m, HC_F_PREV_VX)[c] = pv.x;
    sim_field(m, HC_F_PREV_VY)[c] = pv.y;
    sim_field(m, HC_F_PREV_VZ)[c] = pv.z;
    sim_ufield(m, HC_F_MECH_VOXEL)[c] = mr.kstep[p] - m.bin_off;
}

// Re-sort: every slot's state moves to its record position of the last
// step (into the second set), which becomes its slot.
__global__ __launch_bounds__(256) void k_bm_permute(const EnsSim* __restrict__ sims,
                                                    const AgentTile* __restrict__ tiles, BMirror a, BMirror b) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    if (c >= m.n) return;
    const uint32_t p = m.cell_off + c;
    const uint32_t q = a.rpos[p];
    b.pos[q] = a.pos[p];
    b.prev[q] = a.prev[p];
    b.prop[q] = a.prop[p];
    b.idx[q] = a.idx[p];
    b.key[q] = a.key[p];
    b.slot[q] = a.slot[p];
    b.xslab[q] = a.xslab[p];
    b.kstep[q] = a.kstep[p];
    b.rpos[q] = q;
}

// the second set back into the first, slot by slot
__global__ __launch_bounds__(256) void k_bm_copy(const EnsSim* __restrict__ sims, const AgentTile* __restrict__ tiles,
                                                 BMirror a, BMirror b) {
    const AgentTile tl = tiles[blockIdx.x];
    const EnsSim m = sims[tl.sim];
    const uint32_t c = tl.c0 + threadIdx.x;
    if (c >= m.n) return;
    const uint32_t p = m.cell_off + c;
    a.pos[p] = b.pos[p];
    a.prev[p] = b.prev[p];
    a.prop[p] = b.prop[p];
    a.idx[p] = b.idx[p];
    a.key[p] = b.key[p];
    a.slot[p] = b.slot[p];
    a.xslab[p] = b.xslab[p];
    a.kstep[p] = b.kstep[p];
    a.rpos[p] = b.rpos[p];
}

int probe_kernels_batch(hc_ctx* ctx) {
    hipFuncAttributes attr;
    const void* fns[] = {(const void*)k_b_mech_keys, (const void*)k_b_gather<false>,
                         (const void*)k_b_gather<true>, (const void*)k_b_forces<false>, (const void*)k_b_forces<true, true>, (const void*)k_b_forces<true, false>,
                         (const void*)k_b_forces<true, true, true>, (const void*)k_bm_load, (const void*)k_bm_bin_sort,
                         (const void*)k_bm_gather, (const void*)k_bm_integrate, (const void*)k_bm_sync,
                         (const void*)k_bm_permute, (const void*)k_bm_copy, (const void*)k_bm_scatter,
                         (const void*)k_b_integrate, (const void*)k_b_sec_keys, (const void*)k_b_sec_groups<double>,
                         (const void*)k_b_sec_groups<float>,
                         (const void*)k_b_sec_apply<double>, (const void*)k_b_sec_apply<float>,
                         (const void*)k_b_motility<double>, (const void*)k_b_motility<float>,
                         (const void*)k_b_motility_advance};
    for (const void* f : fns) {
        hipError_t e = hipFuncGetAttributes(&attr, f);
        if (e != hipSuccess) return hip_fail(ctx, e, "hc_load_kernels: ensemble kernel");
    }
    return 0;
}

}  // namespace hc

// ============================================================================
// Host side: the ensemble (members, packing, step loop) and the pe_* C-ABI
// ============================================================================
extern "C" hipStream_t hc_stream_of(hc_ctx* ctx);
extern "C" hipStream_t hc_stream_swap(hc_ctx* ctx, hipStream_t s);

namespace physicell {

using hc::AgentTile;
using hc::BatchSim;
using hc::BatchWork;
using hc::EnsSim;

struct Member {
    uint32_t id = 0;
    SimConfig cfg;
    hc_grid grid{};
    std::vector<hc_substrate> subs;
    hc_mech mech{};
    hc_grid mgrid{};
    uint32_t n = 0, max_cells = 0;
    DeviceBuffer density, tables, soa, rates;
    DeviceBuffer mot;                 // motility state (hc_motility_bytes), empty when off
    bool mot_sens = false;            // chemotaxis by per-substrate sensitivities
    DeviceBuffer vol, mask;           // volume state (hc_volume_bytes) and phenotype mask, empty when off
    // the member's own schedule (the step loop of src/main.mm:275-415)
    double mech_acc = 0.0, pheno_acc = 0.0;
    int64_t substeps = 0, mech_steps = 0, pheno_steps = 0;
    size_t bytes() const {
        return density.size() + tables.size() + soa.size() + rates.size() + mot.size() + vol.size() + mask.size();
    }
};

class Ensemble {
    // sweep work lists, one per (axis, segments per line), in launch order:
    // x (4 segments), x (16 segments), x (8 segments: FP32 lines of 131..258),
    // y, z.  The 4-segment lists hold members of both register windows (16 or
    // 32 values per lane, each member's own: BatchSim::wide), those with the
    // longer lines first; the 16-segment x list holds lines of at most 16
    // values per lane, the 8-segment one lines of at most 32.
    static constexpr int NLIST = 5;
    static constexpr int LIST_AXIS[NLIST] = {0, 0, 0, 1, 2};
    static constexpr int LIST_P[NLIST] = {4, 16, 8, 4, 4};
    static constexpr int LIST_MM[NLIST] = {0, 16, 32, 0, 0};
    static int list_of(int axis, int P) {
        for (int li = 0; li < NLIST; li++)
            if (LIST_AXIS[li] == axis && LIST_P[li] == P) return li;
        return -1;
    }

public:
    Ensemble(int device, bool mixed) : mixed_(mixed) {
        int rc = hc_create(&ctx_, device);
        if (rc) throw std::runtime_error("ensemble: no usable HIP device");
        rc = hc_load_kernels(ctx_, nullptr);
        if (rc == 0 && (hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) != hipSuccess ||
                        hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess ||
                        hipEventCreateWithFlags(&join_, hipEventDisableTiming) != hipSuccess))
            rc = HC_ERR_HIP;
        if (rc) {
            std::string m = rc == HC_ERR_HIP ? "stream/event creation failed" : hc_last_error(ctx_);
            release_streams();
            hc_destroy(ctx_);
            throw std::runtime_error("ensemble: " + m);
        }
    }
    ~Ensemble() {
        drop_graph();
        hc_wait(ctx_);
        members_.clear();
        ws_ = Workspace();
        q_hash_ = DeviceBuffer();
        q_off_ = DeviceBuffer();
        q_idx_ = DeviceBuffer();
        release_streams();
        hc_destroy(ctx_);
    }
    // mechanics on a second stream, concurrent with the period's diffusion
    void set_overlap(bool on) {
        overlap_ = on;
        drop_graph();
    }

    uint32_t add(const SimConfig& c_in, uint32_t n, const double* x, const double* y, const double* z,
                 const double* radius, const double* volume, const double* rep, const double* adh, const double* rel,
                 const double* sec, const double* upt, const double* sat, const double* net, const uint32_t* alive) {
        sync_soa();
        auto m = std::make_unique<Member>();
        m->id = next_id_++;
        m->cfg = c_in;
        m->cfg.mixed_precision = mixed_;
        SimConfig& c = m->cfg;
        hc_grid& g = m->grid;
        g.nx = (uint32_t)((c.x_max - c.x_min) / c.dx);
        g.ny = (uint32_t)((c.y_max - c.y_min) / c.dy);
        g.nz = c.use_2D ? 1u : (uint32_t)((c.z_max - c.z_min) / c.dz);
        g.dx = c.dx; g.dy = c.dy; g.dz = c.dz;
        g.x_min = c.x_min; g.y_min = c.y_min; g.z_min = c.z_min;
        g.dt = c.dt_diffusion;
        g.n_substrates = (uint32_t)c.substrates.size();
        g.field_fp32 = mixed_ ? 1u : 0u;
        if (!g.nx || !g.ny || !g.nz) throw std::runtime_error("ensemble: empty grid");
        if (!grid_fits(g))
            throw std::runtime_error("ensemble: grid edges must be 6..258 (x) and 6..130 (y, z) voxels; "
                                     "run other grids as single simulations");
        check_schedule(c);
        m->subs.resize(g.n_substrates);
        for (uint32_t s = 0; s < g.n_substrates; s++)
            m->subs[s] = hc_substrate{c.substrates[s].diffusion_coeff, c.substrates[s].decay_rate, 0.0,
                                      c.substrates[s].dirichlet_value, c.substrates[s].dirichlet_enabled ? 1u : 0u, 0u};
        const double v = 30.0;
        m->mech.dt = c.dt_mechanics;
        m->mech.cell_cell_repulsion = 10.0;
        m->mech.cell_cell_adhesion = 0.4;
        m->mech.max_interaction_distance = v;
        m->mech.mechanics_voxel_size = v;
        m->mech.mech_grid_nx = std::max(1u, (uint32_t)((c.x_max - c.x_min) / v));
        m->mech.mech_grid_ny = std::max(1u, (uint32_t)((c.y_max - c.y_min) / v));
        m->mech.mech_grid_nz = c.use_2D ? 1u : std::max(1u, (uint32_t)((c.z_max - c.z_min) / v));
        m->mech.forces_fp32 = mixed_ ? 1u : 0u;
        m->mgrid = hc_grid{};
        m->mgrid.nx = m->mech.mech_grid_nx; m->mgrid.ny = m->mech.mech_grid_ny; m->mgrid.nz = m->mech.mech_grid_nz;
        m->mgrid.dx = m->mgrid.dy = m->mgrid.dz = v;
        m->mgrid.x_min = c.x_min; m->mgrid.y_min = c.y_min; m->mgrid.z_min = c.z_min;
        // device state
        const size_t V = (size_t)g.nx * g.ny * g.nz;
        const size_t elem = mixed_ ? sizeof(float) : sizeof(double);
        m->density.allocate(ctx_, std::max<size_t>(8, V * g.n_substrates * elem));
        if (mixed_) {
            float* d = (float*)m->density.host();
            for (uint32_t s = 0; s < g.n_substrates; s++)
                std::fill(d + s * V, d + (s + 1) * V, (float)c.substrates[s].initial_value);
        } else {
            double* d = (double*)m->density.host();
            for (uint32_t s = 0; s < g.n_substrates; s++)
                std::fill(d + s * V, d + (s + 1) * V, c.substrates[s].initial_value);
        }
        m->density.upload_all();
        m->tables.allocate(ctx_, hc_thomas_table_bytes(&g));
        check(ctx_, hc_build_thomas_tables(ctx_, &g, m->subs.data(), m->tables.get()), "ensemble tables");
        m->n = n;
        m->max_cells = std::max(1u, n);
        m->soa.allocate(ctx_, hc_cells_bytes(m->max_cells));
        m->rates.allocate(ctx_, std::max<size_t>(8, hc_rates_bytes(m->max_cells, g.n_substrates)));
        char* h = (char*)m->soa.host();
        auto F = [&](uint32_t f) { return (double*)(h + hc_cells_field_offset(f, m->max_cells)); };
        auto U = [&](uint32_t f) { return (uint32_t*)(h + hc_cells_field_offset(f, m->max_cells)); };
        std::memset(h, 0, m->soa.size());
        // fresh-cell values the caller does not pass, as a single simulation's
        // cell store writes them (src/cell_data.cpp:544-680): nuclear volume
        // 540 um^3 and its equivalent-sphere radius, motility speed 1 and
        // bias 0.5, oncoprotein 1
        const double pi = 3.14159265358979323846;
        const double nuc_rad = std::cbrt(3.0 * 540.0 / (4.0 * pi));
        for (uint32_t i = 0; i < n; i++) {
            F(HC_F_NUCLEAR_VOL)[i] = 540.0; F(HC_F_NUCLEAR_RAD)[i] = nuc_rad;
            F(HC_F_MOT_SPEED)[i] = 1.0; F(HC_F_MOT_BIAS)[i] = 0.5; F(HC_F_ONCOPROTEIN)[i] = 1.0;
            F(HC_F_POS_X)[i] = x[i]; F(HC_F_POS_Y)[i] = y[i]; F(HC_F_POS_Z)[i] = z[i];
            F(HC_F_RADIUS)[i] = radius[i]; F(HC_F_TOTAL_VOL)[i] = volume[i];
            F(HC_F_REPULSION)[i] = rep[i]; F(HC_F_ADHESION)[i] = adh[i]; F(HC_F_REL_MAX_ADH)[i] = rel[i];
            U(HC_F_IS_ALIVE)[i] = alive ? alive[i] : 1u;
            // placement voxel, 3-D formula (src/microenvironment.mm:176-184)
            int ii = (int)((x[i] - g.x_min) / g.dx), jj = (int)((y[i] - g.y_min) / g.dy), kk = (int)((z[i] - g.z_min) / g.dz);
            ii = ii < 0 ? 0 : (ii >= (int)g.nx ? (int)g.nx - 1 : ii);
            jj = jj < 0 ? 0 : (jj >= (int)g.ny ? (int)g.ny - 1 : jj);
            kk = kk < 0 ? 0 : (kk >= (int)g.nz ? (int)g.nz - 1 : kk);
            U(HC_F_VOXEL_IDX)[i] = (uint32_t)kk * g.nx * g.ny + (uint32_t)jj * g.nx + (uint32_t)ii;
        }
        double* r = (double*)m->rates.host();
        const size_t stride = (size_t)m->max_cells * g.n_substrates;
        std::memset(r, 0, m->rates.size());
        for (size_t i = 0; i < (size_t)n * g.n_substrates; i++) {
            r[i] = sec ? sec[i] : 0.0;
            r[stride + i] = upt ? upt[i] : 0.0;
            r[2 * stride + i] = sat ? sat[i] : 0.0;
            r[3 * stride + i] = net ? net[i] : 0.0;
        }
        m->soa.upload_all();
        m->rates.upload_all();
        check(ctx_, hc_wait(ctx_), "wait");
        const uint32_t id = m->id;
        members_.push_back(std::move(m));
        dirty_ = true;
        plan_dirty_ = true;
        return id;
    }

    void remove(uint32_t id) {
        sync_soa();
        auto it = std::find_if(members_.begin(), members_.end(), [&](const std::unique_ptr<Member>& m) { return m->id == id; });
        if (it == members_.end()) throw std::runtime_error("ensemble: unknown member");
        check(ctx_, hc_wait(ctx_), "wait");
        members_.erase(it);
        dirty_ = true;
        plan_dirty_ = true;
    }

    // Motility for one member (pc_sim_set_motility for a single simulation):
    // defaults as MotilityData::setDefaults (src/motility.cpp:137-157) where no
    // array is given; speed goes to the member's cell field.
    void set_motility(uint32_t id, uint64_t seed, const double* persistence, const double* bias, const double* speed,
                      const uint32_t* is_motile, const uint32_t* restrict_2d, const int32_t* chem_index,
                      const int32_t* chem_dir, const double* sensitivity) {
        sync_soa();
        Member& m = member(id);
        check(ctx_, hc_wait(ctx_), "wait");
        const uint32_t ns = m.grid.n_substrates, mc = m.max_cells, n = m.n;
        m.mot.allocate(ctx_, hc_motility_bytes(mc, ns));
        uint8_t* h = (uint8_t*)m.mot.host();
        std::memset(h, 0, m.mot.size());
        uint64_t* hdr = (uint64_t*)h;
        hdr[0] = 0;
        hdr[1] = seed;
        auto F = [&](uint32_t f) { return (double*)(h + hc_motility_field_offset(f, mc, ns)); };
        auto U = [&](uint32_t f) { return (uint32_t*)(h + hc_motility_field_offset(f, mc, ns)); };
        auto I = [&](uint32_t f) { return (int32_t*)(h + hc_motility_field_offset(f, mc, ns)); };
        for (uint32_t i = 0; i < n; i++) {
            F(HC_M_PERSISTENCE)[i] = persistence ? persistence[i] : 15.0;
            F(HC_M_BIAS)[i] = bias ? bias[i] : 0.5;
            U(HC_M_MOTILE)[i] = is_motile ? is_motile[i] : 0u;
            U(HC_M_RESTRICT_2D)[i] = restrict_2d ? restrict_2d[i] : (m.cfg.use_2D ? 1u : 0u);
            I(HC_M_CHEM_INDEX)[i] = chem_index ? chem_index[i] : 0;
            I(HC_M_CHEM_DIR)[i] = chem_dir ? chem_dir[i] : 1;
            for (uint32_t q = 0; q < ns; q++)
                F(HC_M_SENSITIVITY)[(size_t)i * ns + q] = sensitivity ? sensitivity[(size_t)i * ns + q] : 0.0;
        }
        m.mot.upload_all();
        if (speed) {
            // the member's cells have moved on the device: write the speed field alone
            double* sp = (double*)((char*)m.soa.host() + hc_cells_field_offset(HC_F_MOT_SPEED, mc));
            for (uint32_t i = 0; i < n; i++) sp[i] = speed[i];
            check(ctx_, hc_upload(ctx_, m.soa.get(), hc_cells_field_offset(HC_F_MOT_SPEED, mc), (size_t)n * 8),
                  "speed upload");
        }
        check(ctx_, hc_wait(ctx_), "wait");
        m.mot_sens = sensitivity != nullptr;
        dirty_ = true;
    }

    // Volume state for one member's device phenotype step (pc_sim_set_volume
    // for a single simulation: CellData::setDefaults, src/cell_data.cpp:
    // 555-624, where fields[f] is NULL), with room for max_cells cells: the
    // member's cell, rate and motility arrays move to the larger layout.
    void set_volume(uint32_t id, uint64_t seed, const double* const* fields, uint32_t max_cells) {
        sync_soa();
        Member& m = member(id);
        check(ctx_, hc_wait(ctx_), "wait");
        if (max_cells > m.max_cells) grow(m, max_cells);
        const uint32_t mc = m.max_cells, n = m.n;
        m.vol.allocate(ctx_, hc_volume_bytes(mc));
        uint8_t* h = (uint8_t*)m.vol.host();
        std::memset(h, 0, m.vol.size());
        uint64_t* hdr = (uint64_t*)h;
        hdr[0] = 0;
        hdr[1] = seed;
        const double V = 2494.0, Vn = 540.0, ff = 0.75, Vc = V - Vn;
        const double Vns = Vn - ff * Vn, Vcs = Vc - ff * Vc;
        const double fresh[HC_NUM_V_FIELDS] = {ff * V, Vns, Vcs, ff, 0.0, Vcs, Vns, ff, 3.62,
                                               3.0 / 60.0, 0.33 / 60.0, 0.27 / 60.0, 0.0, 2.0 * V, 2.0};
        for (uint32_t f = 0; f < HC_NUM_V_FIELDS; f++) {
            double* a = (double*)(h + hc_volume_field_offset(f, mc));
            for (uint32_t i = 0; i < mc; i++) a[i] = fresh[f];
            if (fields && fields[f]) std::memcpy(a, fields[f], (size_t)n * sizeof(double));
        }
        m.vol.upload_all();
        check(ctx_, hc_wait(ctx_), "wait");
        dirty_ = true;
    }

    // One member's device phenotype step, as pc_sim_update_phenotype: the
    // volume model on the masked cells, divisions, removals, then the
    // member's voxel refresh and the secretion re-plan.
    uint32_t update_phenotype(uint32_t id, double dt, int use_ratio, const uint32_t* mask, const uint32_t* parents,
                              uint32_t n_parents, const uint32_t* removed, uint32_t n_removed) {
        sync_soa();
        Member& m = member(id);
        if (!m.vol.get()) throw std::runtime_error("ensemble: set the member's volume state first");
        check(ctx_, hc_wait(ctx_), "wait");
        const uint32_t mc = m.max_cells, ns = m.grid.n_substrates;
        uint32_t n = m.n;
        hc_buf* mbuf = nullptr;
        if (mask && n) {
            if (!m.mask.get() || m.mask.size() < (size_t)mc * 4) m.mask.allocate(ctx_, (size_t)mc * 4);
            std::memcpy(m.mask.host(), mask, (size_t)n * 4);
            m.mask.upload(0, (size_t)n * 4);
            mbuf = m.mask.get();
        }
        check(ctx_, hc_volume_update(ctx_, m.soa.get(), m.vol.get(), mbuf, n, mc, dt, use_ratio), "volume");
        hc_buf* mot = m.mot.get();
        uint32_t n1 = n, n2 = n;
        check(ctx_, hc_divide(ctx_, m.soa.get(), m.rates.get(), m.vol.get(), mot, ns, parents, n_parents, n, mc, &n1),
              "divide");
        check(ctx_, hc_remove(ctx_, m.soa.get(), m.rates.get(), m.vol.get(), mot, ns, removed, n_removed, n1, mc, &n2),
              "remove");
        m.n = n2;
        check(ctx_, hc_voxel_keys(ctx_, m.soa.get(), &m.grid, m.n, m.max_cells), "voxel refresh");
        check(ctx_, hc_wait(ctx_), "wait");
        dirty_ = true;
        plan_dirty_ = true;
        return n2;
    }

    // Neighbour lists of one member (hc_neighbor_query, kind HC_QUERY_*) over
    // a hash of that member alone, built from its current positions into a
    // query workspace; as pc_sim_neighbor_query for a single simulation.
    uint64_t neighbor_query(uint32_t id, int kind, std::vector<uint32_t>& off, std::vector<uint32_t>& idx) {
        sync_soa();
        Member& m = member(id);
        check(ctx_, hc_wait(ctx_), "wait");
        const uint32_t nmv = m.mech.mech_grid_nx * m.mech.mech_grid_ny * m.mech.mech_grid_nz;
        const size_t hb = hc_hash_bytes(nmv, m.max_cells);
        if (!q_hash_.get() || q_hash_.size() < hb) q_hash_.allocate(ctx_, hb);
        if (!q_off_.get() || q_off_.size() < ((size_t)m.n + 1) * 4) q_off_.allocate(ctx_, ((size_t)m.n + 1) * 4);
        check(ctx_, hc_build_hash(ctx_, m.soa.get(), q_hash_.get(), &m.mech, &m.mgrid, m.n, m.max_cells, nmv),
              "member neighbour query: hash");
        uint64_t total = 0;
        check(ctx_, hc_neighbor_query(ctx_, m.soa.get(), q_hash_.get(), &m.mech, m.n, m.max_cells, nmv, kind,
                                      q_off_.get(), nullptr, &total), "member neighbour query: count");
        if (!q_idx_.get() || q_idx_.size() < std::max<uint64_t>(16, total * 4))
            q_idx_.allocate(ctx_, std::max<uint64_t>(16, total * 4));
        check(ctx_, hc_neighbor_query(ctx_, m.soa.get(), q_hash_.get(), &m.mech, m.n, m.max_cells, nmv, kind,
                                      q_off_.get(), q_idx_.get(), &total), "member neighbour query");
        q_off_.download(0, ((size_t)m.n + 1) * 4);
        if (total) q_idx_.download(0, total * 4);
        check(ctx_, hc_wait(ctx_), "wait");
        off.assign((const uint32_t*)q_off_.host(), (const uint32_t*)q_off_.host() + m.n + 1);
        idx.assign((const uint32_t*)q_idx_.host(), (const uint32_t*)q_idx_.host() + total);
        return total;
    }

    // tryFormAllAttachments for one member (pc_sim_form_attachments)
    uint32_t form_attachments(uint32_t id, double dt, const double* rate, uint32_t* count, uint32_t* targets,
                              uint64_t seed) {
        std::vector<uint32_t> off, idx;
        neighbor_query(id, HC_QUERY_ATTACH, off, idx);
        return attach_greedy(member(id).n, off, idx, dt, rate, count, targets, seed);
    }

    // one period: substeps_per_mech diffusion substeps, one mechanics step,
    // then each member's own phenotype-cadence refresh when it fires
    void run_periods(int periods, bool use_graph) {
        rebuild();
        if (members_.empty()) return;
        for (int p = 0; p < periods; p++) {
            const bool mir = mirror_on();
            if (mir) {
                // eager, before the period (graph replays assume a loaded mirror)
                if (!mirror_valid_) mirror_load();
                else if (steps_since_sort_ >= MIRROR_RESORT_STEPS) mirror_resort();
            }
            if (use_graph) {
                if (!graph_) capture();
                hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_, hc_stream_of(ctx_));
                if (e != hipSuccess) throw std::runtime_error(std::string("ensemble graph launch: ") + hipGetErrorString(e));
            } else {
                period_kernels();
            }
            if (mir) {
                soa_stale_ = true;
                steps_since_sort_++;
            }
            advance_clocks();
        }
    }

    void wait() { check(ctx_, hc_wait(ctx_), "wait"); }

    Member& member(uint32_t id) {
        for (auto& m : members_)
            if (m->id == id) return *m;
        throw std::runtime_error("ensemble: unknown member");
    }

    void get_density(uint32_t id, double* out) {
        Member& m = member(id);
        wait();
        m.density.download_all();
        wait();
        const size_t cnt = (size_t)m.grid.nx * m.grid.ny * m.grid.nz * m.grid.n_substrates;
        if (mixed_) {
            const float* f = (const float*)m.density.host();
            for (size_t i = 0; i < cnt; i++) out[i] = f[i];
        } else {
            std::memcpy(out, m.density.host(), cnt * sizeof(double));
        }
    }

    void get_field(uint32_t id, uint32_t field, void* out) {
        sync_soa();
        Member& m = member(id);
        wait();
        const size_t w = field < HC_NUM_F64_FIELDS ? 8 : 4;
        const size_t off = hc_cells_field_offset(field, m.max_cells);
        m.soa.download(off, (size_t)m.n * w);
        wait();
        std::memcpy(out, (char*)m.soa.host() + off, (size_t)m.n * w);
    }

    size_t member_bytes(uint32_t id) { return member(id).bytes(); }
    size_t workspace_bytes() const { return ws_.bytes; }
    size_t n_members() const { return members_.size(); }
    uint64_t total_cells() const { uint64_t t = 0; for (auto& m : members_) t += m->n; return t; }
    uint64_t total_voxel_values() const {
        uint64_t t = 0;
        for (auto& m : members_) t += (uint64_t)m->grid.nx * m->grid.ny * m->grid.nz * m->grid.n_substrates;
        return t;
    }
    int substeps_per_mech() const {
        if (members_.empty()) return 1;
        double acc = 0.0;
        const double dtd = members_[0]->cfg.dt_diffusion, dtm = members_[0]->cfg.dt_mechanics;
        for (int k = 1; k < 1000000; k++) {
            acc += dtd;
            if (acc >= dtm - 0.5 * dtd) return k;
        }
        return 1;
    }
    void set_grid_blocks(int axis, unsigned b) {
        if (axis < 0 || axis > 2) throw std::runtime_error("ensemble: axis must be 0..2");
        for (int li = 0; li < NLIST; li++)
            if (LIST_AXIS[li] == axis) grid_blocks_[li] = b;
        drop_graph();
    }
    hc_ctx* ctx() { return ctx_; }
    const std::vector<std::unique_ptr<Member>>& members() const { return members_; }

    // Times the batched sweep of each axis at candidate persistent-grid sizes
    // (mean over `reps` launches after one warm-up) and keeps the fastest.
    // The sweeps advance the fields, so they are saved first and put back.
    void autotune(int reps, uint32_t blocks_out[3], double us_out[3]) {
        rebuild();
        drop_graph();
        hipStream_t st = hc_stream_of(ctx_);
        std::vector<DeviceBuffer> saved(members_.size());
        for (size_t i = 0; i < members_.size(); i++) {
            saved[i].allocate(ctx_, members_[i]->density.size());
            hipError_t e = hipMemcpyAsync(hc_device_ptr(saved[i].get()), hc_device_ptr(members_[i]->density.get()),
                                          members_[i]->density.size(), hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) throw std::runtime_error(std::string("autotune save: ") + hipGetErrorString(e));
        }
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            if (a) (void)hipEventDestroy(a);
            throw std::runtime_error("autotune: hipEventCreate failed");
        }
        const unsigned cand[] = {0, 256, 512, 1024, 2048, 4096, 8192, 16384};
        Workspace& w = ws_;
        bool counted[3] = {false, false, false};
        uint32_t busiest[3] = {0, 0, 0};
        for (int ax = 0; ax < 3; ax++) {
            blocks_out[ax] = 0;
            us_out[ax] = 0.0;
        }
        for (int li = 0; li < NLIST; li++) {
            if (!w.nwork[li]) continue;
            double best = 1e300;
            unsigned pick = grid_blocks_[li];
            for (unsigned c : cand) {
                if (c > 2 * w.nwork[li]) break;
                sweep_list(li, c);
                hipError_t e = hipEventRecord(a, st);
                for (int r = 0; r < reps; r++) sweep_list(li, c);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventRecord(b, st);
                if (e == hipSuccess) e = hipEventSynchronize(b);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
                if (e != hipSuccess) {
                    (void)hipEventDestroy(a);
                    (void)hipEventDestroy(b);
                    throw std::runtime_error(std::string("autotune timing: ") + hipGetErrorString(e));
                }
                const double us = 1e3 * ms / std::max(1, reps);
                if (us < best) {
                    best = us;
                    pick = c;
                }
            }
            grid_blocks_[li] = pick;
            const int ax = LIST_AXIS[li];
            // per axis: the grid of the axis's busiest list, the summed time
            if (!counted[ax] || w.nwork[li] > busiest[ax]) {
                blocks_out[ax] = pick;
                busiest[ax] = w.nwork[li];
            }
            us_out[ax] += best;
            counted[ax] = true;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        for (size_t i = 0; i < members_.size(); i++) {
            hipError_t e = hipMemcpyAsync(hc_device_ptr(members_[i]->density.get()), hc_device_ptr(saved[i].get()),
                                          members_[i]->density.size(), hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) throw std::runtime_error(std::string("autotune restore: ") + hipGetErrorString(e));
        }
        check(ctx_, hc_wait(ctx_), "wait");
    }
    std::string last_error;

private:
    DeviceBuffer q_hash_, q_off_, q_idx_;     // member neighbour-query workspace
    struct Workspace {
        DeviceBuffer sims, mots, bsims, tiles, work[NLIST], agent_sim, kglob, slot, tmp, sorted, pos_of, recs, pos32, xoff, xslab, counts,
            offsets, partial, rpart, rmax;
        DeviceBuffer pcounts, poffsets, pkv, pslot, ptmp, psorted, pflag, pgsim, pgvox, pgbegin, pcoef, phdr, ppartial;
        // agent mirror (mixed form without motility): two sets, forces by record
        DeviceBuffer mpos[2], mprev[2], mprop[2], midx[2], mkey[2], mslot[2], mxslab[2], mrpos[2], mkstep[2], mv, mtmp64;
        DeviceBuffer mmark, mflag;          // 8-byte mark entries of the records (padded by MARK32_STEP) and their flag
        uint32_t ntot = 0, nbins = 0, vtot = 0, ntiles = 0, nwork[NLIST] = {}, ns_max = 0, nmax_all = 0;
        uint32_t max_line[NLIST] = {};
        size_t bytes = 0;
    };

    // the batched sweeps hold up to 32 values per segment with 4 segments
    // (y, z, short x lines) and 16 with 16 segments (long x lines); every
    // segment needs at least one unknown
    static bool grid_fits(const hc_grid& g) {
        return g.nx >= 6 && g.nx <= 258 && g.ny >= 6 && g.ny <= 130 && (g.nz == 1 || (g.nz >= 6 && g.nz <= 130));
    }

    static int substeps_until(double dtd, double dt) {
        double acc = 0.0;
        for (int k = 1; k < 1000000; k++) {
            acc += dtd;
            if (acc >= dt - 0.5 * dtd) return k;
        }
        return 1000000;
    }

    // Members step in lockstep periods of k diffusion substeps + one mechanics
    // step, so they share dt_diffusion and dt_mechanics; each member's
    // phenotype cadence (its voxel refresh + secretion re-plan) must fall on a
    // period end, which is where the single-simulation loop runs it too.
    void check_schedule(const SimConfig& c) const {
        if (!members_.empty()) {
            const SimConfig& f = members_.front()->cfg;
            if (f.dt_diffusion != c.dt_diffusion || f.dt_mechanics != c.dt_mechanics)
                throw std::runtime_error("ensemble: members must share dt_diffusion and dt_mechanics");
        }
        const int k = substeps_until(c.dt_diffusion, c.dt_mechanics);
        const int kp = substeps_until(c.dt_diffusion, c.dt_phenotype);
        if (kp % k) throw std::runtime_error("ensemble: dt_phenotype must be a whole number of mechanics periods");
    }

    void drop_graph() {
        if (graph_) {
            (void)hipGraphExecDestroy((hipGraphExec_t)graph_);
            graph_ = nullptr;
        }
    }

    void rebuild() {
        if (!dirty_) return;
        sync_soa();                                  // (every mutator has synced already)
        drop_graph();
        check(ctx_, hc_wait(ctx_), "wait");        // queued work may still read the old workspace
        Workspace& w = ws_;
        std::vector<EnsSim> sims;
        std::vector<BatchSim> bsims;
        std::vector<AgentTile> tiles;
        std::vector<BatchWork> work[NLIST], narrow[NLIST];
        std::vector<uint32_t> agent_sim;
        std::vector<hc::MotPtrs> mots;
        any_motility_ = false;
        uint32_t cell_off = 0, bin_off = 0, vox_off = 0, ns_max = 0, nmax_all = 0;
        uint32_t max_line[NLIST] = {};
        for (uint32_t si = 0; si < members_.size(); si++) {
            Member& m = *members_[si];
            const hc_grid& g = m.grid;
            EnsSim e{};
            e.soa = (double*)hc_device_ptr(m.soa.get());
            e.rates = (const double*)hc_device_ptr(m.rates.get());
            e.rho = hc_device_ptr(m.density.get());
            e.tables = (const uint8_t*)hc_device_ptr(m.tables.get());
            e.max_cells = m.max_cells; e.n = m.n; e.ns = g.n_substrates; e.nmax = std::max({g.nx, g.ny, g.nz});
            e.nx = g.nx; e.ny = g.ny; e.nz = g.nz; e.V = g.nx * g.ny * g.nz;
            e.x_min = g.x_min; e.y_min = g.y_min; e.z_min = g.z_min;
            e.dx = g.dx; e.dy = g.dy; e.dz = g.dz; e.dt_diff = g.dt;
            e.mvoxel = m.mech.mechanics_voxel_size; e.max_dist = m.mech.max_interaction_distance; e.dt_mech = m.mech.dt;
            e.mnx = m.mech.mech_grid_nx; e.mny = m.mech.mech_grid_ny; e.mnz = m.mech.mech_grid_nz;
            e.cell_off = cell_off; e.bin_off = bin_off; e.vox_off = vox_off;
            hc::MotPtrs mp{};
            if (m.mot.get()) {
                mp = hc::mot_ptrs(hc_device_ptr(m.mot.get()), m.max_cells, g.n_substrates);
                e.mot = 1u | (m.mot_sens ? 2u : 0u);
                any_motility_ = true;
            }
            mots.push_back(mp);
            sims.push_back(e);
            BatchSim b{};
            b.rho = e.rho;
            b.tab = e.tables;
            b.nx = g.nx; b.ny = g.ny; b.nz = g.nz; b.ns = g.n_substrates; b.nmax = e.nmax;
            b.last_axis = g.nz > 1 ? 2u : 1u;
            b.xmap = (uint32_t)hc::ctx_x_lane_order(ctx_, g.nx, hc::part_p(0, g.nx, mixed_), mixed_ ? 4 : 8);
            b.wide = 0;
            for (uint32_t c0 = 0; c0 < m.n; c0 += 256) tiles.push_back(AgentTile{si, c0});
            for (uint32_t c = 0; c < m.n; c++) agent_sim.push_back(si);
            const uint32_t nn[3] = {g.nx, g.ny, g.nz};
            for (int ax = 0; ax < 3; ax++) {
                if (ax == 2 && g.nz == 1) continue;
                const int P = hc::part_p(ax, nn[ax], mixed_);
                const int li = list_of(ax, P);
                const int mm = hc::part_window(nn[ax], P);
                if (li < 0 || mm == 0 || mm > 32)
                    throw std::runtime_error("ensemble: a member's line length is outside the batched sweep's "
                                             "register windows (FP64: at most 130 along y and z)");
                if (mm == 32) b.wide |= 1u << ax;
                const uint64_t lines = hc::sweep_lines(ax, g.nx, g.ny, g.nz);
                const uint64_t groups = (lines + 4 * (64 / P) - 1) / (4 * (64 / P));
                std::vector<BatchWork>& dst = mm == 32 ? work[li] : narrow[li];
                for (uint32_t s = 0; s < g.n_substrates; s++)
                    for (uint64_t q = 0; q < groups; q++) dst.push_back(BatchWork{si, s, q});
                max_line[li] = std::max(max_line[li], nn[ax]);
            }
            bsims.push_back(b);
            cell_off += m.n;
            bin_off += m.mech.mech_grid_nx * m.mech.mech_grid_ny * m.mech.mech_grid_nz;
            vox_off += e.V;
            ns_max = std::max(ns_max, g.n_substrates);
            nmax_all = std::max(nmax_all, e.nmax);
        }
        for (int li = 0; li < NLIST; li++) work[li].insert(work[li].end(), narrow[li].begin(), narrow[li].end());
        w.ntot = cell_off; w.nbins = bin_off; w.vtot = vox_off; w.ns_max = ns_max; w.nmax_all = nmax_all;
        w.ntiles = (uint32_t)tiles.size();
        for (int li = 0; li < NLIST; li++) { w.nwork[li] = (uint32_t)work[li].size(); w.max_line[li] = max_line[li]; }
        size_t bytes = 0;
        auto put = [&](DeviceBuffer& b, const void* src, size_t nb) {
            b.allocate(ctx_, std::max<size_t>(16, nb));
            if (nb) {
                std::memcpy(b.host(), src, nb);
                b.upload(0, nb);
            }
            bytes += b.size();
        };
        auto alloc = [&](DeviceBuffer& b, size_t nb) {
            b.allocate(ctx_, std::max<size_t>(16, nb));
            bytes += b.size();
        };
        put(w.sims, sims.data(), sims.size() * sizeof(EnsSim));
        put(w.mots, mots.data(), mots.size() * sizeof(hc::MotPtrs));
        put(w.bsims, bsims.data(), bsims.size() * sizeof(BatchSim));
        put(w.tiles, tiles.data(), tiles.size() * sizeof(AgentTile));
        for (int li = 0; li < NLIST; li++) put(w.work[li], work[li].data(), work[li].size() * sizeof(BatchWork));
        put(w.agent_sim, agent_sim.data(), agent_sim.size() * sizeof(uint32_t));
        const size_t N = std::max<uint32_t>(1, w.ntot);
        alloc(w.kglob, N * 4); alloc(w.slot, N * 4); alloc(w.tmp, N * 4); alloc(w.sorted, N * 4);
        alloc(w.pos_of, N * 4);
        // records and pos32 carry one mark step of padding: the force kernels'
        // mark passes load whole steps past the end of a range
        alloc(w.recs, (N + hc::MARK32_STEP) * (mixed_ ? sizeof(hc::CellRec32) : sizeof(hc::CellRec)));
        if (!mixed_) alloc(w.pos32, (N + hc::MARKX_STEP) * sizeof(float4));
        if (mixed_) {
            alloc(w.xoff, ((size_t)w.nbins + 1) * hc::XSLABS * 4);
            alloc(w.xslab, N);
        }
        if (mirror_on()) {
            for (int k = 0; k < 2; k++) {
                alloc(w.mpos[k], N * sizeof(hc::Dbl3)); alloc(w.mprev[k], N * sizeof(hc::Dbl3));
                alloc(w.mprop[k], N * 16);
                alloc(w.midx[k], N * 4); alloc(w.mkey[k], N * 4); alloc(w.mslot[k], N * 4);
                alloc(w.mxslab[k], N); alloc(w.mrpos[k], N * 4); alloc(w.mkstep[k], N * 4);
            }
            alloc(w.mv, N * 16);
            alloc(w.mtmp64, N * 8);
            alloc(w.mmark, (N + hc::MARK32_STEP) * 8);
            alloc(w.mflag, 256);
        }
        mirror_valid_ = false;
        soa_stale_ = false;
        alloc(w.rpart, (size_t)std::max(1u, w.ntiles) * 4);
        alloc(w.rmax, 256);
        alloc(w.counts, ((size_t)w.nbins + 1) * 4); alloc(w.offsets, ((size_t)w.nbins + 2) * 4);
        alloc(w.partial, hc::scan_partial_bytes((uint64_t)w.nbins + 1));
        alloc(w.pcounts, ((size_t)w.vtot + 2) * 4); alloc(w.poffsets, ((size_t)w.vtot + 3) * 4);
        alloc(w.pkv, N * 4); alloc(w.pslot, N * 4); alloc(w.ptmp, N * 4); alloc(w.psorted, N * 4);
        alloc(w.pflag, (N + 1) * 4); alloc(w.pgsim, (N + 1) * 4); alloc(w.pgvox, (N + 1) * 4);
        alloc(w.pgbegin, (N + 2) * 4); alloc(w.pcoef, (size_t)3 * std::max(1u, w.ns_max) * N * 8);
        alloc(w.phdr, sizeof(hc::PlanHeader));
        alloc(w.ppartial, std::max(hc::scan_partial_bytes((uint64_t)w.vtot + 1), hc::scan_partial_bytes(N)));
        w.bytes = bytes;
        check(ctx_, hc_wait(ctx_), "wait");
        dirty_ = false;
        plan_dirty_ = true;
    }

    hc_ctx* c() { return ctx_; }

    // Moves a member's cell SoA, rate arrays and motility state to the layout
    // of a larger capacity (every field keeps its first n values).
    void grow(Member& m, uint32_t new_max) {
        const uint32_t old_max = m.max_cells, n = m.n, ns = m.grid.n_substrates;
        hipStream_t st = hc_stream_of(ctx_);
        auto copy = [&](DeviceBuffer& to, size_t to_off, DeviceBuffer& from, size_t from_off, size_t nb) {
            if (!nb) return;
            hipError_t e = hipMemcpyAsync((char*)hc_device_ptr(to.get()) + to_off,
                                          (const char*)hc_device_ptr(from.get()) + from_off, nb,
                                          hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) throw std::runtime_error(std::string("ensemble grow: ") + hipGetErrorString(e));
        };
        DeviceBuffer soa, rates, mot;
        soa.allocate(ctx_, hc_cells_bytes(new_max));
        for (uint32_t f = 0; f < HC_NUM_CELL_FIELDS; f++)
            copy(soa, hc_cells_field_offset(f, new_max), m.soa, hc_cells_field_offset(f, old_max),
                 (size_t)n * (f < HC_NUM_F64_FIELDS ? 8 : 4));
        rates.allocate(ctx_, std::max<size_t>(8, hc_rates_bytes(new_max, ns)));
        for (uint32_t q = 0; q < 4; q++)
            copy(rates, (size_t)q * new_max * ns * 8, m.rates, (size_t)q * old_max * ns * 8, (size_t)n * ns * 8);
        if (m.mot.get()) {
            mot.allocate(ctx_, hc_motility_bytes(new_max, ns));
            copy(mot, 0, m.mot, 0, 256);                       // step counter and seed
            for (uint32_t f = 0; f < HC_M_SENSITIVITY; f++)
                copy(mot, hc_motility_field_offset(f, new_max, ns), m.mot, hc_motility_field_offset(f, old_max, ns),
                     (size_t)n * (f <= HC_M_VEC_Z ? 8 : 4));
            copy(mot, hc_motility_field_offset(HC_M_SENSITIVITY, new_max, ns), m.mot,
                 hc_motility_field_offset(HC_M_SENSITIVITY, old_max, ns), (size_t)n * ns * 8);
        }
        check(ctx_, hc_wait(ctx_), "wait");
        m.soa = std::move(soa);
        m.rates = std::move(rates);
        if (mot.get()) m.mot = std::move(mot);
        // host mirrors of the new buffers: the cell SoA's, which set_motility
        // writes single fields through
        m.soa.download_all();
        check(ctx_, hc_wait(ctx_), "wait");
        m.max_cells = new_max;
        m.vol = DeviceBuffer();
        m.mask = DeviceBuffer();
        dirty_ = true;
        plan_dirty_ = true;
    }

    // x, y, [z] over every member; each member re-applies Dirichlet in its
    // own last sweep (z in 3-D, y in 2-D)
    void sweep_list(int li, unsigned blocks) {
        Workspace& w = ws_;
        check(ctx_, hc::launch_sweep_batch(ctx_, LIST_AXIS[li], LIST_P[li], LIST_MM[li], mixed_,
                                           (const BatchSim*)hc_device_ptr(w.bsims.get()),
                                           (const BatchWork*)hc_device_ptr(w.work[li].get()), w.nwork[li], w.nmax_all,
                                           w.max_line[li], blocks),
              "ensemble sweep");
    }
    void sweeps() {
        for (int li = 0; li < NLIST; li++) sweep_list(li, grid_blocks_[li]);
    }

    void plan_secretion() {
        Workspace& w = ws_;
        plan_dirty_ = false;
        if (!w.ntot) return;
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        uint32_t* counts = (uint32_t*)hc_device_ptr(w.pcounts.get());
        uint32_t* offsets = (uint32_t*)hc_device_ptr(w.poffsets.get());
        uint32_t* kv = (uint32_t*)hc_device_ptr(w.pkv.get());
        uint32_t* slot = (uint32_t*)hc_device_ptr(w.pslot.get());
        uint32_t* tmp = (uint32_t*)hc_device_ptr(w.ptmp.get());
        uint32_t* sorted = (uint32_t*)hc_device_ptr(w.psorted.get());
        uint32_t* flag = (uint32_t*)hc_device_ptr(w.pflag.get());
        uint32_t* partial = (uint32_t*)hc_device_ptr(w.ppartial.get());
        hc::PlanHeader* hdr = (hc::PlanHeader*)hc_device_ptr(w.phdr.get());
        hipStream_t st = hc_stream_of(ctx_);
        check(ctx_, hc::fill_u32(ctx_, counts, (uint64_t)w.vtot + 1, 0), "fill");
        check(ctx_, hc::fill_u32(ctx_, (uint32_t*)hdr, sizeof(hc::PlanHeader) / 4, 0), "fill");
        if (w.ntiles)
            hipLaunchKernelGGL(hc::k_b_sec_keys, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, counts, kv, slot, w.vtot);
        check(ctx_, hc::check_launch(ctx_, "k_b_sec_keys"), "sec keys");
        check(ctx_, hc::scan_exclusive(ctx_, counts, offsets, (uint64_t)w.vtot + 1, partial), "scan");
        check(ctx_, hc::launch_scatter(ctx_, kv, slot, offsets, tmp, w.ntot), "scatter");
        check(ctx_, hc::launch_stable_sort(ctx_, kv, offsets, tmp, sorted, w.ntot), "stable");
        check(ctx_, hc::launch_sec_flags(ctx_, kv, sorted, offsets, flag, w.ntot, w.vtot), "flags");
        check(ctx_, hc::scan_exclusive(ctx_, flag, flag, w.ntot, partial), "scan flags");
        const dim3 gg(hc::blocks_for(std::max(1u, w.ntot), 256));
        if (mixed_)
            hipLaunchKernelGGL(hc::k_b_sec_groups<float>, gg, dim3(256), 0, st, sims,
                               (const uint32_t*)hc_device_ptr(w.agent_sim.get()), kv, sorted, offsets, flag,
                               (uint32_t*)hc_device_ptr(w.pgsim.get()), (uint32_t*)hc_device_ptr(w.pgvox.get()),
                               (uint32_t*)hc_device_ptr(w.pgbegin.get()), (float*)hc_device_ptr(w.pcoef.get()), hdr,
                               w.ntot, w.vtot, w.ns_max);
        else
            hipLaunchKernelGGL(hc::k_b_sec_groups<double>, gg, dim3(256), 0, st, sims,
                               (const uint32_t*)hc_device_ptr(w.agent_sim.get()), kv, sorted, offsets, flag,
                               (uint32_t*)hc_device_ptr(w.pgsim.get()), (uint32_t*)hc_device_ptr(w.pgvox.get()),
                               (uint32_t*)hc_device_ptr(w.pgbegin.get()), (double*)hc_device_ptr(w.pcoef.get()), hdr,
                               w.ntot, w.vtot, w.ns_max);
        check(ctx_, hc::check_launch(ctx_, "k_b_sec_groups"), "sec groups");
        plan_dirty_ = false;
    }

    void secretion() {
        Workspace& w = ws_;
        if (!w.ntot) return;
        hipStream_t st = hc_stream_of(ctx_);
        const dim3 grid(hc::blocks_for(w.ntot, 256));
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        hc::KernelTimer tm(ctx_, HC_K_SEC_APPLY);
        if (mixed_)
            hipLaunchKernelGGL(hc::k_b_sec_apply<float>, grid, dim3(256), 0, st, sims,
                               (const uint32_t*)hc_device_ptr(w.pgsim.get()), (const uint32_t*)hc_device_ptr(w.pgvox.get()),
                               (const uint32_t*)hc_device_ptr(w.pgbegin.get()), (const float*)hc_device_ptr(w.pcoef.get()),
                               (const hc::PlanHeader*)hc_device_ptr(w.phdr.get()), w.ntot);
        else
            hipLaunchKernelGGL(hc::k_b_sec_apply<double>, grid, dim3(256), 0, st, sims,
                               (const uint32_t*)hc_device_ptr(w.pgsim.get()), (const uint32_t*)hc_device_ptr(w.pgvox.get()),
                               (const uint32_t*)hc_device_ptr(w.pgbegin.get()), (const double*)hc_device_ptr(w.pcoef.get()),
                               (const hc::PlanHeader*)hc_device_ptr(w.phdr.get()), w.ntot);
        check(ctx_, hc::check_launch(ctx_, "k_b_sec_apply"), "secretion");
    }

    // The agent mirror runs the mixed form's mechanics when no member is
    // motile (motility rewrites the velocities between forces and AB2) and
    // the records take the mark pass with a list.
    bool mirror_on() const {
        return mixed_ && !any_motility_ && ws_.ntot > 0 && ws_.ntot < (1u << hc::MARK32_POS_BITS) &&
               ws_.nbins < (1u << 30);
    }

    hc::BMirror mirror_set(int k) {
        Workspace& w = ws_;
        hc::BMirror m;
        m.pos = (hc::Dbl3*)hc_device_ptr(w.mpos[k].get());
        m.prev = (hc::Dbl3*)hc_device_ptr(w.mprev[k].get());
        m.prop = (float4*)hc_device_ptr(w.mprop[k].get());
        m.idx = (uint32_t*)hc_device_ptr(w.midx[k].get());
        m.key = (uint32_t*)hc_device_ptr(w.mkey[k].get());
        m.slot = (uint32_t*)hc_device_ptr(w.mslot[k].get());
        m.xslab = (uint8_t*)hc_device_ptr(w.mxslab[k].get());
        m.rpos = (uint32_t*)hc_device_ptr(w.mrpos[k].get());
        m.kstep = (uint32_t*)hc_device_ptr(w.mkstep[k].get());
        m.v = (float4*)hc_device_ptr(w.mv.get());
        if (!m.pos || !m.prev || !m.prop || !m.idx || !m.key || !m.slot || !m.xslab || !m.rpos || !m.kstep || !m.v)
            throw std::runtime_error("ensemble: agent mirror not allocated");
        return m;
    }

    // hash of the mirror's coming step: lists, record positions, slab starts
    // and the FP32 records (counts hold the histogram of the mirror's bins)
    // the hash of the slots' coming bins.  at_load: also the largest radius
    // (the slots' radii change only by a reload), and the counts stay for
    // the first step's hash; else the scan clears them for AB2's histogram.
    void mirror_hash(const hc::BMirror& a, bool at_load) {
        Workspace& w = ws_;
        hipStream_t st = hc_stream_of(ctx_);
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        uint32_t* counts = (uint32_t*)hc_device_ptr(w.counts.get());
        uint32_t* offsets = (uint32_t*)hc_device_ptr(w.offsets.get());
        uint64_t* tmp64 = (uint64_t*)hc_device_ptr(w.mtmp64.get());
        uint32_t* sorted = (uint32_t*)hc_device_ptr(w.sorted.get());
        uint32_t* xoff = (uint32_t*)hc_device_ptr(w.xoff.get());
        float* rpart = (float*)hc_device_ptr(w.rpart.get());
        if (!tmp64 || !hc_device_ptr(w.mmark.get()) || !hc_device_ptr(w.mflag.get()))
            throw std::runtime_error("ensemble: agent mirror not allocated");
        check(ctx_, hc::scan_exclusive(ctx_, counts, offsets, w.nbins, (uint32_t*)hc_device_ptr(w.partial.get()),
                                       at_load ? nullptr : counts),
              "scan");
        hipLaunchKernelGGL(hc::k_bm_scatter, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, offsets, tmp64);
        uint32_t* mflag = (uint32_t*)hc_device_ptr(w.mflag.get());
        hipLaunchKernelGGL(hc::k_bm_bin_sort, dim3(hc::blocks_for(w.nbins + 1, 256)), dim3(256), 0, st, offsets,
                           tmp64, a, sorted, xoff, w.nbins, mflag);
        hipLaunchKernelGGL(hc::k_bm_gather, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, offsets, tmp64, sorted,
                           (hc::CellRec32*)hc_device_ptr(w.recs.get()), at_load ? rpart : nullptr,
                           (uint2*)hc_device_ptr(w.mmark.get()), mflag);
        check(ctx_, hc::check_launch(ctx_, "mirror hash"), "mirror hash");
        if (at_load)
            check(ctx_, hc::launch_rmax(ctx_, rpart, w.ntiles, (float*)hc_device_ptr(w.rmax.get())), "largest radius");
    }

    // every slot's state to its record position of the last step
    void mirror_resort() {
        Workspace& w = ws_;
        hipStream_t st = hc_stream_of(ctx_);
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        const hc::BMirror a = mirror_set(0), b = mirror_set(1);
        hipLaunchKernelGGL(hc::k_bm_permute, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, b);
        hipLaunchKernelGGL(hc::k_bm_copy, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, b);
        check(ctx_, hc::check_launch(ctx_, "mirror re-sort"), "mirror re-sort");
        steps_since_sort_ = 0;
    }

    // the mirror from the cell fields, in the record order of the current
    // positions; bins and slots for the coming step in counts
    void mirror_load() {
        Workspace& w = ws_;
        hipStream_t st = hc_stream_of(ctx_);
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        uint32_t* counts = (uint32_t*)hc_device_ptr(w.counts.get());
        const hc::BMirror a = mirror_set(0);
        check(ctx_, hc::fill_u32(ctx_, counts, w.nbins, 0), "fill");
        hipLaunchKernelGGL(hc::k_bm_load, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, counts);
        check(ctx_, hc::check_launch(ctx_, "k_bm_load"), "mirror load");
        mirror_hash(a, true);
        mirror_resort();
        mirror_valid_ = true;
        soa_stale_ = false;
    }

    // the cell fields from the mirror, when a step has run since they were
    // last written
    void sync_soa() {
        if (!soa_stale_) return;
        soa_stale_ = false;
        Workspace& w = ws_;
        hipLaunchKernelGGL(hc::k_bm_sync, dim3(w.ntiles), dim3(256), 0, hc_stream_of(ctx_),
                           (const EnsSim*)hc_device_ptr(w.sims.get()), (const AgentTile*)hc_device_ptr(w.tiles.get()),
                           mirror_set(0));
        check(ctx_, hc::check_launch(ctx_, "k_bm_sync"), "mirror write-back");
    }

    // one mechanics step on the mirror (bins and slots from the last AB2)
    void mechanics_mirror() {
        Workspace& w = ws_;
        hipStream_t st = hc_stream_of(ctx_);
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        uint32_t* counts = (uint32_t*)hc_device_ptr(w.counts.get());
        const hc::BMirror a = mirror_set(0);
        {
            hc::KernelTimer tm(ctx_, HC_K_HASH);
            mirror_hash(a, false);
        }
        {
            hc::KernelTimer tm(ctx_, HC_K_FORCES);
            hipLaunchKernelGGL((hc::k_b_forces<true, true, true>), dim3(w.ntiles), dim3(256), 0, st, sims, tiles,
                               hc_device_ptr(w.recs.get()), nullptr, (const uint32_t*)hc_device_ptr(w.offsets.get()),
                               (const uint32_t*)hc_device_ptr(w.xoff.get()), a.key,
                               (const float*)hc_device_ptr(w.rmax.get()), a.prop, a.v,
                               (const uint2*)hc_device_ptr(w.mmark.get()), (const uint32_t*)hc_device_ptr(w.mflag.get()));
            check(ctx_, hc::check_launch(ctx_, "k_b_forces (mirror)"), "forces");
        }
        {
            hc::KernelTimer tm(ctx_, HC_K_INTEGRATE);
            hipLaunchKernelGGL(hc::k_bm_integrate, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, a, counts);
            check(ctx_, hc::check_launch(ctx_, "k_bm_integrate"), "integrate");
        }
    }

    void mechanics() {
        Workspace& w = ws_;
        if (!w.ntot) return;
        if (mirror_on()) {
            mechanics_mirror();
            return;
        }
        hipStream_t st = hc_stream_of(ctx_);
        const EnsSim* sims = (const EnsSim*)hc_device_ptr(w.sims.get());
        const AgentTile* tiles = (const AgentTile*)hc_device_ptr(w.tiles.get());
        uint32_t* counts = (uint32_t*)hc_device_ptr(w.counts.get());
        uint32_t* offsets = (uint32_t*)hc_device_ptr(w.offsets.get());
        uint32_t* kglob = (uint32_t*)hc_device_ptr(w.kglob.get());
        uint32_t* slot = (uint32_t*)hc_device_ptr(w.slot.get());
        uint32_t* tmp = (uint32_t*)hc_device_ptr(w.tmp.get());
        uint32_t* sorted = (uint32_t*)hc_device_ptr(w.sorted.get());
        uint32_t* pos_of = (uint32_t*)hc_device_ptr(w.pos_of.get());
        const uint32_t* asim = (const uint32_t*)hc_device_ptr(w.agent_sim.get());
        void* recs = hc_device_ptr(w.recs.get());
        float4* pos32 = mixed_ ? nullptr : (float4*)hc_device_ptr(w.pos32.get());
        uint32_t* xoff = mixed_ ? (uint32_t*)hc_device_ptr(w.xoff.get()) : nullptr;
        uint8_t* xslab = mixed_ ? (uint8_t*)hc_device_ptr(w.xslab.get()) : nullptr;
        if (!sims || !tiles || !counts || !offsets || !kglob || !slot || !tmp || !sorted || !pos_of || !asim ||
            !recs || (!mixed_ && !pos32) || (mixed_ && (!xoff || !xslab)) || !hc_device_ptr(w.partial.get()))
            throw std::runtime_error("ensemble: mechanics workspace not allocated");
        {
            hc::KernelTimer tm(ctx_, HC_K_HASH);
            check(ctx_, hc::fill_u32(ctx_, counts, w.nbins, 0), "fill");
            hipLaunchKernelGGL(hc::k_b_mech_keys, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, counts, kglob, slot,
                               xslab);
            check(ctx_, hc::check_launch(ctx_, "k_b_mech_keys"), "keys");
            check(ctx_, hc::scan_exclusive(ctx_, counts, offsets, w.nbins, (uint32_t*)hc_device_ptr(w.partial.get())), "scan");
            check(ctx_, hc::launch_scatter(ctx_, kglob, slot, offsets, tmp, w.ntot), "scatter");
            if (mixed_)
                check(ctx_, hc::launch_bin_sort32(ctx_, offsets, tmp, xslab, sorted, pos_of, xoff, w.nbins), "bin sort");
            else
                check(ctx_, hc::launch_bin_sort(ctx_, offsets, tmp, sorted, pos_of, w.nbins), "bin sort");
            float* rpart = (float*)hc_device_ptr(w.rpart.get());
            if (mixed_)
                hipLaunchKernelGGL(hc::k_b_gather<true>, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, kglob, offsets, tmp,
                                   sorted, pos_of, xslab, recs, nullptr, rpart);
            else
                hipLaunchKernelGGL(hc::k_b_gather<false>, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, kglob, offsets,
                                   tmp, sorted, pos_of, nullptr, recs, pos32, rpart);
            check(ctx_, hc::check_launch(ctx_, "k_b_gather"), "gather");
            check(ctx_, hc::launch_rmax(ctx_, rpart, w.ntiles, (float*)hc_device_ptr(w.rmax.get())), "largest radius");
        }
        {
            hc::KernelTimer tm(ctx_, HC_K_FORCES);
            const dim3 g(w.ntiles);
            // positions in the records' list entries take 27 bits and slab
            // indices 32 (force_agent32x)
            const float* rmax = (const float*)hc_device_ptr(w.rmax.get());
            const bool list = w.ntot < (1u << hc::MARK32_POS_BITS) && w.nbins < (1u << 30);
            if (mixed_)
                hipLaunchKernelGGL((list ? hc::k_b_forces<true, true> : hc::k_b_forces<true, false>), g, dim3(256), 0,
                                   st, sims, tiles, recs, nullptr, offsets, xoff, kglob, rmax, nullptr, nullptr,
                                   nullptr, nullptr);
            else
                hipLaunchKernelGGL(hc::k_b_forces<false>, g, dim3(256), 0, st, sims, tiles, recs, pos32, offsets,
                                   nullptr, nullptr, rmax, nullptr, nullptr, nullptr, nullptr);
            check(ctx_, hc::check_launch(ctx_, "k_b_forces"), "forces");
        }
        if (any_motility_) {
            const hc::MotPtrs* mots = (const hc::MotPtrs*)hc_device_ptr(w.mots.get());
            hc::KernelTimer tm(ctx_, HC_K_MOTILITY);
            if (mixed_)
                hipLaunchKernelGGL(hc::k_b_motility<float>, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, mots);
            else
                hipLaunchKernelGGL(hc::k_b_motility<double>, dim3(w.ntiles), dim3(256), 0, st, sims, tiles, mots);
            const uint32_t nsims = (uint32_t)members_.size();
            hipLaunchKernelGGL(hc::k_b_motility_advance, dim3(hc::blocks_for(nsims, 256)), dim3(256), 0, st, sims, mots,
                               nsims);
            check(ctx_, hc::check_launch(ctx_, "k_b_motility"), "motility");
        }
        {
            hc::KernelTimer tm(ctx_, HC_K_INTEGRATE);
            hipLaunchKernelGGL(hc::k_b_integrate, dim3(w.ntiles), dim3(256), 0, st, sims, tiles);
            check(ctx_, hc::check_launch(ctx_, "k_b_integrate"), "integrate");
        }
    }

    void release_streams() {
        if (join_) (void)hipEventDestroy(join_);
        if (fork_) (void)hipEventDestroy(fork_);
        if (side_) (void)hipStreamDestroy(side_);
        join_ = fork_ = nullptr;
        side_ = nullptr;
    }

    static void hip_ok(hipError_t e, const char* what) {
        if (e != hipSuccess) throw std::runtime_error(std::string("ensemble ") + what + ": " + hipGetErrorString(e));
    }

    // One period.  Within a period the diffusion substeps read and write the
    // fields and read the secretion plan, and the mechanics step reads and
    // writes positions, velocities, pressure and the hash: disjoint data, so
    // with overlap on the mechanics step runs on a second stream while the
    // substeps run, and the period ends when both are done.  The results are
    // the same either way.
    void period_kernels() {
        if (plan_dirty_) plan_secretion();
        const int k = substeps_per_mech();
        if (!overlap_ || any_motility_) {            // motility reads the substeps' fields
            for (int s = 0; s < k; s++) {
                sweeps();
                secretion();
            }
            mechanics();
            return;
        }
        hipStream_t main = hc_stream_of(ctx_);
        hip_ok(hipEventRecord(fork_, main), "fork");
        hip_ok(hipStreamWaitEvent(side_, fork_, 0), "fork wait");
        hc_stream_swap(ctx_, side_);
        try {
            mechanics();
        } catch (...) {
            hc_stream_swap(ctx_, main);
            throw;
        }
        hc_stream_swap(ctx_, main);
        hip_ok(hipEventRecord(join_, side_), "join");
        for (int s = 0; s < k; s++) {
            sweeps();
            secretion();
        }
        hip_ok(hipStreamWaitEvent(main, join_, 0), "join wait");
    }

    void capture() {
        if (plan_dirty_) plan_secretion();
        hipStream_t st = hc_stream_of(ctx_);
        check(ctx_, hc_wait(ctx_), "wait");
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) throw std::runtime_error(std::string("ensemble capture: ") + hipGetErrorString(e));
        try {
            period_kernels();
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);    // already failing: end the capture, rethrow
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) throw std::runtime_error(std::string("ensemble end capture: ") + hipGetErrorString(e));
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw std::runtime_error(std::string("ensemble instantiate: ") + hipGetErrorString(e));
        graph_ = (void*)exec;
    }

    // member clocks: the accumulators of src/main.mm:375-412, per member
    void advance_clocks() {
        const int k = substeps_per_mech();
        bool refreshed = false;
        for (auto& mp : members_) {
            Member& m = *mp;
            const double dtd = m.cfg.dt_diffusion;
            for (int s = 0; s < k; s++) {
                m.substeps++;
                m.mech_acc += dtd;
                if (m.mech_acc >= m.cfg.dt_mechanics - 0.5 * dtd) {
                    m.mech_steps++;
                    m.mech_acc = 0.0;
                }
                m.pheno_acc += dtd;
                if (m.pheno_acc >= m.cfg.dt_phenotype - 0.5 * dtd) {
                    m.pheno_steps++;
                    m.pheno_acc = 0.0;
                    sync_soa();
                    check(ctx_, hc_voxel_keys(ctx_, m.soa.get(), &m.grid, m.n, m.max_cells), "voxel refresh");
                    refreshed = true;
                }
            }
        }
        if (refreshed) {
            // the plan is rebuilt outside any graph; the graph reads the plan
            // buffers, which keep their addresses
            plan_secretion();
        }
    }

    hc_ctx* ctx_ = nullptr;
    hipStream_t side_ = nullptr;
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    bool overlap_ = false;
    bool mixed_ = false, dirty_ = true, plan_dirty_ = true, any_motility_ = false;
    // agent mirror state: loaded, cell fields behind it, steps since its
    // slots were last put in record order (re-sorted every
    // MIRROR_RESORT_STEPS mechanics steps)
    static constexpr int MIRROR_RESORT_STEPS = 64;
    bool mirror_valid_ = false, soa_stale_ = false;
    int steps_since_sort_ = 0;
    uint32_t next_id_ = 0;
    unsigned grid_blocks_[NLIST] = {};                 // 0: one workgroup per work item
    void* graph_ = nullptr;
    std::vector<std::unique_ptr<Member>> members_;
    Workspace ws_;
};

}  // namespace physicell

// ---------------------------------------------------------------- pe_* C-ABI
using physicell::Ensemble;

struct pe_ensemble {
    std::unique_ptr<Ensemble> e;
    std::string err;
};

template <class F>
static int eguard(pe_ensemble* p, F&& f) {
    if (!p || !p->e) return HC_ERR_ARG;
    try {
        f(*p->e);
        return 0;
    } catch (const std::exception& ex) {
        p->err = ex.what();
        return HC_ERR_STATE;
    }
}

static physicell::SimConfig to_simconfig(const pc_config* cfg) {
    physicell::SimConfig c;
    c.x_min = cfg->x_min; c.x_max = cfg->x_max; c.y_min = cfg->y_min; c.y_max = cfg->y_max;
    c.z_min = cfg->z_min; c.z_max = cfg->z_max;
    c.dx = cfg->dx; c.dy = cfg->dy; c.dz = cfg->dz;
    c.use_2D = cfg->use_2D != 0;
    c.dt_diffusion = cfg->dt_diffusion; c.dt_mechanics = cfg->dt_mechanics;
    c.dt_phenotype = cfg->dt_phenotype; c.max_time = cfg->max_time;
    c.max_cells = cfg->max_cells;
    for (uint32_t s = 0; s < cfg->n_substrates; s++) {
        physicell::SubstrateConfig sc;
        sc.diffusion_coeff = cfg->substrates[s].diffusion_coeff;
        sc.decay_rate = cfg->substrates[s].decay_rate;
        sc.initial_value = cfg->substrates[s].initial_value;
        sc.dirichlet_value = cfg->substrates[s].dirichlet_value;
        sc.dirichlet_enabled = cfg->substrates[s].dirichlet_enabled != 0;
        c.substrates.push_back(sc);
    }
    return c;
}

extern "C" {

int pe_create(pe_ensemble** out, int device, uint32_t mixed_precision) {
    if (!out) return HC_ERR_ARG;
    *out = nullptr;
    auto* p = new pe_ensemble();
    try {
        p->e.reset(new Ensemble(device, mixed_precision != 0));
    } catch (const std::exception& ex) {
        fprintf(stderr, "[pe_create] %s\n", ex.what());
        delete p;
        return HC_ERR_HIP;
    }
    *out = p;
    return 0;
}

void pe_destroy(pe_ensemble* p) { delete p; }
const char* pe_last_error(const pe_ensemble* p) { return p ? p->err.c_str() : "null ensemble"; }

int pe_add(pe_ensemble* p, const pc_config* cfg, uint32_t n, const double* x, const double* y, const double* z,
           const double* radius, const double* volume, const double* rep, const double* adh, const double* rel,
           const double* sec, const double* upt, const double* sat, const double* net, const uint32_t* alive,
           uint32_t* member_id) {
    if (!cfg || (cfg->n_substrates && !cfg->substrates) || !member_id) return HC_ERR_ARG;
    if (n && (!x || !y || !z || !radius || !volume || !rep || !adh || !rel)) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) {
        *member_id = e.add(to_simconfig(cfg), n, x, y, z, radius, volume, rep, adh, rel, sec, upt, sat, net, alive);
    });
}

int pe_set_motility(pe_ensemble* p, uint32_t member_id, uint64_t seed, const double* persistence, const double* bias,
                    const double* speed, const uint32_t* is_motile, const uint32_t* restrict_2d,
                    const int32_t* chem_index, const int32_t* chem_dir, const double* sensitivity) {
    return eguard(p, [&](Ensemble& e) {
        e.set_motility(member_id, seed, persistence, bias, speed, is_motile, restrict_2d, chem_index, chem_dir,
                       sensitivity);
    });
}

int pe_set_volume(pe_ensemble* p, uint32_t member_id, uint64_t seed, const double* const* fields,
                  uint32_t max_cells) {
    return eguard(p, [&](Ensemble& e) { e.set_volume(member_id, seed, fields, max_cells); });
}

int pe_update_phenotype(pe_ensemble* p, uint32_t member_id, double dt, int use_ratio, const uint32_t* mask,
                        const uint32_t* parents, uint32_t n_parents, const uint32_t* removed, uint32_t n_removed,
                        uint32_t* n_out) {
    if ((n_parents && !parents) || (n_removed && !removed)) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) {
        const uint32_t n = e.update_phenotype(member_id, dt, use_ratio, mask, parents, n_parents, removed, n_removed);
        if (n_out) *n_out = n;
    });
}

int pe_neighbor_query(pe_ensemble* p, uint32_t member_id, int kind, uint32_t* offsets, uint32_t* idx,
                      uint64_t capacity, uint64_t* total) {
    if (!total) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) {
        std::vector<uint32_t> off, lst;
        *total = e.neighbor_query(member_id, kind, off, lst);
        if (offsets) std::memcpy(offsets, off.data(), off.size() * 4);
        if (idx && capacity >= lst.size() && !lst.empty()) std::memcpy(idx, lst.data(), lst.size() * 4);
    });
}

int pe_form_attachments(pe_ensemble* p, uint32_t member_id, double dt, const double* rate, uint32_t* count,
                        uint32_t* targets, uint64_t seed, uint32_t* formed) {
    if (!rate || !count || !targets) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) {
        const uint32_t f = e.form_attachments(member_id, dt, rate, count, targets, seed);
        if (formed) *formed = f;
    });
}

int pe_remove(pe_ensemble* p, uint32_t member_id) {
    return eguard(p, [&](Ensemble& e) { e.remove(member_id); });
}

int pe_run_periods(pe_ensemble* p, int periods, int use_graph) {
    if (periods < 0) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) { e.run_periods(periods, use_graph != 0); });
}

int pe_wait(pe_ensemble* p) { return eguard(p, [](Ensemble& e) { e.wait(); }); }

int pe_get_density(pe_ensemble* p, uint32_t member_id, double* out) {
    if (!out) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) { e.get_density(member_id, out); });
}

int pe_get_field(pe_ensemble* p, uint32_t member_id, uint32_t field, void* out) {
    if (!out || field >= HC_NUM_CELL_FIELDS) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) { e.get_field(member_id, field, out); });
}

int pe_member_info(pe_ensemble* p, uint32_t member_id, uint64_t* bytes, int64_t* counts) {
    return eguard(p, [&](Ensemble& e) {
        physicell::Member& m = e.member(member_id);
        if (bytes) *bytes = m.bytes();
        if (counts) { counts[0] = m.substeps; counts[1] = m.mech_steps; counts[2] = m.pheno_steps; }
    });
}

int pe_info(pe_ensemble* p, uint32_t* n_members, uint64_t* total_cells, uint64_t* total_voxel_values,
            uint64_t* workspace_bytes) {
    return eguard(p, [&](Ensemble& e) {
        if (n_members) *n_members = (uint32_t)e.n_members();
        if (total_cells) *total_cells = e.total_cells();
        if (total_voxel_values) *total_voxel_values = e.total_voxel_values();
        if (workspace_bytes) *workspace_bytes = e.workspace_bytes();
    });
}

int pe_set_grid_blocks(pe_ensemble* p, int axis, uint32_t blocks) {
    return eguard(p, [&](Ensemble& e) { e.set_grid_blocks(axis, blocks); });
}

int pe_set_overlap(pe_ensemble* p, int on) {
    return eguard(p, [&](Ensemble& e) { e.set_overlap(on != 0); });
}

int pe_autotune(pe_ensemble* p, int reps, uint32_t* blocks_out, double* us_out) {
    if (!blocks_out || !us_out || reps < 1) return HC_ERR_ARG;
    return eguard(p, [&](Ensemble& e) { e.autotune(reps, blocks_out, us_out); });
}

hc_ctx* pe_context(pe_ensemble* p) { return (p && p->e) ? p->e->ctx() : nullptr; }

}  // extern "C"
