// oracle.cpp -- CPU FP64 restatement of the reference's step hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against and
// the timed CPU baseline.  The product never links or calls this file.
//
// Compiled with -ffp-contract=off: every expression below is evaluated with
// the rounding of the operations as written (left to right), which is the
// sequence the reference's kernels spell out.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstring>
#include <random>
#include <vector>
#include <omp.h>

namespace {

inline int nthreads_or_default(int nt) { return nt > 0 ? nt : 1; }

// Decay split per sweep.  The reference is 2-D and hard-wires h = 0.5*lambda*dt
// (src/microenvironment.mm:91, applied in both sweeps).  With a third sweep the
// same splitting principle gives lambda*dt/3 per sweep, so that the product of
// the three diagonal factors reproduces (1+h)^3 ~ 1+lambda*dt.
inline double half_decay(const or_grid* g, const or_sub& s) {
    if (g->nz > 1) return s.decay * g->dt / 3.0;
    return 0.5 * s.decay * g->dt;
}

inline double axis_spacing(const or_grid* g, int axis) {
    return axis == 0 ? g->dx : (axis == 1 ? g->dy : g->dz);
}

// One line of the LOD system.  Restates shaders/diffusion_2d.metal:28-157
// (x-sweep; the y-sweep :165-292 is identical with stride nx).
//   - Dirichlet set before the solve (:55-58) and re-applied after (:151-155)
//   - n <= 2: nothing (:60-63); n == 3: single interior node (:66-84)
//   - n >= 4: unknowns u1..u_{n-2}; the edge rows use the boundary diagonal
//     1+c+h and add c*T to the RHS when Dirichlet is on (:89-108, :126-142)
//   - forward elimination (:111-124), back substitution (:144-148)
void solve_line(double* rho, size_t base, size_t stride, uint32_t n,
                double coeff, double h, bool dir, double T, double* c_prime) {
    auto at = [&](uint32_t i) -> double& { return rho[base + (size_t)i * stride]; };
    const double diag_interior = 1.0 + 2.0 * coeff + h;
    const double diag_boundary = 1.0 + coeff + h;

    if (dir) { at(0) = T; at(n - 1) = T; }
    if (n <= 2) return;

    if (n == 3) {
        double d_i = at(1);
        const double b = diag_interior;
        if (dir) {
            d_i += coeff * T;
            d_i += coeff * T;
        }
        at(1) = d_i / b;
        if (dir) { at(0) = T; at(n - 1) = T; }
        return;
    }

    // i = 1
    double d_i = at(1);
    const double b_1 = diag_boundary;
    const double c_1 = -coeff;
    if (dir) d_i += coeff * T;
    double c_p = c_1 / b_1;
    double d_p = d_i / b_1;
    c_prime[1] = c_p;
    at(1) = d_p;
    double c_prev = c_p;
    double d_prev = d_p;

    // i = 2 .. n-3
    for (uint32_t i = 2; i < n - 2; i++) {
        const double a_i = -coeff;
        const double b_i = diag_interior;
        const double c_i = -coeff;
        d_i = at(i);
        const double denom = b_i - a_i * c_prev;
        c_prev = c_i / denom;
        d_prev = (d_i - a_i * d_prev) / denom;
        c_prime[i] = c_prev;
        at(i) = d_prev;
    }

    // i = n-2
    {
        const uint32_t last = n - 2;
        const double a_last = -coeff;
        const double b_last = diag_boundary;
        d_i = at(last);
        if (dir) d_i += coeff * T;
        const double denom = b_last - a_last * c_prev;
        d_prev = (d_i - a_last * d_prev) / denom;
        at(last) = d_prev;
    }

    // back substitution i = n-3 .. 1
    for (int i = (int)n - 3; i >= 1; i--) {
        at((uint32_t)i) = at((uint32_t)i) - c_prime[i] * at((uint32_t)i + 1);
    }

    if (dir) { at(0) = T; at(n - 1) = T; }
}

}  // namespace

extern "C" {

void or_sweep(int axis, double* rho, const or_grid* g, const or_sub* subs, int nthreads) {
    const uint32_t nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t V = (size_t)nx * ny * nz;
    const uint32_t n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)nx : (size_t)nx * ny);
    // lines of one substrate: (outer, inner) pairs
    const uint32_t n_outer = axis == 0 ? nz : (axis == 1 ? nz : ny);
    const uint32_t n_inner = axis == 0 ? ny : nx;
    const double sp = axis_spacing(g, axis);
    const int64_t n_lines = (int64_t)g->ns * n_outer * n_inner;

#pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        std::vector<double> c_prime(n > 0 ? n : 1);
#pragma omp for schedule(static)
        for (int64_t L = 0; L < n_lines; L++) {
            const uint32_t s = (uint32_t)(L / ((int64_t)n_outer * n_inner));
            const int64_t r = L % ((int64_t)n_outer * n_inner);
            const uint32_t o = (uint32_t)(r / n_inner);
            const uint32_t in = (uint32_t)(r % n_inner);
            size_t base = (size_t)s * V;
            if (axis == 0) base += (size_t)o * nx * ny + (size_t)in * nx;       // (k, j)
            else if (axis == 1) base += (size_t)o * nx * ny + in;               // (k, i)
            else base += (size_t)o * nx + in;                                   // (j, i)
            const or_sub& sb = subs[s];
            const double coeff = sb.D * g->dt / (sp * sp);
            solve_line(rho, base, stride, n, coeff, half_decay(g, sb),
                       sb.dirichlet != 0, sb.target, c_prime.data());
        }
    }
}

void or_dirichlet(double* rho, const or_grid* g, const or_sub* subs) {
    const uint32_t nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t V = (size_t)nx * ny * nz;
    for (uint32_t s = 0; s < g->ns; s++) {
        if (!subs[s].dirichlet) continue;
        const double T = subs[s].target;
        double* b = rho + (size_t)s * V;
        for (uint32_t k = 0; k < nz; k++)
            for (uint32_t j = 0; j < ny; j++)
                for (uint32_t i = 0; i < nx; i++) {
                    const bool face = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 ||
                                      (nz > 1 && (k == 0 || k == nz - 1));
                    if (face) b[(size_t)k * nx * ny + (size_t)j * nx + i] = T;
                }
    }
}

void or_diffusion_substep(double* rho, const or_grid* g, const or_sub* subs, int nthreads) {
    or_sweep(0, rho, g, subs, nthreads);
    or_sweep(1, rho, g, subs, nthreads);
    if (g->nz > 1) or_sweep(2, rho, g, subs, nthreads);
    or_dirichlet(rho, g, subs);
}

// src/microenvironment.mm:176-184 (truncate toward zero, then clamp).
void or_voxel_keys(const double* x, const double* y, const double* z, uint32_t n,
                   const or_grid* g, uint32_t* out) {
    for (uint32_t c = 0; c < n; c++) {
        int i = (int)((x[c] - g->x_min) / g->dx);
        int j = (int)((y[c] - g->y_min) / g->dy);
        int k = (int)((z[c] - g->z_min) / g->dz);
        i = (i < 0) ? 0 : (i >= (int)g->nx ? (int)g->nx - 1 : i);
        j = (j < 0) ? 0 : (j >= (int)g->ny ? (int)g->ny - 1 : j);
        k = (k < 0) ? 0 : (k >= (int)g->nz ? (int)g->nz - 1 : k);
        out[c] = (uint32_t)k * g->nx * g->ny + (uint32_t)j * g->nx + (uint32_t)i;
    }
}

static inline void secrete_one(double* rho, const or_grid* g, size_t V, double dt, uint32_t c,
                               uint32_t vi, const double* volume, const double* sec,
                               const double* upt, const double* sat, const double* net) {
    // src/cell_phenotype.cpp:536-563
    const double V_cell = volume[c];
    const double V_voxel = g->dx * g->dy * g->dz;
    const double internal_constant = dt * V_cell / V_voxel;
    for (uint32_t s = 0; s < g->ns; s++) {
        const size_t off = (size_t)c * g->ns + s;
        const double sec_rate = sec[off];
        const double upt_rate = upt[off];
        const double sat_density = sat[off];
        const double net_export = net[off];
        const size_t voxel_off = (size_t)s * V + vi;
        double r = rho[voxel_off];
        r += dt * net_export / V_voxel;
        const double temp1 = internal_constant * sec_rate * sat_density;
        const double temp2 = 1.0 + internal_constant * (sec_rate + upt_rate);
        r = (r + temp1) / temp2;
        if (r < 0.0) r = 0.0;
        rho[voxel_off] = r;
    }
}

void or_secretion(double* rho, const or_grid* g, double dt, uint32_t n,
                  const uint32_t* voxel, const double* volume,
                  const double* sec, const double* upt, const double* sat,
                  const double* net, const uint32_t* alive) {
    if (g->ns == 0) return;
    const size_t V = (size_t)g->nx * g->ny * g->nz;
    for (uint32_t c = 0; c < n; c++) {
        if (alive && !alive[c]) continue;                 // src/main.mm:360
        const uint32_t vi = voxel[c];
        if (vi >= V) continue;                            // src/cell_phenotype.cpp:534
        secrete_one(rho, g, V, dt, c, vi, volume, sec, upt, sat, net);
    }
}

void or_secretion_grouped(double* rho, const or_grid* g, double dt, uint32_t n,
                          const uint32_t* voxel, const double* volume,
                          const double* sec, const double* upt, const double* sat,
                          const double* net, const uint32_t* alive, int nthreads) {
    if (g->ns == 0 || n == 0) return;
    const size_t V = (size_t)g->nx * g->ny * g->nz;
    // group eligible cells by voxel, ascending index inside each group
    std::vector<uint32_t> order;
    order.reserve(n);
    for (uint32_t c = 0; c < n; c++)
        if ((!alive || alive[c]) && voxel[c] < V) order.push_back(c);
    std::vector<uint32_t> tmp(order.size());
    // stable sort by voxel via counting on a compact key space
    std::vector<uint32_t> keys(order.size());
    for (size_t t = 0; t < order.size(); t++) keys[t] = voxel[order[t]];
    std::vector<size_t> idx(order.size());
    for (size_t t = 0; t < idx.size(); t++) idx[t] = t;
    // radix by voxel (two 16-bit passes keep it O(n) with small tables)
    for (int pass = 0; pass < 2; pass++) {
        std::vector<size_t> cnt(65537, 0);
        const int sh = pass * 16;
        for (size_t t = 0; t < idx.size(); t++) cnt[((keys[idx[t]] >> sh) & 0xFFFF) + 1]++;
        for (int b = 0; b < 65536; b++) cnt[b + 1] += cnt[b];
        std::vector<size_t> out(idx.size());
        for (size_t t = 0; t < idx.size(); t++) out[cnt[(keys[idx[t]] >> sh) & 0xFFFF]++] = idx[t];
        idx.swap(out);
    }
    std::vector<size_t> starts;
    for (size_t t = 0; t < idx.size(); t++)
        if (t == 0 || keys[idx[t]] != keys[idx[t - 1]]) starts.push_back(t);
    starts.push_back(idx.size());
    const int64_t ng = (int64_t)starts.size() - 1;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
    for (int64_t q = 0; q < ng; q++)
        for (size_t t = starts[q]; t < starts[q + 1]; t++) {
            const uint32_t c = order[idx[t]];
            secrete_one(rho, g, V, dt, c, voxel[c], volume, sec, upt, sat, net);
        }
}

// shaders/mechanics.metal:54-72
void or_mech_keys(const double* x, const double* y, const double* z, uint32_t n,
                  const or_mech* m, uint32_t* out) {
    for (uint32_t c = 0; c < n; c++) {
        int vi = (int)((x[c] - m->x_min) / m->voxel);
        int vj = (int)((y[c] - m->y_min) / m->voxel);
        int vk = (int)((z[c] - m->z_min) / m->voxel);
        vi = vi < 0 ? 0 : (vi > (int)m->mnx - 1 ? (int)m->mnx - 1 : vi);
        vj = vj < 0 ? 0 : (vj > (int)m->mny - 1 ? (int)m->mny - 1 : vj);
        vk = vk < 0 ? 0 : (vk > (int)m->mnz - 1 ? (int)m->mnz - 1 : vk);
        out[c] = (uint32_t)vk * m->mny * m->mnx + (uint32_t)vj * m->mnx + (uint32_t)vi;
    }
}

void or_hash(const uint32_t* keys, uint32_t n, uint32_t nbins,
             uint32_t* offsets, uint32_t* sorted) {
    std::memset(offsets, 0, sizeof(uint32_t) * ((size_t)nbins + 1));
    for (uint32_t c = 0; c < n; c++) offsets[keys[c] + 1]++;
    for (uint32_t b = 0; b < nbins; b++) offsets[b + 1] += offsets[b];
    std::vector<uint32_t> pos(offsets, offsets + nbins);
    for (uint32_t c = 0; c < n; c++) sorted[pos[keys[c]]++] = c;
}

// shaders/mechanics.metal:91-216, with the bucket list replaced by the
// ascending-index bin of the counting sort.
void or_forces(const double* x, const double* y, const double* z,
               const double* radius, const double* rep, const double* adh,
               const double* rel_adh, uint32_t n, const uint32_t* keys,
               const uint32_t* offsets, const uint32_t* sorted, const or_mech* m,
               double* vx, double* vy, double* vz, double* pressure, int nthreads) {
    const uint32_t mnx = m->mnx, mny = m->mny, mnz = m->mnz;
    const double max_dist = m->max_dist;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const double px = x[i], py = y[i], pz = z[i];
        const double radius_i = radius[i], rep_i = rep[i], adh_i = adh[i];
        const double rel_max_adh = rel_adh[i];
        const uint32_t my_voxel = keys[i];
        double fx = 0.0, fy = 0.0, fz = 0.0, local_pressure = 0.0;
        const uint32_t vi = my_voxel % mnx;
        const uint32_t vj = (my_voxel / mnx) % mny;
        const uint32_t vk = my_voxel / (mnx * mny);
        const int dk_min = (mnz > 1) ? -1 : 0;
        const int dk_max = (mnz > 1) ? 1 : 0;
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
                    for (uint32_t t = offsets[nb]; t < offsets[nb + 1]; t++) {
                        const uint32_t other = sorted[t];
                        if (other == i) continue;
                        const double ddx = px - x[other];
                        const double ddy = py - y[other];
                        const double ddz = pz - z[other];
                        const double dist_sq = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (dist_sq >= max_dist * max_dist) continue;
                        if (dist_sq < 1e-12) continue;
                        const double dist = std::sqrt(dist_sq);
                        const double inv_dist = 1.0 / dist;
                        const double nxv = ddx * inv_dist;
                        const double nyv = ddy * inv_dist;
                        const double nzv = ddz * inv_dist;
                        const double R_sum = radius_i + radius[other];
                        double force_mag = 0.0;
                        if (dist < R_sum) {
                            const double overlap = 1.0 - dist / R_sum;
                            force_mag += overlap * overlap * std::sqrt(rep_i * rep[other]);
                            local_pressure += overlap * overlap / 0.027288820670331;
                        }
                        const double adhesion_distance = rel_max_adh * R_sum;
                        if (dist < adhesion_distance) {
                            const double t2 = 1.0 - dist / adhesion_distance;
                            force_mag -= t2 * t2 * std::sqrt(adh_i * adh[other]);
                        }
                        fx += force_mag * nxv;
                        fy += force_mag * nyv;
                        fz += force_mag * nzv;
                    }
                }
            }
        }
        vx[i] = fx; vy[i] = fy; vz[i] = fz;
        pressure[i] = local_pressure;
    }
}

void or_neighbors(const double* x, const double* y, const double* z, uint32_t n,
                  const uint32_t* keys, const uint32_t* offsets, const uint32_t* sorted,
                  const or_mech* m, uint32_t cap, uint32_t* counts, uint32_t* idx) {
    const uint32_t mnx = m->mnx, mny = m->mny, mnz = m->mnz;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t cnt = 0;
        const uint32_t my = keys[i];
        const uint32_t vi = my % mnx, vj = (my / mnx) % mny, vk = my / (mnx * mny);
        const int dk_min = (mnz > 1) ? -1 : 0, dk_max = (mnz > 1) ? 1 : 0;
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
                    for (uint32_t t = offsets[nb]; t < offsets[nb + 1]; t++) {
                        const uint32_t o = sorted[t];
                        if (o == i) continue;
                        const double ddx = x[i] - x[o], ddy = y[i] - y[o], ddz = z[i] - z[o];
                        const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (d2 >= m->max_dist * m->max_dist) continue;
                        if (d2 < 1e-12) continue;
                        if (cnt < cap) idx[(size_t)i * cap + cnt] = o;
                        cnt++;
                    }
                }
            }
        }
        counts[i] = cnt;
    }
}

// shaders/integrate.metal:56-79; the clamp extent is the mech grid's
// x_min + n*voxel (src/cell_mechanics.mm:58-70 builds that GridParams).
void or_integrate(double* x, double* y, double* z,
                  const double* vx, const double* vy, const double* vz,
                  double* pvx, double* pvy, double* pvz,
                  uint32_t n, const or_mech* m, double dt, int nthreads) {
    const double x_max = m->x_min + (double)m->mnx * m->voxel;
    const double y_max = m->y_min + (double)m->mny * m->voxel;
    const double z_max = m->z_min + (double)m->mnz * m->voxel;
#pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const size_t i = (size_t)ii;
        double px = x[i], py = y[i], pz = z[i];
        const double ux = vx[i], uy = vy[i], uz = vz[i];
        px += dt * (1.5 * ux - 0.5 * pvx[i]);
        py += dt * (1.5 * uy - 0.5 * pvy[i]);
        pz += dt * (1.5 * uz - 0.5 * pvz[i]);
        px = px < m->x_min ? m->x_min : (px > x_max ? x_max : px);
        py = py < m->y_min ? m->y_min : (py > y_max ? y_max : py);
        pz = pz < m->z_min ? m->z_min : (pz > z_max ? z_max : pz);
        x[i] = px; y[i] = py; z[i] = pz;
        pvx[i] = ux; pvy[i] = uy; pvz[i] = uz;
    }
}

void or_mech_step(or_state* st, const or_mech* m, double dt_mech,
                  uint32_t* hash_offsets, uint32_t* hash_sorted, int nthreads) {
    const uint32_t nb = m->mnx * m->mny * m->mnz;
    or_mech_keys(st->x, st->y, st->z, st->n, m, st->mech_key);
    or_hash(st->mech_key, st->n, nb, hash_offsets, hash_sorted);
    or_forces(st->x, st->y, st->z, st->radius, st->rep, st->adh, st->rel_adh, st->n,
              st->mech_key, hash_offsets, hash_sorted, m, st->vx, st->vy, st->vz,
              st->pressure, nthreads);
    or_integrate(st->x, st->y, st->z, st->vx, st->vy, st->vz, st->pvx, st->pvy, st->pvz,
                 st->n, m, dt_mech, nthreads);
}

// src/main.mm:275-415 (Appendix B of the survey): the loop condition, the
// accumulators and the phase order are kept; output, motility, rules and
// phenotype (other than the voxel refresh) are not on the path.
void or_run(or_state* st, const or_grid* g, const or_sub* subs, const or_mech* m,
            const or_schedule* sch, uint32_t* hash_offsets, uint32_t* hash_sorted,
            int nthreads, or_counts* counts) {
    double current_time = 0.0;
    double mech_acc = 0.0, pheno_acc = 0.0;
    const double dtd = sch->dt_diff;
    int64_t ns = 0, nm = 0, np = 0;
    while (current_time < sch->max_time + 0.5 * dtd) {
        or_diffusion_substep(st->rho, g, subs, nthreads);
        if (nthreads > 1)
            or_secretion_grouped(st->rho, g, dtd, st->n, st->voxel, st->volume, st->sec, st->upt,
                                 st->sat, st->net, st->alive, nthreads);
        else
            or_secretion(st->rho, g, dtd, st->n, st->voxel, st->volume, st->sec, st->upt,
                         st->sat, st->net, st->alive);
        ns++;
        mech_acc += dtd;
        if (mech_acc >= sch->dt_mech - 0.5 * dtd) {
            or_mech_step(st, m, sch->dt_mech, hash_offsets, hash_sorted, nthreads);
            nm++;
            mech_acc = 0.0;
        }
        pheno_acc += dtd;
        if (pheno_acc >= sch->dt_pheno - 0.5 * dtd) {
            or_voxel_keys(st->x, st->y, st->z, st->n, g, st->voxel);
            np++;
            pheno_acc = 0.0;
        }
        current_time += dtd;
    }
    if (counts) { counts->substeps = ns; counts->mech_steps = nm; counts->pheno_steps = np; }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// motility (src/motility.cpp:179-293); the random stream is counter-based
// (SplitMix64 finaliser, Steele, Lea & Flood 2014) so that it is reproducible
// and the same on the device
// ---------------------------------------------------------------------------
static uint64_t or_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double or_uniform(uint64_t seed, uint64_t step, uint32_t cell, uint32_t draw) {
    const uint64_t h = or_mix64(seed ^ or_mix64(step ^ or_mix64(((uint64_t)cell << 2) | draw)));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// central differences of substrate s at the cell's voxel, indices clamped
// (src/motility.cpp:45-86)
static void or_gradient(const double* rho, const or_grid* g, double px, double py, double pz, uint32_t s,
                        double& gx, double& gy, double& gz) {
    const int nx = (int)g->nx, ny = (int)g->ny, nz = (int)g->nz;
    int ix = (int)((px - g->x_min) / g->dx);
    int iy = (int)((py - g->y_min) / g->dy);
    int iz = (int)((pz - g->z_min) / g->dz);
    ix = ix < 0 ? 0 : (ix > nx - 1 ? nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > ny - 1 ? ny - 1 : iy);
    iz = iz < 0 ? 0 : (iz > nz - 1 ? nz - 1 : iz);
    const size_t V = (size_t)g->nx * g->ny * g->nz;
    auto D = [&](int i, int j, int k) {
        i = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
        j = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        k = k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k);
        return rho[(size_t)s * V + (size_t)k * ny * nx + (size_t)j * nx + (size_t)i];
    };
    gx = (D(ix + 1, iy, iz) - D(ix - 1, iy, iz)) / (2.0 * g->dx);
    gy = (D(ix, iy + 1, iz) - D(ix, iy - 1, iz)) / (2.0 * g->dy);
    gz = 0.0;
    if (nz > 1) gz = (D(ix, iy, iz + 1) - D(ix, iy, iz - 1)) / (2.0 * g->dz);
}

void or_motility(const double* x, const double* y, const double* z, double* vx, double* vy, double* vz,
                 const uint32_t* alive, const double* speed, uint32_t n, const or_grid* g, const double* rho,
                 double dt, uint64_t seed, uint64_t step, double* persistence, double* elapsed, const double* bias,
                 double* mvx, double* mvy, double* mvz, const uint32_t* motile, const uint32_t* restrict_2d,
                 const int32_t* chem_index, const int32_t* chem_dir, const double* sensitivity) {
    const double pi = 3.14159265358979323846;
    for (uint32_t i = 0; i < n; i++) {
        if (!alive[i]) continue;
        if (!motile[i]) {
            mvx[i] = mvy[i] = mvz[i] = 0.0;
            continue;
        }
        elapsed[i] += dt;
        const double pt = persistence[i];
        const bool resample = (pt < dt) || (or_uniform(seed, step, i, 0) < dt / pt);
        if (resample) {
            elapsed[i] = 0.0;
            double rx, ry, rz;
            const double theta = 2.0 * pi * or_uniform(seed, step, i, 1);
            if (restrict_2d[i]) {
                rx = cos(theta); ry = sin(theta); rz = 0.0;
            } else {
                const double phi = pi * or_uniform(seed, step, i, 2);
                const double sp = sin(phi);
                rx = sp * cos(theta); ry = sp * sin(theta); rz = cos(phi);
            }
            double bx = 0.0, by = 0.0, bz = 0.0;
            const double b = bias[i];
            if (b > 0.0 && rho) {
                if (sensitivity) {
                    for (uint32_t s = 0; s < g->ns; s++) {
                        const double sens = sensitivity[(size_t)i * g->ns + s];
                        if (fabs(sens) < 1e-16) continue;
                        double gx, gy, gz;
                        or_gradient(rho, g, x[i], y[i], z[i], s, gx, gy, gz);
                        bx += sens * gx; by += sens * gy; bz += sens * gz;
                    }
                } else {
                    double gx, gy, gz;
                    or_gradient(rho, g, x[i], y[i], z[i], (uint32_t)chem_index[i], gx, gy, gz);
                    bx = chem_dir[i] * gx; by = chem_dir[i] * gy; bz = chem_dir[i] * gz;
                }
                const double bmag = sqrt(bx * bx + by * by + bz * bz);
                if (bmag > 1e-16) { bx /= bmag; by /= bmag; bz /= bmag; }
            }
            const double omb = 1.0 - b;
            double mx = omb * rx + b * bx, my = omb * ry + b * by, mz = omb * rz + b * bz;
            const double mag = sqrt(mx * mx + my * my + mz * mz);
            if (mag > 1e-16) { mx /= mag; my /= mag; mz /= mag; }
            mvx[i] = mx * speed[i]; mvy[i] = my * speed[i]; mvz[i] = mz * speed[i];
        }
        vx[i] += mvx[i]; vy[i] += mvy[i]; vz[i] += mvz[i];
    }
}

// ---------------------------------------------------------------------------
// phenotype-side SoA mutation (src/cell_phenotype.cpp:52-113, 571-720,
// 833-844; src/cell_data.cpp:398-542)
// ---------------------------------------------------------------------------
namespace {
enum { F_PX = 0, F_PY = 1, F_PZ = 2, F_RAD = 9, F_NRAD = 10, F_TVOL = 11, F_NVOL = 12, F_REP = 13, F_ADH = 14,
       F_REL = 15, F_SPEED = 16, F_BIAS = 17, F_ONCO = 21 };
enum { U_TYPE = 0, U_PHASE = 1, U_ALIVE = 2, U_VOX = 3, U_MVOX = 4 };
enum { V_FLUID = 0, V_SNUC = 1, V_SCYTO = 2, V_FFRAC = 3, V_CFRAC = 4, V_TSCYTO = 5, V_TSNUC = 6, V_TFFRAC = 7,
       V_RATIO = 8, V_FRATE = 9, V_NRATE = 10, V_CRATE = 11, V_CALCRATE = 12, V_RUPT = 13, V_RELRUPT = 14 };
const double OR_FOUR_THIRDS_PI = 4.188790204786391;    // src/cell_phenotype.cpp:45
}

void or_volume_update(or_cells* c, const uint32_t* mask, uint32_t n, double dt, int use_ratio) {
    for (uint32_t i = 0; i < n; i++) {
        if (mask && !mask[i]) continue;
        double** v = c->vol;
        double total = c->f64[F_TVOL][i];
        double V_fluid = v[V_FLUID][i];
        double nuclear_solid = v[V_SNUC][i];
        double cyto_solid = v[V_SCYTO][i];
        V_fluid += dt * v[V_FRATE][i] * (v[V_TFFRAC][i] * total - V_fluid);
        if (V_fluid < 0.0) V_fluid = 0.0;
        double nuclear = c->f64[F_NVOL][i];
        const double nuclear_fluid = (nuclear / (total + 1e-16)) * V_fluid;
        const double cyto_fluid = V_fluid - nuclear_fluid;
        nuclear_solid += dt * v[V_NRATE][i] * (v[V_TSNUC][i] - nuclear_solid);
        if (nuclear_solid < 0.0) nuclear_solid = 0.0;
        double target_sc = v[V_TSCYTO][i];
        if (use_ratio) {
            target_sc = v[V_RATIO][i] * v[V_TSNUC][i];
            v[V_TSCYTO][i] = target_sc;
        }
        cyto_solid += dt * v[V_CRATE][i] * (target_sc - cyto_solid);
        if (cyto_solid < 0.0) cyto_solid = 0.0;
        double calc_frac = v[V_CFRAC][i];
        calc_frac += dt * v[V_CALCRATE][i] * (1.0 - calc_frac);
        nuclear = nuclear_solid + nuclear_fluid;
        const double cytoplasmic = cyto_solid + cyto_fluid;
        total = cytoplasmic + nuclear;
        const double fluid_fraction = V_fluid / (total + 1e-16);
        c->f64[F_TVOL][i] = total;
        c->f64[F_NVOL][i] = nuclear;
        v[V_FLUID][i] = V_fluid;
        v[V_SNUC][i] = nuclear_solid;
        v[V_SCYTO][i] = cyto_solid;
        v[V_FFRAC][i] = fluid_fraction;
        v[V_CFRAC][i] = calc_frac;
        c->f64[F_RAD][i] = std::cbrt(total / OR_FOUR_THIRDS_PI);
        c->f64[F_NRAD][i] = std::cbrt(nuclear / OR_FOUR_THIRDS_PI);
    }
}

uint32_t or_divide(or_cells* c, const uint32_t* parents, uint32_t count, uint32_t n, uint32_t max_cells,
                   uint64_t seed, uint64_t counter) {
    const double pi = 3.14159265358979323846;
    uint32_t cur = n;
    for (uint32_t k = 0; k < count; k++) {
        if (cur >= max_cells) break;                     // CellData::addCell: no room
        const uint32_t p = parents[k], d = cur++;
        // fresh values (src/cell_data.cpp:544-640) for what divideCell leaves
        for (int f = 0; f < 23; f++) c->f64[f][d] = 0.0;
        c->f64[F_ONCO][d] = 1.0;
        for (int f : {F_TVOL, F_NVOL}) {
            const double h = c->f64[f][p] * 0.5;
            c->f64[f][p] = c->f64[f][d] = h;
        }
        for (int f : {V_FLUID, V_SCYTO, V_SNUC, V_RUPT, V_TSNUC, V_TSCYTO}) {
            const double h = c->vol[f][p] * 0.5;
            c->vol[f][p] = c->vol[f][d] = h;
        }
        for (int f : {V_FFRAC, V_CRATE, V_NRATE, V_FRATE, V_CALCRATE, V_CFRAC, V_TFFRAC, V_RELRUPT})
            c->vol[f][d] = c->vol[f][p];
        c->vol[V_RATIO][d] = 3.62;
        const double r = std::cbrt(c->f64[F_TVOL][p] / OR_FOUR_THIRDS_PI);
        const double nr = std::cbrt(c->f64[F_NVOL][p] / OR_FOUR_THIRDS_PI);
        c->f64[F_RAD][p] = c->f64[F_RAD][d] = r;
        c->f64[F_NRAD][p] = c->f64[F_NRAD][d] = nr;
        const double theta = or_uniform(seed, counter, k, 0) * 2.0 * pi;
        const double offset = r * 0.5;
        c->f64[F_PX][d] = c->f64[F_PX][p] + offset * std::cos(theta);
        c->f64[F_PY][d] = c->f64[F_PY][p] + offset * std::sin(theta);
        c->f64[F_PZ][d] = c->f64[F_PZ][p];
        for (int f : {F_REP, F_ADH, F_REL, F_SPEED, F_BIAS}) c->f64[f][d] = c->f64[f][p];
        const double onco = c->f64[F_ONCO][p];
        const double u1 = 1.0 - or_uniform(seed, counter, k, 1), u2 = or_uniform(seed, counter, k, 2);
        const double noise = 0.02 * onco * (std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * pi * u2));
        c->f64[F_ONCO][d] = std::max(0.0, onco + noise);
        c->f64[F_ONCO][p] = std::max(0.0, onco - noise);
        c->u32[U_TYPE][d] = c->u32[U_TYPE][p];
        c->u32[U_PHASE][p] = c->u32[U_PHASE][d] = 0u;
        c->u32[U_ALIVE][d] = 1u;
        c->u32[U_VOX][d] = c->u32[U_VOX][p];
        c->u32[U_MVOX][d] = c->u32[U_MVOX][p];
        for (int a = 0; a < 4; a++)
            for (uint32_t s = 0; s < c->ns; s++)
                c->rates[a][(size_t)d * c->ns + s] = c->rates[a][(size_t)p * c->ns + s];
        if (c->persistence) {
            c->persistence[d] = c->persistence[p];
            c->elapsed[d] = 0.0;
            c->bias[d] = c->bias[p];
            for (int q = 0; q < 3; q++) c->vec[q][d] = c->vec[q][p];
            c->motile[d] = c->motile[p];
            c->restrict_2d[d] = c->restrict_2d[p];
            c->chem_index[d] = c->chem_index[p];
            c->chem_dir[d] = c->chem_dir[p];
            for (uint32_t s = 0; s < c->ns; s++)
                c->sensitivity[(size_t)d * c->ns + s] = c->sensitivity[(size_t)p * c->ns + s];
        }
    }
    return cur;
}

uint32_t or_remove(or_cells* c, const uint32_t* indices, uint32_t count, uint32_t n) {
    std::vector<uint32_t> idx(indices, indices + count);
    std::sort(idx.begin(), idx.end(), std::greater<uint32_t>());
    uint32_t cur = n;
    for (uint32_t index : idx) {
        if (index >= cur) continue;                      // src/cell_phenotype.cpp:840
        const uint32_t last = cur - 1;
        if (index != last) {
            for (int f = 0; f < 23; f++) c->f64[f][index] = c->f64[f][last];
            for (int f = 0; f < 5; f++) c->u32[f][index] = c->u32[f][last];
            for (int a = 0; a < 4; a++)
                for (uint32_t s = 0; s < c->ns; s++)
                    c->rates[a][(size_t)index * c->ns + s] = c->rates[a][(size_t)last * c->ns + s];
            for (int f = 0; f < 15; f++)
                if (c->vol[f]) c->vol[f][index] = c->vol[f][last];
            if (c->persistence) {
                c->persistence[index] = c->persistence[last];
                c->elapsed[index] = c->elapsed[last];
                c->bias[index] = c->bias[last];
                for (int q = 0; q < 3; q++) c->vec[q][index] = c->vec[q][last];
                c->motile[index] = c->motile[last];
                c->restrict_2d[index] = c->restrict_2d[last];
                c->chem_index[index] = c->chem_index[last];
                c->chem_dir[index] = c->chem_dir[last];
                for (uint32_t s = 0; s < c->ns; s++)
                    c->sensitivity[(size_t)index * c->ns + s] = c->sensitivity[(size_t)last * c->ns + s];
            }
        }
        cur--;
    }
    return cur;
}

// ---------------------------------------------------------------------------
// neighbour queries (src/output_writer.cpp:457-490, src/cell_phenotype.cpp:
// 893-967)
// ---------------------------------------------------------------------------
uint64_t or_contact_graph(const double* x, const double* y, const double* z, const double* radius,
                          const double* volume, const double* rel, const uint32_t* alive, uint32_t n,
                          uint32_t* offsets, uint32_t* idx) {
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = (uint32_t)tot;
        if (!alive[i] || !(volume[i] > 1e-15)) continue;
        for (uint32_t j = i + 1; j < n; j++) {
            if (!alive[j] || !(volume[j] > 1e-15)) continue;
            const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            const double md = std::max(rel[i], rel[j]) * (radius[i] + radius[j]);
            if (dx * dx + dy * dy + dz * dz <= md * md) {
                if (idx) idx[tot] = j;
                tot++;
            }
        }
    }
    offsets[n] = (uint32_t)tot;
    return tot;
}

uint32_t or_form_attachments(const double* x, const double* y, const double* z, const double* radius,
                             const double* rel, const uint32_t* alive, const double* rate, uint32_t n, double dt,
                             uint64_t seed, uint32_t* count, uint32_t* targets) {
    // tryFormAllAttachments (src/cell_interactions.cpp:459-478) calling
    // tryFormAttachment (:411-455) on every pair j > i in order; the draw is
    // a std::mt19937 double in [0, 1) (:293-298), seeded here
    std::mt19937 gen((std::mt19937::result_type)seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    uint32_t formed = 0;
    const uint32_t M = OR_MAX_ATTACHMENTS;
    for (uint32_t i = 0; i < n; i++) {
        if (!alive[i] || rate[i] <= 0.0 || count[i] >= M) continue;
        for (uint32_t j = i + 1; j < n; j++) {
            if (!alive[j] || rate[j] <= 0.0) continue;
            // tryFormAttachment: capacity, already attached, distance, draw
            if (count[i] < M && count[j] < M) {
                bool have = false;
                for (uint32_t a = 0; a < count[i]; a++) have = have || targets[(size_t)i * M + a] == j;
                if (!have) {
                    const double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    const double md = std::max(rel[i], rel[j]) * (radius[i] + radius[j]);
                    if (!(dist > md) && uni(gen) < std::max(rate[i], rate[j]) * dt) {
                        targets[(size_t)i * M + count[i]++] = j;
                        targets[(size_t)j * M + count[j]++] = i;
                        formed++;
                    }
                }
            }
            if (count[i] >= M) break;
        }
    }
    return formed;
}

uint64_t or_interaction_candidates(const double* x, const double* y, const double* radius, const double* volume,
                                   uint32_t n, const uint32_t* keys, const uint32_t* off, const uint32_t* srt,
                                   const or_mech* m, uint32_t* offsets, uint32_t* idx) {
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = (uint32_t)tot;
        const uint32_t key = keys[i];
        const int vx = (int)(key % m->mnx), vy = (int)((key / m->mnx) % m->mny), vz = (int)(key / (m->mnx * m->mny));
        std::vector<uint32_t> cand;
        for (int dk = -1; dk <= 1; dk++) {
            const int nz = vz + dk;
            if (nz < 0 || nz >= (int)m->mnz) continue;
            for (int dj = -1; dj <= 1; dj++) {
                const int ny = vy + dj;
                if (ny < 0 || ny >= (int)m->mny) continue;
                for (int di = -1; di <= 1; di++) {
                    const int nx = vx + di;
                    if (nx < 0 || nx >= (int)m->mnx) continue;
                    const uint32_t b = ((uint32_t)nz * m->mny + (uint32_t)ny) * m->mnx + (uint32_t)nx;
                    for (uint32_t u = off[b]; u < off[b + 1] && cand.size() < 512; u++) {
                        const uint32_t j = srt[u];
                        if (j != i && j < n) cand.push_back(j);
                    }
                }
            }
        }
        const double ri = radius[i];
        for (uint32_t j : cand) {
            if (volume[j] < 1e-15) continue;
            const double dx = x[j] - x[i], dy = y[j] - y[i];
            const double d2 = dx * dx + dy * dy, rj = radius[j];
            const double max_r = 2.5 * ri + rj;
            if (d2 > max_r * max_r) continue;
            if (std::sqrt(d2) > (ri + rj) * 1.25) continue;
            if (idx) idx[tot] = j;
            tot++;
        }
    }
    offsets[n] = (uint32_t)tot;
    return tot;
}
