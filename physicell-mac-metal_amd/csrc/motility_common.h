// motility_common.h -- per-agent motility on the device, shared by the
// single-simulation kernel (motility.hip) and the packed ensemble (batch.hip).
//
// Between the force step and AB2, every live motile cell adds its motility
// vector to its velocity; the vector is resampled (persistent random walk
// biased along substrate gradients) when the persistence time is below dt or
// a uniform draw is below dt/persistence (src/motility.cpp:179-293).
// Gradients are central differences of the substrate-major field at the
// cell's voxel, indices clamped (src/motility.cpp:45-86).
#pragma once
#include "cells_common.h"

namespace hc {

template <typename F>
__device__ inline void gradient(const F* __restrict__ rho, const hc_grid& g, double px, double py, double pz,
                                uint32_t s, double& gx, double& gy, double& gz) {
    const int nx = (int)g.nx, ny = (int)g.ny, nz = (int)g.nz;
    int ix = (int)((px - g.x_min) / g.dx);
    int iy = (int)((py - g.y_min) / g.dy);
    int iz = (int)((pz - g.z_min) / g.dz);
    ix = ix < 0 ? 0 : (ix > nx - 1 ? nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > ny - 1 ? ny - 1 : iy);
    iz = iz < 0 ? 0 : (iz > nz - 1 ? nz - 1 : iz);
    const uint64_t V = (uint64_t)g.nx * g.ny * g.nz;
    auto D = [&](int i, int j, int k) {
        i = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
        j = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        k = k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k);
        return (double)rho[(uint64_t)s * V + (uint64_t)k * ny * nx + (uint64_t)j * nx + (uint64_t)i];
    };
    gx = (D(ix + 1, iy, iz) - D(ix - 1, iy, iz)) / (2.0 * g.dx);
    gy = (D(ix, iy + 1, iz) - D(ix, iy - 1, iz)) / (2.0 * g.dy);
    gz = 0.0;
    if (nz > 1) gz = (D(ix, iy, iz + 1) - D(ix, iy, iz - 1)) / (2.0 * g.dz);
}

// one agent's motility update (src/motility.cpp:179-293); used by the
// single-simulation and the ensemble kernels
template <typename F>
__device__ inline void motility_one(const CellPtrs& cp, const MotPtrs& mp, const F* __restrict__ rho, const hc_grid& g,
                                    double dt, uint32_t i, int use_sensitivity) {
    if (!cp.u[HC_F_IS_ALIVE - HC_NUM_F64_FIELDS][i]) return;
    if (!mp.motile[i]) {
        mp.vec[0][i] = mp.vec[1][i] = mp.vec[2][i] = 0.0;
        return;
    }
    const uint64_t step = mp.hdr[0], seed = mp.hdr[1];
    mp.elapsed[i] += dt;
    const double pt = mp.persistence[i];
    const bool resample = (pt < dt) || (uniform01(seed, step, i, 0) < dt / pt);
    if (resample) {
        const double pi = 3.14159265358979323846;
        mp.elapsed[i] = 0.0;
        double rx, ry, rz;
        const double theta = 2.0 * pi * uniform01(seed, step, i, 1);
        if (mp.restrict_2d[i]) {
            rx = cos(theta); ry = sin(theta); rz = 0.0;
        } else {
            const double phi = pi * uniform01(seed, step, i, 2);
            const double sp = sin(phi);
            rx = sp * cos(theta); ry = sp * sin(theta); rz = cos(phi);
        }
        double bx = 0.0, by = 0.0, bz = 0.0;
        const double b = mp.bias[i];
        if (b > 0.0 && rho) {
            const double px = cp.f[HC_F_POS_X][i], py = cp.f[HC_F_POS_Y][i], pz = cp.f[HC_F_POS_Z][i];
            if (use_sensitivity) {
                for (uint32_t s = 0; s < g.n_substrates; s++) {
                    const double sens = mp.sensitivity[(uint64_t)i * g.n_substrates + s];
                    if (fabs(sens) < 1e-16) continue;
                    double gx, gy, gz;
                    gradient(rho, g, px, py, pz, s, gx, gy, gz);
                    bx += sens * gx; by += sens * gy; bz += sens * gz;
                }
            } else {
                double gx, gy, gz;
                gradient(rho, g, px, py, pz, (uint32_t)mp.chem_index[i], gx, gy, gz);
                const double d = (double)mp.chem_dir[i];
                bx = d * gx; by = d * gy; bz = d * gz;
            }
            const double bmag = sqrt(bx * bx + by * by + bz * bz);
            if (bmag > 1e-16) { bx /= bmag; by /= bmag; bz /= bmag; }
        }
        const double omb = 1.0 - b;
        double mx = omb * rx + b * bx, my = omb * ry + b * by, mz = omb * rz + b * bz;
        const double mag = sqrt(mx * mx + my * my + mz * mz);
        if (mag > 1e-16) { mx /= mag; my /= mag; mz /= mag; }
        const double speed = cp.f[HC_F_MOT_SPEED][i];
        mp.vec[0][i] = mx * speed; mp.vec[1][i] = my * speed; mp.vec[2][i] = mz * speed;
    }
    cp.f[HC_F_VEL_X][i] += mp.vec[0][i];
    cp.f[HC_F_VEL_Y][i] += mp.vec[1][i];
    cp.f[HC_F_VEL_Z][i] += mp.vec[2][i];
}

}  // namespace hc
