// ref_driver.cpp -- thin C entry points into the reference's own compiled
// processSecretion (src/cell_phenotype.cpp:526-564), updateMotility
// (src/motility.cpp:179-293) and .mat writers
// (src/matlab_writer.cpp, src/output_writer.cpp:295-453), used only by tests
// to pin the FP64 oracle's secretion restatement and the library's snapshot
// writer against the reference code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/build_ref.sh into oracle/_ref/
// from the reference sources where they lie; nothing here is product code.
#include "cell_data.h"
#include "cell_phenotype.h"
#include "types.h"

#include <vector>

extern "C" int ref_secretion(uint32_t n, uint32_t ns, uint32_t nx, uint32_t ny, uint32_t nz,
                             float dx, float dy, float dz, double dt,
                             const uint32_t* voxel, const float* volume,
                             const double* sec, const double* upt, const double* sat,
                             const double* net, float* density) {
    CellData cells(n > 0 ? n : 1);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    cells.allocateSubstrateBuffers(ns);
    cells.num_cells = n;
    for (uint32_t i = 0; i < n; i++) {
        cells.voxel_index[i] = voxel[i];
        cells.total_volume[i] = volume[i];
        for (uint32_t s = 0; s < ns; s++) {
            const size_t off = (size_t)i * ns + s;
            cells.secretion_rate[off] = sec[off];
            cells.uptake_rate[off] = upt[off];
            cells.saturation_density[off] = sat[off];
            cells.net_export_rate[off] = net[off];
        }
    }
    GridParams gp{};
    gp.nx = nx; gp.ny = ny; gp.nz = nz; gp.n_substrates = ns;
    gp.dx = dx; gp.dy = dy; gp.dz = dz; gp.dt = (float)dt;
    // one cell after another, ascending index (single-thread run of main.mm:358-362)
    for (uint32_t i = 0; i < n; i++) processSecretion(cells, density, gp, i, dt);
    return 0;
}

// The reference's MATLAB Level-4 writer (src/matlab_writer.cpp:5-49) on a
// matrix given column-major (data[c*rows + r] = M[r][c]).
#include "matlab_writer.h"
#include "output_writer.h"

extern "C" int ref_write_matlab4(const char* path, const char* name, uint32_t rows, uint32_t cols,
                                 const double* data) {
    std::vector<std::vector<double>> m(cols, std::vector<double>(rows));
    for (uint32_t c = 0; c < cols; c++)
        for (uint32_t r = 0; r < rows; r++) m[c][r] = data[(size_t)c * rows + r];
    return write_matlab4(m, path, name) ? 0 : -1;
}

// The reference's snapshot writer (src/output_writer.cpp:295-453) on a cell
// store set up the way src/main.mm:163-213 sets it up (GPU-shared fields, CPU
// buffers, substrate buffers, one cell type, interaction-rate buffers), with
// the path's fields given per cell (FP32 in the reference's store).
extern "C" int ref_write_snapshot(const char* folder, int frame, uint32_t n, const double* x, const double* y,
                                  const double* z, const double* total_volume, const double* nuclear_volume,
                                  const double* radius, const double* nuclear_radius, const double* vx,
                                  const double* vy, const double* vz, const double* pressure,
                                  const double* adhesion, const double* repulsion, const double* rel_adh,
                                  const uint32_t* cell_type, const uint32_t* phase, const uint32_t* alive,
                                  uint32_t ns, const double* sec, const double* upt, const double* sat,
                                  const double* net, uint32_t nx, uint32_t ny, uint32_t nz, float dx, float dy,
                                  float dz, float x_min, float y_min, float z_min, const float* density) {
    CellData cells(n > 0 ? n : 1);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    cells.allocateCPUBuffers();
    cells.allocateSubstrateBuffers(ns);
    cells.num_cell_types = 1;
    cells.allocateInteractionRateBuffers();
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t c = cells.addCell();
        cells.position_x[c] = (float)x[i];
        cells.position_y[c] = (float)y[i];
        cells.position_z[c] = (float)z[i];
        cells.total_volume[c] = (float)total_volume[i];
        cells.nuclear_volume[c] = (float)nuclear_volume[i];
        cells.radius[c] = (float)radius[i];
        cells.nuclear_radius[c] = (float)nuclear_radius[i];
        cells.velocity_x[c] = (float)vx[i];
        cells.velocity_y[c] = (float)vy[i];
        cells.velocity_z[c] = (float)vz[i];
        cells.simple_pressure[c] = (float)pressure[i];
        cells.cell_cell_adhesion[c] = (float)adhesion[i];
        cells.cell_cell_repulsion[c] = (float)repulsion[i];
        cells.relative_max_adhesion_distance[c] = (float)rel_adh[i];
        cells.cell_type[c] = cell_type[i];
        cells.current_phase[c] = phase[i];
        cells.is_alive[c] = alive[i];
        for (uint32_t s = 0; s < ns; s++) {
            const size_t off = (size_t)i * ns + s;
            cells.secretion_rate[off] = sec[off];
            cells.uptake_rate[off] = upt[off];
            cells.saturation_density[off] = sat[off];
            cells.net_export_rate[off] = net[off];
        }
    }
    GridParams gp{};
    gp.nx = nx; gp.ny = ny; gp.nz = nz; gp.n_substrates = ns;
    gp.dx = dx; gp.dy = dy; gp.dz = dz;
    gp.x_min = x_min; gp.y_min = y_min; gp.z_min = z_min;
    writeMatlabSnapshots(cells, density, gp, frame, folder);
    return 0;
}

// The reference's updateMotility (src/motility.cpp:179-293) on n cells.
// Positions, velocities and speeds are the reference's FP32 cell fields;
// density is read the way the reference reads it (density[voxel*ns + s]).
// legacy != 0 takes the single-substrate chemotaxis branch (:237-250) by
// leaving the sensitivity table unallocated.  Its random draws come from its
// own generator, so only cases that do not depend on them (forced resample
// with bias 1, non-motile cells) are comparable with another stream.
#include "motility.h"

extern "C" int ref_motility(uint32_t n, const float* x, const float* y, const float* z, float* vx, float* vy,
                            float* vz, const uint32_t* alive, const float* speed, uint32_t nx, uint32_t ny,
                            uint32_t nz, uint32_t ns, float dx, float dy, float dz, float x_min, float y_min,
                            float z_min, const float* density, double dt, double* persistence, double* elapsed,
                            const double* bias, double* mvx, double* mvy, double* mvz, const uint32_t* motile,
                            const uint32_t* restrict_2d, const int32_t* chem_index, const int32_t* chem_dir,
                            const double* sensitivity, int legacy) {
    CellData cells(n > 0 ? n : 1);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    cells.num_cells = n;
    MotilityData mot;
    mot.allocate(n > 0 ? n : 1);
    if (!legacy) mot.allocateSubstrate(ns);
    for (uint32_t i = 0; i < n; i++) {
        cells.position_x[i] = x[i]; cells.position_y[i] = y[i]; cells.position_z[i] = z[i];
        cells.velocity_x[i] = vx[i]; cells.velocity_y[i] = vy[i]; cells.velocity_z[i] = vz[i];
        cells.is_alive[i] = alive[i];
        cells.motility_speed[i] = speed[i];
        mot.persistence_time[i] = persistence[i];
        mot.motility_elapsed[i] = elapsed[i];
        mot.motility_bias[i] = bias[i];
        mot.motility_vec_x[i] = mvx[i]; mot.motility_vec_y[i] = mvy[i]; mot.motility_vec_z[i] = mvz[i];
        mot.is_motile[i] = motile[i];
        mot.restrict_to_2D[i] = restrict_2d[i];
        mot.chemotaxis_index[i] = chem_index[i];
        mot.chemotaxis_direction[i] = chem_dir[i];
        if (!legacy)
            for (uint32_t s = 0; s < ns; s++)
                mot.chemotactic_sensitivity[(size_t)i * ns + s] = sensitivity[(size_t)i * ns + s];
    }
    GridParams gp{};
    gp.nx = nx; gp.ny = ny; gp.nz = nz; gp.n_substrates = ns;
    gp.dx = dx; gp.dy = dy; gp.dz = dz;
    gp.x_min = x_min; gp.y_min = y_min; gp.z_min = z_min;
    updateMotility(cells, mot, density, gp, dt);
    for (uint32_t i = 0; i < n; i++) {
        vx[i] = cells.velocity_x[i]; vy[i] = cells.velocity_y[i]; vz[i] = cells.velocity_z[i];
        elapsed[i] = mot.motility_elapsed[i];
        mvx[i] = mot.motility_vec_x[i]; mvy[i] = mot.motility_vec_y[i]; mvz[i] = mot.motility_vec_z[i];
    }
    mot.free_all();
    return 0;
}
