// ref_driver.cpp -- thin C entry points into the reference's own compiled
// processSecretion (src/cell_phenotype.cpp:526-564) and .mat writers
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
