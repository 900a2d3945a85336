// ref_driver.cpp -- thin C entry points into the reference's own compiled
// processSecretion (src/cell_phenotype.cpp:526-564), updateMotility
// (src/motility.cpp:179-293), updateVolume / divideCell / removeCell
// (src/cell_phenotype.cpp, src/cell_data.cpp), tryFormAllAttachments
// (src/cell_interactions.cpp:459-478) and .mat writers
// (src/matlab_writer.cpp, src/output_writer.cpp:295-453), used only by tests
// to pin the FP64 oracle's secretion restatement and the library's snapshot
// writer against the reference code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/build_ref.sh into oracle/_ref/
// from the reference sources where they lie; nothing here is product code.
#include "cell_data.h"
#include "cell_interactions.h"
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

// The reference's phenotype-side SoA mutation on a cell store filled from the
// library's field layout (f64: 23 arrays of max_cells in HC_F_* order, stored
// in the reference's FP32 fields; u32: 5 arrays; vol: 15 FP64 arrays in HC_V_*
// order; rates: 4 arrays [cell*ns + s]).  op 0: updateVolume
// (src/cell_phenotype.cpp:52-113) on every cell with list[i] != 0 (list NULL:
// all); use_ratio 0 runs it without the cyto/nuclear ratio array.  op 1:
// divideCell (:571-720) for each parent in list.  op 2: the removal loop
// (:833-844), removeCell in descending order.  Every field is written back
// (FP32 ones widened).  *n_out = the resulting cell count.
#include "cell_phenotype.h"
#include <algorithm>
#include <functional>

extern "C" int ref_phenotype_op(int op, uint32_t n, uint32_t max_cells, uint32_t ns, double* f64, uint32_t* u32,
                                double* vol, double* rates, const uint32_t* list, uint32_t count, double dt,
                                int use_ratio, uint32_t* n_out) {
    CellData cells(max_cells);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    cells.allocateCPUBuffers();
    cells.allocateSubstrateBuffers(ns);
    float* F[23] = {cells.position_x, cells.position_y, cells.position_z, cells.velocity_x, cells.velocity_y,
                    cells.velocity_z, cells.prev_velocity_x, cells.prev_velocity_y, cells.prev_velocity_z,
                    cells.radius, cells.nuclear_radius, cells.total_volume, cells.nuclear_volume,
                    cells.cell_cell_repulsion, cells.cell_cell_adhesion, cells.relative_max_adhesion_distance,
                    cells.motility_speed, cells.motility_bias, cells.motility_dir_x, cells.motility_dir_y,
                    cells.motility_dir_z, cells.oncoprotein, cells.simple_pressure};
    uint32_t* U[5] = {cells.cell_type, cells.current_phase, cells.is_alive, cells.voxel_index, cells.mech_voxel_index};
    double* V[15] = {cells.fluid, cells.solid_nuclear, cells.solid_cytoplasmic, cells.fluid_fraction,
                     cells.calcified_fraction, cells.target_solid_cytoplasmic, cells.target_solid_nuclear,
                     cells.target_fluid_fraction, cells.target_cyto_to_nuclear_ratio, cells.fluid_change_rate,
                     cells.nuclear_biomass_change_rate, cells.cytoplasmic_biomass_change_rate,
                     cells.calcification_rate, cells.rupture_volume, cells.relative_rupture_volume};
    double* R[4] = {cells.secretion_rate, cells.uptake_rate, cells.saturation_density, cells.net_export_rate};
    for (uint32_t i = 0; i < n; i++) cells.addCell();
    for (uint32_t i = 0; i < max_cells; i++) {
        for (int f = 0; f < 23; f++) F[f][i] = (float)f64[(size_t)f * max_cells + i];
        for (int f = 0; f < 5; f++) U[f][i] = u32[(size_t)f * max_cells + i];
        for (int f = 0; f < 15; f++) V[f][i] = vol[(size_t)f * max_cells + i];
        for (int a = 0; a < 4; a++)
            for (uint32_t s = 0; s < ns; s++) R[a][(size_t)i * ns + s] = rates[(size_t)a * max_cells * ns + (size_t)i * ns + s];
    }
    if (op == 0) {
        double* keep = cells.target_cyto_to_nuclear_ratio;
        if (!use_ratio) cells.target_cyto_to_nuclear_ratio = nullptr;
        for (uint32_t i = 0; i < n; i++)
            if (!list || list[i]) updateVolume(cells, i, dt);
        cells.target_cyto_to_nuclear_ratio = keep;
    } else if (op == 1) {
        for (uint32_t k = 0; k < count; k++) divideCell(cells, list[k]);
    } else {
        std::vector<uint32_t> idx(list, list + count);
        std::sort(idx.begin(), idx.end(), std::greater<uint32_t>());
        for (uint32_t i : idx)
            if (i < cells.num_cells) cells.removeCell(i);
    }
    for (uint32_t i = 0; i < max_cells; i++) {
        for (int f = 0; f < 23; f++) f64[(size_t)f * max_cells + i] = (double)F[f][i];
        for (int f = 0; f < 5; f++) u32[(size_t)f * max_cells + i] = U[f][i];
        for (int f = 0; f < 15; f++) vol[(size_t)f * max_cells + i] = V[f][i];
        for (int a = 0; a < 4; a++)
            for (uint32_t s = 0; s < ns; s++) rates[(size_t)a * max_cells * ns + (size_t)i * ns + s] = R[a][(size_t)i * ns + s];
    }
    *n_out = cells.num_cells;
    return 0;
}

// The reference's writeCellNeighborGraph (src/output_writer.cpp:457-490) on n
// cells (FP32 positions, radii, volumes and adhesion scales, as its store
// holds them); writes <folder>/final_cell_neighbor_graph.txt.
extern "C" int ref_write_neighbor_graph(const char* folder, uint32_t n, const float* x, const float* y, const float* z,
                                        const float* radius, const float* volume, const float* rel,
                                        const uint32_t* alive) {
    CellData cells(n > 0 ? n : 1);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    for (uint32_t i = 0; i < n; i++) cells.addCell();
    for (uint32_t i = 0; i < n; i++) {
        cells.position_x[i] = x[i]; cells.position_y[i] = y[i]; cells.position_z[i] = z[i];
        cells.radius[i] = radius[i];
        cells.total_volume[i] = volume[i];
        cells.relative_max_adhesion_distance[i] = rel[i];
        cells.is_alive[i] = alive[i];
    }
    writeCellNeighborGraph(cells, folder);
    return 0;
}

// The reference's tryFormAllAttachments (src/cell_interactions.cpp:459-478)
// on n cells (FP32 positions, radii, adhesion scales and rates, as its store
// holds them) from empty attachment lists; count[n] and targets[n * 12] out.
// Its draws come from an unseeded generator, so only rates with rate * dt >= 1
// (every draw accepted) give a repeatable result.
extern "C" int ref_form_attachments(uint32_t n, const float* x, const float* y, const float* z,
                                    const float* radius, const float* rel, const uint32_t* alive,
                                    const float* rate, double dt, uint32_t* count, uint32_t* targets) {
    CellData cells(n > 0 ? n : 1);
    std::vector<float> gpu(cells.gpuBufferSize() / sizeof(float), 0.0f);
    cells.setGPUBufferBase(gpu.data());
    cells.allocateCPUBuffers();                 // the attachment lists (src/main.mm:170)
    for (uint32_t i = 0; i < n; i++) cells.addCell();
    for (uint32_t i = 0; i < n; i++) {
        cells.position_x[i] = x[i]; cells.position_y[i] = y[i]; cells.position_z[i] = z[i];
        cells.radius[i] = radius[i];
        cells.relative_max_adhesion_distance[i] = rel[i];
        cells.is_alive[i] = alive[i];
        cells.attachment_rate[i] = rate[i];
        cells.attachment_count[i] = 0;
    }
    tryFormAllAttachments(cells, dt);
    for (uint32_t i = 0; i < n; i++) {
        count[i] = cells.attachment_count[i];
        for (uint32_t a = 0; a < CellData::MAX_ATTACHMENTS; a++)
            targets[(size_t)i * CellData::MAX_ATTACHMENTS + a] =
                a < count[i] ? cells.attachment_targets[(size_t)i * CellData::MAX_ATTACHMENTS + a] : 0u;
    }
    return 0;
}
