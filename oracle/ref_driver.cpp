// ref_driver.cpp -- a thin C entry point into the reference's own compiled
// processSecretion (src/cell_phenotype.cpp:526-564), used only by tests to
// pin the FP64 oracle's secretion restatement against the reference code.
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
