// snapshot.cpp -- MATLAB Level-4 snapshots in the reference's output format.
//
// File format (src/matlab_writer.cpp:5-49): five little-endian uint32 --
// type 0 (little-endian, double, full matrix), rows, cols, imaginary flag 0,
// name length including the terminating NUL -- then the name, then
// rows*cols doubles in column-major order.
//
// Snapshot files (src/output_writer.cpp:295-453): output%08d_cells.mat holds
// one column per cell with the reference's row schema; the path's fields
// (position, volumes, radii, velocity, pressure, mechanics parameters, type,
// phase, alive flag, per-substrate rates) come from the snapshot, and every
// other row holds the value a cell of the reference has from its placement
// (CellData::setDefaults, src/cell_data.cpp:544-680), since nothing on this
// path changes them.  One cell type, no custom variables.  The values are
// written in FP64; the one derived row, cytoplasmic volume (total - nuclear),
// is subtracted in FP64 where the reference subtracts in FP32.
// output%08d_microenvironment0.mat holds one column per voxel: its centre,
// volume (both computed in FP32 from the grid, as the reference computes
// them) and the substrate densities.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "physicell_hip.h"

namespace {

bool write_mat4(const char* path, const char* name, uint32_t rows, uint32_t cols, const double* data) {
    if (!path || !name || rows == 0 || cols == 0 || !data) return false;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return false;
    const uint32_t name_len = (uint32_t)std::strlen(name) + 1;
    const uint32_t head[5] = {0u, rows, cols, 0u, name_len};
    bool ok = std::fwrite(head, sizeof(uint32_t), 5, fp) == 5;
    ok = ok && std::fwrite(name, 1, name_len, fp) == name_len;
    ok = ok && std::fwrite(data, sizeof(double), (size_t)rows * cols, fp) == (size_t)rows * cols;
    ok = (std::fclose(fp) == 0) && ok;
    return ok;
}

// Values a freshly placed reference cell holds (src/cell_data.cpp:544-680)
// for the rows of the cells matrix that this path does not own.
struct CellDefaults {
    static constexpr double cycle_model = 5.0, elapsed = 0.0, fluid_fraction = 0.75, calcified = 0.0;
    static constexpr double rate_01 = 0.00072, death_model = 0.0, necrosis_rate = 0.0;
    static constexpr double cyto_biomass_rate = 0.27 / 60.0, nuclear_biomass_rate = 0.33 / 60.0;
    static constexpr double fluid_change_rate = 3.0 / 60.0, calcification_rate = 0.0;
    // volumes of the default cell: total 2494, nuclear 540, fluid fraction 0.75
    static constexpr double target_solid_cyto = (2494.0 - 540.0) - 0.75 * (2494.0 - 540.0);
    static constexpr double target_solid_nuclear = 540.0 - 0.75 * 540.0;
    static constexpr double target_fluid_fraction = 0.75;
    static constexpr double max_attachments = 12.0, persistence_time = 15.0;
    static constexpr float elastic_constant = 0.01f, attach_rate = 0.0f, detach_rate = 0.0f;
    static constexpr float motility_speed = 1.0f, motility_bias = 0.5f;
    static constexpr double attacking_cell = 4294967295.0;   // no attacker (UINT32_MAX)
};

}  // namespace

extern "C" {

int pc_write_matlab4(const char* path, const char* name, uint32_t rows, uint32_t cols, const double* data) {
    return write_mat4(path, name, rows, cols, data) ? 0 : HC_ERR_ARG;
}

int pc_write_snapshot(const char* folder, int frame, const pc_snapshot* s) {
    if (!folder || !s) return HC_ERR_ARG;
    char path[4096];
    const uint32_t n = s->n_cells, ns = s->n_substrates;
    if (n > 0) {
        if (!s->x || !s->y || !s->z || !s->total_volume || !s->nuclear_volume || !s->radius || !s->nuclear_radius)
            return HC_ERR_ARG;
        std::vector<double> col, mat;
        for (uint32_t i = 0; i < n; i++) {
            col.clear();
            auto put = [&](double v) { col.push_back(v); };
            auto rep = [&](uint32_t k, double v) { for (uint32_t q = 0; q < k; q++) col.push_back(v); };
            auto opt = [&](const double* a, double dflt) { return a ? a[i] : dflt; };
            auto rate = [&](const double* a, uint32_t q, double dflt) { return a ? a[(size_t)i * ns + q] : dflt; };
            using D = CellDefaults;
            const double vol = s->total_volume[i], nvol = s->nuclear_volume[i];
            // identity, position, volume, cycle (17 compatibility entries)
            put((double)i); put(s->x[i]); put(s->y[i]); put(s->z[i]); put(vol);
            put(s->cell_type ? (double)s->cell_type[i] : 0.0); put(D::cycle_model);
            put(s->current_phase ? (double)s->current_phase[i] : 0.0); put(D::elapsed);
            put(nvol); put(vol - nvol); put(D::fluid_fraction); put(D::calcified);
            rep(4, 0.0);                                       // orientation, polarity
            // state
            put(opt(s->vx, 0.0)); put(opt(s->vy, 0.0)); put(opt(s->vz, 0.0)); put(opt(s->pressure, 0.0));
            put(1.0); put(0.0); put(0.0);                      // nuclei, attack time, contact with BM
            // cycle and death
            put(D::rate_01); put(D::elapsed);
            put(s->is_alive ? (s->is_alive[i] ? 0.0 : 1.0) : 0.0);
            put(D::death_model); put(0.0); put(D::necrosis_rate);
            // volume model
            put(D::cyto_biomass_rate); put(D::nuclear_biomass_rate); put(D::fluid_change_rate);
            put(D::calcification_rate); put(D::target_solid_cyto); put(D::target_solid_nuclear);
            put(D::target_fluid_fraction);
            // geometry: radius, nuclear radius, surface area (FP32, as the reference)
            const float rf = (float)s->radius[i];
            put(s->radius[i]); put(s->nuclear_radius[i]); put((double)(4.0f * 3.14159f * rf * rf));
            // mechanics
            put(opt(s->adhesion, 0.0)); put(0.0); put(opt(s->repulsion, 0.0)); put(0.0);
            put(1.0);                                          // adhesion affinity (one type)
            put(opt(s->rel_max_adhesion, 0.0)); put(D::max_attachments);
            put((double)D::elastic_constant); put((double)D::attach_rate); put((double)D::detach_rate);
            // motility
            put(D::motility_speed > 0.0f ? 1.0 : 0.0); put(D::persistence_time); put((double)D::motility_speed);
            rep(3, 0.0); put((double)D::motility_bias); rep(3, 0.0); put(0.0); put(1.0);
            rep(ns, 0.0);                                      // chemotactic sensitivities
            // secretion
            for (uint32_t q = 0; q < ns; q++) put(rate(s->secretion, q, 0.0));
            for (uint32_t q = 0; q < ns; q++) put(rate(s->uptake, q, 0.0));
            for (uint32_t q = 0; q < ns; q++) put(rate(s->saturation, q, 1.0));
            for (uint32_t q = 0; q < ns; q++) put(rate(s->net_export, q, 0.0));
            rep(3 * ns, 0.0);                                  // molecular
            // interactions (one type)
            rep(3, 0.0); put(0.0); put(0.0); put(1.0);
            put(D::attacking_cell); put(0.0); put(0.0); put(0.0);
            put(0.0); put(0.0); put(0.0);
            // integrity
            rep(3, 0.0);
            if (i == 0) mat.reserve((size_t)col.size() * n);
            else if (col.size() * (size_t)(i) != mat.size()) return HC_ERR_STATE;
            mat.insert(mat.end(), col.begin(), col.end());
        }
        std::snprintf(path, sizeof(path), "%s/output%08d_cells.mat", folder, frame);
        if (!write_mat4(path, "cells", (uint32_t)col.size(), n, mat.data())) return HC_ERR_ARG;
    }
    const uint64_t V = (uint64_t)s->nx * s->ny * (s->nz ? s->nz : 1u);
    if (V > 0 && s->density) {
        const uint32_t rows = 4 + ns;
        std::vector<double> mat((size_t)V * rows);
        const float dx = (float)s->dx, dy = (float)s->dy, dz = (float)s->dz;
        const float x0 = (float)s->x_min, y0 = (float)s->y_min, z0 = (float)s->z_min;
        const float vvol = dx * dy * dz;
        for (uint64_t v = 0; v < V; v++) {
            const uint32_t iz = (uint32_t)(v / ((uint64_t)s->nx * s->ny));
            const uint32_t rem = (uint32_t)(v % ((uint64_t)s->nx * s->ny));
            const uint32_t iy = rem / s->nx, ix = rem % s->nx;
            double* c = &mat[(size_t)v * rows];
            c[0] = (double)(x0 + ((float)ix + 0.5f) * dx);
            c[1] = (double)(y0 + ((float)iy + 0.5f) * dy);
            c[2] = (double)(z0 + ((float)iz + 0.5f) * dz);
            c[3] = (double)vvol;
            for (uint32_t q = 0; q < ns; q++) c[4 + q] = s->density[(size_t)q * V + v];
        }
        std::snprintf(path, sizeof(path), "%s/output%08d_microenvironment0.mat", folder, frame);
        if (!write_mat4(path, "multiscale_microenvironment", rows, (uint32_t)V, mat.data())) return HC_ERR_ARG;
    }
    return 0;
}

int pc_write_neighbor_graph(const char* path, uint32_t n, const uint32_t* offsets, const uint32_t* idx) {
    if (!path || !offsets || (offsets[n] && !idx)) return HC_ERR_ARG;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return HC_ERR_ARG;
    // the reference streams "i j\n" lines with std::ofstream (decimal, no padding)
    std::vector<char> buf;
    buf.reserve(1 << 20);
    char line[32];
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = offsets[i]; k < offsets[i + 1]; k++) {
            const int len = std::snprintf(line, sizeof(line), "%u %u\n", i, idx[k]);
            buf.insert(buf.end(), line, line + len);
            if (buf.size() >= (1u << 20)) {
                std::fwrite(buf.data(), 1, buf.size(), fp);
                buf.clear();
            }
        }
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    return std::fclose(fp) == 0 && ok ? 0 : HC_ERR_ARG;
}

}  // extern "C"
