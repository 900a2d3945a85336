// physicell_host.h -- plain C++ host API of the step hot path.
//
// The classes keep the reference's names and method meanings
// (Microenvironment: src/microenvironment.h:17-83, CellMechanics:
// src/cell_mechanics.h:15-87, CellData: src/cell_data.h:18-242) and sit on the
// hc_* C-ABI (include/physicell_hip.h) instead of MetalContext.  Memory is
// discrete: each class owns device buffers plus host mirrors, and the
// sync*() calls are the explicit transfer points the reference's shared
// memory did not need.  Errors throw std::runtime_error (the reference prints
// and aborts).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "physicell_hip.h"

namespace physicell {

// Subset of SimConfig (src/config_parser.h) read by the path.
struct SubstrateConfig {
    double diffusion_coeff = 0, decay_rate = 0, initial_value = 0, dirichlet_value = 0;
    bool dirichlet_enabled = false;
};

struct SimConfig {
    double x_min = -500, x_max = 500, y_min = -500, y_max = 500, z_min = -10, z_max = 10;
    double dx = 20, dy = 20, dz = 20;
    bool use_2D = true;
    double dt_diffusion = 0.01, dt_mechanics = 0.1, dt_phenotype = 6.0, max_time = 60.0;
    uint32_t max_cells = 1;
    bool mixed_precision = false;      // FP32-mixed form
    std::vector<SubstrateConfig> substrates;
};

void check(hc_ctx* ctx, int rc, const char* what);

// RAII owner of one hc_buf.
// The greedy pass of tryFormAllAttachments (src/cell_interactions.cpp:
// 411-478) over attachment candidates (HC_QUERY_ATTACH lists off[n + 1],
// idx): i ascending with rate[i] > 0 and fewer than 12 partners, its
// candidates j ascending with rate[j] > 0; skip when either list is full or
// j is already i's partner, else draw u (std::mt19937(seed), uniform [0, 1))
// and attach both ways when u < max(rate_i, rate_j) dt.  Returns the number
// of attachments formed.
uint32_t attach_greedy(uint32_t n, const std::vector<uint32_t>& off, const std::vector<uint32_t>& idx, double dt,
                       const double* rate, uint32_t* count, uint32_t* targets, uint64_t seed);

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(hc_ctx* ctx, size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept;
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    void allocate(hc_ctx* ctx, size_t bytes);
    hc_buf* get() const { return buf_; }
    void* host();
    size_t size() const { return hc_buf_size(buf_); }
    void upload(size_t off, size_t n);
    void download(size_t off, size_t n);
    void upload_all() { upload(0, size()); }
    void download_all() { download(0, size()); }

private:
    hc_ctx* ctx_ = nullptr;
    hc_buf* buf_ = nullptr;
};

// Cell store: the 28-field SoA (fields 0..22 FP64, 23..27 uint32) on the
// device plus host views of its mirror, and the per-substrate rate arrays.
class CellData {
public:
    explicit CellData(uint32_t max_cells);
    void attach(hc_ctx* ctx, uint32_t n_substrates);

    // host views into the mirror (valid after syncToHost, pushed by syncToDevice)
    double *position_x = nullptr, *position_y = nullptr, *position_z = nullptr;
    double *velocity_x = nullptr, *velocity_y = nullptr, *velocity_z = nullptr;
    double *prev_velocity_x = nullptr, *prev_velocity_y = nullptr, *prev_velocity_z = nullptr;
    double *radius = nullptr, *nuclear_radius = nullptr, *total_volume = nullptr, *nuclear_volume = nullptr;
    double *cell_cell_repulsion = nullptr, *cell_cell_adhesion = nullptr;
    double *relative_max_adhesion_distance = nullptr;
    double *simple_pressure = nullptr;
    uint32_t *cell_type = nullptr, *current_phase = nullptr, *is_alive = nullptr;
    uint32_t *voxel_index = nullptr, *mech_voxel_index = nullptr;
    // [cell * n_substrates + s]
    double *secretion_rate = nullptr, *uptake_rate = nullptr, *saturation_density = nullptr,
           *net_export_rate = nullptr;

    uint32_t num_cells = 0;
    uint32_t max_cells;
    uint32_t n_substrates = 0;

    double* field(uint32_t f);
    uint32_t* ufield(uint32_t f);

    // CellData::addCell / removeCell (src/cell_data.cpp:398-542): append a
    // cell with the reference's fresh-cell values (setDefaults: volumes,
    // radii, mechanics, motility, rates), and swap-with-last removal.  Host
    // side; push with syncToDevice.
    uint32_t addCell();
    void removeCell(uint32_t index);

    void syncToDevice();          // SoA + rates
    void syncToHost();            // SoA (rates are host-authored)
    void syncFieldToHost(uint32_t f);

    hc_buf* gpu_buffer() const { return soa_.get(); }
    hc_buf* rates_buffer() const { return rates_.get(); }

private:
    hc_ctx* ctx_ = nullptr;
    DeviceBuffer soa_, rates_;
    void bind_views();
};

class Microenvironment {
public:
    void initialize(const SimConfig& config, hc_ctx* ctx);
    // x, y, [z] sweeps; the Dirichlet re-apply is fused into the last sweep
    void simulateDiffusionDecay();
    // all faces; idempotent after simulateDiffusionDecay
    void applyDirichletBC();
    // processSecretion for every live cell (plan built by planSecretion)
    void applySecretion();
    void planSecretion(CellData& cells, double dt);
    // cell -> voxel refresh with the 3-D formula (the reference refreshes in
    // updatePhenotype, src/cell_phenotype.cpp:778-789)
    void refreshVoxelIndices(CellData& cells);

    double getDensity(uint32_t voxel_index, uint32_t substrate_index) const;
    void setDensity(uint32_t voxel_index, uint32_t substrate_index, double value);
    uint32_t numVoxels() const { return grid_.nx * grid_.ny * grid_.nz; }
    uint32_t voxelIndex(double x, double y, double z) const;
    hc_grid getGridParams() const { return grid_; }
    const std::vector<hc_substrate>& substrateParams() const { return subs_; }

    double* densityBuffer();      // FP64 host copy, refreshed from the device
    void totals(double* out);     // compensated per-substrate totals
    void syncToHost();
    void syncToDevice();
    hc_buf* density() const { return density_.get(); }
    hc_buf* tables() const { return tables_.get(); }
    hc_buf* plan() const { return plan_.get(); }

private:
    hc_ctx* ctx_ = nullptr;
    hc_grid grid_{};
    std::vector<hc_substrate> subs_;
    DeviceBuffer density_, tables_, plan_;
    std::vector<double> host_density_;   // FP64 view (the device field may be FP32)
    uint32_t plan_capacity_ = 0;
};

class CellMechanics {
public:
    void initialize(const SimConfig& config, hc_ctx* ctx, uint32_t max_cells);
    // dispatchMechanicsPipeline: hash -> forces -> AB2
    void update(CellData& cells, double dt);
    // dispatchForcesOnlyPipeline: hash -> forces
    void computeForces(CellData& cells, double dt);
    // dispatchIntegrate
    void integratePositions(CellData& cells, double dt);

    uint32_t totalMechVoxels() const { return total_mech_voxels_; }
    const hc_mech& params() const { return params_; }
    const hc_grid& mechGrid() const { return mgrid_; }
    uint32_t voxelIndexFromPosition(double px, double py, double pz) const;
    // hash accessors (valid after syncHashPointers)
    uint32_t voxelCellCount(uint32_t voxel_idx) const;
    const uint32_t* voxelCells(uint32_t voxel_idx) const;
    void syncHashPointers(uint32_t num_cells);
    uint32_t maxVoxelOccupancy();
    const std::vector<uint32_t>& hostOffsets() const { return host_offsets_; }
    hc_buf* hash() const { return hash_.get(); }
    // hash from the current positions, then hc_neighbor_query (kind
    // HC_QUERY_*); copies out offsets[num_cells + 1] and, when idx is not null
    // and capacity suffices, the lists.  Returns the number of entries.
    uint64_t neighborQuery(CellData& cells, int kind, uint32_t* offsets, uint32_t* idx, uint64_t capacity);
    // the same into host vectors sized to the result
    uint64_t neighborQuery(CellData& cells, int kind, std::vector<uint32_t>& offsets, std::vector<uint32_t>& idx);
    // Hash-ordered mechanics mirror (FP64 forces; hc::mirror_*): loadMirror
    // builds it from the cell fields, mirrorStep runs one mechanics step on it
    // (hash -> forces -> AB2, the cell fields untouched), syncMirror writes
    // positions, velocities, previous velocities, pressure and mechanics
    // voxels back into the cell fields.
    void loadMirror(CellData& cells);
    void mirrorStep(CellData& cells, double dt);
    void syncMirror(CellData& cells);
    void resortMirror(CellData& cells);

private:
    hc_ctx* ctx_ = nullptr;
    hc_mech params_{};
    hc_grid mgrid_{};
    uint32_t total_mech_voxels_ = 0, max_cells_ = 0;
    DeviceBuffer hash_, q_off_, q_idx_, mirror_;
    std::vector<uint32_t> host_offsets_, host_cells_;
};

// The step loop of src/main.mm:275-415 restricted to the path.
class Simulation {
public:
    Simulation(const SimConfig& config, int device);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void setCells(uint32_t n, const double* x, const double* y, const double* z, const double* radius,
                  const double* volume, const double* rep, const double* adh, const double* rel_adh,
                  const double* sec, const double* upt, const double* sat, const double* net,
                  const uint32_t* alive);
    void diffusionSubstep();
    void mechanicsStep();
    void phenotypeStep();
    void run(double max_time, int64_t counts[3]);
    // periods x (substeps + one mechanics step); with overlap on (default
    // off), the mechanics step of a period runs on a second stream
    // concurrently with its substeps (they touch disjoint data; same results).  The
    // phenotype step (voxel refresh + secretion re-plan) fires every
    // dt_phenotype as in run(), on a clock that persists across calls and
    // restarts when the cells are set.
    void runPeriods(int periods, bool use_graph);
    void setOverlap(bool on);
    // motility in the mechanics step (forces -> motility -> AB2); per-cell
    // arrays as MotilityData (src/motility.h); sensitivity may be null
    void setMotility(uint64_t seed, const double* persistence, const double* bias, const double* speed,
                     const uint32_t* is_motile, const uint32_t* restrict_2d, const int32_t* chem_index,
                     const int32_t* chem_dir, const double* sensitivity);
    // Volume state (HC_V_* fields: one array of num_cells values each, or
    // null for CellData's fresh value, src/cell_data.cpp:597-624) and the seed
    // of division's draws.  Required before updatePhenotypeSoA.
    void setVolume(uint64_t seed, const double* const* fields);
    // updatePhenotype's SoA side on the device (src/cell_phenotype.cpp:
    // 792-845): the volume model for cells with mask[i] != 0 (mask: num_cells
    // values, or null for all), then the divisions of `parents` (daughters
    // appended), then the removal of `removed` (indices below the count after
    // the divisions); voxel keys and the secretion plan follow.  Returns the
    // new cell count.
    uint32_t updatePhenotypeSoA(double dt, int use_ratio, const uint32_t* mask, const uint32_t* parents,
                                uint32_t n_parents, const uint32_t* removed, uint32_t n_removed);
    // tryFormAllAttachments (src/cell_interactions.cpp:459-478): candidate
    // pairs from the device hash (HC_QUERY_ATTACH), then the reference's
    // greedy pass in its order on the host arrays rate / count / targets
    // (num_cells, num_cells, num_cells x 12), drawing from std::mt19937(seed).
    // Returns the number of attachments formed.
    uint32_t formAttachments(double dt, const double* rate, uint32_t* count, uint32_t* targets, uint64_t seed);
    // The FP64 mechanics step without motility runs on the hash-ordered
    // mirror (on by default); the cell fields it changes are written back
    // before anything else here reads them (the phenotype step, field reads,
    // snapshots, neighbour queries, attachments) and by syncMirror.  Same
    // results either way.
    void setMechanicsMirror(bool on);
    // writes the mirror back into the cell fields when they are behind it
    void syncMirror();
    void wait();
    int substepsPerMechanicsStep() const;

    hc_ctx* ctx = nullptr;
    SimConfig config;
    CellData cells;
    Microenvironment microenv;
    CellMechanics mechanics;
    std::string last_error;

private:
    void periodBody();
    void launchPeriodGraph();
    std::vector<int> advancePhenotypeClock(int k);
    void periodWithPhenotype(const std::vector<int>& fires);
    double pheno_acc_ = 0.0;       // runPeriods' phenotype accumulator (src/main.mm:394-412)
    void* graph_exec_ = nullptr;   // hipGraphExec_t of one period
    void* side_ = nullptr;         // hipStream_t for the overlapped mechanics step
    void* fork_ = nullptr;         // hipEvent_t
    void* join_ = nullptr;         // hipEvent_t
    bool overlap_ = false;
    bool useMirror() const;
    // 0: the cell fields hold the mechanics state (the mirror is not built);
    // 1: the mirror does and the fields are behind; 2: both agree
    int mirror_state_ = 0;
    // mechanics steps since the mirror's slots were put in record order
    // (re-sorted every MIRROR_RESORT_STEPS steps, between periods)
    int mirror_steps_ = 0;
    static constexpr int MIRROR_RESORT_STEPS = 64;
    void mirrorResortIfDue();
    bool mirror_on_ = true;
    DeviceBuffer motility_;
    DeviceBuffer volume_, mask_;
    bool motility_on_ = false;
    int motility_use_sensitivity_ = 0;
};

}  // namespace physicell
