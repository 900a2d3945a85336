// physicell_host.cpp -- host classes (Microenvironment, CellMechanics,
// CellData, Simulation) over the hc_* C-ABI, and the pc_* C entry points.
#include "physicell_host.h"
#include "hc_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

// the context's stream (internal export of context.hip; not part of the C-ABI header)
extern "C" hipStream_t hc_stream_of(hc_ctx* ctx);
extern "C" hipStream_t hc_stream_swap(hc_ctx* ctx, hipStream_t s);

namespace physicell {

void check(hc_ctx* ctx, int rc, const char* what) {
    if (rc != 0) {
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) +
                                 "): " + (ctx ? hc_last_error(ctx) : ""));
    }
}

// ---------------------------------------------------------------- DeviceBuffer
DeviceBuffer::DeviceBuffer(hc_ctx* ctx, size_t bytes) { allocate(ctx, bytes); }
DeviceBuffer::~DeviceBuffer() { hc_free(buf_); }
DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : ctx_(o.ctx_), buf_(o.buf_) { o.buf_ = nullptr; }
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        hc_free(buf_);
        ctx_ = o.ctx_;
        buf_ = o.buf_;
        o.buf_ = nullptr;
    }
    return *this;
}
void DeviceBuffer::allocate(hc_ctx* ctx, size_t bytes) {
    hc_free(buf_);
    buf_ = nullptr;
    ctx_ = ctx;
    check(ctx, hc_alloc(ctx, bytes, &buf_), "allocateBuffer");
}
void* DeviceBuffer::host() {
    void* h = hc_host_mirror(buf_);
    if (!h && size()) throw std::runtime_error("host mirror allocation failed");
    return h;
}
void DeviceBuffer::upload(size_t off, size_t n) { check(ctx_, hc_upload(ctx_, buf_, off, n), "upload"); }
void DeviceBuffer::download(size_t off, size_t n) { check(ctx_, hc_download(ctx_, buf_, off, n), "download"); }

// -------------------------------------------------------------------- CellData
CellData::CellData(uint32_t max_cells_) : max_cells(max_cells_ ? max_cells_ : 1) {}

void CellData::attach(hc_ctx* ctx, uint32_t ns) {
    ctx_ = ctx;
    n_substrates = ns;
    soa_.allocate(ctx, hc_cells_bytes(max_cells));
    rates_.allocate(ctx, std::max<size_t>(8, hc_rates_bytes(max_cells, ns)));
    bind_views();
}

double* CellData::field(uint32_t f) {
    return (double*)((char*)soa_.host() + hc_cells_field_offset(f, max_cells));
}
uint32_t* CellData::ufield(uint32_t f) {
    return (uint32_t*)((char*)soa_.host() + hc_cells_field_offset(f, max_cells));
}

void CellData::bind_views() {
    position_x = field(HC_F_POS_X); position_y = field(HC_F_POS_Y); position_z = field(HC_F_POS_Z);
    velocity_x = field(HC_F_VEL_X); velocity_y = field(HC_F_VEL_Y); velocity_z = field(HC_F_VEL_Z);
    prev_velocity_x = field(HC_F_PREV_VX); prev_velocity_y = field(HC_F_PREV_VY); prev_velocity_z = field(HC_F_PREV_VZ);
    radius = field(HC_F_RADIUS); nuclear_radius = field(HC_F_NUCLEAR_RAD);
    total_volume = field(HC_F_TOTAL_VOL); nuclear_volume = field(HC_F_NUCLEAR_VOL);
    cell_cell_repulsion = field(HC_F_REPULSION); cell_cell_adhesion = field(HC_F_ADHESION);
    relative_max_adhesion_distance = field(HC_F_REL_MAX_ADH);
    simple_pressure = field(HC_F_PRESSURE);
    cell_type = ufield(HC_F_CELL_TYPE); current_phase = ufield(HC_F_CUR_PHASE);
    is_alive = ufield(HC_F_IS_ALIVE); voxel_index = ufield(HC_F_VOXEL_IDX);
    mech_voxel_index = ufield(HC_F_MECH_VOXEL);
    double* r = (double*)rates_.host();
    const size_t stride = (size_t)max_cells * n_substrates;
    secretion_rate = r;
    uptake_rate = r + stride;
    saturation_density = r + 2 * stride;
    net_export_rate = r + 3 * stride;
}

uint32_t CellData::addCell() {
    if (num_cells >= max_cells) return UINT32_MAX;   // src/cell_data.cpp:398-401
    const uint32_t i = num_cells++;
    for (uint32_t f = 0; f < HC_NUM_F64_FIELDS; f++) field(f)[i] = 0.0;
    for (uint32_t f = HC_NUM_F64_FIELDS; f < HC_NUM_CELL_FIELDS; f++) ufield(f)[i] = 0;
    // the reference's fresh-cell values (src/cell_data.cpp:544-680): MCF-7
    // volumes 2494 / 540 um^3 and their equivalent-sphere radii, mechanics,
    // motility (speed 1, bias 0.5) and per-substrate rates; computed in FP64
    const double pi = 3.14159265358979323846;
    field(HC_F_TOTAL_VOL)[i] = 2494.0;
    field(HC_F_NUCLEAR_VOL)[i] = 540.0;
    field(HC_F_RADIUS)[i] = std::cbrt(3.0 * 2494.0 / (4.0 * pi));
    field(HC_F_NUCLEAR_RAD)[i] = std::cbrt(3.0 * 540.0 / (4.0 * pi));
    field(HC_F_REPULSION)[i] = 10.0;
    field(HC_F_ADHESION)[i] = 0.4;
    field(HC_F_REL_MAX_ADH)[i] = 1.25;
    field(HC_F_MOT_SPEED)[i] = 1.0;
    field(HC_F_MOT_BIAS)[i] = 0.5;
    field(HC_F_ONCOPROTEIN)[i] = 1.0;
    is_alive[i] = 1;
    for (uint32_t s = 0; s < n_substrates; s++) {
        const size_t o = (size_t)i * n_substrates + s;
        secretion_rate[o] = 0.0;
        uptake_rate[o] = 10.0;
        saturation_density[o] = 1.0;
        net_export_rate[o] = 0.0;
    }
    return i;
}

void CellData::removeCell(uint32_t index) {
    if (index >= num_cells) return;
    const uint32_t last = num_cells - 1;
    if (index != last) {
        for (uint32_t f = 0; f < HC_NUM_F64_FIELDS; f++) field(f)[index] = field(f)[last];
        for (uint32_t f = HC_NUM_F64_FIELDS; f < HC_NUM_CELL_FIELDS; f++) ufield(f)[index] = ufield(f)[last];
        for (uint32_t s = 0; s < n_substrates; s++) {
            const size_t oi = (size_t)index * n_substrates + s, ol = (size_t)last * n_substrates + s;
            secretion_rate[oi] = secretion_rate[ol];
            uptake_rate[oi] = uptake_rate[ol];
            saturation_density[oi] = saturation_density[ol];
            net_export_rate[oi] = net_export_rate[ol];
        }
    }
    num_cells = last;
}

void CellData::syncToDevice() {
    soa_.upload_all();
    rates_.upload_all();
}
void CellData::syncToHost() { soa_.download_all(); check(ctx_, hc_wait(ctx_), "wait"); }
void CellData::syncFieldToHost(uint32_t f) {
    const size_t w = f < HC_NUM_F64_FIELDS ? sizeof(double) : sizeof(uint32_t);
    soa_.download(hc_cells_field_offset(f, max_cells), (size_t)num_cells * w);
    check(ctx_, hc_wait(ctx_), "wait");
}

// ------------------------------------------------------------ Microenvironment
void Microenvironment::initialize(const SimConfig& config, hc_ctx* ctx) {
    ctx_ = ctx;
    // grid sizing by truncation (src/microenvironment.mm:33-35)
    grid_.nx = (uint32_t)((config.x_max - config.x_min) / config.dx);
    grid_.ny = (uint32_t)((config.y_max - config.y_min) / config.dy);
    grid_.nz = config.use_2D ? 1u : (uint32_t)((config.z_max - config.z_min) / config.dz);
    grid_.dx = config.dx; grid_.dy = config.dy; grid_.dz = config.dz;
    grid_.x_min = config.x_min; grid_.y_min = config.y_min; grid_.z_min = config.z_min;
    grid_.dt = config.dt_diffusion;
    grid_.n_substrates = (uint32_t)config.substrates.size();
    grid_.field_fp32 = config.mixed_precision ? 1u : 0u;
    if (!grid_.nx || !grid_.ny || !grid_.nz) throw std::runtime_error("Microenvironment: empty grid");
    subs_.resize(grid_.n_substrates);
    for (uint32_t s = 0; s < grid_.n_substrates; s++) {
        const auto& c = config.substrates[s];
        subs_[s] = hc_substrate{c.diffusion_coeff, c.decay_rate, 0.0, c.dirichlet_value,
                                c.dirichlet_enabled ? 1u : 0u, 0u};
    }
    const size_t V = numVoxels();
    const size_t elem = grid_.field_fp32 ? sizeof(float) : sizeof(double);
    density_.allocate(ctx, std::max<size_t>(8, V * grid_.n_substrates * elem));
    host_density_.assign(V * grid_.n_substrates, 0.0);
    for (uint32_t s = 0; s < grid_.n_substrates; s++)
        std::fill(host_density_.begin() + s * V, host_density_.begin() + (s + 1) * V, config.substrates[s].initial_value);
    syncToDevice();
    tables_.allocate(ctx, std::max<size_t>(8, hc_thomas_table_bytes(&grid_)));
    check(ctx, hc_build_thomas_tables(ctx, &grid_, subs_.data(), tables_.get()), "build Thomas tables");
}

void Microenvironment::simulateDiffusionDecay() {
    check(ctx_, hc_diffusion_substep(ctx_, density_.get(), &grid_, subs_.data(), tables_.get()),
          "simulateDiffusionDecay");
}

void Microenvironment::applyDirichletBC() {
    check(ctx_, hc_dirichlet(ctx_, density_.get(), &grid_, subs_.data()), "applyDirichletBC");
}

void Microenvironment::planSecretion(CellData& cells, double dt) {
    if (plan_capacity_ != cells.max_cells) {
        plan_.allocate(ctx_, hc_secretion_plan_bytes(&grid_, cells.max_cells));
        plan_capacity_ = cells.max_cells;
    }
    check(ctx_, hc_secretion_plan(ctx_, cells.gpu_buffer(), cells.rates_buffer(), &grid_, dt, cells.num_cells,
                                  cells.max_cells, plan_.get()),
          "secretion plan");
}

void Microenvironment::applySecretion() {
    if (!plan_.get()) return;
    check(ctx_, hc_secretion_apply(ctx_, density_.get(), plan_.get(), &grid_), "secretion");
}

void Microenvironment::refreshVoxelIndices(CellData& cells) {
    check(ctx_, hc_voxel_keys(ctx_, cells.gpu_buffer(), &grid_, cells.num_cells, cells.max_cells), "voxel keys");
}

double* Microenvironment::densityBuffer() {
    syncToHost();
    return host_density_.data();
}
// device field -> FP64 host copy
void Microenvironment::syncToHost() {
    density_.download_all();
    check(ctx_, hc_wait(ctx_), "wait");
    const size_t cnt = host_density_.size();
    if (grid_.field_fp32) {
        const float* f = (const float*)density_.host();
        for (size_t i = 0; i < cnt; i++) host_density_[i] = (double)f[i];
    } else {
        std::memcpy(host_density_.data(), density_.host(), cnt * sizeof(double));
    }
}
// FP64 host copy -> device field (rounded to FP32 in the mixed form)
void Microenvironment::syncToDevice() {
    const size_t cnt = host_density_.size();
    if (grid_.field_fp32) {
        float* f = (float*)density_.host();
        for (size_t i = 0; i < cnt; i++) f[i] = (float)host_density_[i];
    } else {
        std::memcpy(density_.host(), host_density_.data(), cnt * sizeof(double));
    }
    density_.upload_all();
    check(ctx_, hc_wait(ctx_), "wait");
}

void Microenvironment::totals(double* out) {
    check(ctx_, hc_substrate_totals(ctx_, density_.get(), &grid_, out), "substrate totals");
}

double Microenvironment::getDensity(uint32_t v, uint32_t s) const {
    return host_density_[(size_t)s * numVoxels() + v];
}
void Microenvironment::setDensity(uint32_t v, uint32_t s, double value) {
    host_density_[(size_t)s * numVoxels() + v] = value;
}

// src/microenvironment.mm:176-184
uint32_t Microenvironment::voxelIndex(double x, double y, double z) const {
    int i = (int)((x - grid_.x_min) / grid_.dx);
    int j = (int)((y - grid_.y_min) / grid_.dy);
    int k = (int)((z - grid_.z_min) / grid_.dz);
    i = (i < 0) ? 0 : (i >= (int)grid_.nx ? (int)grid_.nx - 1 : i);
    j = (j < 0) ? 0 : (j >= (int)grid_.ny ? (int)grid_.ny - 1 : j);
    k = (k < 0) ? 0 : (k >= (int)grid_.nz ? (int)grid_.nz - 1 : k);
    return (uint32_t)k * grid_.nx * grid_.ny + (uint32_t)j * grid_.nx + (uint32_t)i;
}

// --------------------------------------------------------------- CellMechanics
void CellMechanics::initialize(const SimConfig& config, hc_ctx* ctx, uint32_t max_cells) {
    ctx_ = ctx;
    max_cells_ = max_cells;
    // src/cell_mechanics.mm:27-39: 30 um voxels, max(1, trunc(L/30)), nz = 1 in 2-D
    const double v = 30.0;
    params_.dt = config.dt_mechanics;
    params_.cell_cell_repulsion = 10.0;
    params_.cell_cell_adhesion = 0.4;
    params_.max_interaction_distance = v;
    params_.mechanics_voxel_size = v;
    params_.mech_grid_nx = std::max(1u, (uint32_t)((config.x_max - config.x_min) / v));
    params_.mech_grid_ny = std::max(1u, (uint32_t)((config.y_max - config.y_min) / v));
    params_.mech_grid_nz = config.use_2D ? 1u : std::max(1u, (uint32_t)((config.z_max - config.z_min) / v));
    params_.forces_fp32 = config.mixed_precision ? 1u : 0u;
    total_mech_voxels_ = params_.mech_grid_nx * params_.mech_grid_ny * params_.mech_grid_nz;
    // mech GridParams (src/cell_mechanics.mm:46-59): origin, mech dims, 30 um spacing
    mgrid_ = hc_grid{};
    mgrid_.nx = params_.mech_grid_nx; mgrid_.ny = params_.mech_grid_ny; mgrid_.nz = params_.mech_grid_nz;
    mgrid_.dx = mgrid_.dy = mgrid_.dz = v;
    mgrid_.x_min = config.x_min; mgrid_.y_min = config.y_min; mgrid_.z_min = config.z_min;
    hash_.allocate(ctx, hc_hash_bytes(total_mech_voxels_, max_cells));
}

void CellMechanics::computeForces(CellData& cells, double) {
    if (cells.num_cells == 0) return;
    check(ctx_, hc_forces_only_pipeline(ctx_, cells.gpu_buffer(), hash_.get(), &params_, &mgrid_, cells.num_cells,
                                        cells.max_cells, total_mech_voxels_),
          "computeForces");
}

void CellMechanics::integratePositions(CellData& cells, double dt) {
    if (cells.num_cells == 0) return;
    check(ctx_, hc_integrate(ctx_, cells.gpu_buffer(), &mgrid_, dt, cells.num_cells, cells.max_cells),
          "integratePositions");
}

void CellMechanics::update(CellData& cells, double dt) {
    if (cells.num_cells == 0) return;
    check(ctx_, hc_mechanics_pipeline(ctx_, cells.gpu_buffer(), hash_.get(), &params_, &mgrid_, dt, cells.num_cells,
                                      cells.max_cells, total_mech_voxels_),
          "mechanics pipeline");
}

uint32_t CellMechanics::voxelIndexFromPosition(double px, double py, double pz) const {
    const double v = params_.mechanics_voxel_size;
    int vi = (int)((px - mgrid_.x_min) / v);
    int vj = (int)((py - mgrid_.y_min) / v);
    int vk = (int)((pz - mgrid_.z_min) / v);
    vi = std::max(0, std::min(vi, (int)params_.mech_grid_nx - 1));
    vj = std::max(0, std::min(vj, (int)params_.mech_grid_ny - 1));
    vk = std::max(0, std::min(vk, (int)params_.mech_grid_nz - 1));
    return (uint32_t)vk * params_.mech_grid_ny * params_.mech_grid_nx + (uint32_t)vj * params_.mech_grid_nx +
           (uint32_t)vi;
}

void CellMechanics::syncHashPointers(uint32_t num_cells) {
    host_offsets_.assign((size_t)total_mech_voxels_ + 1, 0);
    host_cells_.assign(std::max<uint32_t>(1, num_cells), 0);
    check(ctx_, hc_hash_download(ctx_, hash_.get(), total_mech_voxels_, max_cells_, num_cells, host_offsets_.data(),
                                 host_cells_.data()),
          "syncHashPointers");
}

uint32_t CellMechanics::voxelCellCount(uint32_t v) const {
    if (host_offsets_.empty() || v >= total_mech_voxels_) return 0;
    return host_offsets_[v + 1] - host_offsets_[v];
}

const uint32_t* CellMechanics::voxelCells(uint32_t v) const {
    if (host_offsets_.empty() || v >= total_mech_voxels_) return nullptr;
    return host_cells_.data() + host_offsets_[v];
}

uint32_t CellMechanics::maxVoxelOccupancy() {
    uint32_t m = 0;
    check(ctx_, hc_hash_max_bin(ctx_, hash_.get(), total_mech_voxels_, &m), "max bin");
    return m;
}

uint64_t CellMechanics::neighborQuery(CellData& cells, int kind, uint32_t* offsets, uint32_t* idx,
                                      uint64_t capacity) {
    const uint32_t n = cells.num_cells;
    check(ctx_, hc_build_hash(ctx_, cells.gpu_buffer(), hash_.get(), &params_, &mgrid_, n, cells.max_cells,
                              total_mech_voxels_),
          "neighbour query: hash");
    if (!q_off_.get() || q_off_.size() < ((size_t)cells.max_cells + 1) * 4)
        q_off_.allocate(ctx_, ((size_t)cells.max_cells + 1) * 4);
    uint64_t total = 0;
    check(ctx_, hc_neighbor_query(ctx_, cells.gpu_buffer(), hash_.get(), &params_, n, cells.max_cells,
                                  total_mech_voxels_, kind, q_off_.get(), nullptr, &total),
          "neighbour query: count");
    if (!q_idx_.get() || q_idx_.size() < std::max<uint64_t>(16, total * 4))
        q_idx_.allocate(ctx_, std::max<uint64_t>(16, total * 4));
    check(ctx_, hc_neighbor_query(ctx_, cells.gpu_buffer(), hash_.get(), &params_, n, cells.max_cells,
                                  total_mech_voxels_, kind, q_off_.get(), q_idx_.get(), &total),
          "neighbour query");
    if (offsets) {
        q_off_.download(0, ((size_t)n + 1) * 4);
        check(ctx_, hc_wait(ctx_), "wait");
        std::memcpy(offsets, q_off_.host(), ((size_t)n + 1) * 4);
    }
    if (idx && capacity >= total && total) {
        q_idx_.download(0, total * 4);
        check(ctx_, hc_wait(ctx_), "wait");
        std::memcpy(idx, q_idx_.host(), total * 4);
    }
    return total;
}

void CellMechanics::loadMirror(CellData& cells) {
    const size_t bytes = hc::mirror_layout(max_cells_).total;
    if (!mirror_.get() || mirror_.size() < bytes) mirror_.allocate(ctx_, bytes);
    check(ctx_, hc::mirror_load(ctx_, cells.gpu_buffer(), hash_.get(), mirror_.get(), &params_, &mgrid_,
                                cells.num_cells, cells.max_cells, total_mech_voxels_),
          "mechanics mirror: load");
}

void CellMechanics::mirrorStep(CellData& cells, double dt) {
    check(ctx_, hc::mirror_step(ctx_, hash_.get(), mirror_.get(), &params_, &mgrid_, dt, cells.num_cells,
                                cells.max_cells, total_mech_voxels_),
          "mechanics mirror: step");
}

void CellMechanics::resortMirror(CellData& cells) {
    check(ctx_, hc::mirror_resort(ctx_, mirror_.get(), cells.num_cells, cells.max_cells), "mechanics mirror: re-sort");
}

void CellMechanics::syncMirror(CellData& cells) {
    check(ctx_, hc::mirror_sync(ctx_, cells.gpu_buffer(), mirror_.get(), cells.num_cells, cells.max_cells),
          "mechanics mirror: sync");
}

uint64_t CellMechanics::neighborQuery(CellData& cells, int kind, std::vector<uint32_t>& offsets,
                                      std::vector<uint32_t>& idx) {
    offsets.assign((size_t)cells.num_cells + 1, 0u);
    const uint64_t total = neighborQuery(cells, kind, offsets.data(), nullptr, 0);
    idx.resize(total);
    if (total) {
        // the fill pass above left the lists in q_idx_
        q_idx_.download(0, total * 4);
        check(ctx_, hc_wait(ctx_), "wait");
        std::memcpy(idx.data(), q_idx_.host(), total * 4);
    }
    return total;
}

// ------------------------------------------------------------------ Simulation
static hc_ctx* open_context(int device) {
    hc_ctx* ctx = nullptr;
    int rc = hc_create(&ctx, device);
    if (rc) throw std::runtime_error("hip_context: no usable HIP device (" + std::to_string(rc) + ")");
    rc = hc_load_kernels(ctx, nullptr);
    if (rc) {
        std::string m = hc_last_error(ctx);
        hc_destroy(ctx);
        throw std::runtime_error("hip_context: " + m);
    }
    return ctx;
}

Simulation::Simulation(const SimConfig& cfg, int device)
    : ctx(open_context(device)), config(cfg), cells(cfg.max_cells) {
    try {
        microenv.initialize(config, ctx);
        cells.attach(ctx, (uint32_t)config.substrates.size());
        mechanics.initialize(config, ctx, cells.max_cells);
    } catch (...) {
        hc_destroy(ctx);
        throw;
    }
}

Simulation::~Simulation() {
    if (graph_exec_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
    hc_wait(ctx);
    if (join_) (void)hipEventDestroy((hipEvent_t)join_);
    if (fork_) (void)hipEventDestroy((hipEvent_t)fork_);
    if (side_) (void)hipStreamDestroy((hipStream_t)side_);
    // buffers are released by their owners before the context goes
    cells = CellData(1);
    microenv = Microenvironment();
    mechanics = CellMechanics();
    motility_ = DeviceBuffer();
    volume_ = DeviceBuffer();
    mask_ = DeviceBuffer();
    hc_destroy(ctx);
}

void Simulation::setCells(uint32_t n, const double* x, const double* y, const double* z, const double* radius,
                          const double* volume, const double* rep, const double* adh, const double* rel_adh,
                          const double* sec, const double* upt, const double* sat, const double* net,
                          const uint32_t* alive) {
    if (n > cells.max_cells) throw std::runtime_error("setCells: more cells than max_cells");
    pheno_acc_ = 0.0;
    cells.num_cells = 0;
    const uint32_t ns = cells.n_substrates;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t c = cells.addCell();
        cells.position_x[c] = x[i];
        cells.position_y[c] = y[i];
        cells.position_z[c] = z[i];
        cells.radius[c] = radius[i];
        cells.total_volume[c] = volume[i];
        cells.cell_cell_repulsion[c] = rep[i];
        cells.cell_cell_adhesion[c] = adh[i];
        cells.relative_max_adhesion_distance[c] = rel_adh[i];
        cells.is_alive[c] = alive ? alive[i] : 1u;
        // placement sets the voxel with the 3-D formula (src/main.mm:95)
        cells.voxel_index[c] = microenv.voxelIndex(x[i], y[i], z[i]);
        for (uint32_t s = 0; s < ns; s++) {
            const size_t o = (size_t)i * ns + s;
            cells.secretion_rate[o] = sec ? sec[o] : 0.0;
            cells.uptake_rate[o] = upt ? upt[o] : 0.0;
            cells.saturation_density[o] = sat ? sat[o] : 0.0;
            cells.net_export_rate[o] = net ? net[o] : 0.0;
        }
    }
    cells.syncToDevice();
    mirror_state_ = 0;
    microenv.planSecretion(cells, config.dt_diffusion);
    if (graph_exec_) {
        (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
        graph_exec_ = nullptr;
    }
}

void Simulation::diffusionSubstep() {
    microenv.simulateDiffusionDecay();
    microenv.applySecretion();
}

bool Simulation::useMirror() const {
    return mirror_on_ && !motility_on_ && cells.num_cells > 0;
}

void Simulation::setMechanicsMirror(bool on) {
    syncMirror();
    mirror_on_ = on;
    if (!on) mirror_state_ = 0;
    if (graph_exec_) {
        (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
        graph_exec_ = nullptr;
    }
}

void Simulation::syncMirror() {
    if (mirror_state_ != 1) return;
    mechanics.syncMirror(cells);
    mirror_state_ = 2;
}

// the mirror's slots back in record order every MIRROR_RESORT_STEPS steps
// (eager, outside any graph)
void Simulation::mirrorResortIfDue() {
    if (mirror_state_ != 0 && mirror_steps_ >= MIRROR_RESORT_STEPS) {
        mechanics.resortMirror(cells);
        mirror_steps_ = 0;
    }
}

void Simulation::mechanicsStep() {
    if (useMirror()) {
        // while a period graph is being captured the step is recorded, not
        // run: launchPeriodGraph loads, re-sorts and counts around its replays
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(hc_stream_of(ctx), &cs) != hipSuccess)
            throw std::runtime_error("mechanics step: stream capture status");
        const bool eager = cs == hipStreamCaptureStatusNone;
        if (mirror_state_ == 0) {
            mechanics.loadMirror(cells);
            mirror_steps_ = 0;
        }
        if (eager) mirrorResortIfDue();
        mechanics.mirrorStep(cells, config.dt_mechanics);
        mirror_state_ = 1;
        if (eager) mirror_steps_++;
        return;
    }
    syncMirror();
    mirror_state_ = 0;                           // the fields move on; the mirror is rebuilt when next used
    mechanics.computeForces(cells, config.dt_mechanics);
    if (motility_on_) {
        // updateMotility between forces and AB2 (src/main.mm:382-384)
        const hc_grid g = microenv.getGridParams();
        check(ctx, hc_motility(ctx, cells.gpu_buffer(), motility_.get(), microenv.density(), &g,
                               config.dt_mechanics, cells.num_cells, cells.max_cells, motility_use_sensitivity_),
              "motility");
    }
    mechanics.integratePositions(cells, config.dt_mechanics);
}

void Simulation::setMotility(uint64_t seed, const double* persistence, const double* bias, const double* speed,
                             const uint32_t* is_motile, const uint32_t* restrict_2d, const int32_t* chem_index,
                             const int32_t* chem_dir, const double* sensitivity) {
    syncMirror();
    mirror_state_ = 0;
    const uint32_t ns = cells.n_substrates, mc = cells.max_cells, n = cells.num_cells;
    motility_.allocate(ctx, hc_motility_bytes(mc, ns));
    uint8_t* h = (uint8_t*)motility_.host();
    std::memset(h, 0, motility_.size());
    uint64_t* hdr = (uint64_t*)h;
    hdr[0] = 0;          // step counter
    hdr[1] = seed;
    auto F = [&](uint32_t f) { return (double*)(h + hc_motility_field_offset(f, mc, ns)); };
    auto U = [&](uint32_t f) { return (uint32_t*)(h + hc_motility_field_offset(f, mc, ns)); };
    auto I = [&](uint32_t f) { return (int32_t*)(h + hc_motility_field_offset(f, mc, ns)); };
    for (uint32_t i = 0; i < n; i++) {
        // MotilityData::setDefaults (src/motility.cpp:137-157) where no value is given
        F(HC_M_PERSISTENCE)[i] = persistence ? persistence[i] : 15.0;
        F(HC_M_BIAS)[i] = bias ? bias[i] : 0.5;
        U(HC_M_MOTILE)[i] = is_motile ? is_motile[i] : 0u;
        U(HC_M_RESTRICT_2D)[i] = restrict_2d ? restrict_2d[i] : (config.use_2D ? 1u : 0u);
        I(HC_M_CHEM_INDEX)[i] = chem_index ? chem_index[i] : 0;
        I(HC_M_CHEM_DIR)[i] = chem_dir ? chem_dir[i] : 1;
        for (uint32_t q = 0; q < ns; q++)
            F(HC_M_SENSITIVITY)[(size_t)i * ns + q] = sensitivity ? sensitivity[(size_t)i * ns + q] : 0.0;
        if (speed) cells.field(HC_F_MOT_SPEED)[i] = speed[i];
    }
    motility_.upload_all();
    if (speed) cells.syncToDevice();
    check(ctx, hc_wait(ctx), "wait");
    motility_use_sensitivity_ = sensitivity ? 1 : 0;
    motility_on_ = true;
    if (graph_exec_) {
        (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
        graph_exec_ = nullptr;
    }
}

void Simulation::setVolume(uint64_t seed, const double* const* fields) {
    const uint32_t mc = cells.max_cells, n = cells.num_cells;
    volume_.allocate(ctx, hc_volume_bytes(mc));
    uint8_t* h = (uint8_t*)volume_.host();
    std::memset(h, 0, volume_.size());
    uint64_t* hdr = (uint64_t*)h;
    hdr[0] = 0;          // division counter
    hdr[1] = seed;
    // CellData::setDefaults (src/cell_data.cpp:555-624)
    const double V = 2494.0, Vn = 540.0, ff = 0.75, Vc = V - Vn;
    const double Vns = Vn - ff * Vn, Vcs = Vc - ff * Vc;
    const double fresh[HC_NUM_V_FIELDS] = {ff * V, Vns, Vcs, ff, 0.0, Vcs, Vns, ff, 3.62,
                                           3.0 / 60.0, 0.33 / 60.0, 0.27 / 60.0, 0.0, 2.0 * V, 2.0};
    for (uint32_t f = 0; f < HC_NUM_V_FIELDS; f++) {
        double* a = (double*)(h + hc_volume_field_offset(f, mc));
        for (uint32_t i = 0; i < mc; i++) a[i] = fresh[f];
        if (fields && fields[f]) std::memcpy(a, fields[f], (size_t)n * sizeof(double));
    }
    volume_.upload_all();
    check(ctx, hc_wait(ctx), "wait");
}

uint32_t Simulation::updatePhenotypeSoA(double dt, int use_ratio, const uint32_t* mask, const uint32_t* parents,
                                        uint32_t n_parents, const uint32_t* removed, uint32_t n_removed) {
    if (!volume_.get()) throw std::runtime_error("updatePhenotypeSoA: call setVolume first");
    syncMirror();
    mirror_state_ = 0;                           // divisions and removals move agents
    const uint32_t mc = cells.max_cells, ns = cells.n_substrates;
    uint32_t n = cells.num_cells;
    hc_buf* mbuf = nullptr;
    if (mask && n) {
        if (!mask_.get() || mask_.size() < (size_t)mc * 4) mask_.allocate(ctx, (size_t)mc * 4);
        std::memcpy(mask_.host(), mask, (size_t)n * 4);
        mask_.upload(0, (size_t)n * 4);
        mbuf = mask_.get();
    }
    check(ctx, hc_volume_update(ctx, cells.gpu_buffer(), volume_.get(), mbuf, n, mc, dt, use_ratio), "volume");
    hc_buf* mot = motility_on_ ? motility_.get() : nullptr;
    uint32_t n1 = n;
    check(ctx, hc_divide(ctx, cells.gpu_buffer(), cells.rates_buffer(), volume_.get(), mot, ns, parents, n_parents,
                         n, mc, &n1), "divide");
    uint32_t n2 = n1;
    check(ctx, hc_remove(ctx, cells.gpu_buffer(), cells.rates_buffer(), volume_.get(), mot, ns, removed, n_removed,
                         n1, mc, &n2), "remove");
    cells.num_cells = n2;
    // the rates moved on the device: refresh the host copy
    if (ns) check(ctx, hc_download(ctx, cells.rates_buffer(), 0, hc_rates_bytes(mc, ns)), "download rates");
    check(ctx, hc_wait(ctx), "wait");
    phenotypeStep();
    if (graph_exec_) {
        (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
        graph_exec_ = nullptr;
    }
    return n2;
}

uint32_t attach_greedy(uint32_t n, const std::vector<uint32_t>& off, const std::vector<uint32_t>& idx, double dt,
                       const double* rate, uint32_t* count, uint32_t* targets, uint64_t seed) {
    constexpr uint32_t M = 12;    // CellData::MAX_ATTACHMENTS (src/cell_data.h:124)
    if (off.size() != (size_t)n + 1) throw std::runtime_error("form_attachments: candidate lists do not match the cells");
    for (uint32_t i = 0; i < n; i++)
        if (count[i] > M) throw std::runtime_error("form_attachments: count[" + std::to_string(i) + "] exceeds 12");
    std::mt19937 gen((std::mt19937::result_type)seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    uint32_t formed = 0;
    // the reference's i / j loops (src/cell_interactions.cpp:463-477), j over
    // the candidates only: every other j fails the distance test, which comes
    // before the draw in tryFormAttachment (:411-455)
    for (uint32_t i = 0; i < n; i++) {
        if (!(rate[i] > 0.0) || count[i] >= M) continue;
        uint32_t* ti = targets + (size_t)i * M;
        for (uint32_t u = off[i]; u < off[i + 1]; u++) {
            const uint32_t j = idx[u];
            if (!(rate[j] > 0.0)) continue;
            if (count[i] < M && count[j] < M && std::find(ti, ti + count[i], j) == ti + count[i] &&
                uni(gen) < std::max(rate[i], rate[j]) * dt) {
                ti[count[i]++] = j;
                targets[(size_t)j * M + count[j]++] = i;
                formed++;
            }
            if (count[i] >= M) break;
        }
    }
    return formed;
}

uint32_t Simulation::formAttachments(double dt, const double* rate, uint32_t* count, uint32_t* targets,
                                     uint64_t seed) {
    syncMirror();
    std::vector<uint32_t> off, idx;
    mechanics.neighborQuery(cells, HC_QUERY_ATTACH, off, idx);
    return attach_greedy(cells.num_cells, off, idx, dt, rate, count, targets, seed);
}

void Simulation::phenotypeStep() {
    syncMirror();                                // the voxel refresh reads positions
    microenv.refreshVoxelIndices(cells);
    microenv.planSecretion(cells, config.dt_diffusion);
}

// src/main.mm:275-415 (Appendix B): loop condition, accumulators and order.
void Simulation::run(double max_time, int64_t counts[3]) {
    double current_time = 0.0, mech_acc = 0.0, pheno_acc = 0.0;
    const double dtd = config.dt_diffusion;
    int64_t ns = 0, nm = 0, np = 0;
    while (current_time < max_time + 0.5 * dtd) {
        diffusionSubstep();
        ns++;
        mech_acc += dtd;
        if (mech_acc >= config.dt_mechanics - 0.5 * dtd) {
            mechanicsStep();
            nm++;
            mech_acc = 0.0;
        }
        pheno_acc += dtd;
        if (pheno_acc >= config.dt_phenotype - 0.5 * dtd) {
            phenotypeStep();
            np++;
            pheno_acc = 0.0;
        }
        current_time += dtd;
    }
    wait();
    if (counts) { counts[0] = ns; counts[1] = nm; counts[2] = np; }
}

int Simulation::substepsPerMechanicsStep() const {
    double acc = 0.0;
    const double dtd = config.dt_diffusion;
    for (int k = 1; k < 1000000; k++) {
        acc += dtd;
        if (acc >= config.dt_mechanics - 0.5 * dtd) return k;
    }
    return 1;
}

void Simulation::setOverlap(bool on) {
    overlap_ = on;
    if (graph_exec_) {
        (void)hipGraphExecDestroy((hipGraphExec_t)graph_exec_);
        graph_exec_ = nullptr;
    }
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One period.  The substeps read and write the fields and read the secretion
// plan; the mechanics step reads and writes positions, velocities, pressure
// and the hash.  With overlap on, the mechanics step runs on a second stream
// while the substeps run; the period ends when both are done.
void Simulation::periodBody() {
    const int k = substepsPerMechanicsStep();
    if (!overlap_ || motility_on_) {             // motility reads the substeps' density
        for (int s = 0; s < k; s++) diffusionSubstep();
        mechanicsStep();
        return;
    }
    if (!side_) {
        hipStream_t st = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "side stream");
        side_ = st;
        hip_check(hipEventCreateWithFlags(&a, hipEventDisableTiming), "fork event");
        fork_ = a;
        hip_check(hipEventCreateWithFlags(&b, hipEventDisableTiming), "join event");
        join_ = b;
    }
    hipStream_t main = hc_stream_of(ctx), side = (hipStream_t)side_;
    hip_check(hipEventRecord((hipEvent_t)fork_, main), "fork");
    hip_check(hipStreamWaitEvent(side, (hipEvent_t)fork_, 0), "fork wait");
    hc_stream_swap(ctx, side);
    try {
        mechanicsStep();
    } catch (...) {
        hc_stream_swap(ctx, main);
        throw;
    }
    hc_stream_swap(ctx, main);
    hip_check(hipEventRecord((hipEvent_t)join_, side), "join");
    for (int s = 0; s < k; s++) diffusionSubstep();
    hip_check(hipStreamWaitEvent(main, (hipEvent_t)join_, 0), "join wait");
}

// The phenotype accumulator of src/main.mm:394-412 over the k substeps of the
// next period: the substep indices (0..k-1) after which the phenotype step
// fires, in order.  The mechanics step of a period comes at substep k-1, and
// the reference runs it before that iteration's phenotype step.
std::vector<int> Simulation::advancePhenotypeClock(int k) {
    std::vector<int> fires;
    const double dtd = config.dt_diffusion;
    for (int s = 0; s < k; s++) {
        pheno_acc_ += dtd;
        if (pheno_acc_ >= config.dt_phenotype - 0.5 * dtd) {
            fires.push_back(s);
            pheno_acc_ = 0.0;
        }
    }
    return fires;
}

// One period in the exact order of the step loop when the phenotype step
// fires inside it: substeps, the phenotype step after each substep where it
// fires, the mechanics step after the last substep, then a phenotype step
// that fires at the last substep.
void Simulation::periodWithPhenotype(const std::vector<int>& fires) {
    const int k = substepsPerMechanicsStep();
    size_t f = 0;
    for (int s = 0; s < k; s++) {
        diffusionSubstep();
        if (s == k - 1) mechanicsStep();
        while (f < fires.size() && fires[f] == s) {
            phenotypeStep();
            f++;
        }
    }
}

void Simulation::runPeriods(int periods, bool use_graph) {
    const int k = substepsPerMechanicsStep();
    for (int p = 0; p < periods; p++) {
        const std::vector<int> fires = advancePhenotypeClock(k);
        if (fires.empty() || (fires.size() == 1 && fires[0] == k - 1)) {
            // the phenotype step, if due, follows the whole period: the period
            // runs as usual (graph replay or eager) and the phenotype step is
            // issued after it, outside any graph.  Its voxel refresh and
            // secretion re-plan write buffers whose addresses the graph keeps.
            if (use_graph) launchPeriodGraph();
            else periodBody();
            if (!fires.empty()) phenotypeStep();
        } else {
            periodWithPhenotype(fires);
        }
    }
}

void Simulation::launchPeriodGraph() {
    hipStream_t stream = hc_stream_of(ctx);
    if (useMirror() && mirror_state_ == 0) {
        // the captured period steps the mirror; build it first, outside the graph
        mechanics.loadMirror(cells);
        mirror_state_ = 2;
        mirror_steps_ = 0;
    }
    if (useMirror()) mirrorResortIfDue();
    if (!graph_exec_) {
        check(ctx, hc_wait(ctx), "wait");
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) throw std::runtime_error(std::string("graph capture: ") + hipGetErrorString(e));
        try {
            periodBody();
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        e = hipStreamEndCapture(stream, &graph);
        if (e != hipSuccess) throw std::runtime_error(std::string("graph end capture: ") + hipGetErrorString(e));
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw std::runtime_error(std::string("graph instantiate: ") + hipGetErrorString(e));
        graph_exec_ = (void*)exec;
    }
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec_, stream);
    if (e != hipSuccess) throw std::runtime_error(std::string("graph launch: ") + hipGetErrorString(e));
    if (useMirror()) {
        mirror_state_ = 1;                       // the replayed mechanics step moved the mirror on
        mirror_steps_++;
    }
}

void Simulation::wait() { check(ctx, hc_wait(ctx), "wait"); }

}  // namespace physicell

// =================================================================== pc_* C-ABI
using namespace physicell;

struct pc_sim {
    std::unique_ptr<Simulation> sim;
    std::string err;
};

template <class F>
static int guarded(pc_sim* s, F&& f) {
    if (!s || !s->sim) return HC_ERR_ARG;
    try {
        f(*s->sim);
        return 0;
    } catch (const std::exception& e) {
        s->err = e.what();
        return HC_ERR_STATE;
    }
}

extern "C" {

int pc_sim_create(pc_sim** out, const pc_config* cfg, int device) {
    if (!out || !cfg || (cfg->n_substrates && !cfg->substrates)) return HC_ERR_ARG;
    *out = nullptr;
    SimConfig c;
    c.x_min = cfg->x_min; c.x_max = cfg->x_max; c.y_min = cfg->y_min; c.y_max = cfg->y_max;
    c.z_min = cfg->z_min; c.z_max = cfg->z_max;
    c.dx = cfg->dx; c.dy = cfg->dy; c.dz = cfg->dz;
    c.use_2D = cfg->use_2D != 0;
    c.dt_diffusion = cfg->dt_diffusion; c.dt_mechanics = cfg->dt_mechanics;
    c.dt_phenotype = cfg->dt_phenotype; c.max_time = cfg->max_time;
    c.max_cells = cfg->max_cells ? cfg->max_cells : 1;
    c.mixed_precision = cfg->mixed_precision != 0;
    for (uint32_t s = 0; s < cfg->n_substrates; s++) {
        SubstrateConfig sc;
        sc.diffusion_coeff = cfg->substrates[s].diffusion_coeff;
        sc.decay_rate = cfg->substrates[s].decay_rate;
        sc.initial_value = cfg->substrates[s].initial_value;
        sc.dirichlet_value = cfg->substrates[s].dirichlet_value;
        sc.dirichlet_enabled = cfg->substrates[s].dirichlet_enabled != 0;
        c.substrates.push_back(sc);
    }
    auto* s = new pc_sim();
    try {
        s->sim.reset(new Simulation(c, device));
    } catch (const std::exception& e) {
        fprintf(stderr, "[pc_sim_create] %s\n", e.what());
        delete s;
        return HC_ERR_HIP;
    }
    *out = s;
    return 0;
}

void pc_sim_destroy(pc_sim* s) { delete s; }
const char* pc_last_error(const pc_sim* s) { return s ? s->err.c_str() : "null simulation"; }

int pc_sim_set_cells(pc_sim* s, uint32_t n, const double* x, const double* y, const double* z,
                     const double* radius, const double* volume, const double* rep, const double* adh,
                     const double* rel, const double* sec, const double* upt, const double* sat,
                     const double* net, const uint32_t* alive) {
    if (n && (!x || !y || !z || !radius || !volume || !rep || !adh || !rel)) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) { m.setCells(n, x, y, z, radius, volume, rep, adh, rel, sec, upt, sat, net, alive); });
}

int pc_sim_diffusion_substep(pc_sim* s) { return guarded(s, [](Simulation& m) { m.diffusionSubstep(); }); }
int pc_sim_mechanics_step(pc_sim* s) { return guarded(s, [](Simulation& m) { m.mechanicsStep(); }); }
int pc_sim_phenotype_step(pc_sim* s) { return guarded(s, [](Simulation& m) { m.phenotypeStep(); }); }
int pc_sim_run(pc_sim* s, double max_time, int64_t* counts) {
    return guarded(s, [&](Simulation& m) { m.run(max_time, counts); });
}
int pc_sim_run_periods(pc_sim* s, int periods, int use_graph) {
    if (periods < 0) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) { m.runPeriods(periods, use_graph != 0); });
}
int pc_sim_wait(pc_sim* s) { return guarded(s, [](Simulation& m) { m.wait(); }); }

int pc_sim_get_density(pc_sim* s, double* out) {
    if (!out) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        const double* d = m.microenv.densityBuffer();
        const hc_grid g = m.microenv.getGridParams();
        std::memcpy(out, d, (size_t)g.nx * g.ny * g.nz * g.n_substrates * sizeof(double));
    });
}

int pc_sim_set_density(pc_sim* s, const double* in) {
    if (!in) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        m.wait();
        const hc_grid g = m.microenv.getGridParams();
        const size_t V = (size_t)g.nx * g.ny * g.nz;
        for (uint32_t sb = 0; sb < g.n_substrates; sb++)
            for (size_t v = 0; v < V; v++) m.microenv.setDensity((uint32_t)v, sb, in[sb * V + v]);
        m.microenv.syncToDevice();
    });
}

int pc_sim_write_snapshot(pc_sim* s, const char* folder, int frame) {
    if (!folder) return HC_ERR_ARG;
    int rc = 0;
    const int g = guarded(s, [&](Simulation& m) {
        m.syncMirror();
        m.cells.syncToHost();
        CellData& c = m.cells;
        const hc_grid gr = m.microenv.getGridParams();
        pc_snapshot snap{};
        snap.n_cells = c.num_cells;
        snap.n_substrates = c.n_substrates;
        snap.x = c.position_x; snap.y = c.position_y; snap.z = c.position_z;
        snap.total_volume = c.total_volume; snap.nuclear_volume = c.nuclear_volume;
        snap.radius = c.radius; snap.nuclear_radius = c.nuclear_radius;
        snap.vx = c.velocity_x; snap.vy = c.velocity_y; snap.vz = c.velocity_z;
        snap.pressure = c.simple_pressure; snap.adhesion = c.cell_cell_adhesion;
        snap.repulsion = c.cell_cell_repulsion; snap.rel_max_adhesion = c.relative_max_adhesion_distance;
        snap.cell_type = c.cell_type; snap.current_phase = c.current_phase; snap.is_alive = c.is_alive;
        snap.secretion = c.secretion_rate; snap.uptake = c.uptake_rate;
        snap.saturation = c.saturation_density; snap.net_export = c.net_export_rate;
        snap.nx = gr.nx; snap.ny = gr.ny; snap.nz = gr.nz;
        snap.dx = gr.dx; snap.dy = gr.dy; snap.dz = gr.dz;
        snap.x_min = gr.x_min; snap.y_min = gr.y_min; snap.z_min = gr.z_min;
        snap.density = m.microenv.densityBuffer();
        rc = pc_write_snapshot(folder, frame, &snap);
        if (rc) throw std::runtime_error(std::string("pc_sim_write_snapshot: cannot write under ") + folder);
    });
    return g ? g : rc;
}

int pc_sim_neighbor_query(pc_sim* s, int kind, uint32_t* offsets, uint32_t* idx, uint64_t capacity,
                          uint64_t* total) {
    if (!total) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        m.syncMirror();
        *total = m.mechanics.neighborQuery(m.cells, kind, offsets, idx, capacity);
    });
}

int pc_sim_form_attachments(pc_sim* s, double dt, const double* rate, uint32_t* count, uint32_t* targets,
                            uint64_t seed, uint32_t* formed) {
    if (!rate || !count || !targets) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        const uint32_t f = m.formAttachments(dt, rate, count, targets, seed);
        if (formed) *formed = f;
    });
}

int pc_sim_set_volume(pc_sim* s, uint64_t seed, const double* const* fields) {
    return guarded(s, [&](Simulation& m) { m.setVolume(seed, fields); });
}

int pc_sim_update_phenotype(pc_sim* s, double dt, int use_ratio, const uint32_t* mask, const uint32_t* parents,
                            uint32_t n_parents, const uint32_t* removed, uint32_t n_removed, uint32_t* n_out) {
    if ((n_parents && !parents) || (n_removed && !removed)) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        const uint32_t n = m.updatePhenotypeSoA(dt, use_ratio, mask, parents, n_parents, removed, n_removed);
        if (n_out) *n_out = n;
    });
}

int pc_sim_set_motility(pc_sim* s, uint64_t seed, const double* persistence, const double* bias,
                        const double* speed, const uint32_t* is_motile, const uint32_t* restrict_2d,
                        const int32_t* chem_index, const int32_t* chem_dir, const double* sensitivity) {
    return guarded(s, [&](Simulation& m) {
        m.setMotility(seed, persistence, bias, speed, is_motile, restrict_2d, chem_index, chem_dir, sensitivity);
    });
}

int pc_sim_set_overlap(pc_sim* s, int on) {
    return guarded(s, [&](Simulation& m) { m.setOverlap(on != 0); });
}

int pc_sim_set_mechanics_mirror(pc_sim* s, int on) {
    return guarded(s, [&](Simulation& m) { m.setMechanicsMirror(on != 0); });
}


int pc_sim_substrate_totals(pc_sim* s, double* out) {
    if (!out) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) { m.microenv.totals(out); });
}

int pc_sim_get_field(pc_sim* s, uint32_t field, void* out) {
    if (!out || field >= HC_NUM_CELL_FIELDS) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        m.syncMirror();
        m.cells.syncFieldToHost(field);
        if (field < HC_NUM_F64_FIELDS)
            std::memcpy(out, m.cells.field(field), (size_t)m.cells.num_cells * sizeof(double));
        else
            std::memcpy(out, m.cells.ufield(field), (size_t)m.cells.num_cells * sizeof(uint32_t));
    });
}

int pc_sim_get_hash(pc_sim* s, uint32_t* offsets, uint32_t* cells_out) {
    if (!offsets) return HC_ERR_ARG;
    return guarded(s, [&](Simulation& m) {
        m.mechanics.syncHashPointers(m.cells.num_cells);
        const auto& off = m.mechanics.hostOffsets();
        std::memcpy(offsets, off.data(), off.size() * sizeof(uint32_t));
        if (cells_out && m.cells.num_cells)
            std::memcpy(cells_out, m.mechanics.voxelCells(0), (size_t)m.cells.num_cells * sizeof(uint32_t));
    });
}

int pc_sim_info(const pc_sim* s, hc_grid* grid, hc_mech* mech, uint32_t* n_cells) {
    if (!s || !s->sim) return HC_ERR_ARG;
    if (grid) *grid = s->sim->microenv.getGridParams();
    if (mech) *mech = s->sim->mechanics.params();
    if (n_cells) *n_cells = s->sim->cells.num_cells;
    return 0;
}

hc_ctx* pc_sim_context(pc_sim* s) { return (s && s->sim) ? s->sim->ctx : nullptr; }

}  // extern "C"
