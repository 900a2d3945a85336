/*
 * physicell_hip.h -- C-ABI of the MI355X (gfx950) step hot path.
 *
 * Two layers, both plain C (no torch or HIP types in any signature):
 *
 *   hc_*  hip_context: the drop-in for the reference's MetalContext
 *         (src/metal_context.h:14-100).  Device, stream, buffers and one
 *         dispatch call per reference dispatch, plus the kernels the reference
 *         runs on the CPU (Dirichlet, secretion, voxel keys).
 *
 *   pc_*  the host API above it (Microenvironment / CellMechanics / CellData
 *         and the step loop of src/main.mm:275-415, written in C++ in this
 *         library) exposed for callers that drive a whole simulation.
 *
 * Conventions (shared by every entry point):
 *   - return 0 on success, a negative hc_status on failure; the message is
 *     kept in hc_last_error(ctx) (or pc_last_error(sim)).  The reference
 *     prints and aborts on setup errors (src/metal_context.mm:22-35); the C++
 *     host classes in this library keep that behaviour, the C-ABI reports.
 *   - one in-order HIP stream per context; every dispatch is asynchronous and
 *     ordered after the previous one; hc_wait() is the only blocking call
 *     (src/metal_context.mm:448-461).
 *   - device memory is discrete: a host read or write of a buffer is legal
 *     only through hc_download / hc_upload after hc_wait (the reference shares
 *     memory with the GPU, src/metal_context.mm:52-60).
 *   - all fields, positions and forces are IEEE FP64 (the reference is FP32;
 *     FP64 is the parity form).
 */
#ifndef PHYSICELL_HIP_H
#define PHYSICELL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hc_status {
    HC_OK = 0,
    HC_ERR_ARG = -1,       /* invalid argument (null, size, range)         */
    HC_ERR_HIP = -2,       /* a HIP runtime call failed                    */
    HC_ERR_NOMEM = -3,     /* device allocation failed                     */
    HC_ERR_STATE = -4,     /* call out of order (e.g. kernels not loaded)  */
    HC_ERR_CAPACITY = -5   /* a fixed capacity would be exceeded            */
} hc_status;

typedef struct hc_ctx hc_ctx;   /* opaque: device, stream, timing state      */
typedef struct hc_buf hc_buf;   /* opaque: device allocation + host mirror   */

/* Diffusion grid; FP64 form of GridParams (shaders/types.h:23-38).
 * field_fp32 selects the storage of the density field: 0 = FP64 (the plain
 * form), 1 = FP32 (the mixed form: FP32 values and segment recurrences,
 * FP64 elimination tables, Dirichlet terms and interface solves). */
typedef struct hc_grid {
    uint32_t nx, ny, nz, n_substrates;
    uint32_t field_fp32, _pad;
    double dx, dy, dz, dt;
    double x_min, y_min, z_min;
} hc_grid;

/* Per-substrate parameters; FP64 form of SubstrateParams (shaders/types.h:41-49). */
typedef struct hc_substrate {
    double diffusion_coefficient;
    double decay_rate;
    double supply_rate;
    double target_value;
    uint32_t dirichlet_enabled;
    uint32_t _pad;
} hc_substrate;

/* Mechanics parameters; FP64 form of MechanicsParams (shaders/types.h:52-63). */
typedef struct hc_mech {
    double dt;
    double cell_cell_repulsion;       /* carried, not read by the kernels (as in the reference) */
    double cell_cell_adhesion;        /* idem                                                   */
    double max_interaction_distance;  /* 30 um cutoff                                           */
    double mechanics_voxel_size;      /* 30 um                                                  */
    uint32_t mech_grid_nx, mech_grid_ny, mech_grid_nz;
    /* 0: FP64 forces (plain form, 64-byte records); 1: FP32 force arithmetic
     * over 32-byte records holding positions relative to their bin's corner
     * (mixed form); 2: as 1, with every member of the neighbour rows visited
     * (no mark pass; the same results, for testing).  Keys, bins and AB2
     * positions are FP64 in all. */
    uint32_t forces_fp32;
} hc_mech;

/* ---- cell SoA layout (device buffer of hc_cells_bytes(max_cells)) ----------
 * The reference packs 28 four-byte fields x max_cells (src/cell_data.cpp:101-180).
 * Here fields 0..22 are FP64 arrays of max_cells and fields 23..27 are uint32
 * arrays of max_cells that follow them; the field numbers are the reference's
 * CELL_FIELD_* (shaders/types.h:74-102).  Field f (< 23) starts at byte
 * f*8*max_cells; field f (>= 23) at 23*8*max_cells + (f-23)*4*max_cells.     */
enum {
    HC_F_POS_X = 0, HC_F_POS_Y = 1, HC_F_POS_Z = 2,
    HC_F_VEL_X = 3, HC_F_VEL_Y = 4, HC_F_VEL_Z = 5,
    HC_F_PREV_VX = 6, HC_F_PREV_VY = 7, HC_F_PREV_VZ = 8,
    HC_F_RADIUS = 9, HC_F_NUCLEAR_RAD = 10, HC_F_TOTAL_VOL = 11, HC_F_NUCLEAR_VOL = 12,
    HC_F_REPULSION = 13, HC_F_ADHESION = 14, HC_F_REL_MAX_ADH = 15,
    HC_F_MOT_SPEED = 16, HC_F_MOT_BIAS = 17, HC_F_MOT_DIR_X = 18, HC_F_MOT_DIR_Y = 19,
    HC_F_MOT_DIR_Z = 20, HC_F_ONCOPROTEIN = 21, HC_F_PRESSURE = 22,
    HC_F_CELL_TYPE = 23, HC_F_CUR_PHASE = 24, HC_F_IS_ALIVE = 25,
    HC_F_VOXEL_IDX = 26, HC_F_MECH_VOXEL = 27,
    HC_NUM_F64_FIELDS = 23, HC_NUM_U32_FIELDS = 5, HC_NUM_CELL_FIELDS = 28
};
size_t hc_cells_bytes(uint32_t max_cells);
size_t hc_cells_field_offset(uint32_t field, uint32_t max_cells);

/* Per-substrate rate buffer: 4 FP64 arrays [cell*ns + s] of max_cells*ns each,
 * in the order secretion, uptake, saturation, net export
 * (src/cell_data.h:148-151). */
size_t hc_rates_bytes(uint32_t max_cells, uint32_t n_substrates);

/* ---- context and buffers ------------------------------------------------ */

/* MetalContext() (src/metal_context.mm:13-36): opens the device, creates the stream. */
int  hc_create(hc_ctx** out, int device_ordinal);
/* ~MetalContext (src/metal_context.mm:41-47): waits for queued work, then frees. */
void hc_destroy(hc_ctx* ctx);
/* loadShaders (src/metal_context.mm:66-168).  The gfx950 code object is linked
 * into this library; the call checks that every kernel is loadable on the
 * device.  code_object_or_null must be NULL (no external code objects). */
int  hc_load_kernels(hc_ctx* ctx, const char* code_object_or_null);
/* allocateBuffer (src/metal_context.mm:52-60); device memory, zero-filled. */
int  hc_alloc(hc_ctx* ctx, size_t bytes, hc_buf** out);
void hc_free(hc_buf* buf);
/* Host staging copy of the buffer (pinned, allocated on first use).  Stands
 * in for the reference's [buffer contents] view. */
void* hc_host_mirror(hc_buf* buf);
void* hc_device_ptr(hc_buf* buf);
size_t hc_buf_size(const hc_buf* buf);
/* Explicit copies between the mirror and the device at the reference's sync
 * points; asynchronous on the context stream (hc_wait completes them). */
int  hc_upload(hc_ctx* ctx, hc_buf* buf, size_t offset, size_t bytes);
int  hc_download(hc_ctx* ctx, hc_buf* buf, size_t offset, size_t bytes);
/* waitForCompletion (src/metal_context.mm:448-461). */
int  hc_wait(hc_ctx* ctx);
const char* hc_last_error(const hc_ctx* ctx);

/* ---- diffusion ----------------------------------------------------------- */

/* Size of the line-independent Thomas tables (c' and denominator per
 * substrate, axis and node); and their construction (host compute, upload). */
size_t hc_thomas_table_bytes(const hc_grid* grid);
int hc_build_thomas_tables(hc_ctx* ctx, const hc_grid* grid, const hc_substrate* subs,
                           hc_buf* tables);
/* dispatchDiffusionX / dispatchDiffusionY (src/metal_context.mm:223-272),
 * plus axis 2 (z).  Solves every line of every substrate along `axis` in
 * place (shaders/diffusion_2d.metal:28-292; edge rows, n<=2 and n==3 cases
 * as there).  fuse_dirichlet != 0 also re-applies Dirichlet values on every
 * face (valid only for the last sweep of a substep). */
int hc_diffusion_sweep(hc_ctx* ctx, int axis, hc_buf* density, const hc_grid* grid,
                       const hc_substrate* subs, hc_buf* tables, int fuse_dirichlet);
/* Microenvironment::applyDirichletBC (src/microenvironment.mm:127-152), all
 * faces (z faces included when nz > 1). */
int hc_dirichlet(hc_ctx* ctx, hc_buf* density, const hc_grid* grid, const hc_substrate* subs);
/* Sweep algorithm.  0 (default): lines long enough are solved by the
 * partitioned elimination (segments held in registers -- 16 per x line,
 * 4 per y/z line -- joined by a precomputed interface solve; every value is
 * read and written once; x lines are copied through LDS); other lengths use
 * the serial elimination.  1: always the serial elimination, which evaluates
 * the reference's forward/backward recurrence in its order
 * (shaders/diffusion_2d.metal:111-148) and is bit-identical to it in FP64.
 * 2: as 0 with x segments loaded directly instead of through LDS.  All agree
 * to rounding (well below 1e-12 relative). */
int hc_set_sweep_mode(hc_ctx* ctx, int mode);
/* Order of the sweeps inside hc_diffusion_substep (results are identical):
 * 0 default (= 1), 1 axis-major (x of every substrate, then y, then z),
 * 2 substrate-major (x, y, z of substrate 0, then substrate 1, ...). */
int hc_set_substep_order(hc_ctx* ctx, int order);
/* One substep: x, y, [z] sweeps with Dirichlet fused into the last sweep
 * (src/microenvironment.mm:106-115 + src/main.mm:343-347). */
int hc_diffusion_substep(hc_ctx* ctx, hc_buf* density, const hc_grid* grid,
                         const hc_substrate* subs, hc_buf* tables);

/* ---- cells: voxel keys and secretion ------------------------------------ */

/* Cell -> diffusion voxel, 3-D formula (src/microenvironment.mm:176-184);
 * writes field HC_F_VOXEL_IDX.  Replaces the host refresh of
 * src/cell_phenotype.cpp:778-789. */
int hc_voxel_keys(hc_ctx* ctx, hc_buf* cells, const hc_grid* grid, uint32_t n, uint32_t max_cells);

/* Secretion plan: live cells with a valid voxel grouped by voxel (ascending
 * cell index inside a group) and their folded coefficients.  Rebuild it when
 * voxel keys, volumes or rates change. */
size_t hc_secretion_plan_bytes(const hc_grid* grid, uint32_t max_cells);
int hc_secretion_plan(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, const hc_grid* grid,
                      double dt, uint32_t n, uint32_t max_cells, hc_buf* plan);
/* processSecretion for every planned cell (src/cell_phenotype.cpp:526-564
 * called from src/main.mm:352-363): within a voxel, cells are applied one
 * after another in ascending index; no floating-point atomics. */
int hc_secretion_apply(hc_ctx* ctx, hc_buf* density, hc_buf* plan, const hc_grid* grid);
/* plan + apply in one call (the per-substep call of the reference). */
int hc_secretion(hc_ctx* ctx, hc_buf* density, hc_buf* cells, hc_buf* rates,
                 const hc_grid* grid, double dt, uint32_t n, uint32_t max_cells, hc_buf* plan);

/* ---- mechanics ----------------------------------------------------------- */

/* Hash workspace: counting-sort bins over n_mvox mech voxels (offsets,
 * ascending-index cell list, hash-sorted cell records). */
size_t hc_hash_bytes(uint32_t n_mvox, uint32_t max_cells);
/* clear_spatial_hash + build_spatial_hash (shaders/mechanics.metal:27-85) as
 * a counting sort: key (writes HC_F_MECH_VOXEL) -> wave-aggregated histogram
 * -> prefix scan -> stable scatter.  No 64-per-voxel cap. */
int hc_build_hash(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech,
                  const hc_grid* mgrid, uint32_t n, uint32_t max_cells, uint32_t n_mvox);
/* compute_forces (shaders/mechanics.metal:91-216) over the hash: writes
 * velocity = force and simple pressure. */
int hc_forces(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech,
              uint32_t n, uint32_t max_cells, uint32_t n_mvox);
/* dispatchForcesOnlyPipeline (src/metal_context.mm:390-443): hash + forces. */
int hc_forces_only_pipeline(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech,
                            const hc_grid* mgrid, uint32_t n, uint32_t max_cells, uint32_t n_mvox);
/* dispatchIntegrate (src/metal_context.mm:360-383): AB2, clamp to the
 * mech-grid extent [min, min + n*voxel], v_prev <- v (shaders/integrate.metal:31-80). */
int hc_integrate(hc_ctx* ctx, hc_buf* cells, const hc_grid* mgrid, double dt,
                 uint32_t n, uint32_t max_cells);
/* dispatchMechanicsPipeline (src/metal_context.mm:468-541): hash, forces, AB2. */
int hc_mechanics_pipeline(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech,
                          const hc_grid* mgrid, double dt, uint32_t n, uint32_t max_cells,
                          uint32_t n_mvox);
/* Hash accessors of CellMechanics (src/cell_mechanics.h:53-62): downloads
 * the bin offsets (n_mvox+1) and the ascending-index cell list (n) so that
 * voxelCellCount(v) = off[v+1]-off[v] and voxelCells(v) = &list[off[v]]. */
int hc_hash_download(hc_ctx* ctx, hc_buf* hash, uint32_t n_mvox, uint32_t max_cells, uint32_t n,
                     uint32_t* offsets_out, uint32_t* cells_out);
/* Largest bin population of the last hash build (the reference drops cells
 * beyond 64 per bucket, shaders/mechanics.metal:82-84; here none are dropped
 * and this reports how far past 64 a bucket went). */
int hc_hash_max_bin(hc_ctx* ctx, hc_buf* hash, uint32_t n_mvox, uint32_t* max_out);
/* Neighbour lists through the hash: for each cell the candidates passing the
 * force kernel's predicates (same or adjacent mech voxel, d^2 < cutoff^2,
 * d^2 >= 1e-12), in visiting order.  counts[n] (device), idx[n*cap] (device);
 * entries past cap are counted but not stored. */
int hc_neighbor_lists(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech,
                      uint32_t n, uint32_t max_cells, uint32_t n_mvox, uint32_t cap,
                      hc_buf* counts, hc_buf* idx);

/* Neighbour queries on the hash for the phenotype side, as CSR lists per cell:
 * offsets[n + 1] (device) and idx[*total] (device); cell i's neighbours are
 * idx[offsets[i] .. offsets[i+1]).  The hash must have been built from the
 * current positions (hc_build_hash, FP64 or FP32 records).  Positions and
 * radii come from the cell SoA in FP64.
 *   HC_QUERY_CONTACT: writeCellNeighborGraph (src/output_writer.cpp:457-490,
 *     an O(N^2) loop there): for live cells with volume > 1e-15, the cells
 *     j > i (live, volume > 1e-15) with |d|^2 <= (max(rel_i, rel_j) (r_i +
 *     r_j))^2, ascending j.  Fails (HC_ERR_CAPACITY) when the largest such
 *     distance, max rel x 2 max r, exceeds the mechanics voxel (the bins
 *     would not hold every pair).
 *   HC_QUERY_INTERACT: the candidate scan of standardCellCellInteractions
 *     (src/cell_phenotype.cpp:893-967): the first 512 cells j != i of the
 *     3x3x3 bins in visiting order (dk, dj, di, then ascending index), kept
 *     when volume_j >= 1e-15, 2-D distance^2 <= (2.5 r_i + r_j)^2 and
 *     distance <= 1.25 (r_i + r_j); visiting order kept.  Which cells query
 *     (death model, interaction rates) is the caller's.
 *   HC_QUERY_ATTACH: the pair scan of tryFormAllAttachments
 *     (src/cell_interactions.cpp:411-478, an O(N^2) loop there): for live
 *     cells, the live cells j > i with sqrt(|p_j - p_i|^2) <= max(rel_i,
 *     rel_j) (r_i + r_j), ascending j.  Rates, capacity and draws are the
 *     caller's (pc_sim_form_attachments).  Same reach check as CONTACT.
 * idx may be NULL: then only offsets and *total are produced (size the idx
 * buffer and call again); an idx buffer smaller than *total fails. */
enum { HC_QUERY_CONTACT = 0, HC_QUERY_INTERACT = 1, HC_QUERY_ATTACH = 2 };
int hc_neighbor_query(hc_ctx* ctx, hc_buf* cells, hc_buf* hash, const hc_mech* mech, uint32_t n, uint32_t max_cells,
                      uint32_t n_mvox, int kind, hc_buf* offsets, hc_buf* idx, uint64_t* total);

/* Per-substrate totals sum_v rho[s][v] with compensated (Neumaier) FP64
 * summation in a fixed order (deterministic); out[n_substrates]. */
int hc_substrate_totals(hc_ctx* ctx, hc_buf* density, const hc_grid* grid, double* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * totals as hc_substrate_totals, bit for bit, with the partial sums and
 * results in a file-scope device array of 256.5 KiB instead of a per-call
 * hipMalloc/hipFree.  Each GPU the library runs on holds its own copy of the
 * array until the process ends.  n_substrates above
 * HC_TOTALS_NOALLOC_MAX_SUBSTRATES fails with HC_ERR_CAPACITY.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (from
 * several host threads or streams) race. */
#define HC_TOTALS_NOALLOC_MAX_SUBSTRATES 64
int hc_substrate_totals_noalloc(hc_ctx* ctx, hc_buf* density, const hc_grid* grid, double* out);

/* ---- timing ------------------------------------------------------------- */
/* HIP events on the context stream around every launch of the selected
 * kernels (bit mask of hc_kernel_id); accumulated per kernel. */
typedef enum hc_kernel_id {
    HC_K_SWEEP_X = 0, HC_K_SWEEP_Y = 1, HC_K_SWEEP_Z = 2, HC_K_SWEEP_SMALL = 3,
    HC_K_DIRICHLET = 4, HC_K_VOXEL_KEYS = 5, HC_K_SEC_PLAN = 6, HC_K_SEC_APPLY = 7,
    HC_K_HASH = 8, HC_K_FORCES = 9, HC_K_INTEGRATE = 10, HC_K_MOTILITY = 11, HC_K_VOLUME = 12,
    HC_K_DIVIDE = 13, HC_K_REMOVE = 14, HC_K_COUNT = 15
} hc_kernel_id;
int hc_timing_enable(hc_ctx* ctx, uint32_t kernel_mask);
/* Waits, then returns the summed event time (ms) and launch count of `kernel`. */
int hc_timing_get(hc_ctx* ctx, int kernel, double* total_ms, int64_t* launches);
int hc_timing_reset(hc_ctx* ctx);

/* ==== pc_*: the host simulation (Microenvironment + CellMechanics + CellData
 * + the step loop), C++ in this library ==================================== */

typedef struct pc_sim pc_sim;

typedef struct pc_substrate_cfg {
    double diffusion_coeff, decay_rate, initial_value, dirichlet_value;
    uint32_t dirichlet_enabled, _pad;
} pc_substrate_cfg;

/* Subset of SimConfig (src/config_parser.h) that the path reads. */
typedef struct pc_config {
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double dx, dy, dz;
    uint32_t use_2D, n_substrates;
    double dt_diffusion, dt_mechanics, dt_phenotype, max_time;
    uint32_t max_cells;
    /* 0: plain FP64 form; 1: FP32-mixed form (FP32 density and sweeps, FP32
     * force arithmetic; FP64 tables, Dirichlet terms, interface solves,
     * positions/AB2 and substrate totals). */
    uint32_t mixed_precision;
    const pc_substrate_cfg* substrates;
} pc_config;

int  pc_sim_create(pc_sim** out, const pc_config* cfg, int device_ordinal);
void pc_sim_destroy(pc_sim* sim);
const char* pc_last_error(const pc_sim* sim);
/* Places n cells (CellData::addCell + the fields the path reads); rate arrays
 * are [cell*ns + s]; alive may be NULL (all alive).  Sets voxel keys with the
 * 3-D formula, uploads, and plans secretion. */
int pc_sim_set_cells(pc_sim* sim, uint32_t n, const double* x, const double* y, const double* z,
                     const double* radius, const double* volume, const double* repulsion,
                     const double* adhesion, const double* rel_max_adhesion,
                     const double* secretion, const double* uptake, const double* saturation,
                     const double* net_export, const uint32_t* alive);
/* One diffusion substep: sweeps + Dirichlet + secretion. */
int pc_sim_diffusion_substep(pc_sim* sim);
/* One mechanics step: hash + forces + AB2 (dt = dt_mechanics). */
int pc_sim_mechanics_step(pc_sim* sim);
/* Phenotype-cadence hook on the path: voxel-key refresh + secretion re-plan. */
int pc_sim_phenotype_step(pc_sim* sim);
/* The step loop of src/main.mm:275-415 up to max_time (Appendix B order);
 * counts may be NULL (int64[3]: substeps, mechanics, phenotype). */
int pc_sim_run(pc_sim* sim, double max_time, int64_t* counts);
/* Run `periods` x (steps_per_mech diffusion substeps + 1 mechanics step);
 * use_graph != 0 captures one period as a hipGraph and replays it.  The
 * phenotype step (voxel refresh + secretion re-plan) fires every
 * dt_phenotype in the order of src/main.mm:394-412, outside the graph, on a
 * clock that persists across calls and restarts at pc_sim_set_cells; so
 * P periods from fresh cells equal pc_sim_run(sim, (P*k - 1)*dt_diffusion). */
int pc_sim_run_periods(pc_sim* sim, int periods, int use_graph);
/* on != 0 (default off): in pc_sim_run_periods, each period's mechanics step
 * runs on a second stream concurrently with its diffusion substeps (they
 * touch disjoint data; the results are the same).  pc_sim_run keeps the
 * reference's serial order. */
int pc_sim_set_overlap(pc_sim* sim, int on);
/* No counterpart in the reference (whose cell buffer is the only copy).
 * on != 0 (default): the FP64 mechanics step without motility runs on a
 * hash-ordered mirror of the agents' mechanics state (records and previous
 * velocities in the order of the last hash build), so its per-agent passes
 * read and write in spatial order; positions, velocities, previous
 * velocities, pressure and mechanics voxels are written back into the cell
 * buffer before anything in pc_sim_* reads them (the phenotype step, field
 * reads, snapshots, neighbour queries, attachments, phenotype updates).  The
 * results are the same bit for bit with the mirror on or off. */
int pc_sim_set_mechanics_mirror(pc_sim* sim, int on);
/* Enables motility in the mechanics step (forces -> motility -> AB2, as
 * src/main.mm:376-387).  Per-cell arrays as MotilityData (src/motility.h);
 * sensitivity [cell*ns + s] may be NULL (then chem_index / chem_dir bias).
 * While motility is on, periods run in the serial order (motility reads the
 * density the substeps write). */
int pc_sim_set_motility(pc_sim* sim, uint64_t seed, const double* persistence, const double* bias,
                        const double* speed, const uint32_t* is_motile, const uint32_t* restrict_2d,
                        const int32_t* chem_index, const int32_t* chem_dir, const double* sensitivity);
/* Volume state for the device phenotype step: fields[HC_V_*] points at
 * n_cells values or is NULL (CellData's fresh value, src/cell_data.cpp:
 * 597-624; fields itself may be NULL); seed keys division's draws. */
int pc_sim_set_volume(pc_sim* sim, uint64_t seed, const double* const* fields);
/* updatePhenotype's SoA side on the device (src/cell_phenotype.cpp:792-845),
 * with the decisions from the host: hc_volume_update on the cells with
 * mask[i] != 0 (mask: n_cells values, or NULL for all), hc_divide of
 * parents, hc_remove of removed (indices below the count after the
 * divisions), then the voxel refresh and secretion re-plan of
 * pc_sim_phenotype_step.  *n_out (may be NULL) = the new cell count. */
int pc_sim_update_phenotype(pc_sim* sim, double dt, int use_ratio, const uint32_t* mask, const uint32_t* parents,
                            uint32_t n_parents, const uint32_t* removed, uint32_t n_removed, uint32_t* n_out);
/* Builds the hash from the current positions and runs hc_neighbor_query;
 * copies out offsets[n_cells + 1] and, when idx is not NULL and capacity >=
 * *total, idx[*total] (host arrays).  *total is always set. */
int pc_sim_neighbor_query(pc_sim* sim, int kind, uint32_t* offsets, uint32_t* idx, uint64_t capacity,
                          uint64_t* total);
/* tryFormAllAttachments (src/cell_interactions.cpp:459-478) with the pair
 * scan on the device (HC_QUERY_ATTACH over the hash built from the current
 * positions) and the reference's greedy pass on the host: for i ascending
 * with rate[i] > 0 and count[i] < 12, its candidates j ascending with
 * rate[j] > 0; skip when either list is full or j is already i's partner,
 * else draw u and attach both ways when u < max(rate_i, rate_j) dt.  The
 * draws come from std::mt19937 seeded with seed (the reference's generator
 * type, src/cell_interactions.cpp:293-298; pass a new seed per call).
 * rate[n_cells], count[n_cells] and targets[n_cells * 12] are host arrays,
 * count and targets updated in place; *formed (may be NULL) = the number of
 * attachments made. */
int pc_sim_form_attachments(pc_sim* sim, double dt, const double* rate, uint32_t* count, uint32_t* targets,
                            uint64_t seed, uint32_t* formed);
int pc_sim_wait(pc_sim* sim);
/* Copies out (waits first).  density: ns*nx*ny*nz FP64 (converted from FP32
 * in the mixed form). */
int pc_sim_get_density(pc_sim* sim, double* out);
int pc_sim_set_density(pc_sim* sim, const double* in);
/* field: an HC_F_* number; out has n entries (double or uint32 by field). */
int pc_sim_get_field(pc_sim* sim, uint32_t field, void* out);
int pc_sim_get_hash(pc_sim* sim, uint32_t* offsets, uint32_t* cells);
/* Compensated per-substrate totals of the density field (hc_substrate_totals). */
int pc_sim_substrate_totals(pc_sim* sim, double* out);
int pc_sim_info(const pc_sim* sim, hc_grid* grid, hc_mech* mech, uint32_t* n_cells);
hc_ctx* pc_sim_context(pc_sim* sim);

/* ==== motility on the device (src/motility.cpp:179-293) ==================== */

/* Per-cell motility state (MotilityData, src/motility.h), one buffer:
 * a 256-byte header (uint64 step counter, uint64 seed), then FP64
 * persistence time, elapsed time, bias, motility vector x/y/z; uint32
 * is-motile and restrict-to-2D flags; int32 chemotaxis substrate index and
 * direction (each max_cells long); then FP64 chemotactic sensitivities
 * [cell*ns + s]. */
enum {
    HC_M_PERSISTENCE = 0, HC_M_ELAPSED = 1, HC_M_BIAS = 2, HC_M_VEC_X = 3, HC_M_VEC_Y = 4, HC_M_VEC_Z = 5,
    HC_M_MOTILE = 6, HC_M_RESTRICT_2D = 7, HC_M_CHEM_INDEX = 8, HC_M_CHEM_DIR = 9, HC_M_SENSITIVITY = 10
};
size_t hc_motility_bytes(uint32_t max_cells, uint32_t n_substrates);
size_t hc_motility_field_offset(uint32_t field, uint32_t max_cells, uint32_t n_substrates);
/* updateMotility: for each live motile cell, elapsed += dt; resample the
 * direction when persistence < dt or uniform < dt/persistence (random unit
 * vector on the circle if restricted to 2-D, else the sphere, blended with
 * the normalised chemotactic bias: the sensitivity-weighted gradient sum if
 * use_sensitivity, else direction * gradient of the indexed substrate),
 * normalise, scale by the cell's speed (field HC_F_MOT_SPEED); add the vector
 * to the velocity.  Non-motile live cells get a zero vector.  Uniform draws
 * come from a counter-based stream keyed by (seed, step, cell, draw); the step
 * counter in the header advances by one per call.  density may be NULL (no
 * bias). */
int hc_motility(hc_ctx* ctx, hc_buf* cells, hc_buf* motility, hc_buf* density, const hc_grid* grid, double dt,
                uint32_t n, uint32_t max_cells, int use_sensitivity);

/* ==== phenotype-side SoA mutation on the device ============================
 * The volume model, division and removal of updatePhenotype
 * (src/cell_phenotype.cpp:52-113, 571-720, 833-844; src/cell_data.cpp:398-542).
 * Which cells divide or are removed is decided on the host (cycle and death
 * models) and passed as index lists. */

/* Per-cell volume state (the Volume fields of CellData, src/cell_data.h), one
 * buffer: a 256-byte header (uint64 division counter, uint64 seed), then one
 * FP64 array of max_cells values per field. */
enum {
    HC_V_FLUID = 0, HC_V_SOLID_NUC = 1, HC_V_SOLID_CYTO = 2, HC_V_FLUID_FRAC = 3, HC_V_CALC_FRAC = 4,
    HC_V_TARGET_SOLID_CYTO = 5, HC_V_TARGET_SOLID_NUC = 6, HC_V_TARGET_FLUID_FRAC = 7, HC_V_TARGET_CN_RATIO = 8,
    HC_V_FLUID_RATE = 9, HC_V_NUC_RATE = 10, HC_V_CYTO_RATE = 11, HC_V_CALC_RATE = 12, HC_V_RUPTURE_VOL = 13,
    HC_V_REL_RUPTURE_VOL = 14, HC_NUM_V_FIELDS = 15
};
size_t hc_volume_bytes(uint32_t max_cells);
size_t hc_volume_field_offset(uint32_t field, uint32_t max_cells);
/* updateVolume (src/cell_phenotype.cpp:52-113) for every cell i < n whose
 * mask[i] != 0 (mask: a device buffer of n uint32, or NULL for every cell):
 * fluid, nuclear and cytoplasmic solid relax toward their targets, the
 * calcified fraction grows, total and nuclear volume (HC_F_TOTAL_VOL,
 * HC_F_NUCLEAR_VOL) are rebuilt from the compartments and the radii
 * (HC_F_RADIUS, HC_F_NUCLEAR_RAD) follow as cbrt(V / (4/3 pi)).  use_ratio
 * != 0 recomputes the cytoplasmic solid target as ratio * nuclear solid
 * target first (:80-87; the reference does so whenever the ratio array
 * exists). */
int hc_volume_update(hc_ctx* ctx, hc_buf* cells, hc_buf* volume, hc_buf* mask, uint32_t n, uint32_t max_cells,
                     double dt, int use_ratio);
/* divideCell (src/cell_phenotype.cpp:571-720) for parents[0..count) (host
 * array, distinct indices < n), in list order: parent k's daughter is cell
 * n + k.  Both halve their volumes and volume compartments and recompute
 * their radii; the daughter sits radius/2 from the parent at a random angle
 * in the x-y plane, inherits type, mechanics, motility speed and bias,
 * volume rates, rates [cell*ns + s] and voxels, and takes CellData's fresh
 * values (src/cell_data.cpp:544-640) elsewhere; the oncoprotein level gets
 * +noise in the daughter and -noise in the parent, noise ~ N(0, 0.02 level),
 * both clamped at 0.  The angle and the noise come from the counter-based
 * stream keyed by (seed, division counter, k, draw); the counter advances by
 * one per call.  As in the reference, divisions beyond max_cells are skipped.
 * motility may be NULL; else the daughter takes the parent's motility state
 * with zero elapsed time (the reference leaves it unset).  *n_out = the new
 * cell count. */
int hc_divide(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, hc_buf* volume, hc_buf* motility, uint32_t ns,
              const uint32_t* parents, uint32_t count, uint32_t n, uint32_t max_cells, uint32_t* n_out);
/* The reference's removal loop (src/cell_phenotype.cpp:833-844): the
 * indices (host array, distinct, < n, any order) are removed in descending
 * order, each by CellData::removeCell (src/cell_data.cpp:408-542: the last
 * cell is copied into the hole).  The resulting arrangement is computed on
 * the host from the index list alone, then every moved cell's fields (cell
 * SoA, rates, volume state, and, when motility is not NULL, the motility
 * state as MotilityData::swapRemove, src/motility.cpp:154-175) are copied in
 * one parallel pass.  *n_out = n - count. */
int hc_remove(hc_ctx* ctx, hc_buf* cells, hc_buf* rates, hc_buf* volume, hc_buf* motility, uint32_t ns,
              const uint32_t* indices, uint32_t count, uint32_t n, uint32_t max_cells, uint32_t* n_out);
/* The moves hc_remove makes, host only: after removing `indices` from n
 * cells, cell dst[k] holds what was cell src[k] (dst < n - count <= src);
 * every other cell below n - count keeps its index.  dst and src hold count
 * entries (or are NULL); *n_pairs = the number of moves.  A host that keeps
 * per-cell data of its own (cycle and death state) applies the same moves. */
int hc_remove_plan(const uint32_t* indices, uint32_t count, uint32_t n, uint32_t* dst, uint32_t* src,
                   uint32_t* n_pairs);

/* ==== MATLAB Level-4 snapshots in the reference's output format =========== */

/* One matrix file as src/matlab_writer.cpp:5-49 writes it: uint32 header
 * {type 0 (little-endian double, full), rows, cols, 0, name length + 1}, the
 * name with its NUL, then rows*cols doubles, column-major
 * (data[c*rows + r] = M[r][c]).  Host only. */
int pc_write_matlab4(const char* path, const char* name, uint32_t rows, uint32_t cols, const double* data);
/* final_cell_neighbor_graph.txt as writeCellNeighborGraph writes it
 * (src/output_writer.cpp:457-490): one "i j" line per pair, i ascending, then
 * j ascending; from a contact query's CSR lists (host arrays). */
int pc_write_neighbor_graph(const char* path, uint32_t n, const uint32_t* offsets, const uint32_t* idx);

/* Host view of one snapshot of the path's state. */
typedef struct pc_snapshot {
    uint32_t n_cells, n_substrates;
    const double *x, *y, *z, *total_volume, *nuclear_volume, *radius, *nuclear_radius;
    const double *vx, *vy, *vz, *pressure, *adhesion, *repulsion, *rel_max_adhesion; /* may be NULL */
    const uint32_t *cell_type, *current_phase, *is_alive;                          /* may be NULL */
    const double *secretion, *uptake, *saturation, *net_export;     /* [cell*ns + s], may be NULL */
    uint32_t nx, ny, nz, _pad;
    double dx, dy, dz, x_min, y_min, z_min;
    const double* density;      /* ns*nx*ny*nz substrate-major; NULL: no microenvironment file */
} pc_snapshot;

/* writeMatlabSnapshots (src/output_writer.cpp:295-453): writes
 * <folder>/output%08d_cells.mat (one column per cell, the reference's row
 * schema with one cell type and no custom variables; rows outside the path
 * hold a placed reference cell's values) and
 * <folder>/output%08d_microenvironment0.mat (one column per voxel: centre and
 * volume computed in FP32 as the reference does, then the densities).  Host
 * only. */
int pc_write_snapshot(const char* folder, int frame, const pc_snapshot* snapshot);
/* The simulation's current state through pc_write_snapshot (waits, copies the
 * SoA and density to the host). */
int pc_sim_write_snapshot(pc_sim* sim, const char* folder, int frame);

/* ==== pe_*: packed ensembles of independent simulations ====================
 * No counterpart in the reference (which runs one simulation per process):
 * many simulations of uneven grid and population size are stepped together.
 * Each member keeps its own fields, tables and cells on the device; the step
 * walks packed work lists (sweep line groups, agent tiles, one hash and one
 * secretion plan over the concatenated bin/voxel spaces).  A member's results
 * do not depend on the other members: they equal the pc_sim results of the
 * same inputs.  Members share dt_diffusion and dt_mechanics; a period is
 * (substeps per mechanics step) diffusion substeps + one mechanics step; each
 * member's phenotype refresh (voxel keys + secretion plan) runs at the end of
 * the periods where its own dt_phenotype clock fires.  Grid limits: x edge
 * 6..258, y and z edges 6..130 (or nz = 1) voxels. */
typedef struct pe_ensemble pe_ensemble;

int  pe_create(pe_ensemble** out, int device_ordinal, uint32_t mixed_precision);
void pe_destroy(pe_ensemble* e);
const char* pe_last_error(const pe_ensemble* e);
/* Adds a member (cells as pc_sim_set_cells; max_cells = n); returns its id. */
int pe_add(pe_ensemble* e, const pc_config* cfg, uint32_t n, const double* x, const double* y, const double* z,
           const double* radius, const double* volume, const double* repulsion, const double* adhesion,
           const double* rel_max_adhesion, const double* secretion, const double* uptake,
           const double* saturation, const double* net_export, const uint32_t* alive, uint32_t* member_id);
/* Removes a member and frees its device state (waits for queued work). */
int pe_remove(pe_ensemble* e, uint32_t member_id);
/* Motility for one member, as pc_sim_set_motility for a single simulation
 * (updateMotility, src/motility.cpp:179-293, between the forces and AB2 of
 * every mechanics step; NULL arrays take MotilityData's defaults,
 * src/motility.cpp:137-157).  A member with motility steps exactly as a
 * pc_sim with the same motility; members that have it run their mechanics
 * after the substeps rather than beside them. */
int pe_set_motility(pe_ensemble* e, uint32_t member_id, uint64_t seed, const double* persistence,
                    const double* bias, const double* speed, const uint32_t* is_motile,
                    const uint32_t* restrict_2d, const int32_t* chem_index, const int32_t* chem_dir,
                    const double* sensitivity);
/* The device phenotype step for one member, as pc_sim_set_volume and
 * pc_sim_update_phenotype for a single simulation.  pe_set_volume also
 * makes room for max_cells cells (ignored when not larger than the member's
 * capacity; the member's arrays move to the larger layout), which the
 * divisions of pe_update_phenotype may fill.  pe_update_phenotype refreshes
 * the member's voxel keys and re-plans secretion; *n_out is its new cell
 * count. */
int pe_set_volume(pe_ensemble* e, uint32_t member_id, uint64_t seed, const double* const* fields,
                  uint32_t max_cells);
int pe_update_phenotype(pe_ensemble* e, uint32_t member_id, double dt, int use_ratio, const uint32_t* mask,
                        const uint32_t* parents, uint32_t n_parents, const uint32_t* removed, uint32_t n_removed,
                        uint32_t* n_out);
/* Neighbour queries and attachments for one member, as pc_sim_neighbor_query
 * and pc_sim_form_attachments for a single simulation: a hash of the member
 * alone is built from its current positions (waits for queued work first);
 * host arrays sized by the member's cell count. */
int pe_neighbor_query(pe_ensemble* e, uint32_t member_id, int kind, uint32_t* offsets, uint32_t* idx,
                      uint64_t capacity, uint64_t* total);
int pe_form_attachments(pe_ensemble* e, uint32_t member_id, double dt, const double* rate, uint32_t* count,
                        uint32_t* targets, uint64_t seed, uint32_t* formed);
/* Runs `periods` periods of every member; use_graph != 0 captures one period
 * as a hipGraph (re-captured after membership changes) and replays it. */
int pe_run_periods(pe_ensemble* e, int periods, int use_graph);
int pe_wait(pe_ensemble* e);
/* Member results (wait first): density as pc_sim_get_density, field as
 * pc_sim_get_field. */
int pe_get_density(pe_ensemble* e, uint32_t member_id, double* out);
int pe_get_field(pe_ensemble* e, uint32_t member_id, uint32_t field, void* out);
/* bytes: the member's device bytes; counts (int64[3]): substeps, mechanics
 * steps, phenotype refreshes done. */
int pe_member_info(pe_ensemble* e, uint32_t member_id, uint64_t* bytes, int64_t* counts);
int pe_info(pe_ensemble* e, uint32_t* n_members, uint64_t* total_cells, uint64_t* total_voxel_values,
            uint64_t* workspace_bytes);
/* Grid of the batched sweeps along axis 0..2: 0 (default) launches one
 * workgroup per work item; a positive value launches a persistent grid of
 * that many workgroups, each walking a contiguous run of items. */
int pe_set_grid_blocks(pe_ensemble* e, int axis, uint32_t blocks);
/* Autotuner: times each batched sweep list at candidate grids (0 and
 * 256..16384 workgroups) on the current members (fields restored
 * afterwards) and keeps the fastest; writes, per axis, the grid of its
 * busiest list and the summed mean launch times (us). */
int pe_autotune(pe_ensemble* e, int reps, uint32_t* blocks_out, double* us_out);
/* on != 0 (default off): each period's batched mechanics step runs on a
 * second stream concurrently with the batched diffusion substeps (same
 * results). */
int pe_set_overlap(pe_ensemble* e, int on);
hc_ctx* pe_context(pe_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* PHYSICELL_HIP_H */
