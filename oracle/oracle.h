/*
 * oracle.h -- CPU FP64 restatement of the reference's step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), and the cpu_baseline leg
 * of bench.py).  Nothing in the product links, loads or calls it.
 *
 * Every function states the reference file:line it restates.  The semantic
 * decisions it encodes (FP64 arithmetic, 3-D extension, deterministic
 * secretion order, stable counting-sort hash with no 64-slot cap) are listed
 * in DESIGN.md.  All arithmetic is IEEE FP64, compiled with
 * -ffp-contract=off so that the operation sequence written here is the one
 * executed (no fused multiply-add).
 */
#ifndef PHYSICELL_ORACLE_H
#define PHYSICELL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Diffusion grid (restates GridParams, shaders/types.h:23-38, in FP64). */
typedef struct or_grid {
    uint32_t nx, ny, nz, ns;      /* voxels per axis, number of substrates */
    double dx, dy, dz, dt;        /* spacing (um), diffusion dt (min)      */
    double x_min, y_min, z_min;   /* domain origin (um)                    */
} or_grid;

/* Per-substrate parameters (restates SubstrateParams, shaders/types.h:41-49). */
typedef struct or_sub {
    double D;          /* diffusion coefficient (um^2/min) */
    double decay;      /* lambda (1/min)                   */
    double target;     /* Dirichlet value                  */
    int32_t dirichlet; /* nonzero: Dirichlet on all faces  */
    int32_t _pad;
} or_sub;

/* Mechanics grid (restates MechanicsParams + the mech GridParams,
 * shaders/types.h:52-63 and src/cell_mechanics.mm:23-82). */
typedef struct or_mech {
    uint32_t mnx, mny, mnz, _pad;
    double voxel;      /* mechanics voxel edge (30 um in the reference) */
    double max_dist;   /* interaction cutoff (30 um in the reference)   */
    double x_min, y_min, z_min;
} or_mech;

/* One Thomas sweep along `axis` (0=x,1=y,2=z) over every line of every
 * substrate.  Restates shaders/diffusion_2d.metal:28-157 (x) and :165-292 (y);
 * z is the same per-line system with stride nx*ny.  nthreads>1 runs lines in
 * parallel (results are identical: lines are independent). */
void or_sweep(int axis, double* rho, const or_grid* g, const or_sub* subs, int nthreads);

/* Dirichlet re-apply on all faces (restates src/microenvironment.mm:127-152;
 * in 3-D the z faces are added). */
void or_dirichlet(double* rho, const or_grid* g, const or_sub* subs);

/* One diffusion substep: x, y, [z if nz>1] sweeps, then Dirichlet
 * (src/microenvironment.mm:106-115 + src/main.mm:343-347). */
void or_diffusion_substep(double* rho, const or_grid* g, const or_sub* subs, int nthreads);

/* Cell -> diffusion voxel index, 3-D formula (src/microenvironment.mm:176-184). */
void or_voxel_keys(const double* x, const double* y, const double* z, uint32_t n,
                   const or_grid* g, uint32_t* out);

/* Secretion/uptake for one diffusion substep, cells applied one after another
 * in ascending index (a single-thread run of src/main.mm:352-363 calling
 * src/cell_phenotype.cpp:526-564).  Rate arrays are [cell*ns + s]. */
void or_secretion(double* rho, const or_grid* g, double dt, uint32_t n,
                  const uint32_t* voxel, const double* volume,
                  const double* sec, const double* upt, const double* sat,
                  const double* net, const uint32_t* alive);

/* Same result, computed with cells grouped by voxel and groups processed in
 * parallel (the CPU-baseline form). */
void or_secretion_grouped(double* rho, const or_grid* g, double dt, uint32_t n,
                          const uint32_t* voxel, const double* volume,
                          const double* sec, const double* upt, const double* sat,
                          const double* net, const uint32_t* alive, int nthreads);

/* Mechanics voxel of each cell (shaders/mechanics.metal:54-72). */
void or_mech_keys(const double* x, const double* y, const double* z, uint32_t n,
                  const or_mech* m, uint32_t* out);

/* Stable counting sort of cells by key: offsets[nbins+1], sorted[n]
 * (ascending cell index inside each bin). */
void or_hash(const uint32_t* keys, uint32_t n, uint32_t nbins,
             uint32_t* offsets, uint32_t* sorted);

/* Repulsion/adhesion forces and simple pressure (shaders/mechanics.metal:91-216).
 * Writes v = F and pressure. */
void or_forces(const double* x, const double* y, const double* z,
               const double* radius, const double* rep, const double* adh,
               const double* rel_adh, uint32_t n, const uint32_t* keys,
               const uint32_t* offsets, const uint32_t* sorted, const or_mech* m,
               double* vx, double* vy, double* vz, double* pressure, int nthreads);

/* Neighbour lists: for each cell, the candidates that pass the force
 * kernel's distance predicates, in visiting order.  counts[n]; idx[n*cap]. */
void or_neighbors(const double* x, const double* y, const double* z, uint32_t n,
                  const uint32_t* keys, const uint32_t* offsets, const uint32_t* sorted,
                  const or_mech* m, uint32_t cap, uint32_t* counts, uint32_t* idx);

/* Adams-Bashforth-2 update, clamp to the mech-grid extent, v_prev <- v
 * (shaders/integrate.metal:31-80). */
void or_integrate(double* x, double* y, double* z,
                  const double* vx, const double* vy, const double* vz,
                  double* pvx, double* pvy, double* pvz,
                  uint32_t n, const or_mech* m, double dt, int nthreads);

/* Whole-simulation state for the scheduler restatement. */
typedef struct or_state {
    uint32_t n;
    double *x, *y, *z, *vx, *vy, *vz, *pvx, *pvy, *pvz;
    double *radius, *volume, *rep, *adh, *rel_adh, *pressure;
    double *sec, *upt, *sat, *net;     /* [cell*ns + s] */
    uint32_t *alive, *voxel, *mech_key;
    double* rho;                        /* [s][nz][ny][nx] */
} or_state;

typedef struct or_schedule {
    double dt_diff, dt_mech, dt_pheno, max_time;
} or_schedule;

typedef struct or_counts {
    int64_t substeps, mech_steps, pheno_steps;
} or_counts;

/* The step loop of src/main.mm:275-415 restricted to the hot path:
 * every iteration a diffusion substep + secretion; every dt_mech the hash,
 * forces and AB2; every dt_pheno the cell->voxel refresh (3-D formula).
 * Hash scratch: offsets[nbins+1], sorted[n]. */
void or_run(or_state* st, const or_grid* g, const or_sub* subs, const or_mech* m,
            const or_schedule* sch, uint32_t* hash_offsets, uint32_t* hash_sorted,
            int nthreads, or_counts* counts);

/* Single pieces of the loop, for benchmarks that time them separately. */
void or_mech_step(or_state* st, const or_mech* m, double dt_mech,
                  uint32_t* hash_offsets, uint32_t* hash_sorted, int nthreads);

/* Motility (restates updateMotility, src/motility.cpp:179-293): for every
 * live motile cell, the elapsed time grows by dt; the direction is resampled
 * when the persistence time is below dt or a uniform draw is below
 * dt/persistence; a resampled direction blends a random unit vector (circle
 * in 2-D, sphere otherwise) with the normalised chemotactic bias, then is
 * normalised and scaled by the cell's speed; the motility vector is added to
 * the velocity.  Non-motile live cells get a zero motility vector.
 * The bias is the sensitivity-weighted sum of central-difference gradients
 * over substrates (src/motility.cpp:220-236) when `sensitivity` is given
 * ([cell*ns + s]), else direction * gradient of substrate `chem_index`
 * (:237-250).  Gradients read the substrate-major field (the reference reads
 * density[voxel*ns + s], which equals it for one substrate).
 * Uniform draws are a counter-based stream: or_uniform(seed, step, cell, k)
 * for draw k = 0 (resample), 1 (theta), 2 (phi). */
double or_uniform(uint64_t seed, uint64_t step, uint32_t cell, uint32_t draw);
void or_motility(const double* x, const double* y, const double* z, double* vx, double* vy, double* vz,
                 const uint32_t* alive, const double* speed, uint32_t n, const or_grid* g, const double* rho,
                 double dt, uint64_t seed, uint64_t step, double* persistence, double* elapsed, const double* bias,
                 double* mvx, double* mvy, double* mvz, const uint32_t* motile, const uint32_t* restrict_2d,
                 const int32_t* chem_index, const int32_t* chem_dir, const double* sensitivity);

/* Phenotype-side SoA mutation (src/cell_phenotype.cpp:52-113, 571-720,
 * 833-844; src/cell_data.cpp:398-542, 544-640).  The cells are the library's
 * field arrays: f64[23] and u32[5] in HC_F_* order, rates[4] (secretion,
 * uptake, saturation, net export; [cell*ns + s]), vol[15] in HC_V_* order,
 * and optionally the motility state (persistence == NULL: none). */
typedef struct or_cells {
    double* f64[23];
    uint32_t* u32[5];
    double* rates[4];
    double* vol[15];
    double *persistence, *elapsed, *bias, *vec[3];
    uint32_t *motile, *restrict_2d;
    int32_t *chem_index, *chem_dir;
    double* sensitivity;
    uint32_t ns, _pad;
} or_cells;
/* updateVolume (src/cell_phenotype.cpp:52-113) on every i < n with mask[i]
 * != 0 (mask NULL: all); use_ratio as the reference's ratio-array test. */
void or_volume_update(or_cells* c, const uint32_t* mask, uint32_t n, double dt, int use_ratio);
/* divideCell (src/cell_phenotype.cpp:571-720) for parents[k] in order,
 * daughter n + k; the angle is 2 pi or_uniform(seed, counter, k, 0) and the
 * oncoprotein noise 0.02 level sqrt(-2 ln(1 - u1)) cos(2 pi u2), u1, u2 the
 * draws 1 and 2.  Divisions beyond max_cells are skipped.  Returns the new n. */
uint32_t or_divide(or_cells* c, const uint32_t* parents, uint32_t count, uint32_t n, uint32_t max_cells,
                   uint64_t seed, uint64_t counter);
/* The removal loop (src/cell_phenotype.cpp:833-844): indices in descending
 * order, each CellData::removeCell (src/cell_data.cpp:408-542; motility as
 * MotilityData::swapRemove, src/motility.cpp:154-175).  Returns n - count. */
uint32_t or_remove(or_cells* c, const uint32_t* indices, uint32_t count, uint32_t n);

/* writeCellNeighborGraph (src/output_writer.cpp:457-490), the O(N^2) pair
 * loop: for i < j, both alive with volume > 1e-15, |d|^2 <= (max(rel_i,
 * rel_j) (r_i + r_j))^2.  CSR by i (offsets[n+1]), j ascending; returns the
 * number of pairs (idx may be NULL to count). */
uint64_t or_contact_graph(const double* x, const double* y, const double* z, const double* radius,
                          const double* volume, const double* rel, const uint32_t* alive, uint32_t n,
                          uint32_t* offsets, uint32_t* idx);
/* tryFormAllAttachments (src/cell_interactions.cpp:411-478), the O(N^2)
 * pair loop: for live i with rate > 0 and fewer than OR_MAX_ATTACHMENTS
 * partners, every live j > i with rate > 0, ascending: skip when either is
 * full or j is already i's partner; else when |d| <= max(rel_i, rel_j)
 * (r_i + r_j) draw u (std::mt19937 seeded with seed, uniform [0, 1)) and
 * attach both ways when u < max(rate_i, rate_j) dt.  count[n] and
 * targets[n * OR_MAX_ATTACHMENTS] are updated in place; returns the number
 * of attachments formed. */
#define OR_MAX_ATTACHMENTS 12u
uint32_t or_form_attachments(const double* x, const double* y, const double* z, const double* radius,
                             const double* rel, const uint32_t* alive, const double* rate, uint32_t n, double dt,
                             uint64_t seed, uint32_t* count, uint32_t* targets);
/* The candidate scan of standardCellCellInteractions (src/cell_phenotype.cpp:
 * 893-967) over the hash (keys, off, srt of or_mech_keys / or_hash): the
 * first 512 cells j != i of the 3x3x3 bins around key[i] in visiting order,
 * kept when volume_j >= 1e-15, 2-D d^2 <= (2.5 r_i + r_j)^2 and d <= 1.25
 * (r_i + r_j).  CSR by i; returns the number of entries. */
uint64_t or_interaction_candidates(const double* x, const double* y, const double* radius, const double* volume,
                                   uint32_t n, const uint32_t* keys, const uint32_t* off, const uint32_t* srt,
                                   const or_mech* m, uint32_t* offsets, uint32_t* idx);

#ifdef __cplusplus
}
#endif

#endif
