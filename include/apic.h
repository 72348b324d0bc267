/*
 * apic.h -- C ABI of the MI355X-native APIC time step (libapic_hip).
 *
 * The reference exposes a Rust module API, `FLIP::apic` (src/lib.rs:1, crate
 * name Cargo.toml:2), with no FFI.  Each entry point below replaces one public
 * item of that API; the comment names the item it stands in for.
 *
 * Conventions
 *  - Every call returns an apic_status (0 = ok).  apic_last_error() gives the
 *    message of the most recent failure on the calling thread.
 *  - A handle owns all device memory of one simulation and one HIP stream.
 *    One thread at a time per handle (the reference takes `&mut self`).
 *  - Caller buffers stay caller-owned; reads and writes are synchronous copies.
 *  - Grid indexing follows the reference: cell (x,y) -> y*W+x, u face (i,j) ->
 *    j*(W+1)+i, v face (i,j) -> j*W+i (src/apic.rs:977-991).  Solver arrays
 *    (DIAG..MATRIX_TIMES_SEARCH) are padded by one row on each side:
 *    index = raw + W, length ceil8(W*H + 2W) (src/apic.rs:428-453).
 *  - Particle arrays are flat SoA of length num_chunks*8 (num_chunks =
 *    ceil(N/8)); particle i of the reference's AoSoA layout is chunk i/8,
 *    lane i%8 (src/apic.rs:757-767, 876-898).
 */
#ifndef APIC_H
#define APIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum apic_status {
    APIC_OK = 0,
    APIC_ERR_INVALID_ARGUMENT = 1,
    APIC_ERR_UNSUPPORTED_SIZE = 2, /* no multigrid level can be built (the reference panics, src/apic.rs:910) */
    APIC_ERR_OUT_OF_MEMORY = 3,
    APIC_ERR_HIP = 4,
    APIC_ERR_NO_DEVICE = 5
} apic_status;

typedef struct apic_handle apic_handle;

/* WorldProperties (src/apic.rs:707-713). */
typedef struct apic_world_properties {
    float gravity;
    float border_damping;
    float cfl;
    float solve_error;
} apic_world_properties;

/* DebugOptions (src/apic.rs:397-400). */
typedef struct apic_debug_options {
    uint8_t disable_particle_sort;
} apic_debug_options;

/* Pressure-solver rules.  mode 0 ("reference", the default) reproduces
 * solve_pcg exactly: absolute 1e-5 exits on r.z and s.As, infinity-norm
 * residual test against WorldProperties.solve_error, 100-iteration cap, FP32
 * reductions (src/apic.rs:8, 1865-1956).  mode 1 ("relative") is an extension
 * the reference cannot express: the same multigrid-preconditioned CG, stopped
 * when ||r||_2 <= rel_tol * ||b||_2 or after max_iterations (1..100000), with
 * breakdown exits when s.As <= 0 or r.z <= 0; the preconditioner's vectors
 * stay FP32, the dot products, ||r||, ||b|| and the step scalars are FP64
 * (fp64_reductions must be 1), each row of A s is summed in FP64 and kept
 * as its FP32 rounding (APIC_MATRIX_TIMES_SEARCH) plus the FP32 rest, the
 * pressure is accumulated in FP64 (APIC_PRESSURE_F64), and the residual is
 * the FP32 r that the preconditioner reads plus a per-cell FP32 low part
 * that keeps what r's roundings drop, so the test sees the true residual
 * b - A p of the FP64 pressure to FP64 rounding; b - A p itself (FP64)
 * replaces it every recompute_every iterations (0 = never).
 * pcg_rel_residual reports ||r + r_lo||_2 / ||b||_2 of the residual the
 * test used.  The FP32 pressure (APIC_PRESSURE) cannot have a true residual
 * much below eps32 * || |A| |p| || / ||b||. */
typedef struct apic_solver_config {
    int32_t mode;
    float rel_tol;
    int32_t max_iterations;
    int32_t fp64_reductions;
    int32_t recompute_every;
} apic_solver_config;

/* Counters of the most recent apic_update call. */
typedef struct apic_step_stats {
    uint32_t substeps;              /* substeps taken by the last update */
    uint32_t pcg_iterations;        /* PCG iterations of the last substep */
    uint32_t pcg_iterations_total;  /* summed over the last update's substeps */
    uint32_t pcg_exit_reason;       /* 0 running, 1 gamma<eps at start, 2 |delta|<eps,
                                       3 residual below tolerance, 4 gamma'<eps, 5 iteration cap */
    float pcg_max_residual;         /* infinity norm of r at exit */
    double pcg_rel_residual;        /* ||r||2/||b||2 at exit (relative mode), else -1 */
    uint32_t fluid_cells;           /* base-level fluid cells of the last substep */
    uint32_t timestamp;             /* Apic::timestamp after the update */
    float prev_deltatime;           /* Apic::prev_deltatime after the update */
    float max_particle_velocity;    /* value that set the last substep's dt */
} apic_step_stats;

/* Fields readable / writable for host mirrors and per-stage tests.
 * Element type F32 unless noted: U32 for the sort record, U8 (0/1) for flags,
 * F64 for APIC_PRESSURE_F64. */
typedef enum apic_field {
    APIC_POS_X = 0, APIC_POS_Y, APIC_VEL_X, APIC_VEL_Y,
    APIC_C_U_X, APIC_C_U_Y, APIC_C_V_X, APIC_C_V_Y, APIC_TIME_RESIDUAL, /* num_chunks*8 */
    APIC_SORT_KEY,            /* U32, num_particles */
    APIC_SORT_VAL,            /* U32, num_particles */
    APIC_MAC_U,               /* (W+1)*H */
    APIC_MAC_V,               /* W*(H+1) */
    APIC_VALID_U,             /* U8, (W+1)*H */
    APIC_VALID_V,             /* U8, W*(H+1) */
    APIC_DENSITY,             /* W*H */
    APIC_SMOOTHED_DENSITY,    /* W*H */
    APIC_PRESSURE,            /* W*H */
    APIC_FLUID,               /* U8, one per cell of the level */
    APIC_OBSTACLE,            /* U8, W*H */
    APIC_OPEN_U,              /* U8, (W+1)*H; the reference's mac_u_blocked bit (1 = open) */
    APIC_OPEN_V,              /* U8, W*(H+1) */
    APIC_DIAG, APIC_INV_DIAG, APIC_PLUS_X, APIC_PLUS_Y,
    APIC_RESIDUAL, APIC_ERROR,                 /* padded, per level */
    APIC_SEARCH, APIC_MATRIX_TIMES_SEARCH,     /* padded, base level */
    APIC_EXT_FORCE_U, APIC_EXT_FORCE_V,        /* u faces, v faces */
    APIC_POST_FORCE_U, APIC_POST_FORCE_V,
    APIC_PRESSURE_F64,        /* F64, W*H: the relative mode's FP64-accumulated pressure (extension) */
    APIC_RHS,                 /* padded, base level: the relative mode's copy of b (extension) */
    APIC_FIELD_COUNT
} apic_field;

/* ---- lifetime ---------------------------------------------------------- */

/* Apic::new (src/apic.rs:779-916): walls one cell thick, 3x3 stratified
 * particles in cells 4<x<W-4, 4<y<H/2, multigrid levels while both sides
 * are even and > 4, open-face masks (update_world, :918-944). */
apic_status apic_create(uint32_t width, uint32_t height, float cell_size,
                        const apic_world_properties* world_properties,
                        apic_handle** out);

/* Apic::new followed by replacing the particle set with n explicit positions
 * (xy interleaved x0,y0,x1,y1,...), zero velocity, C and time residual; the
 * last chunk's tail lanes repeat the last particle as the constructor does
 * (src/apic.rs:851-898).  The reference can only be driven this way by
 * overwriting its pub particle fields after new. */
apic_status apic_create_with_particles(uint32_t width, uint32_t height, float cell_size,
                                       const apic_world_properties* world_properties,
                                       const float* xy, uint32_t n,
                                       apic_handle** out);

/* Drop of Apic: frees device memory and the stream. */
void apic_destroy(apic_handle* h);

/* Message of the last failed call on this thread ("" if none). */
const char* apic_last_error(void);

/* ---- the step ------------------------------------------------------------ */

/* Apic::update (src/apic.rs:2429-2474). */
apic_status apic_update(apic_handle* h, const apic_debug_options* debug_options,
                        float frame_deltatime);

/* Apic::add_radial_force (src/apic.rs:2352-2410). */
apic_status apic_add_radial_force(apic_handle* h, float world_x, float world_y,
                                  float radius, float strength);

/* Apic::get_max_particle_velocity (src/apic.rs:2412-2427). */
apic_status apic_get_max_particle_velocity(apic_handle* h, float* out);

/* Apic::is_fluid (src/apic.rs:1015-1022). */
apic_status apic_is_fluid(apic_handle* h, int32_t i, int32_t j, uint64_t* out);

/* The reference's pub particle fields in its own layout (src/apic.rs:757-767):
 * Vec<Vec2x8> for positions, velocities, c_u and c_v -- per chunk of eight
 * particles eight x then eight y, 16 floats -- and Vec<f32x8> for the time
 * residual, 8 floats per chunk.  count = 16 * num_chunks or 8 * num_chunks.
 * Lanes past num_particles in the last chunk are included, as in the
 * reference's vectors.  Reads synchronise the library's stream. */
typedef enum apic_particle_group {
    APIC_PART_POSITIONS = 0,
    APIC_PART_VELOCITIES = 1,
    APIC_PART_C_U = 2,
    APIC_PART_C_V = 3,
    APIC_PART_TIME_RESIDUAL = 4
} apic_particle_group;
apic_status apic_read_particles_aosoa(apic_handle* h, apic_particle_group group, float* dst, size_t count);
apic_status apic_write_particles_aosoa(apic_handle* h, apic_particle_group group, const float* src, size_t count);

/* Solver rules (extension; default is mode 0, the reference's rules). */
apic_status apic_set_solver_config(apic_handle* h, const apic_solver_config* cfg);

apic_status apic_get_stats(apic_handle* h, apic_step_stats* out);

/* Stage timers (extension, for benchmarking; no counterpart in the reference).
 * With profiling on, the library records HIP events on its own stream around
 * each stage of every substep, and around the two transfer kernels alone, and
 * accumulates the elapsed milliseconds.  Enabling resets the accumulators. */
typedef enum apic_timer {
    APIC_T_SORT = 0, APIC_T_P2G, APIC_T_PRESSURE_BUILD, APIC_T_MG_BUILD, APIC_T_PCG,
    APIC_T_GRADIENT, APIC_T_EXTRAPOLATE, APIC_T_G2P,
    APIC_T_K_P2G,   /* the P2G kernel alone */
    APIC_T_K_G2P,   /* the G2P gather/advect kernel alone */
    APIC_T_K_SPMV,  /* the PCG's SpMV + dot kernel alone (iteration 1 of each solve) */
    APIC_T_K_PCG_ITER, /* one whole PCG iteration: SpMV, update and V-cycle (iteration 1 of each solve) */
    APIC_TIMER_COUNT
} apic_timer;

/* enable: 0 off, 1 the kernel timers only (APIC_T_K_P2G, APIC_T_K_G2P,
 * APIC_T_K_SPMV, APIC_T_K_PCG_ITER), 2 every timer.  Resets the timers. */
apic_status apic_set_profiling(apic_handle* h, int enable);
/* ms[i] and counts[i] for i < n (n <= APIC_TIMER_COUNT); either may be null. */
apic_status apic_get_timers(apic_handle* h, double* ms, uint32_t* counts, uint32_t n);

/* Launch shapes that do not change a bit of the result (extension; the
 * reference has no counterpart: its loops run on host threads).
 *   radix_bits   widest digit of the Morton radix sort, 8..11: a key of b
 *                bits takes ceil(b / radix_bits) passes (the sort is
 *                stable, so every width gives the same order);
 *   few_tiles    a multigrid level above the base with at most this many
 *                32 x 32 tiles runs as 16 x 16 tiles, and two such levels
 *                above the bottom share one launch each way (the per-cell
 *                arithmetic is the same for every tile side).
 * sort_ms and vcycle_ms are the autotuner's timings of the chosen shapes
 * (0 when set by hand).  Defaults: radix_bits 11, few_tiles 1024. */
typedef struct apic_tuning {
    uint32_t radix_bits;
    uint32_t few_tiles;
    float sort_ms;
    float vcycle_ms;
} apic_tuning;
apic_status apic_get_tuning(apic_handle* h, apic_tuning* out);
apic_status apic_set_tuning(apic_handle* h, const apic_tuning* t);
/* Times the radix sort of the current particles' keys for radix_bits 8..11
 * and apply_preconditioner on the current pressure system for few_tiles in
 * {0, 64, 256, 1024, 4096, 16384}, reps times each after one untimed run,
 * keeps the fastest of each and writes them (with their timings) to out if
 * not null.  The particle state is left as it was; the preconditioner's
 * error vectors are overwritten.  Call it after a few updates, when the
 * particles and the pressure system are representative of the run. */
apic_status apic_autotune(apic_handle* h, uint32_t reps, apic_tuning* out);

/* Horizontal gravity (extension; the reference applies gravity along y only,
 * src/apic.rs:1604-1626).  gx is added, times dt, to every u face with a fluid
 * cell on either side, as WorldProperties.gravity is to v faces; 0 (the
 * default) leaves the step identical to the reference's. */
apic_status apic_set_gravity_x(apic_handle* h, float gx);

/* ---- batched stepping (extension; no counterpart in the reference, whose
 * simulations are stepped one at a time) ------------------------------------- */

/* A batch owns one HIP stream.  Simulations created in it share that stream
 * and accept every apic_* call; apic_batch_update steps any subset of them
 * together: each kind of kernel launch of the step is issued once for all
 * of them, over every simulation's workgroups.  Each simulation computes
 * exactly what it computes when stepped alone with apic_update (the same
 * kernels on the same data; the PCG iterations run in lockstep, and a
 * simulation whose solve has ended skips the rest).  Calls on one batch and
 * its simulations must come from one thread at a time. */
typedef struct apic_batch apic_batch;

apic_status apic_batch_create(apic_batch** out);
/* Fails with APIC_ERR_INVALID_ARGUMENT while simulations of the batch exist. */
apic_status apic_batch_destroy(apic_batch* b);
/* apic_create / apic_create_with_particles in the batch (xy null: the
 * reference scene, n ignored). */
apic_status apic_batch_create_sim(apic_batch* b, uint32_t width, uint32_t height, float cell_size,
                                  const apic_world_properties* world_properties, const float* xy,
                                  uint32_t n, apic_handle** out);
/* Apic::update(opts, frame_dt[i]) for sims[0..count), count <= 4096, every
 * simulation of this batch and none twice.  Returns when the last substep's
 * work is queued (as apic_update). */
apic_status apic_batch_update(apic_batch* b, apic_handle* const* sims, uint32_t count,
                              const apic_debug_options* opts, const float* frame_dt);
/* apic_batch_update without any device allocation (extension; no counterpart
 * in the reference).  For callers that cannot allocate device memory while
 * they step: serving loops that must not stall on hipMalloc/hipFree, code
 * that forbids allocation after set-up.  apic_batch_update uploads each
 * substep's launch tables into a device arena that the batch grows with
 * hipMalloc; this variant puts them in a static device array of
 * APIC_NOALLOC_ARENA_BYTES bytes that the library holds on every GPU it runs
 * on until the process ends, and computes the same bits.  An update whose
 * tables need more fails with APIC_ERR_INVALID_ARGUMENT before it steps any
 * simulation (tables take about 20-80 KB per simulation, by grid size).  Every
 * call on one GPU shares that array: calls on one GPU are serialised on the
 * host, and each call's stream waits for the previous call's kernels before
 * it rewrites the array, so calls from several threads or batches are
 * ordered, not overlapped. */
#define APIC_NOALLOC_ARENA_BYTES (32u << 20)
apic_status apic_batch_update_noalloc(apic_batch* b, apic_handle* const* sims, uint32_t count,
                                      const apic_debug_options* opts, const float* frame_dt);
/* Device memory (bytes) the simulation holds, for admission by free HBM. */
apic_status apic_device_bytes(apic_handle* h, size_t* bytes);
/* Free and total memory of the current HIP device (hipMemGetInfo). */
apic_status apic_device_memory(size_t* free_bytes, size_t* total_bytes);

/* ---- pub fields (src/apic.rs:715-776) --------------------------------------- */

/* width, height, num_particles, num_chunks, multigrid level count (excluding
 * the base grid), fluid_cells, timestamp. */
apic_status apic_get_dims(apic_handle* h, uint32_t* width, uint32_t* height,
                          uint32_t* num_particles, uint32_t* num_chunks,
                          uint32_t* num_levels);
/* Multigrid levels (excluding the base grid) of each solver mode's V-cycle:
 * `reference` is the reference's halving while both sides are even and > 4
 * (src/apic.rs:900-908, the num_levels above); `relative` adds the relative
 * mode's levels below it, each side halved and rounded up while the coarsest
 * has more than 64 cells and both sides exceed 2 (a side of 2 x odd stops the
 * reference's halving early).  Their fields are read and written by level
 * index like the others.  No counterpart in the reference. */
apic_status apic_get_level_counts(apic_handle* h, uint32_t* reference, uint32_t* relative);
apic_status apic_get_timestamp(apic_handle* h, uint32_t* timestamp, float* prev_deltatime);
apic_status apic_set_timestamp(apic_handle* h, uint32_t timestamp, float prev_deltatime);

/* Element count of a field; level 0 is the base grid, level k>0 is
 * multigrid_levels[k-1] (APIC_FLUID and the padded solver fields only). */
apic_status apic_field_length(apic_handle* h, apic_field field, uint32_t level, size_t* count);
apic_status apic_read_field(apic_handle* h, apic_field field, uint32_t level,
                            void* dst, size_t count);
apic_status apic_write_field(apic_handle* h, apic_field field, uint32_t level,
                             const void* src, size_t count);

/* ---- pub stage methods (run one stage on the device state) ------------------ */

apic_status apic_update_spatial_lookup(apic_handle* h);                       /* :1043-1111 */
apic_status apic_transfer_particles_to_grid(apic_handle* h, float deltatime,
                                            uint32_t timestamp);              /* :1218-1497 */
apic_status apic_apply_external_forces(apic_handle* h, float deltatime);     /* :1604-1626 */
apic_status apic_build_pressure_system(apic_handle* h);                      /* :1640-1785 */
apic_status apic_update_multilevel_grid(apic_handle* h);                     /* :1499-1602 */
apic_status apic_apply_preconditioner(apic_handle* h);                       /* :1788-1821 */
apic_status apic_solve_pcg(apic_handle* h);                                  /* :1865-1956 */
apic_status apic_apply_pressure_gradient(apic_handle* h);                    /* :1958-2055 */
apic_status apic_extrapolate_velocity(apic_handle* h);                       /* :2266-2350 */
apic_status apic_transfer_grid_to_particles_and_advect(apic_handle* h,
                                                       float deltatime);     /* :2057-2264 */

#ifdef __cplusplus
}
#endif

#endif /* APIC_H */
