/*
 * itr_hip.h -- C ABI of the MI355X (gfx950) batch evaluator for the inverse terrain
 * engine's f32 shallow-water candidate scoring.
 *
 * This is the drop-in boundary for the reference's population batch backend:
 *   reference trait  BatchBackend<f32>        crates/itr-opt/src/evaluator.rs:131-135
 *   reference impl   GpuSolver (wgpu/WGSL)    crates/itr-gpu/src/lib.rs:62-373
 *   reference glue   Gpu(GpuSolver)           crates/itr-cli/src/backend.rs:44-64
 * A Rust FFI crate binds these functions and implements BatchBackend<f32> with them.
 *
 * Conventions: plain pointers and sizes only; inputs are caller-owned and read only
 * during the call (copied to the device); outputs go to caller-allocated buffers.
 * A context is not thread-safe: use one per host thread.  Every entry point returns an
 * itr_status; ITR_DECLINED means "not supported here" and maps to the reference's
 * Err(...) decline, on which the evaluator falls back to the CPU for that level
 * (evaluator.rs:231-237).  err/errlen receive a NUL-terminated message (may be NULL).
 *
 * Results are bitwise equal to the reference's CPU f32 fused kernel
 * (crates/itr-hydro/src/kernel.rs:573-727 driven by solver.rs:346-660): state, step
 * counts, t_end, monitor maxima and times, ledger rows and abort decisions.  Grids run
 * with the local-inertial physics (crates/itr-hydro/src/inertial.rs) are bitwise equal
 * to the reference's f32 local-inertial run in all of these except two ledger sums that
 * the reference adds cell by cell in f64 (infiltration and the CFL-rate sum behind lts):
 * those agree to within f64 rounding.
 */
#ifndef ITR_HIP_H
#define ITR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ITR_OK = 0,
    ITR_DECLINED = 1, /* unsupported configuration: caller runs this level on the CPU */
    ITR_EINVAL = 2,   /* invalid arguments */
    ITR_EDEVICE = 3,  /* HIP runtime error */
    ITR_ENOMEM = 4    /* device allocation failed */
} itr_status;

typedef struct itr_ctx itr_ctx;
typedef struct itr_batch itr_batch;

/* Candidate designs materialized on the device: DesignSpace::materialize
 * (crates/itr-core/src/design.rs:316-385: 4x4 sub-sample quadrature of each primitive's
 * profile, design.rs:108-134, summed in primitive order, clamped, masked) on the fine grid,
 * and PreparedTerrain::apply_edit (crates/itr-hydro/src/terrain.rs:460-505): signed faces at
 * level 1, or at level >= 2 the coarse cells (coarse_cell, terrain.rs:357-373) and the
 * face-crest rises (terrain.rs:378-429) of the candidate's fine terrain.  The per-primitive
 * scalars that need libm (Primitive::segment, design.rs:97-105) and the cell windows
 * (design.rs:320-337) come from the host; every per-cell value is computed on the device
 * with the same f64 operations.  Arrays, windows and volumes are on the fine grid. */
typedef struct {
    const int32_t* prim_start; /* [n_cand + 1] offsets into prim */
    const double* prim;        /* [n][12]: profile (0 cut, 1 berm/swale/mound), segment x0, y0, x1,
                                  y1, width, height, cell window c0, r0, c1, r1 (half-open), 0 */
    const int32_t* bbox;       /* [n_cand][4]: union window c0, r0, w, h of the candidate (w or h 0: no edit) */
    const double* z_fine;      /* [ny][nx] of the fine grid: FineTerrain::z (re-based) */
    const uint8_t* wall;       /* [ny][nx] of the fine grid: 1 = wall cell */
    const uint8_t* editable;   /* [ny][nx] of the fine grid: 1 = editable, or NULL: every cell */
    double origin_x, origin_y, dx, dy; /* GeoTransform: x = origin_x + col dx, y = origin_y - row dy */
    double max_abs_dz, z_wall;
    double* volumes;           /* out: [n_cand][2] cut_m3, fill_m3 of each edit, or NULL */
    int32_t level;             /* fidelity level lv: 0 or 1, the grid is the fine grid (signed faces);
                                  >= 2, the grid is the face-crest level lv of the fine grid */
    int32_t fine_nx, fine_ny;  /* the fine grid at level >= 2: grid nx = ceil(fine_nx / lv), ny likewise */
    int32_t reserved;
} itr_design;

/* SolverParams (crates/itr-hydro/src/solver.rs:26-49).  Rain points are (t [s],
 * rate [m/s]) of the piecewise-linear hyetograph (crates/itr-core/src/scenario.rs:527-589).
 * locked_dt: baseline-locked dt schedule (solver.rs:41-44, 452-455) or NULL; when a
 * step's locked dt fails the hard-CFL check the candidate restarts adaptively from its
 * last sync point, as the reference does (solver.rs:456-480; itr_run_out.dt_fallback_t).
 * max_steps: stop each candidate after this many steps (0 = no limit; test control). */
typedef struct {
    double duration_s, sync_interval_s, cfl, dt_max_s, h_dry, h_eps;
    double infil_capacity; /* m/s */
    double initial_depth;
    double initial_stage; /* used when has_initial_stage */
    double cfl_hard;
    double inflow_scale; /* ensemble scaling of every inflow hydrograph (solver.rs:47-48, boundary.rs:26) */
    const double* rain_t;
    const double* rain_r;
    const double* locked_dt;
    int64_t n_locked;
    int64_t max_steps;
    int32_t n_rain;
    int32_t has_initial_stage;
    int32_t skip_dry; /* dry-row accounting rule of kernel.rs:605-612 (cell_updates) */
    int32_t physics;  /* KernelKind (solver.rs:19-24): ITR_PHYSICS_HLL (Fused) or
                         ITR_PHYSICS_LOCAL_INERTIAL (inertial.rs, the screening physics of
                         coarse fidelity levels, crates/itr-opt/src/problem.rs:178-181) */
} itr_solver_params;

enum { ITR_PHYSICS_HLL = 0, ITR_PHYSICS_LOCAL_INERTIAL = 1 };

/* Subdomain replay (crates/itr-hydro/src/replay.rs, spec §7.11.3).
 * mode 1 (record, Tape::record_ring / record_band): the grid's run records, for its
 *   candidate 0, the 1-cell ghost ring around the region of interest (c0, r0, w, h, in the
 *   grid's cells) at every step start, after the ghost fill, and the ROI's 1-cell inner
 *   band after every sync step: ring [ring_steps][2(w+2)+2h][3] and band
 *   [band_syncs][n_band][3] (h, qx, qy; replay.rs ring and band orders), capacities in
 *   steps / sync points.  *n_steps_out / *n_syncs_out receive the counts the run reached
 *   (a count above the capacity means the tape overflowed).
 * mode 2 (replay, evaluator.rs:427-440): the grid IS the w x h region (faces from
 *   sub_terrain, replay.rs:181-241: dzx / dzy include the faces between the region and its
 *   ring), its ghosts replay the ring of step k at step k, and after every sync step the
 *   band must stay within tol_h / tol_q of the tape (problem.rs:318: 2e-3 m, 1e-3 m^2/s),
 *   else aborted = 5 (ReplayInvalid; also when the tape is exhausted or a locked dt fails
 *   its check).  HLL physics only (the reference does not replay local-inertial levels). */
typedef struct {
    int32_t mode;
    int32_t c0, r0, w, h;
    int32_t reserved;
    float* ring;
    float* band;
    int64_t ring_steps, band_syncs;
    double tol_h, tol_q;
    int64_t* n_steps_out;
    int64_t* n_syncs_out;
} itr_replay;

/* One grid (a TerrainBase, crates/itr-hydro/src/terrain.rs:237-257) and the candidates
 * run on it (PreparedTerrain overlays flattened as in crates/itr-gpu/src/lib.rs:177-194).
 * Several grids of different shapes may be passed to one batch (fidelity ladder). */
typedef struct {
    int32_t nx, ny;
    double dx, dy, g, z_ref;
    int32_t crest;       /* FaceMode::Crest (coarse levels) */
    int32_t wet_count;   /* TerrainBase::wet_count */
    const float* wet;    /* [ny*nx] 1 = non-wall cell, 0 = wall (row-major, row 0 north) */
    const float* gn2;    /* [ny*nx] g*n^2 rounded to f32 */
    const int32_t* ghost; /* [2*nx + 2*ny] north[nx], south[nx], west[ny], east[ny];
                             0 wall, 1 transmissive, 2 stage, 3 inflow */
    int32_t n_cand;
    int32_t n_mon;
    const float* dzx;    /* [n_cand][ny][nx+1] fx[0] per face (signed jump, or crest d_l) */
    const float* dzy;    /* [n_cand][ny+1][nx] fy[0] */
    const float* dzx_r;  /* crest: [n_cand][ny][nx+1] fx[1] (d_r); NULL for signed */
    const float* dzy_r;  /* crest: [n_cand][ny+1][nx] fy[1] */
    const double* z_level; /* [n_cand][ny*nx] level z incl. edits (initial_stage), or NULL; not read
                              for a grid given as designs (the device builds the levels) */
    const uint32_t* mon_cells; /* MonitorSet: sorted unique interior cells row*nx+col */
    /* Exact guard early abort (Objective::guard_breach, crates/itr-core/src/objective.rs:132-144),
     * evaluated on the device at every sync point.  baseline_max NULL: disabled. */
    const double* baseline_max; /* [n_mon] */
    const int32_t* guard_pos;   /* positions into mon_cells, in guard order */
    int32_t n_guard;
    int32_t reserved;
    double guard_tol;
    /* Stage and Inflow boundary cells (Ghost::Stage / Ghost::Inflow as built by
     * Boundaries::from_segments, crates/itr-hydro/src/terrain.rs:122-187, and filled as
     * boundary.rs:43-47, 66-70).  May be NULL when no ghost code is 2 or 3.
     * ghost_value: [2*nx + 2*ny] in the order of `ghost`: stage level minus z_ref [m] for
     *              code 2 (the ghost depth uses z_level, which must then be given unless the
     *              grid is given as designs), or
     *              1 / segment length [1/m] for code 3.
     * ghost_hydro: [2*nx + 2*ny] hydrograph index of each code-3 cell.
     * hydro_start: [n_hydro + 1] offsets into hydro_t / hydro_q, the piecewise-linear
     *              hydrographs (t [s], Q [m^3/s]; Hyetograph rules, scenario.rs:540-563). */
    const double* ghost_value;
    const int32_t* ghost_hydro;
    const int32_t* hydro_start;
    const double* hydro_t;
    const double* hydro_q;
    int32_t n_hydro;
    int32_t reserved2;
    /* This grid's own solver parameters, or NULL for the call's.  An ensemble member
     * (crates/itr-opt/src/problem.rs:42-65: scaled rain, inflow_scale, its own locked dt
     * schedule and guard baseline) is one grid entry with the member's parameters, so all
     * members of a level run in one batch (the reference runs them on the CPU,
     * evaluator.rs:212-215, 494-560). */
    const itr_solver_params* params;
    /* Candidates given as designs (itr_design) instead of prepared faces, or NULL.  With a
     * design, dzx / dzy hold the grid's base faces (one copy: [ny][nx+1] / [ny+1][nx]); each
     * candidate's faces are built on the device.  Level-1 signed or level >= 2 face-crest
     * grids; z_level may be NULL (Stage ghosts and initial_stage read the levels built on
     * the device).  Not with replay (ITR_DECLINED: prepare the faces on the host). */
    const itr_design* design;
    /* Subdomain replay recording or replay for this grid (itr_replay), or NULL. */
    const itr_replay* replay;
} itr_grid;

/* Per-candidate result: GpuRun (crates/itr-gpu/src/lib.rs:49-60) + BatchRun
 * (evaluator.rs:118-125) + the RunSummary fields (crates/itr-core/src/model.rs:94-117). */
typedef struct {
    double t_end, volume_m3, dt_min, dt_max, dt_mean, lts, v0, rel_residual, abort_value;
    uint64_t steps, cell_updates, rows_skipped;
    int32_t bad;        /* Unhealthy (non-finite or beyond-roundoff negative depth) */
    int32_t aborted;    /* 0 none, 1 guard, 2 sync callback, 3 unhealthy, 5 replay invalid */
    int32_t abort_cell; /* guard: offending cell (row*nx+col) */
    int32_t abort_sync; /* sync index at the abort */
    int32_t sync_count;
    int32_t n_ledger;   /* ledger rows written */
    double* monitor_max;  /* [n_mon] or NULL */
    double* monitor_tmax; /* [n_mon] or NULL */
    float* depth;         /* [ny*nx] final depth or NULL */
    float* qx;            /* [ny*nx] or NULL (internal convention: +col east, +row south) */
    float* qy;
    double* ledger;       /* [max_ledger][8] rows (t, volume, rain, inflow, outflow, infil, fix, residual) */
    int32_t max_ledger;
    int32_t reserved;
    double dt_fallback_t; /* RunSummary::dt_fallback_t: the time a locked dt schedule failed its
                             hard-CFL check and the run restarted adaptively from its last sync
                             point (solver.rs:456-480), or -1 */
} itr_run_out;

/* Observer::on_sync (crates/itr-core/src/model.rs:59-61) for one candidate, called on
 * the calling thread whenever a candidate reaches a sync point (the candidate waits on
 * the device until the callback returns).  Return 0 to continue or non-zero to abort
 * the candidate (its result then reports aborted = 2 and *abort_value). */
typedef int32_t (*itr_sync_fn)(void* user, int32_t grid, int32_t cand, double t, uint64_t step,
                               uint32_t sync_index, const double* monitor_max, int32_t n_mon,
                               double* abort_value);

typedef struct {
    double wall_s;       /* host wall time of the last itr_hip_batch_run */
    double device_ms;    /* HIP-event time of the step loop on the library stream */
    double cell_steps;   /* sum over candidates of steps * cells actually run */
    uint64_t launches;   /* step-kernel launches */
    uint64_t syncs;      /* host round trips */
    int32_t band_rows, waves_per_task, steps_per_chunk, use_graph;
    double sim_s;        /* sum over candidates of the simulated time advanced [s] */
    int32_t launch_lists; /* step launch lists, each on its own stream (a class of grids or a
                            part of one, split by candidates) */
    int32_t reserved;
} itr_stats;

/* GpuSolver::new (lib.rs:85-124).  device: HIP device ordinal. */
itr_status itr_hip_create(int32_t device, itr_ctx** out, char* err, size_t errlen);
void itr_hip_destroy(itr_ctx* ctx);
/* GpuSolver::adapter_info (lib.rs:126-128). */
const char* itr_hip_device_info(itr_ctx* ctx);
/* Tuning knobs: "band_rows" (rows per workgroup task), "steps_per_chunk" (step launches
 * per host round trip), "use_graph" (0/1: capture the chunk in a hipGraph), "min_lds"
 * (bytes of LDS reserved per step workgroup, which caps occupancy), "adjoint_mode"
 * (reverse sweep of itr_hip_adjoint: 1 band march, 2 per-cell kernels; both give the
 * same bits), "adjoint_rows" (rows per band of the reverse band march; 0: 16; the same
 * bits at any height), "control_task" (who runs each candidate's ledger sums and writes its
 * control block: 1 a workgroup of its own beside the bands, 0 its first band after its
 * march, -1 chosen per launch list: 1 when the list's bands fit the chip at once; both
 * give the same bits).  0 = default (adjoint_mode: band march for large batches;
 * control_task: -1). */
itr_status itr_hip_set_option(itr_ctx* ctx, const char* name, int64_t value);

/* GpuSolver::run_batch (lib.rs:149-372) / BatchBackend::run (evaluator.rs:134):
 * upload, run every candidate of every grid to its end (or abort), download.
 * outs: one entry per candidate, grids concatenated in order. */
itr_status itr_hip_run_batch(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                             itr_sync_fn on_sync, void* user, itr_run_out* outs, char* err, size_t errlen);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * itr_hip_run_batch with every device buffer of the batch (state, faces, control blocks,
 * partial results, monitors, ledgers) carved from a static 64 MiB array that each device
 * holds from module load until the process ends.  It calls no hipMalloc, hipFree,
 * hipHostMalloc, graph instantiation or event creation: steps launch directly and the
 * run is not timed on the device (itr_stats is not reported).  Results are bitwise
 * those of itr_hip_run_batch.  A batch whose buffers exceed the array returns
 * ITR_ENOMEM with a message (use itr_hip_run_batch).  Every call on one device shares
 * that device's array: calls that overlap on one device race. */
itr_status itr_hip_run_batch_noalloc(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                                     itr_sync_fn on_sync, void* user, itr_run_out* outs, char* err, size_t errlen);

/* The same, split so that inputs can stay resident in HBM across runs (serving loops,
 * benchmarks): create uploads, run executes the step loop, results downloads. */
itr_status itr_hip_batch_create(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                                itr_batch** out, char* err, size_t errlen);
itr_status itr_hip_batch_run(itr_batch* b, itr_sync_fn on_sync, void* user, char* err, size_t errlen);
itr_status itr_hip_batch_results(itr_batch* b, itr_run_out* outs, char* err, size_t errlen);
/* Launch exactly n_steps step kernels (no sync callbacks; finished candidates idle) and
 * wait for them: times a fixed number of steps of a resident batch.  itr_hip_batch_stats
 * then reports the HIP-event time and the cell-steps actually run. */
itr_status itr_hip_batch_advance(itr_batch* b, int64_t n_steps, char* err, size_t errlen);
itr_status itr_hip_batch_stats(itr_batch* b, itr_stats* out);
void itr_hip_batch_destroy(itr_batch* b);

/* ---------------------------------------------------------------- discrete adjoint
 * No counterpart in the reference (its search is gradient-free): the gradient of the
 * peak-depth risk term J_risk = sum_a w_a tau softplus((peak_a - eta_a) / tau)
 * (crates/itr-core/src/objective.rs:70-128) with respect to the bed jump of every face
 * and, mapped through z_R - z_L, of every cell's elevation, for each candidate of one
 * grid.  The forward runs on the locked dt schedule p->locked_dt (solver.rs:452-455; a
 * frozen schedule makes the discrete step a fixed function of the state), stores
 * checkpoints every ceil(sqrt(N)) steps, replays each segment and sweeps it backwards;
 * adjoint arithmetic is f32, face-gradient accumulation f64.  p->max_steps > 0 bounds the
 * horizon (the run then ends after that many steps).  Signed levels and face-crest levels
 * (numerics.rs:117-129 with the faces of terrain.rs:378-429), wall, transmissive, Stage and
 * Inflow boundaries (a Stage ghost's depth moves with its boundary cell's level, which
 * dJ_dz_cell includes) and an initial stage (the initial depth moves with each cell's level
 * where the stage term is the maximum; needs z_level); local inertial physics, replay
 * grids and device designs return ITR_DECLINED.  A replayed segment that differs bitwise from its
 * checkpoint returns ITR_EDEVICE (no gradient).  The earthwork term and the chain rule to
 * the fine raster and to primitive heights are host work on the returned gradients. */
typedef struct {
    int32_t n_assets;
    int32_t reserved;
    const int32_t* asset_start;   /* [n_assets + 1]: offsets into asset_pos */
    const int32_t* asset_pos;     /* positions into the grid's monitor list (itr_grid.mon_cells) */
    const double* threshold_m;    /* [n_assets] eta_a */
    const double* weight;         /* [n_assets] w_a */
    double tau_m;
} itr_objective;

/* Earthwork term of one edit: TerrainEdit::cut_m3 / fill_m3 (crates/itr-core/src/design.rs:372-379)
 * and the ObjectiveDef costs and weight lambda (objective.rs:35-45). */
typedef struct {
    double cut_m3, fill_m3;
    double cut_cost, fill_cost;
    double earth_weight;
} itr_earthwork;

/* Objective::evaluate's J (objective.rs:95-114) of one candidate from its monitor maxima
 * (monitor_max[n_mon], positions as in itr_objective), which with running maxima is
 * Objective::running_lower_bound (objective.rs:146-149), the incumbent abort's bound
 * (crates/itr-opt/src/evaluator.rs:100-105).  Host only (no device work): a host without
 * the reference's Objective (this library's Python mirror, sync callbacks) uses it, and the
 * adjoint computes its J_risk with the same code (include/itr_objective.h).  earth may be
 * NULL (no earthwork term); j_risk_out and peaks_out[n_assets] may be NULL. */
itr_status itr_hip_objective_j(const itr_objective* obj, const itr_earthwork* earth, const double* monitor_max,
                               int32_t n_mon, double* j_out, double* j_risk_out, double* peaks_out, char* err,
                               size_t errlen);

typedef struct {
    double j_risk;                /* J_risk of the forward run */
    double* peak_m;               /* [n_assets] asset peaks, or NULL */
    double* dJ_dz_face_x;         /* [ny][nx + 1] dJ_risk / d(z_R - z_L) of x-faces; on a crest level
                                     dJ_risk / d(crest line of the face), 0 where a cell's level is
                                     the face's maximum; or NULL */
    double* dJ_dz_face_y;         /* [ny + 1][nx] of y-faces, or NULL */
    double* dJ_dz_cell;           /* [ny][nx] dJ_risk / dz of each cell's level (0 on walls; on a crest
                                     level with the crest lines held), or NULL */
    uint64_t steps;               /* N */
    int32_t segment_steps;        /* checkpoint spacing ceil(sqrt(N)) */
    int32_t checkpoints;
    uint64_t replay_mismatches;   /* replayed segment ends that differ bitwise from the checkpoint (0) */
    double plain_forward_ms;      /* the same forward without checkpoints (device time) */
    double forward_ms, replay_ms, reverse_ms;
    double* dJ_drise_face_x;      /* crest levels: [ny][nx + 1][2] dJ_risk / d(d_l), d(d_r) of x-faces,
                                     or NULL (not written on signed levels) */
    double* dJ_drise_face_y;      /* crest levels: [ny + 1][nx][2] of y-faces, or NULL */
} itr_adjoint_out;

itr_status itr_hip_adjoint(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grid, const itr_objective* obj,
                           itr_adjoint_out* outs, char* err, size_t errlen);

/* No counterpart in the reference (its wgpu kernel has fixed workgroup shapes): times
 * the step on these grids for band heights 1..16, 20, 24, 28, 32 (rows per workgroup) per grid shape,
 * by coordinate descent over shapes (largest work first), `trial_steps` launches per
 * trial.  The chosen heights are kept in the context and used by later batches whose
 * grids have the same shape (unless the "band_rows" option fixes one height). */
typedef struct {
    int32_t band_rows;
    int32_t reserved;
    double us_per_step;   /* device time per step of the whole batch at the chosen shapes */
} itr_tune_result;

itr_status itr_hip_autotune(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                            int32_t trial_steps, itr_tune_result* out, char* err, size_t errlen);

/* No counterpart in the reference: runs the step kernel's short correctly-rounded
 * reciprocal and square-root sequences on every f32 bit pattern and counts results that
 * differ from the IEEE operations (both counts are 0 for a correct build). */
itr_status itr_hip_math_selfcheck(itr_ctx* ctx, uint64_t* recip_mismatches, uint64_t* sqrt_mismatches, char* err,
                                  size_t errlen);

/* No counterpart in the reference: the dt sequence of one candidate's free forward (at
 * most `capacity` steps, or p->max_steps), i.e. the schedule solver.rs:452-455 locks to
 * when a baseline's dt is reused.  Writes *n_out values to dts. */
itr_status itr_hip_dt_schedule(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grid, int32_t cand,
                               double* dts, int64_t capacity, int64_t* n_out, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
