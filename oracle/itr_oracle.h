/*
 * itr_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a C++ restatement of the reference
 * engine's CPU solver path (Rust, not buildable here: no cargo/rustc, crates not
 * vendored).  It exists so the HIP path can be checked against it bitwise.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does.
 *
 * Every run takes f64 terrain inputs and rounds them to the working precision
 * exactly where the reference does (crates/itr-hydro/src/terrain.rs:262-321).
 */
#ifndef ITR_ORACLE_H
#define ITR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary segment (crates/itr-core/src/scenario.rs BoundarySegment).
 * side: 0 north, 1 south, 2 west, 3 east.  kind: 0 transmissive, 1 stage, 2 inflow. */
typedef struct {
    double range_a, range_b;   /* metres along the side; has_range = 0 => whole side */
    double stage_m;            /* absolute water-surface elevation (stage) */
    const double* hydro_t;     /* inflow hydrograph (t [s], q [m^3/s]) */
    const double* hydro_q;
    int32_t n_hydro;
    int32_t side;
    int32_t kind;
    int32_t has_range;
} orc_segment;

/* Sparse fine-resolution edit (crates/itr-core/src/design.rs TerrainEdit). */
typedef struct {
    const double* dz;          /* h*w, row-major */
    double cut_m3, fill_m3;
    int32_t c0, r0, w, h;
} orc_edit;

/* SolverParams (crates/itr-hydro/src/solver.rs:26-49) + test controls. */
typedef struct {
    double duration_s, sync_interval_s, cfl, dt_max_s, h_dry, h_eps;
    double infil_capacity;     /* m/s */
    double initial_depth;
    double initial_stage;      /* used when has_initial_stage */
    double cfl_hard;
    double inflow_scale;
    const double* rain_t;      /* hyetograph points (t [s], rate [m/s]) */
    const double* rain_r;
    const double* locked_dt;   /* baseline-locked dt schedule, or NULL */
    int64_t n_locked;
    int64_t max_steps;         /* test control: stop after this many steps (0 = none) */
    int32_t n_rain;
    int32_t has_initial_stage;
    int32_t kernel;            /* 0 = oracle (step_oracle), 1 = fused (step_fused) */
    int32_t skip_dry;
    int32_t strip_rows;        /* fused: rows per strip (<= 0: one strip) */
    int32_t precision;         /* 32 or 64 */
} orc_params;

/* Warm/custom start (crates/itr-hydro/src/solver.rs Checkpoint, subset). */
typedef struct {
    const double* h;           /* interior nx*ny, row-major */
    const double* qx;
    const double* qy;
    double t, dt_next, smax, v0;
    int64_t step;
    int32_t sync_index;
    int32_t pad_;
} orc_start;

/* Observer = the optimizer's GuardObserver (crates/itr-opt/src/evaluator.rs:88-108). */
typedef struct {
    const double* baseline_max;   /* per monitor */
    const int32_t* guard_pos;     /* positions into the monitor list, guard order */
    const int32_t* asset_pos;     /* concatenated asset monitor positions */
    const int32_t* asset_len;     /* cells per asset */
    const double* asset_thr;
    const double* asset_w;
    double guard_tol, tau, cut_cost, fill_cost, earth_weight;
    double edit_cut_m3, edit_fill_m3;
    double incumbent;
    int32_t n_guard, n_assets;
    int32_t require_no_offsite_worsening;
    int32_t has_incumbent;
} orc_observer;

typedef struct {
    double t_end, dt_min, dt_max, dt_mean, lts;
    double v0, rel_residual, abs_residual, gross;
    double abort_value;          /* guard exceedance or incumbent lower bound */
    double dt_fallback_t;        /* < 0 when no fallback happened */
    int64_t steps, cell_updates, rows_skipped;
    int32_t sync_count;
    int32_t n_rows;              /* ledger rows written */
    int32_t aborted;             /* 0 none, 1 guard, 2 incumbent, 3 unhealthy, 5 replay invalid */
    int32_t abort_cell;
    int32_t abort_sync;
    int32_t n_dts;               /* per-step dt values written */
} orc_summary;

typedef struct orc_terrain orc_terrain;

/* FineTerrain::new + TerrainBase::build (terrain.rs:204-321). */
orc_terrain* orc_terrain_build(int32_t nx, int32_t ny, double origin_x, double origin_y, double dx, double dy,
                               const double* dem, const uint8_t* walls, const double* manning, int32_t level,
                               const orc_segment* segs, int32_t nsegs);
void orc_terrain_free(orc_terrain* t);
/* info[0..3] = nx, ny, crest, wet_count ; dinfo[0..4] = dx, dy, z_ref, z_wall, g */
void orc_terrain_info(const orc_terrain* t, int32_t* info, double* dinfo);
/* Prepared faces (base + optional edit), rounded to `precision`, returned as f64 in a
 * padded (ny+2)*(nx+2) layout; wet/gn2 likewise; z_level interior nx*ny. Any NULL skipped. */
void orc_terrain_faces(const orc_terrain* t, const orc_edit* edit, int32_t precision, double* fx0, double* fx1,
                       double* fy0, double* fy1, double* wet, double* gn2, double* z_level);
/* Per-side ghost codes: 0 wall, 1 transmissive, 2 stage, 3 inflow; lengths nx,nx,ny,ny. */
void orc_terrain_ghosts(const orc_terrain* t, int32_t* north, int32_t* south, int32_t* west, int32_t* east);

/* Solver::run_impl (solver.rs:346-660).  Returns 0 on success. */
int32_t orc_run(const orc_terrain* base, const orc_edit* edit, const orc_params* p, const uint32_t* mon_cells,
                int32_t n_mon, const orc_observer* obs, const orc_start* start, orc_summary* sum,
                double* ledger_rows /* max_rows*8 */, int32_t max_rows, double* mon_max, double* mon_tmax,
                double* out_h, double* out_qx, double* out_qy, double* out_dts, int64_t max_dts);

/* Subdomain replay (crates/itr-hydro/src/replay.rs).  roi4 = c0, r0, w, h in level-grid
 * cells.  orc_record_tape runs the baseline (no edit) and returns the ghost ring
 * (3 values per ring cell per step, replay.rs ring order) and the inner band (3 values per
 * band cell per sync point).  orc_run_replay runs a candidate on the ROI grid with the
 * tape as its ghosts; aborted = 5 is ReplayInvalid (band beyond tol_h / tol_q, tape
 * exhausted, or a failed locked-dt check). */
int32_t orc_record_tape(const orc_terrain* base, const orc_params* p, const int32_t* roi4, const uint32_t* mon_cells,
                        int32_t n_mon, orc_summary* sum, double* ring, int64_t ring_cap, double* band, int64_t band_cap,
                        int64_t* n_ring, int64_t* n_band);
int32_t orc_run_replay(const orc_terrain* base, const orc_edit* edit, const orc_params* p, const int32_t* roi4,
                       const double* ring, int64_t n_steps, const double* band, int64_t n_syncs, double tol_h, double tol_q,
                       const uint32_t* mon_sub, int32_t n_mon, const orc_observer* obs, orc_summary* sum,
                       double* ledger_rows, int32_t max_rows, double* mon_max, double* mon_tmax, double* out_h,
                       double* out_qx, double* out_qy);

/* Candidate-parallel CPU evaluation: one candidate per worker thread, each worker running
 * the fused kernel with a single-threaded workspace (evaluator.rs:337-352).  Fills the
 * per-candidate step counts and returns the wall time in seconds. */
double orc_run_candidates(const orc_terrain* base, const orc_edit* edits, int32_t n_cand, const orc_params* p,
                          const uint32_t* mon_cells, int32_t n_mon, int32_t threads, int64_t* steps_out,
                          int64_t* cell_updates_out);

/* Primitives (design.rs:95-142) and materialisation (design.rs:316-385). kind: 0 berm,
 * 1 swale, 2 cut, 3 mound. prims: n*7 doubles (kind, cx, cy, length, width, angle_deg,
 * height).  Writes bbox (c0, r0, w, h) and dz (capacity max_dz_len); returns cells written
 * or -1 when dz_cap is too small. */
int64_t orc_materialize(const double* prims, int32_t n_prims, double origin_x, double origin_y, double dx, double dy,
                        int32_t nx, int32_t ny, const uint8_t* editable, double max_abs_dz, int32_t* bbox,
                        double* dz, int64_t dz_cap, double* cut_fill);
double orc_profile(const double* prim, double x, double y);

/* Hyetograph::rate / integral (scenario.rs:540-585). */
double orc_hyeto_rate(const double* ts, const double* rs, int32_t n, double t);
double orc_hyeto_integral(const double* ts, const double* rs, int32_t n, double a, double b);

/* Real primitives (real.rs) exposed for known-answer tests. */
float orc_rcbrt_f32(float x);
double orc_rcbrt_f64(double x);
float orc_fixedsum_f32(const float* v, int64_t n, int32_t mode /* 0 add, 1 add_slice */);
float orc_max_slice_f32(const float* v, int64_t n);
float orc_rmax_f32(float a, float b);

/* Objective::evaluate J (objective.rs:95-128) with C libm exp/log1p. */
double orc_objective_j(const orc_observer* o, const double* monitor_max);

#ifdef __cplusplus
}
#endif
#endif
