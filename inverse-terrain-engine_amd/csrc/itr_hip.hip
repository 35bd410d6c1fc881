// itr_hip.hip -- gfx950 batch evaluator for the f32 shallow-water candidate scoring,
// behind the C ABI of include/itr_hip.h.
//
// Execution model
//   * One step of every active candidate of every grid is ONE launch of the step kernel
//     (grids of different widths and face modes share the launch).  A workgroup owns a
//     band of R full rows of one candidate and marches down it (register-resident
//     sliding window in y); waves tile the row in x with 62 updated columns and one halo
//     column on each side, exchanging neighbour values with DPP (no LDS traffic, no
//     barrier inside the march).
//   * The launch boundary is the only cross-workgroup synchronisation (no device-scope
//     fences: on gfx950 those write back and invalidate the L2).  Each workgroup starts
//     by completing its candidate's previous step from that step's results (CFL max and
//     health flag, monitors): f64 time, dt, sync clipping, exact hyetograph integral,
//     guard abort, termination and the next step's f32 constants.  Every workgroup of
//     the candidate computes the same decision from the same inputs; the candidate's
//     control task (a workgroup with no band) also does the row-ordered ledger sums and
//     writes the new control block, while the bands march.
//   * Everything a launch reads from the previous one is double-buffered (control
//     blocks by launch parity; state, row partials, boundary masses and monitors by the
//     candidate's state parity; per-step atomics in a ring of four by launch index).
//   * The host launches chunks of steps (a captured hipGraph) and polls a status word
//     per candidate between chunks (finished, or paused for a sync callback).
//
// HBM layout (per grid; P = row pitch in floats, multiple of 64):
//   state  [cand][2][3][ny][P]  h, qx, qy (two copies, by state parity)
//   dzx    [cand][ny][P]        west-face bed jump of cell (col); col 0 = 0
//   dzy    [cand][ny][P]        north-face bed jump of row (row); row 0 = 0
//   dzx2/dzy2                   crest mode second rises, same layout
//   wet, gn2 [ny][P]            shared by every candidate of the grid
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/itr_hip.h"
#include "../../include/itr_objective.h"
#include "itr_numerics.h"
#include "itr_adjoint.h"

using namespace itr;

namespace {

constexpr int kWave = 64;
constexpr int kColsPerWave = 62;   // updated columns per wave (lanes 1..62)
// Row groups the last wave of an HLL grid marches side by side: a lane group holds the
// last wave's t columns plus a halo lane each side, so 64 / (t + 2) groups fit (at most 8).
constexpr int tail_group_count(int nx, int waves) {
    const int t = nx - (waves - 1) * kColsPerWave;
    const int s = 64 / (t + 2);
    return s < 2 ? 1 : (s > 8 ? 8 : s);
}
constexpr int kMaxWaves = 16;      // 1024 threads: rows up to 992 columns

constexpr uint32_t F_DONE = 1u;
constexpr uint32_t F_PAUSED = 2u;

// Row partial results of one step, consumed by the candidate's control.
struct RowAcc {
    float infil, fix;      // FixedSum totals (real.rs:132-175) of the row
    float bw, be;          // west / east boundary face mass flux (kernel.rs:436-450)
    float vol, sum_s;      // sync steps: FixedSum of new depth and of the CFL rate
    uint32_t wet;          // any new depth > 0 (kernel.rs:297)
    uint32_t pad;
};

// The same 32 bytes for a local-inertial grid: the band's f64 sums of infiltration and
// CFL rate sit in its first row (0 in the others), the FixedSum of the new depth and the
// mask of waves whose positivity-fix values are in the grid's fix64 buffer in every row.
struct LiRowAcc {
    double infil, sum_s;
    float vol;
    uint32_t fixmask;
    uint32_t wet;          // unused (no dry-row rule in the local-inertial step)
    uint32_t pad;
};
static_assert(sizeof(LiRowAcc) == sizeof(RowAcc), "row partials share one buffer");

struct DevParams {
    double duration, sync, cfl, dt_max, cfl_hard, infil_cap, inflow_scale;
    const double* rain_t;
    const double* rain_r;
    const double* locked;
    int64_t n_locked;
    int64_t max_steps;
    int32_t n_rain;
    uint32_t n_sync;
    int32_t skip_dry;
    int32_t has_cb;
    int32_t max_ledger;
    float h_dry, heps2, q_min;
};

struct DevGrid {
    DevParams prm;             // this grid's solver parameters (ensemble members differ)
    int32_t nx, ny, P, R, nbands, cand0, ncand, n_mon, n_guard, crest, wet_count, waves;
    int32_t nslot;             // max(ncand, 1): stride of the double-buffered per-candidate arrays
    int32_t max_log;           // capacity of steplog per candidate (0: no log)
    int32_t li;                // local-inertial physics (inertial.rs) instead of HLL
    // HLL grids whose last wave updates few columns: that wave marches tail_groups row
    // groups of each band side by side, in lane groups of 64 / tail_groups lanes (1: off).
    int32_t tail_groups;
    size_t plane;              // ny * P
    size_t plane_y;            // (ny + 1) * P: dzy rows 0..ny (rows 0 and ny are zero)
    float* state;              // [ncand][2][3][plane]
    const float* dzx;
    const float* dzy;
    const float* dzx2;
    const float* dzy2;
    const float* wet;
    const float* gn2;
    const int32_t* mon_row;    // [ny + 1] first position in mon_cells of each row (cells sorted row-major)
    const uint32_t* mon_off;   // [n_mon] offset row * P + col of each monitored cell in a state plane
    const int32_t* ghost;
    // Stage / Inflow ghosts (codes 2 / 3); special = 0 when the grid has none.
    int32_t special, n_hydro;
    const float* gstage;       // [nslot][2nx + 2ny] f32 stage depth of code-2 cells per candidate
    const double* ginv;        // [2nx + 2ny] 1 / segment length of code-3 cells
    const int32_t* ghydro;     // [2nx + 2ny] hydrograph of code-3 cells
    const int32_t* hstart;     // [n_hydro + 1] offsets into hyd_t / hyd_q
    const double* hyd_t;
    const double* hyd_q;
    float* mon_max;            // [2][nslot][n_mon]   by state parity
    double* mon_tmax;          // [2][nslot][n_mon]
    RowAcc* rowacc;            // [2][nslot][ny]      by state parity
    float* ybnd;               // [2][nslot][2][nx]   top / bottom boundary y-face mass
    float* sbuf;               // [nslot][plane] CFL rate at sync steps (within one workgroup)
    float* fixbuf;             // [nslot][plane] positivity-fix values of rows that have one
    double* ledger;            // [nslot][max_ledger][8]
    float* lif;                // local-inertial: [2][nslot][ny + nx] west then north boundary faces
    // Locked dt schedule: the state at the last sync point, for the adaptive restart
    // (solver.rs:446-480).  NULL when the grid's parameters have no schedule.
    float* snap_state;         // [nslot][3][plane]
    float* snap_mon;           // [nslot][n_mon]
    double* snap_tmax;         // [nslot][n_mon]
    RowAcc* snap_rows;         // [nslot][ny]
    float* snap_lif;           // local-inertial: [nslot][ny + nx]
    struct CandCtl* snap_ctl;  // [nslot]
    // Subdomain replay (replay.rs).  tape_mode 1: this run records the tape of the ROI
    // (roi_*) for its candidate 0; 2: this grid is the ROI and its ghosts (code 4) replay
    // the tape, with the band checked at every sync point.
    int32_t tape_mode, roi_c0, roi_r0, roi_w, roi_h, n_ring, n_band;
    float* tape_ring;          // [tape_steps][n_ring][3] h, qx, qy
    float* tape_band;          // [tape_syncs][n_band][3]
    int64_t tape_steps, tape_syncs;
    double tol_h, tol_q;
    double* fix64;             // local-inertial: [2][nslot][plane] positivity-fix values (f64), by state parity
    struct StepRec* steplog;   // [nslot][max_log] constants of every step run (adjoint replay)
    const double* baseline_max;
    const int32_t* guard_pos;
    const uint32_t* mon_cells;
    double guard_tol, dx, dy, g, area;
    float g32, g_half32, inv_dx32, inv_dy32;
};

__host__ __device__ inline size_t slot(const DevGrid& G, int par, int local) { return (size_t)par * G.nslot + local; }

// The f32 constants and new time of one step, as the forward used them.
struct StepRec {
    float dt32, dtdx, dtdy, rain32, cap32, pad;
    double t_new, dt;
};

// Per-candidate control block.  Two copies per candidate, selected by launch parity:
// launch k reads copy (k+1)&1, written by launch k-1, and writes copy k&1, so no
// launch reads what it writes and the launch boundary orders the two.
struct CandCtl {
    float dt32, dtdx, dtdy, rain32, cap32;   // constants of the prepared step
    int32_t hits_sync, parity, grid;
    double t, t_new, dt, rain, dt_cfl, s_cur;
    uint64_t step, cell_updates, rows_skipped;
    uint32_t sync_index, flags, skipped_now, n_ledger;
    double v0, acc_rain, inflow, outflow, infil, fix, dt_min, dt_max, dt_sum, lts;
    int32_t abort_kind, abort_cell, abort_sync, local;
    double abort_value;
    uint32_t in_flight;   // a step was launched whose results are not yet accounted
    uint32_t locked_off;  // the locked dt schedule was abandoned (adaptive from dt_fallback_t on)
    uint32_t restore;     // the prepared step reads its old state from the sync snapshot
    uint32_t pad;
    double dt_fallback_t; // RunSummary::dt_fallback_t, -1 when the schedule held
};

// Cross-workgroup results of one step of one candidate (atomics).  Four copies,
// selected by launch index & 3: written in launch k, read in k+1, cleared in k+2.
struct StepAcc {
    int32_t smax_bits;   // max CFL rate as ordered int bits (rates are >= 0)
    uint32_t bad;
    unsigned long long smax64;   // local inertial: the f64 CFL rate max, as ordered bits
    unsigned long long dh64, dq64;   // replay: band deviation maxima (f64 bits, >= 0)
};

struct KArgs {
    const DevGrid* grids;
    const int4* tasks;         // (grid, local cand, band, batch cand)
    CandCtl* snap;             // [2][ncand]
    StepAcc* acc;              // [4][ncand]
    uint32_t* status;          // [ncand] flags of the latest control block, for the host
    int32_t ntasks;
    int32_t ncand;
    int32_t ctl_task;          // 1: each candidate has a control task (band -1); 0: its first band is it
};

// Dynamic LDS of the step kernel, in floats (offsets 16-byte aligned).
// Positivity-fix values are zero almost everywhere: a wave that has one in a row writes
// its lanes' values to the grid's fix buffer and sets its bit in the row's mask, so only
// infiltration is staged per cell.  Grids with Stage / Inflow ghosts add the band's
// special ghost values: (hv, qv) pairs for the west / east ghosts of rows a0-1 .. b0 and
// the north / south ghost rows.
struct LdsLayout { size_t infil, fixmask, bw, be, wet, red, acc, moff, gsp, total; };
__host__ __device__ inline size_t gsp_pairs(int R, int nx) { return (size_t)2 * (R + 2) + (size_t)2 * nx; }
__host__ __device__ inline LdsLayout lds_layout(int R, int nx, bool special) {
    auto al = [](size_t v) { return (v + 3) & ~(size_t)3; };
    LdsLayout L;
    L.infil = 0;
    L.fixmask = al(L.infil + (size_t)R * nx);
    L.bw = al(L.fixmask + (size_t)R);
    L.be = al(L.bw + R);
    L.wet = al(L.be + R);
    L.red = al(L.wet + R);
    L.acc = al(L.red + 2 * kMaxWaves);
    const size_t acc_n = (size_t)R * 64;
    L.moff = al(L.acc + acc_n);                       // a monitored cell's plane offset per thread
    L.gsp = al(L.moff + (size_t)kMaxWaves * 64);
    L.total = al(L.gsp + (special ? 4 * gsp_pairs(R, nx) : 0));
    // the control epilogue stages 2*nx floats and rows of 8 words in the same space
    if (L.total < (size_t)2 * nx) L.total = al((size_t)2 * nx);
    if (L.total < 64) L.total = 64;
    return L;
}

// ------------------------------------------------------------------ device helpers
// Lanes without a source (lane 0 / lane 63) read 0 (bound_ctrl), as the halo lanes expect.
__device__ __forceinline__ float from_left(float x) {   // lane i <- lane i-1 (DPP wave_shr:1)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_right(float x) {  // lane i <- lane i+1 (DPP wave_shl:1)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}

template <bool B> struct BoolC { static constexpr bool value = B; };

// Wave-uniform values into scalar registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ double uni(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ int wave_max_i32(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Pointers into device global memory (address space 1), so hot accesses compile to
// global_load/global_store rather than flat instructions.
#define ITR_GLOBAL __attribute__((address_space(1)))
template <class T> __device__ __forceinline__ ITR_GLOBAL T* gp(T* p) { return (ITR_GLOBAL T*)p; }
template <class T> __device__ __forceinline__ const ITR_GLOBAL T* gp(const T* p) { return (const ITR_GLOBAL T*)p; }

// Raw buffer access (V# with a byte bound): a scalar base + scalar row offset + 32-bit
// lane offset, no 64-bit address arithmetic per access, and hardware bounds checking
// (reads past `bytes` return 0, writes past it are dropped).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, size_t bytes) {
    const uint32_t n = bytes > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)n, 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ int bload_i(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// Cell (row, col) of the old state with ghost cells built on the fly from the
// boundary codes, as fill_ghosts does once per step (crates/itr-hydro/src/boundary.rs:12-87):
// wall mirrors the normal discharge, transmissive copies, corners are zero.  code_n /
// code_s: the lane's north / south codes; the west / east code is looked up per row.
// Split in two so the loads can be issued a row ahead and the (value-dependent)
// ghost transform applied only when the row is used.
struct Cell { float h, qx, qy; int code_we; };
__device__ __forceinline__ Cell fetch_cell(const ITR_GLOBAL float* __restrict__ H, const DevGrid& G,
                                           const ITR_GLOBAL int32_t* __restrict__ ghost, int row, int col) {
    const int nx = G.nx, ny = G.ny;
    const int rr = min(max(row, 0), ny - 1);
    const int cc = min(max(col, 0), nx - 1);
    const size_t o = (size_t)rr * G.P + cc;
    Cell v;
    v.h = H[o];
    v.qx = H[G.plane + o];
    v.qy = H[2 * G.plane + o];
    v.code_we = ghost[2 * nx + (col < 0 ? 0 : ny) + rr];
    return v;
}
__device__ __forceinline__ Cell ghosted(Cell v, int nx, int ny, int row, int col, int code_n, int code_s) {
    const bool rin = row >= 0 && row < ny;
    const bool cin = col >= 0 && col < nx;
    const bool wall_x = !cin && v.code_we == 0;
    const bool wall_y = cin && !rin && (row < 0 ? code_n : code_s) == 0;
    v.qx = wall_x ? -v.qx : v.qx;
    v.qy = wall_y ? -v.qy : v.qy;
    const bool corner = !rin && !cin;
    v.h = corner ? 0.0f : v.h;
    v.qx = corner ? 0.0f : v.qx;
    v.qy = corner ? 0.0f : v.qy;
    return v;
}
__device__ __forceinline__ Cell load_cell(const ITR_GLOBAL float* __restrict__ H, const DevGrid& G,
                                          const ITR_GLOBAL int32_t* __restrict__ ghost, int row, int col,
                                          int code_n, int code_s) {
    return ghosted(fetch_cell(H, G, ghost, row, col), G.nx, G.ny, row, col, code_n, code_s);
}

// Hyetograph::rate / integral (crates/itr-core/src/scenario.rs:540-585), f64.
__device__ __forceinline__ double hy_rate(const double* rt, const double* rr, int n, double x) {
    if (n == 0) return 0.0;
    if (x <= rt[0]) return rr[0];
    for (int k = 0; k + 1 < n; k++) {
        if (x <= rt[k + 1]) {
            const double t0 = rt[k], r0 = rr[k], t1 = rt[k + 1], r1 = rr[k + 1];
            if (t1 == t0) return r1;
            return r0 + (r1 - r0) * (x - t0) / (t1 - t0);
        }
    }
    return rr[n - 1];
}
__device__ __forceinline__ double hy_integral(const double* rt, const double* rr, int n, double a, double b) {
    if (b <= a) return 0.0;
    if (n == 0) return 0.0;
    double total = 0.0, t = a, r = hy_rate(rt, rr, n, a);
    for (int k = 0; k < n; k++) {
        const double tk = rt[k];
        if (tk > a && tk < b) {
            const double rk = hy_rate(rt, rr, n, tk);
            total += 0.5 * (r + rk) * (tk - t);
            t = tk;
            r = rk;
        }
    }
    const double rb = hy_rate(rt, rr, n, b);
    total += 0.5 * (r + rb) * (b - t);
    return total;
}

// Real::rcbrt for f64 (crates/itr-core/src/real.rs:79-94): exponent-bias seed minus
// bits / 3, then 4 steps r <- r*(4/3) - (x/3)*r^4, no fused operations.
__device__ __forceinline__ double rcbrt64(double x) {
    const uint64_t K = 0x5540000000000000ull - (0x8782Bull << 29);
    const uint64_t bits = (uint64_t)__double_as_longlong(x);
    double r = __longlong_as_double((long long)(K - bits / 3u));
    const double third = 1.0 / 3.0;
    const double x3 = x * third;
    const double c43 = 4.0 * third;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double r2 = r * r;
        r = r * c43 - x3 * (r2 * r2);
    }
    return r;
}

// Stage / Inflow ghost of boundary cell k (index into the grid's ghost codes) of
// candidate `local` at step start t (boundary.rs:43-47, 66-70).  Stage: hv is the ghost
// depth f32(max(stage - z, 0)) (prepared on the host), the discharge copies the
// interior.  Inflow: q = rate(t) / length * inflow_scale per unit width, hv is the
// critical depth (q^2/g)^(1/3) = x * rcbrt(x) * rcbrt(x) rounded to f32 (the ghost depth
// is max(h_interior, critical)), qv = f32(inward * q) is the normal discharge and the
// tangential one is 0.  Both kinds make the ghost wet exactly when hv > 0 (or the
// interior is wet).
struct GhostSp { float hv, qv, qt; };
// Ring position (replay.rs:27-41 order) of ghost cell k of a w x h replay grid:
// north / south columns after the ring row's corner, west / east alternating per row.
__host__ __device__ inline int ring_of_ghost(int k, int w, int h) {
    if (k < w) return 1 + k;
    if (k < 2 * w) return (w + 2) + 1 + (k - w);
    if (k < 2 * w + h) return 2 * (w + 2) + 2 * (k - 2 * w);
    return 2 * (w + 2) + 2 * (k - 2 * w - h) + 1;
}
__device__ __forceinline__ GhostSp special_ghost(const DevGrid& G, double inflow_scale, int code, int k, int local,
                                                 double t, bool inward_pos, uint64_t step) {
    GhostSp s;
    s.hv = 0.0f;
    s.qv = 0.0f;
    s.qt = 0.0f;
    if (code == 4) {   // Tape::fill (replay.rs:136-159): the recorded ring cell, as it was
        if (step < (uint64_t)G.tape_steps) {
            const float* v = G.tape_ring + ((size_t)step * G.n_ring + ring_of_ghost(k, G.nx, G.ny)) * 3;
            s.hv = v[0];
            s.qv = v[1];
            s.qt = v[2];
        }
    } else if (code == 2) {
        s.hv = G.gstage[(size_t)local * (2 * G.nx + 2 * G.ny) + k];
    } else if (code == 3) {
        const int hi = G.ghydro[k];
        const int a = G.hstart[hi], n = G.hstart[hi + 1] - a;
        const double q = hy_rate(G.hyd_t + a, G.hyd_q + a, n, t) * G.ginv[k] * inflow_scale;
        const double x = q * q / G.g;
        const double crit = x > 0.0 ? x * rcbrt64(x) * rcbrt64(x) : 0.0;
        s.hv = (float)crit;
        s.qv = (float)((inward_pos ? 1.0 : -1.0) * q);
    }
    return s;
}
// f32(max(f64(h), crit)) with f64::max's NaN rule (a NaN operand yields the other one),
// from hc = f32(crit): rounding is monotone and h is an f32, so the f64 comparison and
// the f32 one agree.  Interior depths are never -0.
__device__ __forceinline__ float inflow_depth(float h, float hc) { return (h != h || hc > h) ? hc : h; }

// Head of one loop iteration of Solver::run_impl (solver.rs:449-489 and 269-285): the
// step's dt (CFL, dt_max, locked schedule, sync clipping), exact rain depth, new time
// and the f32 constants.  Returns false if the locked-dt check failed.
__device__ __forceinline__ bool prepare_step(const DevParams& p, const DevGrid& G, CandCtl* C, const double* rt,
                                             const double* rr) {
    const double t = C->t;
    const double t_sync = fmin((double)(C->sync_index + 1) * p.sync, p.duration);
    double dt = fmin(C->dt_cfl, p.dt_max);
    if (p.locked && !C->locked_off) {
        if (C->step < (uint64_t)p.n_locked && p.locked[C->step] * C->s_cur <= p.cfl_hard) dt = p.locked[C->step];
        else return false;
    }
    const bool hits = t + dt >= t_sync - 1e-9 * p.sync;
    if (hits) dt = t_sync - t;
    const double rain = hy_integral(rt, rr, p.n_rain, t, t + dt);
    C->t_new = hits ? t_sync : t + dt;
    C->dt = dt;
    C->rain = rain;
    C->hits_sync = hits ? 1 : 0;
    C->dt32 = (float)dt;
    C->dtdx = (float)(dt / G.dx);
    C->dtdy = (float)(dt / G.dy);
    C->rain32 = (float)rain;
    C->cap32 = (float)(p.infil_cap * dt);
    return true;
}

// Dry-row accounting of the coming step (kernel.rs:605-612): rows j (padded 1..ny)
// with rows j-1, j, j+1 exactly dry and no rain.  Ghost rows mirror their neighbour for
// wall and transmissive boundaries; a wet Stage / Inflow ghost makes its row (west /
// east) or the ghost row (north / south) wet (boundary.rs:50-52, 73-75).  Computed by one
// wave at the step's start time C->t.
__device__ __forceinline__ bool we_ghost_wet(const DevParams& p, const DevGrid& G, int local, double t, int r, uint64_t step) {
    const int nx = G.nx, ny = G.ny;
    bool w = false;
    for (int side = 0; side < 2; side++) {
        const int k = 2 * nx + side * ny + r;
        const int code = G.ghost[k];
        if (code >= 2) w = w || special_ghost(G, p.inflow_scale, code, k, local, t, side == 0, step).hv > 0.0f;
    }
    return w;
}
// A replay grid's ghost-row corners (ring cells 0, w+1 / w+2, 2w+3) also make the ghost
// row wet (replay.rs:148-153).
__device__ __forceinline__ bool tape_corner_wet(const DevGrid& G, bool north, uint64_t step) {
    if (G.tape_mode != 2 || step >= (uint64_t)G.tape_steps) return false;
    const float* v = G.tape_ring + (size_t)step * G.n_ring * 3;
    const int a = north ? 0 : G.nx + 2, b = north ? G.nx + 1 : 2 * G.nx + 3;
    return v[3 * a] > 0.0f || v[3 * b] > 0.0f;
}
template <bool SPECIAL>
__device__ __forceinline__ uint32_t count_skippable(const DevParams& p, const DevGrid& G, const CandCtl* C, const RowAcc* ra,
                                    int lane) {
    if (!p.skip_dry || C->rain32 > 0.0f) return 0;
    const int ny = G.ny;
    uint32_t n = 0;
    if (!SPECIAL || !G.special) {
        for (int r = lane; r < ny; r += kWave) {
            const bool d0 = r > 0 ? !ra[r - 1].wet : !ra[0].wet;
            const bool d1 = !ra[r].wet;
            const bool d2 = r + 1 < ny ? !ra[r + 1].wet : !ra[ny - 1].wet;
            n += (d0 && d1 && d2) ? 1u : 0u;
        }
    } else {
        // Each lane walks a run of consecutive rows, so every row's ghosts are evaluated
        // about once.  The north / south ghost rows count as dry unless a special ghost
        // cell of theirs is wet (a wall / transmissive ghost is wet only with its row).
        const int nx = G.nx, local = C->local;
        const double t = C->t;
        const uint64_t step = C->step;
        bool top = tape_corner_wet(G, true, step), bot = tape_corner_wet(G, false, step);
        for (int c = lane; c < nx; c += kWave) {
            const int cn = G.ghost[c], cs = G.ghost[nx + c];
            if (cn >= 2) top = top || special_ghost(G, p.inflow_scale, cn, c, local, t, true, step).hv > 0.0f;
            if (cs >= 2) bot = bot || special_ghost(G, p.inflow_scale, cs, nx + c, local, t, false, step).hv > 0.0f;
        }
        top = __any(top);
        bot = __any(bot);
        const int per = (ny + kWave - 1) / kWave;
        const int r0 = lane * per, r1 = min(r0 + per, ny);
        if (r0 < r1) {
            auto dry = [&](int r) -> bool {
                if (r < 0) return !top;
                if (r >= ny) return !bot;
                return !ra[r].wet && !we_ghost_wet(p, G, local, t, r, step);
            };
            bool d0 = dry(r0 - 1), d1 = dry(r0);
            for (int r = r0; r < r1; r++) {
                const bool d2 = dry(r + 1);
                n += (d0 && d1 && d2) ? 1u : 0u;
                d0 = d1;
                d1 = d2;
            }
        }
    }
    for (int o = 32; o >= 1; o >>= 1) n += __shfl_xor(n, o, 64);
    return n;
}


// Objective::guard_breach (objective.rs:132-144) over the guard cells, one wave: the
// first guard position (in guard order) of the largest positive exceedance
// w = max_k - baseline_k - tol.  Returns the position, or -1 when nothing exceeds.
__device__ __forceinline__ int guard_breach(const DevGrid& G, const float* mm, int lane, double* worst_out) {
    bool have = false;
    double worst = 0.0;
    int wq = 0x7fffffff;
    for (int q = lane; q < G.n_guard; q += kWave) {
        const int k = G.guard_pos[q];
        const double w = (double)mm[k] - G.baseline_max[k] - G.guard_tol;
        if (w > 0.0 && (!have || w > worst)) { have = true; worst = w; wq = q; }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const int oh = __shfl_xor(have ? 1 : 0, o, 64);
        const double ow = __shfl_xor(worst, o, 64);
        const int oq = __shfl_xor(wq, o, 64);
        if (oh && (!have || ow > worst || (ow == worst && oq < wq))) { have = true; worst = ow; wq = oq; }
    }
    *worst_out = worst;
    return have ? wq : -1;
}

// The completed step's CFL rate maximum: f32 for the HLL step (kernel.rs), f64 for the
// local-inertial step (inertial.rs:186-196).
__device__ __forceinline__ double step_smax(const DevGrid& G, const StepAcc& A) {
    return G.li ? __longlong_as_double((long long)A.smax64) : (double)__int_as_float(max(A.smax_bits, 0));
}

// Completion of the candidate's step launched by the previous launch (solver.rs:529-545
// and the sync point, 554-603, except the ledger sums), then the termination test and
// the head of the next step (prepare_step).  Every wave of every workgroup of the
// candidate evaluates it on the same inputs, so all reach the same decision; only the
// candidate's control task writes the result.  Returns true if this launch runs a step.
__device__ __forceinline__ bool complete_step(const DevParams& p, const DevGrid& G, CandCtl& S, const StepAcc& A,
                                              int lane) {
    if (S.flags & F_DONE) return false;
    if (S.in_flight) {
        const bool was_sync = S.hits_sync != 0;
        S.in_flight = 0;
        S.restore = 0;
        S.step += 1;
        S.t = S.t_new;
        S.dt_min = fmin(S.dt_min, S.dt);
        S.dt_max = fmax(S.dt_max, S.dt);
        S.dt_sum += S.dt;
        S.parity ^= 1;
        uint32_t flags = S.flags;
        bool go = true;
        if (A.bad) {
            S.abort_kind = 3;
            flags |= F_DONE;
            go = false;
        } else {
            const double smax = step_smax(G, A);
            S.dt_cfl = smax > 0.0 ? p.cfl / smax : p.dt_max;
            S.s_cur = smax;
            if (S.hits_sync) {
                S.sync_index += 1;
                if (G.tape_mode == 2) {   // Tape::band_diff within tolerance, else replay-invalid (solver.rs:578-590)
                    const double dh = __longlong_as_double((long long)A.dh64), dq = __longlong_as_double((long long)A.dq64);
                    if ((int64_t)S.sync_index > G.tape_syncs || !(dh <= G.tol_h && dq <= G.tol_q)) {
                        S.abort_kind = 5;
                        S.abort_sync = (int)S.sync_index;
                        flags |= F_DONE;
                        go = false;
                    }
                }
                if (go && G.baseline_max && G.n_guard > 0) {
                    double worst = 0.0;
                    const int q = guard_breach(G, G.mon_max + slot(G, S.parity, S.local) * G.n_mon, lane, &worst);
                    if (q >= 0) {
                        S.abort_kind = 1;
                        S.abort_cell = (int)G.mon_cells[G.guard_pos[q]];
                        S.abort_value = worst;
                        S.abort_sync = (int)S.sync_index;
                        flags |= F_DONE;
                        go = false;
                    }
                }
                if (go && p.has_cb) flags |= F_PAUSED;
            }
            if (go && !(S.t < p.duration && S.sync_index < p.n_sync)) { flags |= F_DONE; go = false; }
            if (go && p.max_steps > 0 && S.step >= (uint64_t)p.max_steps) { flags |= F_DONE; go = false; }
            if (go && G.tape_mode == 2 && (S.step >= (uint64_t)G.tape_steps || !prepare_step(p, G, &S, p.rain_t, p.rain_r))) {
                S.abort_kind = 5;   // replay-invalid: tape exhausted or locked-dt check failed (solver.rs:456-460, 484-489)
                S.abort_sync = (int)S.sync_index;
                flags |= F_DONE;
                go = false;
            }
            if (go && G.tape_mode != 2 && !prepare_step(p, G, &S, p.rain_t, p.rain_r)) {
                // The locked step fails the hard-CFL check: restart adaptively from the
                // last sync snapshot (solver.rs:456-480).  After a sync step the snapshot
                // is the state just completed, so only the schedule is dropped.
                if (!was_sync) {
                    CandCtl K = G.snap_ctl[S.local];
                    K.parity = S.parity;
                    K.flags = flags;
                    K.restore = 1;
                    S = K;
                }
                S.locked_off = 1;
                S.dt_fallback_t = S.t;
                if (p.max_steps > 0 && S.step >= (uint64_t)p.max_steps) { flags |= F_DONE; go = false; }
                else (void)prepare_step(p, G, &S, p.rain_t, p.rain_r);   // adaptive: cannot fail
            }
        }
        S.flags = flags;
    }
    return !(S.flags & (F_DONE | F_PAUSED));
}

// One FixedSum lane over a row: value(c) for c = c0, c0 + 16, ... < n, added in that
// order (real.rs:152-160), with the values of four columns fetched before they are added.
template <class F>
__device__ __forceinline__ float lane_chain(int c0, int n, F value) {
    float s = 0.0f;
    int c = c0;
    for (; c + 3 * 16 < n; c += 4 * 16) {
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = value(c + 16 * k);
#pragma unroll
        for (int k = 0; k < 4; k++) s += v[k];
    }
    for (; c < n; c += 16) s += value(c);
    return s;
}

// Stage / Inflow ghosts of a band at the step start into LDS, as (hv, qv) pairs: west /
// east of rows a0-1 .. b0 (pairs 0 .. 2R+3), then the north ghost row when a0 <= north_upto
// and the south one when the band ends at the last row.  Returns whether any ghost of
// the band is wet (all threads).
__device__ __forceinline__ bool stage_special_ghosts(const DevGrid& G, int local, int a0, int b0, double t, float4* l_gsp,
                                                     int north_upto, uint64_t step) {
    const int nx = G.nx, ny = G.ny, R = G.R;
    const int tid = threadIdx.x;
    const int nthr = G.waves * kWave;
    const int nwe = 2 * (R + 2);
    bool fw = false;
    for (int i = tid; i < nwe + 2 * nx; i += nthr) {
        int k = -1;
        bool inward = true;
        if (i < nwe) {
            const int side = i / (R + 2), row = a0 - 1 + i % (R + 2);
            if (row >= 0 && row < ny) { k = 2 * nx + side * ny + row; inward = side == 0; }
        } else {
            const int side = (i - nwe) / nx, c = (i - nwe) % nx;
            if (side == 0 ? a0 <= north_upto : b0 == ny) { k = side * nx + c; inward = side == 0; }
        }
        GhostSp gs;
        gs.hv = 0.0f;
        gs.qv = 0.0f;
        gs.qt = 0.0f;
        if (k >= 0) gs = special_ghost(G, G.prm.inflow_scale, G.ghost[k], k, local, t, inward, step);
        l_gsp[i] = make_float4(gs.hv, gs.qv, gs.qt, 0.0f);
        fw = fw || gs.hv > 0.0f;
    }
    if (tid == 0 && ((a0 == 0 && tape_corner_wet(G, true, step)) || (b0 == ny && tape_corner_wet(G, false, step)))) fw = true;
    return __syncthreads_or(fw ? 1 : 0) != 0;
}

// One step of one band of one candidate (kernel.rs:590-727 semantics; the sweep order
// differs but every face and cell value is the same IEEE expression of the old state).
// Reads state copy `par`, writes copy par^1 and the row partials / monitors of parity
// par^1; the band's CFL max and health flag go to the candidate's StepAcc.
// Position of cell (i, j) (1-based in the ROI, on its inner band) in the band's
// row-major order (replay.rs:43-54).
__device__ __forceinline__ int band_index(int i, int j, int w, int h) {
    if (j == 1) return i - 1;
    const int mid = w == 1 ? 1 : 2;   // band cells of a middle row
    if (j < h) return w + mid * (j - 2) + (i == 1 ? 0 : 1);
    return w + mid * (h - 2) + (i - 1);
}

// A band's monitor range (positions q0 .. q1-1 of the sorted monitor list) and the plane
// offset of the calling thread's first monitored cell: two dependent loads, which the
// kernel issues before its control prologue so that their latency passes under it.
struct MonRange { int q0, q1; uint32_t off0; };
__device__ __forceinline__ MonRange mon_range(const DevGrid& G, int band, int tid) {
    const int a0 = band * G.R, b0 = min(a0 + G.R, G.ny);
    MonRange m;
    m.q0 = G.mon_row[a0];
    m.q1 = G.mon_row[b0];
    const int qm = m.q0 + tid;
    m.off0 = qm < m.q1 ? G.mon_off[qm] : 0u;
    return m;
}

// The band march's lane geometry: the lane's column and its group's first row relative to
// the band (the tail wave, when it updates few columns, splits into `tail_groups` lane
// groups of 64 / tail_groups lanes; group g marches rows a0 + g * t_rows ..; every other
// wave is one group of 64 lanes, whose first and last lanes are halos).
struct LaneGeo { int tail_s, tgw, tg, gl, t_rows, lrow, col; bool tail; };
template <bool SPECIAL>
__device__ __forceinline__ LaneGeo lane_geo(const DevGrid& G, int band, int tid) {
    const int a0 = band * G.R, nrows = min(a0 + G.R, G.ny) - a0;
    const int lane = tid & 63, wave = tid >> 6;
    LaneGeo q;
    q.tail_s = SPECIAL ? 1 : G.tail_groups;
    q.tail = q.tail_s > 1 && uni(wave) == G.waves - 1;
    q.tgw = q.tail ? 64 / q.tail_s : 64;
    q.tg = q.tail ? lane / q.tgw : 0;
    q.gl = lane - q.tg * q.tgw;
    q.t_rows = q.tail ? (nrows + q.tail_s - 1) / q.tail_s : nrows;
    q.lrow = q.tg * q.t_rows;
    q.col = wave * kColsPerWave - 1 + q.gl;
    return q;
}

// The old state of the band's first two rows (a0-1 and a0, plus the lane group's row
// offset) at the lane's column, read at the kernel's start from the state copy the step
// will most likely read (the copy the previous launch wrote), so that their latency
// passes under the control prologue.  The march uses them when the step does read that
// copy (`par` equal, no restart from a snapshot) and loads the rows itself otherwise.
// (Reading more of the march's first inputs this way cost registers and time: measured.)
struct FirstRows { float h0, qx0, qy0, h1, qx1, qy1; int cwe0, cwe1, par; };
template <bool SPECIAL>
__device__ __forceinline__ FirstRows first_rows(const DevGrid& G, int local, int band, int par, int tid) {
    const LaneGeo q = lane_geo<SPECIAL>(G, band, tid);
    const int nx = G.nx, ny = G.ny;
    const int a0 = band * G.R;
    const float* S = G.state + (size_t)local * 6 * G.plane + (size_t)par * 3 * G.plane;
    const int cc = min(max(q.col, 0), nx - 1);
    const int r0 = min(max(a0 - 1 + q.lrow, 0), ny - 1), r1 = min(max(a0 + q.lrow, 0), ny - 1);
    const size_t o0 = (size_t)r0 * G.P + cc, o1 = (size_t)r1 * G.P + cc;
    FirstRows f;
    f.h0 = S[o0];
    f.qx0 = S[G.plane + o0];
    f.qy0 = S[2 * G.plane + o0];
    f.h1 = S[o1];
    f.qx1 = S[G.plane + o1];
    f.qy1 = S[2 * G.plane + o1];
    f.cwe0 = G.ghost[2 * nx + (q.col < 0 ? 0 : ny) + r0];
    f.cwe1 = G.ghost[2 * nx + (q.col < 0 ? 0 : ny) + r1];
    f.par = par;
    return f;
}

template <bool CREST, bool SPECIAL>
__device__ __forceinline__ void band_step(const KArgs& a, const DevGrid& G, int local, int band, const StepK& K,
                                          double t, double t_new, bool sync, int par, bool from_snap, uint64_t step,
                                          uint32_t sync_idx, StepAcc* acc_cur, float* lds, const MonRange& mr,
                                          const FirstRows& pre) {
    const int nx = G.nx, ny = G.ny, P = G.P, R = G.R;
    const int a0 = band * R;
    const int b0 = min(a0 + R, ny);
    const int nrows = b0 - a0;
    const int tid = threadIdx.x;
    const int nthr = G.waves * kWave;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = G.waves;
    // The tail wave's lane groups (lane_geo): group tg of tgw lanes marches rows
    // a0 + lrow .. of the band; gl is the lane within its group.
    const LaneGeo lg = lane_geo<SPECIAL>(G, band, tid);
    const int tail_s = lg.tail_s, tgw = lg.tgw, tg = lg.tg, gl = lg.gl, t_rows = lg.t_rows, lrow = lg.lrow, col = lg.col;
    const bool tail = lg.tail;
    const bool upd = gl >= 1 && gl <= tgw - 2 && tg < tail_s && col < nx;
    const int npar = par ^ 1;

    const size_t cstride = 6 * G.plane;
    float* base = G.state + (size_t)local * cstride;
    ITR_GLOBAL float* __restrict__ Ho = gp(base + (size_t)npar * 3 * G.plane);
    const ITR_GLOBAL int32_t* __restrict__ GHOST = gp(G.ghost);
    // The old state: copy `par`, or the sync snapshot after a locked-dt restart.
    const float* old_state = from_snap ? G.snap_state + (size_t)local * 3 * G.plane : base + (size_t)par * 3 * G.plane;
    const ITR_GLOBAL float* __restrict__ mon_in =
        gp(from_snap ? G.snap_mon + (size_t)local * G.n_mon : G.mon_max + slot(G, par, local) * G.n_mon);
    const ITR_GLOBAL double* __restrict__ tmax_in =
        gp(from_snap ? G.snap_tmax + (size_t)local * G.n_mon : G.mon_tmax + slot(G, par, local) * G.n_mon);
    ITR_GLOBAL float* __restrict__ mon_out = gp(G.mon_max + slot(G, npar, local) * G.n_mon);
    ITR_GLOBAL double* __restrict__ tmax_out = gp(G.mon_tmax + slot(G, npar, local) * G.n_mon);
    // The band's monitors: positions mon_q0 .. mon_q1-1 of the sorted monitor list; this
    // thread's first one (qm) has its plane offset staged in LDS for the epilogue.
    const int mon_q0 = mr.q0, mon_q1 = mr.q1;
    const ITR_GLOBAL uint32_t* __restrict__ MONOFF = gp(G.mon_off);
    const int qm = mon_q0 + tid;
    const bool has_m = qm < mon_q1;
    const uint32_t m_off0 = mr.off0;

    ITR_GLOBAL float* __restrict__ sbuf = gp(G.sbuf + (size_t)local * G.plane);
    ITR_GLOBAL float* __restrict__ fixbuf = gp(G.fixbuf + (size_t)local * G.plane);
    const RowAcc* ra_in = from_snap ? G.snap_rows + (size_t)local * ny : G.rowacc + slot(G, par, local) * ny;
    RowAcc* ra_out = G.rowacc + slot(G, npar, local) * ny;
    float* yb_out = G.ybnd + slot(G, npar, local) * 2 * nx;
    // Lane columns for addressing (unsigned, so loads use a scalar row base + 32-bit
    // lane offset): ccol clamps to the interior, xcol to the x-face range 0..nx.
    const uint32_t ccol = (uint32_t)min(max(col, 0), nx - 1);
    const uint32_t xcol = (uint32_t)min(max(col, 0), nx);
    const uint32_t cb = ccol * 4u, xb = xcol * 4u;   // lane byte offsets within a row
    const uint32_t rowb = (uint32_t)P * 4u;           // row pitch in bytes
    const uint32_t planeb = (uint32_t)(G.plane * 4u);
    const __amdgpu_buffer_rsrc_t rs_ho = make_rsrc(base + (size_t)npar * 3 * G.plane, 3 * G.plane * 4);
    const bool in_y = col >= 0 && col < nx;   // columns that own a y-face
    const bool owner = in_y && gl >= 1 && gl <= tgw - 2 && tg < tail_s;
    // Everything read from the grid descriptor is read here, before the first store:
    // later loads of it could not be proven unclobbered and would become per-lane
    // loads (and per-lane buffer descriptors).
    const __amdgpu_buffer_rsrc_t rs_h = make_rsrc(old_state, 3 * G.plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzx = make_rsrc(G.dzx + (size_t)local * G.plane, G.plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzy = make_rsrc(G.dzy + (size_t)local * G.plane_y, G.plane_y * 4);
    const __amdgpu_buffer_rsrc_t rs_dzx2 = make_rsrc(CREST ? G.dzx2 + (size_t)local * G.plane : G.dzx, G.plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzy2 = make_rsrc(CREST ? G.dzy2 + (size_t)local * G.plane_y : G.dzy, G.plane_y * 4);
    const __amdgpu_buffer_rsrc_t rs_wet = make_rsrc(G.wet, G.plane * 4);
    const __amdgpu_buffer_rsrc_t rs_gn2 = make_rsrc(G.gn2, G.plane * 4);
    const int code_n = GHOST[ccol];
    const int code_s = GHOST[nx + ccol];
    const bool edge_wave = wave == 0 || wave == nwaves - 1;   // only they hold ghost columns
    const bool skip_dry = G.prm.skip_dry != 0;
    const bool special = SPECIAL && G.special != 0;
    const LdsLayout LL = lds_layout(R, nx, G.special != 0);
    float4* l_gsp = (float4*)(lds + LL.gsp);

    bool forced_wet = false;
    if (special) forced_wet = stage_special_ghosts(G, local, a0, b0, t, l_gsp, 0, step);

    // Dry band (kernel.rs:605-612 rule for every row of the band: rows a0-1 .. b0 exactly
    // dry, ghost rows mirroring their neighbour or holding no wet special ghost, and no
    // rain).  Every face of such a row is dry-dry (exact zero flux) and every cell stays
    // (0, 0, 0), so the band's new state, row partials and boundary masses are zeros and
    // its monitors carry over.
    // (not while recording a tape, nor at a replay's sync steps: those read the band's cells)
    if (skip_dry && !(K.rain > 0.0f) && !forced_wet && !(SPECIAL && (G.tape_mode == 1 || (G.tape_mode == 2 && sync)))) {
        bool dry = true;
        for (int r = max(a0 - 1, 0); r <= min(b0, ny - 1); r++) dry = dry && ra_in[r].wet == 0u;
        if (dry) {
            for (int j = a0; j < b0; j++) {
                if (upd) {
                    const uint32_t so = (uint32_t)j * rowb;
                    bstore(0.0f, rs_ho, cb, so);
                    bstore(0.0f, rs_ho, cb, so + planeb);
                    bstore(0.0f, rs_ho, cb, so + 2 * planeb);
                }
            }
            for (int q = qm; q < mon_q1; q += nthr) {
                mon_out[q] = mon_in[q];
                tmax_out[q] = tmax_in[q];
            }
            if (a0 == 0 && owner) yb_out[col] = 0.0f;
            if (b0 == ny && owner) yb_out[nx + col] = 0.0f;
            for (int r = tid; r < nrows; r += nthr) {
                RowAcc q;
                q.infil = q.fix = q.bw = q.be = q.vol = q.sum_s = 0.0f;
                q.wet = 0u;
                q.pad = 0u;
                ra_out[a0 + r] = q;
            }
            return;
        }
    }


    // LDS: per-row staging of infil (R x nx) + row scalars (lds_layout).
    float* l_infil = lds + LL.infil;
    uint32_t* l_fixmask = (uint32_t*)(lds + LL.fixmask);
    float* l_bw = lds + LL.bw;
    float* l_be = lds + LL.be;
    uint32_t* l_wet = (uint32_t*)(lds + LL.wet);
    int* l_red = (int*)(lds + LL.red);            // [2 * kMaxWaves]
    float* l_acc = lds + LL.acc;                  // [R][16][4]
    // (l_wet, l_fixmask, l_bw and l_be were zeroed before the control prologue's barrier;
    // the special ghosts staged above are read by other threads)
    if (special) __syncthreads();

    // Old-state cell of this lane's column in `row`: raw loads (scalar row base + lane
    // offset) and, on ghost rows or in the edge waves only, the boundary transform.
    auto fetch = [&](int row) -> Cell {
        const int rr = min(max(row, 0), ny - 1);
        const uint32_t so = (uint32_t)rr * rowb;
        Cell v;
        v.h = bload(rs_h, cb, so);
        v.qx = bload(rs_h, cb, so + planeb);
        v.qy = bload(rs_h, cb, so + 2 * planeb);
        // The row's west / east boundary codes are wave-uniform: scalar loads, selected per
        // lane (only the edge lanes use them), so no vector load waits in the march.
        const int code_w = GHOST[2 * nx + rr], code_e = GHOST[2 * nx + ny + rr];
        v.code_we = col < 0 ? code_w : code_e;
        return v;
    };
    // Stage / Inflow ghosts over the wall / transmissive transform (boundary.rs:43-47, 66-70).
    auto ghost_special = [&](Cell w, const Cell& v, int row) -> Cell {
        const bool rin = row >= 0 && row < ny;
        const bool cin = col >= 0 && col < nx;
        if (rin && !cin) {
            if (v.code_we >= 2) {
                const float4 g = l_gsp[(col < 0 ? 0 : R + 2) + (row - (a0 - 1))];
                if (v.code_we == 2) {
                    w.h = g.x;
                } else if (v.code_we == 4) {
                    w.h = g.x;
                    w.qx = g.y;
                    w.qy = g.z;
                } else {
                    w.h = inflow_depth(v.h, g.x);
                    w.qx = g.y;
                    w.qy = 0.0f;
                }
            }
        } else if (!rin && cin) {
            const int code = row < 0 ? code_n : code_s;
            if (code >= 2) {
                const float4 g = l_gsp[2 * (R + 2) + (row < 0 ? 0 : nx) + col];
                if (code == 2) {
                    w.h = g.x;
                } else if (code == 4) {
                    w.h = g.x;
                    w.qx = g.y;
                    w.qy = g.z;
                } else {
                    w.h = inflow_depth(v.h, g.x);
                    w.qx = 0.0f;
                    w.qy = g.y;
                }
            }
        }
        return w;
    };
    auto fix = [&](const Cell& v, int row) -> Cell {
        if (edge_wave || row < 0 || row >= ny) {
            const Cell w = ghosted(v, nx, ny, row, col, code_n, code_s);
            return special ? ghost_special(w, v, row) : w;
        }
        return v;
    };
    // Face jumps are stored with zeros on boundary faces (row 0 and row ny of dzy, columns
    // 0 and nx of dzx), so no masking is needed; lanes outside a face's role read an
    // in-bounds clamped entry whose value is never used.
    auto dzy_at = [&](__amdgpu_buffer_rsrc_t D, int row) -> float { return bload(D, cb, (uint32_t)min(row, ny) * rowb); };
    auto dzx_at = [&](__amdgpu_buffer_rsrc_t D, int row) -> float { return bload(D, xb, (uint32_t)row * rowb); };
    auto row_at = [&](__amdgpu_buffer_rsrc_t A, int row) -> float { return bload(A, cb, (uint32_t)min(row, ny - 1) * rowb); };
    // The same reads for the tail wave, whose lanes read row `row + lrow` (their group's
    // row): per-lane row offsets, clamped per lane.
    auto fetch_t = [&](int row) -> Cell {
        const int rr = min(max(row + lrow, 0), ny - 1);
        const uint32_t vo = cb + (uint32_t)rr * rowb;
        Cell v;
        v.h = bload(rs_h, vo, 0u);
        v.qx = bload(rs_h, vo, planeb);
        v.qy = bload(rs_h, vo, 2 * planeb);
        v.code_we = GHOST[2 * nx + (col < 0 ? 0 : ny) + rr];
        return v;
    };
    auto dzy_t = [&](__amdgpu_buffer_rsrc_t D, int row) -> float {
        return bload(D, cb + (uint32_t)min(row + lrow, ny) * rowb, 0u);
    };
    auto dzx_t = [&](__amdgpu_buffer_rsrc_t D, int row) -> float {
        return bload(D, xb + (uint32_t)min(row + lrow, ny - 1) * rowb, 0u);
    };
    auto row_t = [&](__amdgpu_buffer_rsrc_t A, int row) -> float {
        return bload(A, cb + (uint32_t)min(row + lrow, ny - 1) * rowb, 0u);
    };

    // Subdomain replay tape (replay.rs:112-176; special kernel only).  Record mode: the
    // ROI's ghost ring from the old state with ghosts applied (Tape::record_ring, after the
    // ghost fill) and, at sync steps, its inner band from the new state; replay mode: the
    // band measured against the tape at sync steps (Tape::band_diff).
    const bool rec = SPECIAL && G.tape_mode == 1 && local == 0 && step < (uint64_t)G.tape_steps;
    const bool rec_band = rec && sync && (int64_t)sync_idx < G.tape_syncs;
    const bool chk = SPECIAL && G.tape_mode == 2 && sync && (int64_t)sync_idx < G.tape_syncs;
    const int rc0 = G.roi_c0, rr0 = G.roi_r0, rw = G.roi_w, rh = G.roi_h;
    const bool ring_owner = upd || (col == -1 && wave == 0 && lane == 0) || (col == nx && lane >= 1);
    auto ring_put = [&](int row, float h, float qx, float qy) {
        int k = -1;
        if (row == rr0 - 1 || row == rr0 + rh) {
            if (col >= rc0 - 1 && col <= rc0 + rw) k = (row == rr0 - 1 ? 0 : rw + 2) + col - (rc0 - 1);
        } else if (row >= rr0 && row < rr0 + rh) {
            if (col == rc0 - 1) k = 2 * (rw + 2) + 2 * (row - rr0);
            else if (col == rc0 + rw) k = 2 * (rw + 2) + 2 * (row - rr0) + 1;
        }
        if (k >= 0) {
            float* d = G.tape_ring + ((size_t)step * G.n_ring + k) * 3;
            d[0] = h;
            d[1] = qx;
            d[2] = qy;
        }
    };
    auto band_at = [&](int row) -> int {   // this lane's band cell in `row`, or -1
        if (!upd || row < rr0 || row >= rr0 + rh || col < rc0 || col >= rc0 + rw) return -1;
        const int i = col - rc0 + 1, jj = row - rr0 + 1;
        return (i == 1 || jj == 1 || i == rw || jj == rh) ? band_index(i, jj, rw, rh) : -1;
    };
    double dh_l = 0.0, dq_l = 0.0;
    uint32_t* l_moff = (uint32_t*)(lds + LL.moff);
    if (has_m) l_moff[tid] = m_off0;   // read back after the march

    int smax_bits = 0;
    // Health flag of the band, as a wave mask (ballots keep it in scalar registers).
    uint64_t bad_mask = 0;
    // Lanes that update a cell, and per-row lane masks of a comparison: v_cmp straight into
    // a lane mask (LLVM fcmp predicates: 2 ogt, 12 ult, 8 uno, 14 une), with every lane active.
    const uint64_t upd_mask = __builtin_amdgcn_ballot_w64(upd);
    auto lanes_ogt0 = [](float x) -> uint64_t { return __builtin_amdgcn_fcmpf(x, 0.0f, 2); };
    auto lanes_olt0 = [](float x) -> uint64_t { return __builtin_amdgcn_fcmpf(x, 0.0f, 4); };
    // The march down the band.  TL: the tail wave's form, whose lanes are at row j + lrow
    // (their group's row) when the march is at row j; everything else is the same code.
    auto march = [&](auto tag) {
        constexpr bool TL = decltype(tag)::value;
        auto FETCH = [&](int row) -> Cell {
            if constexpr (TL) return fetch_t(row);
            else return fetch(row);
        };
        auto DZY = [&](__amdgpu_buffer_rsrc_t D, int row) -> float {
            if constexpr (TL) return dzy_t(D, row);
            else return dzy_at(D, row);
        };
        auto DZX = [&](__amdgpu_buffer_rsrc_t D, int row) -> float {
            if constexpr (TL) return dzx_t(D, row);
            else return dzx_at(D, row);
        };
        auto ROW = [&](__amdgpu_buffer_rsrc_t A, int row) -> float {
            if constexpr (TL) return row_t(A, row);
            else return row_at(A, row);
        };
        const int lr = TL ? lrow : 0;

        // Prologue: rows a0-1 and a0 (read at the kernel's start when the step reads the copy
        // predicted there), and the y-face between them.
        Cell rp, r0;
        if (pre.par == par && !from_snap) {
            rp.h = pre.h0;
            rp.qx = pre.qx0;
            rp.qy = pre.qy0;
            rp.code_we = pre.cwe0;
            r0.h = pre.h1;
            r0.qx = pre.qx1;
            r0.qy = pre.qy1;
            r0.code_we = pre.cwe1;
        } else {
            rp = FETCH(a0 - 1);
            r0 = FETCH(a0);
        }
        const Cell cp = fix(rp, a0 - 1 + lr);
        const Cell c0 = fix(r0, a0 + lr);
        float ch = c0.h, cqx = c0.qx, cqy = c0.qy;
        if (!TL && rec && ring_owner && a0 == 0) ring_put(-1, cp.h, cp.qx, cp.qy);
        const Vel vp = cell_velocity(cp.h, cp.qx, cp.qy, K);
        Vel vc = cell_velocity(ch, cqx, cqy, K);
        Flux fn;
        if (CREST) fn = face_crest<false>(cp.h, vp, ch, vc, DZY(rs_dzy, a0), DZY(rs_dzy2, a0), K);
        else fn = face_signed<false>(cp.h, vp, ch, vc, DZY(rs_dzy, a0), K);
        if (a0 == 0 && owner && lr == 0) yb_out[col] = fn.m;

        // Row inputs of iteration j: the old state of row j+1, the y-face jumps of row j+1
        // (south faces of row j), the x-face jumps, wet and g n^2 of row j.
        struct RowIn {
            Cell c;
            float dzs, dzs2, dzw, dzw2, wet, gn2;
        };
        auto load_in = [&](int jj) -> RowIn {
            RowIn q;
            q.c = FETCH(jj + 1);
            q.dzs = DZY(rs_dzy, jj + 1);
            q.dzs2 = CREST ? DZY(rs_dzy2, jj + 1) : 0.0f;
            q.dzw = DZX(rs_dzx, min(jj, ny - 1));
            q.dzw2 = CREST ? DZX(rs_dzx2, min(jj, ny - 1)) : 0.0f;
            q.wet = ROW(rs_wet, jj);
            q.gn2 = ROW(rs_gn2, jj);
            return q;
        };

        // The new state of a row is stored one row late, right after the next row's wait:
        // the vector-memory counter counts loads and stores together, and with loads and
        // stores both pending every wait is a full drain, so a row's stores issued at its end
        // would hold up the next row's start by their whole latency.  Issued after the wait,
        // they complete under a row of work like the prefetched loads.
        float held_h = 0.0f, held_qx = 0.0f, held_qy = 0.0f;
        int held_row = -1;        // the row whose new state is held (-1: none)
        bool held_rv = false;     // tail wave: the lane's held row is a row of the band
        auto flush = [&]() {
            if (held_row >= 0 && upd && (!TL || held_rv)) {
                if constexpr (TL) {
                    const uint32_t vo = cb + (uint32_t)(held_row + lr) * rowb;
                    bstore(held_h, rs_ho, vo, 0u);
                    bstore(held_qx, rs_ho, vo, planeb);
                    bstore(held_qy, rs_ho, vo, 2 * planeb);
                } else {
                    const uint32_t so = (uint32_t)held_row * rowb;
                    bstore(held_h, rs_ho, cb, so);
                    bstore(held_qx, rs_ho, cb, so + planeb);
                    bstore(held_qy, rs_ho, cb, so + 2 * planeb);
                }
            }
        };
        // One row of the march: faces, update, row flags; the row below becomes the current
        // one.  `prefetch` issues the next row's loads once this row's inputs have arrived,
        // then the previous row's stores go out (the scheduling barrier keeps both after
        // that wait).
        auto row_step = [&](int j, const RowIn& in, auto&& prefetch) {
            const int jl = j + lr;          // this lane's row
            const bool rv = !TL || jl < b0;   // a row of the band (a tail group's last rows may not be)
            const Cell cn = fix(in.c, jl + 1);
            __builtin_amdgcn_sched_barrier(0);
            prefetch();
            flush();
            __builtin_amdgcn_sched_barrier(0);
            const float nh = cn.h, nqx = cn.qx, nqy = cn.qy;
            const float dzs = in.dzs, dzs2 = in.dzs2, dzw = in.dzw, dzw2 = in.dzw2;
            const float wet = in.wet, gn2 = in.gn2;
            const Vel vn = cell_velocity(nh, nqx, nqy, K);
            Flux fs;
            if (CREST) fs = face_crest<false>(ch, vc, nh, vn, dzs, dzs2, K);
            else fs = face_signed<false>(ch, vc, nh, vn, dzs, K);
            // West face of this lane's column, from the left neighbour's old state.
            Vel vw;
            const float hw = from_left(ch);
            vw.u = from_left(vc.u);
            vw.v = from_left(vc.v);
            vw.c = from_left(vc.c);
            Flux fw;
            if (CREST) fw = face_crest<true>(hw, vw, ch, vc, dzw, dzw2, K);
            else fw = face_signed<true>(hw, vw, ch, vc, dzw, K);
            const float em = from_right(fw.m);
            const float enl = from_right(fw.nl);
            const float et = from_right(fw.t);
            const CellNew u = cell_update(ch, cqx, cqy, fw.m, fw.nr, fw.t, em, enl, et, fn.m, fn.nr, fn.t, fs.m, fs.nl,
                                          fs.t, wet, gn2, K);
            const int r = jl - a0;
            if (!TL && SPECIAL && (rec || chk)) {
                if (rec && ring_owner) ring_put(j, ch, cqx, cqy);
                const int bi = (rec_band || chk) ? band_at(j) : -1;
                if (bi >= 0) {
                    float* bt = G.tape_band + ((size_t)sync_idx * G.n_band + bi) * 3;
                    if (rec_band) {
                        bt[0] = u.h;
                        bt[1] = u.qx;
                        bt[2] = u.qy;
                    } else {
                        dh_l = fmax(dh_l, fabs((double)u.h - (double)bt[0]));
                        dq_l = fmax(fmax(dq_l, fabs((double)u.qx - (double)bt[1])), fabs((double)u.qy - (double)bt[2]));
                    }
                }
            }
            held_h = u.h;
            held_qx = u.qx;
            held_qy = u.qy;
            held_row = j;
            held_rv = rv;
            if (upd && rv) {
                l_infil[r * nx + col] = u.infil;
                smax_bits = max(smax_bits, __float_as_int(u.s));
                if (sync) sbuf[(size_t)jl * P + ccol] = u.s;
            }
            if constexpr (TL) bad_mask |= __builtin_amdgcn_ballot_w64(upd && rv && u.bad);
            else   // u.bad = !(h1 >= -h_dry) || qm is NaN
                bad_mask |= (__builtin_amdgcn_fcmpf(u.h1, -K.h_dry, 12) | __builtin_amdgcn_fcmpf(u.qm, u.qm, 8)) & upd_mask;
            if (edge_wave) {
                if (col == 0 && gl >= 1 && rv) l_bw[r] = fw.m;
                if (col == nx && gl >= 1 && rv) l_be[r] = fw.m;
            }
            if constexpr (TL) {   // the groups' rows differ: per-lane row flags
                if (upd && rv && u.h > 0.0f) atomicOr(&l_wet[r], 1u);
                if (__builtin_expect(__builtin_amdgcn_ballot_w64(upd && rv && u.h1 < 0.0f) != 0ull, 0)) {
                    if (upd && rv) fixbuf[(size_t)jl * P + ccol] = u.fix;
                    if (upd && rv && u.fix != 0.0f) atomicOr(&l_fixmask[r], 1u << wave);
                }
                if (b0 == ny && owner && jl == ny - 1) yb_out[nx + col] = fs.m;
            } else {
                // (a plain store: every writer of a row's word writes 1, the tail wave with an
                // atomic or of 1, so the word ends at 1 in any order)
                const uint64_t wb = lanes_ogt0(u.h) & upd_mask;
                if (lane == 0 && wb != 0ull) l_wet[r] = 1u;
                // (fix = -h1 where h1 < 0, else +0: nonzero exactly where h1 < 0)
                if (__builtin_expect((lanes_olt0(u.h1) & upd_mask) != 0ull, 0)) {   // rare: a cell went negative
                    if (upd) fixbuf[(size_t)j * P + ccol] = u.fix;
                    if (lane == 0) atomicOr(&l_fixmask[r], 1u << wave);
                }
            }
            ch = nh;
            cqx = nqx;
            cqy = nqy;
            vc = vn;
            fn = fs;
        };

        // Two input sets in ping-pong: a row's loads are issued one row of work ahead, and
        // the unrolled pair needs no register copies to rotate the pipeline.
        const int j_end = TL ? a0 + t_rows : b0;
        RowIn in0 = load_in(a0), in1;
        int j = a0;
        for (; j + 1 < j_end; j += 2) {
            row_step(j, in0, [&] { in1 = load_in(j + 1); });
            row_step(j + 1, in1, [&] { in0 = load_in(j + 2); });   // (clamped loads; unused past the band)
        }
        if (j < j_end) row_step(j, in0, [] {});
        flush();
        if constexpr (!TL) {
            if (b0 == ny && owner) yb_out[nx + col] = fn.m;
            if (SPECIAL && rec && ring_owner && b0 == ny) ring_put(ny, ch, cqx, cqy);   // the south ghost row
        }
    };
    if constexpr (SPECIAL) march(BoolC<false>{});
    else if (tail) march(BoolC<true>{});
    else march(BoolC<false>{});
    if (SPECIAL && chk) {   // band deviation maxima (f64 >= 0, never NaN: max ignores NaN)
        unsigned long long bh = (unsigned long long)__double_as_longlong(dh_l);
        unsigned long long bq = (unsigned long long)__double_as_longlong(dq_l);
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned long long oh = __shfl_xor(bh, o, 64), oq = __shfl_xor(bq, o, 64);
            bh = oh > bh ? oh : bh;
            bq = oq > bq ? oq : bq;
        }
        if (lane == 0) {
            if (bh > 0ull) atomicMax(&acc_cur->dh64, bh);
            if (bq > 0ull) atomicMax(&acc_cur->dq64, bq);
        }
    }

    // Reductions of the band: CFL rate max (as ordered int bits; -0.0 never wins over
    // +0.0) and the health flag.
    smax_bits = wave_max_i32(smax_bits);
    const bool bad_any = bad_mask != 0ull;
    if (lane == 0) {
        l_red[wave] = smax_bits;
        l_red[kMaxWaves + wave] = bad_any ? 1 : 0;
    }
    __syncthreads();

    // Monitors of the band (kernel.rs:311-321 on the new state: strict >, time of the new
    // state).  Done after the march so that no wave of the march waits on their loads:
    // each thread reads its first monitored cell's new depth and old maximum now and stores
    // the update after the row sums, which hide the loads' latency.
    float m_h = 0.0f, m_old = 0.0f;
    double m_t = 0.0;
    if (has_m) {
        m_h = Ho[l_moff[tid]];
        m_old = mon_in[qm];
        m_t = tmax_in[qm];
    }

    // Row FixedSums (real.rs:132-175): lane L = padded column & 15, sequential in column
    // order inside a lane, then the fixed pairwise tree.
    for (int t = tid; t < nrows * 16; t += nthr) {
        const int r = t >> 4, L = t & 15;
        const int j = a0 + r;
        const int c0 = (L == 0 ? 16 : L) - 1;   // first column of this lane
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const float* rin = l_infil + (size_t)r * nx;
        acc[0] = lane_chain(c0, nx, [&](int c) { return rin[c]; });
        const uint32_t fm = l_fixmask[r];   // waves whose fix values are in fixbuf (others hold +0)
        if (fm != 0u)
            acc[1] = lane_chain(c0, nx, [&](int c) {
                return ((fm >> (c / kColsPerWave)) & 1u) ? fixbuf[(size_t)j * P + c] : 0.0f;
            });
        if (sync) {
            acc[2] = lane_chain(c0, nx, [&](int c) { return Ho[(size_t)j * P + c]; });
            acc[3] = lane_chain(c0, nx, [&](int c) { return sbuf[(size_t)j * P + c]; });
        }
#pragma unroll
        for (int q = 0; q < 4; q++) l_acc[((size_t)r * 16 + L) * 4 + q] = acc[q];
    }
    __syncthreads();
    for (int r = tid; r < nrows; r += nthr) {
        float tot[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float* A = l_acc + (size_t)r * 64 + q;
            float l8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) l8[k] = A[(2 * k) * 4] + A[(2 * k + 1) * 4];
            const float l4[4] = {l8[0] + l8[1], l8[2] + l8[3], l8[4] + l8[5], l8[6] + l8[7]};
            tot[q] = (l4[0] + l4[1]) + (l4[2] + l4[3]);
        }
        if (!sync) { tot[2] = 0.0f; tot[3] = 0.0f; }
        RowAcc q;
        q.infil = tot[0];
        q.fix = tot[1];
        q.bw = l_bw[r];
        q.be = l_be[r];
        q.vol = tot[2];
        q.sum_s = tot[3];
        q.wet = l_wet[r];
        q.pad = 0;
        ra_out[a0 + r] = q;
    }
    if (has_m) {
        const bool up = m_h > m_old;
        mon_out[qm] = up ? m_h : m_old;
        tmax_out[qm] = up ? t_new : m_t;
    }
    for (int q = qm + nthr; q < mon_q1; q += nthr) {   // bands with more monitored cells than threads
        const float hn = Ho[MONOFF[q]];
        const float old = mon_in[q];
        const bool up = hn > old;
        mon_out[q] = up ? hn : old;
        tmax_out[q] = up ? t_new : tmax_in[q];
    }
    if (tid == 0) {
        int m = 0, bd = 0;
        for (int w = 0; w < nwaves; w++) {
            m = max(m, l_red[w]);
            bd |= l_red[kMaxWaves + w];
        }
        if (m > 0) atomicMax(&acc_cur->smax_bits, m);
        if (bd) atomicOr(&acc_cur->bad, 1u);
    }
    // No device-scope fence: on gfx950 that is an L2 write-back + invalidate.  The
    // candidate's next launch reads these results; the launch boundary orders them.
}

// The completed step's sums (shd[40..45]: infiltration, fix, boundary in, boundary out,
// volume, CFL rate sum) into the new control block and, at a sync point, the ledger row
// (solver.rs:529-540, 554-592).  One thread.
__device__ __forceinline__ void ledger_finish(const DevGrid& G, const CandCtl& Si, const StepAcc& A, CandCtl& Oc, bool sync,
                                              const double* shd) {
    CandCtl* O = &Oc;
    const DevParams& p = G.prm;
    const int local = Si.local;
    const double st_infil = shd[40], st_fix = shd[41], st_in = shd[42], st_out = shd[43];
    const double st_vol = shd[44], st_sum_s = shd[45];
    const double area = G.area;
    const uint64_t cells = (uint64_t)G.nx * (uint64_t)G.ny;
    O->acc_rain += Si.rain * area * (double)G.wet_count;
    O->inflow += st_in;
    O->outflow += st_out;
    O->infil += st_infil * area;
    O->fix += st_fix * area;
    O->cell_updates += cells - (uint64_t)Si.skipped_now * (uint64_t)G.nx;
    O->rows_skipped += Si.skipped_now;
    if (sync) {
        const double smax = step_smax(G, A);
        if (st_sum_s > 0.0) O->lts = fmax(O->lts, smax * (double)cells / st_sum_s);
        const double v = st_vol * area;
        const double expected = O->v0 + O->acc_rain + O->inflow - O->outflow - O->infil + O->fix;
        if ((int)O->n_ledger < p.max_ledger) {
            double* row = G.ledger + ((size_t)local * p.max_ledger + O->n_ledger) * 8;
            row[0] = O->t; row[1] = v; row[2] = O->acc_rain; row[3] = O->inflow;
            row[4] = O->outflow; row[5] = O->infil; row[6] = O->fix; row[7] = v - expected;
        }
        O->n_ledger += 1;
    }
}

// Ledger accounting of the completed step (kernel.rs:453-500 sums, solver.rs:529-540,
// 554-592), run by the candidate's control task while the bands march the next step.  Row-ordered f64
// sums of the row partials and the FixedSum<f64> of the boundary y-faces, then thread
// 0 adds them to the new control block (in LDS).  `Sin` is the block of the previous
// launch (the completed step's dt, rain, dry-row count).
__device__ __forceinline__ void ledger_epilogue(const KArgs& a, const DevGrid& G, const CandCtl* Sin, const StepAcc& A,
                                                CandCtl& Oc, float* stage, size_t stage_floats, double* shd) {
    const CandCtl& Si = *Sin;
    const int tid = threadIdx.x;
    const int nthr = G.waves * kWave;
    const int local = Si.local;
    const int par = Si.parity ^ 1;   // parity the completed step wrote
    const RowAcc* ra = G.rowacc + slot(G, par, local) * G.ny;
    const int nx = G.nx, ny = G.ny;
    const double dt = Si.dt;
    const double dt_dy = dt * G.dy;
    const double dt_dx = dt * G.dx;
    __syncthreads();   // the band step is done with the LDS
    {
        // y-boundary faces (kernel.rs:453-472): FixedSum<f64> over padded columns 1..nx.
        const float* yb_src = G.ybnd + slot(G, par, local) * 2 * nx;
        for (int i = tid; i < 2 * nx; i += nthr) stage[i] = yb_src[i];
        __syncthreads();
        if (tid < 32) {
            const int L = tid & 15;
            const float* row = stage + (tid >> 4) * nx;
            double yb = 0.0;
            for (int i = (L == 0 ? 16 : L); i <= nx; i += 16) yb += (double)row[i - 1];
            shd[tid] = yb;
        }
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const double* acc = shd + 16 * s;
            double l8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) l8[k] = acc[2 * k] + acc[2 * k + 1];
            const double l4[4] = {l8[0] + l8[1], l8[2] + l8[3], l8[4] + l8[5], l8[6] + l8[7]};
            const double v = ((l4[0] + l4[1]) + (l4[2] + l4[3])) * dt_dx;
            double fin = 0.0, fout = 0.0;
            if (s == 0) { if (v > 0.0) fin = v; else fout = -v; }
            else if (v > 0.0) fout = v;
            else fin = -v;
            shd[32 + 2 * s] = fin;
            shd[33 + 2 * s] = fout;
        }
    }
    __syncthreads();
    const bool sync = Si.hits_sync != 0 && !A.bad;
    {
        // Row-ordered f64 sums of the row partials (finalize, kernel.rs:485-500), staged
        // through LDS in chunks; lanes 0-5 of wave 0 each run one sequential chain:
        // infil, fix, bnd_in, bnd_out, volume, sum_s.
        double s = 0.0;
        if (tid == 2) s = shd[32] + shd[34];   // st.bnd_in = top.0 + bot.0
        if (tid == 3) s = shd[33] + shd[35];   // st.bnd_out = top.1 + bot.1
        const int chunk = max(1, (int)(stage_floats / 8));
        for (int r0 = 0; r0 < ny; r0 += chunk) {
            const int nr = min(chunk, ny - r0);
            const uint4* src = (const uint4*)(ra + r0);   // a RowAcc is two uint4
            uint4* st4 = (uint4*)stage;
            for (int i = tid; i < nr * 2; i += nthr) st4[i] = src[i];
            __syncthreads();
            if (tid <= 1 || (sync && (tid == 4 || tid == 5))) {
                const int f = tid;   // RowAcc word: 0 infil, 1 fix, 4 vol, 5 sum_s
#pragma unroll 8
                for (int r = 0; r < nr; r++) s += (double)stage[r * 8 + f];
            } else if (tid == 2 || tid == 3) {
#pragma unroll 4
                for (int r = 0; r < nr; r++) {
                    const double w = (double)stage[r * 8 + 2] * dt_dy;
                    const double e = (double)stage[r * 8 + 3] * dt_dy;
                    double ai = 0.0, ao = 0.0;
                    if (w > 0.0) ai += w; else ao -= w;
                    if (e > 0.0) ao += e; else ai -= e;
                    s += tid == 2 ? ai : ao;
                }
            }
            __syncthreads();
        }
        if (tid < 6) shd[40 + tid] = s;
    }
    __syncthreads();
    if (tid == 0) ledger_finish(G, Si, A, Oc, sync, shd);
}

// ------------------------------------------------------------------ local-inertial screening step
// crates/itr-hydro/src/inertial.rs:60-219, the physics the reference runs on coarse
// fidelity levels (crates/itr-opt/src/problem.rs:178-181).  Staggered discharges on the
// grid's state planes: the qx plane holds the east face of each cell (face f = col + 1)
// and the qy plane the south face of each row (face k = row + 1), as the reference's
// qx[at(i, j)] / qy[at(i, j)]; the west boundary faces (f = 0) and north boundary faces
// (k = 0) live in the grid's `lif` array.  Every face and cell is the reference's f64
// expression of the f32 state, evaluated in its order without fused operations.
// A wave tiles 61 columns: lanes 2..62 update, lanes 0, 1 and 63 are halos (lane 1's
// east face is lane 2's west face; the theta weighting of a face reads its neighbours'
// old discharges, so each computed face needs one lane on either side).
constexpr int kLiCols = 61;
constexpr double kLiTheta = 0.8;   // de Almeida et al. (2012) weighting (inertial.rs:85-98)

struct LiK {
    double dt, rx, ry, rain, cap, h_dry, g, dx, dy, dmin, inv_dx, inv_dy, inv_dmin;
    bool p2;   // dx and dy are powers of two
};

// x / d for a grid spacing d.  When d is a power of two, 1/d is exact and x * (1/d)
// rounds the same exact quotient as x / d: the same bits, without an f64 division.
__device__ __forceinline__ double li_div(double x, double d, double inv_d, const LiK& K) {
    double q;
    if (K.p2) q = x * inv_d;
    else q = x / d;
    return q;
}
__device__ __forceinline__ bool pow2_spacing(double d) {
    return d > 0.0 && d < 1e300 && (__double_as_longlong(d) & 0x000FFFFFFFFFFFFFll) == 0;
}

// face_q (inertial.rs:46-58): zero when the face depth is at most h_dry.
__device__ __forceinline__ double li_face_q(double q, double hl, double hr, double dl, double dr, double d, double inv_d,
                                            double gn2, const LiK& K) {
    const double el = hl - dl, er = hr - dr;
    const double hf = fmax(fmax(el, er), 0.0);
    double out = 0.0;
    if (hf > K.h_dry) {
        const double slope = li_div(er - el, d, inv_d, K);
        const double r = rcbrt64(hf);
        const double r7 = r * r * r * r * r * r * r;
        out = (q - K.g * hf * K.dt * slope) / (1.0 + gn2 * K.dt * fabs(q) * r7);
    }
    return out;
}
// rises / jump (inertial.rs:25-43): signed faces store the jump z_R - z_L, crest faces
// the two rises.
template <bool CREST> __device__ __forceinline__ void li_rises(float f0, float f1, double& dl, double& dr) {
    if (CREST) {
        dl = (double)f0;
        dr = (double)f1;
    } else {
        const double dz = (double)f0;
        dl = fmax(dz, 0.0);
        dr = fmax(-dz, 0.0);
    }
}
template <bool CREST> __device__ __forceinline__ double li_jump(float f0, float f1) {
    return CREST ? (double)f0 - (double)f1 : (double)f0;
}
// A boundary face (inertial.rs:104-124, 143-157).  lo: west / north face (the interior
// cell is on its right), else east / south.  qold: the face's discharge before the step;
// hc, n2c: the interior cell's depth and g n^2; hg: the stage ghost depth; qv: the inflow
// ghost discharge; f0, f1: the adjacent interior face (transmissive bed slope).
template <bool CREST>
__device__ __forceinline__ double li_boundary(int code, bool lo, float qold, float hc, float hg, float qv, float f0, float f1,
                                              bool has_face, double d, double inv_d, float n2c, const LiK& K) {
    double q = 0.0;
    if (code == 1) {   // normal-depth outflow only
        double dz = has_face ? li_jump<CREST>(f0, f1) : 0.0;
        dz = lo ? fmax(dz, 0.0) : fmin(dz, 0.0);
        const double h = (double)hc;
        q = li_face_q((double)qold, h, h, fmax(dz, 0.0), fmax(-dz, 0.0), d, inv_d, (double)n2c, K);
        q = lo ? fmin(q, 0.0) : fmax(q, 0.0);
    } else if (code == 3) {
        q = (lo ? 1.0 : -1.0) * fabs((double)qv);
    } else if (code == 2) {
        q = lo ? li_face_q((double)qold, (double)hg, (double)hc, 0.0, 0.0, d, inv_d, (double)n2c, K)
               : li_face_q((double)qold, (double)hc, (double)hg, 0.0, 0.0, d, inv_d, (double)n2c, K);
    }
    return q;
}

// Dynamic LDS of the local-inertial step, in floats.
struct LiLds { size_t fixmask, red, smx, dsum, acc, moff, bnd, gsp, total; };
__host__ __device__ inline LiLds li_lds_layout(int R, int nx, bool special) {
    auto al = [](size_t v) { return (v + 3) & ~(size_t)3; };
    LiLds L;
    L.fixmask = 0;
    L.red = al(L.fixmask + (size_t)R);
    L.smx = al(L.red + kMaxWaves);                  // kMaxWaves u64
    L.dsum = al(L.smx + 2 * kMaxWaves);             // kMaxWaves x 2 f64
    L.acc = al(L.dsum + 4 * kMaxWaves);             // R x 16 FixedSum lanes
    L.moff = al(L.acc + (size_t)R * 16);
    L.bnd = al(L.moff + (size_t)kMaxWaves * 64);      // the band's west / east boundary faces
    L.gsp = al(L.bnd + (size_t)2 * R);
    L.total = al(L.gsp + (special ? 4 * gsp_pairs(R, nx) : 0));
    if (L.total < 64) L.total = 64;
    return L;
}

// One local-inertial step of one band (rows a0 .. b0-1) of one candidate: reads state
// copy `par` (and the boundary faces of that parity), writes copy par^1.  Per row: the
// row's x-faces, its south y-face (the north one is carried from the row above, or
// computed in the prologue), then continuity and sources (inertial.rs:160-196).
template <bool CREST, bool SPECIAL>
__device__ __forceinline__ void li_band_step(const KArgs& a, const DevGrid& G, int local, int band, const LiK& K, double t,
                                             double t_new, bool sync, int par, bool from_snap, StepAcc* acc_cur,
                                             float* lds, const MonRange& mr) {
    const int nx = G.nx, ny = G.ny, P = G.P, R = G.R;
    const int a0 = band * R;
    const int b0 = min(a0 + R, ny);
    const int nrows = b0 - a0;
    const int tid = threadIdx.x;
    const int nthr = G.waves * kWave;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int col = wave * kLiCols - 2 + lane;
    const bool upd = lane >= 2 && lane <= kLiCols + 1 && col < nx;
    const bool west_lane = col == -1;     // computes face 0 of each row
    const bool east_lane = col == nx - 1; // computes face nx
    const int npar = par ^ 1;
    const size_t plane = G.plane;
    float* base = G.state + (size_t)local * 6 * plane;
    ITR_GLOBAL float* __restrict__ Ho = gp(base + (size_t)npar * 3 * plane);
    const ITR_GLOBAL int32_t* __restrict__ GHOST = gp(G.ghost);
    // The old state: copy `par`, or the sync snapshot after a locked-dt restart.
    const float* old_state = from_snap ? G.snap_state + (size_t)local * 3 * plane : base + (size_t)par * 3 * plane;
    const ITR_GLOBAL float* __restrict__ lif_in =
        gp(from_snap ? G.snap_lif + (size_t)local * (ny + nx) : G.lif + slot(G, par, local) * (size_t)(ny + nx));
    ITR_GLOBAL float* __restrict__ lif_out = gp(G.lif + slot(G, npar, local) * (size_t)(ny + nx));
    const ITR_GLOBAL float* __restrict__ mon_in =
        gp(from_snap ? G.snap_mon + (size_t)local * G.n_mon : G.mon_max + slot(G, par, local) * G.n_mon);
    const ITR_GLOBAL double* __restrict__ tmax_in =
        gp(from_snap ? G.snap_tmax + (size_t)local * G.n_mon : G.mon_tmax + slot(G, par, local) * G.n_mon);
    ITR_GLOBAL float* __restrict__ mon_out = gp(G.mon_max + slot(G, npar, local) * G.n_mon);
    ITR_GLOBAL double* __restrict__ tmax_out = gp(G.mon_tmax + slot(G, npar, local) * G.n_mon);
    const int mon_q0 = mr.q0, mon_q1 = mr.q1;
    const ITR_GLOBAL uint32_t* __restrict__ MONOFF = gp(G.mon_off);
    const int qm = mon_q0 + tid;
    const bool has_m = qm < mon_q1;
    const uint32_t m_off0 = mr.off0;
    ITR_GLOBAL double* __restrict__ fix64 = gp(G.fix64 + slot(G, npar, local) * plane);
    LiRowAcc* ra_out = (LiRowAcc*)(G.rowacc + slot(G, npar, local) * ny);
    const uint32_t ccol = (uint32_t)min(max(col, 0), nx - 1);
    const uint32_t fcol = (uint32_t)max(min(max(col + 1, 1), nx - 1), 1);   // the lane's x-face jump column
    const uint32_t cb = ccol * 4u, xb = fcol * 4u;
    const uint32_t rowb = (uint32_t)P * 4u;
    const uint32_t planeb = (uint32_t)(plane * 4u);
    const __amdgpu_buffer_rsrc_t rs_h = make_rsrc(old_state, 3 * plane * 4);
    const __amdgpu_buffer_rsrc_t rs_ho = make_rsrc(base + (size_t)npar * 3 * plane, 3 * plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzx = make_rsrc(G.dzx + (size_t)local * plane, plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzy = make_rsrc(G.dzy + (size_t)local * G.plane_y, G.plane_y * 4);
    const __amdgpu_buffer_rsrc_t rs_dzx2 = make_rsrc(CREST ? G.dzx2 + (size_t)local * plane : G.dzx, plane * 4);
    const __amdgpu_buffer_rsrc_t rs_dzy2 = make_rsrc(CREST ? G.dzy2 + (size_t)local * G.plane_y : G.dzy, G.plane_y * 4);
    const __amdgpu_buffer_rsrc_t rs_wet = make_rsrc(G.wet, plane * 4);
    const __amdgpu_buffer_rsrc_t rs_gn2 = make_rsrc(G.gn2, plane * 4);
    const int code_n = GHOST[ccol];
    const int code_s = GHOST[nx + ccol];
    const int we_base = 2 * nx + (west_lane ? 0 : ny);   // this lane's west / east ghost codes
    const bool special = SPECIAL && G.special != 0;
    const LiLds LL = li_lds_layout(R, nx, G.special != 0);
    float4* l_gsp = (float4*)(lds + LL.gsp);
    uint32_t* l_fixmask = (uint32_t*)(lds + LL.fixmask);
    int* l_red = (int*)(lds + LL.red);
    unsigned long long* l_smx = (unsigned long long*)(lds + LL.smx);
    double* l_dsum = (double*)(lds + LL.dsum);
    float* l_acc = lds + LL.acc;
    uint32_t* l_moff = (uint32_t*)(lds + LL.moff);
    // (the band starting at row 1 reads the north inflow discharge for the theta
    // weighting of face 1)
    if (special) (void)stage_special_ghosts(G, local, a0, b0, t, l_gsp, 1, 0);
    for (int r = tid; r < R; r += nthr) l_fixmask[r] = 0u;
    if (has_m) l_moff[tid] = m_off0;
    __syncthreads();
    // The band's west and east boundary faces (inertial.rs:104-124), one thread each, so
    // that the march's edge waves carry no boundary branch.
    float* l_bnd = lds + LL.bnd;
    {
        const float* S0 = old_state;
        for (int i = tid; i < 2 * nrows; i += nthr) {
            const bool west = i < nrows;
            const int r = a0 + (west ? i : i - nrows);
            const size_t ro = (size_t)r * P;
            const int code = GHOST[2 * nx + (west ? 0 : ny) + r];
            float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (special && code >= 2) gs = l_gsp[(west ? 0 : R + 2) + (r - (a0 - 1))];
            const int fc = west ? min(1, nx - 1) : nx - 1;   // the adjacent interior face (nx > 1)
            const int cc = west ? 0 : nx - 1;                // the interior cell
            float qold;
            if (west) {
                qold = lif_in[r];
                if (special && code == 3) qold = gs.y;
            } else {
                qold = S0[plane + ro + cc];
            }
            const float f0 = G.dzx[(size_t)local * plane + ro + max(fc, 1)];
            const float f1 = CREST ? G.dzx2[(size_t)local * plane + ro + max(fc, 1)] : 0.0f;
            const double q = li_boundary<CREST>(code, west, qold, S0[ro + cc], gs.x, gs.y, f0, f1, nx > 1, K.dx,
                                                    K.inv_dx, G.gn2[ro + cc], K);
            l_bnd[i] = canon((float)q);
        }
    }
    __syncthreads();
    // Special ghost pairs of this lane: west / east of row r, north / south of its column.
    auto gs_we = [&](int r) -> float4 { return l_gsp[(west_lane ? 0 : R + 2) + (r - (a0 - 1))]; };
    auto gs_ns = [&](bool north) -> float4 { return l_gsp[2 * (R + 2) + (north ? 0 : nx) + (int)ccol]; };

    struct Row { float h, qx, qy, gn2, wet, dzx, dzx2, dzy, dzy2; };
    // Row r of the old state (clamped rows are never used); the west lane's qx is the
    // old west boundary face (the inflow ghost discharge at an inflow cell, as the
    // reference's ghost fill leaves it there).
    auto fetch = [&](int r) -> Row {
        const int rr = min(max(r, 0), ny - 1);
        const uint32_t so = (uint32_t)rr * rowb;
        const uint32_t sy = (uint32_t)min(max(r, 0), ny) * rowb;
        Row v;
        v.h = bload(rs_h, cb, so);
        v.qx = bload(rs_h, cb, so + planeb);
        v.qy = bload(rs_h, cb, so + 2 * planeb);
        v.gn2 = bload(rs_gn2, cb, so);
        v.wet = bload(rs_wet, cb, so);
        v.dzx = bload(rs_dzx, xb, so);
        v.dzy = bload(rs_dzy, cb, sy);
        v.dzx2 = CREST ? bload(rs_dzx2, xb, so) : 0.0f;
        v.dzy2 = CREST ? bload(rs_dzy2, cb, sy) : 0.0f;
        if (west_lane) {
            float w = lif_in[rr];
            if (special && GHOST[we_base + rr] == 3) w = gs_we(rr).y;
            v.qx = w;
        }
        return v;
    };
    // Old discharge of y-face k (0 ..= ny) in this lane's column.
    auto qy_old = [&](int k) -> float {
        if (k > 0) return bload(rs_h, cb, (uint32_t)(k - 1) * rowb + 2 * planeb);
        float v = lif_in[ny + ccol];
        if (special && code_n == 3) v = gs_ns(true).y;
        return v;
    };
    // New x-face f = col + 1 of row r (c: this lane's cell; right-neighbour values by DPP).
    auto xface = [&](const Row& c, int r) -> float {
        const float h_r = from_right(c.h);
        const float g2_r = from_right(c.gn2);
        const float q_l = from_left(c.qx);
        const float q_r = from_right(c.qx);
        double dl, dr;
        li_rises<CREST>(c.dzx, c.dzx2, dl, dr);
        const double n2 = 0.5 * ((double)c.gn2 + (double)g2_r);
        double qt = (double)c.qx;
        if (col + 2 < nx) qt = kLiTheta * qt + 0.5 * (1.0 - kLiTheta) * ((double)q_l + (double)q_r);
        const float qi = canon((float)li_face_q(qt, (double)c.h, (double)h_r, dl, dr, K.dx, K.inv_dx, n2, K));
        // boundary faces come from the band's prologue
        return west_lane ? l_bnd[r - a0] : east_lane ? l_bnd[nrows + r - a0] : qi;
    };
    // New y-face k between rows k-1 (u) and k (d), 1 <= k <= ny-1: u.qy is its old
    // discharge, qp and d.qy those of faces k-1 and k+1.
    auto yface = [&](const Row& u, const Row& d, float qp, int k) -> float {
        double dl, dr;
        li_rises<CREST>(d.dzy, d.dzy2, dl, dr);
        const double n2 = 0.5 * ((double)u.gn2 + (double)d.gn2);
        double qt = (double)u.qy;
        if (k + 1 < ny) qt = kLiTheta * qt + 0.5 * (1.0 - kLiTheta) * ((double)qp + (double)d.qy);
        return canon((float)li_face_q(qt, (double)u.h, (double)d.h, dl, dr, K.dy, K.inv_dy, n2, K));
    };
    auto yface_south = [&](const Row& u) -> float {   // k = ny, u = row ny-1
        float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (special && code_s >= 2) gs = gs_ns(false);
        return canon((float)li_boundary<CREST>(code_s, false, u.qy, u.h, gs.x, gs.y, u.dzy, u.dzy2, ny > 1, K.dy,
                                                   K.inv_dy, u.gn2, K));
    };
    auto yface_north = [&](const Row& d, const Row& d1) -> float {   // k = 0, d = row 0, d1 = row 1
        float4 gs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (special && code_n >= 2) gs = gs_ns(true);
        return canon((float)li_boundary<CREST>(code_n, true, lif_in[ny + ccol], d.h, gs.x, gs.y, d1.dzy, d1.dzy2,
                                                   ny > 1, K.dy, K.inv_dy, d.gn2, K));
    };

    Row cur = fetch(a0);
    Row nxt = fetch(a0 + 1);
    float qn_new, qprev;   // new north face of the row; old discharge of that face
    if (a0 == 0) {
        qn_new = yface_north(cur, nxt);
        qprev = qy_old(0);
        if (upd) lif_out[ny + col] = qn_new;
    } else {
        const Row up = fetch(a0 - 1);
        qn_new = yface(up, cur, qy_old(a0 - 1), a0);
        qprev = up.qy;
    }
    // The faces of a row depend on the old state only, so each iteration computes the next
    // row's faces beside this row's continuity (two independent f64 chains per lane).
    // Row loads are issued two rows ahead of the row they complete (rows j+2 and j+3 are
    // in flight while row j is finished).
    float qe = xface(cur, a0);
    float qs = a0 + 1 < ny ? yface(cur, nxt, qprev, a0 + 1) : yface_south(cur);
    double smax_l = 0.0, inf_l = 0.0, ss_l = 0.0;
    bool bad = false;
    Row pre = fetch(a0 + 2);
    for (int j = a0; j < b0; j++) {
        const Row pre2 = fetch(j + 3);
        float qe_n = 0.0f, qs_n = 0.0f;   // faces of row j + 1
        if (j + 1 < b0) {
            qe_n = xface(nxt, j + 1);
            qs_n = j + 2 < ny ? yface(nxt, pre, cur.qy, j + 2) : yface_south(nxt);
        }
        const float qw = from_left(qe);
        // Continuity and sources (inertial.rs:166-196).
        const double wet = (double)cur.wet;
        const double dw = (double)qw, de = (double)qe, dn = (double)qn_new, ds = (double)qs;
        double hn = (double)cur.h + K.rx * (dw - de) + K.ry * (dn - ds) + K.rain * wet;
        const double infil = fmin(K.cap, fmax(hn, 0.0)) * wet;
        hn -= infil;
        double fixv = 0.0;
        if (hn < 0.0) {
            fixv = -hn;
            hn = 0.0;
        }
        const float hnew = canon((float)(hn * wet));
        double s = 0.0;
        if (hn > K.h_dry) {
            // u.max(v) with u = A / hn, v = B / hn: correctly rounded division by a finite
            // positive hn is monotone, so max(A / hn, B / hn) = max(A, B) / hn bit for bit
            // (a NaN operand: fmax keeps the other one on both sides), one f64 division
            // instead of two.  An infinite hn takes the two divisions.
            const double A = 0.5 * (fabs(dw) + fabs(de)), B = 0.5 * (fabs(dn) + fabs(ds));
            double uv;
            if (__builtin_expect(hn < __builtin_inf(), 1)) uv = fmax(A, B) / hn;
            else uv = fmax(A / hn, B / hn);
            s = li_div((sqrt(K.g * hn) + uv) * 0.7, K.dmin, K.inv_dmin, K);
        }
        const int r = j - a0;
        if (upd) {
            const uint32_t so = (uint32_t)j * rowb;
            bstore(hnew, rs_ho, cb, so);
            bstore(qe, rs_ho, cb, so + planeb);
            bstore(qs, rs_ho, cb, so + 2 * planeb);
            bad = bad || !__builtin_isfinite(hn);
            smax_l = fmax(smax_l, s);
            inf_l += infil;
            if (sync) ss_l += s;
        }
        if (west_lane) lif_out[j] = qe;
        if (__builtin_expect(__ballot(upd && fixv != 0.0) != 0ull, 0)) {   // rare: a cell went negative
            if (upd) fix64[(size_t)j * P + col] = fixv;
            if (lane == 0) atomicOr(&l_fixmask[r], 1u << wave);
        }
        qn_new = qs;
        qe = qe_n;
        qs = qs_n;
        cur = nxt;
        nxt = pre;
        pre = pre2;
    }

    // Band reductions: f64 CFL max (ordered bits; rates are >= 0 and never NaN here),
    // health flag, and the band's infiltration and CFL-rate sums (fixed lane tree, then
    // waves in order).
    unsigned long long sm = (unsigned long long)__double_as_longlong(smax_l);
    double si = inf_l, ss = ss_l;
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long om = __shfl_xor(sm, o, 64);
        sm = om > sm ? om : sm;
        si += __shfl_xor(si, o, 64);
        ss += __shfl_xor(ss, o, 64);
    }
    const bool bad_any = __any(bad);
    if (lane == 0) {
        l_smx[wave] = sm;
        l_red[wave] = bad_any ? 1 : 0;
        l_dsum[2 * wave] = si;
        l_dsum[2 * wave + 1] = ss;
    }
    __syncthreads();
    // Monitors (inertial.rs:202-209 on the new state), loads issued before the row sums.
    float m_h = 0.0f, m_old = 0.0f;
    double m_t = 0.0;
    if (has_m) {
        m_h = Ho[l_moff[tid]];
        m_old = mon_in[qm];
        m_t = tmax_in[qm];
    }
    if (sync) {   // FixedSum lanes of the new depth per row (volume, solver.rs:250-262)
        for (int q = tid; q < nrows * 16; q += nthr) {
            const int r = q >> 4, L = q & 15;
            const int j = a0 + r;
            const int c0 = (L == 0 ? 16 : L) - 1;
            l_acc[r * 16 + L] = lane_chain(c0, nx, [&](int c) { return Ho[(size_t)j * P + c]; });
        }
    }
    __syncthreads();
    for (int r = tid; r < nrows; r += nthr) {
        LiRowAcc q;
        q.infil = 0.0;
        q.sum_s = 0.0;
        if (r == 0) {
            for (int w = 0; w < G.waves; w++) {
                q.infil += l_dsum[2 * w];
                q.sum_s += l_dsum[2 * w + 1];
            }
            if (!sync) q.sum_s = 0.0;
        }
        q.vol = 0.0f;
        if (sync) {
            const float* A = l_acc + (size_t)r * 16;
            float l8[8];
#pragma unroll
            for (int k = 0; k < 8; k++) l8[k] = A[2 * k] + A[2 * k + 1];
            const float l4[4] = {l8[0] + l8[1], l8[2] + l8[3], l8[4] + l8[5], l8[6] + l8[7]};
            q.vol = (l4[0] + l4[1]) + (l4[2] + l4[3]);
        }
        q.fixmask = l_fixmask[r];
        q.wet = 0u;
        q.pad = 0u;
        ra_out[a0 + r] = q;
    }
    if (has_m) {
        const bool up = m_h > m_old;
        mon_out[qm] = up ? m_h : m_old;
        tmax_out[qm] = up ? t_new : m_t;
    }
    for (int q = qm + nthr; q < mon_q1; q += nthr) {
        const float hn = Ho[MONOFF[q]];
        const float old = mon_in[q];
        const bool up = hn > old;
        mon_out[q] = up ? hn : old;
        tmax_out[q] = up ? t_new : tmax_in[q];
    }
    if (tid == 0) {
        unsigned long long m = 0ull;
        int bd = 0;
        for (int w = 0; w < G.waves; w++) {
            m = l_smx[w] > m ? l_smx[w] : m;
            bd |= l_red[w];
        }
        if (m > 0ull) atomicMax(&acc_cur->smax64, m);
        if (bd) atomicOr(&acc_cur->bad, 1u);
    }
}

// Ledger sums of a completed local-inertial step (inertial.rs:171-216), by the
// candidate's designated workgroup (its control task or first band): the positivity fixes and the boundary discharges are the
// reference's sequential f64 sums in its order (rows, then the north / south faces by
// column); infiltration and CFL rate are f64 sums of the band sums in band order (the
// reference adds cell by cell: equal to within rounding); the volume is the row FixedSums
// (solver.rs:250-262).
__device__ __forceinline__ void li_ledger_epilogue(const DevGrid& G, const CandCtl* Sin, const StepAcc& A, CandCtl& Oc,
                                                   float* stage, size_t stage_floats, double* shd) {
    const CandCtl& Si = *Sin;
    const int tid = threadIdx.x;
    const int nthr = G.waves * kWave;
    const int local = Si.local;
    const int par = Si.parity ^ 1;   // parity the completed step wrote
    const int nx = G.nx, ny = G.ny;
    const size_t P = G.P, plane = G.plane;
    const LiRowAcc* ra = (const LiRowAcc*)(G.rowacc + slot(G, par, local) * ny);
    const float* lf = G.lif + slot(G, par, local) * (size_t)(ny + nx);
    const float* S = G.state + (size_t)local * 6 * plane + (size_t)par * 3 * plane;
    const double* fx = G.fix64 + slot(G, par, local) * plane;
    const double dt = Si.dt;
    const bool sync = Si.hits_sync != 0 && !A.bad;
    // Four sequential chains, each on its own wave when the workgroup has four (else on
    // lanes 0..3 of wave 0): 0 infiltration, 1 fixes, 2 boundary in / out, 3 volume and
    // CFL-rate sum.  Their terms are staged through LDS in chunks of rows, then columns.
    const int ch = nthr >= 4 * kWave ? ((tid & 63) == 0 ? (tid >> 6) : -1) : (tid < 4 ? tid : -1);
    const bool chain = ch >= 0 && ch < 4;
    double s0 = 0.0, s1 = 0.0;
    double* T = (double*)stage;
    const int chunk = max(1, (int)(stage_floats / 12));   // rows of 6 f64 terms
    __syncthreads();   // the band step is done with the LDS
    for (int r0 = 0; r0 < ny; r0 += chunk) {
        const int nr = min(chunk, ny - r0);
        for (int i = tid; i < nr; i += nthr) {
            const int r = r0 + i;
            const LiRowAcc q = ra[r];
            double* t = T + 6 * i;
            t[0] = q.infil;
            t[1] = (double)lf[r] * dt * G.dy;                                  // west face (inertial.rs:198-201)
            t[2] = (double)S[plane + (size_t)r * P + (nx - 1)] * dt * G.dy;    // east face
            t[3] = (double)q.vol;
            t[4] = q.sum_s;
            t[5] = __longlong_as_double((long long)q.fixmask);
        }
        __syncthreads();
        if (chain) {
            if (ch == 0) {
#pragma unroll 8
                for (int i = 0; i < nr; i++) s0 += T[6 * i];
            } else if (ch == 1) {
                for (int i = 0; i < nr; i++) {
                    uint32_t m = (uint32_t)__double_as_longlong(T[6 * i + 5]);
                    const size_t rb = (size_t)(r0 + i) * P;
                    while (m != 0u) {   // waves of the row that had a positivity fix
                        const int w = __builtin_ctz(m);
                        m &= m - 1u;
                        const int c1 = min(w * kLiCols + kLiCols, nx);
                        for (int c = w * kLiCols; c < c1; c++) s0 += fx[rb + c];
                    }
                }
            } else if (ch == 2) {
#pragma unroll 4
                for (int i = 0; i < nr; i++) {
                    const double w = T[6 * i + 1], e = T[6 * i + 2];
                    if (w > 0.0) s0 += w; else s1 -= w;
                    if (e > 0.0) s1 += e; else s0 -= e;
                }
            } else if (sync) {
#pragma unroll 8
                for (int i = 0; i < nr; i++) {
                    s0 += T[6 * i + 3];
                    s1 += T[6 * i + 4];
                }
            }
        }
        __syncthreads();
    }
    const int cchunk = max(1, (int)(stage_floats / 4));   // columns of 2 f64 terms
    for (int c0 = 0; c0 < nx; c0 += cchunk) {
        const int nc = min(cchunk, nx - c0);
        for (int i = tid; i < nc; i += nthr) {
            const int c = c0 + i;
            T[2 * i] = (double)lf[ny + c] * dt * G.dx;                                  // north face
            T[2 * i + 1] = (double)S[2 * plane + (size_t)(ny - 1) * P + c] * dt * G.dx;  // south face
        }
        __syncthreads();
        if (chain && ch == 2) {
#pragma unroll 4
            for (int i = 0; i < nc; i++) {
                const double n = T[2 * i], sv = T[2 * i + 1];
                if (n > 0.0) s0 += n; else s1 -= n;
                if (sv > 0.0) s1 += sv; else s0 -= sv;
            }
        }
        __syncthreads();
    }
    if (chain) {
        if (ch == 0) shd[40] = s0;
        if (ch == 1) shd[41] = s0;
        if (ch == 2) { shd[42] = s0; shd[43] = s1; }
        if (ch == 3) { shd[44] = s0; shd[45] = s1; }
    }
    __syncthreads();
    if (tid == 0) ledger_finish(G, Si, A, Oc, sync, shd);
}

// The band's part of the sync snapshot of a locked-dt run (solver.rs:446-448, 609-612):
// the new state rows, row partials, monitors and (local inertial) boundary faces the band
// just wrote as copy `npar`.
__device__ __forceinline__ void snapshot_band(const DevGrid& G, int local, int band, int npar) {
    const int nx = G.nx, ny = G.ny, R = G.R, P = G.P;
    const int a0 = band * R, b0 = min(a0 + R, ny), nrows = b0 - a0;
    const int tid = threadIdx.x, nthr = G.waves * kWave;
    const size_t plane = G.plane;
    __syncthreads();   // the band's stores are done
    const float* src = G.state + (size_t)local * 6 * plane + (size_t)npar * 3 * plane;
    float* dst = G.snap_state + (size_t)local * 3 * plane;
    const int per = nrows * nx;
    for (int i = tid; i < 3 * per; i += nthr) {
        const int q = i / per, r = (i % per) / nx, c = i % nx;
        const size_t o = (size_t)q * plane + (size_t)(a0 + r) * P + c;
        dst[o] = src[o];
    }
    const RowAcc* ra = G.rowacc + slot(G, npar, local) * ny;
    for (int r = tid; r < nrows; r += nthr) G.snap_rows[(size_t)local * ny + a0 + r] = ra[a0 + r];
    const float* mm = G.mon_max + slot(G, npar, local) * G.n_mon;
    const double* mt = G.mon_tmax + slot(G, npar, local) * G.n_mon;
    for (int q = G.mon_row[a0] + tid; q < G.mon_row[b0]; q += nthr) {
        G.snap_mon[(size_t)local * G.n_mon + q] = mm[q];
        G.snap_tmax[(size_t)local * G.n_mon + q] = mt[q];
    }
    if (G.li) {
        const float* lf = G.lif + slot(G, npar, local) * (size_t)(ny + nx);
        float* sl = G.snap_lif + (size_t)local * (ny + nx);
        for (int r = tid; r < nrows; r += nthr) sl[a0 + r] = lf[a0 + r];
        if (a0 == 0)
            for (int c = tid; c < nx; c += nthr) sl[ny + c] = lf[ny + c];
    }
}

// One launch = one step of every active candidate.  A workgroup first completes the
// candidate's previous step from its results (identically in every workgroup of the
// candidate), then, unless the candidate finished or waits for the host, runs its band
// of the next step.  The candidate's designated workgroup also runs the ledger accounting
// of the completed step and writes the new control block: in a launch whose bands fit the
// chip at once (latency-bound), a control task of its own (band index -1, no band) that
// runs side by side with the bands (the ledger reads the completed step's row partials,
// the bands write the other parity); in a launch of several rounds of workgroups
// (throughput-bound), the first band, after its march.  The control block is worked on in
// LDS (it is 60 words; held in registers it would set the kernel's register budget).
// lk = launch index & 3.
// MODES: face modes present in the batch (1 signed, 2 crest, 3 both); a batch of one
// mode gets a kernel without the other mode's code and its register budget.  SPECIAL:
// the batch has Stage / Inflow ghosts (their code stays out of the other kernels).
constexpr int kCtlWords = (int)(sizeof(CandCtl) / 4);
template <int MODES, bool SPECIAL, bool LI>
__device__ __forceinline__ void step_body(KArgs a, uint32_t lk) {
    extern __shared__ float lds[];
    __shared__ double shd[48];
    __shared__ CandCtl sS;
    __shared__ int s_go;
    // XCD-aware task order: each XCD takes a contiguous share of the task list (whole runs
    // of a candidate's bands, so band halo rows stay in one L2), ordered band-major by
    // order_tasks so that resident workgroups read the same rows of the shared planes.
    const int nblk = gridDim.x;
    const int per = nblk >> 3;
    const int task = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (task >= a.ntasks) return;
    const int4 tk = a.tasks[task];
    const DevGrid& G = a.grids[tk.x];
    const int tid = threadIdx.x;
    if ((tid >> 6) >= G.waves) return;   // the launch is sized for the widest grid
    const int nthr = G.waves * kWave;
    const int local = tk.y;
    const int cand = tk.w;
    const bool ctl_only = tk.z < 0;   // a control task (no band)
    const bool designated = a.ctl_task ? ctl_only : tk.z == 0;
    const DevParams& p = G.prm;
    const CandCtl* Sin = a.snap + (size_t)((lk + 1) & 1) * a.ncand + cand;
    CandCtl* Sout = a.snap + (size_t)(lk & 1) * a.ncand + cand;
    const MonRange mr = ctl_only ? MonRange{0, 0, 0u} : mon_range(G, tk.z, tid);
    if ((Sin->flags & F_DONE) && !designated) return;
    // The band's first rows from the copy the previous launch wrote (the completed step's
    // state: parity flipped when a step is in flight), under the control prologue.
    FirstRows pre{};
    pre.par = -1;
    if (!LI && !ctl_only && !(Sin->flags & F_DONE))
        pre = first_rows<SPECIAL>(G, local, tk.z, Sin->parity ^ (Sin->in_flight ? 1 : 0), tid);
    const StepAcc A = a.acc[(size_t)((lk + 3) & 3) * a.ncand + cand];
    const bool completed = Sin->in_flight != 0 && !(Sin->flags & F_DONE);
    for (int w = tid; w < kCtlWords; w += nthr) ((uint32_t*)&sS)[w] = ((const uint32_t*)Sin)[w];
    if (!LI && !ctl_only) {   // a band's row flags and boundary masses (band_step), zeroed before this barrier
        const LdsLayout LL = lds_layout(G.R, G.nx, G.special != 0);
        uint32_t* l_fixmask = (uint32_t*)(lds + LL.fixmask);
        uint32_t* l_wet = (uint32_t*)(lds + LL.wet);
        float* l_bw = lds + LL.bw;
        float* l_be = lds + LL.be;
        for (int r = tid; r < G.R; r += nthr) { l_wet[r] = 0u; l_fixmask[r] = 0u; l_bw[r] = 0.0f; l_be[r] = 0.0f; }
    }
    __syncthreads();
    if (tid < kWave) {
        const bool go = complete_step(p, G, sS, A, tid);
        // dry-row count of the prepared step (rows of the state just completed)
        if (designated && completed && !(sS.flags & F_DONE))
            sS.skipped_now = LI ? 0u
                                : count_skippable<SPECIAL>(p, G, &sS,
                                                           sS.restore ? G.snap_rows + (size_t)local * G.ny
                                                                      : G.rowacc + slot(G, sS.parity, local) * G.ny,
                                                           tid);
        if (tid == 0) {
            sS.in_flight = go ? 1u : 0u;
            s_go = go ? 1 : 0;
            if (go && designated && G.steplog && sS.step < (uint64_t)G.max_log) {
                StepRec q;
                q.dt32 = sS.dt32;
                q.dtdx = sS.dtdx;
                q.dtdy = sS.dtdy;
                q.rain32 = sS.rain32;
                q.cap32 = sS.cap32;
                q.pad = 0.0f;
                q.t_new = sS.t_new;
                q.dt = sS.dt;
                G.steplog[(size_t)local * G.max_log + sS.step] = q;
            }
        }
    }
    __syncthreads();
    if (LI && s_go && !ctl_only) {
        LiK K;
        K.dt = uni(sS.dt);
        K.rx = K.dt / G.dx;
        K.ry = K.dt / G.dy;
        K.rain = (double)uni(sS.rain32);
        K.cap = (double)uni(sS.cap32);
        K.h_dry = (double)p.h_dry;
        K.g = G.g;
        K.dx = G.dx;
        K.dy = G.dy;
        K.dmin = fmin(G.dx, G.dy);
        K.inv_dx = 1.0 / G.dx;
        K.inv_dy = 1.0 / G.dy;
        K.inv_dmin = 1.0 / K.dmin;
        K.p2 = pow2_spacing(G.dx) && pow2_spacing(G.dy);
        const double t_new = uni(sS.t_new);
        const double t_now = uni(sS.t);
        const bool sync = uni(sS.hits_sync) != 0;
        const int par = uni(sS.parity);
        const bool from_snap = uni((int)sS.restore) != 0;
        StepAcc* acc_cur = a.acc + (size_t)(lk & 3) * a.ncand + cand;
        if (MODES == 3) {
            if (G.crest) li_band_step<true, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap, acc_cur, lds, mr);
            else li_band_step<false, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap, acc_cur, lds, mr);
        } else {
            li_band_step<MODES == 2, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap, acc_cur, lds, mr);
        }
        if (sync && G.snap_state && !sS.locked_off) snapshot_band(G, local, tk.z, par ^ 1);
    } else if (!LI && s_go && !ctl_only) {
        // The step's constants are wave-uniform: keep them in scalar registers.
        StepK K;
        K.g = G.g32;
        K.g_half = G.g_half32;
        K.inv_dx = G.inv_dx32;
        K.inv_dy = G.inv_dy32;
        K.h_dry = p.h_dry;
        K.heps2 = p.heps2;
        K.q_min = p.q_min;
        K.dt = uni(sS.dt32);
        K.dtdx = uni(sS.dtdx);
        K.dtdy = uni(sS.dtdy);
        K.rain = uni(sS.rain32);
        K.cap = uni(sS.cap32);
        const double t_new = uni(sS.t_new);
        const double t_now = uni(sS.t);
        const bool sync = uni(sS.hits_sync) != 0;
        const int par = uni(sS.parity);
        const bool from_snap = uni((int)sS.restore) != 0;
        StepAcc* acc_cur = a.acc + (size_t)(lk & 3) * a.ncand + cand;
        const uint64_t step = sS.step;
        const uint32_t sync_idx = sS.sync_index;
        if (MODES == 3) {
            if (G.crest)
                band_step<true, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap, step, sync_idx, acc_cur,
                                            lds, mr, pre);
            else
                band_step<false, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap, step, sync_idx,
                                             acc_cur, lds, mr, pre);
        } else {
            band_step<MODES == 2, SPECIAL>(a, G, local, tk.z, K, t_now, t_new, sync, par, from_snap,
                                                              step, sync_idx, acc_cur, lds, mr, pre);
        }
        if (sync && G.snap_state && !sS.locked_off) snapshot_band(G, local, tk.z, par ^ 1);
    } else if (!designated) {
        return;
    }
    if (!designated) return;
    if (completed && !sS.restore) {   // (a restart discards the completed step)
        if (LI) {
            const LiLds LL = li_lds_layout(G.R, G.nx, G.special != 0);
            li_ledger_epilogue(G, Sin, A, sS, lds, LL.total, shd);
        } else {
            const LdsLayout LL = lds_layout(G.R, G.nx, G.special != 0);
            ledger_epilogue(a, G, Sin, A, sS, lds, LL.total, shd);
        }
    }
    __syncthreads();
    for (int w = tid; w < kCtlWords; w += nthr) ((uint32_t*)Sout)[w] = ((const uint32_t*)&sS)[w];
    // The control part of the sync snapshot, once the sync step is accounted.
    if (completed && Sin->hits_sync && G.snap_ctl && !sS.locked_off && !(sS.flags & F_DONE))
        for (int w = tid; w < kCtlWords; w += nthr) ((uint32_t*)(G.snap_ctl + local))[w] = ((const uint32_t*)&sS)[w];
    if (tid == 0) {
        StepAcc z;
        z.smax_bits = 0;
        z.bad = 0u;
        z.smax64 = 0ull;
        z.dh64 = 0ull;
        z.dq64 = 0ull;
        a.acc[(size_t)((lk + 2) & 3) * a.ncand + cand] = z;
        a.status[cand] = sS.flags;
    }
}

// The launch entry points.  The signed-face kernel (the common case) is held to 80 VGPRs
// (6 waves per SIMD); the crest variants keep their natural budget.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
itr_step_signed(KArgs a, uint32_t lk) { step_body<1, false, false>(a, lk); }
__global__ void __launch_bounds__(1024) itr_step_crest(KArgs a, uint32_t lk) { step_body<2, false, false>(a, lk); }
__global__ void __launch_bounds__(1024) itr_step_mixed(KArgs a, uint32_t lk) { step_body<3, false, false>(a, lk); }
__global__ void __launch_bounds__(1024) itr_step_special(KArgs a, uint32_t lk) { step_body<3, true, false>(a, lk); }
// Local-inertial grids (launch lists of their own).  The signed and crest forms are held to
// 96 VGPRs (5 waves per SIMD): their f64 chains are latency-bound, and at 4 waves the
// step ran 22% slower.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5)))
itr_step_li_signed(KArgs a, uint32_t lk) { step_body<1, false, true>(a, lk); }
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(5)))
itr_step_li_crest(KArgs a, uint32_t lk) { step_body<2, false, true>(a, lk); }
__global__ void __launch_bounds__(1024) itr_step_li_general(KArgs a, uint32_t lk) { step_body<3, true, true>(a, lk); }

// Head of the first step of every candidate: control copy 1 (read by launch 0).
__global__ void itr_init_kernel(KArgs a) {
    const int c = blockIdx.x;
    CandCtl* C = a.snap + (size_t)a.ncand + c;
    const DevGrid& G = a.grids[C->grid];
    if (C->flags & F_DONE) {
        if (threadIdx.x == 0) a.status[c] = C->flags;
        return;
    }
    CandCtl S = *C;
    bool ok = prepare_step(G.prm, G, &S, G.prm.rain_t, G.prm.rain_r);
    if (G.tape_mode == 2 && (!ok || G.tape_steps == 0)) {   // replay-invalid at the first step
        S.abort_kind = 5;
        S.abort_sync = 0;
        S.flags |= F_DONE;
        ok = true;
    }
    if (!ok) {   // the first locked step fails its check: adaptive from t = 0 (solver.rs:456-480)
        S.locked_off = 1;
        S.dt_fallback_t = S.t;
        ok = prepare_step(G.prm, G, &S, G.prm.rain_t, G.prm.rain_r);
    }
    {
        S.skipped_now = G.li ? 0u
                             : count_skippable<true>(G.prm, G, &S, G.rowacc + slot(G, S.parity, S.local) * G.ny, threadIdx.x & 63);
    }
    if (threadIdx.x == 0) {
        *C = S;
        a.status[c] = S.flags;
    }
}

// Exhaustive check of csqrt / crecip against the IEEE operations over every f32 bit
// pattern: one thread per mantissa, looping over sign and exponent.
__global__ void __launch_bounds__(256) itr_math_check_kernel(unsigned long long* bad) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;   // 23-bit mantissa
    if (m >= (1u << 23)) return;
    unsigned long long nr = 0, ns = 0;
    for (uint32_t se = 0; se < 512; se++) {   // sign and biased exponent
        const float x = __uint_as_float((se << 23) | m);
        const float r = crecip(x), r0 = 1.0f / x;
        const float q = csqrt(x), q0 = __builtin_sqrtf(x);
        nr += (__float_as_uint(r) != __float_as_uint(r0) && !(r != r && r0 != r0)) ? 1ull : 0ull;
        ns += (__float_as_uint(q) != __float_as_uint(q0) && !(q != q && q0 != q0)) ? 1ull : 0ull;
    }
    if (nr) atomicAdd(&bad[0], nr);
    if (ns) atomicAdd(&bad[1], ns);
}

// ------------------------------------------------------------------ design materialization
// Candidate edits built on the device from their primitives (itr_design): Δz over each
// candidate's union window (DesignSpace::materialize, design.rs:316-385), the level
// z = 0 + (z_fine + Δz) of open cells (coarse_cell / fine_z at level 1,
// terrain.rs:353-373) and every signed face jump f32(z_R - z_L) (terrain.rs:460-505).
struct DesignDev {
    const double* prim;        // [n][12]
    const int32_t* prim_start; // [ncand + 1]
    const int32_t* bbox;       // [ncand][4] c0, r0, w, h
    const int64_t* dz_off;     // [ncand] offset of the candidate's window in dzbuf
    double* dzbuf;
    const double* z;           // [fny][fnx] fine-grid z
    const uint8_t* wall;       // [fny][fnx]
    const uint8_t* editable;   // [fny][fnx] or nullptr
    double ox, oy, dx, dy, max_dz, z_wall, area;
    int32_t nx, ny, P, ncand;  // the grid (a coarse level when lv >= 2)
    int32_t fnx, fny, lv, pad; // the fine grid the designs are drawn on (the grid itself when lv == 1)
    float* dzx;                // [ncand][plane]
    float* dzy;                // [ncand][plane_y]
    float* dzx2;               // face-crest second rises (lv >= 2)
    float* dzy2;
    double* zc;                // [ncand][ny][nx] coarse level z with the edit (lv >= 2)
    size_t plane, plane_y;
    double* volumes;           // [ncand][2]
};

// Primitive::profile (design.rs:108-134): value in [0, 1], exactly 0 outside the support.
__device__ double prim_profile(const double* pr, double x, double y) {
    const double x0 = pr[1], y0 = pr[2], x1 = pr[3], y1 = pr[4];
    const double vx = x1 - x0, vy = y1 - y0;
    const double l2 = vx * vx + vy * vy;
    double t = 0.0;
    if (l2 > 0.0) {
        t = ((x - x0) * vx + (y - y0) * vy) / l2;
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
    const double px = x0 + t * vx - x, py = y0 + t * vy - y;
    const double d = __builtin_sqrt(px * px + py * py);
    const double s = d / (0.5 * pr[5]);
    if (s >= 1.0) return 0.0;
    if (pr[0] == 0.0) {   // cut: flat bottom, quartic shoulder
        if (s <= 0.5) return 1.0;
        const double u = (s - 0.5) / 0.5;
        const double w = 1.0 - u * u;
        return w * w;
    }
    const double w = 1.0 - s * s;
    return w * w;
}

// Δz of every cell of each candidate's window: 4x4 sub-sample mean of each primitive's
// profile times its height, summed in primitive order (a primitive adds exactly 0 outside
// its own cell window, so it is skipped there), clamped, 0 where not editable.
__global__ void __launch_bounds__(256) itr_materialize_kernel(DesignDev d) {
    const int cand = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = d.bbox[4 * cand], r0 = d.bbox[4 * cand + 1], w = d.bbox[4 * cand + 2], h = d.bbox[4 * cand + 3];
    if (i >= w * h) return;
    const int col = c0 + i % w, row = r0 + i / w;
    double v = 0.0;
    if (!d.editable || d.editable[(size_t)row * d.fnx + col]) {
        double acc = 0.0;
        for (int q = d.prim_start[cand]; q < d.prim_start[cand + 1]; q++) {
            const double* pr = d.prim + (size_t)12 * q;
            if (pr[6] == 0.0) continue;
            if (col < (int)pr[7] || row < (int)pr[8] || col >= (int)pr[9] || row >= (int)pr[10]) continue;
            double sum = 0.0;
            for (int sy = 0; sy < 4; sy++)
                for (int sx = 0; sx < 4; sx++) {
                    const double x = d.ox + ((double)col + ((double)sx + 0.5) / 4.0) * d.dx;
                    const double y = d.oy - ((double)row + ((double)sy + 0.5) / 4.0) * d.dy;
                    sum += prim_profile(pr, x, y);
                }
            acc += pr[6] * sum * (1.0 / 16.0);
        }
        v = acc < -d.max_dz ? -d.max_dz : (acc > d.max_dz ? d.max_dz : acc);
    }
    d.dzbuf[d.dz_off[cand] + i] = v;
}

// Fine cell (c, r) of a candidate: z + its edit (terrain.rs:353-355, fine_z; 0 outside the window).
__device__ __forceinline__ double design_fine_z(const DesignDev& d, int cand, int c, int r) {
    const int c0 = d.bbox[4 * cand], r0 = d.bbox[4 * cand + 1], w = d.bbox[4 * cand + 2], h = d.bbox[4 * cand + 3];
    double e = 0.0;
    if (c >= c0 && c < c0 + w && r >= r0 && r < r0 + h) e = d.dzbuf[d.dz_off[cand] + (size_t)(r - r0) * w + (c - c0)];
    return d.z[(size_t)r * d.fnx + c] + e;
}
// Level z of cell (c, r) of a candidate with its edit (level 1).
__device__ __forceinline__ double design_z(const DesignDev& d, int cand, int c, int r) {
    if (d.wall[(size_t)r * d.fnx + c]) return d.z_wall;
    return 0.0 + design_fine_z(d, cand, c, r);
}

// Every face jump of every candidate (boundary faces stay 0): x-face i of row r joins
// cells i-1 and i; y-face f of column c joins rows f-1 and f.
__global__ void __launch_bounds__(256) itr_design_faces_kernel(DesignDev d) {
    const int cand = blockIdx.z, r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > d.nx) return;
    if (r < d.ny && i >= 1 && i <= d.nx - 1)
        d.dzx[(size_t)cand * d.plane + (size_t)r * d.P + i] = (float)(design_z(d, cand, i, r) - design_z(d, cand, i - 1, r));
    if (r >= 1 && r <= d.ny - 1 && i < d.nx)
        d.dzy[(size_t)cand * d.plane_y + (size_t)r * d.P + i] = (float)(design_z(d, cand, i, r) - design_z(d, cand, i, r - 1));
}

// Coarse level z of every cell of each candidate (lv >= 2): coarse_cell
// (terrain.rs:357-373) over the candidate's fine z, the non-wall cells of the lv x lv block
// added row-major from 0.0 and divided by their count; z_wall when every cell is a wall.
// Outside the edit window this is the base level's z bit for bit (z + 0.0 == z).
__global__ void __launch_bounds__(256) itr_design_coarse_kernel(DesignDev d) {
    const int cand = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx * d.ny) return;
    const int c = i % d.nx, r = i / d.nx, lv = d.lv;
    double s = 0.0;
    int cnt = 0;
    for (int rr = r * lv; rr < min((r + 1) * lv, d.fny); rr++)
        for (int cc = c * lv; cc < min((c + 1) * lv, d.fnx); cc++)
            if (!d.wall[(size_t)rr * d.fnx + cc]) {
                s += design_fine_z(d, cand, cc, rr);
                cnt++;
            }
    d.zc[(size_t)cand * d.nx * d.ny + i] = cnt == 0 ? d.z_wall : s / (double)cnt;
}

// Face-crest rises of every face of each candidate (lv >= 2; terrain.rs:378-429): the
// crest is the highest non-wall fine cell on either side of the face (z_wall when all are
// walls), raised to both coarse levels; the rises crest - z_left and crest - z_right are
// rounded to f32 once.  x-face i of row r joins coarse cells i-1 and i; y-face f of column
// c joins rows f-1 and f.  Boundary faces stay 0.
__global__ void __launch_bounds__(256) itr_design_crest_kernel(DesignDev d) {
    const int cand = blockIdx.z, r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = d.nx, ny = d.ny, lv = d.lv;
    const double* zc = d.zc + (size_t)cand * nx * ny;
    if (r < ny && i >= 1 && i <= nx - 1) {
        const double zl = zc[(size_t)r * nx + i - 1], zr = zc[(size_t)r * nx + i];
        const int cl = i * lv - 1, cr = min(i * lv, d.fnx - 1);
        double cst = -DBL_MAX;
        bool all_wall = true;
        for (int rr = r * lv; rr < min((r + 1) * lv, d.fny); rr++) {
            if (!d.wall[(size_t)rr * d.fnx + cl]) {
                all_wall = false;
                cst = fmax(cst, design_fine_z(d, cand, cl, rr));
            }
            if (!d.wall[(size_t)rr * d.fnx + cr]) {
                all_wall = false;
                cst = fmax(cst, design_fine_z(d, cand, cr, rr));
            }
        }
        if (all_wall) cst = d.z_wall;
        const double zf = fmax(fmax(cst, zl), zr);
        const size_t o = (size_t)cand * d.plane + (size_t)r * d.P + i;
        d.dzx[o] = (float)(zf - zl);
        d.dzx2[o] = (float)(zf - zr);
    }
    if (r >= 1 && r <= ny - 1 && i < nx) {
        const int c = i;
        const double zt = zc[(size_t)(r - 1) * nx + c], zb = zc[(size_t)r * nx + c];
        const int rt = r * lv - 1, rb = min(r * lv, d.fny - 1);
        double cst = -DBL_MAX;
        bool all_wall = true;
        for (int cc = c * lv; cc < min((c + 1) * lv, d.fnx); cc++) {
            if (!d.wall[(size_t)rt * d.fnx + cc]) {
                all_wall = false;
                cst = fmax(cst, design_fine_z(d, cand, cc, rt));
            }
            if (!d.wall[(size_t)rb * d.fnx + cc]) {
                all_wall = false;
                cst = fmax(cst, design_fine_z(d, cand, cc, rb));
            }
        }
        if (all_wall) cst = d.z_wall;
        const double zf = fmax(fmax(cst, zt), zb);
        const size_t o = (size_t)cand * d.plane_y + (size_t)r * d.P + c;
        d.dzy[o] = (float)(zf - zt);
        d.dzy2[o] = (float)(zf - zb);
    }
}

// Cut and fill volumes of each edit: row-major over the window, as design.rs:373-380 sums.
// One workgroup per candidate: the window's dz values are staged through LDS with
// coalesced loads, then one thread adds them in the reference's (row-major) order.
constexpr int kVolChunk = 4096;
__global__ void __launch_bounds__(256) itr_design_volumes_kernel(DesignDev d) {
    __shared__ double buf[kVolChunk];
    const int cand = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t n = (int64_t)d.bbox[4 * cand + 2] * d.bbox[4 * cand + 3];
    const double* v = d.dzbuf + d.dz_off[cand];
    const double area = d.area;
    double cut = 0.0, fill = 0.0;
    for (int64_t b = 0; b < n; b += kVolChunk) {
        const int cnt = (int)min((int64_t)kVolChunk, n - b);
        for (int i = tid; i < cnt; i += blockDim.x) buf[i] = v[b + i];
        __syncthreads();
        if (tid == 0) {
#pragma unroll 8
            for (int i = 0; i < cnt; i++) {
                const double x = buf[i];
                if (x > 0.0) fill += x * area;
                else cut -= x * area;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        d.volumes[2 * cand] = cut;
        d.volumes[2 * cand + 1] = fill;
    }
}

// ------------------------------------------------------------------ discrete adjoint
// Per-cell kernels of the reverse sweep (one thread per cell, grid x = column blocks,
// y = rows, z = candidates).  The forward replay recomputes a step cell by cell with
// the same IEEE operations as the band kernel (every face value is a function of its
// two cells only), so the replayed states are bitwise the forward's; the built-in
// check compares each replayed segment end with the stored checkpoint.
struct AdjArgs {
    const float* U;        // old state [cand][3][plane] (h, qx, qy)
    float* Uo;             // replay: new state
    float* pre;            // replay writes / sources adjoint reads: (h1, qx1, qy1) after the flux divergence
    const float* lam;      // adjoint of the new state [cand][3][plane]
    float* lam1;           // adjoint after the sources, before the flux divergence
    float* lamo;           // adjoint of the old state
    double* gx;            // [cand][ny][P] dJ/d(dzx) (west face of each cell)
    double* gy;            // [cand][ny+1][P] dJ/d(dzy) (north face of each cell)
    const float* dzx;      // [cand][ny][P] signed jump, or the left rise d_l of a crest face
    const float* dzy;      // [cand][ny+1][P]
    const float* dzx2;     // crest levels: [cand][ny][P] the right rise d_r
    const float* dzy2;     // crest levels: [cand][ny+1][P]
    double* gx2;           // crest levels: dJ/d(d_r) of the west faces
    double* gy2;           // crest levels: dJ/d(d_r) of the north faces
    int32_t crest;         // face-crest level (numerics.rs:117-129)
    int32_t special;       // the grid has Stage / Inflow ghosts
    const DevGrid* G;      // the grid's descriptor (its ghost data), for special ghosts
    double t;              // the step's start time (Inflow discharges)
    double* gz;            // [cand][2nx + 2ny] dJ/d(level z of the boundary cell) through its Stage ghost
    const float* wet;      // [ny][P]
    const float* gn2;      // [ny][P]
    const int32_t* ghost;  // [2nx + 2ny]
    int32_t nx, ny, P, R;    // R: rows per band of the reverse march
    int32_t waves;           // waves per row (62 columns each)
    int32_t tail_groups;     // row groups of the last wave (as the forward step; 1: off)
    size_t plane, cstride; // plane = ny*P, cstride = 3*plane
    size_t dzx_stride, dzy_stride;
    StepK K;
};

struct Nb { float h, qx, qy; };

// Old-state value of cell (r, c) or of the ghost next to the interior (boundary.rs:12-87:
// wall mirrors the normal discharge, transmissive copies); never a corner here.
// A special ghost (boundary.rs:43-47, 66-70) as the forward step builds it from the
// interior neighbour v: Stage holds the candidate's stage depth f32(max(stage - z, 0)) with
// the interior discharges; Inflow the larger of the interior depth and the critical depth
// of the hydrograph's discharge at the step's start, with that discharge normal to the
// boundary.  k: ghost index (north[nx], south[nx], west[ny], east[ny]).
__device__ __forceinline__ Nb special_nb(const AdjArgs& a, int cand, int k, int code, Nb v) {
    const bool xdir = k >= 2 * a.nx;
    const bool inward = k < a.nx || (xdir && k < 2 * a.nx + a.ny);   // north or west
    const GhostSp g = special_ghost(*a.G, a.G->prm.inflow_scale, code, k, cand, a.t, inward, 0);
    if (code == 2) {
        v.h = g.hv;
        return v;
    }
    v.h = inflow_depth(v.h, g.hv);
    v.qx = xdir ? g.qv : 0.0f;
    v.qy = xdir ? 0.0f : g.qv;
    return v;
}

// Old-state value of cell (r, c) or of the ghost next to the interior (boundary.rs:12-87:
// wall mirrors the normal discharge, transmissive copies, Stage / Inflow as special_nb);
// never a corner here.  *code / *kg: the ghost's code and index (-1 for an interior cell).
__device__ __forceinline__ Nb nb_at(const float* U, const AdjArgs& a, int cand, int r, int c, int* code, int* kg) {
    int rr = r, cc = c, cd = -1, k = -1;
    if (c < 0 || c >= a.nx) {
        cc = c < 0 ? 0 : a.nx - 1;
        k = 2 * a.nx + (c < 0 ? 0 : a.ny) + r;
        cd = a.ghost[k];
    } else if (r < 0 || r >= a.ny) {
        rr = r < 0 ? 0 : a.ny - 1;
        k = (r < 0 ? 0 : a.nx) + c;
        cd = a.ghost[k];
    }
    const size_t o = (size_t)rr * a.P + cc;
    Nb v{U[o], U[a.plane + o], U[2 * a.plane + o]};
    if (cd == 0 && k >= 2 * a.nx) v.qx = -v.qx;
    if (cd == 0 && k >= 0 && k < 2 * a.nx) v.qy = -v.qy;
    if (cd >= 2) v = special_nb(a, cand, k, cd, v);
    if (code) *code = cd;
    if (kg) *kg = k;
    return v;
}

__device__ __forceinline__ adj::Consts<float> adj_consts(const StepK& K) {
    adj::Consts<float> k;
    k.g = K.g;
    k.g_half = K.g_half;
    k.heps2 = K.heps2;
    k.h_dry = K.h_dry;
    k.q_min = K.q_min;
    k.dt = K.dt;
    k.dtdx = K.dtdx;
    k.dtdy = K.dtdy;
    k.rain = K.rain;
    k.cap = K.cap;
    return k;
}

struct Stencil {
    Nb C, W, E, N, S;
    int cW, cE, cN, cS;   // ghost codes of the neighbours (-1: interior)
    int kW, kE, kN, kS;   // their ghost indices
    Vel vC, vW, vE, vN, vS;
    float dzW, dzE, dzN, dzS, wet, gn2;
    float dzW2, dzE2, dzN2, dzS2;   // crest levels: the right rises
};

__device__ __forceinline__ Stencil stencil(const AdjArgs& a, const float* U, int cand, int r, int c) {
    Stencil s;
    s.C = nb_at(U, a, cand, r, c, nullptr, nullptr);
    s.W = nb_at(U, a, cand, r, c - 1, &s.cW, &s.kW);
    s.E = nb_at(U, a, cand, r, c + 1, &s.cE, &s.kE);
    s.N = nb_at(U, a, cand, r - 1, c, &s.cN, &s.kN);
    s.S = nb_at(U, a, cand, r + 1, c, &s.cS, &s.kS);
    s.vC = cell_velocity(s.C.h, s.C.qx, s.C.qy, a.K);
    s.vW = cell_velocity(s.W.h, s.W.qx, s.W.qy, a.K);
    s.vE = cell_velocity(s.E.h, s.E.qx, s.E.qy, a.K);
    s.vN = cell_velocity(s.N.h, s.N.qx, s.N.qy, a.K);
    s.vS = cell_velocity(s.S.h, s.S.qx, s.S.qy, a.K);
    const float* dzx = a.dzx + (size_t)cand * a.dzx_stride;
    const float* dzy = a.dzy + (size_t)cand * a.dzy_stride;
    s.dzW = dzx[(size_t)r * a.P + c];
    s.dzE = dzx[(size_t)r * a.P + c + 1];
    s.dzN = dzy[(size_t)r * a.P + c];
    s.dzS = dzy[(size_t)(r + 1) * a.P + c];
    s.dzW2 = s.dzE2 = s.dzN2 = s.dzS2 = 0.0f;
    if (a.crest) {
        const float* dzx2 = a.dzx2 + (size_t)cand * a.dzx_stride;
        const float* dzy2 = a.dzy2 + (size_t)cand * a.dzy_stride;
        s.dzW2 = dzx2[(size_t)r * a.P + c];
        s.dzE2 = dzx2[(size_t)r * a.P + c + 1];
        s.dzN2 = dzy2[(size_t)r * a.P + c];
        s.dzS2 = dzy2[(size_t)(r + 1) * a.P + c];
    }
    s.wet = a.wet[(size_t)r * a.P + c];
    s.gn2 = a.gn2[(size_t)r * a.P + c];
    return s;
}

// The adjoint of a ghost (bh: of its depth; q: of its discharges) into the boundary cell:
// wall and transmissive ghosts mirror or copy the cell; a Stage ghost copies the
// discharges and its depth f32(max(stage - z, 0)) moves with the cell's level z (-1 while
// positive), which goes to gz[k]; an Inflow ghost's depth is the cell's own where that
// exceeds the critical depth (hc: the ghost's depth; hi: the cell's), its discharges are
// the hydrograph's.
__device__ __forceinline__ void fold_ghost(const AdjArgs& a, int cand, int code, int kg, bool xdir, float hg, float hi,
                                           float bh, const adj::S3<float>& q, adj::S3<float>& bC) {
    if (code == 2) {
        if (hg > 0.0f && bh != 0.0f) a.gz[(size_t)cand * (2 * a.nx + 2 * a.ny) + kg] -= (double)bh;
        bC.qx += q.qx;
        bC.qy += q.qy;
    } else if (code == 3) {
        if (hg == hi) bC.h += bh;
    } else {
        const bool wall = code == 0;
        bC.h += bh;
        bC.qx += (xdir && wall) ? -q.qx : q.qx;
        bC.qy += (!xdir && wall) ? -q.qy : q.qy;
    }
}

// Forward replay of one step (kernel.rs:590-727 semantics, bitwise).
__global__ void __launch_bounds__(256) itr_replay_kernel(AdjArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, cand = blockIdx.z;
    if (c >= a.nx) return;
    const float* U = a.U + (size_t)cand * a.cstride;
    const Stencil s = stencil(a, U, cand, r, c);
    Flux fw, fe, fn, fs;
    if (a.crest) {
        fw = face_crest<true>(s.W.h, s.vW, s.C.h, s.vC, s.dzW, s.dzW2, a.K);
        fe = face_crest<true>(s.C.h, s.vC, s.E.h, s.vE, s.dzE, s.dzE2, a.K);
        fn = face_crest<false>(s.N.h, s.vN, s.C.h, s.vC, s.dzN, s.dzN2, a.K);
        fs = face_crest<false>(s.C.h, s.vC, s.S.h, s.vS, s.dzS, s.dzS2, a.K);
    } else {
        fw = face_signed<true>(s.W.h, s.vW, s.C.h, s.vC, s.dzW, a.K);
        fe = face_signed<true>(s.C.h, s.vC, s.E.h, s.vE, s.dzE, a.K);
        fn = face_signed<false>(s.N.h, s.vN, s.C.h, s.vC, s.dzN, a.K);
        fs = face_signed<false>(s.C.h, s.vC, s.S.h, s.vS, s.dzS, a.K);
    }
    const CellNew u = cell_update(s.C.h, s.C.qx, s.C.qy, fw.m, fw.nr, fw.t, fe.m, fe.nl, fe.t, fn.m, fn.nr, fn.t, fs.m,
                                  fs.nl, fs.t, s.wet, s.gn2, a.K);
    float* Uo = a.Uo + (size_t)cand * a.cstride;
    const size_t o = (size_t)r * a.P + c;
    Uo[o] = u.h;
    Uo[a.plane + o] = u.qx;
    Uo[2 * a.plane + o] = u.qy;
    // The state after the flux divergence (numerics.rs:148-153, as cell_update forms it),
    // kept for the reverse of the sources.
    const StepK& K = a.K;
    float* Pr = a.pre + (size_t)cand * a.cstride;
    Pr[o] = s.C.h - K.dtdx * (fe.m - fw.m) - K.dtdy * (fs.m - fn.m);
    Pr[a.plane + o] = s.C.qx - K.dtdx * (fe.nl - fw.nr) - K.dtdy * (fs.t - fn.t);
    Pr[2 * a.plane + o] = s.C.qy - K.dtdx * (fe.t - fw.t) - K.dtdy * (fs.nl - fn.nr);
}

// Reverse of the sources (numerics.rs:158-195): lam1 = (d new / d (h1, qx1, qy1))^T lam,
// at the post-divergence state the replay kept.
__global__ void __launch_bounds__(256) itr_adj_sources_kernel(AdjArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, cand = blockIdx.z;
    if (c >= a.nx) return;
    const size_t o = (size_t)cand * a.cstride + (size_t)r * a.P + c;
    const size_t oc = (size_t)r * a.P + c;
    const float h1 = a.pre[o], qx1 = a.pre[a.plane + o], qy1 = a.pre[2 * a.plane + o];
    const adj::S3<float> b{a.lam[o], a.lam[a.plane + o], a.lam[2 * a.plane + o]};
    const adj::S3<float> l1 = adj::sources_adj(h1, qx1, qy1, a.wet[oc], a.gn2[oc], adj_consts(a.K), b);
    a.lam1[o] = l1.h;
    a.lam1[a.plane + o] = l1.qx;
    a.lam1[2 * a.plane + o] = l1.qy;
}

// Reverse of the flux divergence, the four faces and the cell velocities: the adjoint
// of the old state, and the west / north face jump gradients (each interior face is
// owned by the cell to its east / south; boundary faces have fixed zero jumps).
__global__ void __launch_bounds__(256) itr_adj_faces_kernel(AdjArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, cand = blockIdx.z;
    if (c >= a.nx) return;
    const float* U = a.U + (size_t)cand * a.cstride;
    const Stencil s = stencil(a, U, cand, r, c);
    const float* L1 = a.lam1 + (size_t)cand * a.cstride;
    auto lam1_at = [&](int rr, int cc) -> adj::S3<float> {
        if (rr < 0 || rr >= a.ny || cc < 0 || cc >= a.nx) return adj::S3<float>{0.0f, 0.0f, 0.0f};   // ghosts are not updated
        const size_t o = (size_t)rr * a.P + cc;
        return adj::S3<float>{L1[o], L1[a.plane + o], L1[2 * a.plane + o]};
    };
    const adj::S3<float> lC = lam1_at(r, c), lW = lam1_at(r, c - 1), lE = lam1_at(r, c + 1), lN = lam1_at(r - 1, c),
                         lS = lam1_at(r + 1, c);
    const adj::Consts<float> k = adj_consts(a.K);
    const float dtdx = a.K.dtdx, dtdy = a.K.dtdy;
    auto V3 = [](const Vel& v) { return adj::V<float>{v.u, v.v, v.c}; };
    adj::S3<float> bC = lC;   // h1 = h - ..., qx1 = qx - ..., qy1 = qy - ...
    adj::V<float> bv{0.0f, 0.0f, 0.0f};
    // A ghost's adjoint folds back into this cell (its mirror image), or, for a Stage
    // ghost's depth, into the boundary cell's level.
    auto fold = [&](const Nb& g, int code, int kg, bool xdir, float bh, const adj::V<float>& gv) {
        adj::S3<float> q = adj::vel_adj(g.h, g.qx, g.qy, k, gv);
        fold_ghost(a, cand, code, kg, xdir, g.h, s.C.h, bh + q.h, q, bC);
    };
    double* gx = a.gx + (size_t)cand * a.dzx_stride;
    double* gy = a.gy + (size_t)cand * a.dzy_stride;
    double* gx2 = a.crest ? a.gx2 + (size_t)cand * a.dzx_stride : nullptr;
    double* gy2 = a.crest ? a.gy2 + (size_t)cand * a.dzy_stride : nullptr;
    // the face's adjoint: signed jump, or the two rises of a crest face
    auto fadj_x = [&](const Nb& L, const Vel& vl, const Nb& R, const Vel& vr, float d, float d2, const adj::F<float>& b) {
        return a.crest ? adj::face_cr_adj<true>(L.h, V3(vl), R.h, V3(vr), d, d2, k, b)
                       : adj::face_adj<true>(L.h, V3(vl), R.h, V3(vr), d, k, b);
    };
    auto fadj_y = [&](const Nb& L, const Vel& vl, const Nb& R, const Vel& vr, float d, float d2, const adj::F<float>& b) {
        return a.crest ? adj::face_cr_adj<false>(L.h, V3(vl), R.h, V3(vr), d, d2, k, b)
                       : adj::face_adj<false>(L.h, V3(vl), R.h, V3(vr), d, k, b);
    };
    {   // west face: L = W, R = C
        const adj::F<float> b{dtdx * (lC.h - lW.h), -dtdx * lW.qx, dtdx * lC.qx, dtdx * (lC.qy - lW.qy)};
        const adj::FaceBar<float> f = fadj_x(s.W, s.vW, s.C, s.vC, s.dzW, s.dzW2, b);
        bC.h += f.hr;
        bv.u += f.ur; bv.v += f.vr; bv.c += f.cr;
        if (c == 0) fold(s.W, s.cW, s.kW, true, f.hl, adj::V<float>{f.ul, f.vl, f.cl});
        else {
            if (f.dz != 0.0f) gx[(size_t)r * a.P + c] += (double)f.dz;
            if (gx2 && f.dz2 != 0.0f) gx2[(size_t)r * a.P + c] += (double)f.dz2;
        }
    }
    {   // east face: L = C, R = E
        const adj::F<float> b{dtdx * (lE.h - lC.h), -dtdx * lC.qx, dtdx * lE.qx, dtdx * (lE.qy - lC.qy)};
        const adj::FaceBar<float> f = fadj_x(s.C, s.vC, s.E, s.vE, s.dzE, s.dzE2, b);
        bC.h += f.hl;
        bv.u += f.ul; bv.v += f.vl; bv.c += f.cl;
        if (c == a.nx - 1) fold(s.E, s.cE, s.kE, true, f.hr, adj::V<float>{f.ur, f.vr, f.cr});
    }
    {   // north face: L = N, R = C (normal velocity v)
        const adj::F<float> b{dtdy * (lC.h - lN.h), -dtdy * lN.qy, dtdy * lC.qy, dtdy * (lC.qx - lN.qx)};
        const adj::FaceBar<float> f = fadj_y(s.N, s.vN, s.C, s.vC, s.dzN, s.dzN2, b);
        bC.h += f.hr;
        bv.u += f.ur; bv.v += f.vr; bv.c += f.cr;
        if (r == 0) fold(s.N, s.cN, s.kN, false, f.hl, adj::V<float>{f.ul, f.vl, f.cl});
        else {
            if (f.dz != 0.0f) gy[(size_t)r * a.P + c] += (double)f.dz;
            if (gy2 && f.dz2 != 0.0f) gy2[(size_t)r * a.P + c] += (double)f.dz2;
        }
    }
    {   // south face: L = C, R = S
        const adj::F<float> b{dtdy * (lS.h - lC.h), -dtdy * lC.qy, dtdy * lS.qy, dtdy * (lS.qx - lC.qx)};
        const adj::FaceBar<float> f = fadj_y(s.C, s.vC, s.S, s.vS, s.dzS, s.dzS2, b);
        bC.h += f.hl;
        bv.u += f.ul; bv.v += f.vl; bv.c += f.cl;
        if (r == a.ny - 1) fold(s.S, s.cS, s.kS, false, f.hr, adj::V<float>{f.ur, f.vr, f.cr});
    }
    const adj::S3<float> q = adj::vel_adj(s.C.h, s.C.qx, s.C.qy, k, bv);
    const size_t o = (size_t)cand * a.cstride + (size_t)r * a.P + c;
    a.lamo[o] = bC.h + q.h;
    a.lamo[a.plane + o] = bC.qx + q.qx;
    a.lamo[2 * a.plane + o] = bC.qy + q.qy;
}

// Reverse of one step for a band of R rows of one candidate, marched like the forward
// band (lanes are columns with a halo lane on each side, neighbours by DPP).  Per row:
// the reverse of the sources, one velocity, the west face (whose
// left-side bars reach the left lane by DPP as its east face) and the south face
// (carried to the next row as its north face); each face is differentiated once where
// itr_adj_faces_kernel does it twice.  Each cell adds its faces' contributions in that
// kernel's order (west, east, north, south, ghost folds included), so lamo is bitwise
// the same.  Each interior face's jump gradient is written by the lane that owns its
// east / south cell.  A row's loads are issued one row ahead.  A zero face adjoint is not
// added: the f64 sums start at +0 and adding ±0 leaves any value but −0 unchanged, so
// skipping it changes no bit and saves the read-modify-write where λ has not reached.
struct NbW { float h, qx, qy, code; };   // code: the ghost code of a ghost cell, -1 in the grid
struct RevIn { float h, qx, qy, lh, lqx, lqy, ph, pqx, pqy, wet, gn2, dzs, dzw, dzs2, dzw2; int code; };
__global__ void __launch_bounds__(1024) itr_adj_band_kernel(AdjArgs a) {
    const int band = blockIdx.x, cand = blockIdx.y;
    const int nx = a.nx, ny = a.ny;
    const int a0 = band * a.R, b0 = min(a0 + a.R, ny);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // The last wave of a row that updates few columns marches tail_groups row groups of the
    // band side by side (lane groups with their own halo lanes), as the forward step does:
    // group g covers rows a0 + g*t_rows .. of the band; lr is its offset.
    const int S = a.tail_groups;
    const bool tail = S > 1 && uni(wave) == a.waves - 1;
    const int tgw = tail ? 64 / S : 64;
    const int tg = tail ? lane / tgw : 0;
    const int gl = lane - tg * tgw;
    const int t_rows = tail ? (b0 - a0 + S - 1) / S : b0 - a0;
    const int lr = tg * t_rows;
    const int col = wave * kColsPerWave - 1 + gl;
    const bool upd = gl >= 1 && gl <= tgw - 2 && tg < S && col < nx;
    const bool cin = col >= 0 && col < nx;
    const int cc = min(max(col, 0), nx - 1), xc = min(max(col, 0), nx);
    const float* U = a.U + (size_t)cand * a.cstride;
    const float* LAM = a.lam + (size_t)cand * a.cstride;
    const float* PRE = a.pre + (size_t)cand * a.cstride;
    float* LO = a.lamo + (size_t)cand * a.cstride;
    const float* dzx = a.dzx + (size_t)cand * a.dzx_stride;
    const float* dzy = a.dzy + (size_t)cand * a.dzy_stride;
    double* gx = a.gx + (size_t)cand * a.dzx_stride;
    double* gy = a.gy + (size_t)cand * a.dzy_stride;
    const bool crest = a.crest != 0;
    const float* dzx2 = crest ? a.dzx2 + (size_t)cand * a.dzx_stride : nullptr;
    const float* dzy2 = crest ? a.dzy2 + (size_t)cand * a.dzy_stride : nullptr;
    double* gx2 = crest ? a.gx2 + (size_t)cand * a.dzx_stride : nullptr;
    double* gy2 = crest ? a.gy2 + (size_t)cand * a.dzy_stride : nullptr;
    const adj::Consts<float> k = adj_consts(a.K);
    const float dtdx = a.K.dtdx, dtdy = a.K.dtdy;
    // Raw inputs of row k1: its old state (the interior neighbour for a ghost), the adjoint
    // of its new state and its post-divergence state (for the reverse of the sources), and
    // the jumps of its north face and of its west face.
    auto load = [&](int k1) -> RevIn {
        RevIn v;
        const int rr = min(max(k1, 0), ny - 1);
        const size_t o = (size_t)rr * a.P + cc;
        v.h = U[o];
        v.qx = U[a.plane + o];
        v.qy = U[2 * a.plane + o];
        const bool rin = k1 >= 0 && k1 < ny;
        v.code = 1;
        if (rin && !cin) v.code = a.ghost[2 * nx + (col < 0 ? 0 : ny) + k1];
        else if (!rin && cin) v.code = a.ghost[(k1 < 0 ? 0 : nx) + col];
        v.lh = LAM[o];
        v.lqx = LAM[a.plane + o];
        v.lqy = LAM[2 * a.plane + o];
        v.ph = PRE[o];
        v.pqx = PRE[a.plane + o];
        v.pqy = PRE[2 * a.plane + o];
        const size_t ow = (size_t)rr * a.P + cc;
        v.wet = a.wet[ow];
        v.gn2 = a.gn2[ow];
        v.dzs = dzy[(size_t)min(max(k1, 0), ny) * a.P + cc];
        v.dzw = dzx[(size_t)min(max(k1, 0), ny - 1) * a.P + xc];
        v.dzs2 = crest ? dzy2[(size_t)min(max(k1, 0), ny) * a.P + cc] : 0.0f;
        v.dzw2 = crest ? dzx2[(size_t)min(max(k1, 0), ny - 1) * a.P + xc] : 0.0f;
        return v;
    };
    // The ghost rule (boundary.rs:12-87) on a loaded row: wall mirrors the normal
    // discharge, transmissive copies, Stage / Inflow as special_nb; corners are 0.
    auto cell = [&](const RevIn& v, int k1) -> NbW {
        const bool rin = k1 >= 0 && k1 < ny;
        if (!rin && !cin) return NbW{0.0f, 0.0f, 0.0f, -1.0f};
        const bool ghost = !rin || !cin;
        const bool wall = ghost && v.code == 0;
        NbW w{v.h, v.qx, v.qy, ghost ? (float)v.code : -1.0f};
        if (!cin && wall) w.qx = -w.qx;
        if (!rin && wall) w.qy = -w.qy;
        if (ghost && v.code >= 2) {
            const int kg = !cin ? 2 * nx + (col < 0 ? 0 : ny) + k1 : (k1 < 0 ? 0 : nx) + col;
            const Nb g = special_nb(a, cand, kg, v.code, Nb{w.h, w.qx, w.qy});
            w.h = g.h;
            w.qx = g.qx;
            w.qy = g.qy;
        }
        return w;
    };
    auto velo = [&](const NbW& c) {
        const Vel w = cell_velocity(c.h, c.qx, c.qy, a.K);
        return adj::V<float>{w.u, w.v, w.c};
    };
    // Adjoint after the reverse of the sources (lam1); 0 outside the grid (ghosts are not
    // updated).
    auto lam1 = [&](const RevIn& v, int k1) -> adj::S3<float> {
        if (!(k1 >= 0 && k1 < ny && cin)) return adj::S3<float>{0.0f, 0.0f, 0.0f};
        return adj::sources_adj(v.ph, v.pqx, v.pqy, v.wet, v.gn2, k, adj::S3<float>{v.lh, v.lqx, v.lqy});
    };
    auto yface = [&](const NbW& L, const adj::V<float>& vl, const adj::S3<float>& ll, const NbW& R,
                     const adj::V<float>& vr, const adj::S3<float>& lr, float dz, float dz2) {
        const adj::F<float> b{dtdy * (lr.h - ll.h), -dtdy * ll.qy, dtdy * lr.qy, dtdy * (lr.qx - ll.qx)};
        return crest ? adj::face_cr_adj<false>(L.h, vl, R.h, vr, dz, dz2, k, b) : adj::face_adj<false>(L.h, vl, R.h, vr, dz, k, b);
    };
    // Rows where λ has not reached: with λ1 of a cell and of its four neighbours exactly 0,
    // every bar of its faces and of its update is a product of zeros with finite values,
    // so its adjoint and its faces' contributions are 0.  A wave whose three rows hold no
    // non-zero λ1 writes zeros for the row and carries zero bars (exact: only the sign of
    // a zero adjoint can differ, and zero face contributions are not added).
    auto zero3 = [](const adj::S3<float>& l) { return l.h == 0.0f && l.qx == 0.0f && l.qy == 0.0f; };
    const int g0 = a0 + lr;   // this lane's first row
    const RevIn ip = load(g0 - 1), ic = load(g0);
    RevIn nxt = load(g0 + 1);
    NbW Pn = cell(ip, g0 - 1), Cn = cell(ic, g0);
    adj::V<float> vC = velo(Cn);
    adj::S3<float> lC = lam1(ic, g0);
    bool zN = zero3(lam1(ip, g0 - 1)), zC = zero3(lC);
    float dzw = ic.dzw, dzw2 = ic.dzw2;   // west face jump (rises) of the current row
    adj::FaceBar<float> fN =
        yface(Pn, velo(Pn), lam1(ip, g0 - 1), Cn, vC, lC, ic.dzs, ic.dzs2);
    // (a group's north face inside the band is the previous group's south face, counted there)
    if (upd && a0 >= 1 && lr == 0) {
        if (fN.dz != 0.0f) gy[(size_t)a0 * a.P + col] += (double)fN.dz;
        if (crest && fN.dz2 != 0.0f) gy2[(size_t)a0 * a.P + col] += (double)fN.dz2;
    }
    for (int jb = a0; jb < a0 + t_rows; jb++) {
        const int j = jb + lr;   // this lane's row
        const bool rv = j < b0;
        const RevIn in = nxt;   // row j+1
        if (j + 1 < b0) nxt = load(j + 2);
        const NbW Sn = cell(in, j + 1);
        const adj::V<float> vS = velo(Sn);
        const bool raw0 = in.lh == 0.0f && in.lqx == 0.0f && in.lqy == 0.0f;
        const adj::S3<float> lS = raw0 ? adj::S3<float>{0.0f, 0.0f, 0.0f} : lam1(in, j + 1);
        const bool zS = zero3(lS);
        if (__builtin_amdgcn_ballot_w64(!(zN && zC && zS)) == 0ull) {   // (wave-uniform)
            if (upd && rv) {
                const size_t o = (size_t)j * a.P + col;
                LO[o] = 0.0f;
                LO[a.plane + o] = 0.0f;
                LO[2 * a.plane + o] = 0.0f;
            }
            dzw = in.dzw;
            dzw2 = in.dzw2;
            Pn = Cn;
            Cn = Sn;
            vC = vS;
            lC = lS;
            fN = adj::FaceBar<float>{};
            zN = zC;
            zC = zS;
            continue;
        }
        const adj::FaceBar<float> fS = yface(Cn, vC, lC, Sn, vS, lS, in.dzs, in.dzs2);
        if (upd && j + 1 <= ny - 1 && j + 1 < b0) {
            if (fS.dz != 0.0f) gy[(size_t)(j + 1) * a.P + col] += (double)fS.dz;
            if (crest && fS.dz2 != 0.0f) gy2[(size_t)(j + 1) * a.P + col] += (double)fS.dz2;
        }
        // West face: L = the left lane's cell, R = this lane's.
        const NbW Wn{from_left(Cn.h), from_left(Cn.qx), from_left(Cn.qy), from_left(Cn.code)};
        const adj::V<float> vW{from_left(vC.u), from_left(vC.v), from_left(vC.c)};
        const adj::S3<float> lW{from_left(lC.h), from_left(lC.qx), from_left(lC.qy)};
        const adj::F<float> bw{dtdx * (lC.h - lW.h), -dtdx * lW.qx, dtdx * lC.qx, dtdx * (lC.qy - lW.qy)};
        const adj::FaceBar<float> fW = crest ? adj::face_cr_adj<true>(Wn.h, vW, Cn.h, vC, dzw, dzw2, k, bw)
                                             : adj::face_adj<true>(Wn.h, vW, Cn.h, vC, dzw, k, bw);
        if (upd && rv && col >= 1 && col <= nx - 1) {
            if (fW.dz != 0.0f) gx[(size_t)j * a.P + col] += (double)fW.dz;
            if (crest && fW.dz2 != 0.0f) gx2[(size_t)j * a.P + col] += (double)fW.dz2;
        }
        dzw = in.dzw;   // row j+1's west face
        dzw2 = in.dzw2;
        // East face: the right lane's west face.
        adj::FaceBar<float> fE;
        fE.hl = from_right(fW.hl); fE.ul = from_right(fW.ul); fE.vl = from_right(fW.vl); fE.cl = from_right(fW.cl);
        fE.hr = from_right(fW.hr); fE.ur = from_right(fW.ur); fE.vr = from_right(fW.vr); fE.cr = from_right(fW.cr);
        const NbW En{from_right(Cn.h), from_right(Cn.qx), from_right(Cn.qy), from_right(Cn.code)};
        adj::S3<float> bC = lC;
        adj::V<float> bv{0.0f, 0.0f, 0.0f};
        auto fold = [&](const NbW& g, int kg, bool xdir, float bh, const adj::V<float>& gv) {
            const adj::S3<float> q = adj::vel_adj(g.h, g.qx, g.qy, k, gv);
            fold_ghost(a, cand, (int)g.code, kg, xdir, g.h, Cn.h, bh + q.h, q, bC);
        };
        bC.h += fW.hr;
        bv.u += fW.ur; bv.v += fW.vr; bv.c += fW.cr;
        if (col == 0) fold(Wn, 2 * nx + j, true, fW.hl, adj::V<float>{fW.ul, fW.vl, fW.cl});
        bC.h += fE.hl;
        bv.u += fE.ul; bv.v += fE.vl; bv.c += fE.cl;
        if (col == nx - 1) fold(En, 2 * nx + ny + j, true, fE.hr, adj::V<float>{fE.ur, fE.vr, fE.cr});
        bC.h += fN.hr;
        bv.u += fN.ur; bv.v += fN.vr; bv.c += fN.cr;
        if (j == 0) fold(Pn, col, false, fN.hl, adj::V<float>{fN.ul, fN.vl, fN.cl});
        bC.h += fS.hl;
        bv.u += fS.ul; bv.v += fS.vl; bv.c += fS.cl;
        if (j == ny - 1) fold(Sn, nx + col, false, fS.hr, adj::V<float>{fS.ur, fS.vr, fS.cr});
        const adj::S3<float> q = adj::vel_adj(Cn.h, Cn.qx, Cn.qy, k, bv);
        if (upd && rv) {
            const size_t o = (size_t)j * a.P + col;
            LO[o] = bC.h + q.h;
            LO[a.plane + o] = bC.qx + q.qx;
            LO[2 * a.plane + o] = bC.qy + q.qy;
        }
        Pn = Cn;
        Cn = Sn;
        vC = vS;
        lC = lS;
        fN = fS;
        zN = zC;
        zC = zS;
    }
}

// Forward replay of one step for a band of R rows of one candidate (large batches),
// marched like the forward band: lanes are columns with a halo lane on each side,
// neighbours by DPP; each face is evaluated once (the west face's values reach the left
// lane by DPP as its east face, the south face carries to the next row as its north
// face) and each cell velocity once, where itr_replay_kernel evaluates four faces and
// five velocities per cell.  Every value is the same IEEE expression of the same inputs
// (ghosts by the same rule, corners unused), so both give the same bits; the segment
// check against the checkpoints holds either to it.
struct FwdIn { float h, qx, qy, wet, gn2, dzs, dzw, dzs2, dzw2; int code; };
__global__ void __launch_bounds__(1024) itr_replay_band_kernel(AdjArgs a) {
    const int band = blockIdx.x, cand = blockIdx.y;
    const int nx = a.nx, ny = a.ny;
    const int a0 = band * a.R, b0 = min(a0 + a.R, ny);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int S = a.tail_groups;   // the last wave's row groups, as the forward step
    const bool tail = S > 1 && uni(wave) == a.waves - 1;
    const int tgw = tail ? 64 / S : 64;
    const int tg = tail ? lane / tgw : 0;
    const int gl = lane - tg * tgw;
    const int t_rows = tail ? (b0 - a0 + S - 1) / S : b0 - a0;
    const int lr = tg * t_rows;
    const int col = wave * kColsPerWave - 1 + gl;
    const bool upd = gl >= 1 && gl <= tgw - 2 && tg < S && col < nx;
    const bool cin = col >= 0 && col < nx;
    const int cc = min(max(col, 0), nx - 1), xc = min(max(col, 0), nx);
    const float* U = a.U + (size_t)cand * a.cstride;
    float* Uo = a.Uo + (size_t)cand * a.cstride;
    float* Pr = a.pre + (size_t)cand * a.cstride;
    const float* dzx = a.dzx + (size_t)cand * a.dzx_stride;
    const float* dzy = a.dzy + (size_t)cand * a.dzy_stride;
    const bool crest = a.crest != 0;
    const float* dzx2 = crest ? a.dzx2 + (size_t)cand * a.dzx_stride : nullptr;
    const float* dzy2 = crest ? a.dzy2 + (size_t)cand * a.dzy_stride : nullptr;
    const StepK& K = a.K;
    // Raw inputs of row k1: its old state (the interior neighbour for a ghost), wet and
    // g n^2, the jumps of its north face and of its west face.
    auto load = [&](int k1) -> FwdIn {
        FwdIn v;
        const int rr = min(max(k1, 0), ny - 1);
        const size_t o = (size_t)rr * a.P + cc;
        v.h = U[o];
        v.qx = U[a.plane + o];
        v.qy = U[2 * a.plane + o];
        const bool rin = k1 >= 0 && k1 < ny;
        v.code = 1;
        if (rin && !cin) v.code = a.ghost[2 * nx + (col < 0 ? 0 : ny) + k1];
        else if (!rin && cin) v.code = a.ghost[(k1 < 0 ? 0 : nx) + col];
        v.wet = a.wet[o];
        v.gn2 = a.gn2[o];
        v.dzs = dzy[(size_t)min(max(k1, 0), ny) * a.P + cc];
        v.dzw = dzx[(size_t)min(max(k1, 0), ny - 1) * a.P + xc];
        v.dzs2 = crest ? dzy2[(size_t)min(max(k1, 0), ny) * a.P + cc] : 0.0f;
        v.dzw2 = crest ? dzx2[(size_t)min(max(k1, 0), ny - 1) * a.P + xc] : 0.0f;
        return v;
    };
    // The ghost rule (boundary.rs:12-87), as nb_at: wall mirrors the normal discharge,
    // transmissive copies, Stage / Inflow as special_nb; corners (never read by an
    // interior cell) are 0.
    auto cell = [&](const FwdIn& v, int k1) -> Nb {
        const bool rin = k1 >= 0 && k1 < ny;
        if (!rin && !cin) return Nb{0.0f, 0.0f, 0.0f};
        Nb w{v.h, v.qx, v.qy};
        const bool ghost = !rin || !cin;
        const bool wall = ghost && v.code == 0;
        if (!cin && wall) w.qx = -w.qx;
        if (!rin && wall) w.qy = -w.qy;
        if (ghost && v.code >= 2)
            w = special_nb(a, cand, !cin ? 2 * nx + (col < 0 ? 0 : ny) + k1 : (k1 < 0 ? 0 : nx) + col, v.code, w);
        return w;
    };
    auto yface = [&](const Nb& L, const Vel& vl, const Nb& R, const Vel& vr, float dz, float dz2) -> Flux {
        return crest ? face_crest<false>(L.h, vl, R.h, vr, dz, dz2, K) : face_signed<false>(L.h, vl, R.h, vr, dz, K);
    };
    const int g0 = a0 + lr;   // this lane's first row
    const FwdIn ip = load(g0 - 1), ic = load(g0);
    FwdIn nxt = load(g0 + 1);
    const Nb Pn = cell(ip, g0 - 1);
    Nb Cn = cell(ic, g0);
    Vel vC = cell_velocity(Cn.h, Cn.qx, Cn.qy, K);
    float wet = ic.wet, gn2 = ic.gn2, dzw = ic.dzw, dzw2 = ic.dzw2;
    Flux fN;
    {
        const Vel vP = cell_velocity(Pn.h, Pn.qx, Pn.qy, K);
        fN = yface(Pn, vP, Cn, vC, ic.dzs, ic.dzs2);
    }
    for (int jb = a0; jb < a0 + t_rows; jb++) {
        const int j = jb + lr;   // this lane's row
        const bool rv = j < b0;
        const FwdIn in = nxt;    // row j+1
        if (j + 1 < b0) nxt = load(j + 2);
        const Nb Sn = cell(in, j + 1);
        const Vel vS = cell_velocity(Sn.h, Sn.qx, Sn.qy, K);
        const Flux fS = yface(Cn, vC, Sn, vS, in.dzs, in.dzs2);
        // West face: L = the left lane's cell, R = this lane's; the east face is the right
        // lane's west face.
        const float hw = from_left(Cn.h);
        Vel vW;
        vW.u = from_left(vC.u);
        vW.v = from_left(vC.v);
        vW.c = from_left(vC.c);
        const Flux fW = crest ? face_crest<true>(hw, vW, Cn.h, vC, dzw, dzw2, K) : face_signed<true>(hw, vW, Cn.h, vC, dzw, K);
        Flux fE;
        fE.m = from_right(fW.m);
        fE.nl = from_right(fW.nl);
        fE.nr = from_right(fW.nr);
        fE.t = from_right(fW.t);
        const CellNew u = cell_update(Cn.h, Cn.qx, Cn.qy, fW.m, fW.nr, fW.t, fE.m, fE.nl, fE.t, fN.m, fN.nr, fN.t, fS.m,
                                      fS.nl, fS.t, wet, gn2, K);
        if (upd && rv) {
            const size_t o = (size_t)j * a.P + col;
            Uo[o] = u.h;
            Uo[a.plane + o] = u.qx;
            Uo[2 * a.plane + o] = u.qy;
            // the state after the flux divergence (numerics.rs:148-153), as itr_replay_kernel
            Pr[o] = Cn.h - K.dtdx * (fE.m - fW.m) - K.dtdy * (fS.m - fN.m);
            Pr[a.plane + o] = Cn.qx - K.dtdx * (fE.nl - fW.nr) - K.dtdy * (fS.t - fN.t);
            Pr[2 * a.plane + o] = Cn.qy - K.dtdx * (fE.t - fW.t) - K.dtdy * (fS.nl - fN.nr);
        }
        Cn = Sn;
        vC = vS;
        fN = fS;
        wet = in.wet;
        gn2 = in.gn2;
        dzw = in.dzw;
        dzw2 = in.dzw2;
    }
}

// lam[cand][plane h][cell] += v (a peak-depth seed at the step that set the peak).
__global__ void itr_adj_seed_kernel(float* lam, size_t cstride, int cand, size_t cell, float v) {
    if (threadIdx.x == 0) lam[(size_t)cand * cstride + cell] += v;
}

// Count of cells where two states differ bitwise (replay check).
__global__ void __launch_bounds__(256) itr_adj_compare_kernel(const float* A, const float* B, size_t n,
                                                               unsigned long long* count) {
    unsigned long long bad = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad += __float_as_uint(A[i]) != __float_as_uint(B[i]) ? 1ull : 0ull;
    if (bad) atomicAdd(count, bad);
}

// ------------------------------------------------------------------ static scratch
// Device memory of itr_hip_run_batch_noalloc: every buffer of that entry point's batch
// (state, faces, control blocks, partial results, monitors, ledgers, staging) is carved
// from this array, so the call allocates nothing.  Each device the library runs on
// holds its own copy until the process ends.  Host code never names it after set-up:
// itr_hip_create maps it once per device with hipGetSymbolAddress and keeps the address.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// ------------------------------------------------------------------ host side
void set_err(char* err, size_t errlen, const char* fmt, ...) {
    if (!err || errlen == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_err(err, errlen, "%s failed: %s", #expr, hipGetErrorString(e_));            \
            return e_ == hipErrorOutOfMemory ? ITR_ENOMEM : ITR_EDEVICE;                     \
        }                                                                                    \
    } while (0)

// Two timing events destroyed on every return path.
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    EventPair() = default;
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    hipError_t create() {
        for (auto& x : e) {
            const hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    ~EventPair() {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Bump allocator over the static scratch of one device.  While an Arena is installed on
// the calling thread (ArenaScope), dev_alloc carves from it instead of calling hipMalloc,
// and the carved buffers are never freed (the next call reuses the array).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
};
thread_local Arena* tl_arena = nullptr;
struct ArenaScope {
    Arena* prev;
    explicit ArenaScope(Arena* a) : prev(tl_arena) { tl_arena = a; }
    ~ArenaScope() { tl_arena = prev; }
};

// The scratch address of each device, mapped once (hipGetSymbolAddress with the device
// current) by the first itr_hip_create on it.
constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch[kMaxDevices];

// FixedSum<f32> of one row, then the f64 sum over rows (solver.rs:250-262).
double host_volume(const std::vector<float>& h, int nx, int ny, double area) {
    double v = 0.0;
    for (int r = 0; r < ny; r++) {
        float acc[16];
        for (auto& x : acc) x = 0.0f;
        for (int i = 1; i <= nx; i++) acc[i & 15] += h[(size_t)r * nx + (i - 1)];
        float l8[8];
        for (int k = 0; k < 8; k++) l8[k] = acc[2 * k] + acc[2 * k + 1];
        const float l4[4] = {l8[0] + l8[1], l8[2] + l8[3], l8[4] + l8[5], l8[6] + l8[7]};
        v += (double)((l4[0] + l4[1]) + (l4[2] + l4[3]));
    }
    return v * area;
}

struct GridHost {
    itr_grid in;
    DevGrid d;
    DevParams prm{};             // host copy of this grid's parameters
    int max_ledger = 0;
    std::vector<void*> allocs;
    int waves = 1;
    std::vector<uint32_t> mon;   // monitored cells (row * nx + col)
};


}  // namespace

constexpr int kMaxLists = 4;   // launch lists per batch (streams per context)

struct itr_ctx {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipStream_t aux[kMaxLists - 1] = {};   // the streams of launch lists 1.. (default priority)
    hipEvent_t ev_fork = nullptr;          // fork / join of the lists' streams (no timing)
    hipEvent_t ev_join[kMaxLists - 1] = {};
    std::string info;
    int band_rows = 0;
    int steps_per_chunk = 0;
    int use_graph = 1;
    int64_t min_lds = 0;   // reserve at least this much LDS per step workgroup (caps occupancy)
    int adjoint_mode = 0;  // reverse sweep: 0 by batch size, 1 band march, 2 per-cell kernels
    int adjoint_rows = 0;  // rows per band of the reverse band march (0: 16)
    int control_task = -1; // ledger and control block: -1 by launch shape, 0 first band, 1 control task
    int tail_rows = -1;     // rows at a grid's end marched in shorter bands (0: none, -1: by launch shape)
    int tail_band_rows = 0; // their band height (0: a third of the band height)
    int list_split = 0;     // launch lists per class of grids, by candidates (0: two for a list of several rounds)
    struct Tuned { int nx, ny, crest, R; };
    std::vector<Tuned> tuned;   // band rows chosen by itr_hip_autotune, per grid shape
    void* scratch = nullptr;    // this device's static scratch (itr_hip_run_batch_noalloc)
};

namespace {

// Band rows of a grid: the fixed option, else the tuned value for the shape, else enough
// bands to give the whole batch a few workgroups per CU (capped at 12 rows).
int choose_band_rows(const itr_ctx* ctx, const itr_grid* grids, int n_grids, int g) {
    const itr_grid& G = grids[g];
    int R = 0;
    if (ctx->band_rows > 0) R = ctx->band_rows;
    else {
        for (const auto& t : ctx->tuned)
            if (t.nx == G.nx && t.ny == G.ny && t.crest == (G.crest ? 1 : 0)) R = t.R;
        if (R == 0) {
            double rows = 0.0;
            for (int k = 0; k < n_grids; k++) rows += (double)grids[k].ny * (double)std::max(grids[k].n_cand, 0);
            R = (int)std::lround(std::min(12.0, std::max(2.0, rows / 768.0)));
        }
    }
    return std::max(1, std::min(R, G.ny));
}

// Shorter bands at a grid's end.  A launch dispatches its tasks in band-major order, so the
// bands at the end of each grid are the launch's last workgroups, and in a launch of
// several rounds of workgroups the compute units that have no task left wait while they
// run: with the last rows of a grid (tail_rows or more) in bands of tail_band_rows rows,
// that tail is made of short workgroups (128 x 512^2 batch, 30-row bands: the last 32 rows
// in 10-row bands, 365.6 -> 354.8 us per step).  Such a band is a band of a copy of the
// grid's descriptor with R = tail_band_rows (a shadow grid, placed after the batch's grids
// in the device array): the step kernel is unchanged.  The grid's own bands
// 0 .. nb_main-1 end on a multiple of the short height, where the shadow's band `first`
// starts.  By default (tail_rows -1) a grid in a launch of several rounds gets a tail of
// about one band (R rows) in bands of the largest divisor of R up to R/3 rows.
struct TailLayout { bool on = false; int nb_main = 0, R2 = 0, first = 0, count = 0; };
TailLayout tail_layout(const DevGrid& D, const itr_ctx* ctx, bool several_rounds) {
    TailLayout t;
    const int R = D.R;
    const int T = ctx->tail_rows >= 0 ? ctx->tail_rows : (several_rounds ? R : 0);
    int R2 = ctx->tail_band_rows;
    if (R2 <= 0) {   // the largest divisor of R up to a third of it (so that the tail starts on one)
        R2 = 0;
        for (int d = (R + 2) / 3; d >= 2 && R2 == 0; d--)
            if (R % d == 0) R2 = d;
    }
    if (T <= 0 || R2 <= 0 || R2 >= R || D.ny - T < R) return t;
    int nb = (D.ny - T) / R;
    while (nb > 0 && (nb * R) % R2 != 0) nb--;
    if (nb <= 0) return t;
    t.on = true;
    t.nb_main = nb;
    t.R2 = R2;
    t.first = nb * R / R2;
    t.count = (D.ny + R2 - 1) / R2 - t.first;
    return t;
}

// The device array of grid descriptors: the batch's grids, then a shadow for each grid
// with a tail layout (tails[g], in grid order).
std::vector<DevGrid> device_grids(const std::vector<DevGrid>& dg, const std::vector<TailLayout>& tails) {
    std::vector<DevGrid> all = dg;
    for (size_t g = 0; g < dg.size(); g++) {
        const DevGrid& D = dg[g];
        const TailLayout t = g < tails.size() ? tails[g] : TailLayout{};
        if (!t.on) continue;
        DevGrid S = D;
        S.R = t.R2;
        S.nbands = t.first + t.count;
        all.push_back(S);
    }
    return all;
}

}  // namespace

// The step tasks of one class of grids (physics, face mode, ghosts, waves per row), run
// as one kernel launch per step.  No candidate is in two lists, so the lists of a batch
// are independent chains of launches: each list after the first runs on a stream of its
// own, forked from and joined to the context's stream around every run of launches, and
// a list of small grids does not wait each step for the slowest band of a large one.
struct LaunchList {
    KArgs args{};
    std::vector<int> grids;
    std::vector<int> cbeg, cend;    // per grid: its local candidates cbeg .. cend-1 in this list
    bool part = false;              // one of the lists a class of grids was split into
    size_t task_off = 0;            // its slice of the batch's task buffer
    int ntasks = 0;
    int block = 64;
    size_t lds = 0;
    int modes = 1;                  // face modes present: 1 signed, 2 crest, 3 both; | 4 Stage / Inflow ghosts or tape
    bool li = false;                // local-inertial physics
    bool ctl_task = false;          // control tasks (band -1) carry the ledger and control block
};

struct itr_batch {
    itr_ctx* ctx = nullptr;
    std::vector<GridHost> grids;
    DevParams prm{};                // the first grid's parameters (single-grid users: adjoint, schedules)
    KArgs args{};                   // grids, control and accumulators (the lists add their tasks)
    std::vector<LaunchList> lists;
    std::vector<TailLayout> tails;  // per grid: shorter bands at its end (fill_lists)
    void* d_grids = nullptr;
    void* d_tasks = nullptr;
    void* d_snap = nullptr;         // CandCtl [2][ncand]
    void* d_acc = nullptr;          // StepAcc [4][ncand]
    void* d_status = nullptr;
    uint32_t* h_status = nullptr;   // pinned
    int ncand = 0;
    int ntasks = 0;                 // over all lists
    uint64_t launches = 0;          // step launches so far; control copy (launches-1)&1 is current
    // Captured chunks of step launches, longest first: graph[0] holds the batch's chunk
    // (graph_steps[0] launches), the others 16, 8 and 4 launches where shorter, so that an
    // advance by a few steps also runs as graphs (back-to-back graph launches start each
    // step kernel sooner than separate launches do).  One graph per launch list, launched
    // on the list's stream, so that the lists' chunks follow one another on each stream
    // without a join between chunks.
    static constexpr int kGraphs = 4;
    hipGraphExec_t graph[kGraphs][kMaxLists] = {};
    int graph_steps[kGraphs] = {};
    itr_stats stats{};
    std::vector<std::vector<double>> cb_mon;   // scratch for callbacks
    // Built by itr_hip_run_batch_noalloc: device buffers live in the static scratch (never
    // freed), the status word is polled into host memory, launches are direct (no graph)
    // and untimed (no events).
    bool no_alloc = false;
    std::vector<uint32_t> h_status_host;
};

namespace {

void drop_graphs(itr_batch* b) {
    for (int i = 0; i < itr_batch::kGraphs; i++) {
        for (int l = 0; l < kMaxLists; l++) {
            if (b->graph[i][l]) (void)hipGraphExecDestroy(b->graph[i][l]);
            b->graph[i][l] = nullptr;
        }
        b->graph_steps[i] = 0;
    }
}

void free_batch_mem(itr_batch* b) {
    // (buffers carved from the static scratch were never added to these lists)
    for (auto& g : b->grids)
        for (void* p : g.allocs) (void)hipFree(p);
    if (!b->no_alloc) {
        void* ps[] = {b->d_grids, b->d_tasks, b->d_snap, b->d_acc, b->d_status};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (b->h_status) (void)hipHostFree(b->h_status);
    }
    drop_graphs(b);
}

template <class T>
itr_status dev_alloc(std::vector<void*>& allocs, T** out, size_t count, char* err, size_t errlen) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (Arena* A = tl_arena) {
        const size_t off = (A->used + 255) & ~(size_t)255;
        if (off > A->cap || bytes > A->cap - off) {
            set_err(err, errlen,
                    "batch does not fit the %zu MiB static scratch of the no-allocation entry point "
                    "(a buffer of %zu bytes at offset %zu); use itr_hip_run_batch",
                    A->cap >> 20, bytes, off);
            return ITR_ENOMEM;
        }
        A->used = off + bytes;
        *out = (T*)(A->base + off);
        return ITR_OK;
    }
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, bytes));
    allocs.push_back(p);
    *out = (T*)p;
    return ITR_OK;
}

// Task order within each XCD's share of the launch (the kernel maps block i to XCD i % 8
// and XCD x to tasks [x*per, (x+1)*per)): band-major, so that workgroups resident together
// on one XCD read the same rows of the shared wet / gn2 planes.
void order_tasks(std::vector<int4>& tasks) {
    const size_t n = tasks.size();
    if (n == 0) return;
    const size_t per = (n + 7) / 8;
    for (size_t x0 = 0; x0 < n; x0 += per) {
        const size_t x1 = std::min(n, x0 + per);
        std::stable_sort(tasks.begin() + x0, tasks.begin() + x1, [](const int4& a, const int4& b) {
            return a.x != b.x ? a.x < b.x : a.z < b.z;
        });
    }
}

// Launch lists of a batch: one per class of grids (physics, face mode, Stage / Inflow
// ghosts or tape, waves per row) when there are at most kMaxLists classes, else one per
// physics (whose kernels then carry every face mode present).
bool use_control_task(int cus, size_t nbands, int waves, size_t lds, int modes, bool li);

// A list's workgroup width, LDS and whether its bands fit the chip in one round of
// workgroups (use_control_task), for the grids' current band heights.
bool list_one_round(const itr_batch* b, const std::vector<DevGrid>& dg, LaunchList& L, int* waves_out, size_t* lds_out) {
    int waves = 1;
    size_t lds = (size_t)std::max<int64_t>(b->ctx->min_lds, 0), nbands = 0;
    L.modes = 0;
    for (size_t i = 0; i < L.grids.size(); i++) {
        const DevGrid& D = dg[L.grids[i]];
        nbands += (size_t)(L.cend[i] - L.cbeg[i]) * (size_t)D.nbands;
        waves = std::max(waves, D.waves);
        L.modes |= (D.crest ? 2 : 1) | (D.special || D.tape_mode ? 4 : 0);
        lds = std::max(lds, (D.li ? li_lds_layout(D.R, D.nx, D.special != 0).total
                                  : lds_layout(D.R, D.nx, D.special != 0).total) * sizeof(float));
    }
    if (waves_out) *waves_out = waves;
    if (lds_out) *lds_out = lds;
    return use_control_task(b->ctx->cus, nbands, waves, lds, L.modes, L.li);
}

void assign_lists(itr_batch* b, const std::vector<DevGrid>& dg) {
    b->lists.clear();
    std::vector<uint64_t> keys;
    std::vector<int> which(dg.size());
    for (size_t g = 0; g < dg.size(); g++) {
        const DevGrid& D = dg[g];
        const uint64_t k = (uint64_t)D.li | (uint64_t)(D.crest ? 1 : 0) << 1 | (uint64_t)(D.special || D.tape_mode ? 1 : 0) << 2 |
                           (uint64_t)D.waves << 3;
        size_t i = 0;
        while (i < keys.size() && keys[i] != k) i++;
        if (i == keys.size()) keys.push_back(k);
        which[g] = (int)i;
    }
    if (keys.size() > (size_t)kMaxLists) {
        keys.clear();
        for (size_t g = 0; g < dg.size(); g++) {
            const uint64_t k = (uint64_t)dg[g].li;
            size_t i = 0;
            while (i < keys.size() && keys[i] != k) i++;
            if (i == keys.size()) keys.push_back(k);
            which[g] = (int)i;
        }
    }
    b->lists.resize(keys.size());
    for (size_t g = 0; g < dg.size(); g++) {
        LaunchList& L = b->lists[which[g]];
        L.grids.push_back((int)g);
        L.cbeg.push_back(0);
        L.cend.push_back(std::max(dg[g].ncand, 0));
        L.li = dg[g].li != 0;
    }
    // Lists split by candidates: a list's step launches then run beside another's on its
    // own stream, and the end of one launch, when few of its workgroups remain and the
    // compute units run out of work, overlaps the other list's work (128 x 512^2 batch:
    // 356 -> 338 us per step at 30-row bands, 334 us at 32).  list_split 0 (the default)
    // splits each list of several rounds of workgroups in two while streams remain; 1 keeps
    // the lists whole; k splits every list in k (within kMaxLists lists).
    std::vector<int> parts(b->lists.size(), 1);
    if (b->ctx->list_split >= 2) {
        const int k = std::min(b->ctx->list_split, kMaxLists / std::max<int>(1, (int)b->lists.size()));
        for (int& q : parts) q = std::max(k, 1);
    } else if (b->ctx->list_split == 0) {
        int n = (int)b->lists.size();
        for (size_t i = 0; i < b->lists.size() && n < kMaxLists; i++)
            if (!list_one_round(b, dg, b->lists[i], nullptr, nullptr)) {
                parts[i] = 2;
                n++;
            }
    }
    if (*std::max_element(parts.begin(), parts.end()) > 1) {
        std::vector<LaunchList> out;
        for (size_t li = 0; li < b->lists.size(); li++) {
            const LaunchList& L = b->lists[li];
            const int np = parts[li];
            for (int q = 0; q < np; q++) {
                LaunchList M;
                M.li = L.li;
                M.part = np > 1;
                for (size_t i = 0; i < L.grids.size(); i++) {
                    const int n = L.cend[i] - L.cbeg[i];
                    const int a = L.cbeg[i] + (int)((int64_t)n * q / np), e = L.cbeg[i] + (int)((int64_t)n * (q + 1) / np);
                    if (e <= a) continue;
                    M.grids.push_back(L.grids[i]);
                    M.cbeg.push_back(a);
                    M.cend.push_back(e);
                }
                if (!M.grids.empty()) out.push_back(M);
            }
        }
        b->lists = out;
    }
}

// Whether a list's candidates get control tasks: when its bands fit the chip in one round
// of workgroups (by waves at the kernel's register budget and by LDS), the step is
// latency-bound and the ledger would sit on its critical path.  A list of several rounds
// is throughput-bound: a control task would only hold a workgroup's LDS for the length of
// the ledger, which the first band runs as well after its march.
bool use_control_task(int cus, size_t nbands, int waves, size_t lds, int modes, bool li) {
    const int waves_per_simd = (!li && modes == 1) ? 6 : 4;   // 80 VGPRs (signed HLL), else 128
    const int by_waves = 4 * waves_per_simd / std::max(1, waves);
    const int by_lds = lds > 0 ? (int)((size_t)160 * 1024 / lds) : 64;
    const size_t per_cu = (size_t)std::max(1, std::min(by_waves, by_lds));
    return nbands <= (size_t)cus * per_cu;
}

// The lists' tasks (concatenated into `all`, each list's slice in XCD order), workgroup
// size, face modes, LDS and control-task mode for the grids' current band heights.
// Returns the largest LDS.
size_t fill_lists(itr_batch* b, const std::vector<DevGrid>& dg, std::vector<int4>& all) {
    all.clear();
    size_t lds_max = 0;
    // Each list's shape (its grids' bands fit the chip in one round of workgroups or not)
    // and its grids' tail layouts, then the shadow indices (device_grids' order).
    std::vector<TailLayout>& tl = b->tails;
    tl.assign(dg.size(), TailLayout{});
    std::vector<int> shadow(dg.size(), -1);
    std::vector<char> several(dg.size(), 0);   // a grid in a list of several rounds
    for (LaunchList& L : b->lists) {
        int waves = 1;
        size_t lds = 0;
        const bool one_round = list_one_round(b, dg, L, &waves, &lds);
        L.ctl_task = b->ctx->control_task >= 0 ? b->ctx->control_task == 1 : one_round;
        L.block = waves * kWave;
        L.lds = lds;
        // (a list split by candidates has its launches' ends covered by the other part's
        // work: short end bands only cost it their extra workgroups)
        for (int g : L.grids) several[g] = several[g] || (!one_round && !L.part);
    }
    for (size_t g = 0; g < dg.size(); g++) tl[g] = tail_layout(dg[g], b->ctx, several[g] != 0);
    int nsh = 0;
    for (size_t g = 0; g < dg.size(); g++)
        if (tl[g].on) shadow[g] = (int)dg.size() + nsh++;
    for (LaunchList& L : b->lists) {
        std::vector<int4> t;
        for (size_t i = 0; i < L.grids.size(); i++) {
            const int g = L.grids[i];
            const DevGrid& D = dg[g];
            for (int k = L.cbeg[i]; k < L.cend[i]; k++) {
                if (L.ctl_task) t.push_back(make_int4(g, k, -1, D.cand0 + k));
                const TailLayout& T = tl[g];
                for (int band = 0; band < (T.on ? T.nb_main : D.nbands); band++) t.push_back(make_int4(g, k, band, D.cand0 + k));
                for (int j = 0; j < (T.on ? T.count : 0); j++) t.push_back(make_int4(shadow[g], k, T.first + j, D.cand0 + k));
            }
        }
        order_tasks(t);
        L.task_off = all.size();
        L.ntasks = (int)t.size();
        all.insert(all.end(), t.begin(), t.end());
        lds_max = std::max(lds_max, L.lds);
    }
    b->ntasks = (int)all.size();
    return lds_max;
}

// Point each list's kernel arguments at its slice of the task buffer.
void bind_lists(itr_batch* b) {
    for (LaunchList& L : b->lists) {
        L.args = b->args;
        L.args.tasks = (const int4*)b->d_tasks + L.task_off;
        L.args.ntasks = L.ntasks;
        L.args.ctl_task = L.ctl_task ? 1 : 0;
    }
}

// A grid given as designs (itr_design), built on the device in three parts: design_begin
// uploads the designs and computes every candidate's fine dz (and at level >= 2 its coarse
// level z); design_levels hands the candidates' level z to the host (Stage ghosts and
// initial_stage depend on it); design_finish builds the faces into the grid's zero-filled
// face planes (boundary faces stay 0) and returns the edits' volumes.  The scratch buffers
// live as long as the DesignWork.
struct DesignWork {
    std::vector<void*> tmp;
    DesignDev dd{};
    int nc = 0;
    int64_t maxwh = 0;
    ~DesignWork() {
        for (void* q : tmp) (void)hipFree(q);
    }
};

itr_status design_begin(itr_ctx* ctx, const itr_grid& I, const DevGrid& D, DesignWork& W, char* err, size_t errlen) {
    const itr_design& Ds = *I.design;
    const int nx = I.nx, ny = I.ny, nc = I.n_cand;
    W.nc = nc;
    if (nc <= 0) return ITR_OK;
    const int lv = std::max(Ds.level, 1);
    const int fnx = lv == 1 ? nx : Ds.fine_nx, fny = lv == 1 ? ny : Ds.fine_ny;
    const size_t fn = (size_t)fnx * fny;
    const int np = Ds.prim_start[nc];
    std::vector<int64_t> off(nc);
    int64_t tot = 0;
    for (int k = 0; k < nc; k++) {
        off[k] = tot;
        const int64_t wh = (int64_t)Ds.bbox[4 * k + 2] * Ds.bbox[4 * k + 3];
        tot += wh;
        W.maxwh = std::max(W.maxwh, wh);
    }
    double *prim = nullptr, *dzbuf = nullptr, *z = nullptr, *vol = nullptr, *zc = nullptr;
    int32_t *ps = nullptr, *bb = nullptr;
    int64_t* doff = nullptr;
    uint8_t *wall = nullptr, *edit = nullptr;
    itr_status s;
#define DTRY(expr)                  \
    if ((s = (expr)) != ITR_OK) return s;
    DTRY(dev_alloc(W.tmp, &prim, (size_t)std::max(np, 1) * 12, err, errlen));
    DTRY(dev_alloc(W.tmp, &ps, (size_t)nc + 1, err, errlen));
    DTRY(dev_alloc(W.tmp, &bb, (size_t)nc * 4, err, errlen));
    DTRY(dev_alloc(W.tmp, &doff, (size_t)nc, err, errlen));
    DTRY(dev_alloc(W.tmp, &dzbuf, (size_t)std::max<int64_t>(tot, 1), err, errlen));
    DTRY(dev_alloc(W.tmp, &z, fn, err, errlen));
    DTRY(dev_alloc(W.tmp, &wall, fn, err, errlen));
    if (Ds.editable) DTRY(dev_alloc(W.tmp, &edit, fn, err, errlen));
    DTRY(dev_alloc(W.tmp, &vol, (size_t)nc * 2, err, errlen));
    DTRY(dev_alloc(W.tmp, &zc, (size_t)nc * nx * ny, err, errlen));
#undef DTRY
    const hipStream_t st = ctx->stream;
    bool ok = (np == 0 || hipMemcpy(prim, Ds.prim, sizeof(double) * 12 * np, hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemcpy(ps, Ds.prim_start, sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(bb, Ds.bbox, sizeof(int32_t) * 4 * nc, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(doff, off.data(), sizeof(int64_t) * nc, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(z, Ds.z_fine, sizeof(double) * fn, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(wall, Ds.wall, fn, hipMemcpyHostToDevice) == hipSuccess &&
              (!Ds.editable || hipMemcpy(edit, Ds.editable, fn, hipMemcpyHostToDevice) == hipSuccess);
    if (!ok) {
        set_err(err, errlen, "design upload failed");
        return ITR_EDEVICE;
    }
    DesignDev& dd = W.dd;
    dd.prim = prim;
    dd.prim_start = ps;
    dd.bbox = bb;
    dd.dz_off = doff;
    dd.dzbuf = dzbuf;
    dd.z = z;
    dd.wall = wall;
    dd.editable = edit;
    dd.ox = Ds.origin_x;
    dd.oy = Ds.origin_y;
    dd.dx = Ds.dx;
    dd.dy = Ds.dy;
    dd.max_dz = Ds.max_abs_dz;
    dd.z_wall = Ds.z_wall;
    dd.area = Ds.dx * Ds.dy;
    dd.nx = nx;
    dd.ny = ny;
    dd.P = D.P;
    dd.ncand = nc;
    dd.fnx = fnx;
    dd.fny = fny;
    dd.lv = lv;
    dd.zc = zc;
    dd.plane = D.plane;
    dd.plane_y = D.plane_y;
    dd.volumes = vol;
    if (W.maxwh > 0)
        hipLaunchKernelGGL(itr_materialize_kernel, dim3((unsigned)((W.maxwh + 255) / 256), (unsigned)nc), dim3(256), 0, st,
                           dd);
    hipLaunchKernelGGL(itr_design_coarse_kernel, dim3((unsigned)(((size_t)nx * ny + 255) / 256), (unsigned)nc), dim3(256),
                       0, st, dd);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_err(err, errlen, "design materialization failed: %s", hipGetErrorString(e));
        return ITR_EDEVICE;
    }
    return ITR_OK;
}

itr_status design_levels(itr_ctx* ctx, const DesignWork& W, std::vector<double>& zlev, char* err, size_t errlen) {
    const size_t n = (size_t)W.nc * W.dd.nx * W.dd.ny;
    zlev.assign(std::max<size_t>(n, 1), 0.0);
    if (W.nc <= 0) return ITR_OK;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(zlev.data(), W.dd.zc, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_err(err, errlen, "design levels failed: %s", hipGetErrorString(e));
        return ITR_EDEVICE;
    }
    return ITR_OK;
}

itr_status design_finish(itr_ctx* ctx, const itr_grid& I, DesignWork& W, float* d_dzx, float* d_dzy, float* d_dzx2,
                         float* d_dzy2, char* err, size_t errlen) {
    const int nc = W.nc;
    if (nc <= 0) return ITR_OK;
    DesignDev& dd = W.dd;
    dd.dzx = d_dzx;
    dd.dzy = d_dzy;
    dd.dzx2 = d_dzx2;
    dd.dzy2 = d_dzy2;
    const hipStream_t st = ctx->stream;
    const int nx = dd.nx, ny = dd.ny;
    if (dd.lv == 1)
        hipLaunchKernelGGL(itr_design_faces_kernel, dim3((unsigned)((nx + 1 + 255) / 256), (unsigned)ny, (unsigned)nc),
                           dim3(256), 0, st, dd);
    else
        hipLaunchKernelGGL(itr_design_crest_kernel, dim3((unsigned)((nx + 1 + 255) / 256), (unsigned)ny, (unsigned)nc),
                           dim3(256), 0, st, dd);
    hipLaunchKernelGGL(itr_design_volumes_kernel, dim3((unsigned)nc), dim3(256), 0, st, dd);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && I.design->volumes)
        e = hipMemcpy(I.design->volumes, dd.volumes, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_err(err, errlen, "design materialization failed: %s", hipGetErrorString(e));
        return ITR_EDEVICE;
    }
    return ITR_OK;
}

// Device parameters of one grid (SolverParams, solver.rs:26-49, with the f32 constants of
// Solver::consts, solver.rs:269-285), the hyetograph and locked schedule uploaded into
// `allocs`.  max_ledger: ledger rows per candidate (one per sync point plus one).
itr_status set_params(const itr_solver_params* p, DevParams& P, int& max_ledger, std::vector<void*>& allocs, char* err,
                      size_t errlen) {
    const double dur = p->duration_s, syn = p->sync_interval_s;
    const uint32_t n_sync = (uint32_t)std::ceil(dur / syn);
    max_ledger = (int)std::min<uint64_t>(n_sync + 1, 1u << 20);
    P = DevParams{};
    P.duration = dur;
    P.sync = syn;
    P.cfl = p->cfl;
    P.dt_max = p->dt_max_s;
    P.cfl_hard = p->cfl_hard;
    P.infil_cap = p->infil_capacity;
    P.inflow_scale = p->inflow_scale;
    P.n_rain = p->n_rain;
    P.n_sync = n_sync;
    P.skip_dry = p->skip_dry;
    P.max_steps = p->max_steps;
    P.max_ledger = max_ledger;
    P.h_dry = (float)p->h_dry;
    P.heps2 = (float)(p->h_eps * p->h_eps);
    P.q_min = (float)1e-30;
    // The step's short reciprocal of the velocity denominator h^2 + max(h^2, h_eps^2) relies
    // on h_eps^2 being a normal f32 (h_eps >= about 1.1e-19 m; the reference's default is 1e-5).
    if (!(P.heps2 >= FLT_MIN)) {
        set_err(err, errlen, "h_eps = %g: h_eps^2 is below the normal f32 range", p->h_eps);
        return ITR_DECLINED;
    }
    double *rt = nullptr, *rr = nullptr;
    itr_status s;
    if ((s = dev_alloc(allocs, &rt, (size_t)p->n_rain, err, errlen)) != ITR_OK) return s;
    if ((s = dev_alloc(allocs, &rr, (size_t)p->n_rain, err, errlen)) != ITR_OK) return s;
    if (p->n_rain > 0 && (hipMemcpy(rt, p->rain_t, sizeof(double) * p->n_rain, hipMemcpyHostToDevice) != hipSuccess ||
                          hipMemcpy(rr, p->rain_r, sizeof(double) * p->n_rain, hipMemcpyHostToDevice) != hipSuccess)) {
        set_err(err, errlen, "rain upload failed");
        return ITR_EDEVICE;
    }
    P.rain_t = rt;
    P.rain_r = rr;
    if (p->locked_dt && p->n_locked > 0) {
        double* ld = nullptr;
        if ((s = dev_alloc(allocs, &ld, (size_t)p->n_locked, err, errlen)) != ITR_OK) return s;
        if (hipMemcpy(ld, p->locked_dt, sizeof(double) * p->n_locked, hipMemcpyHostToDevice) != hipSuccess) {
            set_err(err, errlen, "locked dt upload failed");
            return ITR_EDEVICE;
        }
        P.locked = ld;
        P.n_locked = p->n_locked;
    }
    return ITR_OK;
}

// The sync-callback flag in every grid's parameters: candidates pause at sync points for
// the host only while a callback is attached.
itr_status set_callback_flag(itr_batch* b, int has_cb, char* err, size_t errlen) {
    std::vector<DevGrid> h(b->grids.size());
    for (size_t g = 0; g < b->grids.size(); g++) {
        b->grids[g].prm.has_cb = has_cb;
        b->grids[g].d.prm.has_cb = has_cb;
        h[g] = b->grids[g].d;
    }
    b->prm.has_cb = has_cb;
    h = device_grids(h, b->tails);
    if (!h.empty()) HIP_TRY(hipMemcpy(b->d_grids, h.data(), sizeof(DevGrid) * h.size(), hipMemcpyHostToDevice));
    return ITR_OK;
}

// One step launch of a list.
void launch_list(const LaunchList& L, hipStream_t st, uint32_t lk) {
    const dim3 grid((unsigned)(((L.ntasks + 7) / 8) * 8)), blk((unsigned)L.block);
    if (L.li) {
        if (L.modes == 1)
            hipLaunchKernelGGL(itr_step_li_signed, grid, blk, L.lds, st, L.args, lk);
        else if (L.modes == 2)
            hipLaunchKernelGGL(itr_step_li_crest, grid, blk, L.lds, st, L.args, lk);
        else
            hipLaunchKernelGGL(itr_step_li_general, grid, blk, L.lds, st, L.args, lk);
    } else if (L.modes & 4) {
        hipLaunchKernelGGL(itr_step_special, grid, blk, L.lds, st, L.args, lk);
    } else if (L.modes == 1) {
        hipLaunchKernelGGL(itr_step_signed, grid, blk, L.lds, st, L.args, lk);
    } else if (L.modes == 2) {
        hipLaunchKernelGGL(itr_step_crest, grid, blk, L.lds, st, L.args, lk);
    } else {
        hipLaunchKernelGGL(itr_step_mixed, grid, blk, L.lds, st, L.args, lk);
    }
}

// The stream of launch list l: the context's stream for the first list, an auxiliary
// stream (default priority) for each further one.
hipStream_t list_stream(const itr_batch* b, int l) { return l == 0 ? b->ctx->stream : b->ctx->aux[l - 1]; }

// Lists after the first run on their own streams between a fork from and a join to the
// context's stream (around a whole run of launches, not each chunk).
itr_status fork_lists(itr_batch* b, char* err, size_t errlen) {
    const int nl = (int)b->lists.size();
    if (nl <= 1) return ITR_OK;
    HIP_TRY(hipEventRecord(b->ctx->ev_fork, b->ctx->stream));
    for (int i = 1; i < nl; i++) HIP_TRY(hipStreamWaitEvent(b->ctx->aux[i - 1], b->ctx->ev_fork, 0));
    return ITR_OK;
}
itr_status join_lists(itr_batch* b, char* err, size_t errlen) {
    const int nl = (int)b->lists.size();
    for (int i = 1; i < nl; i++) {
        HIP_TRY(hipEventRecord(b->ctx->ev_join[i - 1], b->ctx->aux[i - 1]));
        HIP_TRY(hipStreamWaitEvent(b->ctx->stream, b->ctx->ev_join[i - 1], 0));
    }
    return ITR_OK;
}

// Enqueue n step launches of every list on its own stream, launch indices from lk0
// (mod 4); the caller forks and joins the streams.
itr_status enqueue_steps(itr_batch* b, int n, uint64_t lk0, char* err, size_t errlen) {
    for (int i = 0; i < (int)b->lists.size(); i++) {
        const LaunchList& L = b->lists[i];
        if (L.ntasks == 0) continue;
        for (int s = 0; s < n; s++) launch_list(L, list_stream(b, i), (uint32_t)((lk0 + (uint64_t)s) & 3u));
    }
    HIP_TRY(hipGetLastError());
    return ITR_OK;
}

// n launches continuing the batch's launch sequence: while the sequence is at a multiple
// of 4, the longest captured chunk that fits; single launches otherwise.
itr_status launch_steps(itr_batch* b, int64_t n, char* err, size_t errlen) {
    itr_status s = fork_lists(b, err, errlen);
    if (s != ITR_OK) return s;
    while (n > 0) {
        int gi = -1;
        if ((b->launches & 3u) == 0)
            for (int i = 0; i < itr_batch::kGraphs && gi < 0; i++)
                if (b->graph_steps[i] > 0 && b->graph_steps[i] <= n) gi = i;
        if (gi >= 0) {
            for (int l = 0; l < (int)b->lists.size(); l++)
                if (b->graph[gi][l]) HIP_TRY(hipGraphLaunch(b->graph[gi][l], list_stream(b, l)));
            b->launches += (uint64_t)b->graph_steps[gi];
            n -= b->graph_steps[gi];
        } else {
            int64_t k = n;   // single launches: the remainder, or up to the next multiple of 4
            if (b->graph_steps[0] > 0 && (b->launches & 3u) != 0) k = std::min<int64_t>(n, 4 - (int64_t)(b->launches & 3u));
            if ((s = enqueue_steps(b, (int)k, b->launches, err, errlen)) != ITR_OK) return s;
            b->launches += (uint64_t)k;
            n -= k;
        }
    }
    return join_lists(b, err, errlen);
}

// The current control copy of every candidate.
itr_status read_ctl(itr_batch* b, std::vector<CandCtl>& out, char* err, size_t errlen) {
    out.resize(b->ncand);
    if (b->ncand == 0) return ITR_OK;
    const size_t half = b->launches == 0 ? 1 : (size_t)((b->launches - 1) & 1u);
    HIP_TRY(hipMemcpy(out.data(), (CandCtl*)b->d_snap + half * b->ncand, sizeof(CandCtl) * b->ncand,
                      hipMemcpyDeviceToHost));
    return ITR_OK;
}
CandCtl* cur_ctl_dev(itr_batch* b, int c) {
    const size_t half = b->launches == 0 ? 1 : (size_t)((b->launches - 1) & 1u);
    return (CandCtl*)b->d_snap + half * b->ncand + c;
}

// Capture n launches (launch indices 0..n-1 mod 4) in a hipGraph.
// One graph per launch list of n step launches (launch indices 0 .. n-1 mod 4), captured
// on the list's stream; out[l] stays null for a list without tasks.
itr_status capture_steps(itr_batch* b, int n, hipGraphExec_t* out, char* err, size_t errlen) {
    for (int l = 0; l < (int)b->lists.size(); l++) {
        const LaunchList& L = b->lists[l];
        if (L.ntasks == 0) continue;
        const hipStream_t st = list_stream(b, l);
        hipGraph_t g;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < n; s++) launch_list(L, st, (uint32_t)(s & 3));
        const hipError_t le = hipGetLastError();
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (le != hipSuccess) {
            if (ce == hipSuccess) (void)hipGraphDestroy(g);
            set_err(err, errlen, "graph capture: %s", hipGetErrorString(le));
            return ITR_EDEVICE;
        }
        if (ce != hipSuccess) {
            set_err(err, errlen, "graph capture failed: %s", hipGetErrorString(ce));
            return ITR_EDEVICE;
        }
        hipError_t ie = hipGraphInstantiate(&out[l], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            out[l] = nullptr;
            set_err(err, errlen, "graph instantiate failed: %s", hipGetErrorString(ie));
            return ITR_EDEVICE;
        }
    }
    return ITR_OK;
}

// Capture chunks of K launches and, where shorter than K, of 16, 8 and 4 launches;
// K = 0 drops them.
itr_status prepare_graph(itr_batch* b, int K, char* err, size_t errlen) {
    if (K == b->graph_steps[0] && (K == 0 || b->graph_steps[0] > 0)) return ITR_OK;
    drop_graphs(b);
    if (K == 0 || b->ntasks == 0) return ITR_OK;
    int sizes[itr_batch::kGraphs] = {K, 16, 8, 4};
    int ng = 0;
    for (int i = 0; i < itr_batch::kGraphs; i++) {
        if (i > 0 && sizes[i] >= K) continue;
        itr_status s = capture_steps(b, sizes[i], b->graph[ng], err, errlen);
        if (s != ITR_OK) {
            drop_graphs(b);
            return s;
        }
        b->graph_steps[ng++] = sizes[i];
    }
    return ITR_OK;
}

// Cell-steps run between two control snapshots: completed steps plus the step in
// flight, each of nx*ny cells (the reference counts steps * cells, gpubench.rs:28).
double cell_steps_between(const itr_batch* b, const std::vector<CandCtl>& c0, const std::vector<CandCtl>& c1) {
    double cs = 0.0;
    for (size_t c = 0; c < c1.size(); c++) {
        const double n1 = (double)(c1[c].step + c1[c].in_flight), n0 = (double)(c0[c].step + c0[c].in_flight);
        cs += (n1 - n0) * (double)b->grids[c1[c].grid].d.nx * (double)b->grids[c1[c].grid].d.ny;
    }
    return cs;
}

// Simulated time advanced between two control snapshots, summed over candidates (the
// step in flight counted to its end time, as cell_steps_between counts it).
double sim_s_between(const std::vector<CandCtl>& c0, const std::vector<CandCtl>& c1) {
    double s = 0.0;
    for (size_t c = 0; c < c1.size(); c++) {
        const double t1 = c1[c].in_flight ? c1[c].t_new : c1[c].t, t0 = c0[c].in_flight ? c0[c].t_new : c0[c].t;
        s += t1 - t0;
    }
    return s;
}

}  // namespace

extern "C" {

itr_status itr_hip_create(int32_t device, itr_ctx** out, char* err, size_t errlen) {
    if (!out) return ITR_EINVAL;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_err(err, errlen, "no HIP device %d (%d present)", device, n);
        return ITR_DECLINED;
    }
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
        set_err(err, errlen, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
        return ITR_DECLINED;
    }
    // The static scratch of this device: mapped once, with the device current.
    void* sp = device < kMaxDevices ? g_scratch[device].load() : nullptr;
    if (device < kMaxDevices && !sp) {
        HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        void* expected = nullptr;
        if (!g_scratch[device].compare_exchange_strong(expected, sp)) sp = expected;
    }
    auto* c = new itr_ctx();
    c->device = device;
    c->cus = std::max(1, prop.multiProcessorCount);
    c->scratch = sp;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; i < kMaxLists - 1 && e == hipSuccess; i++) e = hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    for (int i = 0; i < kMaxLists - 1 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        itr_hip_destroy(c);
        set_err(err, errlen, "stream / event creation: %s", hipGetErrorString(e));
        return ITR_EDEVICE;
    }
    char buf[512];
    snprintf(buf, sizeof buf, "%s (%s, %d CUs, %.0f GiB HBM, HIP %d)", prop.name[0] ? prop.name : "AMD GPU",
             prop.gcnArchName,
             prop.multiProcessorCount, (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0), HIP_VERSION);
    c->info = buf;
    *out = c;
    return ITR_OK;
}

void itr_hip_destroy(itr_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < kMaxLists - 1; i++) {
        if (ctx->aux[i]) (void)hipStreamDestroy(ctx->aux[i]);
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* itr_hip_device_info(itr_ctx* ctx) { return ctx ? ctx->info.c_str() : ""; }

itr_status itr_hip_set_option(itr_ctx* ctx, const char* name, int64_t value) {
    if (!ctx || !name) return ITR_EINVAL;
    std::string n(name);
    if (n == "band_rows") ctx->band_rows = (int)value;
    else if (n == "steps_per_chunk") ctx->steps_per_chunk = (int)value;
    else if (n == "use_graph") ctx->use_graph = value != 0;
    else if (n == "min_lds") ctx->min_lds = value;
    else if (n == "adjoint_mode" && value >= 0 && value <= 2) ctx->adjoint_mode = (int)value;
    else if (n == "control_task" && value >= -1 && value <= 1) ctx->control_task = (int)value;
    else if (n == "adjoint_rows" && value >= 0 && value <= 1 << 20) ctx->adjoint_rows = (int)value;
    else if (n == "tail_rows" && value >= -1 && value <= 1 << 20) ctx->tail_rows = (int)value;
    else if (n == "tail_band_rows" && value >= 0 && value <= 1 << 20) ctx->tail_band_rows = (int)value;
    else if (n == "list_split" && value >= 0 && value <= kMaxLists) ctx->list_split = (int)value;
    else return ITR_EINVAL;
    return ITR_OK;
}

itr_status batch_create_impl(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                             itr_batch** out, int64_t log_steps, char* err, size_t errlen);

itr_status itr_hip_batch_create(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                                itr_batch** out, char* err, size_t errlen) {
    return batch_create_impl(ctx, p, grids, n_grids, out, 0, err, errlen);
}

}  // extern "C"

itr_status batch_create_impl(itr_ctx* ctx, const itr_solver_params* p_call, const itr_grid* grids, int32_t n_grids,
                             itr_batch** out, int64_t log_steps, char* err, size_t errlen) {
    const itr_solver_params* p = p_call;
    if (!ctx || !p || !grids || n_grids <= 0 || !out) {
        set_err(err, errlen, "null argument");
        return ITR_EINVAL;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // ---- validation and declines (lib.rs:155-171)
    for (int g = 0; g < n_grids; g++) {
        const itr_grid& G = grids[g];
        const itr_solver_params* q = G.params ? G.params : p;
        if (!(q->sync_interval_s > 0.0) || !(q->duration_s >= 0.0) || q->n_rain < 0 ||
            (q->n_rain > 0 && (!q->rain_t || !q->rain_r)) || q->n_locked < 0) {
            set_err(err, errlen, "grid %d: invalid solver parameters", g);
            return ITR_EINVAL;
        }
        if (G.nx <= 0 || G.ny <= 0 || G.n_cand < 0 || !G.wet || !G.gn2 || !G.ghost ||
            (!G.design && (!G.dzx || !G.dzy))) {
            set_err(err, errlen, "grid %d: missing arrays or empty shape", g);
            return ITR_EINVAL;
        }
        if (q->physics != ITR_PHYSICS_HLL && q->physics != ITR_PHYSICS_LOCAL_INERTIAL) {
            set_err(err, errlen, "grid %d: unknown physics %d", g, q->physics);
            return ITR_EINVAL;
        }
        const int max_cols = kMaxWaves * (q->physics == ITR_PHYSICS_LOCAL_INERTIAL ? kLiCols : kColsPerWave);
        if (G.nx > max_cols) {
            set_err(err, errlen, "grid %d: nx=%d exceeds %d columns", g, G.nx, max_cols);
            return ITR_DECLINED;
        }
        if (G.crest && !G.design && (!G.dzx_r || !G.dzy_r)) {
            set_err(err, errlen, "grid %d: crest mode needs dzx_r and dzy_r", g);
            return ITR_EINVAL;
        }
        for (int i = 0; i < 2 * G.nx + 2 * G.ny; i++) {
            const int code = G.ghost[i];
            if (code < 0 || code > 3) {
                set_err(err, errlen, "grid %d: unknown boundary code %d", g, code);
                return ITR_EINVAL;
            }
            if (code == 2 && (!G.ghost_value || (!G.z_level && !G.design))) {
                set_err(err, errlen, "grid %d: stage boundary cells need ghost_value and z_level", g);
                return ITR_EINVAL;
            }
            if (code == 3) {
                if (!G.ghost_value || !G.ghost_hydro || !G.hydro_start || G.n_hydro <= 0 || G.ghost_hydro[i] < 0 ||
                    G.ghost_hydro[i] >= G.n_hydro) {
                    set_err(err, errlen, "grid %d: inflow boundary cell %d needs ghost_value and a hydrograph", g, i);
                    return ITR_EINVAL;
                }
            }
        }
        if (G.n_hydro > 0) {
            if (!G.hydro_start || G.hydro_start[0] != 0 || (G.hydro_start[G.n_hydro] > 0 && (!G.hydro_t || !G.hydro_q))) {
                set_err(err, errlen, "grid %d: invalid hydrograph arrays", g);
                return ITR_EINVAL;
            }
            for (int h = 0; h < G.n_hydro; h++)
                if (G.hydro_start[h + 1] < G.hydro_start[h]) {
                    set_err(err, errlen, "grid %d: hydrograph offsets must not decrease", g);
                    return ITR_EINVAL;
                }
        }
        for (int q = 0; q < G.n_mon; q++) {
            if (G.mon_cells[q] >= (uint32_t)(G.nx * G.ny) || (q > 0 && G.mon_cells[q] <= G.mon_cells[q - 1])) {
                set_err(err, errlen, "grid %d: monitor cells must be sorted, unique and inside the grid", g);
                return ITR_EINVAL;
            }
        }
        if (G.baseline_max)
            for (int q = 0; q < G.n_guard; q++)
                if (G.guard_pos[q] < 0 || G.guard_pos[q] >= G.n_mon) {
                    set_err(err, errlen, "grid %d: guard position out of range", g);
                    return ITR_EINVAL;
                }
        if (G.replay) {
            const itr_replay& Rp = *G.replay;
            if (q->physics != ITR_PHYSICS_HLL || G.design) {
                set_err(err, errlen, "grid %d: subdomain replay runs HLL levels with prepared faces only", g);
                return ITR_DECLINED;
            }
            if (Rp.mode == 1) {
                if (Rp.c0 < 0 || Rp.r0 < 0 || Rp.w <= 0 || Rp.h <= 0 || Rp.c0 + Rp.w > G.nx || Rp.r0 + Rp.h > G.ny ||
                    Rp.ring_steps < 0 || Rp.band_syncs < 0 || (Rp.ring_steps > 0 && !Rp.ring) ||
                    (Rp.band_syncs > 0 && !Rp.band) || G.n_cand < 1) {
                    set_err(err, errlen, "grid %d: invalid replay recording (region, buffers)", g);
                    return ITR_EINVAL;
                }
            } else if (Rp.mode == 2) {
                if (Rp.w != G.nx || Rp.h != G.ny || Rp.ring_steps < 0 || Rp.band_syncs < 0 ||
                    (Rp.ring_steps > 0 && !Rp.ring) || (Rp.band_syncs > 0 && !Rp.band)) {
                    set_err(err, errlen, "grid %d: a replay grid is the w x h region with its tape", g);
                    return ITR_EINVAL;
                }
            } else {
                set_err(err, errlen, "grid %d: replay mode must be 1 (record) or 2 (replay)", g);
                return ITR_EINVAL;
            }
        }
        if (G.design) {
            const itr_design& Ds = *G.design;
            const int lv = std::max(Ds.level, 1);
            if ((lv >= 2) != (G.crest != 0)) {
                set_err(err, errlen, "grid %d: device designs need a level-1 signed or level >= 2 face-crest grid", g);
                return ITR_DECLINED;
            }
            if (lv >= 2 && (Ds.fine_nx <= 0 || Ds.fine_ny <= 0 || (Ds.fine_nx + lv - 1) / lv != G.nx ||
                            (Ds.fine_ny + lv - 1) / lv != G.ny)) {
                set_err(err, errlen, "grid %d: the design's fine grid %dx%d is not level %d of this %dx%d grid", g,
                        Ds.fine_nx, Ds.fine_ny, lv, G.nx, G.ny);
                return ITR_EINVAL;
            }
            const int fnx = lv >= 2 ? Ds.fine_nx : G.nx, fny = lv >= 2 ? Ds.fine_ny : G.ny;
            if (!Ds.prim_start || !Ds.bbox || !Ds.z_fine || !Ds.wall || G.n_cand > 65535 ||
                (G.n_cand > 0 && Ds.prim_start[G.n_cand] > 0 && !Ds.prim)) {
                set_err(err, errlen, "grid %d: incomplete design arrays", g);
                return ITR_EINVAL;
            }
            for (int c = 0; c < G.n_cand; c++) {
                const int32_t* bb = Ds.bbox + 4 * c;
                if (bb[0] < 0 || bb[1] < 0 || bb[2] < 0 || bb[3] < 0 || bb[0] + bb[2] > fnx || bb[1] + bb[3] > fny ||
                    Ds.prim_start[c + 1] < Ds.prim_start[c]) {
                    set_err(err, errlen, "grid %d: design window of candidate %d outside the grid", g, c);
                    return ITR_EINVAL;
                }
            }
        }
        if (q->has_initial_stage && !G.z_level && !G.design) {
            set_err(err, errlen, "grid %d: initial_stage needs z_level", g);
            return ITR_EINVAL;
        }
    }
    auto* b = new itr_batch();
    b->ctx = ctx;
    b->no_alloc = tl_arena != nullptr;
    auto fail = [&](itr_status s) {
        free_batch_mem(b);
        delete b;
        return s;
    };
    // ---- grids
    int total = 0;
    std::vector<CandCtl> ctl;
    std::vector<DevGrid> dgrids;
    b->grids.resize(n_grids);
    for (int g = 0; g < n_grids; g++) {
        const itr_grid& I = grids[g];
        GridHost& GH = b->grids[g];
        GH.in = I;
        DevGrid& D = GH.d;
        // ---- this grid's parameters (the call's, or an ensemble member's own)
        const itr_solver_params* p = I.params ? I.params : p_call;
        const double dur = p->duration_s, syn = p->sync_interval_s;
        const uint32_t n_sync = (uint32_t)std::ceil(dur / syn);
        {
            itr_status ps = set_params(p, GH.prm, GH.max_ledger, GH.allocs, err, errlen);
            if (ps != ITR_OK) return fail(ps);
            D.prm = GH.prm;
        }
        const int nx = I.nx, ny = I.ny, nc = I.n_cand;
        D.nx = nx;
        D.ny = ny;
        D.P = (nx + 1 + 63) / 64 * 64;   // >= nx + 1: column nx holds the east boundary face
        D.li = p->physics == ITR_PHYSICS_LOCAL_INERTIAL ? 1 : 0;
        D.waves = D.li ? (nx + kLiCols - 1) / kLiCols : (nx + kColsPerWave - 1) / kColsPerWave;
        D.tail_groups = D.li ? 1 : tail_group_count(nx, D.waves);
        const int R = choose_band_rows(ctx, grids, n_grids, g);
        D.R = R;
        D.nbands = (ny + R - 1) / R;
        D.cand0 = total;
        D.ncand = nc;
        D.nslot = std::max(nc, 1);
        D.n_mon = I.n_mon;
        D.n_guard = I.baseline_max ? I.n_guard : 0;
        D.crest = I.crest;
        D.wet_count = I.wet_count;
        D.plane = (size_t)ny * D.P;
        D.plane_y = (size_t)(ny + 1) * D.P;
        D.dx = I.dx;
        D.dy = I.dy;
        D.g = I.g;
        D.area = I.dx * I.dy;
        D.guard_tol = I.guard_tol;
        D.g32 = (float)I.g;
        D.g_half32 = (float)(0.5 * I.g);
        D.inv_dx32 = (float)(1.0 / I.dx);
        D.inv_dy32 = (float)(1.0 / I.dy);
        // Stage / Inflow ghosts: the stage depth of every code-2 cell per candidate
        // (boundary.rs:43, 66: f32(max(stage - z_level(interior neighbour), 0)), with the
        // candidate's own levels); inflow cells keep 1 / length and their hydrograph.
        const int nb = 2 * nx + 2 * ny;
        D.special = 0;
        for (int i = 0; i < nb; i++) D.special |= I.ghost[i] >= 2 ? 1 : 0;
        // Subdomain replay: a replay grid's ghosts are all the tape's (code 4).
        std::vector<int32_t> ghost_codes(I.ghost, I.ghost + nb);
        D.tape_mode = I.replay ? I.replay->mode : 0;
        if (D.tape_mode == 2) {
            std::fill(ghost_codes.begin(), ghost_codes.end(), 4);
            D.special = 1;
        }
        {
            const int rw = D.tape_mode == 2 ? nx : (I.replay ? I.replay->w : 0);
            const int rh = D.tape_mode == 2 ? ny : (I.replay ? I.replay->h : 0);
            D.roi_c0 = D.tape_mode == 1 ? I.replay->c0 : 0;
            D.roi_r0 = D.tape_mode == 1 ? I.replay->r0 : 0;
            D.roi_w = rw;
            D.roi_h = rh;
            D.n_ring = I.replay ? 2 * (rw + 2) + 2 * rh : 0;
            int nband = 0;
            for (int j = 1; j <= rh; j++)
                for (int i = 1; i <= rw; i++) nband += (i == 1 || j == 1 || i == rw || j == rh) ? 1 : 0;
            D.n_band = nband;
            D.tape_steps = I.replay ? I.replay->ring_steps : 0;
            D.tape_syncs = I.replay ? I.replay->band_syncs : 0;
            D.tol_h = I.replay ? I.replay->tol_h : 0.0;
            D.tol_q = I.replay ? I.replay->tol_q : 0.0;
        }
        // Candidates given as designs: their dz and level z on the device now; the Stage
        // ghosts and initial_stage below read the level z.
        DesignWork dwork;
        std::vector<double> dz_levels;
        const double* z_level = I.z_level;
        if (I.design) {
            itr_status ds = design_begin(ctx, I, D, dwork, err, errlen);
            if (ds == ITR_OK && (D.special || p->has_initial_stage)) {
                ds = design_levels(ctx, dwork, dz_levels, err, errlen);
                z_level = dz_levels.data();
            }
            if (ds != ITR_OK) return fail(ds);
        }
        std::vector<float> gstage;
        if (D.special) {
            gstage.assign((size_t)std::max(nc, 1) * nb, 0.0f);
            for (int k = 0; k < nc; k++)
                for (int i = 0; i < nb; i++) {
                    if (I.ghost[i] != 2) continue;
                    int r, c;
                    if (i < nx) { r = 0; c = i; }
                    else if (i < 2 * nx) { r = ny - 1; c = i - nx; }
                    else if (i < 2 * nx + ny) { r = i - 2 * nx; c = 0; }
                    else { r = i - 2 * nx - ny; c = nx - 1; }
                    const double zl = z_level[(size_t)k * nx * ny + (size_t)r * nx + c];
                    gstage[(size_t)k * nb + i] = (float)std::fmax(I.ghost_value[i] - zl, 0.0);
                }
        }
        const size_t plane = D.plane;
        // Host staging of the shared planes only: faces and initial states go to the device
        // as pitched copies from the caller's arrays (and zero fills).
        const size_t plane_y = D.plane_y;
        std::vector<float> wet(plane, 0.0f), gn2(plane, 0.0f);
        GH.mon.assign(I.mon_cells, I.mon_cells + I.n_mon);
        std::vector<uint32_t> mon_off((size_t)std::max(I.n_mon, 1), 0u);
        for (int q = 0; q < I.n_mon; q++)
            mon_off[q] = (I.mon_cells[q] / (uint32_t)nx) * (uint32_t)D.P + I.mon_cells[q] % (uint32_t)nx;
        std::vector<int32_t> mon_row((size_t)ny + 1, 0);   // monitor cells are sorted row-major
        for (int r = 0, q = 0; r <= ny; r++) {
            while (q < I.n_mon && I.mon_cells[q] < (uint32_t)r * (uint32_t)nx) q++;
            mon_row[r] = q;
        }
        for (int r = 0; r < ny; r++)
            for (int c = 0; c < nx; c++) {
                wet[(size_t)r * D.P + c] = I.wet[(size_t)r * nx + c];
                gn2[(size_t)r * D.P + c] = I.gn2[(size_t)r * nx + c];
            }
        std::vector<RowAcc> rowacc((size_t)std::max(nc, 1) * ny);   // parity 0 half
        std::memset(rowacc.data(), 0, rowacc.size() * sizeof(RowAcc));
        // init_state (solver.rs:287-312) and the initial CFL rate (solver.rs:315-332): one
        // initial depth field per candidate with initial_stage (it depends on the levels),
        // else one for all.
        const int n_h0 = p->has_initial_stage ? std::max(nc, 1) : 1;
        std::vector<std::vector<float>> h0s(n_h0);
        std::vector<double> s0s(n_h0, 0.0), v0s(n_h0, 0.0);
        std::vector<uint32_t> wet_rows((size_t)n_h0 * ny, 0u);
        for (int k = 0; k < n_h0; k++) {
            std::vector<float>& h0 = h0s[k];
            h0.assign((size_t)nx * ny, 0.0f);
            double s0 = 0.0;
            for (int r = 0; r < ny; r++)
                for (int c = 0; c < nx; c++) {
                    const size_t q = (size_t)r * nx + c;
                    if (I.wet[q] == 0.0f) continue;
                    double d = p->initial_depth;
                    if (p->has_initial_stage) d = std::fmax(d, p->initial_stage - I.z_ref - z_level[(size_t)k * nx * ny + q]);
                    float hv = (float)std::fmax(d, 0.0);
                    if (hv == 0.0f) hv = 0.0f;
                    h0[q] = hv;
                    const double hd = (double)hv;
                    if (D.li) {   // initial_smax_li (solver.rs:334-344): the deepest cell
                        s0 = std::fmax(s0, hd);
                        continue;
                    }
                    if (hd <= p->h_dry) continue;
                    const double cc = std::sqrt(I.g * hd);
                    const double rate = (0.0 / hd + cc) / I.dx + (0.0 / hd + cc) / I.dy;
                    s0 = std::fmax(s0, rate);
                }
            for (int r = 0; r < ny; r++) {
                uint32_t w = 0;
                for (int c = 0; c < nx; c++) w |= h0[(size_t)r * nx + c] > 0.0f ? 1u : 0u;
                wet_rows[(size_t)k * ny + r] = w;
            }
            if (D.li) s0 = std::sqrt(I.g * s0) * 0.7 / std::fmin(I.dx, I.dy);
            s0s[k] = s0;
            v0s[k] = host_volume(h0, nx, ny, D.area);
        }
        for (int k = 0; k < nc; k++) {
            const int hk = p->has_initial_stage ? k : 0;
            const double s0 = s0s[hk];
            for (int r = 0; r < ny; r++) rowacc[(size_t)k * ny + r].wet = wet_rows[(size_t)hk * ny + r];
            CandCtl C;
            std::memset(&C, 0, sizeof C);
            C.grid = g;
            C.local = k;
            C.v0 = v0s[hk];
            C.t = 0.0;
            C.s_cur = s0;
            C.dt_cfl = s0 > 0.0 ? std::fmin(p->cfl / s0, p->dt_max_s) : p->dt_max_s;
            C.dt_min = 1.79769313486231570815e+308;
            C.dt_fallback_t = -1.0;
            C.flags = 0;
            if (!(0.0 < dur && 0u < n_sync)) C.flags = F_DONE;
            if (p->max_steps < 0) C.flags = F_DONE;
            ctl.push_back(C);
        }
        total += nc;
        // device buffers
        float *d_state, *d_dzx, *d_dzy, *d_dzx2 = nullptr, *d_dzy2 = nullptr, *d_wet, *d_gn2, *d_monmax, *d_ybnd, *d_sbuf,
            *d_fixbuf;
        double *d_montmax, *d_ledger, *d_base = nullptr;
        int32_t *d_monrow, *d_ghost, *d_gpos = nullptr;
        uint32_t* d_monoff;
        uint32_t* d_moncells;
        RowAcc* d_rowacc;
        const size_t ncx = (size_t)std::max(nc, 1);
        itr_status s;
#define ALLOC(ptr, n) \
    if ((s = dev_alloc(GH.allocs, &ptr, (n), err, errlen)) != ITR_OK) return fail(s);
        ALLOC(d_state, ncx * 6 * plane);
        ALLOC(d_dzx, ncx * plane);
        ALLOC(d_dzy, ncx * plane_y);
        if (I.crest) {
            ALLOC(d_dzx2, ncx * plane);
            ALLOC(d_dzy2, ncx * plane_y);
        }
        ALLOC(d_wet, plane);
        ALLOC(d_gn2, plane);
        ALLOC(d_monrow, (size_t)ny + 1);
        ALLOC(d_monoff, mon_off.size());
        ALLOC(d_ghost, (size_t)(2 * nx + 2 * ny));
        ALLOC(d_monmax, 2 * ncx * std::max(I.n_mon, 1));
        ALLOC(d_montmax, 2 * ncx * std::max(I.n_mon, 1));
        ALLOC(d_rowacc, 2 * ncx * ny);
        ALLOC(d_ybnd, 2 * ncx * 2 * nx);
        ALLOC(d_sbuf, D.li ? 1 : ncx * plane);
        ALLOC(d_fixbuf, ncx * plane * (D.li ? 4 : 1));   // f64 fix values by parity on local-inertial grids
        float* d_lif = nullptr;
        if (D.li) ALLOC(d_lif, 2 * ncx * (size_t)(ny + nx));
        float *d_tring = nullptr, *d_tband = nullptr;
        if (D.tape_mode) {
            ALLOC(d_tring, std::max<size_t>((size_t)D.tape_steps * D.n_ring * 3, 1));
            ALLOC(d_tband, std::max<size_t>((size_t)D.tape_syncs * D.n_band * 3, 1));
        }
        // Sync snapshot of a locked-dt grid (the adaptive restart, solver.rs:446-480).
        float *d_snap_state = nullptr, *d_snap_mon = nullptr, *d_snap_lif = nullptr;
        double* d_snap_tmax = nullptr;
        RowAcc* d_snap_rows = nullptr;
        CandCtl* d_snap_ctl = nullptr;
        if (GH.prm.locked) {
            ALLOC(d_snap_state, ncx * 3 * plane);
            ALLOC(d_snap_mon, ncx * std::max(I.n_mon, 1));
            ALLOC(d_snap_tmax, ncx * std::max(I.n_mon, 1));
            ALLOC(d_snap_rows, ncx * ny);
            ALLOC(d_snap_ctl, ncx);
            if (D.li) ALLOC(d_snap_lif, ncx * (size_t)(ny + nx));
        }
        ALLOC(d_ledger, ncx * GH.max_ledger * 8);
        ALLOC(d_moncells, (size_t)std::max(I.n_mon, 1));
        StepRec* d_steplog = nullptr;
        if (log_steps > 0) ALLOC(d_steplog, ncx * (size_t)log_steps);
        float* d_gstage = nullptr;
        double *d_ginv = nullptr, *d_hyd_t = nullptr, *d_hyd_q = nullptr;
        int32_t *d_ghydro = nullptr, *d_hstart = nullptr;
        const int n_hydro = D.special && I.n_hydro > 0 ? I.n_hydro : 0;
        const size_t n_pts = n_hydro > 0 ? (size_t)I.hydro_start[n_hydro] : 0;
        if (D.special) {
            ALLOC(d_gstage, gstage.size());
            ALLOC(d_ginv, (size_t)nb);
            ALLOC(d_ghydro, (size_t)nb);
            ALLOC(d_hstart, (size_t)n_hydro + 1);
            ALLOC(d_hyd_t, std::max<size_t>(n_pts, 1));
            ALLOC(d_hyd_q, std::max<size_t>(n_pts, 1));
        }
        if (I.baseline_max) {
            ALLOC(d_base, (size_t)std::max(I.n_mon, 1));
            ALLOC(d_gpos, (size_t)std::max(I.n_guard, 1));
        }
#undef ALLOC
        auto up = [&](void* dst, const void* src, size_t bytes) -> bool {
            return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
        };
        // Pitched planes go through one reusable host plane: rows of P floats holding
        // columns c0 .. c1-1 of the caller's [rows][src_w] block (everything else 0).
        std::vector<float> pit(std::max(plane, plane_y));
        auto up_plane = [&](float* dst, const float* src, int src_w, int c0, int c1, int r0, int r1, int rows) -> bool {
            float* q = pit.data();
            for (int r = 0; r < rows; r++) {
                float* row = q + (size_t)r * D.P;
                if (r < r0 || r >= r1) {
                    std::fill(row, row + D.P, 0.0f);
                    continue;
                }
                std::fill(row, row + c0, 0.0f);
                std::memcpy(row + c0, src + (size_t)r * src_w + c0, sizeof(float) * (size_t)std::max(c1 - c0, 0));
                std::fill(row + std::max(c1, c0), row + D.P, 0.0f);
            }
            return hipMemcpy(dst, q, sizeof(float) * (size_t)rows * D.P, hipMemcpyHostToDevice) == hipSuccess;
        };
        bool ok = hipMemset(d_state, 0, ncx * 6 * plane * 4) == hipSuccess;
        if (I.design)
            ok = ok && hipMemset(d_dzx, 0, ncx * plane * 4) == hipSuccess &&
                 hipMemset(d_dzy, 0, ncx * plane_y * 4) == hipSuccess &&
                 (!I.crest || (hipMemset(d_dzx2, 0, ncx * plane * 4) == hipSuccess &&
                               hipMemset(d_dzy2, 0, ncx * plane_y * 4) == hipSuccess));
        for (int k = 0; ok && k < nc; k++) {
            // initial depth (parity 0, h plane): uploaded once when every candidate shares it
            if (p->has_initial_stage || k == 0) ok = up_plane(d_state + (size_t)k * 6 * plane, h0s[p->has_initial_stage ? k : 0].data(), nx, 0, nx, 0, ny, ny);
            else ok = hipMemcpy(d_state + (size_t)k * 6 * plane, d_state, sizeof(float) * plane, hipMemcpyDeviceToDevice) == hipSuccess;
            if (I.design) continue;   // faces built on the device
            // interior faces only: x-faces 1 .. nx-1, y-face rows 1 .. ny-1 (boundary faces 0);
            // a replay grid keeps its boundary faces (they join the region to its ring)
            const int f0 = D.tape_mode == 2 ? 0 : 1, fx1 = D.tape_mode == 2 ? nx + 1 : nx, fy1 = D.tape_mode == 2 ? ny + 1 : ny;
            ok = ok && up_plane(d_dzx + (size_t)k * plane, I.dzx + (size_t)k * ny * (nx + 1), nx + 1, f0, fx1, 0, ny, ny) &&
                 up_plane(d_dzy + (size_t)k * plane_y, I.dzy + (size_t)k * (ny + 1) * nx, nx, 0, nx, f0, fy1, ny + 1);
            if (I.crest)
                ok = ok && up_plane(d_dzx2 + (size_t)k * plane, I.dzx_r + (size_t)k * ny * (nx + 1), nx + 1, f0, fx1, 0, ny, ny) &&
                     up_plane(d_dzy2 + (size_t)k * plane_y, I.dzy_r + (size_t)k * (ny + 1) * nx, nx, 0, nx, f0, fy1, ny + 1);
        }
        ok = ok && up(d_wet, wet.data(), plane * 4) &&
                  up(d_gn2, gn2.data(), plane * 4) && up(d_monrow, mon_row.data(), mon_row.size() * 4) && up(d_monoff, mon_off.data(), mon_off.size() * 4) &&
                  up(d_ghost, ghost_codes.data(), sizeof(int32_t) * (2 * nx + 2 * ny)) &&
                  up(d_rowacc, rowacc.data(), rowacc.size() * sizeof(RowAcc)) &&
                  up(d_moncells, I.mon_cells, sizeof(uint32_t) * I.n_mon);
        if (I.baseline_max)
            ok = ok && up(d_base, I.baseline_max, sizeof(double) * I.n_mon) &&
                 up(d_gpos, I.guard_pos, sizeof(int32_t) * I.n_guard);
        if (D.special) {
            std::vector<double> ginv(nb, 0.0);
            std::vector<int32_t> ghydro(nb, 0), hstart(n_hydro + 1, 0);
            for (int i = 0; i < nb; i++)
                if (I.ghost[i] == 3) {
                    ginv[i] = I.ghost_value[i];
                    ghydro[i] = I.ghost_hydro[i];
                }
            if (n_hydro > 0)   // (grids with stage or replayed ghosts only have no hydrographs)
                for (int h = 0; h <= n_hydro; h++) hstart[h] = I.hydro_start[h];
            ok = ok && up(d_gstage, gstage.data(), gstage.size() * 4) && up(d_ginv, ginv.data(), nb * 8) &&
                 up(d_ghydro, ghydro.data(), nb * 4) && up(d_hstart, hstart.data(), hstart.size() * 4) &&
                 up(d_hyd_t, I.hydro_t, n_pts * 8) && up(d_hyd_q, I.hydro_q, n_pts * 8);
        }
        ok = ok && hipMemset(d_monmax, 0, 2 * ncx * std::max(I.n_mon, 1) * 4) == hipSuccess &&
             hipMemset(d_montmax, 0, 2 * ncx * std::max(I.n_mon, 1) * 8) == hipSuccess &&
             hipMemset(d_ybnd, 0, 2 * ncx * 2 * nx * 4) == hipSuccess &&
             hipMemset(d_rowacc + ncx * ny, 0, ncx * ny * sizeof(RowAcc)) == hipSuccess &&
             (!d_lif || hipMemset(d_lif, 0, 2 * ncx * (size_t)(ny + nx) * 4) == hipSuccess);
        if (D.tape_mode == 2)
            ok = ok && up(d_tring, I.replay->ring, sizeof(float) * (size_t)D.tape_steps * D.n_ring * 3) &&
                 up(d_tband, I.replay->band, sizeof(float) * (size_t)D.tape_syncs * D.n_band * 3);
        if (d_snap_state) {   // the first snapshot: the initial state and control
            for (int k = 0; ok && k < nc; k++)
                ok = hipMemcpy(d_snap_state + (size_t)k * 3 * plane, d_state + (size_t)k * 6 * plane, sizeof(float) * 3 * plane,
                               hipMemcpyDeviceToDevice) == hipSuccess;
            ok = ok && hipMemset(d_snap_mon, 0, ncx * std::max(I.n_mon, 1) * 4) == hipSuccess &&
                 hipMemset(d_snap_tmax, 0, ncx * std::max(I.n_mon, 1) * 8) == hipSuccess &&
                 up(d_snap_rows, rowacc.data(), ncx * ny * sizeof(RowAcc)) &&
                 (nc == 0 || up(d_snap_ctl, ctl.data() + (ctl.size() - (size_t)nc), (size_t)nc * sizeof(CandCtl))) &&
                 (!d_snap_lif || hipMemset(d_snap_lif, 0, ncx * (size_t)(ny + nx) * 4) == hipSuccess);
        }
        if (!ok) {
            set_err(err, errlen, "upload of grid %d failed", g);
            return fail(ITR_EDEVICE);
        }
        if (I.design) {
            itr_status ms = design_finish(ctx, I, dwork, d_dzx, d_dzy, d_dzx2, d_dzy2, err, errlen);
            if (ms != ITR_OK) return fail(ms);
        }
        D.state = d_state;
        D.dzx = d_dzx;
        D.dzy = d_dzy;
        D.dzx2 = d_dzx2;
        D.dzy2 = d_dzy2;
        D.wet = d_wet;
        D.gn2 = d_gn2;
        D.mon_row = d_monrow;
        D.mon_off = d_monoff;
        D.ghost = d_ghost;
        D.mon_max = d_monmax;
        D.mon_tmax = d_montmax;
        D.rowacc = d_rowacc;
        D.ybnd = d_ybnd;
        D.sbuf = d_sbuf;
        D.fixbuf = d_fixbuf;
        D.fix64 = D.li ? (double*)d_fixbuf : nullptr;
        D.lif = d_lif;
        D.snap_state = d_snap_state;
        D.snap_mon = d_snap_mon;
        D.snap_tmax = d_snap_tmax;
        D.snap_rows = d_snap_rows;
        D.snap_lif = d_snap_lif;
        D.snap_ctl = d_snap_ctl;
        D.tape_ring = d_tring;
        D.tape_band = d_tband;
        D.ledger = d_ledger;
        D.baseline_max = d_base;
        D.guard_pos = d_gpos;
        D.mon_cells = d_moncells;
        D.steplog = d_steplog;
        D.n_hydro = n_hydro;
        D.gstage = d_gstage;
        D.ginv = d_ginv;
        D.ghydro = d_ghydro;
        D.hstart = d_hstart;
        D.hyd_t = d_hyd_t;
        D.hyd_q = d_hyd_q;
        D.max_log = log_steps > 0 ? (int32_t)std::min<int64_t>(log_steps, 1 << 30) : 0;
        dgrids.push_back(D);
    }
    b->ncand = total;
    std::vector<int4> tasks;
    assign_lists(b, dgrids);
    const size_t lds_max = fill_lists(b, dgrids, tasks);
    if (lds_max > 160 * 1024 - 1024) {   // static shared memory of the kernel: < 1 KiB
        set_err(err, errlen, "band of %zu bytes of LDS does not fit; lower band_rows", lds_max);
        return fail(ITR_DECLINED);
    }
    std::vector<void*> misc;
    itr_status s;
    DevGrid* d_grids;
    int4* d_tasks;
    CandCtl* d_snap;
    StepAcc* d_acc;
    uint32_t* d_status;
    if ((s = dev_alloc(misc, &d_grids, 2 * dgrids.size(), err, errlen)) != ITR_OK) return fail(s);   // room for shadows
    b->d_grids = d_grids;
    size_t task_cap = 0;   // room for any band height (re-tiling never reallocates)
    for (auto& D : dgrids) task_cap += (size_t)std::max(D.ncand, 0) * (size_t)(D.ny + 1);
    if ((s = dev_alloc(misc, &d_tasks, std::max(task_cap, tasks.size()), err, errlen)) != ITR_OK) return fail(s);
    b->d_tasks = d_tasks;
    if ((s = dev_alloc(misc, &d_snap, 2 * std::max<size_t>(ctl.size(), 1), err, errlen)) != ITR_OK) return fail(s);
    b->d_snap = d_snap;
    if ((s = dev_alloc(misc, &d_acc, 4 * std::max<size_t>(ctl.size(), 1), err, errlen)) != ITR_OK) return fail(s);
    b->d_acc = d_acc;
    b->prm = b->grids[0].prm;
    if ((s = dev_alloc(misc, &d_status, std::max(total, 1), err, errlen)) != ITR_OK) return fail(s);
    b->d_status = d_status;
    if (b->no_alloc) {
        b->h_status_host.assign((size_t)std::max(total, 1), 0u);
        b->h_status = b->h_status_host.data();
    } else if (hipHostMalloc((void**)&b->h_status, sizeof(uint32_t) * std::max(total, 1), hipHostMallocDefault) !=
               hipSuccess) {
        set_err(err, errlen, "pinned allocation failed");
        return fail(ITR_ENOMEM);
    }
    const std::vector<DevGrid> dall = device_grids(dgrids, b->tails);
    bool ok = hipMemcpy(d_grids, dall.data(), dall.size() * sizeof(DevGrid), hipMemcpyHostToDevice) == hipSuccess &&
              (tasks.empty() || hipMemcpy(d_tasks, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice) == hipSuccess) &&
              (ctl.empty() || (hipMemcpy(d_snap, ctl.data(), ctl.size() * sizeof(CandCtl), hipMemcpyHostToDevice) == hipSuccess &&
                               hipMemcpy(d_snap + ctl.size(), ctl.data(), ctl.size() * sizeof(CandCtl),
                                         hipMemcpyHostToDevice) == hipSuccess)) &&
              hipMemset(d_acc, 0, 4 * std::max<size_t>(ctl.size(), 1) * sizeof(StepAcc)) == hipSuccess;
    if (!ok) {
        set_err(err, errlen, "control upload failed");
        return fail(ITR_EDEVICE);
    }
    b->args.grids = d_grids;
    b->args.tasks = d_tasks;
    b->args.snap = d_snap;
    b->args.acc = d_acc;
    b->args.status = d_status;
    b->args.ntasks = b->ntasks;
    b->args.ncand = total;
    bind_lists(b);
    if (total > 0) {
        hipLaunchKernelGGL(itr_init_kernel, dim3(total), dim3(64), 0, ctx->stream, b->args);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    b->stats.band_rows = dgrids.empty() ? 0 : dgrids[0].R;
    b->stats.waves_per_task = 1;
    for (const LaunchList& L : b->lists) b->stats.waves_per_task = std::max(b->stats.waves_per_task, L.block / kWave);
    b->stats.launch_lists = (int32_t)b->lists.size();
    *out = b;
    return ITR_OK;
}

extern "C" {

itr_status itr_hip_batch_run(itr_batch* b, itr_sync_fn on_sync, void* user, char* err, size_t errlen) {
    if (!b) return ITR_EINVAL;
    itr_ctx* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const int has_cb = on_sync ? 1 : 0;
    if (b->prm.has_cb != has_cb) {
        itr_status ps = set_callback_flag(b, has_cb, err, errlen);
        if (ps != ITR_OK) return ps;
    }
    int K = ctx->steps_per_chunk > 0 ? ctx->steps_per_chunk : 32;
    K = (K + 3) / 4 * 4;   // chunks keep the launch sequence aligned to the 4-deep ring
    // (a no-allocation batch launches directly and is not timed on the device: graph
    // instantiation and events are runtime allocations)
    const bool graph = ctx->use_graph != 0 && !b->no_alloc;
    const bool timed = !b->no_alloc;
    b->stats.steps_per_chunk = K;
    b->stats.use_graph = graph ? 1 : 0;
    b->stats.launches = 0;
    b->stats.syncs = 0;
    if (b->ncand == 0 || b->ntasks == 0) return ITR_OK;
    itr_status s = prepare_graph(b, graph ? K : 0, err, errlen);
    if (s != ITR_OK) return s;
    std::vector<CandCtl> c0;
    if ((s = read_ctl(b, c0, err, errlen)) != ITR_OK) return s;
    EventPair ev;
    if (timed) HIP_TRY(ev.create());
    const auto t0 = std::chrono::steady_clock::now();
    if (timed) HIP_TRY(hipEventRecord(ev.e[0], ctx->stream));
    for (;;) {
        // Complete every in-flight step first: a chunk always ends on the launch grid.
        if ((s = launch_steps(b, K, err, errlen)) != ITR_OK) return s;
        b->stats.launches += (uint64_t)K;
        HIP_TRY(hipMemcpyAsync(b->h_status, b->d_status, sizeof(uint32_t) * b->ncand, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        b->stats.syncs += 1;
        bool any_paused = false;
        for (int c = 0; c < b->ncand; c++)
            if (b->h_status[c] & F_PAUSED) any_paused = true;
        if (any_paused && on_sync) {
            std::vector<CandCtl> snap;
            if ((s = read_ctl(b, snap, err, errlen)) != ITR_OK) return s;
            for (int c = 0; c < b->ncand; c++) {
                if (!(b->h_status[c] & F_PAUSED)) continue;
                CandCtl& C = snap[c];
                GridHost& GH = b->grids[C.grid];
                const int nm = GH.d.n_mon;
                std::vector<float> mm((size_t)std::max(nm, 1));
                if (nm > 0)
                    HIP_TRY(hipMemcpy(mm.data(), GH.d.mon_max + slot(GH.d, C.parity, C.local) * nm, sizeof(float) * nm,
                                      hipMemcpyDeviceToHost));
                std::vector<double> md(mm.begin(), mm.begin() + nm);
                double av = 0.0;
                const int32_t r = on_sync(user, C.grid, C.local, C.t, C.step, C.sync_index, md.data(), nm, &av);
                C.flags &= ~F_PAUSED;
                if (r != 0) {
                    C.flags |= F_DONE;
                    C.abort_kind = 2;
                    C.abort_value = av;
                    C.abort_sync = (int)C.sync_index;
                }
                HIP_TRY(hipMemcpy(cur_ctl_dev(b, c), &C, sizeof(CandCtl), hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy((uint32_t*)b->d_status + c, &C.flags, sizeof(uint32_t), hipMemcpyHostToDevice));
                b->h_status[c] = C.flags;
            }
        }
        bool all_done = true;
        for (int c = 0; c < b->ncand; c++)
            if (!(b->h_status[c] & F_DONE)) all_done = false;
        if (all_done) break;
    }
    float ms = 0.0f;
    if (timed) {
        HIP_TRY(hipEventRecord(ev.e[1], ctx->stream));
        HIP_TRY(hipEventSynchronize(ev.e[1]));
        HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    } else {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    const auto t1 = std::chrono::steady_clock::now();
    b->stats.device_ms = ms;
    b->stats.wall_s = std::chrono::duration<double>(t1 - t0).count();
    std::vector<CandCtl> c1;
    if ((s = read_ctl(b, c1, err, errlen)) != ITR_OK) return s;
    b->stats.cell_steps = cell_steps_between(b, c0, c1);
    b->stats.sim_s = sim_s_between(c0, c1);
    return ITR_OK;
}

itr_status itr_hip_batch_advance(itr_batch* b, int64_t n_steps, char* err, size_t errlen) {
    if (!b || n_steps < 0) return ITR_EINVAL;
    itr_ctx* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    if (b->prm.has_cb) {
        itr_status ps = set_callback_flag(b, 0, err, errlen);
        if (ps != ITR_OK) return ps;
    }
    int K = ctx->steps_per_chunk > 0 ? ctx->steps_per_chunk : 32;
    K = (K + 3) / 4 * 4;
    itr_status s = prepare_graph(b, ctx->use_graph ? K : 0, err, errlen);
    if (s != ITR_OK) return s;
    std::vector<CandCtl> c0, c1;
    if ((s = read_ctl(b, c0, err, errlen)) != ITR_OK) return s;
    EventPair ev;
    HIP_TRY(ev.create());
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(ev.e[0], ctx->stream));
    if (b->ntasks > 0 && n_steps > 0)
        if ((s = launch_steps(b, n_steps, err, errlen)) != ITR_OK) return s;
    HIP_TRY(hipEventRecord(ev.e[1], ctx->stream));
    HIP_TRY(hipEventSynchronize(ev.e[1]));
    const auto t1 = std::chrono::steady_clock::now();
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    if ((s = read_ctl(b, c1, err, errlen)) != ITR_OK) return s;
    HIP_TRY(hipMemcpy(b->h_status, b->d_status, sizeof(uint32_t) * b->ncand, hipMemcpyDeviceToHost));
    b->stats.device_ms = ms;
    b->stats.wall_s = std::chrono::duration<double>(t1 - t0).count();
    b->stats.cell_steps = cell_steps_between(b, c0, c1);
    b->stats.sim_s = sim_s_between(c0, c1);
    b->stats.launches = (uint64_t)n_steps;
    b->stats.syncs = 1;
    b->stats.steps_per_chunk = K;
    b->stats.use_graph = ctx->use_graph ? 1 : 0;
    return ITR_OK;
}

itr_status itr_hip_batch_results(itr_batch* b, itr_run_out* outs, char* err, size_t errlen) {
    if (!b || (!outs && b->ncand > 0)) return ITR_EINVAL;
    HIP_TRY(hipSetDevice(b->ctx->device));
    std::vector<CandCtl> cc;
    itr_status s = read_ctl(b, cc, err, errlen);
    if (s != ITR_OK) return s;
    for (int c = 0; c < b->ncand; c++) {
        const CandCtl& C = cc[c];
        const GridHost& GH = b->grids[C.grid];
        const DevGrid& D = GH.d;
        itr_run_out& o = outs[c];
        const int nx = D.nx, ny = D.ny, nm = D.n_mon;
        o.t_end = C.t;
        o.steps = C.step;
        o.cell_updates = C.cell_updates;
        o.rows_skipped = C.rows_skipped;
        o.dt_min = C.dt_min == 1.79769313486231570815e+308 ? 0.0 : C.dt_min;
        o.dt_max = C.dt_max;
        o.dt_mean = C.step > 0 ? C.dt_sum / (double)C.step : 0.0;
        o.lts = C.lts;
        o.v0 = C.v0;
        o.bad = C.abort_kind == 3 ? 1 : 0;
        o.aborted = C.abort_kind;
        o.abort_cell = C.abort_cell;
        o.abort_sync = C.abort_sync;
        o.abort_value = C.abort_value;
        o.dt_fallback_t = C.dt_fallback_t;
        if (D.tape_mode == 1 && C.local == 0) {   // the recorded tape
            const itr_replay& Rp = *GH.in.replay;
            const int64_t ns = std::min<int64_t>((int64_t)C.step, D.tape_steps);
            const int64_t nsy = std::min<int64_t>((int64_t)C.sync_index, D.tape_syncs);
            if (ns > 0) HIP_TRY(hipMemcpy(Rp.ring, D.tape_ring, sizeof(float) * (size_t)ns * D.n_ring * 3, hipMemcpyDeviceToHost));
            if (nsy > 0) HIP_TRY(hipMemcpy(Rp.band, D.tape_band, sizeof(float) * (size_t)nsy * D.n_band * 3, hipMemcpyDeviceToHost));
            if (Rp.n_steps_out) *Rp.n_steps_out = (int64_t)C.step;
            if (Rp.n_syncs_out) *Rp.n_syncs_out = (int64_t)C.sync_index;
        }
        o.sync_count = (int32_t)C.sync_index;
        const int nl = (int)std::min<uint32_t>(C.n_ledger, (uint32_t)GH.max_ledger);
        std::vector<double> led((size_t)std::max(nl, 1) * 8);
        if (nl > 0)
            HIP_TRY(hipMemcpy(led.data(), D.ledger + (size_t)C.local * GH.max_ledger * 8, sizeof(double) * 8 * nl,
                              hipMemcpyDeviceToHost));
        const double gross = C.v0 + C.acc_rain + C.inflow;
        const double last_res = nl > 0 ? led[(size_t)(nl - 1) * 8 + 7] : 0.0;
        o.rel_residual = std::fabs(last_res) / std::fmax(gross, 1e-300);
        o.n_ledger = nl;
        if (o.ledger) {
            const int m = std::min(nl, o.max_ledger);
            std::memcpy(o.ledger, led.data(), sizeof(double) * 8 * (size_t)m);
        }
        if (nm > 0 && (o.monitor_max || o.monitor_tmax)) {
            std::vector<float> mm(nm);
            std::vector<double> mt(nm);
            const size_t off = slot(D, C.parity, C.local) * nm;
            HIP_TRY(hipMemcpy(mm.data(), D.mon_max + off, sizeof(float) * nm, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(mt.data(), D.mon_tmax + off, sizeof(double) * nm, hipMemcpyDeviceToHost));
            for (int q = 0; q < nm; q++) {
                if (o.monitor_max) o.monitor_max[q] = (double)mm[q];
                if (o.monitor_tmax) o.monitor_tmax[q] = mt[q];
            }
        }
        std::vector<float> pl(3 * D.plane);
        HIP_TRY(hipMemcpy(pl.data(), D.state + (size_t)C.local * 6 * D.plane + (size_t)C.parity * 3 * D.plane,
                          sizeof(float) * 3 * D.plane, hipMemcpyDeviceToHost));
        std::vector<float> hh((size_t)nx * ny);
        for (int r = 0; r < ny; r++)
            for (int col = 0; col < nx; col++) {
                const size_t q = (size_t)r * D.P + col, k = (size_t)r * nx + col;
                hh[k] = pl[q];
                if (o.depth) o.depth[k] = pl[q];
                if (o.qx) o.qx[k] = pl[D.plane + q];
                if (o.qy) o.qy[k] = pl[2 * D.plane + q];
            }
        o.volume_m3 = host_volume(hh, nx, ny, D.area);
    }
    return ITR_OK;
}

itr_status itr_hip_batch_stats(itr_batch* b, itr_stats* out) {
    if (!b || !out) return ITR_EINVAL;
    *out = b->stats;
    return ITR_OK;
}

void itr_hip_batch_destroy(itr_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    free_batch_mem(b);
    delete b;
}

itr_status itr_hip_run_batch(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                             itr_sync_fn on_sync, void* user, itr_run_out* outs, char* err, size_t errlen) {
    itr_batch* b = nullptr;
    itr_status s = itr_hip_batch_create(ctx, p, grids, n_grids, &b, err, errlen);
    if (s != ITR_OK) return s;
    s = itr_hip_batch_run(b, on_sync, user, err, errlen);
    if (s == ITR_OK) s = itr_hip_batch_results(b, outs, err, errlen);
    itr_hip_batch_destroy(b);
    return s;
}

itr_status itr_hip_run_batch_noalloc(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                                     itr_sync_fn on_sync, void* user, itr_run_out* outs, char* err, size_t errlen) {
    if (!ctx) {
        set_err(err, errlen, "null argument");
        return ITR_EINVAL;
    }
    if (!ctx->scratch) {
        set_err(err, errlen, "device %d has no static scratch mapped", ctx->device);
        return ITR_EDEVICE;
    }
    // Every device buffer of the batch comes from the static scratch; the calls of one
    // device share it, so calls that overlap on one device race.
    Arena A;
    A.base = (char*)ctx->scratch;
    A.cap = kScratchBytes;
    ArenaScope scope(&A);
    itr_batch* b = nullptr;
    itr_status s = batch_create_impl(ctx, p, grids, n_grids, &b, 0, err, errlen);
    if (s != ITR_OK) return s;
    s = itr_hip_batch_run(b, on_sync, user, err, errlen);
    if (s == ITR_OK) s = itr_hip_batch_results(b, outs, err, errlen);
    itr_hip_batch_destroy(b);
    return s;
}

itr_status itr_hip_objective_j(const itr_objective* obj, const itr_earthwork* earth, const double* monitor_max,
                               int32_t n_mon, double* j_out, double* j_risk_out, double* peaks_out, char* err,
                               size_t errlen) {
    if (!obj || !j_out || (n_mon > 0 && !monitor_max) || obj->n_assets < 0 ||
        (obj->n_assets > 0 && (!obj->asset_start || !obj->threshold_m || !obj->weight))) {
        set_err(err, errlen, "null argument");
        return ITR_EINVAL;
    }
    for (int a = 0; a < obj->n_assets; a++)
        for (int q = obj->asset_start[a]; q < obj->asset_start[a + 1]; q++)
            if (!obj->asset_pos || obj->asset_pos[q] < 0 || obj->asset_pos[q] >= n_mon) {
                set_err(err, errlen, "asset position out of the monitor list");
                return ITR_EINVAL;
            }
    *j_out = itr_objective_value(obj, monitor_max, earth, peaks_out, j_risk_out);
    return ITR_OK;
}

itr_status itr_hip_adjoint(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grid, const itr_objective* obj,
                           itr_adjoint_out* outs, char* err, size_t errlen) {
    if (!ctx || !p || !grid || !obj || !outs) {
        set_err(err, errlen, "null argument");
        return ITR_EINVAL;
    }
    if (grid->params) p = grid->params;   // the grid's own parameters win, as in a batch
    if (!p->locked_dt || p->n_locked <= 0) {
        set_err(err, errlen, "the adjoint runs on a locked dt schedule (locked_dt)");
        return ITR_EINVAL;
    }
    if (p->physics != ITR_PHYSICS_HLL) {
        set_err(err, errlen, "the adjoint differentiates the HLL step only");
        return ITR_DECLINED;
    }
    // An initial stage makes the initial depth f32(max(h0, stage - z_ref - z, 0)) a function
    // of the level z: the adjoint of the initial state adds -lambda_h(0) where the stage
    // term is the maximum (it needs the levels the initial state was built from).
    if (p->has_initial_stage && !grid->z_level) {
        set_err(err, errlen, "initial_stage needs z_level");
        return ITR_EINVAL;
    }
    if (grid->replay || grid->design) {
        set_err(err, errlen, "the adjoint runs on prepared faces of a full grid (no replay, no device designs)");
        return ITR_DECLINED;
    }
    bool special = false;
    if (grid->ghost)
        for (int i = 0; i < 2 * grid->nx + 2 * grid->ny; i++) {
            if (grid->ghost[i] > 3) {
                set_err(err, errlen, "the adjoint supports wall, transmissive, Stage and Inflow boundaries");
                return ITR_DECLINED;
            }
            special = special || grid->ghost[i] >= 2;
        }
    if (grid->n_cand <= 0 || !(obj->tau_m > 0.0) || obj->n_assets < 0 || (obj->n_assets > 0 && !obj->asset_start)) {
        set_err(err, errlen, "invalid adjoint arguments");
        return ITR_EINVAL;
    }
    for (int a = 0; a < obj->n_assets; a++)
        for (int q = obj->asset_start[a]; q < obj->asset_start[a + 1]; q++)
            if (obj->asset_pos[q] < 0 || obj->asset_pos[q] >= grid->n_mon) {
                set_err(err, errlen, "asset position out of the monitor list");
                return ITR_EINVAL;
            }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nc = grid->n_cand, nx = grid->nx, ny = grid->ny;
    itr_solver_params q = *p;
    // 1. the plain forward: its device time and the step count N
    double plain_ms = 0.0;
    uint64_t N = 0;
    {
        itr_batch* b0 = nullptr;
        itr_status s = batch_create_impl(ctx, &q, grid, 1, &b0, 0, err, errlen);
        if (s != ITR_OK) return s;
        s = itr_hip_batch_run(b0, nullptr, nullptr, err, errlen);
        std::vector<CandCtl> cc;
        if (s == ITR_OK) s = read_ctl(b0, cc, err, errlen);
        plain_ms = b0->stats.device_ms;
        itr_hip_batch_destroy(b0);
        if (s != ITR_OK) return s;
        for (int c = 0; c < nc; c++) {
            if (cc[c].abort_kind != 0) {
                set_err(err, errlen, "candidate %d: forward aborted (kind %d) on the locked schedule", c, cc[c].abort_kind);
                return ITR_EINVAL;
            }
            if (cc[c].locked_off) {
                set_err(err, errlen, "candidate %d: the locked schedule fails the hard-CFL check at t = %.17g", c,
                        cc[c].dt_fallback_t);
                return ITR_EINVAL;
            }
            if (c > 0 && cc[c].step != N) {
                set_err(err, errlen, "candidates ran different step counts");
                return ITR_EINVAL;
            }
            N = cc[c].step;
        }
    }
    if (N == 0) {
        set_err(err, errlen, "the forward ran no step");
        return ITR_EINVAL;
    }
    const int C = (int)std::ceil(std::sqrt((double)N));
    const int nseg = (int)((N + (uint64_t)C - 1) / (uint64_t)C);
    // 2. the forward with checkpoints every C steps and the log of step constants
    itr_batch* b = nullptr;
    itr_status s = batch_create_impl(ctx, &q, grid, 1, &b, (int64_t)N + 1, err, errlen);
    if (s != ITR_OK) return s;
    std::vector<void*> mem;
    auto cleanup = [&]() {
        for (void* m : mem) (void)hipFree(m);
        itr_hip_batch_destroy(b);
    };
    auto fail = [&](itr_status e) {
        cleanup();
        return e;
    };
#define ADJ_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_err(err, errlen, "%s failed: %s", #expr, hipGetErrorString(e_));           \
            return fail(e_ == hipErrorOutOfMemory ? ITR_ENOMEM : ITR_EDEVICE);              \
        }                                                                                   \
    } while (0)
    const DevGrid& D = b->grids[0].d;
    const size_t plane = D.plane, cstride = 3 * plane, P = (size_t)D.P;
    const size_t state_n = (size_t)nc * cstride;
    auto dalloc = [&](size_t bytes, void** out) -> hipError_t {
        hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 1));
        if (e == hipSuccess) mem.push_back(*out);
        return e;
    };
    float *ckpt, *seg, *pre, *lam, *lam1;
    double *gx, *gy, *gx2 = nullptr, *gy2 = nullptr, *gz = nullptr;
    unsigned long long* d_mis;
    const bool crest = grid->crest != 0;
    const size_t nb = 2 * (size_t)grid->nx + 2 * (size_t)grid->ny;
    ADJ_TRY(dalloc(sizeof(float) * state_n * (size_t)(nseg + 1), (void**)&ckpt));
    ADJ_TRY(dalloc(sizeof(float) * state_n * (size_t)(C + 1), (void**)&seg));
    ADJ_TRY(dalloc(sizeof(float) * state_n * (size_t)C, (void**)&pre));   // post-divergence state of each step
    ADJ_TRY(dalloc(sizeof(float) * state_n * 2, (void**)&lam));
    ADJ_TRY(dalloc(sizeof(float) * state_n, (void**)&lam1));
    ADJ_TRY(dalloc(sizeof(double) * (size_t)nc * plane, (void**)&gx));
    ADJ_TRY(dalloc(sizeof(double) * (size_t)nc * D.plane_y, (void**)&gy));
    if (crest) {
        ADJ_TRY(dalloc(sizeof(double) * (size_t)nc * plane, (void**)&gx2));
        ADJ_TRY(dalloc(sizeof(double) * (size_t)nc * D.plane_y, (void**)&gy2));
    }
    if (special) ADJ_TRY(dalloc(sizeof(double) * (size_t)nc * nb, (void**)&gz));
    ADJ_TRY(dalloc(sizeof(unsigned long long), (void**)&d_mis));
    auto copy_states = [&](float* dst, int parity) -> hipError_t {
        for (int c = 0; c < nc; c++) {
            hipError_t e = hipMemcpyAsync(dst + (size_t)c * cstride, D.state + (size_t)c * 6 * plane + (size_t)parity * cstride,
                                          sizeof(float) * cstride, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    EventPair evp;   // destroyed on every return path
    ADJ_TRY(evp.create());
    hipEvent_t* ev = evp.e;
    auto elapsed = [&](float* ms) -> hipError_t {
        hipError_t e = hipEventSynchronize(ev[1]);
        return e == hipSuccess ? hipEventElapsedTime(ms, ev[0], ev[1]) : e;
    };
    float ms = 0.0f;
    ADJ_TRY(hipEventRecord(ev[0], st));
    ADJ_TRY(copy_states(ckpt, 0));
    for (int k = 1; k <= nseg; k++) {
        const uint64_t n0 = (uint64_t)(k - 1) * C, n1 = std::min<uint64_t>((uint64_t)k * C, N);
        if ((s = launch_steps(b, (int64_t)(n1 - n0), err, errlen)) != ITR_OK) return fail(s);
        ADJ_TRY(copy_states(ckpt + (size_t)k * state_n, (int)(n1 & 1u)));   // step n writes state copy n & 1
    }
    if ((s = launch_steps(b, 1, err, errlen)) != ITR_OK) return fail(s);   // completes step N
    ADJ_TRY(hipEventRecord(ev[1], st));
    ADJ_TRY(elapsed(&ms));
    const double fwd_ms = ms;
    std::vector<CandCtl> cc;
    if ((s = read_ctl(b, cc, err, errlen)) != ITR_OK) return fail(s);
    for (int c = 0; c < nc; c++)
        if (cc[c].step != N || !(cc[c].flags & F_DONE)) {
            set_err(err, errlen, "candidate %d: checkpointed forward differs from the plain forward", c);
            return fail(ITR_EDEVICE);
        }
    std::vector<StepRec> log((size_t)N);
    ADJ_TRY(hipMemcpy(log.data(), D.steplog, sizeof(StepRec) * N, hipMemcpyDeviceToHost));
    // 3. J_risk and the peak seeds (objective.rs:95-101; softplus at :71-74)
    struct Seed { uint64_t step; int cand; size_t cell; float v; };
    std::vector<Seed> seeds;
    const int nm = D.n_mon;
    std::vector<float> mm((size_t)std::max(nm, 1));
    std::vector<double> mt((size_t)std::max(nm, 1));
    std::vector<uint32_t> mcells((size_t)std::max(nm, 1));
    if (nm > 0) ADJ_TRY(hipMemcpy(mcells.data(), grid->mon_cells, sizeof(uint32_t) * nm, hipMemcpyHostToHost));
    for (int c = 0; c < nc; c++) {
        itr_adjoint_out& o = outs[c];
        o.j_risk = 0.0;
        if (nm > 0) {
            const size_t off = slot(D, cc[c].parity, c) * nm;
            ADJ_TRY(hipMemcpy(mm.data(), D.mon_max + off, sizeof(float) * nm, hipMemcpyDeviceToHost));
            ADJ_TRY(hipMemcpy(mt.data(), D.mon_tmax + off, sizeof(double) * nm, hipMemcpyDeviceToHost));
        }
        std::vector<double> md(mm.begin(), mm.begin() + std::max(nm, 1));
        o.j_risk = itr_j_risk(obj, md.data(), o.peak_m);
        for (int a = 0; a < obj->n_assets; a++) {
            int32_t kstar = -1;
            const double pk = itr_asset_peak(obj, a, md.data(), &kstar);
            const double x = (pk - obj->threshold_m[a]) / obj->tau_m;
            if (kstar < 0) continue;   // never wet: the peak does not move to first order
            uint64_t nstar = 0;
            for (uint64_t i = 0; i < N; i++)
                if (log[i].t_new == mt[kstar]) { nstar = i + 1; break; }
            if (nstar == 0) continue;
            const double dpk = obj->weight[a] / (1.0 + std::exp(-x));
            const uint32_t cell = mcells[kstar];
            seeds.push_back(Seed{nstar, c, (size_t)(cell / (uint32_t)nx) * P + cell % (uint32_t)nx, (float)dpk});
        }
    }
    // 4. reverse sweep, segment by segment
    ADJ_TRY(hipMemsetAsync(lam, 0, sizeof(float) * state_n * 2, st));
    ADJ_TRY(hipMemsetAsync(gx, 0, sizeof(double) * (size_t)nc * plane, st));
    ADJ_TRY(hipMemsetAsync(gy, 0, sizeof(double) * (size_t)nc * D.plane_y, st));
    if (crest) {
        ADJ_TRY(hipMemsetAsync(gx2, 0, sizeof(double) * (size_t)nc * plane, st));
        ADJ_TRY(hipMemsetAsync(gy2, 0, sizeof(double) * (size_t)nc * D.plane_y, st));
    }
    if (special) ADJ_TRY(hipMemsetAsync(gz, 0, sizeof(double) * (size_t)nc * nb, st));
    ADJ_TRY(hipMemsetAsync(d_mis, 0, sizeof(unsigned long long), st));
    AdjArgs A{};
    A.dzx = D.dzx;
    A.dzy = D.dzy;
    A.crest = crest ? 1 : 0;
    A.dzx2 = crest ? D.dzx2 : nullptr;
    A.dzy2 = crest ? D.dzy2 : nullptr;
    A.gx2 = gx2;
    A.gy2 = gy2;
    A.special = special ? 1 : 0;
    A.G = (const DevGrid*)b->d_grids;
    A.gz = gz;
    A.wet = D.wet;
    A.gn2 = D.gn2;
    A.ghost = D.ghost;
    A.nx = nx;
    A.ny = ny;
    A.P = D.P;
    A.plane = plane;
    A.cstride = cstride;
    A.dzx_stride = plane;
    A.dzy_stride = D.plane_y;
    A.gx = gx;
    A.gy = gy;
    A.lam1 = lam1;
    // the start time of step n (1-based): the previous step's end, bitwise as the forward
    auto t_start = [&](uint64_t n) { return n <= 1 ? 0.0 : log[(size_t)(n - 2)].t_new; };
    auto consts = [&](uint64_t n) {   // constants of step n (1-based)
        StepK K;
        const StepRec& r = log[(size_t)(n - 1)];
        K.g = D.g32;
        K.g_half = D.g_half32;
        K.inv_dx = D.inv_dx32;
        K.inv_dy = D.inv_dy32;
        K.h_dry = b->prm.h_dry;
        K.heps2 = b->prm.heps2;
        K.q_min = b->prm.q_min;
        K.dt = r.dt32;
        K.dtdx = r.dtdx;
        K.dtdy = r.dtdy;
        K.rain = r.rain32;
        K.cap = r.cap32;
        return K;
    };
    const dim3 grd((unsigned)((nx + 63) / 64), (unsigned)ny, (unsigned)nc), blk(64);
    // The band march needs enough bands in flight to hide its row-to-row latency; small
    // batches keep the per-cell faces kernel (one thread per cell).
    A.R = std::min(ctx->adjoint_rows > 0 ? ctx->adjoint_rows : 16, ny);
    A.waves = (nx + kColsPerWave - 1) / kColsPerWave;
    A.tail_groups = tail_group_count(nx, A.waves);
    const dim3 bgrd((unsigned)((ny + A.R - 1) / A.R), (unsigned)nc);
    const bool band_mode = ctx->adjoint_mode == 1 ||
                           (ctx->adjoint_mode == 0 && (size_t)bgrd.x * bgrd.y * (size_t)A.waves >= 2048);
    EventPair erp0, erp1;
    ADJ_TRY(erp0.create());
    ADJ_TRY(erp1.create());
    hipEvent_t er[4] = {erp0.e[0], erp0.e[1], erp1.e[0], erp1.e[1]};
    double replay_ms = 0.0, reverse_ms = 0.0;
    int li = 0;
    for (int k = nseg; k >= 1; k--) {
        const uint64_t n0 = (uint64_t)(k - 1) * C, n1 = std::min<uint64_t>((uint64_t)k * C, N);
        ADJ_TRY(hipEventRecord(er[0], st));
        ADJ_TRY(hipMemcpyAsync(seg, ckpt + (size_t)(k - 1) * state_n, sizeof(float) * state_n, hipMemcpyDeviceToDevice, st));
        for (uint64_t n = n0 + 1; n <= n1; n++) {   // replay; the segment's last step is the check
            A.U = seg + (size_t)(n - 1 - n0) * state_n;
            A.Uo = seg + (size_t)(n - n0) * state_n;
            A.pre = pre + (size_t)(n - 1 - n0) * state_n;
            A.K = consts(n);
            A.t = t_start(n);
            if (band_mode)
                hipLaunchKernelGGL(itr_replay_band_kernel, bgrd, dim3((unsigned)(A.waves * 64)), 0, st, A);
            else
                hipLaunchKernelGGL(itr_replay_kernel, grd, blk, 0, st, A);
        }
        hipLaunchKernelGGL(itr_adj_compare_kernel, dim3(256), dim3(256), 0, st, seg + (size_t)(n1 - n0) * state_n,
                           ckpt + (size_t)k * state_n, state_n, d_mis);
        ADJ_TRY(hipGetLastError());
        ADJ_TRY(hipEventRecord(er[1], st));
        for (uint64_t n = n1; n > n0; n--) {
            float* lam_n = lam + (size_t)li * state_n;
            for (const Seed& sd : seeds)
                if (sd.step == n) hipLaunchKernelGGL(itr_adj_seed_kernel, dim3(1), dim3(64), 0, st, lam_n, cstride, sd.cand, sd.cell, sd.v);
            A.U = seg + (size_t)(n - 1 - n0) * state_n;
            A.pre = pre + (size_t)(n - 1 - n0) * state_n;
            A.lam = lam_n;
            A.lamo = lam + (size_t)(li ^ 1) * state_n;
            A.K = consts(n);
            A.t = t_start(n);
            if (band_mode) {
                hipLaunchKernelGGL(itr_adj_band_kernel, bgrd, dim3((unsigned)(A.waves * 64)), 0, st, A);
            } else {
                hipLaunchKernelGGL(itr_adj_sources_kernel, grd, blk, 0, st, A);
                hipLaunchKernelGGL(itr_adj_faces_kernel, grd, blk, 0, st, A);
            }
            li ^= 1;
        }
        ADJ_TRY(hipGetLastError());
        ADJ_TRY(hipEventRecord(er[2], st));
        ADJ_TRY(hipEventSynchronize(er[2]));
        float a_ms = 0.0f, b_ms = 0.0f;
        ADJ_TRY(hipEventElapsedTime(&a_ms, er[0], er[1]));
        ADJ_TRY(hipEventElapsedTime(&b_ms, er[1], er[2]));
        replay_ms += a_ms;
        reverse_ms += b_ms;
    }
    unsigned long long mis = 0;
    ADJ_TRY(hipMemcpy(&mis, d_mis, sizeof mis, hipMemcpyDeviceToHost));
    // lambda of the initial state (the last sweep's output), for an initial stage
    std::vector<float> lam0h;
    if (p->has_initial_stage) {
        lam0h.resize((size_t)nc * plane);
        const float* L0 = lam + (size_t)li * state_n;
        for (int c = 0; c < nc; c++)
            ADJ_TRY(hipMemcpy(lam0h.data() + (size_t)c * plane, L0 + (size_t)c * cstride, sizeof(float) * plane,
                              hipMemcpyDeviceToHost));
    }
    // 5. face and cell gradients
    std::vector<double> hx((size_t)nc * plane), hy((size_t)nc * D.plane_y), hx2, hy2;
    ADJ_TRY(hipMemcpy(hx.data(), gx, sizeof(double) * hx.size(), hipMemcpyDeviceToHost));
    ADJ_TRY(hipMemcpy(hy.data(), gy, sizeof(double) * hy.size(), hipMemcpyDeviceToHost));
    if (crest) {
        hx2.resize(hx.size());
        hy2.resize(hy.size());
        ADJ_TRY(hipMemcpy(hx2.data(), gx2, sizeof(double) * hx2.size(), hipMemcpyDeviceToHost));
        ADJ_TRY(hipMemcpy(hy2.data(), gy2, sizeof(double) * hy2.size(), hipMemcpyDeviceToHost));
    }
    std::vector<double> hz;
    if (special) {
        hz.resize((size_t)nc * nb);
        ADJ_TRY(hipMemcpy(hz.data(), gz, sizeof(double) * hz.size(), hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < nc; c++) {
        itr_adjoint_out& o = outs[c];
        const double* X = hx.data() + (size_t)c * plane;
        const double* Y = hy.data() + (size_t)c * D.plane_y;
        // interior faces only: boundary faces have fixed zero jumps / rises
        auto fx = [&](int r, int f) { return (f >= 1 && f <= nx - 1) ? X[(size_t)r * P + f] : 0.0; };
        auto fy = [&](int f, int col) { return (f >= 1 && f <= ny - 1) ? Y[(size_t)f * P + col] : 0.0; };
        if (!crest) {
            if (o.dJ_dz_face_x)
                for (int r = 0; r < ny; r++)
                    for (int f = 0; f <= nx; f++) o.dJ_dz_face_x[(size_t)r * (nx + 1) + f] = fx(r, f);
            if (o.dJ_dz_face_y)
                for (int f = 0; f <= ny; f++)
                    for (int col = 0; col < nx; col++) o.dJ_dz_face_y[(size_t)f * nx + col] = fy(f, col);
            if (o.dJ_dz_cell)
                for (int r = 0; r < ny; r++)
                    for (int col = 0; col < nx; col++) {
                        const size_t i = (size_t)r * nx + col;
                        o.dJ_dz_cell[i] = grid->wet[i] == 0.0f ? 0.0
                                          : fx(r, col) - fx(r, col + 1) + fy(r, col) - fy(r + 1, col);
                    }
        } else {
            // Crest faces (terrain.rs:378-429): zf = max(crest, z_l, z_r), d_l = zf - z_l,
            // d_r = zf - z_r.  The rises tell which term is the maximum (a zero rise: that
            // side's level), so the rise gradients split into the crest line's gradient (it
            // moves zf only while it is the strict maximum) and the two cells' levels.
            const double* X2 = hx2.data() + (size_t)c * plane;
            const double* Y2 = hy2.data() + (size_t)c * D.plane_y;
            const float* dl_x = grid->dzx + (size_t)c * ny * (nx + 1);
            const float* dr_x = grid->dzx_r + (size_t)c * ny * (nx + 1);
            const float* dl_y = grid->dzy + (size_t)c * (ny + 1) * nx;
            const float* dr_y = grid->dzy_r + (size_t)c * (ny + 1) * nx;
            std::vector<double> cell((size_t)nx * ny, 0.0);
            // one face: rise gradients (bl, br), rises (dl, dr); returns the crest gradient
            // and adds the levels' gradients to cells l and r
            auto split = [&](double bl, double br, float dl, float dr, size_t l, size_t r) -> double {
                if (dl > 0.0f && dr > 0.0f) {
                    cell[l] -= bl;
                    cell[r] -= br;
                    return bl + br;
                }
                if (dl == 0.0f) {   // zf = z_l: d_l stays 0, d_r = z_l - z_r
                    cell[l] += br;
                    cell[r] -= br;
                } else {            // zf = z_r: d_l = z_r - z_l
                    cell[r] += bl;
                    cell[l] -= bl;
                }
                return 0.0;
            };
            for (int r = 0; r < ny; r++)
                for (int f = 0; f <= nx; f++) {
                    const size_t i = (size_t)r * (nx + 1) + f;
                    const bool in = f >= 1 && f <= nx - 1;
                    const double bl = in ? X[(size_t)r * P + f] : 0.0, br = in ? X2[(size_t)r * P + f] : 0.0;
                    const double gc = in ? split(bl, br, dl_x[i], dr_x[i], (size_t)r * nx + f - 1, (size_t)r * nx + f) : 0.0;
                    if (o.dJ_dz_face_x) o.dJ_dz_face_x[i] = gc;
                    if (o.dJ_drise_face_x) {
                        o.dJ_drise_face_x[2 * i] = bl;
                        o.dJ_drise_face_x[2 * i + 1] = br;
                    }
                }
            for (int f = 0; f <= ny; f++)
                for (int col = 0; col < nx; col++) {
                    const size_t i = (size_t)f * nx + col;
                    const bool in = f >= 1 && f <= ny - 1;
                    const double bl = in ? Y[(size_t)f * P + col] : 0.0, br = in ? Y2[(size_t)f * P + col] : 0.0;
                    const double gc = in ? split(bl, br, dl_y[i], dr_y[i], (size_t)(f - 1) * nx + col, (size_t)f * nx + col) : 0.0;
                    if (o.dJ_dz_face_y) o.dJ_dz_face_y[i] = gc;
                    if (o.dJ_drise_face_y) {
                        o.dJ_drise_face_y[2 * i] = bl;
                        o.dJ_drise_face_y[2 * i + 1] = br;
                    }
                }
            if (o.dJ_dz_cell)
                for (size_t i = 0; i < (size_t)nx * ny; i++) o.dJ_dz_cell[i] = grid->wet[i] == 0.0f ? 0.0 : cell[i];
        }
        // Initial stage (solver.rs:292-304): h0 = f32(max(h_init, stage - z_ref - z, 0))
        // moves with the level z (-1) where the stage term is the maximum.
        if (p->has_initial_stage && o.dJ_dz_cell) {
            const double* zl = grid->z_level + (size_t)c * nx * ny;
            const float* L0 = lam0h.data() + (size_t)c * plane;
            for (int r = 0; r < ny; r++)
                for (int col = 0; col < nx; col++) {
                    const size_t i = (size_t)r * nx + col;
                    const double ds = p->initial_stage - grid->z_ref - zl[i];
                    if (grid->wet[i] != 0.0f && ds > p->initial_depth && ds > 0.0)
                        o.dJ_dz_cell[i] -= (double)L0[(size_t)r * P + col];
                }
        }
        // Stage ghosts: their depth moves with the boundary cell's level (gz, per ghost).
        if (special && o.dJ_dz_cell) {
            const double* Z = hz.data() + (size_t)c * nb;
            for (size_t kg = 0; kg < nb; kg++) {
                if (Z[kg] == 0.0) continue;
                const size_t k = kg, snx = (size_t)nx, sny = (size_t)ny;
                const size_t i = k < snx ? k : k < 2 * snx ? (sny - 1) * snx + (k - snx)
                               : k < 2 * snx + sny ? (k - 2 * snx) * snx : (k - 2 * snx - sny) * snx + snx - 1;
                if (grid->wet[i] != 0.0f) o.dJ_dz_cell[i] += Z[kg];
            }
        }
        o.steps = N;
        o.segment_steps = C;
        o.checkpoints = nseg + 1;
        o.replay_mismatches = mis;
        o.plain_forward_ms = plain_ms;
        o.forward_ms = fwd_ms;
        o.replay_ms = replay_ms;
        o.reverse_ms = reverse_ms;
    }
#undef ADJ_TRY
    cleanup();
    if (mis != 0) {   // a replayed segment left the checkpointed trajectory: no gradient
        set_err(err, errlen, "%llu replayed segment-end values differ from the checkpoints", mis);
        return ITR_EDEVICE;
    }
    return ITR_OK;
}

}  // extern "C"

// Re-tile a resident batch: new band heights per grid (tasks, band monitor masks and
// LDS size follow).  State and control blocks are untouched.
itr_status retile(itr_batch* b, const std::vector<int>& Rs, char* err, size_t errlen) {
    std::vector<DevGrid> dg;
    for (size_t g = 0; g < b->grids.size(); g++) {
        DevGrid D = b->grids[g].d;
        const int R = std::max(1, std::min(Rs[g], D.ny));
        D.R = R;
        D.nbands = (D.ny + R - 1) / R;
        dg.push_back(D);
    }
    std::vector<int4> tasks;
    std::vector<LaunchList> keep = b->lists;
    std::vector<TailLayout> keep_tails = b->tails;
    const size_t lds_max = fill_lists(b, dg, tasks);
    if (lds_max > 160 * 1024 - 1024) {
        b->lists = keep;
        b->tails = keep_tails;
        b->ntasks = 0;
        for (const LaunchList& L : keep) b->ntasks += L.ntasks;
        set_err(err, errlen, "band of %zu bytes of LDS does not fit", lds_max);
        return ITR_DECLINED;
    }
    for (size_t g = 0; g < b->grids.size(); g++) b->grids[g].d = dg[g];
    const std::vector<DevGrid> dall = device_grids(dg, b->tails);
    HIP_TRY(hipMemcpy(b->d_grids, dall.data(), dall.size() * sizeof(DevGrid), hipMemcpyHostToDevice));
    if (!tasks.empty()) HIP_TRY(hipMemcpy(b->d_tasks, tasks.data(), tasks.size() * sizeof(int4), hipMemcpyHostToDevice));
    b->args.ntasks = b->ntasks;
    bind_lists(b);
    b->stats.band_rows = dg.empty() ? 0 : dg[0].R;
    drop_graphs(b);
    return ITR_OK;
}

extern "C" {

itr_status itr_hip_autotune(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grids, int32_t n_grids,
                            int32_t trial_steps, itr_tune_result* out, char* err, size_t errlen) {
    if (!ctx || !p || !grids || n_grids <= 0 || !out) {
        set_err(err, errlen, "null argument");
        return ITR_EINVAL;
    }
    // Trials run whole graph chunks, as batches run (launch overheads matter for small grids).
    const int K = ctx->steps_per_chunk > 0 ? (ctx->steps_per_chunk + 3) / 4 * 4 : 32;
    const int T = std::max(1, ((int)trial_steps + K - 1) / K) * K;
    const int saved = ctx->band_rows;
    ctx->band_rows = 0;
    itr_batch* b = nullptr;
    itr_solver_params q = *p;   // trials time the kernel: free dt, no step bound, and a
    q.max_steps = 0;            // duration no trial sequence reaches (a finished candidate
    q.locked_dt = nullptr;      // would make later trials time empty launches)
    q.n_locked = 0;
    q.duration_s = std::max(q.duration_s, 1e9);
    itr_status s = batch_create_impl(ctx, &q, grids, n_grids, &b, 0, err, errlen);
    ctx->band_rows = saved;
    if (s != ITR_OK) return s;
    std::vector<int> Rs;
    for (auto& GH : b->grids) Rs.push_back(GH.d.R);
    auto timed = [&](double* ms) -> itr_status {   // best of two after a warm-up chunk
        itr_status e = itr_hip_batch_advance(b, K, err, errlen);   // graph capture, caches
        double best = 1e300;
        for (int rep = 0; rep < 2 && e == ITR_OK; rep++) {
            e = itr_hip_batch_advance(b, T, err, errlen);
            best = std::min(best, b->stats.device_ms);
        }
        *ms = best;
        return e;
    };
    double best_all = 0.0;
    s = timed(&best_all);
    // Coordinate descent over grid shapes, the largest work first.
    std::vector<int> order(n_grids);
    for (int g = 0; g < n_grids; g++) order[g] = g;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return (double)grids[x].nx * grids[x].ny * grids[x].n_cand > (double)grids[y].nx * grids[y].ny * grids[y].n_cand;
    });
    // Band heights tried: every height to 16 (the launch's tail wave depends on the
    // band count modulo the resident workgroup slots), then coarser steps, with multiples
    // of 3 among them (the last wave of a 512-column row marches 3 row groups).
    static const int cand_R[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 21, 24, 27, 28, 30, 32, 33};
    std::vector<char> done(n_grids, 0);
    for (int gi = 0; gi < n_grids && s == ITR_OK; gi++) {
        const int g = order[gi];
        if (done[g]) continue;
        std::vector<int> same;   // grids of the same shape share one band height
        for (int k = 0; k < n_grids; k++)
            if (grids[k].nx == grids[g].nx && grids[k].ny == grids[g].ny && grids[k].crest == grids[g].crest) same.push_back(k);
        int bestR = Rs[g];
        for (int R : cand_R) {
            if (R > grids[g].ny || R == bestR) continue;
            std::vector<int> trial = Rs;
            for (int k : same) trial[k] = R;
            if (retile(b, trial, err, errlen) != ITR_OK) continue;   // does not fit the LDS
            double ms = 0.0;
            if ((s = timed(&ms)) != ITR_OK) break;
            if (ms < best_all) {
                best_all = ms;
                bestR = R;
            }
        }
        for (int k : same) {
            Rs[k] = bestR;
            done[k] = 1;
        }
        if (s == ITR_OK) s = retile(b, Rs, err, errlen);
    }
    if (s == ITR_OK) {
        for (int g = 0; g < n_grids; g++) {
            out[g].band_rows = Rs[g];
            out[g].reserved = 0;
            out[g].us_per_step = best_all * 1e3 / T;
            bool found = false;
            for (auto& t : ctx->tuned)
                if (t.nx == grids[g].nx && t.ny == grids[g].ny && t.crest == (grids[g].crest ? 1 : 0)) {
                    t.R = Rs[g];
                    found = true;
                }
            if (!found) ctx->tuned.push_back(itr_ctx::Tuned{grids[g].nx, grids[g].ny, grids[g].crest ? 1 : 0, Rs[g]});
        }
    }
    itr_hip_batch_destroy(b);
    return s;
}

itr_status itr_hip_math_selfcheck(itr_ctx* ctx, uint64_t* recip_mismatches, uint64_t* sqrt_mismatches, char* err,
                                  size_t errlen) {
    if (!ctx || !recip_mismatches || !sqrt_mismatches) return ITR_EINVAL;
    HIP_TRY(hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(itr_math_check_kernel, dim3((1u << 23) / 256), dim3(256), 0, ctx->stream, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) {
        set_err(err, errlen, "math self-check failed: %s", hipGetErrorString(e));
        return ITR_EDEVICE;
    }
    *recip_mismatches = h[0];
    *sqrt_mismatches = h[1];
    return ITR_OK;
}

itr_status itr_hip_dt_schedule(itr_ctx* ctx, const itr_solver_params* p, const itr_grid* grid, int32_t cand,
                               double* dts, int64_t capacity, int64_t* n_out, char* err, size_t errlen) {
    if (!ctx || !p || !grid || !dts || !n_out || capacity <= 0 || cand < 0 || cand >= grid->n_cand) {
        set_err(err, errlen, "invalid argument");
        return ITR_EINVAL;
    }
    if (grid->params) p = grid->params;   // the grid's own parameters win, as in a batch
    HIP_TRY(hipSetDevice(ctx->device));
    itr_grid one = *grid;
    one.n_cand = 1;
    const size_t sx = (size_t)grid->ny * (grid->nx + 1), sy = (size_t)(grid->ny + 1) * grid->nx;
    one.dzx = grid->dzx + (size_t)cand * sx;
    one.dzy = grid->dzy + (size_t)cand * sy;
    if (grid->crest) {
        one.dzx_r = grid->dzx_r + (size_t)cand * sx;
        one.dzy_r = grid->dzy_r + (size_t)cand * sy;
    }
    if (grid->z_level) one.z_level = grid->z_level + (size_t)cand * grid->nx * grid->ny;
    one.baseline_max = nullptr;
    one.n_guard = 0;
    one.params = nullptr;   // q below carries them
    itr_solver_params q = *p;
    if (q.max_steps <= 0 || q.max_steps > capacity) q.max_steps = capacity;
    itr_batch* b = nullptr;
    itr_status s = batch_create_impl(ctx, &q, &one, 1, &b, capacity, err, errlen);
    if (s != ITR_OK) return s;
    s = itr_hip_batch_run(b, nullptr, nullptr, err, errlen);
    std::vector<CandCtl> cc;
    if (s == ITR_OK) s = read_ctl(b, cc, err, errlen);
    if (s == ITR_OK) {
        const uint64_t n = std::min<uint64_t>(cc[0].step, (uint64_t)capacity);
        std::vector<StepRec> log(std::max<uint64_t>(n, 1));
        if (n > 0 && hipMemcpy(log.data(), b->grids[0].d.steplog, sizeof(StepRec) * n, hipMemcpyDeviceToHost) != hipSuccess) {
            set_err(err, errlen, "step log read failed");
            s = ITR_EDEVICE;
        } else {
            for (uint64_t i = 0; i < n; i++) dts[i] = log[i].dt;
            *n_out = (int64_t)n;
        }
    }
    itr_hip_batch_destroy(b);
    return s;
}

}  // extern "C"
