// itr_oracle.cpp -- CPU oracle for the batched shallow-water step.
//
// TEST INFRASTRUCTURE ONLY.  A C++17 restatement of the reference engine's CPU solver
// path, used as the checker for the HIP path (bitwise for f32) and as the f64 scalar
// oracle.  The reference is Rust and cannot be compiled here (no cargo/rustc, crates not
// vendored), so every function below restates the Rust source it cites, keeping its
// evaluation order so that IEEE results are identical.  Build with -ffp-contract=off
// (GCC contracts a*b+c into FMA by default in GNU mode on FMA-capable targets).
//
// Reference files (paths relative to the reference repository):
//   crates/itr-core/src/real.rs          Real, FixedSum, max_slice
//   crates/itr-hydro/src/numerics.rs     cell_vel, hll, face_signed, face_crest, ...
//   crates/itr-hydro/src/kernel.rs       row kernels, step_oracle, step_fused
//   crates/itr-hydro/src/boundary.rs     fill_ghosts
//   crates/itr-hydro/src/terrain.rs      FineTerrain, TerrainBase, apply_edit
//   crates/itr-hydro/src/solver.rs       Solver::run_impl
//   crates/itr-core/src/scenario.rs      Hyetograph
//   crates/itr-core/src/objective.rs     Objective (J and the incumbent lower bound come
//                                        from include/itr_objective.h, the one
//                                        implementation of J, pinned by closed forms)
//   crates/itr-core/src/design.rs        Primitive, materialize
//   crates/itr-hydro/src/inertial.rs     step_local_inertial (kernel 2, screening physics)
//
// Pins of the local-inertial step: the reference's own tests (lake at rest, rain mass,
// tilted-plane outflow, peak depths within 15 % of HLL), the kinematic normal depth
// (friction exponent, g n^2 and its face mean), the time step of a still lake and of a
// steady plane, a one-step hand expansion of the theta = 0.8 face update, and the dt of
// a state whose fastest cells are v-bound (max(u, v) against u, u + v, |(u, v)|).
// Stage ghosts: a lake at rest at the stage level over a wavy bed, a basin filling from
// dry to the stage level, and a one-step expansion of a ghost clamped to depth 0.

#include "itr_oracle.h"
#include "../include/itr_objective.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- real.rs:11-191
template <class T> inline T rmax(T a, T b) { return a > b ? a : b; }   // real.rs:44-46
template <class T> inline T rmin(T a, T b) { return a < b ? a : b; }   // real.rs:48-50
template <class T> inline T rabs(T x) { return x < T(0) ? -x : x; }    // real.rs:39-41
template <class T> inline T canon(T x) { return x == T(0) ? T(0) : x; } // real.rs:56-58

__attribute__((always_inline)) inline float rcbrt(float x) {  // real.rs:116-129
    uint32_t bits;
    std::memcpy(&bits, &x, 4);
    // bits / 3 as a multiply-high (exact for every 32-bit value: 0xAAAAAAAB = ceil(2^33 / 3)),
    // which vectorises where a vector integer divide does not exist.
    uint32_t s = 0x54A23280u - (uint32_t)(((uint64_t)bits * 0xAAAAAAABull) >> 33);
    float r;
    std::memcpy(&r, &s, 4);
    const float x3 = x * (float)(1.0 / 3.0);
    const float c43 = (float)(4.0 * (1.0 / 3.0));
    for (int i = 0; i < 3; i++) {
        float r2 = r * r;
        r = r * c43 - x3 * (r2 * r2);
    }
    return r;
}

inline double rcbrt(double x) {  // real.rs:81-95
    const uint64_t K = 0x5540000000000000ull - (0x8782Bull << 29);
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    uint64_t s = K - bits / 3u;
    double r;
    std::memcpy(&r, &s, 8);
    const double third = 1.0 / 3.0;
    const double x3 = x * third;
    const double c43 = 4.0 * third;
    for (int i = 0; i < 4; i++) {
        double r2 = r * r;
        r = r * c43 - x3 * (r2 * r2);
    }
    return r;
}

template <class T> struct FixedSum {  // real.rs:132-175
    T acc[16];
    FixedSum() { for (auto& a : acc) a = T(0); }
    inline void add(size_t col, T v) { acc[col & 15] += v; }
    void add_slice(const T* v, size_t n) { for (size_t k = 0; k < n; k++) acc[k & 15] += v[k]; }
    T total() const {
        T l8[8];
        for (int k = 0; k < 8; k++) l8[k] = acc[2 * k] + acc[2 * k + 1];
        T l4[4] = {l8[0] + l8[1], l8[2] + l8[3], l8[4] + l8[5], l8[6] + l8[7]};
        return (l4[0] + l4[1]) + (l4[2] + l4[3]);
    }
};

template <class T> T max_slice(const T* v, size_t n) {  // real.rs:177-191
    T m[16];
    for (auto& x : m) x = T(0);
    for (size_t k = 0; k < n; k++) m[k & 15] = rmax(m[k & 15], v[k]);
    T a = T(0);
    for (int l = 0; l < 16; l++) a = rmax(a, m[l]);
    return a;
}

// ---------------------------------------------------------------- scenario.rs:527-589
struct Hyeto {
    std::vector<double> t, r;
    double rate(double x) const {
        size_t n = t.size();
        if (n == 0) return 0.0;
        if (x <= t[0]) return r[0];
        for (size_t k = 0; k + 1 < n; k++) {
            if (x <= t[k + 1]) {
                double t0 = t[k], r0 = r[k], t1 = t[k + 1], r1 = r[k + 1];
                if (t1 == t0) return r1;
                return r0 + (r1 - r0) * (x - t0) / (t1 - t0);
            }
        }
        return r[n - 1];
    }
    double integral(double a, double b) const {
        if (b <= a) return 0.0;
        if (t.empty()) return 0.0;
        double total = 0.0, tt = a, rr = rate(a);
        for (double tk : t) {
            if (tk > a && tk < b) {
                double rk = rate(tk);
                total += 0.5 * (rr + rk) * (tk - tt);
                tt = tk;
                rr = rk;
            }
        }
        double rb = rate(b);
        total += 0.5 * (rr + rb) * (b - tt);
        return total;
    }
};

// ---------------------------------------------------------------- numerics.rs:11-214
template <class T> struct StepConsts {
    T g, g_half, dt, dtdx, dtdy, inv_dx, inv_dy, h_dry, heps2, q_min, rain, cap;
};
template <class T> struct Vel { T u, v, c; };
template <class T> struct Face { T m, n_l, n_r, t; };

template <class T> __attribute__((always_inline)) inline Vel<T> cell_vel(T h, T qx, T qy, const StepConsts<T>& k) {  // numerics.rs:36-43
    T h2 = h * h;
    T inv = T(1) / (h2 + rmax(h2, k.heps2));
    T t = (h + h) * inv;
    return Vel<T>{qx * t, qy * t, std::sqrt(k.g * h)};
}

template <class T>
__attribute__((always_inline)) inline Face<T> hll(T hl, T ul, T tl, T cl, T hr, T ur, T tr, T cr, const StepConsts<T>& k) {  // numerics.rs:58-93
    const T z = T(0);
    const T half = T(0.5), quarter = T(0.25), two = T(2.0);
    // The selects are written as nested selects on the ordered comparisons hl > 0 and
    // hr > 0 (dry_x = !(h_x > 0): same values, NaN depths count as dry) so that they
    // vectorise.
    T ustar = half * (ul + ur) + cl - cr;
    T cstar = half * (cl + cr) + quarter * (ul - ur);
    T slw = rmin(ul - cl, ustar - cstar);
    T srw = rmax(ur + cr, ustar + cstar);
    T sl = hl > z ? (hr > z ? slw : ul - cl) : ur - two * cr;
    T sr = hl > z ? (hr > z ? srw : ul + two * cl) : ur + cr;
    T ml = hl * ul;
    T mr = hr * ur;
    T pl_ = k.g_half * hl * hl;
    T pr_ = k.g_half * hr * hr;
    T pl = ml * ul + pl_;
    T pr = mr * ur + pr_;
    const T any_wet = hl > z ? T(1) : (hr > z ? T(1) : z);   // 0 when both sides are dry
    const T span = any_wet > z ? sr - sl : T(1);
    T den = sl >= z ? T(1) : (sr <= z ? T(1) : span);
    T inv = T(1) / den;
    T slsr = sl * sr;
    T fm_star = (sr * ml - sl * mr + slsr * (hr - hl)) * inv;
    T fn_star = (sr * pl - sl * pr + slsr * (mr - ml)) * inv;
    const T fm_w = sl >= z ? ml : (sr <= z ? mr : fm_star);
    const T fn_w = sl >= z ? pl : (sr <= z ? pr : fn_star);
    T fm = any_wet > z ? fm_w : z;
    T fn_ = any_wet > z ? fn_w : z;
    T ft = fm * (fm > z ? tl : tr);
    return Face<T>{fm, fn_ - pl_, fn_ - pr_, ft};
}

template <class T>
__attribute__((always_inline)) inline Face<T> face_signed(T hl, Vel<T> vl, T hr, Vel<T> vr, T dz, bool normal_x, const StepConsts<T>& k) {  // numerics.rs:97-112
    const T z = T(0);
    bool pos = dz > z;
    T hsl = rmax(hl - rmax(dz, z), z);
    T hsr = rmax(hr - rmax(-dz, z), z);
    T hred = pos ? hsl : hsr;
    T cred = std::sqrt(k.g * hred);
    T csl = pos ? cred : vl.c;
    T csr = pos ? vr.c : cred;
    if (normal_x) return hll(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

template <class T>
__attribute__((always_inline)) inline Face<T> face_crest(T hl, Vel<T> vl, T hr, Vel<T> vr, T d_l, T d_r, bool normal_x, const StepConsts<T>& k) {  // numerics.rs:117-129
    const T z = T(0);
    T hsl = rmax(hl - d_l, z);
    T hsr = rmax(hr - d_r, z);
    T csl = d_l > z ? std::sqrt(k.g * hsl) : vl.c;
    T csr = d_r > z ? std::sqrt(k.g * hsr) : vr.c;
    if (normal_x) return hll(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

template <class T> struct CellOut { T h, qx, qy, s, infil, fix; bool ok; };   // ok: !bad

template <class T>
__attribute__((always_inline)) inline void flux_update(T h, T qx, T qy, const Face<T>& w, const Face<T>& e, const Face<T>& n, const Face<T>& s,
                        const StepConsts<T>& k, T& h1, T& qx1, T& qy1) {  // numerics.rs:148-153
    h1 = h - k.dtdx * (e.m - w.m) - k.dtdy * (s.m - n.m);
    qx1 = qx - k.dtdx * (e.n_l - w.n_r) - k.dtdy * (s.t - n.t);
    qy1 = qy - k.dtdx * (e.t - w.t) - k.dtdy * (s.n_l - n.n_r);
}

template <class T> inline T rcbrt_t(T x) { return rcbrt(x); }

template <class T>
__attribute__((always_inline)) inline CellOut<T> sources(T h1, T qx1, T qy1, T wet, T gn2, const StepConsts<T>& k) {  // numerics.rs:158-195
    const T z = T(0);
    bool ok = h1 >= -k.h_dry;   // bad = !ok (ordered comparisons vectorise)
    T fix = h1 < z ? -h1 : z;
    h1 = h1 < z ? z : h1;
    T h2 = h1 + k.rain * wet;
    T infil = rmin(k.cap * wet, h2);
    T h3 = h2 - infil;
    bool dry = h3 < k.h_dry;
    T r = rcbrt_t<T>(dry ? T(1) : h3);
    T r3 = r * r * r;
    T r7 = r3 * r3 * r;
    T qm = std::sqrt(qx1 * qx1 + qy1 * qy1);
    T inv = T(1) / (T(1) + k.dt * gn2 * qm * r7);
    T keep = dry ? z : T(1);
    T qx2 = qx1 * inv * keep;
    T qy2 = qy1 * inv * keep;
    qx2 = rabs(qx2) < k.q_min ? z : qx2;
    qy2 = rabs(qy2) < k.q_min ? z : qy2;
    T c = std::sqrt(k.g * h3);
    T s_rate = (rabs(qx2) * r3 + c) * k.inv_dx + (rabs(qy2) * r3 + c) * k.inv_dy;
    ok = ok & (qm == qm);
    return CellOut<T>{canon(h3), canon(qx2), canon(qy2), s_rate, infil, fix, ok};
}

// ---------------------------------------------------------------- terrain.rs
enum GhostKind { G_WALL = 0, G_TRANS = 1, G_STAGE = 2, G_INFLOW = 3 };
struct Ghost { int kind = G_WALL; double stage = 0; int idx = 0; double inv_len = 0; };

struct Fine {  // terrain.rs:190-235
    int nx, ny;
    double ox, oy, dx, dy;
    std::vector<double> z;
    std::vector<uint8_t> wall;
    double z_ref, z_wall;
    std::vector<double> n;
};

struct TerrainBase {  // terrain.rs:237-321 (faces kept in f64, rounded at run time)
    int nx, ny, level;
    bool crest;
    double dx, dy, g;
    std::shared_ptr<Fine> fine;
    std::vector<double> z;           // coarse z, unpadded
    std::vector<uint8_t> wall;       // coarse wall mask
    std::vector<double> nval;        // coarse Manning n
    int wet_count;
    std::vector<Ghost> north, south, west, east;
    std::vector<Hyeto> hydro;
    size_t P() const { return (size_t)(nx + 2); }
    size_t at(int i, int j) const { return (size_t)j * P() + (size_t)i; }
};

const double G_ACC = 9.81;  // terrain.rs:259

std::shared_ptr<Fine> fine_new(int nx, int ny, double ox, double oy, double dx, double dy, const double* dem,
                               const uint8_t* walls, const double* manning) {  // terrain.rs:204-234
    auto f = std::make_shared<Fine>();
    f->nx = nx; f->ny = ny; f->ox = ox; f->oy = oy; f->dx = dx; f->dy = dy;
    double zmin = DBL_MAX, zmax = -DBL_MAX;
    size_t N = (size_t)nx * ny;
    for (size_t i = 0; i < N; i++) {
        bool w = walls ? walls[i] != 0 : false;
        if (!w && std::isfinite(dem[i])) {
            zmin = std::fmin(zmin, dem[i]);
            zmax = std::fmax(zmax, dem[i]);
        }
    }
    if (zmin > zmax) { zmin = 0.0; zmax = 0.0; }
    f->z_ref = std::floor(zmin);
    f->z_wall = (zmax - f->z_ref) + 100.0;
    f->z.resize(N);
    f->wall.resize(N);
    for (size_t i = 0; i < N; i++) {
        bool w = walls ? walls[i] != 0 : false;
        f->z[i] = (w || !std::isfinite(dem[i])) ? f->z_wall : dem[i] - f->z_ref;
        f->wall[i] = (w || !std::isfinite(dem[i])) ? 1 : 0;
    }
    f->n.assign(manning, manning + N);
    return f;
}

struct EditView {  // design.rs TerrainEdit (get)
    const orc_edit* e;
    double get(int col, int row) const {
        if (!e) return 0.0;
        if (col < e->c0 || row < e->r0 || col >= e->c0 + e->w || row >= e->r0 + e->h) return 0.0;
        return e->dz[(size_t)(row - e->r0) * e->w + (col - e->c0)];
    }
    bool empty() const {  // design.rs:158-160
        if (!e || e->w == 0 || e->h == 0) return true;
        for (int64_t k = 0; k < (int64_t)e->w * e->h; k++) if (e->dz[k] != 0.0) return false;
        return true;
    }
};

inline double fine_z(const Fine& f, int c, int r, const EditView& ed) {  // terrain.rs:353-355
    return f.z[(size_t)r * f.nx + c] + ed.get(c, r);
}

void coarse_cell(const Fine& f, int lv, int c, int r, const EditView& ed, double& zz, bool& w, double& nn) {  // terrain.rs:357-373
    double s = 0.0, sn = 0.0, sn_all = 0.0;
    size_t cnt = 0, tot = 0;
    for (int rr = r * lv; rr < std::min((r + 1) * lv, f.ny); rr++) {
        for (int cc = c * lv; cc < std::min((c + 1) * lv, f.nx); cc++) {
            tot++;
            sn_all += f.n[(size_t)rr * f.nx + cc];
            if (!f.wall[(size_t)rr * f.nx + cc]) {
                s += fine_z(f, cc, rr, ed);
                sn += f.n[(size_t)rr * f.nx + cc];
                cnt++;
            }
        }
    }
    if (cnt == 0) { zz = f.z_wall; w = true; nn = sn_all / (double)tot; }
    else { zz = s / (double)cnt; w = false; nn = sn / (double)cnt; }
}

void face_x(const Fine& f, int lv, bool crest, double zl, double zr, int i, int r, const EditView& ed, double& a,
            double& b) {  // terrain.rs:378-402
    if (!crest) { a = zr - zl; b = 0.0; return; }
    int cl = i * lv - 1;
    int cr = std::min(i * lv, f.nx - 1);
    double cst = -DBL_MAX;
    bool all_wall = true;
    for (int rr = r * lv; rr < std::min((r + 1) * lv, f.ny); rr++) {
        for (int cc : {cl, cr}) {
            if (!f.wall[(size_t)rr * f.nx + cc]) { all_wall = false; cst = std::fmax(cst, fine_z(f, cc, rr, ed)); }
        }
    }
    if (all_wall) cst = f.z_wall;
    double zf = std::fmax(std::fmax(cst, zl), zr);
    a = zf - zl;
    b = zf - zr;
}

void face_y(const Fine& f, int lv, bool crest, double zt, double zb, int c, int rb, const EditView& ed, double& a,
            double& b) {  // terrain.rs:406-429
    if (!crest) { a = zb - zt; b = 0.0; return; }
    int rt_f = rb * lv - 1;
    int rb_f = std::min(rb * lv, f.ny - 1);
    double cst = -DBL_MAX;
    bool all_wall = true;
    for (int cc = c * lv; cc < std::min((c + 1) * lv, f.nx); cc++) {
        for (int rr : {rt_f, rb_f}) {
            if (!f.wall[(size_t)rr * f.nx + cc]) { all_wall = false; cst = std::fmax(cst, fine_z(f, cc, rr, ed)); }
        }
    }
    if (all_wall) cst = f.z_wall;
    double zf = std::fmax(std::fmax(cst, zt), zb);
    a = zf - zt;
    b = zf - zb;
}

// Prepared (per-candidate) terrain in working precision: padded (ny+2)*(nx+2) arrays.
template <class T> struct Prepared {  // terrain.rs:431-515
    const TerrainBase* b;
    std::vector<T> fx0, fx1, fy0, fy1, wet, gn2;
    std::vector<double> zlev;   // z_level(): coarse z incl. edit overrides
};

template <class T> Prepared<T> prepare(const TerrainBase& b, const orc_edit* edit) {
    Prepared<T> p;
    p.b = &b;
    size_t len = (size_t)(b.ny + 2) * b.P();
    p.fx0.assign(len, T(0)); p.fx1.assign(len, T(0)); p.fy0.assign(len, T(0)); p.fy1.assign(len, T(0));
    p.wet.assign(len, T(0)); p.gn2.assign(len, T(0));
    const int nx = b.nx, ny = b.ny, lv = b.level;
    EditView none{nullptr};
    auto zc = [&](int c, int r) { return b.z[(size_t)r * nx + c]; };
    for (int j = 1; j <= ny; j++)   // terrain.rs:276-282
        for (int i = 1; i < nx; i++) {
            double a, bb;
            face_x(*b.fine, lv, b.crest, zc(i - 1, j - 1), zc(i, j - 1), i, j - 1, none, a, bb);
            p.fx0[b.at(i, j)] = (T)a;
            p.fx1[b.at(i, j)] = (T)bb;
        }
    for (int j = 1; j < ny; j++)    // terrain.rs:283-289
        for (int i = 1; i <= nx; i++) {
            double a, bb;
            face_y(*b.fine, lv, b.crest, zc(i - 1, j - 1), zc(i - 1, j), i - 1, j, none, a, bb);
            p.fy0[b.at(i, j)] = (T)a;
            p.fy1[b.at(i, j)] = (T)bb;
        }
    for (int j = 1; j <= ny; j++)   // terrain.rs:293-301
        for (int i = 1; i <= nx; i++) {
            size_t k = (size_t)(j - 1) * nx + (i - 1);
            p.wet[b.at(i, j)] = b.wall[k] ? T(0) : T(1);
            p.gn2[b.at(i, j)] = (T)(G_ACC * b.nval[k] * b.nval[k]);
        }
    p.zlev = b.z;
    EditView ev{edit};
    if (ev.empty()) return p;
    // apply_edit (terrain.rs:460-505)
    const Fine& f = *b.fine;
    int c0 = edit->c0 / lv, r0 = edit->r0 / lv;
    int c1 = std::min((edit->c0 + edit->w + lv - 1) / lv, nx);
    int r1 = std::min((edit->r0 + edit->h + lv - 1) / lv, ny);
    std::map<std::pair<int, int>, double> zloc;
    for (int r = r0; r < r1; r++)
        for (int c = c0; c < c1; c++) {
            double z; bool w; double nn;
            coarse_cell(f, lv, c, r, ev, z, w, nn);
            zloc[{c, r}] = z;
            p.zlev[(size_t)r * nx + c] = z;
        }
    auto zce = [&](int c, int r) {
        auto it = zloc.find({c, r});
        return it == zloc.end() ? b.z[(size_t)r * nx + c] : it->second;
    };
    for (int r = r0; r < r1; r++)
        for (int cl = std::max(c0 - 1, 0); cl < std::min(c1, nx - 1); cl++) {
            int i = cl + 1;
            double a, bb;
            face_x(f, lv, b.crest, zce(cl, r), zce(cl + 1, r), i, r, ev, a, bb);
            p.fx0[b.at(i, r + 1)] = (T)a;
            if (b.crest) p.fx1[b.at(i, r + 1)] = (T)bb;
        }
    for (int rt = std::max(r0 - 1, 0); rt < std::min(r1, ny - 1); rt++)
        for (int c = c0; c < c1; c++) {
            double a, bb;
            face_y(f, lv, b.crest, zce(c, rt), zce(c, rt + 1), c, rt + 1, ev, a, bb);
            p.fy0[b.at(c + 1, rt + 1)] = (T)a;
            if (b.crest) p.fy1[b.at(c + 1, rt + 1)] = (T)bb;
        }
    return p;
}

TerrainBase* base_build(std::shared_ptr<Fine> fine, int level, const orc_segment* segs, int nsegs) {  // terrain.rs:262-321
    auto* b = new TerrainBase();
    int lv = std::max(level, 1);
    b->level = lv;
    b->nx = (fine->nx + lv - 1) / lv;
    b->ny = (fine->ny + lv - 1) / lv;
    b->dx = fine->dx * lv;
    b->dy = fine->dy * lv;
    b->g = G_ACC;
    b->crest = lv != 1;
    b->fine = fine;
    int nx = b->nx, ny = b->ny;
    b->z.resize((size_t)nx * ny);
    b->wall.resize((size_t)nx * ny);
    b->nval.resize((size_t)nx * ny);
    EditView none{nullptr};
    for (int r = 0; r < ny; r++)
        for (int c = 0; c < nx; c++) {
            double zz, nn; bool w;
            coarse_cell(*fine, lv, c, r, none, zz, w, nn);
            b->z[(size_t)r * nx + c] = zz;
            b->wall[(size_t)r * nx + c] = w;
            b->nval[(size_t)r * nx + c] = nn;
        }
    b->wet_count = 0;
    for (auto w : b->wall) b->wet_count += w ? 0 : 1;
    // Boundaries::from_segments (terrain.rs:151-187)
    b->north.assign(nx, Ghost()); b->south.assign(nx, Ghost());
    b->west.assign(ny, Ghost()); b->east.assign(ny, Ghost());
    for (int s = 0; s < nsegs; s++) {
        const orc_segment& sg = segs[s];
        bool ns = sg.side == 0 || sg.side == 1;
        int n = ns ? nx : ny;
        double d = ns ? b->dx : b->dy;
        int a = 0, e = n;
        if (sg.has_range) {
            a = std::min((int)(size_t)std::fmax(std::floor(sg.range_a / d), 0.0), n);
            e = std::min((int)(size_t)std::fmax(std::ceil(sg.range_b / d), 0.0), n);
        }
        double len_m = (double)(e > a ? e - a : 0) * d;
        Ghost g;
        if (sg.kind == 1) { g.kind = G_STAGE; g.stage = sg.stage_m - fine->z_ref; }
        else if (sg.kind == 0) { g.kind = G_TRANS; }
        else {
            Hyeto h;
            h.t.assign(sg.hydro_t, sg.hydro_t + sg.n_hydro);
            h.r.assign(sg.hydro_q, sg.hydro_q + sg.n_hydro);
            b->hydro.push_back(h);
            g.kind = G_INFLOW;
            g.idx = (int)b->hydro.size() - 1;
            g.inv_len = len_m > 0.0 ? 1.0 / len_m : 0.0;
        }
        std::vector<Ghost>& v = sg.side == 0 ? b->north : sg.side == 1 ? b->south : sg.side == 2 ? b->west : b->east;
        for (int k = a; k < e && k < (int)v.size(); k++) v[k] = g;
    }
    return b;
}

// ---------------------------------------------------------------- boundary.rs:12-87
template <class T>
void fill_ghosts(const TerrainBase& b, const std::vector<double>& zlev, double t, double inflow_scale, T* h, T* qx,
                 T* qy, std::vector<uint8_t>& ghost_wet) {
    const int nx = b.nx, ny = b.ny;
    const double g = b.g;
    auto inflow_q = [&](int idx, double inv_len) { return b.hydro[idx].rate(t) * inv_len * inflow_scale; };
    auto crit = [&](double q) {
        double x = q * q / g;
        return x > 0.0 ? x * rcbrt(x) * rcbrt(x) : 0.0;
    };
    std::fill(ghost_wet.begin(), ghost_wet.end(), 0);
    for (int j = 1; j <= ny; j++) {
        for (int side = 0; side < 2; side++) {
            int gi = side == 0 ? 0 : nx + 1, ii = side == 0 ? 1 : nx;
            double inward = side == 0 ? 1.0 : -1.0;
            const Ghost& sp = side == 0 ? b.west[j - 1] : b.east[j - 1];
            size_t gk = b.at(gi, j), ik = b.at(ii, j);
            T hi = h[ik], qxi = qx[ik], qyi = qy[ik];
            T hg, qxg, qyg;
            switch (sp.kind) {
                case G_WALL: hg = hi; qxg = -qxi; qyg = qyi; break;
                case G_TRANS: hg = hi; qxg = qxi; qyg = qyi; break;
                case G_STAGE: hg = (T)std::fmax(sp.stage - zlev[(size_t)(j - 1) * nx + (ii - 1)], 0.0); qxg = qxi; qyg = qyi; break;
                default: {
                    double q = inflow_q(sp.idx, sp.inv_len);
                    double hh = std::fmax((double)hi, crit(q));
                    hg = (T)hh; qxg = (T)(inward * q); qyg = T(0);
                }
            }
            h[gk] = hg; qx[gk] = qxg; qy[gk] = qyg;
            if (hg > T(0)) ghost_wet[j] = 1;
        }
    }
    for (int i = 1; i <= nx; i++) {
        for (int side = 0; side < 2; side++) {
            int gj = side == 0 ? 0 : ny + 1, ij = side == 0 ? 1 : ny;
            double inward = side == 0 ? 1.0 : -1.0;
            const Ghost& sp = side == 0 ? b.north[i - 1] : b.south[i - 1];
            size_t gk = b.at(i, gj), ik = b.at(i, ij);
            T hi = h[ik], qxi = qx[ik], qyi = qy[ik];
            T hg, qxg, qyg;
            switch (sp.kind) {
                case G_WALL: hg = hi; qxg = qxi; qyg = -qyi; break;
                case G_TRANS: hg = hi; qxg = qxi; qyg = qyi; break;
                case G_STAGE: hg = (T)std::fmax(sp.stage - zlev[(size_t)(ij - 1) * nx + (i - 1)], 0.0); qxg = qxi; qyg = qyi; break;
                default: {
                    double q = inflow_q(sp.idx, sp.inv_len);
                    double hh = std::fmax((double)hi, crit(q));
                    hg = (T)hh; qxg = T(0); qyg = (T)(inward * q);
                }
            }
            h[gk] = hg; qx[gk] = qxg; qy[gk] = qyg;
            if (hg > T(0)) ghost_wet[gj] = 1;
        }
    }
    const std::pair<int, int> corners[4] = {{0, 0}, {nx + 1, 0}, {0, ny + 1}, {nx + 1, ny + 1}};
    for (const auto& ij : corners) {
        size_t k = b.at(ij.first, ij.second);
        h[k] = T(0); qx[k] = T(0); qy[k] = T(0);
    }
}

// ---------------------------------------------------------------- kernel.rs
template <class T> struct VelRow {
    std::vector<T> u, v, c;
    void init(size_t n) { u.assign(n, T(0)); v.assign(n, T(0)); c.assign(n, T(0)); }
};
template <class T> struct FaceRow {
    std::vector<T> m, nl, nr, t;
    void init(size_t n) { m.assign(n, T(0)); nl.assign(n, T(0)); nr.assign(n, T(0)); t.assign(n, T(0)); }
    void zero() { std::fill(m.begin(), m.end(), T(0)); std::fill(nl.begin(), nl.end(), T(0));
                  std::fill(nr.begin(), nr.end(), T(0)); std::fill(t.begin(), t.end(), T(0)); }
};

struct RowAcc {  // kernel.rs:73-84
    double infil = 0, fix = 0, bnd_in = 0, bnd_out = 0, smax = 0, sum_s = 0;
    bool wet = false, bad = false, skipped = false;
};

struct StepStats {  // kernel.rs:323-334
    double smax = 0, sum_s = 0, infil = 0, fix = 0, bnd_in = 0, bnd_out = 0;
    bool bad = false;
    uint64_t rows_skipped = 0;
};

struct Monitors {  // kernel.rs:86-113
    int nx;
    std::vector<uint32_t> cells;
    std::vector<size_t> row_start;
    std::vector<double> max, tmax;
    void init(const uint32_t* c, int n, int nx_, int ny) {
        nx = nx_;
        cells.assign(c, c + n);
        row_start.assign(ny + 1, 0);
        size_t k = 0;
        for (int r = 0; r < ny; r++) {
            row_start[r] = k;
            while (k < cells.size() && (int)(cells[k] / nx) == r) k++;
        }
        row_start[ny] = cells.size();
        max.assign(n, 0.0);
        tmax.assign(n, 0.0);
    }
};

// The row loops take raw __restrict__ pointers and compile-time face modes so that the
// compiler vectorises them (IEEE single operations lane by lane: the same results as the
// scalar loop, since nothing is contracted or reassociated).
template <class T> void vel_row(const T* h, const T* qx, const T* qy, VelRow<T>& out, size_t n, const StepConsts<T>& k) {  // kernel.rs:115-126
    const StepConsts<T> K = k;
    const T* __restrict__ H = h;
    const T* __restrict__ QX = qx;
    const T* __restrict__ QY = qy;
    T* __restrict__ U = out.u.data();
    T* __restrict__ V = out.v.data();
    T* __restrict__ Cc = out.c.data();
#pragma GCC ivdep
    for (size_t i = 0; i < n; i++) {
        Vel<T> o = cell_vel(H[i], QX[i], QY[i], K);
        U[i] = o.u; V[i] = o.v; Cc[i] = o.c;
    }
}

template <class T, bool CREST, bool NX>
void face_loop(const T* __restrict__ hl, const T* __restrict__ ul, const T* __restrict__ vl, const T* __restrict__ cl,
               const T* __restrict__ hr, const T* __restrict__ ur, const T* __restrict__ vr, const T* __restrict__ cr,
               const T* __restrict__ f0, const T* __restrict__ f1, int i0, int i1, T* __restrict__ om,
               T* __restrict__ onl, T* __restrict__ onr, T* __restrict__ ot, const StepConsts<T> K) {
#pragma GCC ivdep
    for (int i = i0; i <= i1; i++) {
        Vel<T> a{ul[i], vl[i], cl[i]}, b{ur[i], vr[i], cr[i]};
        Face<T> f = CREST ? face_crest(hl[i], a, hr[i], b, f0[i], f1[i], NX, K)
                          : face_signed(hl[i], a, hr[i], b, f0[i], NX, K);
        om[i] = f.m; onl[i] = f.n_l; onr[i] = f.n_r; ot[i] = f.t;
    }
}

template <class T>
void xface_row(bool crest, const T* h, const VelRow<T>& v, const T* f0, const T* f1, int nx, FaceRow<T>& out,
               const StepConsts<T>& k) {  // kernel.rs:128-149: face i joins padded cells i and i+1
    auto* o = &out;
    if (crest)
        face_loop<T, true, true>(h, v.u.data(), v.v.data(), v.c.data(), h + 1, v.u.data() + 1, v.v.data() + 1,
                                 v.c.data() + 1, f0, f1, 0, nx, o->m.data(), o->nl.data(), o->nr.data(), o->t.data(), k);
    else
        face_loop<T, false, true>(h, v.u.data(), v.v.data(), v.c.data(), h + 1, v.u.data() + 1, v.v.data() + 1,
                                  v.c.data() + 1, f0, f1, 0, nx, o->m.data(), o->nl.data(), o->nr.data(), o->t.data(), k);
}

template <class T>
void yface_row(bool crest, const T* ht, const VelRow<T>& vt, const T* hb, const VelRow<T>& vb, const T* f0,
               const T* f1, int nx, FaceRow<T>& out, const StepConsts<T>& k) {  // kernel.rs:151-179
    if (crest)
        face_loop<T, true, false>(ht, vt.u.data(), vt.v.data(), vt.c.data(), hb, vb.u.data(), vb.v.data(), vb.c.data(),
                                  f0, f1, 1, nx, out.m.data(), out.nl.data(), out.nr.data(), out.t.data(), k);
    else
        face_loop<T, false, false>(ht, vt.u.data(), vt.v.data(), vt.c.data(), hb, vb.u.data(), vb.v.data(), vb.c.data(),
                                   f0, f1, 1, nx, out.m.data(), out.nl.data(), out.nr.data(), out.t.data(), k);
}

template <class T> struct RowScratch {
    std::vector<T> infil, fix, s;
    void init(int nx) { size_t n = ((size_t)nx + 2 + 15) / 16 * 16; infil.assign(n, T(0)); fix.assign(n, T(0)); s.assign(n, T(0)); }
};

// update_cells + update_row (kernel.rs:181-309): returns RowAcc and sum_s.
template <class T>
RowAcc update_row(T* h, T* qx, T* qy, const FaceRow<T>& fx, const FaceRow<T>& gn, const FaceRow<T>& gs, const T* wet,
                  const T* gn2, int nx, const StepConsts<T>& k, RowScratch<T>& rs, bool want_sum_s, double& sum_s) {
    const T z = T(0);
    const StepConsts<T> K = k;
    T* __restrict__ H = h;
    T* __restrict__ QX = qx;
    T* __restrict__ QY = qy;
    {
        const T* __restrict__ xm = fx.m.data();
        const T* __restrict__ xnl = fx.nl.data();
        const T* __restrict__ xnr = fx.nr.data();
        const T* __restrict__ xt = fx.t.data();
        const T* __restrict__ nm = gn.m.data();
        const T* __restrict__ nnr = gn.nr.data();
        const T* __restrict__ nt = gn.t.data();
        const T* __restrict__ sm = gs.m.data();
        const T* __restrict__ snl = gs.nl.data();
        const T* __restrict__ st = gs.t.data();
#pragma GCC ivdep
        for (int c = 1; c <= nx; c++) {   // loop 1 (kernel.rs:214-223)
            Face<T> fw{xm[c - 1], z, xnr[c - 1], xt[c - 1]};
            Face<T> fe{xm[c], xnl[c], z, xt[c]};
            Face<T> fnn{nm[c], z, nnr[c], nt[c]};
            Face<T> fs{sm[c], snl[c], z, st[c]};
            T a, b, cc;
            flux_update(H[c], QX[c], QY[c], fw, fe, fnn, fs, K, a, b, cc);
            H[c] = a; QX[c] = b; QY[c] = cc;
        }
    }
    int ok_all = 1;
    {
        const T* __restrict__ WET = wet;
        const T* __restrict__ GN2 = gn2;
        T* __restrict__ RI = rs.infil.data();
        T* __restrict__ RF = rs.fix.data();
        T* __restrict__ RS = rs.s.data();
#pragma GCC ivdep
        for (int c = 1; c <= nx; c++) {   // loop 2 (kernel.rs:225-234)
            CellOut<T> o = sources(H[c], QX[c], QY[c], WET[c], GN2[c], K);
            H[c] = o.h; QX[c] = o.qx; QY[c] = o.qy;
            RI[c] = o.infil; RF[c] = o.fix; RS[c] = o.s;
            ok_all &= o.ok ? 1 : 0;
        }
    }
    const bool bad = ok_all == 0;
    size_t n = (size_t)nx + 2;
    FixedSum<T> si, sf;
    si.add_slice(rs.infil.data(), n);
    sf.add_slice(rs.fix.data(), n);
    sum_s = 0.0;
    if (want_sum_s) { FixedSum<T> ss; ss.add_slice(rs.s.data(), n); sum_s = (double)ss.total(); }
    RowAcc a;
    a.wet = max_slice(h + 1, n - 2) > T(0);
    a.infil = (double)si.total();
    a.fix = (double)sf.total();
    a.smax = (double)max_slice(rs.s.data(), n);
    a.bad = bad;
    return a;
}

template <class T> void monitors_row(Monitors& mon, int row, const T* h, double t_new) {  // kernel.rs:311-321
    for (size_t kk = mon.row_start[row]; kk < mon.row_start[row + 1]; kk++) {
        int col = (int)(mon.cells[kk] % (uint32_t)mon.nx);
        double v = (double)h[col + 1];
        if (v > mon.max[kk]) { mon.max[kk] = v; mon.tmax[kk] = t_new; }
    }
}

template <class T> void x_boundary(const FaceRow<T>& fx, int nx, double dt_dy, RowAcc& acc) {  // kernel.rs:436-450
    double w = (double)fx.m[0] * dt_dy;
    double e = (double)fx.m[nx] * dt_dy;
    if (w > 0.0) acc.bnd_in += w; else acc.bnd_out -= w;
    if (e > 0.0) acc.bnd_out += e; else acc.bnd_in -= e;
}

template <class T> std::pair<double, double> y_boundary(const FaceRow<T>& row, int nx, double dt_dx, bool top) {  // kernel.rs:453-472
    FixedSum<double> s;
    for (int i = 1; i <= nx; i++) s.add((size_t)i, (double)row.m[i]);
    double v = s.total() * dt_dx;
    double fin = 0.0, fout = 0.0;
    if (top) { if (v > 0.0) fin = v; else fout = -v; }
    else if (v > 0.0) fout = v;
    else fin = -v;
    return {fin, fout};
}

StepStats finalize(const std::vector<RowAcc>& accs, std::pair<double, double> top, std::pair<double, double> bot,
                   const std::vector<double>& sum_s_rows) {  // kernel.rs:485-500
    StepStats st;
    st.bnd_in = top.first + bot.first;
    st.bnd_out = top.second + bot.second;
    for (const RowAcc& a : accs) {
        st.infil += a.infil;
        st.fix += a.fix;
        st.bnd_in += a.bnd_in;
        st.bnd_out += a.bnd_out;
        st.smax = std::fmax(st.smax, a.smax);
        st.bad = st.bad || a.bad;
        st.rows_skipped += a.skipped ? 1 : 0;
    }
    for (double s : sum_s_rows) st.sum_s += s;
    return st;
}

template <class T> struct Workspace {
    int nx, ny;
    size_t P;
    std::vector<T> h, qx, qy;
    std::vector<uint8_t> row_dry, ghost_wet, new_dry;
    std::vector<double> sum_s_rows;
    int strip_rows;
    // fused scratch
    VelRow<T> vel_cur, vel_next, tmp_vt, tmp_vb;
    FaceRow<T> fx, g_prev, g_next;
    std::vector<FaceRow<T>> bnd_rows;
    RowScratch<T> rs;
    std::vector<RowAcc> accs;
    // oracle scratch
    std::vector<VelRow<T>> o_vel;
    std::vector<FaceRow<T>> o_fx, o_fy;
    Monitors mon;
    void init(int nx_, int ny_, int strip) {
        nx = nx_; ny = ny_; P = (size_t)nx + 2;
        size_t len = (size_t)(ny + 2) * P;
        h.assign(len, T(0)); qx.assign(len, T(0)); qy.assign(len, T(0));
        row_dry.assign(ny + 2, 0); ghost_wet.assign(ny + 2, 0); new_dry.assign(ny + 2, 0);
        sum_s_rows.assign(ny, 0.0);
        strip_rows = std::min(std::max(strip, 1), std::max(ny, 1));
        vel_cur.init(P); vel_next.init(P); tmp_vt.init(P); tmp_vb.init(P);
        fx.init(P); g_prev.init(P); g_next.init(P);
        int n_strips = (ny + strip_rows - 1) / strip_rows;
        bnd_rows.resize(n_strips + 1);
        for (auto& r : bnd_rows) r.init(P);
        rs.init(nx);
        accs.assign(ny, RowAcc());
    }
    T* row(std::vector<T>& a, int j) { return a.data() + (size_t)j * P; }
};

struct StepCtx { double t_new, dt; bool want_sum_s, skip_dry; };

template <class T>
StepStats step_oracle(const Prepared<T>& pt, Workspace<T>& ws, const StepConsts<T>& k, const StepCtx& cx) {  // kernel.rs:510-570
    const int nx = ws.nx, ny = ws.ny;
    const bool crest = pt.b->crest;
    if ((int)ws.o_vel.size() != ny + 2) {
        ws.o_vel.resize(ny + 2); for (auto& v : ws.o_vel) v.init(ws.P);
        ws.o_fx.resize(ny + 2); for (auto& v : ws.o_fx) v.init(ws.P);
        ws.o_fy.resize(ny + 1); for (auto& v : ws.o_fy) v.init(ws.P);
    }
    for (int j = 0; j < ny + 2; j++) vel_row(ws.row(ws.h, j), ws.row(ws.qx, j), ws.row(ws.qy, j), ws.o_vel[j], ws.P, k);
    for (int j = 1; j <= ny; j++)
        xface_row(crest, ws.row(ws.h, j), ws.o_vel[j], pt.fx0.data() + j * ws.P, pt.fx1.data() + j * ws.P, nx, ws.o_fx[j], k);
    for (int j = 0; j <= ny; j++)
        yface_row(crest, ws.row(ws.h, j), ws.o_vel[j], ws.row(ws.h, j + 1), ws.o_vel[j + 1], pt.fy0.data() + j * ws.P,
                  pt.fy1.data() + j * ws.P, nx, ws.o_fy[j], k);
    double dt_dy = cx.dt * pt.b->dy;
    double dt_dx = cx.dt * pt.b->dx;
    std::vector<RowAcc> accs;
    accs.reserve(ny);
    for (int j = 1; j <= ny; j++) {
        double ss;
        RowAcc a = update_row(ws.row(ws.h, j), ws.row(ws.qx, j), ws.row(ws.qy, j), ws.o_fx[j], ws.o_fy[j - 1], ws.o_fy[j],
                              pt.wet.data() + j * ws.P, pt.gn2.data() + j * ws.P, nx, k, ws.rs, cx.want_sum_s, ss);
        x_boundary(ws.o_fx[j], nx, dt_dy, a);
        monitors_row(ws.mon, j - 1, ws.row(ws.h, j), cx.t_new);
        ws.row_dry[j] = !a.wet;
        ws.sum_s_rows[j - 1] = ss;
        accs.push_back(a);
    }
    ws.row_dry[0] = 1;
    ws.row_dry[ny + 1] = 1;
    auto top = y_boundary(ws.o_fy[0], nx, dt_dx, true);
    auto bot = y_boundary(ws.o_fy[ny], nx, dt_dx, false);
    return finalize(accs, top, bot, ws.sum_s_rows);
}

template <class T>
StepStats step_fused(const Prepared<T>& pt, Workspace<T>& ws, const StepConsts<T>& k, const StepCtx& cx) {  // kernel.rs:590-727
    const int nx = ws.nx, ny = ws.ny;
    const bool crest = pt.b->crest;
    const int s_rows = ws.strip_rows;
    const int n_strips = (ny + s_rows - 1) / s_rows;
    const bool rain_zero = !(k.rain > T(0));
    std::vector<uint8_t>& dry = ws.row_dry;
    for (int j = 0; j < ny + 2; j++) dry[j] = dry[j] && !ws.ghost_wet[j];
    auto skippable = [&](int j) { return cx.skip_dry && rain_zero && dry[j - 1] && dry[j] && dry[j + 1]; };
    const size_t P = ws.P;
    auto F0 = [&](const std::vector<T>& a, int j) { return a.data() + (size_t)j * P; };
    for (int s = 0; s <= n_strips; s++) {  // pre-phase (kernel.rs:614-622)
        int a = s < n_strips ? 1 + s * s_rows : ny + 1;
        vel_row(ws.row(ws.h, a - 1), ws.row(ws.qx, a - 1), ws.row(ws.qy, a - 1), ws.tmp_vt, P, k);
        vel_row(ws.row(ws.h, a), ws.row(ws.qx, a), ws.row(ws.qy, a), ws.tmp_vb, P, k);
        yface_row(crest, ws.row(ws.h, a - 1), ws.tmp_vt, ws.row(ws.h, a), ws.tmp_vb, F0(pt.fy0, a - 1), F0(pt.fy1, a - 1), nx,
                  ws.bnd_rows[s], k);
    }
    double dt_dy = cx.dt * pt.b->dy;
    double dt_dx = cx.dt * pt.b->dx;
    auto top = y_boundary(ws.bnd_rows[0], nx, dt_dx, true);
    auto bot = y_boundary(ws.bnd_rows[n_strips], nx, dt_dx, false);
    for (int s = 0; s < n_strips; s++) {   // strips, in order (results do not depend on it)
        int a = 1 + s * s_rows;
        int b = std::min(a + s_rows, ny + 1);
        FaceRow<T>* gp = &ws.g_prev;
        FaceRow<T>* gn = &ws.g_next;
        VelRow<T>* vc = &ws.vel_cur;
        VelRow<T>* vn = &ws.vel_next;
        *gp = ws.bnd_rows[s];
        bool cur_valid = false;
        for (int j = a; j < b; j++) {
            if (skippable(j)) {
                ws.accs[j - 1] = RowAcc();
                ws.accs[j - 1].skipped = true;
                ws.new_dry[j] = 1;
                ws.sum_s_rows[j - 1] = 0.0;
                gn->zero();
                std::swap(gp, gn);
                cur_valid = false;
                continue;
            }
            if (!cur_valid) vel_row(ws.row(ws.h, j), ws.row(ws.qx, j), ws.row(ws.qy, j), *vc, P, k);
            if (j + 1 < b) {
                vel_row(ws.row(ws.h, j + 1), ws.row(ws.qx, j + 1), ws.row(ws.qy, j + 1), *vn, P, k);
                yface_row(crest, ws.row(ws.h, j), *vc, ws.row(ws.h, j + 1), *vn, F0(pt.fy0, j), F0(pt.fy1, j), nx, *gn, k);
                cur_valid = true;
            } else {
                *gn = ws.bnd_rows[s + 1];
                cur_valid = false;
            }
            xface_row(crest, ws.row(ws.h, j), *vc, F0(pt.fx0, j), F0(pt.fx1, j), nx, ws.fx, k);
            double ss;
            RowAcc acc = update_row(ws.row(ws.h, j), ws.row(ws.qx, j), ws.row(ws.qy, j), ws.fx, *gp, *gn, F0(pt.wet, j),
                                    F0(pt.gn2, j), nx, k, ws.rs, cx.want_sum_s, ss);
            x_boundary(ws.fx, nx, dt_dy, acc);
            monitors_row(ws.mon, j - 1, ws.row(ws.h, j), cx.t_new);
            ws.new_dry[j] = !acc.wet;
            ws.sum_s_rows[j - 1] = ss;
            ws.accs[j - 1] = acc;
            std::swap(gp, gn);
            if (cur_valid) std::swap(vc, vn);
        }
    }
    ws.new_dry[0] = 1;
    ws.new_dry[ny + 1] = 1;
    std::swap(ws.new_dry, ws.row_dry);
    return finalize(ws.accs, top, bot, ws.sum_s_rows);
}

// ---------------------------------------------------------------- objective.rs:70-149
// ---------------------------------------------------------------- inertial.rs (screening physics)
// Local-inertial step on staggered discharges: qx at (i, j) is the east face of padded
// cell i (i = 0: west boundary face), qy at (i, j) the south face of padded row j
// (j = 0: north boundary face).  f64 arithmetic on the stored state.
inline void li_rises(bool crest, double f0, double f1, double& dl, double& dr) {  // inertial.rs:25-35
    if (crest) { dl = f0; dr = f1; return; }
    dl = std::fmax(f0, 0.0);
    dr = std::fmax(-f0, 0.0);
}
inline double li_jump(bool crest, double f0, double f1) { return crest ? f0 - f1 : f0; }  // inertial.rs:37-43

inline double li_face_q(double q, double hl, double hr, double dl, double dr, double d, double gn2, double g, double dt,
                        double h_dry) {  // inertial.rs:46-58
    const double el = hl - dl, er = hr - dr;
    const double hf = std::fmax(std::fmax(el, er), 0.0);
    if (hf <= h_dry) return 0.0;
    const double slope = (er - el) / d;
    const double r = rcbrt(hf);
    const double r7 = r * r * r * r * r * r * r;
    return (q - g * hf * dt * slope) / (1.0 + gn2 * dt * std::fabs(q) * r7);
}

template <class T>
StepStats step_local_inertial(const Prepared<T>& pt, Workspace<T>& ws, const StepConsts<T>& k, const StepCtx& cx,
                              std::vector<double>& old) {  // inertial.rs:60-219
    const TerrainBase& b = *pt.b;
    const int nx = b.nx, ny = b.ny;
    const double dx = b.dx, dy = b.dy, g = b.g, dt = cx.dt;
    const double h_dry = (double)k.h_dry;
    const bool crest = b.crest;
    auto at = [&](int i, int j) { return b.at(i, j); };
    auto gn2 = [&](int i, int j) { return (double)pt.gn2[at(i, j)]; };
    auto hv = [&](int i, int j) { return (double)ws.h[at(i, j)]; };
    const size_t len = ws.h.size();
    old.resize(2 * len);
    for (size_t q = 0; q < len; q++) {
        old[q] = (double)ws.qx[q];
        old[len + q] = (double)ws.qy[q];
    }
    const double* ox = old.data();
    const double* oy = old.data() + len;
    const double TH = 0.8;   // de Almeida et al. (2012) weighting of the old discharge
    auto qtx = [&](int i, int j) {
        const double c = ox[at(i, j)];
        return (i >= 1 && i + 1 < nx) ? TH * c + 0.5 * (1.0 - TH) * (ox[at(i - 1, j)] + ox[at(i + 1, j)]) : c;
    };
    auto qty = [&](int i, int j) {
        const double c = oy[at(i, j)];
        return (j >= 1 && j + 1 < ny) ? TH * c + 0.5 * (1.0 - TH) * (oy[at(i, j - 1)] + oy[at(i, j + 1)]) : c;
    };
    for (int j = 1; j <= ny; j++)   // x-faces 0 ..= nx
        for (int i = 0; i <= nx; i++) {
            const size_t kk = at(i, j);
            double q;
            if (i == 0 || i == nx) {
                const Ghost& spec = i == 0 ? b.west[j - 1] : b.east[j - 1];
                const double inward = i == 0 ? 1.0 : -1.0;
                if (spec.kind == G_WALL) {
                    q = 0.0;
                } else if (spec.kind == G_TRANS) {   // normal-depth outflow only
                    const int ci = i == 0 ? 1 : nx, fi = i == 0 ? 1 : nx - 1;
                    double dz = nx > 1 ? li_jump(crest, (double)pt.fx0[at(fi, j)], (double)pt.fx1[at(fi, j)]) : 0.0;
                    dz = i == 0 ? std::fmax(dz, 0.0) : std::fmin(dz, 0.0);
                    const double hc = hv(ci, j);
                    q = li_face_q((double)ws.qx[kk], hc, hc, std::fmax(dz, 0.0), std::fmax(-dz, 0.0), dx, gn2(ci, j), g, dt,
                                  h_dry);
                    q = i == 0 ? std::fmin(q, 0.0) : std::fmax(q, 0.0);
                } else if (spec.kind == G_INFLOW) {
                    q = inward * std::fabs((double)ws.qx[at(i == 0 ? 0 : nx + 1, j)]);
                } else {   // stage
                    q = li_face_q((double)ws.qx[kk], hv(i, j), hv(i + 1, j), 0.0, 0.0, dx, gn2(std::min(std::max(i, 1), nx), j),
                                  g, dt, h_dry);
                }
            } else {
                double dl, dr;
                li_rises(crest, (double)pt.fx0[kk], (double)pt.fx1[kk], dl, dr);
                const double n2 = 0.5 * (gn2(i, j) + gn2(i + 1, j));
                q = li_face_q(qtx(i, j), hv(i, j), hv(i + 1, j), dl, dr, dx, n2, g, dt, h_dry);
            }
            ws.qx[kk] = canon((T)q);
        }
    for (int j = 0; j <= ny; j++)   // y-faces 0 ..= ny, +q southward
        for (int i = 1; i <= nx; i++) {
            const size_t kk = at(i, j);
            double q;
            if (j == 0 || j == ny) {
                const Ghost& spec = j == 0 ? b.north[i - 1] : b.south[i - 1];
                const double inward = j == 0 ? 1.0 : -1.0;
                if (spec.kind == G_WALL) {
                    q = 0.0;
                } else if (spec.kind == G_TRANS) {
                    const int cj = j == 0 ? 1 : ny, fj = j == 0 ? 1 : ny - 1;
                    double dz = ny > 1 ? li_jump(crest, (double)pt.fy0[at(i, fj)], (double)pt.fy1[at(i, fj)]) : 0.0;
                    dz = j == 0 ? std::fmax(dz, 0.0) : std::fmin(dz, 0.0);
                    const double hc = hv(i, cj);
                    q = li_face_q((double)ws.qy[kk], hc, hc, std::fmax(dz, 0.0), std::fmax(-dz, 0.0), dy, gn2(i, cj), g, dt,
                                  h_dry);
                    q = j == 0 ? std::fmin(q, 0.0) : std::fmax(q, 0.0);
                } else if (spec.kind == G_INFLOW) {
                    q = inward * std::fabs((double)ws.qy[at(i, j == 0 ? 0 : ny + 1)]);
                } else {
                    q = li_face_q((double)ws.qy[kk], hv(i, j), hv(i, j + 1), 0.0, 0.0, dy, gn2(i, std::min(std::max(j, 1), ny)),
                                  g, dt, h_dry);
                }
            } else {
                double dl, dr;
                li_rises(crest, (double)pt.fy0[kk], (double)pt.fy1[kk], dl, dr);
                const double n2 = 0.5 * (gn2(i, j) + gn2(i, j + 1));
                q = li_face_q(qty(i, j), hv(i, j), hv(i, j + 1), dl, dr, dy, n2, g, dt, h_dry);
            }
            ws.qy[kk] = canon((T)q);
        }
    // continuity and sources, ledger, CFL rate, monitors
    StepStats st;
    const double rain = (double)k.rain, cap = (double)k.cap;
    const double rx = dt / dx, ry = dt / dy;
    for (int j = 1; j <= ny; j++) {
        bool any = false;
        for (int i = 1; i <= nx; i++) {
            const size_t kk = at(i, j);
            const double wet = (double)pt.wet[kk];
            const double qw = (double)ws.qx[at(i - 1, j)], qe = (double)ws.qx[kk];
            const double qn = (double)ws.qy[at(i, j - 1)], qs = (double)ws.qy[kk];
            double hn = (double)ws.h[kk] + rx * (qw - qe) + ry * (qn - qs) + rain * wet;
            const double infil = std::fmin(cap, std::fmax(hn, 0.0)) * wet;
            hn -= infil;
            st.infil += infil;
            if (hn < 0.0) {
                st.fix -= hn;
                hn = 0.0;
            }
            if (!std::isfinite(hn)) st.bad = true;
            ws.h[kk] = canon((T)(hn * wet));
            any = any || hn > 0.0;
            double s = 0.0;
            if (hn > h_dry) {
                const double u = 0.5 * (std::fabs(qw) + std::fabs(qe)) / hn;
                const double v = 0.5 * (std::fabs(qn) + std::fabs(qs)) / hn;
                s = (std::sqrt(g * hn) + std::fmax(u, v)) * 0.7 / std::fmin(dx, dy);
            }
            st.smax = std::fmax(st.smax, s);
            if (cx.want_sum_s) st.sum_s += s;
        }
        ws.row_dry[j] = !any;
        const double w = (double)ws.qx[at(0, j)] * dt * dy, e = (double)ws.qx[at(nx, j)] * dt * dy;
        if (w > 0.0) st.bnd_in += w; else st.bnd_out -= w;
        if (e > 0.0) st.bnd_out += e; else st.bnd_in -= e;
        for (size_t m = ws.mon.row_start[j - 1]; m < ws.mon.row_start[j]; m++) {
            const int col = (int)(ws.mon.cells[m] % (uint32_t)nx);
            const double v = (double)ws.h[at(col + 1, j)];
            if (v > ws.mon.max[m]) {
                ws.mon.max[m] = v;
                ws.mon.tmax[m] = cx.t_new;
            }
        }
    }
    for (int i = 1; i <= nx; i++) {
        const double n = (double)ws.qy[at(i, 0)] * dt * dx, s = (double)ws.qy[at(i, ny)] * dt * dx;
        if (n > 0.0) st.bnd_in += n; else st.bnd_out -= n;
        if (s > 0.0) st.bnd_out += s; else st.bnd_in -= s;
    }
    return st;
}

// Objective::running_lower_bound (objective.rs:146-149) of the observer's edit.
double objective_j(const orc_observer* o, const double* mon) {
    std::vector<int32_t> start((size_t)o->n_assets + 1, 0);
    for (int a = 0; a < o->n_assets; a++) start[a + 1] = start[a] + o->asset_len[a];
    itr_objective ob{};
    ob.n_assets = o->n_assets;
    ob.asset_start = start.data();
    ob.asset_pos = o->asset_pos;
    ob.threshold_m = o->asset_thr;
    ob.weight = o->asset_w;
    ob.tau_m = o->tau;
    itr_earthwork ew{o->edit_cut_m3, o->edit_fill_m3, o->cut_cost, o->fill_cost, o->earth_weight};
    return itr_objective_value(&ob, mon, &ew, nullptr, nullptr);
}

// ---------------------------------------------------------------- solver.rs:250-660
template <class T> double volume(Workspace<T>& ws, double area) {  // solver.rs:250-262
    double v = 0.0;
    for (int j = 1; j <= ws.ny; j++) {
        FixedSum<T> s;
        const T* row = ws.row(ws.h, j);
        for (int i = 1; i <= ws.nx; i++) s.add((size_t)i, row[i]);
        v += (double)s.total();
    }
    return v * area;
}

struct Accum {
    double v0 = 0, rain = 0, inflow = 0, outflow = 0, infil = 0, fix = 0, dt_min = DBL_MAX, dt_max = 0, dt_sum = 0, lts = 0;
    uint64_t cell_updates = 0, rows_skipped = 0;
};

template <class T> struct Snap {
    std::vector<T> h, qx, qy;
    std::vector<uint8_t> row_dry;
    std::vector<double> mon_max, mon_tmax;
    Accum acc;
    std::vector<double> rows;
    double t, dt_next, smax;
    uint64_t step;
    uint32_t sync_index;
};

// ---------------------------------------------------------------- replay.rs:17-181
// Region of interest in level-grid cells (half-open), its ghost ring and inner band in
// the ROI's padded layout, in the reference's fixed orders.
struct Roi { int c0, r0, w, h; };
std::vector<std::pair<int, int>> roi_ring(const Roi& r) {  // replay.rs:27-41
    std::vector<std::pair<int, int>> v;
    for (int i = 0; i < r.w + 2; i++) v.push_back({i, 0});
    for (int i = 0; i < r.w + 2; i++) v.push_back({i, r.h + 1});
    for (int j = 1; j <= r.h; j++) {
        v.push_back({0, j});
        v.push_back({r.w + 1, j});
    }
    return v;
}
std::vector<std::pair<int, int>> roi_band(const Roi& r) {  // replay.rs:43-54
    std::vector<std::pair<int, int>> v;
    for (int j = 1; j <= r.h; j++)
        for (int i = 1; i <= r.w; i++)
            if (i == 1 || j == 1 || i == r.w || j == r.h) v.push_back({i, j});
    return v;
}

// Tape recording during a full-domain run, or a replay on the ROI grid (solver.rs:484-493,
// 573-591).  Values are stored as f64 of the run's precision.
struct ReplayCtx {
    Roi roi{0, 0, 0, 0};
    bool record = false, replay = false;
    std::vector<size_t> ring_idx, band_idx;   // padded indices in the run's own layout
    std::vector<double> ring, band;           // record: appended; replay: the tape
    int64_t n_ring_steps = 0, n_band_syncs = 0;
    double tol_h = 0.0, tol_q = 0.0;
};

template <class T>
int run_impl(const Prepared<T>& pt, const orc_params* p, Workspace<T>& ws, const orc_observer* obs, const orc_start* st,
             orc_summary* sum, std::vector<double>* rows_out, std::vector<double>* dts_out, ReplayCtx* rp = nullptr) {
    const TerrainBase& b = *pt.b;
    const int nx = b.nx, ny = b.ny;
    const double area = b.dx * b.dy;
    Hyeto rain;
    rain.t.assign(p->rain_t, p->rain_t + p->n_rain);
    rain.r.assign(p->rain_r, p->rain_r + p->n_rain);
    Accum acc;
    double t, dt_cfl, s_cur;
    uint64_t step;
    uint32_t sync_index;
    std::vector<double> rows;   // ledger rows, 8 per sync
    if (st) {
        std::fill(ws.h.begin(), ws.h.end(), T(0));
        std::fill(ws.qx.begin(), ws.qx.end(), T(0));
        std::fill(ws.qy.begin(), ws.qy.end(), T(0));
        for (int r = 0; r < ny; r++)
            for (int c = 0; c < nx; c++) {
                size_t k = (size_t)r * nx + c, q = b.at(c + 1, r + 1);
                ws.h[q] = (T)st->h[k]; ws.qx[q] = (T)st->qx[k]; ws.qy[q] = (T)st->qy[k];
            }
        std::fill(ws.row_dry.begin(), ws.row_dry.end(), 0);
        acc.v0 = st->v0;
        t = st->t; step = (uint64_t)st->step; sync_index = (uint32_t)st->sync_index;
        dt_cfl = st->dt_next; s_cur = st->smax;
    } else {
        // init_state (solver.rs:287-312)
        std::fill(ws.h.begin(), ws.h.end(), T(0));
        std::fill(ws.qx.begin(), ws.qx.end(), T(0));
        std::fill(ws.qy.begin(), ws.qy.end(), T(0));
        for (int r = 0; r < ny; r++)
            for (int c = 0; c < nx; c++) {
                size_t k = (size_t)r * nx + c;
                if (b.wall[k]) continue;
                double d = p->initial_depth;
                if (p->has_initial_stage) d = std::fmax(d, p->initial_stage - b.fine->z_ref - pt.zlev[k]);
                ws.h[b.at(c + 1, r + 1)] = canon((T)std::fmax(d, 0.0));
            }
        for (int j = 0; j < ny + 2; j++) {
            bool all0 = true;
            const T* hr = ws.row(ws.h, j);
            for (size_t i = 0; i < ws.P; i++) if (!(hr[i] == T(0))) { all0 = false; break; }
            ws.row_dry[j] = all0;
        }
        std::fill(ws.mon.max.begin(), ws.mon.max.end(), 0.0);
        std::fill(ws.mon.tmax.begin(), ws.mon.tmax.end(), 0.0);
        acc.v0 = volume(ws, area);
        t = 0.0; step = 0; sync_index = 0;
        // initial_smax (solver.rs:315-332), initial_smax_li (solver.rs:334-344)
        double s0 = 0.0;
        if (p->kernel == 2) {
            double hmax = 0.0;
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) hmax = std::fmax(hmax, (double)ws.h[b.at(i, j)]);
            s0 = std::sqrt(b.g * hmax) * 0.7 / std::fmin(b.dx, b.dy);
        } else {
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t k = b.at(i, j);
                    double h = (double)ws.h[k];
                    if (h <= p->h_dry) continue;
                    double c = std::sqrt(b.g * h);
                    double r = (std::fabs((double)ws.qx[k]) / h + c) / b.dx + (std::fabs((double)ws.qy[k]) / h + c) / b.dy;
                    s0 = std::fmax(s0, r);
                }
        }
        s_cur = s0;
        dt_cfl = s0 > 0.0 ? std::fmin(p->cfl / s0, p->dt_max_s) : p->dt_max_s;
    }
    const uint32_t n_sync = (uint32_t)std::ceil(p->duration_s / p->sync_interval_s);
    const uint64_t cells = (uint64_t)nx * ny;
    int aborted = 0, abort_cell = 0, abort_sync = 0;
    double abort_value = 0.0, dt_fallback_t = -1.0;
    bool locked = p->locked_dt != nullptr;
    Snap<T> snap;
    auto take_snap = [&]() {
        snap.h = ws.h; snap.qx = ws.qx; snap.qy = ws.qy; snap.row_dry = ws.row_dry;
        snap.mon_max = ws.mon.max; snap.mon_tmax = ws.mon.tmax; snap.acc = acc; snap.rows = rows;
        snap.t = t; snap.step = step; snap.sync_index = sync_index; snap.dt_next = dt_cfl; snap.smax = s_cur;
    };
    if (locked) take_snap();
    std::vector<T> li_faces;
    std::vector<double> li_old;
    StepConsts<T> k;
    k.g = (T)b.g;
    k.g_half = (T)(0.5 * b.g);
    k.inv_dx = (T)(1.0 / b.dx);
    k.inv_dy = (T)(1.0 / b.dy);
    k.h_dry = (T)p->h_dry;
    k.heps2 = (T)(p->h_eps * p->h_eps);
    k.q_min = (T)1e-30;
    while (t < p->duration_s && sync_index < n_sync) {
        if (p->max_steps > 0 && step >= (uint64_t)p->max_steps) break;
        double t_sync = std::fmin((double)(sync_index + 1) * p->sync_interval_s, p->duration_s);
        double dt = std::fmin(dt_cfl, p->dt_max_s);
        if (locked) {   // solver.rs:452-482
            if (step < (uint64_t)p->n_locked && p->locked_dt[step] * s_cur <= p->cfl_hard) {
                dt = p->locked_dt[step];
            } else if (rp && rp->replay) {   // replay-invalid: locked-dt verification failed
                aborted = 5;
                abort_sync = (int)sync_index;
                break;
            } else {
                ws.h = snap.h; ws.qx = snap.qx; ws.qy = snap.qy; ws.row_dry = snap.row_dry;
                ws.mon.max = snap.mon_max; ws.mon.tmax = snap.mon_tmax; acc = snap.acc; rows = snap.rows;
                t = snap.t; step = snap.step; sync_index = snap.sync_index; dt_cfl = snap.dt_next; s_cur = snap.smax;
                if (dts_out) dts_out->resize(step);
                dt_fallback_t = t;
                locked = false;
                continue;
            }
        }
        bool hits_sync = t + dt >= t_sync - 1e-9 * p->sync_interval_s;
        if (hits_sync) dt = t_sync - t;
        double rn = rain.integral(t, t + dt);
        double t_new = hits_sync ? t_sync : t + dt;
        if (rp && rp->replay) {   // Tape::fill (replay.rs:136-159): the recorded ring is the ghost ring
            const size_t n = rp->ring_idx.size();
            if ((int64_t)step >= rp->n_ring_steps) {   // tape exhausted
                aborted = 5;
                abort_sync = (int)sync_index;
                break;
            }
            std::fill(ws.ghost_wet.begin(), ws.ghost_wet.end(), 0);
            const double* src = rp->ring.data() + (size_t)step * 3 * n;
            const int w = rp->roi.w;
            for (size_t k = 0; k < n; k++) {
                const size_t q = rp->ring_idx[k];
                ws.h[q] = (T)src[3 * k];
                ws.qx[q] = (T)src[3 * k + 1];
                ws.qy[q] = (T)src[3 * k + 2];
                if (ws.h[q] > T(0)) {
                    const int kk = (int)k;
                    const int j = kk < w + 2 ? 0 : kk < 2 * (w + 2) ? rp->roi.h + 1 : 1 + (kk - 2 * (w + 2)) / 2;
                    ws.ghost_wet[j] = 1;
                }
            }
        } else {
        if (p->kernel == 2) {   // the staggered boundary faces survive the ghost fill (inertial.rs:219-242)
            li_faces.clear();
            for (int j = 1; j <= ny; j++) li_faces.push_back(ws.qx[b.at(0, j)]);
            for (int i = 1; i <= nx; i++) li_faces.push_back(ws.qy[b.at(i, 0)]);
        }
        fill_ghosts(b, pt.zlev, t, p->inflow_scale, ws.h.data(), ws.qx.data(), ws.qy.data(), ws.ghost_wet);
        if (p->kernel == 2) {
            for (int j = 1; j <= ny; j++)
                if (b.west[j - 1].kind != G_INFLOW) ws.qx[b.at(0, j)] = li_faces[j - 1];
            for (int i = 1; i <= nx; i++)
                if (b.north[i - 1].kind != G_INFLOW) ws.qy[b.at(i, 0)] = li_faces[ny + i - 1];
        }
        }
        if (rp && rp->record)   // Tape::record_ring (replay.rs:112-125), after the ghost fill
            for (size_t q : rp->ring_idx) {
                rp->ring.push_back((double)ws.h[q]);
                rp->ring.push_back((double)ws.qx[q]);
                rp->ring.push_back((double)ws.qy[q]);
            }
        if (dts_out) dts_out->push_back(dt);
        k.dt = (T)dt;
        k.dtdx = (T)(dt / b.dx);
        k.dtdy = (T)(dt / b.dy);
        k.rain = (T)rn;
        k.cap = (T)(p->infil_capacity * dt);
        StepCtx cx{t_new, dt, hits_sync, p->skip_dry != 0};
        StepStats s;
        if (p->kernel == 2) {
            s = step_local_inertial(pt, ws, k, cx, li_old);
            if (!hits_sync) s.sum_s = 0.0;
        } else {
            s = p->kernel == 0 ? step_oracle(pt, ws, k, cx) : step_fused(pt, ws, k, cx);
        }
        step += 1;
        t = t_new;
        acc.rain += rn * area * (double)b.wet_count;
        acc.inflow += s.bnd_in;
        acc.outflow += s.bnd_out;
        acc.infil += s.infil * area;
        acc.fix += s.fix * area;
        acc.dt_min = std::fmin(acc.dt_min, dt);
        acc.dt_max = std::fmax(acc.dt_max, dt);
        acc.dt_sum += dt;
        acc.cell_updates += cells - s.rows_skipped * (uint64_t)nx;
        acc.rows_skipped += s.rows_skipped;
        if (s.bad) { aborted = 3; break; }
        dt_cfl = s.smax > 0.0 ? p->cfl / s.smax : p->dt_max_s;
        s_cur = s.smax;
        if (hits_sync) {
            sync_index += 1;
            if (s.sum_s > 0.0) acc.lts = std::fmax(acc.lts, s.smax * (double)cells / s.sum_s);
            double v = volume(ws, area);
            double expected = acc.v0 + acc.rain + acc.inflow - acc.outflow - acc.infil + acc.fix;
            double row[8] = {t, v, acc.rain, acc.inflow, acc.outflow, acc.infil, acc.fix, v - expected};
            rows.insert(rows.end(), row, row + 8);
            if (rp && rp->record)   // Tape::record_band (replay.rs:126-133)
                for (size_t q : rp->band_idx) {
                    rp->band.push_back((double)ws.h[q]);
                    rp->band.push_back((double)ws.qx[q]);
                    rp->band.push_back((double)ws.qy[q]);
                }
            if (rp && rp->replay) {   // Tape::band_diff (replay.rs:161-176) against the tolerances
                const size_t n = rp->band_idx.size();
                bool ok = (int64_t)sync_index <= rp->n_band_syncs;
                if (ok) {
                    const double* s0 = rp->band.data() + (size_t)(sync_index - 1) * 3 * n;
                    double dh = 0.0, dq = 0.0;
                    for (size_t k = 0; k < n; k++) {
                        const size_t q = rp->band_idx[k];
                        dh = std::fmax(dh, std::fabs((double)ws.h[q] - (double)(T)s0[3 * k]));
                        dq = std::fmax(std::fmax(dq, std::fabs((double)ws.qx[q] - (double)(T)s0[3 * k + 1])),
                                       std::fabs((double)ws.qy[q] - (double)(T)s0[3 * k + 2]));
                    }
                    ok = dh <= rp->tol_h && dq <= rp->tol_q;
                }
                if (!ok) {
                    aborted = 5;
                    abort_sync = (int)sync_index;
                    break;
                }
            }
            if (obs) {   // GuardObserver::on_sync (evaluator.rs:95-107)
                if (obs->require_no_offsite_worsening) {
                    bool have = false; double wv = 0; int wc = 0;
                    for (int q = 0; q < obs->n_guard; q++) {
                        int kk = obs->guard_pos[q];
                        double w = ws.mon.max[kk] - obs->baseline_max[kk] - obs->guard_tol;
                        if (w > 0.0 && (!have || w > wv)) { have = true; wv = w; wc = (int)ws.mon.cells[kk]; }
                    }
                    if (have) { aborted = 1; abort_cell = wc; abort_value = wv; abort_sync = (int)sync_index; break; }
                }
                if (obs->has_incumbent) {
                    double lb = objective_j(obs, ws.mon.max.data());
                    if (lb >= obs->incumbent) { aborted = 2; abort_value = lb; abort_sync = (int)sync_index; break; }
                }
            }
            if (locked) take_snap();
        }
    }
    double gross = acc.v0 + acc.rain + acc.inflow;
    double last_res = rows.empty() ? 0.0 : rows[rows.size() - 1];
    sum->t_end = t;
    sum->steps = (int64_t)step;
    sum->sync_count = (int32_t)sync_index;
    sum->aborted = aborted;
    sum->abort_cell = abort_cell;
    sum->abort_sync = abort_sync;
    sum->abort_value = abort_value;
    sum->dt_fallback_t = dt_fallback_t;
    sum->v0 = acc.v0;
    sum->rel_residual = std::fabs(last_res) / std::fmax(gross, 1e-300);
    sum->abs_residual = std::fabs(last_res);
    sum->gross = gross;
    sum->dt_min = acc.dt_min == DBL_MAX ? 0.0 : acc.dt_min;
    sum->dt_max = acc.dt_max;
    sum->dt_mean = step > 0 ? acc.dt_sum / (double)step : 0.0;
    sum->cell_updates = (int64_t)acc.cell_updates;
    sum->rows_skipped = (int64_t)acc.rows_skipped;
    sum->lts = acc.lts;
    sum->n_rows = (int32_t)(rows.size() / 8);
    if (rows_out) *rows_out = rows;
    return 0;
}

// sub_terrain (replay.rs:181-241): the ROI grid with every face of the candidate's full
// prepared terrain on and around it, wall boundaries (the tape supplies the ghosts).
template <class T> struct SubTerrain {
    TerrainBase base;
    Prepared<T> pt;
};
template <class T> void sub_terrain(const TerrainBase& b, const Prepared<T>& full, const Roi& r, SubTerrain<T>& out) {
    TerrainBase& s = out.base;
    s.nx = r.w; s.ny = r.h; s.level = b.level; s.crest = b.crest;
    s.dx = b.dx; s.dy = b.dy; s.g = b.g; s.fine = b.fine;
    s.z.assign((size_t)r.w * r.h, 0.0);
    s.wall.assign((size_t)r.w * r.h, 0);
    s.nval.assign((size_t)r.w * r.h, 0.0);
    s.wet_count = 0;
    for (int rr = 0; rr < r.h; rr++)
        for (int c = 0; c < r.w; c++) {
            const size_t k = (size_t)(rr + r.r0) * b.nx + (size_t)(c + r.c0), q = (size_t)rr * r.w + c;
            s.z[q] = full.zlev[k];
            s.wall[q] = b.wall[k];
            s.nval[q] = b.nval[k];
            s.wet_count += b.wall[k] ? 0 : 1;
        }
    s.north.assign(r.w, Ghost()); s.south.assign(r.w, Ghost());
    s.west.assign(r.h, Ghost()); s.east.assign(r.h, Ghost());
    Prepared<T>& q = out.pt;
    q.b = &s;
    const size_t len = (size_t)(r.h + 2) * s.P();
    q.fx0.assign(len, T(0)); q.fx1.assign(len, T(0)); q.fy0.assign(len, T(0)); q.fy1.assign(len, T(0));
    q.wet.assign(len, T(0)); q.gn2.assign(len, T(0));
    for (int js = 0; js < r.h + 2; js++)
        for (int is = 0; is < r.w + 2; is++) {
            const size_t ks = s.at(is, js), kf = b.at(is + r.c0, js + r.r0);
            q.fx0[ks] = full.fx0[kf]; q.fx1[ks] = full.fx1[kf];
            q.fy0[ks] = full.fy0[kf]; q.fy1[ks] = full.fy1[kf];
            q.wet[ks] = full.wet[kf]; q.gn2[ks] = full.gn2[kf];
        }
    q.zlev = s.z;
}

template <class T>
void copy_outputs(const TerrainBase& b, Workspace<T>& ws, const std::vector<double>& rows, int n_mon, double* ledger_rows,
                  int max_rows, double* mon_max, double* mon_tmax, double* out_h, double* out_qx, double* out_qy) {
    if (ledger_rows) {
        int n = std::min((int)(rows.size() / 8), max_rows);
        std::memcpy(ledger_rows, rows.data(), sizeof(double) * 8 * (size_t)n);
    }
    if (mon_max) for (int q = 0; q < n_mon; q++) mon_max[q] = ws.mon.max[q];
    if (mon_tmax) for (int q = 0; q < n_mon; q++) mon_tmax[q] = ws.mon.tmax[q];
    for (int r = 0; r < b.ny; r++)
        for (int c = 0; c < b.nx; c++) {
            size_t k = (size_t)r * b.nx + c, q = b.at(c + 1, r + 1);
            if (out_h) out_h[k] = (double)ws.h[q];
            if (out_qx) out_qx[k] = (double)ws.qx[q];
            if (out_qy) out_qy[k] = (double)ws.qy[q];
        }
}

// A baseline run recording the ROI tape (problem.rs:306-319, record_ring / record_band).
template <class T>
int record_typed(const TerrainBase& b, const orc_params* p, const Roi& roi, const uint32_t* mon_cells, int n_mon,
                 orc_summary* sum, std::vector<double>& ring, std::vector<double>& band) {
    Prepared<T> pt = prepare<T>(b, nullptr);
    Workspace<T> ws;
    ws.init(b.nx, b.ny, p->strip_rows > 0 ? p->strip_rows : b.ny);
    ws.mon.init(mon_cells, n_mon, b.nx, b.ny);
    ReplayCtx rp;
    rp.roi = roi;
    rp.record = true;
    for (auto& ij : roi_ring(roi)) rp.ring_idx.push_back(b.at(ij.first + roi.c0, ij.second + roi.r0));
    for (auto& ij : roi_band(roi)) rp.band_idx.push_back(b.at(ij.first + roi.c0, ij.second + roi.r0));
    std::vector<double> rows;
    int rc = run_impl(pt, p, ws, nullptr, nullptr, sum, &rows, nullptr, &rp);
    ring.swap(rp.ring);
    band.swap(rp.band);
    return rc;
}

// A candidate replayed on the ROI (evaluator.rs:427-440): cold start, the tape's ring as
// ghosts, the band checked at every sync point.
template <class T>
int replay_typed(const TerrainBase& b, const orc_edit* edit, const orc_params* p, const Roi& roi, const double* ring,
                 int64_t n_steps, const double* band, int64_t n_syncs, double tol_h, double tol_q, const uint32_t* mon_sub,
                 int n_mon, const orc_observer* obs, orc_summary* sum, double* ledger_rows, int max_rows, double* mon_max,
                 double* mon_tmax, double* out_h, double* out_qx, double* out_qy) {
    Prepared<T> full = prepare<T>(b, edit);
    SubTerrain<T> st;
    sub_terrain(b, full, roi, st);
    const TerrainBase& sb = st.base;
    Workspace<T> ws;
    ws.init(sb.nx, sb.ny, p->strip_rows > 0 ? p->strip_rows : sb.ny);
    ws.mon.init(mon_sub, n_mon, sb.nx, sb.ny);
    ReplayCtx rp;
    rp.roi = roi;
    rp.replay = true;
    for (auto& ij : roi_ring(roi)) rp.ring_idx.push_back(sb.at(ij.first, ij.second));
    for (auto& ij : roi_band(roi)) rp.band_idx.push_back(sb.at(ij.first, ij.second));
    rp.ring.assign(ring, ring + (size_t)n_steps * 3 * rp.ring_idx.size());
    rp.band.assign(band, band + (size_t)n_syncs * 3 * rp.band_idx.size());
    rp.n_ring_steps = n_steps;
    rp.n_band_syncs = n_syncs;
    rp.tol_h = tol_h;
    rp.tol_q = tol_q;
    std::vector<double> rows;
    int rc = run_impl(st.pt, p, ws, obs, nullptr, sum, &rows, nullptr, &rp);
    copy_outputs(sb, ws, rows, n_mon, ledger_rows, max_rows, mon_max, mon_tmax, out_h, out_qx, out_qy);
    return rc;
}

template <class T>
int run_typed(const TerrainBase& b, const orc_edit* edit, const orc_params* p, const uint32_t* mon_cells, int n_mon,
              const orc_observer* obs, const orc_start* st, orc_summary* sum, double* ledger_rows, int max_rows,
              double* mon_max, double* mon_tmax, double* out_h, double* out_qx, double* out_qy, double* out_dts,
              int64_t max_dts) {
    Prepared<T> pt = prepare<T>(b, edit);
    Workspace<T> ws;
    ws.init(b.nx, b.ny, p->strip_rows > 0 ? p->strip_rows : b.ny);
    ws.mon.init(mon_cells, n_mon, b.nx, b.ny);
    std::vector<double> rows, dts;
    int rc = run_impl(pt, p, ws, obs, st, sum, &rows, out_dts ? &dts : nullptr);
    if (ledger_rows) {
        int n = std::min((int)(rows.size() / 8), max_rows);
        std::memcpy(ledger_rows, rows.data(), sizeof(double) * 8 * (size_t)n);
    }
    if (mon_max) for (int q = 0; q < n_mon; q++) mon_max[q] = ws.mon.max[q];
    if (mon_tmax) for (int q = 0; q < n_mon; q++) mon_tmax[q] = ws.mon.tmax[q];
    for (int r = 0; r < b.ny; r++)
        for (int c = 0; c < b.nx; c++) {
            size_t k = (size_t)r * b.nx + c, q = b.at(c + 1, r + 1);
            if (out_h) out_h[k] = (double)ws.h[q];
            if (out_qx) out_qx[k] = (double)ws.qx[q];
            if (out_qy) out_qy[k] = (double)ws.qy[q];
        }
    if (out_dts) {
        int64_t n = std::min((int64_t)dts.size(), max_dts);
        std::memcpy(out_dts, dts.data(), sizeof(double) * (size_t)n);
        sum->n_dts = (int32_t)n;
    }
    return rc;
}

// ---------------------------------------------------------------- design.rs:95-142
struct Prim { int kind; double cx, cy, length, width, angle_deg, height; };

void prim_segment(const Prim& p, double& x0, double& y0, double& x1, double& y1) {  // design.rs:97-105
    if (p.kind == 3) { x0 = x1 = p.cx; y0 = y1 = p.cy; return; }
    double a = p.angle_deg * (M_PI / 180.0);
    double s = std::sin(a), c = std::cos(a);
    double hl = 0.5 * p.length;
    x0 = p.cx - c * hl; y0 = p.cy - s * hl; x1 = p.cx + c * hl; y1 = p.cy + s * hl;
}

double prim_profile(const Prim& p, double x, double y) {  // design.rs:108-134
    double x0, y0, x1, y1;
    prim_segment(p, x0, y0, x1, y1);
    double vx = x1 - x0, vy = y1 - y0;
    double l2 = vx * vx + vy * vy;
    double t = 0.0;
    if (l2 > 0.0) { t = ((x - x0) * vx + (y - y0) * vy) / l2; t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t); }
    double px = x0 + t * vx - x, py = y0 + t * vy - y;
    double d = std::sqrt(px * px + py * py);
    double s = d / (0.5 * p.width);
    if (s >= 1.0) return 0.0;
    if (p.kind == 2) {
        if (s <= 0.5) return 1.0;
        double u = (s - 0.5) / 0.5;
        double w = 1.0 - u * u;
        return w * w;
    }
    double w = 1.0 - s * s;
    return w * w;
}

Prim prim_of(const double* v) { return Prim{(int)v[0], v[1], v[2], v[3], v[4], v[5], v[6]}; }

}  // namespace

// ================================================================ C API
struct orc_terrain { TerrainBase* b; };

extern "C" {

orc_terrain* orc_terrain_build(int32_t nx, int32_t ny, double origin_x, double origin_y, double dx, double dy,
                               const double* dem, const uint8_t* walls, const double* manning, int32_t level,
                               const orc_segment* segs, int32_t nsegs) {
    auto fine = fine_new(nx, ny, origin_x, origin_y, dx, dy, dem, walls, manning);
    auto* t = new orc_terrain();
    t->b = base_build(fine, level, segs, nsegs);
    return t;
}

void orc_terrain_free(orc_terrain* t) {
    if (!t) return;
    delete t->b;
    delete t;
}

void orc_terrain_info(const orc_terrain* t, int32_t* info, double* dinfo) {
    const TerrainBase& b = *t->b;
    info[0] = b.nx; info[1] = b.ny; info[2] = b.crest ? 1 : 0; info[3] = b.wet_count;
    dinfo[0] = b.dx; dinfo[1] = b.dy; dinfo[2] = b.fine->z_ref; dinfo[3] = b.fine->z_wall; dinfo[4] = b.g;
}

void orc_terrain_faces(const orc_terrain* t, const orc_edit* edit, int32_t precision, double* fx0, double* fx1,
                       double* fy0, double* fy1, double* wet, double* gn2, double* z_level) {
    const TerrainBase& b = *t->b;
    auto out = [&](auto& pt) {
        size_t len = pt.fx0.size();
        for (size_t i = 0; i < len; i++) {
            if (fx0) fx0[i] = (double)pt.fx0[i];
            if (fx1) fx1[i] = (double)pt.fx1[i];
            if (fy0) fy0[i] = (double)pt.fy0[i];
            if (fy1) fy1[i] = (double)pt.fy1[i];
            if (wet) wet[i] = (double)pt.wet[i];
            if (gn2) gn2[i] = (double)pt.gn2[i];
        }
        if (z_level) std::memcpy(z_level, pt.zlev.data(), sizeof(double) * pt.zlev.size());
    };
    if (precision == 64) { auto pt = prepare<double>(b, edit); out(pt); }
    else { auto pt = prepare<float>(b, edit); out(pt); }
}

void orc_terrain_ghosts(const orc_terrain* t, int32_t* north, int32_t* south, int32_t* west, int32_t* east) {
    const TerrainBase& b = *t->b;
    for (int i = 0; i < b.nx; i++) { north[i] = b.north[i].kind; south[i] = b.south[i].kind; }
    for (int j = 0; j < b.ny; j++) { west[j] = b.west[j].kind; east[j] = b.east[j].kind; }
}

int32_t orc_run(const orc_terrain* base, const orc_edit* edit, const orc_params* p, const uint32_t* mon_cells,
                int32_t n_mon, const orc_observer* obs, const orc_start* start, orc_summary* sum, double* ledger_rows,
                int32_t max_rows, double* mon_max, double* mon_tmax, double* out_h, double* out_qx, double* out_qy,
                double* out_dts, int64_t max_dts) {
    std::memset(sum, 0, sizeof(*sum));
    for (int q = 1; q < n_mon; q++) if (mon_cells[q] <= mon_cells[q - 1]) return -2;  // MonitorSet is sorted, unique
    for (int q = 0; q < n_mon; q++) if (mon_cells[q] >= (uint32_t)(base->b->nx * base->b->ny)) return -3;
    if (p->precision == 64)
        return run_typed<double>(*base->b, edit, p, mon_cells, n_mon, obs, start, sum, ledger_rows, max_rows, mon_max,
                                 mon_tmax, out_h, out_qx, out_qy, out_dts, max_dts);
    return run_typed<float>(*base->b, edit, p, mon_cells, n_mon, obs, start, sum, ledger_rows, max_rows, mon_max,
                            mon_tmax, out_h, out_qx, out_qy, out_dts, max_dts);
}

int32_t orc_record_tape(const orc_terrain* base, const orc_params* p, const int32_t* roi4, const uint32_t* mon_cells,
                        int32_t n_mon, orc_summary* sum, double* ring, int64_t ring_cap, double* band, int64_t band_cap,
                        int64_t* n_ring, int64_t* n_band) {
    const Roi roi{roi4[0], roi4[1], roi4[2], roi4[3]};
    std::vector<double> r, bd;
    std::memset(sum, 0, sizeof(*sum));
    int rc = p->precision == 64 ? record_typed<double>(*base->b, p, roi, mon_cells, n_mon, sum, r, bd)
                                : record_typed<float>(*base->b, p, roi, mon_cells, n_mon, sum, r, bd);
    *n_ring = (int64_t)r.size();
    *n_band = (int64_t)bd.size();
    if (ring) std::memcpy(ring, r.data(), sizeof(double) * (size_t)std::min<int64_t>(ring_cap, *n_ring));
    if (band) std::memcpy(band, bd.data(), sizeof(double) * (size_t)std::min<int64_t>(band_cap, *n_band));
    return rc;
}

int32_t orc_run_replay(const orc_terrain* base, const orc_edit* edit, const orc_params* p, const int32_t* roi4,
                       const double* ring, int64_t n_steps, const double* band, int64_t n_syncs, double tol_h, double tol_q,
                       const uint32_t* mon_sub, int32_t n_mon, const orc_observer* obs, orc_summary* sum,
                       double* ledger_rows, int32_t max_rows, double* mon_max, double* mon_tmax, double* out_h,
                       double* out_qx, double* out_qy) {
    const Roi roi{roi4[0], roi4[1], roi4[2], roi4[3]};
    std::memset(sum, 0, sizeof(*sum));
    if (p->precision == 64)
        return replay_typed<double>(*base->b, edit, p, roi, ring, n_steps, band, n_syncs, tol_h, tol_q, mon_sub, n_mon,
                                    obs, sum, ledger_rows, max_rows, mon_max, mon_tmax, out_h, out_qx, out_qy);
    return replay_typed<float>(*base->b, edit, p, roi, ring, n_steps, band, n_syncs, tol_h, tol_q, mon_sub, n_mon, obs,
                               sum, ledger_rows, max_rows, mon_max, mon_tmax, out_h, out_qx, out_qy);
}

double orc_run_candidates(const orc_terrain* base, const orc_edit* edits, int32_t n_cand, const orc_params* p,
                          const uint32_t* mon_cells, int32_t n_mon, int32_t threads, int64_t* steps_out,
                          int64_t* cell_updates_out) {
    const TerrainBase& b = *base->b;
    // Preparation (terrain overlays) is outside the timed region, like the evaluator's
    // prepare step; the timed region is the candidate-parallel simulation only.
    std::vector<Prepared<float>> pts;
    pts.reserve(n_cand);
    for (int c = 0; c < n_cand; c++) pts.push_back(prepare<float>(b, edits ? &edits[c] : nullptr));
    std::atomic<int> next{0};
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        Workspace<float> ws;
        ws.init(b.nx, b.ny, b.ny);
        for (;;) {
            int c = next.fetch_add(1);
            if (c >= n_cand) break;
            ws.mon.init(mon_cells, n_mon, b.nx, b.ny);
            orc_summary s;
            std::memset(&s, 0, sizeof(s));
            orc_params q = *p;
            if (q.kernel != 2) q.kernel = 1;   // the fused HLL kernel, or the local-inertial one
            q.precision = 32;
            run_impl(pts[c], &q, ws, nullptr, nullptr, &s, nullptr, nullptr);
            steps_out[c] = s.steps;
            cell_updates_out[c] = s.cell_updates;
        }
    };
    std::vector<std::thread> th;
    int nt = std::max(1, threads);
    for (int i = 0; i < nt; i++) th.emplace_back(worker);
    for (auto& x : th) x.join();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int64_t orc_materialize(const double* prims, int32_t n_prims, double origin_x, double origin_y, double dx, double dy,
                        int32_t nx, int32_t ny, const uint8_t* editable, double max_abs_dz, int32_t* bbox, double* dz,
                        int64_t dz_cap, double* cut_fill) {  // design.rs:318-385
    std::vector<Prim> ps;
    for (int i = 0; i < n_prims; i++) ps.push_back(prim_of(prims + 7 * i));
    size_t c0 = SIZE_MAX, r0 = SIZE_MAX, c1 = 0, r1 = 0;
    auto to_grid = [&](double x, double y, double& fc, double& fr) { fc = (x - origin_x) / dx; fr = (origin_y - y) / dy; };
    for (const Prim& p : ps) {
        if (p.height == 0.0) continue;
        double x0, y0, x1, y1;
        prim_segment(p, x0, y0, x1, y1);
        double r = 0.5 * p.width;
        double bx0 = std::fmin(x0, x1) - r, by0 = std::fmin(y0, y1) - r, bx1 = std::fmax(x0, x1) + r, by1 = std::fmax(y0, y1) + r;
        double fc0, fr1, fc1, fr0;
        to_grid(bx0, by0, fc0, fr1);
        to_grid(bx1, by1, fc1, fr0);
        size_t a = std::min((size_t)std::fmax(std::floor(fc0), 0.0), (size_t)nx);
        size_t bb = std::min((size_t)std::fmax(std::floor(fr0), 0.0), (size_t)ny);
        size_t c = std::min((size_t)std::fmax(std::ceil(fc1), 0.0), (size_t)nx);
        size_t d = std::min((size_t)std::fmax(std::ceil(fr1), 0.0), (size_t)ny);
        if (a < c && bb < d) { c0 = std::min(c0, a); r0 = std::min(r0, bb); c1 = std::max(c1, c); r1 = std::max(r1, d); }
    }
    cut_fill[0] = 0.0;
    cut_fill[1] = 0.0;
    if (c0 >= c1 || r0 >= r1) { bbox[0] = bbox[1] = bbox[2] = bbox[3] = 0; return 0; }
    size_t w = c1 - c0, h = r1 - r0;
    if ((int64_t)(w * h) > dz_cap) return -1;
    const int S = 4;
    const double inv = 1.0 / (double)(S * S);
    for (size_t r = 0; r < h; r++)
        for (size_t c = 0; c < w; c++) {
            size_t col = c0 + c, row = r0 + r;
            dz[r * w + c] = 0.0;
            if (editable && !editable[row * nx + col]) continue;
            double acc = 0.0;
            for (const Prim& p : ps) {
                if (p.height == 0.0) continue;
                double s = 0.0;
                for (int sy = 0; sy < S; sy++)
                    for (int sx = 0; sx < S; sx++) {
                        double x = origin_x + ((double)col + ((double)sx + 0.5) / (double)S) * dx;
                        double y = origin_y - ((double)row + ((double)sy + 0.5) / (double)S) * dy;
                        s += prim_profile(p, x, y);
                    }
                acc += p.height * s * inv;
            }
            dz[r * w + c] = acc < -max_abs_dz ? -max_abs_dz : (acc > max_abs_dz ? max_abs_dz : acc);
        }
    double a = dx * dy;
    for (size_t i = 0; i < w * h; i++) {
        if (dz[i] > 0.0) cut_fill[1] += dz[i] * a;
        else cut_fill[0] -= dz[i] * a;
    }
    bbox[0] = (int32_t)c0; bbox[1] = (int32_t)r0; bbox[2] = (int32_t)w; bbox[3] = (int32_t)h;
    return (int64_t)(w * h);
}

double orc_profile(const double* prim, double x, double y) { return prim_profile(prim_of(prim), x, y); }

double orc_hyeto_rate(const double* ts, const double* rs, int32_t n, double t) {
    Hyeto h; h.t.assign(ts, ts + n); h.r.assign(rs, rs + n);
    return h.rate(t);
}
double orc_hyeto_integral(const double* ts, const double* rs, int32_t n, double a, double b) {
    Hyeto h; h.t.assign(ts, ts + n); h.r.assign(rs, rs + n);
    return h.integral(a, b);
}

float orc_rcbrt_f32(float x) { return rcbrt(x); }
double orc_rcbrt_f64(double x) { return rcbrt(x); }
float orc_fixedsum_f32(const float* v, int64_t n, int32_t mode) {
    FixedSum<float> s;
    if (mode == 1) s.add_slice(v, (size_t)n);
    else for (int64_t i = 0; i < n; i++) s.add((size_t)i, v[i]);
    return s.total();
}
float orc_max_slice_f32(const float* v, int64_t n) { return max_slice(v, (size_t)n); }
float orc_rmax_f32(float a, float b) { return rmax(a, b); }
double orc_objective_j(const orc_observer* o, const double* monitor_max) { return objective_j(o, monitor_max); }

}  // extern "C"
