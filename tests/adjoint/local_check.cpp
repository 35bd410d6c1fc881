// Host check of the hand-written adjoints in itr_adjoint.h: for each block and a set of
// operating points covering its branches, <adj(b), d> must equal the central difference
// of <b, f(x + eps d)> in double precision.  Prints one line per case and exits non-zero
// on a mismatch.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "itr_adjoint.h"

using namespace itr::adj;
using T = double;

static int failures = 0;
static std::mt19937_64 rng(12345);
static double urand(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static Consts<T> consts() {
    Consts<T> k;
    k.g = 9.80665; k.g_half = 0.5 * 9.80665; k.heps2 = 1e-10; k.h_dry = 1e-6; k.q_min = 1e-30;
    k.dt = 0.2; k.dtdx = 0.2 / 5.0; k.dtdy = 0.2 / 5.0; k.rain = 2e-6; k.cap = 3e-7;
    return k;
}

static void check(const char* name, double ad, double fd) {
    const double scale = std::fmax(1e-12, std::fmax(std::fabs(ad), std::fabs(fd)));
    const double rel = std::fabs(ad - fd) / scale;
    const bool ok = rel < 2e-6;
    std::printf("%-40s adj % .10e  fd % .10e  rel %.2e %s\n", name, ad, fd, rel, ok ? "ok" : "MISMATCH");
    if (!ok) failures++;
}

// f: vector -> vector; adj: (x, b) -> xbar
// `fixed`: inputs held constant (a dry side stays exactly dry: its depth and celerity
// sit on a kink, where only one-sided derivatives exist).
template <class Fwd, class Adj>
static void dir_check(const char* name, std::vector<double> x, int nout, Fwd f, Adj a, double eps = 1e-7,
                      std::vector<int> fixed = {}) {
    const int n = (int)x.size();
    for (int trial = 0; trial < 3; trial++) {
        std::vector<double> d(n), b(nout);
        for (auto& v : d) v = urand(-1, 1);
        for (int i : fixed) d[i] = 0.0;
        for (auto& v : b) v = urand(-1, 1);
        std::vector<double> xp(x), xm(x);
        for (int i = 0; i < n; i++) { xp[i] += eps * d[i]; xm[i] -= eps * d[i]; }
        std::vector<double> yp = f(xp), ym = f(xm);
        double fd = 0;
        for (int j = 0; j < nout; j++) fd += b[j] * (yp[j] - ym[j]) / (2 * eps);
        std::vector<double> xb = a(x, b);
        double ad = 0;
        for (int i = 0; i < n; i++) ad += xb[i] * d[i];
        char nm[96];
        std::snprintf(nm, sizeof nm, "%s #%d", name, trial);
        check(nm, ad, fd);
    }
}

int main() {
    const Consts<T> k = consts();
    // cell velocity: wet (h^2 > heps2) and thin (h^2 < heps2)
    for (double h : {0.3, 4e-6}) {
        dir_check(h > 1e-3 ? "vel wet" : "vel thin", {h, 0.2, -0.05}, 3,
                  [&](const std::vector<double>& x) { V<T> v = vel(x[0], x[1], x[2], k); return std::vector<double>{v.u, v.v, v.c}; },
                  [&](const std::vector<double>& x, const std::vector<double>& b) {
                      S3<T> o = vel_adj(x[0], x[1], x[2], k, V<T>{b[0], b[1], b[2]});
                      return std::vector<double>{o.h, o.qx, o.qy};
                  }, h > 1e-3 ? 1e-7 : 1e-12);
    }
    // HLL regimes: (hl, ul, tl, cl, hr, ur, tr, cr)
    struct P { const char* n; std::vector<double> x; std::vector<int> fixed; };
    auto c = [&](double h) { return std::sqrt(k.g * h); };
    std::vector<P> hp = {
        {"hll subcritical star", {0.5, 0.3, 0.1, c(0.5), 0.3, -0.2, 0.05, c(0.3)}},
        {"hll star, tangential right", {0.2, -0.4, 0.1, c(0.2), 0.4, -0.6, 0.3, c(0.4)}},
        {"hll supercritical use_l", {0.2, 3.0, 0.1, c(0.2), 0.15, 2.5, 0.0, c(0.15)}},
        {"hll supercritical use_r", {0.2, -3.0, 0.1, c(0.2), 0.15, -2.8, 0.2, c(0.15)}},
        {"hll dry left", {0.0, 0.0, 0.0, 0.0, 0.3, -0.1, 0.05, c(0.3)}, {0, 3}},
        {"hll dry right", {0.4, 0.2, 0.1, c(0.4), 0.0, 0.0, 0.0, 0.0}, {4, 7}},
    };
    for (auto& p : hp)
        dir_check(p.n, p.x, 4,
                  [&](const std::vector<double>& x) {
                      F<T> f = hll(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], k);
                      return std::vector<double>{f.m, f.nl, f.nr, f.t};
                  },
                  [&](const std::vector<double>& x, const std::vector<double>& b) {
                      HllBar<T> o = hll_adj(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], k, F<T>{b[0], b[1], b[2], b[3]});
                      return std::vector<double>{o.hl, o.ul, o.tl, o.cl, o.hr, o.ur, o.tr, o.cr};
                  }, 1e-8, p.fixed);
    // faces: (hl, qxl, qyl, hr, qxr, qyr, dz) through cell velocities, x and y normals, +/- dz
    for (int nx = 0; nx < 2; nx++)
        for (double dz : {0.12, -0.08, 0.7}) {
            char nm[64];
            std::snprintf(nm, sizeof nm, "face %s dz=%+.2f", nx ? "x" : "y", dz);
            auto fwd = [&, nx](const std::vector<double>& x) {
                V<T> vl = vel(x[0], x[1], x[2], k), vr = vel(x[3], x[4], x[5], k);
                F<T> f = nx ? face<true>(x[0], vl, x[3], vr, x[6], k) : face<false>(x[0], vl, x[3], vr, x[6], k);
                return std::vector<double>{f.m, f.nl, f.nr, f.t};
            };
            auto adj = [&, nx](const std::vector<double>& x, const std::vector<double>& b) {
                V<T> vl = vel(x[0], x[1], x[2], k), vr = vel(x[3], x[4], x[5], k);
                F<T> fb{b[0], b[1], b[2], b[3]};
                FaceBar<T> o = nx ? face_adj<true>(x[0], vl, x[3], vr, x[6], k, fb) : face_adj<false>(x[0], vl, x[3], vr, x[6], k, fb);
                S3<T> l = vel_adj(x[0], x[1], x[2], k, V<T>{o.ul, o.vl, o.cl});
                S3<T> r = vel_adj(x[3], x[4], x[5], k, V<T>{o.ur, o.vr, o.cr});
                return std::vector<double>{l.h + o.hl, l.qx, l.qy, r.h + o.hr, r.qx, r.qy, o.dz};
            };
            dir_check(nm, {0.45, 0.08, -0.03, 0.3, 0.05, 0.02, dz}, 4, fwd, adj, 1e-8);
        }
    // crest faces: (hl, qxl, qyl, hr, qxr, qyr, d_l, d_r), x and y normals; a rise on one
    // side (the other at the crest), on both, and a left side drained below the crest
    struct CP { const char* n; double dl, dr, hl; };
    std::vector<CP> cp = {{"left rise", 0.12, 0.0, 0.45}, {"right rise", 0.0, 0.09, 0.45},
                          {"both rises", 0.05, 0.2, 0.45}, {"left below crest", 0.6, 0.1, 0.45}};
    for (int nx = 0; nx < 2; nx++)
        for (auto& q : cp) {
            char nm[64];
            std::snprintf(nm, sizeof nm, "crest face %s %s", nx ? "x" : "y", q.n);
            auto fwd = [&, nx](const std::vector<double>& x) {
                V<T> vl = vel(x[0], x[1], x[2], k), vr = vel(x[3], x[4], x[5], k);
                F<T> f = nx ? face_cr<true>(x[0], vl, x[3], vr, x[6], x[7], k) : face_cr<false>(x[0], vl, x[3], vr, x[6], x[7], k);
                return std::vector<double>{f.m, f.nl, f.nr, f.t};
            };
            auto adj = [&, nx](const std::vector<double>& x, const std::vector<double>& b) {
                V<T> vl = vel(x[0], x[1], x[2], k), vr = vel(x[3], x[4], x[5], k);
                F<T> fb{b[0], b[1], b[2], b[3]};
                FaceBar<T> o = nx ? face_cr_adj<true>(x[0], vl, x[3], vr, x[6], x[7], k, fb)
                                  : face_cr_adj<false>(x[0], vl, x[3], vr, x[6], x[7], k, fb);
                S3<T> l = vel_adj(x[0], x[1], x[2], k, V<T>{o.ul, o.vl, o.cl});
                S3<T> r = vel_adj(x[3], x[4], x[5], k, V<T>{o.ur, o.vr, o.cr});
                return std::vector<double>{l.h + o.hl, l.qx, l.qy, r.h + o.hr, r.qx, r.qy, o.dz, o.dz2};
            };
            // a zero rise sits on the select's switch: held fixed (one-sided there)
            std::vector<int> fixed;
            if (q.dl == 0.0) fixed.push_back(6);
            if (q.dr == 0.0) fixed.push_back(7);
            dir_check(nm, {q.hl, 0.08, -0.03, 0.3, 0.05, 0.02, q.dl, q.dr}, 4, fwd, adj, 1e-8, fixed);
        }
    // sources: (h1, qx1, qy1) with wet cells, capacity-limited and not, q flushing
    struct SP { const char* n; std::vector<double> x; double rain, cap; };
    std::vector<SP> sp = {{"sources wet capped", {0.3, 0.05, -0.02}, 2e-6, 3e-7},
                          {"sources wet, shallow", {0.004, 0.001, 0.0005}, 0.0, 1e-7},
                          {"sources deep fast", {1.2, 0.9, 0.4}, 1e-5, 3e-7}};
    for (auto& p : sp) {
        Consts<T> kk = k;
        kk.rain = p.rain;
        kk.cap = p.cap;
        dir_check(p.n, p.x, 3,
                  [&](const std::vector<double>& x) { S3<T> s = sources(x[0], x[1], x[2], 1.0, 0.012, kk); return std::vector<double>{s.h, s.qx, s.qy}; },
                  [&](const std::vector<double>& x, const std::vector<double>& b) {
                      S3<T> o = sources_adj(x[0], x[1], x[2], 1.0, 0.012, kk, S3<T>{b[0], b[1], b[2]});
                      return std::vector<double>{o.h, o.qx, o.qy};
                  });
    }
    std::printf("%s: %d mismatches\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}
