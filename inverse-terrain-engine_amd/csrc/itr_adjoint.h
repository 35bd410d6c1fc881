// itr_adjoint.h -- hand-written reverse-mode derivatives of the shallow-water step's
// building blocks (crates/itr-hydro/src/numerics.rs:36-195), for the discrete adjoint.
//
// Each *_adj function takes the same inputs as the forward function plus the adjoint
// ("bar") of the forward outputs, re-evaluates the forward branch decisions, and returns
// the adjoint of the inputs.  Derivatives follow the branch the forward takes: a select
// contributes only its chosen operand; max(x, 0) / min(a, b) switch with the forward
// comparison.  Where the exact derivative is unbounded (sqrt at 0) the adjoint uses 0:
// a dry cell's celerity carries no first-order sensitivity.
//
// Templated on the arithmetic type: float on the device, double in the host checks that
// compare every block with central differences.
#pragma once

#include <hip/hip_runtime.h>

namespace itr {
namespace adj {

template <class T> struct Consts {
    T g, g_half, heps2, h_dry, q_min, dt, dtdx, dtdy, rain, cap;
};
template <class T> struct V { T u, v, c; };
template <class T> struct F { T m, nl, nr, t; };
template <class T> struct S3 { T h, qx, qy; };
template <class T> struct HllBar { T hl, ul, tl, cl, hr, ur, tr, cr; };
// dz: the adjoint of the signed jump, or of the left rise d_l on a crest face; dz2: of
// the right rise d_r (crest faces only).
template <class T> struct FaceBar { T hl, ul, vl, cl, hr, ur, vr, cr, dz, dz2; };

template <class T> __host__ __device__ inline T smx(T a, T b) { return a > b ? a : b; }   // real.rs:44-46
template <class T> __host__ __device__ inline T smn(T a, T b) { return a < b ? a : b; }   // real.rs:48-50
template <class T> __host__ __device__ inline T sab(T x) { return x < T(0) ? -x : x; }    // real.rs:39-41
template <class T> __host__ __device__ inline T rcp(T x) { return T(1) / x; }
template <class T> __host__ __device__ inline T rt(T x) { return sqrt(x); }
// x^(-1/3) to working precision; the derivative blocks use its exact derivative -r/(3x).
template <class T> __host__ __device__ inline T icbrt(T x) { return T(1) / cbrt(x); }
// On the device (f32) the reciprocal and square root are the hardware estimates
// (v_rcp_f32, v_sqrt_f32, about 1 ulp): a derivative needs no correct rounding (gradients
// are checked against finite differences to 1e-3), and the forward replay, which must
// equal the forward bitwise, does not use these blocks.  The host checks run in double.
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline float icbrt(float x) { return __builtin_amdgcn_rcpf(cbrtf(x)); }

// ------------------------------------------------------------ cell_vel (numerics.rs:36-43)
template <class T> __host__ __device__ inline V<T> vel(T h, T qx, T qy, const Consts<T>& k) {
    const T h2 = h * h;
    const T inv = rcp(h2 + smx(h2, k.heps2));
    const T t = (h + h) * inv;
    return V<T>{qx * t, qy * t, rt(k.g * h)};
}
template <class T>
__host__ __device__ inline S3<T> vel_adj(T h, T qx, T qy, const Consts<T>& k, const V<T>& b) {
    const T h2 = h * h;
    const bool big = h2 > k.heps2;
    const T D = h2 + (big ? h2 : k.heps2);
    const T inv = rcp(D);
    const T t = (h + h) * inv;
    const T dD = (h + h) + (big ? h + h : T(0));
    const T bt = b.u * qx + b.v * qy;
    S3<T> o;
    o.qx = b.u * t;
    o.qy = b.v * t;
    o.h = bt * (T(2) * inv - (h + h) * dD * inv * inv);
    const T c = rt(k.g * h);
    if (c > T(0)) o.h += b.c * k.g * rcp(T(2) * c);
    return o;
}

// ------------------------------------------------------------ hll (numerics.rs:58-93)
template <class T>
__host__ __device__ inline F<T> hll(T hl, T ul, T tl, T cl, T hr, T ur, T tr, T cr, const Consts<T>& k) {
    const bool dry_l = !(hl > T(0)), dry_r = !(hr > T(0));
    const T ustar = T(0.5) * (ul + ur) + cl - cr;
    const T cstar = T(0.5) * (cl + cr) + T(0.25) * (ul - ur);
    const T slw = smn(ul - cl, ustar - cstar);
    const T srw = smx(ur + cr, ustar + cstar);
    const T sl = dry_l ? ur - T(2) * cr : (dry_r ? ul - cl : slw);
    const T sr = dry_l ? ur + cr : (dry_r ? ul + T(2) * cl : srw);
    const T ml = hl * ul, mr = hr * ur;
    const T pl_ = k.g_half * hl * hl, pr_ = k.g_half * hr * hr;
    const T pl = ml * ul + pl_, pr = mr * ur + pr_;
    const bool both_dry = dry_l && dry_r, use_l = sl >= T(0), use_r = sr <= T(0);
    const T den = (both_dry || use_l || use_r) ? T(1) : sr - sl;
    const T inv = rcp(den);
    const T slsr = sl * sr;
    const T fm_star = (sr * ml - sl * mr + slsr * (hr - hl)) * inv;
    const T fn_star = (sr * pl - sl * pr + slsr * (mr - ml)) * inv;
    const T fm = both_dry ? T(0) : (use_l ? ml : (use_r ? mr : fm_star));
    const T fn = both_dry ? T(0) : (use_l ? pl : (use_r ? pr : fn_star));
    return F<T>{fm, fn - pl_, fn - pr_, fm * (fm > T(0) ? tl : tr)};
}

template <class T>
__host__ __device__ inline HllBar<T> hll_adj(T hl, T ul, T tl, T cl, T hr, T ur, T tr, T cr, const Consts<T>& k,
                                             const F<T>& b) {
    // forward values the derivative needs
    const bool dry_l = !(hl > T(0)), dry_r = !(hr > T(0));
    const T ustar = T(0.5) * (ul + ur) + cl - cr;
    const T cstar = T(0.5) * (cl + cr) + T(0.25) * (ul - ur);
    const bool slw_first = (ul - cl) < (ustar - cstar);
    const bool srw_first = (ur + cr) > (ustar + cstar);
    const T slw = slw_first ? ul - cl : ustar - cstar;
    const T srw = srw_first ? ur + cr : ustar + cstar;
    const T sl = dry_l ? ur - T(2) * cr : (dry_r ? ul - cl : slw);
    const T sr = dry_l ? ur + cr : (dry_r ? ul + T(2) * cl : srw);
    const T ml = hl * ul, mr = hr * ur;
    const T pl_ = k.g_half * hl * hl, pr_ = k.g_half * hr * hr;
    const T pl = ml * ul + pl_, pr = mr * ur + pr_;
    const bool both_dry = dry_l && dry_r, use_l = sl >= T(0), use_r = sr <= T(0);
    const bool star = !(both_dry || use_l || use_r);
    const T inv = star ? rcp(sr - sl) : T(1);
    const T slsr = sl * sr;
    const T A = sr * ml - sl * mr + slsr * (hr - hl);
    const T B = sr * pl - sl * pr + slsr * (mr - ml);
    const T fm = both_dry ? T(0) : (use_l ? ml : (use_r ? mr : A * inv));

    HllBar<T> o{};
    // outputs: m = fm, nl = fn - pl_, nr = fn - pr_, t = fm * (fm > 0 ? tl : tr)
    const T bfn = b.nl + b.nr;
    T bpl_ = -b.nl, bpr_ = -b.nr;
    T bfm = b.m + b.t * (fm > T(0) ? tl : tr);
    if (fm > T(0)) o.tl += b.t * fm;
    else o.tr += b.t * fm;
    T bml = 0, bmr = 0, bpl = 0, bpr = 0, bsl = 0, bsr = 0;
    if (!both_dry) {
        if (use_l) {
            bml += bfm;
            bpl += bfn;
        } else if (use_r) {
            bmr += bfm;
            bpr += bfn;
        } else {
            const T bA = bfm * inv, bB = bfn * inv;
            const T binv = bfm * A + bfn * B;
            bsr -= binv * inv * inv;
            bsl += binv * inv * inv;
            T bslsr = bA * (hr - hl) + bB * (mr - ml);
            bsr += bA * ml + bB * pl;
            bml += bA * sr - bB * slsr;
            bsl -= bA * mr + bB * pr;
            bmr += -bA * sl + bB * slsr;
            o.hr += bA * slsr;
            o.hl -= bA * slsr;
            bpl += bB * sr;
            bpr -= bB * sl;
            bsl += bslsr * sr;
            bsr += bslsr * sl;
        }
    }
    // pl = ml*ul + pl_, pr = mr*ur + pr_
    bml += bpl * ul;
    o.ul += bpl * ml;
    bpl_ += bpl;
    bmr += bpr * ur;
    o.ur += bpr * mr;
    bpr_ += bpr;
    o.hl += bpl_ * T(2) * k.g_half * hl;
    o.hr += bpr_ * T(2) * k.g_half * hr;
    o.hl += bml * ul;
    o.ul += bml * hl;
    o.hr += bmr * ur;
    o.ur += bmr * hr;
    // wave speeds
    if (dry_l) {
        o.ur += bsl + bsr;
        o.cr += -T(2) * bsl + bsr;
    } else if (dry_r) {
        o.ul += bsl + bsr;
        o.cl += -bsl + T(2) * bsr;
    } else {
        T bus = 0, bcs = 0;
        if (slw_first) {
            o.ul += bsl;
            o.cl -= bsl;
        } else {
            bus += bsl;
            bcs -= bsl;
        }
        if (srw_first) {
            o.ur += bsr;
            o.cr += bsr;
        } else {
            bus += bsr;
            bcs += bsr;
        }
        o.ul += T(0.5) * bus + T(0.25) * bcs;
        o.ur += T(0.5) * bus - T(0.25) * bcs;
        o.cl += bus + T(0.5) * bcs;
        o.cr += -bus + T(0.5) * bcs;
    }
    return o;
}

// ------------------------------------------------------------ face_signed (numerics.rs:97-112)
// dz = z_R - z_L; NX: x-face (normal velocity u), else y-face (normal v).
template <bool NX, class T>
__host__ __device__ inline F<T> face(T hl, const V<T>& vl, T hr, const V<T>& vr, T dz, const Consts<T>& k) {
    const bool pos = dz > T(0);
    const T hsl = smx(hl - smx(dz, T(0)), T(0));
    const T hsr = smx(hr - smx(-dz, T(0)), T(0));
    const T cred = rt(k.g * (pos ? hsl : hsr));
    const T csl = pos ? cred : vl.c;
    const T csr = pos ? vr.c : cred;
    if (NX) return hll(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

template <bool NX, class T>
__host__ __device__ inline FaceBar<T> face_adj(T hl, const V<T>& vl, T hr, const V<T>& vr, T dz, const Consts<T>& k,
                                               const F<T>& b) {
    const bool pos = dz > T(0);
    const T al = hl - smx(dz, T(0)), ar = hr - smx(-dz, T(0));
    const T hsl = smx(al, T(0)), hsr = smx(ar, T(0));
    const T cred = rt(k.g * (pos ? hsl : hsr));
    const T csl = pos ? cred : vl.c;
    const T csr = pos ? vr.c : cred;
    HllBar<T> hb = NX ? hll_adj(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k, b)
                      : hll_adj(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k, b);
    FaceBar<T> o{};
    if (NX) {
        o.ul = hb.ul; o.vl = hb.tl; o.ur = hb.ur; o.vr = hb.tr;
    } else {
        o.vl = hb.ul; o.ul = hb.tl; o.vr = hb.ur; o.ur = hb.tr;
    }
    T bcred = 0;
    if (pos) {
        bcred = hb.cl;
        o.cr = hb.cr;
    } else {
        bcred = hb.cr;
        o.cl = hb.cl;
    }
    const T bhred = cred > T(0) ? bcred * k.g * rcp(T(2) * cred) : T(0);
    T bhsl = hb.hl, bhsr = hb.hr;
    if (pos) bhsl += bhred;
    else bhsr += bhred;
    if (al > T(0)) {
        o.hl += bhsl;
        if (dz > T(0)) o.dz -= bhsl;
    }
    if (ar > T(0)) {
        o.hr += bhsr;
        if (-dz > T(0)) o.dz += bhsr;
    }
    return o;
}

// ------------------------------------------------------------ face_crest (numerics.rs:117-129)
// dl, dr: the rises from the left / right bed to the face crest (both >= 0); a side with
// a rise reconstructs its celerity from its reduced depth.
template <bool NX, class T>
__host__ __device__ inline F<T> face_cr(T hl, const V<T>& vl, T hr, const V<T>& vr, T dl, T dr, const Consts<T>& k) {
    const T hsl = smx(hl - dl, T(0));
    const T hsr = smx(hr - dr, T(0));
    const T csl = dl > T(0) ? rt(k.g * hsl) : vl.c;
    const T csr = dr > T(0) ? rt(k.g * hsr) : vr.c;
    if (NX) return hll(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

template <bool NX, class T>
__host__ __device__ inline FaceBar<T> face_cr_adj(T hl, const V<T>& vl, T hr, const V<T>& vr, T dl, T dr,
                                                  const Consts<T>& k, const F<T>& b) {
    const T al = hl - dl, ar = hr - dr;
    const T hsl = smx(al, T(0)), hsr = smx(ar, T(0));
    const T csl = dl > T(0) ? rt(k.g * hsl) : vl.c;
    const T csr = dr > T(0) ? rt(k.g * hsr) : vr.c;
    HllBar<T> hb = NX ? hll_adj(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k, b)
                      : hll_adj(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k, b);
    FaceBar<T> o{};
    if (NX) {
        o.ul = hb.ul; o.vl = hb.tl; o.ur = hb.ur; o.vr = hb.tr;
    } else {
        o.vl = hb.ul; o.ul = hb.tl; o.vr = hb.ur; o.ur = hb.tr;
    }
    T bhsl = hb.hl, bhsr = hb.hr;
    if (dl > T(0)) {
        if (csl > T(0)) bhsl += hb.cl * k.g * rcp(T(2) * csl);
    } else {
        o.cl = hb.cl;
    }
    if (dr > T(0)) {
        if (csr > T(0)) bhsr += hb.cr * k.g * rcp(T(2) * csr);
    } else {
        o.cr = hb.cr;
    }
    if (al > T(0)) {
        o.hl += bhsl;
        o.dz -= bhsl;
    }
    if (ar > T(0)) {
        o.hr += bhsr;
        o.dz2 -= bhsr;
    }
    return o;
}

// ------------------------------------------------------------ sources (numerics.rs:158-195)
// (h1, qx1, qy1) after the flux divergence -> new (h, qx, qy).
template <class T>
__host__ __device__ inline S3<T> sources(T h1, T qx1, T qy1, T wet, T gn2, const Consts<T>& k) {
    if (h1 < T(0)) h1 = T(0);
    const T h2 = h1 + k.rain * wet;
    const T h3 = h2 - smn(k.cap * wet, h2);
    const bool dry = h3 < k.h_dry;
    const T r = icbrt(dry ? T(1) : h3);
    const T r3 = r * r * r, r7 = r3 * r3 * r;
    const T qm = rt(qx1 * qx1 + qy1 * qy1);
    const T inv = rcp(T(1) + k.dt * gn2 * qm * r7);
    const T keep = dry ? T(0) : T(1);
    T qx2 = qx1 * inv * keep, qy2 = qy1 * inv * keep;
    if (sab(qx2) < k.q_min) qx2 = T(0);
    if (sab(qy2) < k.q_min) qy2 = T(0);
    return S3<T>{h3, qx2, qy2};
}

template <class T>
__host__ __device__ inline S3<T> sources_adj(T h1, T qx1, T qy1, T wet, T gn2, const Consts<T>& k, const S3<T>& b) {
    const bool pos1 = !(h1 < T(0));
    const T h1c = pos1 ? h1 : T(0);
    const T h2 = h1c + k.rain * wet;
    const bool capped = k.cap * wet < h2;   // infiltration at capacity: h3 = h2 - cap
    const T h3 = h2 - (capped ? k.cap * wet : h2);
    const bool dry = h3 < k.h_dry;
    S3<T> o{};
    T bh3 = b.h;
    if (!dry) {
        const T r = icbrt(h3);
        const T r3 = r * r * r, r7 = r3 * r3 * r;
        const T qm = rt(qx1 * qx1 + qy1 * qy1);
        const T Kf = k.dt * gn2;
        const T inv = rcp(T(1) + Kf * qm * r7);
        const T qx2 = qx1 * inv, qy2 = qy1 * inv;
        const T bqx2 = sab(qx2) < k.q_min ? T(0) : b.qx;
        const T bqy2 = sab(qy2) < k.q_min ? T(0) : b.qy;
        const T binv = bqx2 * qx1 + bqy2 * qy1;
        o.qx = bqx2 * inv;
        o.qy = bqy2 * inv;
        const T bden = -binv * inv * inv;
        const T bqm = bden * Kf * r7;
        const T br7 = bden * Kf * qm;
        if (qm > T(0)) {
            const T rq = rcp(qm);
            o.qx += bqm * qx1 * rq;
            o.qy += bqm * qy1 * rq;
        }
        bh3 += br7 * (T(-7) / T(3)) * r7 * rcp(h3);
    }
    const T bh2 = capped ? bh3 : T(0);
    o.h = pos1 ? bh2 : T(0);
    return o;
}

}  // namespace adj
}  // namespace itr
