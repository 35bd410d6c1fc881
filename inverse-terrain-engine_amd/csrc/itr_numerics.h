// itr_numerics.h -- per-face and per-cell arithmetic of the f32 shallow-water step,
// as device functions for gfx950.
//
// Each function computes exactly the IEEE single-precision operation sequence of the
// reference's CPU kernel (crates/itr-hydro/src/numerics.rs), so the GPU state is bitwise
// equal to the reference's f32 fused kernel.  Rules that keep it so:
//   * built with -ffp-contract=off: no a*b+c -> fma contraction;
//   * '/' and sqrt stay IEEE correctly rounded (hipcc's default
//     -fhip-fp32-correctly-rounded-divide-sqrt, or the exhaustively verified short
//     sequences crecip / csqrt below), f32 denormals preserved;
//   * max/min are selects with the reference's NaN rule (a > b ? a : b), never
//     fmaxf/fminf; abs is a select that keeps -0.0 (fabsf only where the sign of a zero
//     cannot reach the result);
//   * every expression keeps the reference's association order (left to right).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace itr {

// Per-step constants, each rounded once from an f64 value on the control side
// (crates/itr-hydro/src/solver.rs:269-285).
struct StepK {
    float g, g_half, dt, dtdx, dtdy, inv_dx, inv_dy, h_dry, heps2, q_min, rain, cap;
};

struct Vel { float u, v, c; };
struct Flux { float m, nl, nr, t; };

__device__ __forceinline__ float smax(float a, float b) { return a > b ? a : b; }   // real.rs:44-46
__device__ __forceinline__ float smin(float a, float b) { return a < b ? a : b; }   // real.rs:48-50
__device__ __forceinline__ float sabs(float x) { return x < 0.0f ? -x : x; }        // real.rs:39-41
__device__ __forceinline__ float canon(float x) { return x == 0.0f ? 0.0f : x; }    // real.rs:56-58
// Correctly rounded f32 square root and reciprocal in few instructions.  Each takes the
// hardware estimate and one Newton step whose residual is computed with a fused
// multiply-add (Markstein's construction, "IA-64 and Elementary Functions", 2000); an
// exhaustive run over all 2^32 inputs (itr_hip_math_selfcheck) shows the short sequence
// equals the IEEE operation on the ranges guarded below, and every other input (denormal
// or tiny, huge, inf, NaN, negative) takes the IEEE operation itself, which the step
// never needs on a healthy state.
//
//
// The branch to the IEEE operation is taken on a ballot of the guard (a wave-uniform
// branch on a scalar mask): on the common path that costs one scalar test and branch,
// where a per-lane branch saves, sets and restores the execution mask around it.
__device__ __forceinline__ bool any_lane(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }

// sqrt: y = rsq(x) (clamped so that x = +0 gives +0), s = x y, s + (x - s^2) y / 2;
// exact for x = +0 and 2^-102 <= x < inf.
__device__ __forceinline__ float csqrt(float x) {
    float y = __builtin_amdgcn_rsqf(x);
    y = y < 0x1p64f ? y : 0x1p64f;
    const float s = x * y;
    const float e = __builtin_fmaf(-s, s, x);
    float r = __builtin_fmaf(e, 0.5f * y, s);
    const uint32_t b = __float_as_uint(x);
    const bool slow = b - 0x0C800000u >= 0x7F800000u - 0x0C800000u && b != 0u;
    if (__builtin_expect(any_lane(slow), 0)) r = slow ? __builtin_sqrtf(x) : r;
    return r;
}

// 1/x: r = rcp(x), r + (1 - x r) r; exact for 2^-126 <= x < 2^126.
__device__ __forceinline__ float crecip(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    float q = __builtin_fmaf(e, r, r);
    const uint32_t b = __float_as_uint(x);
    const bool slow = b - 0x00800000u >= 0x7E000000u - 0x00800000u;
    if (__builtin_expect(any_lane(slow), 0)) q = slow ? 1.0f / x : q;
    return q;
}

// 1/x for operands the step bounds below: x >= 2^-126 whenever x is not NaN (the
// velocity denominator h^2 + max(h^2, h_eps^2) >= h_eps^2, which the host requires to be a
// normal f32; the friction denominator 1 + dt g n^2 |q| r^7 >= 1).  Then one comparison
// separates the verified range [2^-126, 2^126) from NaN and values at or above 2^126,
// which take the IEEE operation.
__device__ __forceinline__ float crecip_lower_bounded(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    float q = __builtin_fmaf(e, r, r);
    const bool slow = !(x < 0x1p126f);
    if (__builtin_expect(any_lane(slow), 0)) q = slow ? 1.0f / x : q;
    return q;
}

// x^(-1/3): integer seed then three multiplication-only Newton steps (real.rs:116-129).
__device__ __forceinline__ float inv_cbrt(float x) {
    const uint32_t seed = 0x54A23280u - __float_as_uint(x) / 3u;
    float r = __uint_as_float(seed);
    const float third_x = x * (float)(1.0 / 3.0);
    const float four_thirds = (float)(4.0 * (1.0 / 3.0));
#pragma unroll
    for (int it = 0; it < 3; it++) {
        const float r2 = r * r;
        r = r * four_thirds - third_x * (r2 * r2);
    }
    return r;
}

// Desingularised velocity and celerity of one cell (numerics.rs:36-43).
__device__ __forceinline__ Vel cell_velocity(float h, float qx, float qy, const StepK& k) {
    const float h2 = h * h;
    // smax(h2, heps2): h2 is an arithmetic result (no signalling NaN) and never -0, and
    // heps2 > 0, so the hardware max (NaN -> heps2) is the same select.
    const float inv = crecip_lower_bounded(h2 + __builtin_fmaxf(h2, k.heps2));
    const float t = (h + h) * inv;
    Vel o;
    o.u = qx * t;
    o.v = qy * t;
    o.c = csqrt(k.g * h);
    return o;
}

// Two-rarefaction HLL on reconstructed states (numerics.rs:58-93).  un/ut: normal and
// tangential velocity of each side.
__device__ __forceinline__ Flux hll_flux(float hl, float ul, float tl, float cl, float hr, float ur, float tr,
                                         float cr, const StepK& k) {
    const bool dry_l = !(hl > 0.0f);
    const bool dry_r = !(hr > 0.0f);
    const float ustar = 0.5f * (ul + ur) + cl - cr;
    const float cstar = 0.5f * (cl + cr) + 0.25f * (ul - ur);
    const float slw = smin(ul - cl, ustar - cstar);
    const float srw = smax(ur + cr, ustar + cstar);
    const float sl = dry_l ? ur - 2.0f * cr : (dry_r ? ul - cl : slw);
    const float sr = dry_l ? ur + cr : (dry_r ? ul + 2.0f * cl : srw);
    const float ml = hl * ul;
    const float mr = hr * ur;
    const float pl_ = k.g_half * hl * hl;
    const float pr_ = k.g_half * hr * hr;
    const float pl = ml * ul + pl_;
    const float pr = mr * ur + pr_;
    const bool both_dry = dry_l && dry_r;
    const bool use_l = sl >= 0.0f;
    const bool use_r = sr <= 0.0f;
    const float den = (both_dry || use_l || use_r) ? 1.0f : sr - sl;
    const float inv = crecip(den);
    const float slsr = sl * sr;
    const float fm_star = (sr * ml - sl * mr + slsr * (hr - hl)) * inv;
    const float fn_star = (sr * pl - sl * pr + slsr * (mr - ml)) * inv;
    const float fm = both_dry ? 0.0f : (use_l ? ml : (use_r ? mr : fm_star));
    const float fn = both_dry ? 0.0f : (use_l ? pl : (use_r ? pr : fn_star));
    Flux f;
    f.m = fm;
    f.nl = fn - pl_;
    f.nr = fn - pr_;
    f.t = fm * (fm > 0.0f ? tl : tr);
    return f;
}

// Depth-only hydrostatic reconstruction, signed bed jump dz = z_R - z_L (numerics.rs:97-112).
template <bool NORMAL_X>
__device__ __forceinline__ Flux face_signed(float hl, const Vel& vl, float hr, const Vel& vr, float dz,
                                            const StepK& k) {
    // smax(dz, 0) and smax(-dz, 0) as selects on pos (bit selects, so that they stay one
    // instruction each).  For every non-NaN dz (bed jumps are finite) they give the
    // reference's values, except smax(-dz, 0) = -0 instead of +0 at dz = 0, which
    // smax(hr - rise_r, 0) below cannot tell apart.
    const bool pos = dz > 0.0f;
    const float rise_l = __int_as_float(pos ? __float_as_int(dz) : 0);
    const float rise_r = pos ? 0.0f : -dz;
    const float hsl = smax(hl - rise_l, 0.0f);
    const float hsr = smax(hr - rise_r, 0.0f);
    const float hred = pos ? hsl : hsr;
    const float cred = csqrt(k.g * hred);
    const float csl = pos ? cred : vl.c;
    const float csr = pos ? vr.c : cred;
    if (NORMAL_X) return hll_flux(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll_flux(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

// Face-crest reconstruction: two non-negative rises to the crest (numerics.rs:117-129).
template <bool NORMAL_X>
__device__ __forceinline__ Flux face_crest(float hl, const Vel& vl, float hr, const Vel& vr, float d_l, float d_r,
                                           const StepK& k) {
    const float hsl = smax(hl - d_l, 0.0f);
    const float hsr = smax(hr - d_r, 0.0f);
    const float csl = d_l > 0.0f ? csqrt(k.g * hsl) : vl.c;
    const float csr = d_r > 0.0f ? csqrt(k.g * hsr) : vr.c;
    if (NORMAL_X) return hll_flux(hsl, vl.u, vl.v, csl, hsr, vr.u, vr.v, csr, k);
    return hll_flux(hsl, vl.v, vl.u, csl, hsr, vr.v, vr.u, csr, k);
}

struct CellNew {
    float h, qx, qy, s, infil, fix;
    float h1, qm;   // the depth after the flux divergence and |q| (the health test's operands)
    bool bad;
};

// Flux divergence then sources (numerics.rs:148-153 and 158-195).  Face roles: west
// supplies (m, n_r, t), east (m, n_l, t), north (m, n_r, t), south (m, n_l, t).
__device__ __forceinline__ CellNew cell_update(float h, float qx, float qy, float wm, float wnr, float wt, float em,
                                               float enl, float et, float nm, float nnr, float nt, float sm,
                                               float snl, float st, float wet, float gn2, const StepK& k) {
    float h1 = h - k.dtdx * (em - wm) - k.dtdy * (sm - nm);
    const float qx1 = qx - k.dtdx * (enl - wnr) - k.dtdy * (st - nt);
    const float qy1 = qy - k.dtdx * (et - wt) - k.dtdy * (snl - nnr);
    CellNew o;
    o.h1 = h1;
    bool bad = !(h1 >= -k.h_dry);
    o.fix = h1 < 0.0f ? -h1 : 0.0f;
    if (h1 < 0.0f) h1 = 0.0f;
    const float h2 = h1 + k.rain * wet;
    o.infil = smin(k.cap * wet, h2);
    const float h3 = h2 - o.infil;
    const bool dry = h3 < k.h_dry;
    const float r = inv_cbrt(dry ? 1.0f : h3);
    const float r3 = r * r * r;
    const float r7 = r3 * r3 * r;
    const float qm = csqrt(qx1 * qx1 + qy1 * qy1);
    const float inv = crecip_lower_bounded(1.0f + k.dt * gn2 * qm * r7);
    const float keep = dry ? 0.0f : 1.0f;
    float qx2 = qx1 * inv * keep;
    float qy2 = qy1 * inv * keep;
    if (sabs(qx2) < k.q_min) qx2 = 0.0f;
    if (sabs(qy2) < k.q_min) qy2 = 0.0f;
    const float c = csqrt(k.g * h3);
    // |qx2| with the sign bit cleared rather than sabs: they differ only at -0 (and in a
    // NaN's sign), and (-0) r3 + c = (+0) r3 + c for the finite r3 > 0 and c >= +0 here.
    o.s = (__builtin_fabsf(qx2) * r3 + c) * k.inv_dx + (__builtin_fabsf(qy2) * r3 + c) * k.inv_dy;
    bad = bad || !(qm == qm);
    o.qm = qm;
    // canon_zero (real.rs:56-58) of the new state is the identity here, so it is not
    // evaluated: h3 is never -0 (the old depth is never -0, x - y is -0 only for x = -0,
    // and the rain and infiltration terms keep that), and qx2 / qy2 are +0 or at least
    // q_min in magnitude (or NaN) after the q_min test.
    o.h = h3;
    o.qx = qx2;
    o.qy = qy2;
    o.bad = bad;
    return o;
}

}  // namespace itr
