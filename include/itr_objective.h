/*
 * itr_objective.h -- the search objective on monitor maxima, header-only C.
 *
 * One implementation of J for every host of the library: the adjoint driver's J_risk
 * and peak seeds (itr_hip.hip), itr_hip_objective_j (the Python mirror's J and the
 * incumbent lower bound its sync callbacks compute), and the CPU oracle's incumbent
 * abort.  It restates crates/itr-core/src/objective.rs:
 *   softplus              objective.rs:70-74   (x > 30: x; x < -30: e^x; else log1p(e^x))
 *   asset_peaks           objective.rs:90-92   (max over the asset's monitors, folded from 0
 *                                               with f64::max, which ignores NaN)
 *   evaluate (J terms)    objective.rs:95-114  (J = J_risk + earth_weight * J_earth,
 *                                               J_risk = sum_a w_a tau softplus((peak_a - eta_a) / tau),
 *                                               J_earth = cut_cost cut_m3 + fill_cost fill_m3)
 *   running_lower_bound   objective.rs:146-149 (J of the running maxima)
 * The reference evaluates exp / log1p with the libm crate 0.2.16 (not in this image);
 * here they are the C library's, so J agrees with the reference's to the last ulps of
 * those two functions.
 */
#ifndef ITR_OBJECTIVE_H
#define ITR_OBJECTIVE_H

#include <math.h>
#include <stdint.h>

#include "itr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

static inline double itr_softplus(double x) {
    return x > 30.0 ? x : (x < -30.0 ? exp(x) : log1p(exp(x)));
}

/* f64::max (NaN-ignoring): the other operand when one is NaN. */
static inline double itr_fmax_rs(double a, double b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}

/* Peak of asset a: max over its monitor positions, folded from 0.  *argmax receives the
 * first position that raised the fold (-1 when none did: the peak is the initial 0). */
static inline double itr_asset_peak(const itr_objective* o, int a, const double* monitor_max, int32_t* argmax) {
    double pk = 0.0;
    int32_t k_at = -1;
    for (int32_t q = o->asset_start[a]; q < o->asset_start[a + 1]; q++) {
        const int32_t k = o->asset_pos[q];
        const double v = itr_fmax_rs(pk, monitor_max[k]);
        if (v != pk) k_at = k;
        pk = v;
    }
    if (argmax) *argmax = k_at;
    return pk;
}

/* J_risk, and each asset's peak into peaks[n_assets] (may be NULL). */
static inline double itr_j_risk(const itr_objective* o, const double* monitor_max, double* peaks) {
    double j = 0.0;
    for (int a = 0; a < o->n_assets; a++) {
        const double pk = itr_asset_peak(o, a, monitor_max, 0);
        j += o->weight[a] * o->tau_m * itr_softplus((pk - o->threshold_m[a]) / o->tau_m);
        if (peaks) peaks[a] = pk;
    }
    return j;
}

/* J = J_risk + earth_weight * J_earth (earth NULL: no earthwork term). */
static inline double itr_objective_value(const itr_objective* o, const double* monitor_max, const itr_earthwork* earth,
                                         double* peaks, double* j_risk_out) {
    const double jr = itr_j_risk(o, monitor_max, peaks);
    if (j_risk_out) *j_risk_out = jr;
    if (!earth) return jr + 0.0;
    const double je = earth->cut_cost * earth->cut_m3 + earth->fill_cost * earth->fill_m3;
    return jr + earth->earth_weight * je;
}

#ifdef __cplusplus
}
#endif
#endif
