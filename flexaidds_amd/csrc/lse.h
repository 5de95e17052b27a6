// lse.h — FP64 online log-sum-exp state for canonical-ensemble reductions.
//
// A state (m, s, s1, s2) stands for the samples folded so far with
// log-weights lw_i and energies E_i: m = max lw, s = sum exp(lw - m),
// s1 = sum E exp(lw - m), s2 = sum E^2 exp(lw - m).  States merge by
// rescaling to the larger maximum, so any reduction tree gives ln Z, <E>,
// <E^2> (LIB/statmech.cpp:73-173) without overflow.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

#include "fa_types.h"

namespace fa {

constexpr double KB_KCAL = 0.001987206;  // LIB/statmech.h:26

struct LseState {
    double m, s, s1, s2;
};

__device__ __forceinline__ void lse_merge(LseState& a, const LseState& b)
{
    if (b.s == 0.0) return;
    if (a.s == 0.0) {
        a = b;
        return;
    }
    if (b.m > a.m) {
        const double f = exp(a.m - b.m);
        a.s = a.s * f + b.s;
        a.s1 = a.s1 * f + b.s1;
        a.s2 = a.s2 * f + b.s2;
        a.m = b.m;
    } else {
        const double f = exp(b.m - a.m);
        a.s += b.s * f;
        a.s1 += b.s1 * f;
        a.s2 += b.s2 * f;
    }
}

// A sample of log-weight -inf (an energy of +inf: a pose the evaluator could
// not score) has weight exactly 0: it adds nothing to Z, <E> or <E^2> and is
// skipped rather than folded as 0 * inf.
__device__ __forceinline__ void lse_push(LseState& a, double lw, double e)
{
    if (lw == -INFINITY) return;
    if (a.s == 0.0) {
        a.m = lw;
        a.s = 1.0;
        a.s1 = e;
        a.s2 = e * e;
    } else if (lw > a.m) {
        const double f = exp(a.m - lw);
        a.s = a.s * f + 1.0;
        a.s1 = a.s1 * f + e;
        a.s2 = a.s2 * f + e * e;
        a.m = lw;
    } else {
        const double w = exp(lw - a.m);
        a.s += w;
        a.s1 += w * e;
        a.s2 += w * e * e;
    }
}

// thermodynamics of a state at temperature T (log-weights were -E / kT)
__device__ __forceinline__ fa_thermo thermo_of(const LseState& st, double T)
{
    const double lnZ = st.m + log(st.s);
    const double e1 = st.s1 / st.s;
    const double e2 = st.s2 / st.s;
    const double kT = KB_KCAL * T;
    const double var = e2 - e1 * e1;
    fa_thermo th;
    th.temperature = T;
    th.log_Z = lnZ;
    th.free_energy = -kT * lnZ;
    th.mean_energy = e1;
    th.mean_energy_sq = e2;
    th.heat_capacity = var / (kT * kT);
    th.entropy = (e1 - th.free_energy) / T;
    th.std_energy = sqrt(var > 0.0 ? var : 0.0);
    return th;
}

}  // namespace fa
