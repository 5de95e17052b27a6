/*
 * statmech_oracle.cpp — CPU oracle for the ensemble reduction.
 *
 * TEST INFRASTRUCTURE ONLY: the shipped library never links this file.
 *
 *   fa_oracle_log_sum_exp   LIB/UnifiedHardwareDispatch.cpp:800-843 (max shift, Σexp)
 *   fa_oracle_thermo        LIB/statmech.cpp:73-173 (log w = ln n − βE, LSE,
 *                           moments, F, S, Cv, σ_E)
 *   fa_oracle_boltzmann     LIB/UnifiedHardwareDispatch.cpp:730-747 and
 *                           LIB/statmech.cpp:178-201 (weights with multiplicity)
 *   fa_oracle_boltzmann_batch  LIB/UnifiedHardwareDispatch.cpp:730-746, 768-794
 *                           (compute_boltzmann_batch: unnormalised weights
 *                           exp(−β(E−Emin)), log_Z = ln Σw − β Emin)
 *   fa_oracle_log_xi        LIB/GrandPartitionFunction.cpp:24-40, 112-160
 *
 * Parity status: pinned by closed forms (two-state and equal-energy
 * ensembles, ln 2 sums) and the reference's analytic test values in
 * tests/test_statmech_oracle.py.
 */
#include "fa_types.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace {
constexpr double KB_KCAL = 0.001987206;  // LIB/statmech.h:26
}

extern "C" {

double fa_oracle_log_sum_exp(const double* x, int n)
{
    if (n <= 0) return -std::numeric_limits<double>::infinity();
    double m = x[0];
    for (int i = 1; i < n; ++i) m = std::max(m, x[i]);
    if (!std::isfinite(m)) return m;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::exp(x[i] - m);
    return m + std::log(s);
}

int fa_oracle_thermo(const double* energy, const int* count, int n, double T, fa_thermo* out)
{
    if (n <= 0 || T <= 0.0 || !energy || !out) return -1;
    const double beta = 1.0 / (KB_KCAL * T);
    std::vector<double> lw(n);
    for (int i = 0; i < n; ++i) lw[i] = std::log((double)(count ? count[i] : 1)) - beta * energy[i];
    const double lnZ = fa_oracle_log_sum_exp(lw.data(), n);
    double e1 = 0.0, e2 = 0.0;
    for (int i = 0; i < n; ++i) {
        const double p = std::exp(lw[i] - lnZ);
        e1 += p * energy[i];
        e2 += p * energy[i] * energy[i];
    }
    const double kT = KB_KCAL * T;
    const double var = e2 - e1 * e1;
    out->temperature = T;
    out->log_Z = lnZ;
    out->free_energy = -kT * lnZ;
    out->mean_energy = e1;
    out->mean_energy_sq = e2;
    out->heat_capacity = var / (kT * kT);
    out->entropy = (e1 - out->free_energy) / T;
    out->std_energy = std::sqrt(std::max(0.0, var));
    return 0;
}

/* Normalised Boltzmann weights with multiplicities; returns ln Z of the
 * multiplicity-free batch (log Σ exp(−β(E−Emin)) − β Emin). */
double fa_oracle_boltzmann(const double* energy, const int* count, int n, double T, double* w)
{
    if (n <= 0) return 0.0;
    const double beta = 1.0 / (KB_KCAL * T);
    double emin = energy[0];
    for (int i = 1; i < n; ++i) emin = std::min(emin, energy[i]);
    double sum = 0.0, zm = 0.0;
    std::vector<double> raw(n);
    for (int i = 0; i < n; ++i) {
        raw[i] = std::exp(-beta * (energy[i] - emin));
        sum += raw[i];
        zm += (count ? count[i] : 1) * raw[i];
    }
    if (w) {
        for (int i = 0; i < n; ++i) w[i] = (zm > 0.0) ? (count ? count[i] : 1) * raw[i] / zm : 0.0;
    }
    return std::log(sum) - beta * emin;
}

/* compute_boltzmann_batch (LIB/UnifiedHardwareDispatch.cpp:768-794, the
 * Eigen path :730-746) on n >= 1 energies at inverse temperature beta:
 * w[i] = exp(−β (E_i − E_min)) (not normalised), *log_Z = ln Σ w − β E_min,
 * *e_min = E_min.  n = 0 gives log_Z = 0, e_min = 0 as the reference's
 * empty-span result does. */
int fa_oracle_boltzmann_batch(const double* energy, int n, double beta, double* w, double* log_Z, double* e_min)
{
    if (n <= 0) {
        if (log_Z) *log_Z = 0.0;
        if (e_min) *e_min = 0.0;
        return 0;
    }
    double emin = energy[0];
    for (int i = 1; i < n; ++i) emin = std::min(emin, energy[i]);
    double sum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double wi = std::exp(-beta * (energy[i] - emin));
        if (w) w[i] = wi;
        sum += wi;
    }
    if (log_Z) *log_Z = std::log(sum) - beta * emin;
    if (e_min) *e_min = emin;
    return 0;
}

/* ln Ξ = LSE(0, ln c_i + ln Z_i); p_i = exp(ln c_i + ln Z_i − ln Ξ);
 * p_empty = exp(−ln Ξ). */
double fa_oracle_log_xi(const double* log_Z, const double* conc_M, int n, double* p_bound, double* p_empty)
{
    if (n <= 0) {
        if (p_empty) *p_empty = 1.0;
        return 0.0;
    }
    std::vector<double> lzz(n);
    double m = 0.0;
    for (int i = 0; i < n; ++i) {
        lzz[i] = std::log(conc_M[i]) + log_Z[i];
        m = std::max(m, lzz[i]);
    }
    double s = std::exp(0.0 - m);
    for (int i = 0; i < n; ++i) s += std::exp(lzz[i] - m);
    const double lx = m + std::log(s);
    if (p_bound)
        for (int i = 0; i < n; ++i) p_bound[i] = std::exp(lzz[i] - lx);
    if (p_empty) *p_empty = std::exp(-lx);
    return lx;
}

}  // extern "C"
