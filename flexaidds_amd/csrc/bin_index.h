/* bin_index.h — the bin count of the reference's genetoic, without its loop.
 *
 * genetoic (LIB/gaboom.cpp:2887-2906) counts the additions of `bin` to a
 * running sum that starts at `bin` until the sum reaches u:
 *
 *     tot = bin; i = 0; while (tot < u) { tot += bin; ++i; }
 *
 * Every addition is rounded, so i is not simply ceil(u / bin) - 1, and the
 * loop runs up to one iteration per bin (thousands for a cleft-grid gene).
 *
 * Inside one binade [2^(e-1), 2^e) the doubles are the multiples of
 * ulp = 2^(e-53).  A sum s in the binade is such a multiple, so when
 * s + bin stays below 2^e the rounded sum is s + c, with c = bin rounded to a
 * multiple of ulp, the same at every such step (bin / ulp is exact; when its
 * fraction is exactly one half the rounding depends on the last bit of s,
 * and those steps are taken one at a time).  Runs of such steps are jumped in
 * one go; every value compared along the way (s + k c, 2^e - s - k c) is a
 * multiple of ulp inside the binade, so the comparisons are exact.  The step
 * that leaves a binade is the reference's own addition.  The count equals the
 * loop's for every bin > 0 (checked against the loop on the host over bins
 * 1/n and random bins, at and around every reachable sum).
 */
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define FA_BIN_FN __host__ __device__ inline
#else
#define FA_BIN_FN static inline
#endif

FA_BIN_FN int fa_bin_index(double u, double bin)
{
    double tot = bin;
    int i = 0;
    while (tot < u) {
        int e;
        frexp(tot, &e);  // tot in [2^(e-1), 2^e)
        const double top = ldexp(1.0, e);
        const double ulp = ldexp(1.0, e - 53);
        const double q = bin / ulp;  // exact: a power-of-two scaling
        const double qf = floor(q);
        const double fr = q - qf;
        const double c = (fr < 0.5 ? qf : qf + 1.0) * ulp;
        if (fr != 0.5 && c > 0.0) {
            const double D = top - tot;  // exact
            // uniform steps from tot: starts k = 0 .. n-1 with tot + k c + bin < top
            double n = 0.0;
            if (bin < D) {
                n = ceil((D - bin) / c);
                if (n < 1.0) n = 1.0;
                while (n > 1.0 && !(bin < D - (n - 1.0) * c)) n -= 1.0;
                while (bin < D - n * c) n += 1.0;
            }
            if (n > 0.0) {
                // the first k >= 1 with tot + k c >= u
                double k = ceil((u - tot) / c);
                if (k < 1.0) k = 1.0;
                while (k > 1.0 && tot + (k - 1.0) * c >= u) k -= 1.0;
                while (tot + k * c < u && k <= n) k += 1.0;
                if (k <= n) return i + (int)k;
                tot += n * c;
                i += (int)n;
            }
        }
        tot += bin;  // one addition as the reference makes it
        ++i;
    }
    return i;
}
