// Host check of fa_bin_index against the reference's genetoic loop
// (LIB/gaboom.cpp:2887-2906): every bin 1/n for n = 1..N and random bins,
// with u at, just below and just above every sum the loop reaches, plus
// random u.  Prints the number of cases and of mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "bin_index.h"

static int loop_index(double u, double bin)
{
    double tot = bin;
    int i = 0;
    while (tot < u) {
        tot += bin;
        ++i;
    }
    return i;
}

int main(int argc, char** argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 3000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long cases = 0, bad = 0;
    auto check = [&](double u, double bin) {
        ++cases;
        const int a = fa_bin_index(u, bin), b = loop_index(u, bin);
        if (a != b) {
            if (++bad <= 10) std::printf("mismatch u=%.17g bin=%.17g fast=%d loop=%d\n", u, bin, a, b);
        }
    };
    std::vector<double> bins;
    for (int n = 1; n <= N; ++n) bins.push_back(1.0 / n);
    for (int r = 0; r < 300; ++r) bins.push_back(std::ldexp(1.0 + U(rng), -1 - (int)(U(rng) * 14)));
    for (int r = 0; r < 50; ++r) bins.push_back(std::ldexp(std::floor(U(rng) * 64.0) + 1.0, -12));  // few bits
    for (double bin : bins) {
        // every sum the loop reaches up to 1, and its neighbours
        double tot = bin;
        int steps = 0;
        while (tot < 1.0 && steps < 200000) {
            check(tot, bin);
            check(std::nextafter(tot, 0.0), bin);
            check(std::nextafter(tot, 2.0), bin);
            tot += bin;
            ++steps;
        }
        for (int r = 0; r < 200; ++r) check(U(rng), bin);
        // genes as the GA draws them: u = gene / 2^31
        for (int r = 0; r < 200; ++r) check((double)(1 + (rng() % 2147483647u)) / 2147483648.0, bin);
        check(0.0, bin);
        check(1.0, bin);
    }
    std::printf("%lld cases %lld mismatches\n", cases, bad);
    return bad != 0;
}
