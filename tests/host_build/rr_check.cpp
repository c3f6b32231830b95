// Host build of the device Rachford-Rice: the certified-sign fast path (enclosure,
// dyadic jump, certified interior test) against plain bisection, bit for bit.
#include "../../mpmc_hnu_amd/csrc/flash_device.hpp"
#include <cstdio>
#include <random>
using namespace mpmc::dev;
template <int NC> int run(int n, unsigned seed) {
    std::mt19937_64 g(seed); std::uniform_real_distribution<double> U(0, 1); std::normal_distribution<double> N(0, 1);
    int diff = 0, interior = 0; long jumped = 0, total = 0;
    for (int c = 0; c < n; ++c) {
        double z[NC], lk[NC], s = 0;
        for (int i = 0; i < NC; ++i) { z[i] = -std::log(U(g)); s += z[i]; }
        for (int i = 0; i < NC; ++i) { z[i] /= s; lk[i] = N(g) * (c % 3 == 0 ? 0.05 : (c % 3 == 1 ? 1.0 : 6.0)); }
        // renormalise in the checked order
        { double t = ksum<NC>(z); for (int i = 0; i < NC; ++i) z[i] /= t; }
        RRState<NC> a{}, b{}; Counters ca{0,0,0,0}, cb{0,0,0,0}; bool ra = false, rb = false;
        const double guess = U(g);
        int sa = rachford_rice<NC>(z, lk, a, ca, ra, guess);
        int sb = rachford_rice<NC>(z, lk, b, cb, rb, -1.0);
        bool same = sa == sb && ca.rr == cb.rr;
        if (sa == RR_INTERIOR && sb == RR_INTERIOR) {
            ++interior;
            same = same && a.beta == b.beta;
            for (int i = 0; i < NC; ++i) same = same && a.x[i] == b.x[i] && a.y[i] == b.y[i];
        }
        if (!same) { if (diff < 5) printf("diff c=%d st %d %d rr %u %u beta %.17g %.17g\n", c, sa, sb, ca.rr, cb.rr, a.beta, b.beta); ++diff; }
    }
    printf("NC=%d n=%d interior=%d diffs=%d\n", NC, n, interior, diff);
    return diff;
}
int main() { int d = run<2>(200000, 1) + run<6>(200000, 2) + run<12>(100000, 3); return d != 0; }
