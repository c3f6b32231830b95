// Host build of the device serial max3 path (solve_max3, one cell at a time): reads a
// mixture and cells from a binary file and prints per cell: index, max3 status, base split
// status, phase count, EOS evaluations.
#include "../../mpmc_hnu_amd/csrc/flash_device.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
using namespace mpmc::dev;
constexpr int NC = 6;
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    int nc = 0; long n = 0;
    if (std::fread(&nc, 4, 1, f) != 1 || std::fread(&n, 8, 1, f) != 1 || nc != NC) { return 2; }
    std::vector<double> tc(nc), pc(nc), om(nc), kij(nc * nc), p(n), t(n), z(nc * n);
    size_t got = std::fread(tc.data(), 8, nc, f) + std::fread(pc.data(), 8, nc, f) + std::fread(om.data(), 8, nc, f);
    got += std::fread(kij.data(), 8, nc * nc, f) + std::fread(p.data(), 8, n, f) + std::fread(t.data(), 8, n, f);
    got += std::fread(z.data(), 8, nc * n, f);
    std::fclose(f);
    if (got != static_cast<size_t>(3 * nc + nc * nc + 2 * n + nc * n)) { return 3; }
    Mix<NC> m;
    const double r_gas = static_cast<double>(8.31446261815324L);
    for (int i = 0; i < NC; ++i) {
        m.sqrt_tc[i] = std::sqrt(tc[i]);
        m.kappa[i] = static_cast<double>(0.37464L) + om[i] * (static_cast<double>(1.54226L) - static_cast<double>(0.26992L) * om[i]);
        const double rtc = r_gas * tc[i];
        m.ac[i] = (static_cast<double>(0.45724L) * rtc) * (rtc / pc[i]);
        m.b[i] = (static_cast<double>(0.07780L) * rtc) / pc[i];
    }
    for (int k = 0; k < NC * NC; ++k) { m.factor[k] = 1.0 - kij[k]; }
    m.r_gas = r_gas;
    const long first = argc > 2 ? std::atol(argv[2]) : 0, last = argc > 3 ? std::atol(argv[3]) : n;
    for (long c = first; c < last && c < n; ++c) {
        double feed[NC];
        for (int i = 0; i < NC; ++i) { feed[i] = z[i * n + c]; }
        Max3Out<NC> M;
        Counters cnt{0, 0, 0, 0};
        bool rej = false;
        solve_max3<NC>(m, p[c], t[c], feed, 1e-11, M, cnt, rej);
        int np = 0;
        if (!rej) {
            if (M.status == M_THREE) { np = 3; }
            else if (M.status == M_TWO || M.base.status == V_TWO) { np = 2; }
            else if (M.status == M_SINGLE) { np = 1; }
        }
        std::printf("%ld %d %d %d %u\n", c, rej ? -1 : M.status, M.base.status, np, cnt.prop);
    }
    return 0;
}
