// Host check of the kernels' voxel-index and division helpers (device_math.hpp):
// axis_index_fast must give axis_index's result for every input, udiv_by_inv the
// integer quotient.  Random inputs plus positions one ulp either side of every kind
// of voxel face, NaN and huge values.  Prints the counts; exit status 1 on a mismatch.
#include <cmath>
#include <cstdio>
#include <random>

#include "device_math.hpp"

int main(int argc, char** argv) {
    const long n_cases = argc > 1 ? atol(argv[1]) : 2000000;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0;
    for (long t = 0; t < n_cases; t++) {
        const double mn = (U(g) - 0.5) * 1000.0, dx = 0.01 + U(g) * 20.0;
        const uint32_t n = 1 + (uint32_t)(U(g) * 5000.0);
        double p;
        switch (t % 4) {
            case 0: p = mn + U(g) * dx * (n + 2) - dx; break;                        // inside and around the grid
            case 1: p = std::nextafter(mn + (uint32_t)(U(g) * n) * dx, (t & 8) ? 1e300 : -1e300); break;
            case 2: p = mn + (uint32_t)(U(g) * n) * dx; break;                       // on a face
            default: p = mn + (U(g) - 0.5) * 4.0 * dx; break;                       // around the lower bound
        }
        uint32_t a = 7, b = 7;
        const bool x = crm::axis_index(p, mn, dx, n, a);
        const bool y = crm::axis_index_fast(p, mn, dx, 1.0 / dx, n, b);
        if (x != y || (x && a != b)) bad++;
    }
    const double odd[] = {NAN, 1e20, -1e20, INFINITY, -0.5, -0.0};
    for (double p : odd) {
        uint32_t a = 0, b = 0;
        if (crm::axis_index(p, 0.0, 1.0, 10, a) != crm::axis_index_fast(p, 0.0, 1.0, 1.0, 10, b)) bad++;
    }
    long badd = 0;
    for (long t = 0; t < n_cases; t++) {
        const uint32_t d = 1 + (uint32_t)(g() % ((t & 1) ? 100000ull : 4294967295ull));
        const uint64_t q = g() % (4294967295ull / d + 1), av = q * d + g() % d;
        if (av > 4294967295ull) continue;
        if (crm::udiv_by_inv((uint32_t)av, d, 1.0 / (double)d) != (uint32_t)av / d) badd++;
    }
    std::printf("axis_index_fast: %ld cases, %ld mismatches; udiv_by_inv: %ld mismatches\n", n_cases, bad, badd);
    return bad || badd ? 1 : 0;
}
