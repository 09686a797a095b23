// Host check of row_inside_bits (sdfibm_amd/csrc/shapes.h): the inside bits of random rows of
// vertices against the exact per-vertex test inside_world, for random spheres, ellipsoids,
// circles and ellipses (axis ratios to 10:1, any rotation, centres far from the origin), one row
// in five with a vertex bisected onto the surface.  Exit status 1 on any mismatch.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <cmath>
#include "../../sdfibm_amd/csrc/shapes.h"
using namespace sdfibm_dev;
int main(int argc, char** argv) {
    long long nrows = argc > 1 ? atoll(argv[1]) : 2000000;
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long bad = 0, exact_calls = 0, vert = 0;
    const int shapes[4] = {SDFIBM_SHAPE_SPHERE, SDFIBM_SHAPE_ELLIPSOID, SDFIBM_SHAPE_CIRCLE, SDFIBM_SHAPE_ELLIPSE};
    for (long long r = 0; r < nrows; ++r) {
        sdfibm_solid s{};
        s.shape = shapes[r % 4];
        const double h = std::pow(10.0, -3 + 3 * U(g));  // cell size
        const double R = h * (1 + 20 * U(g));
        double ax = R, bx = R * (0.1 + 0.9 * U(g)), cx = R * (0.1 + 0.9 * U(g));
        s.prm[0] = ax; s.prm[1] = bx; s.prm[2] = cx;
        if (s.shape == SDFIBM_SHAPE_SPHERE || s.shape == SDFIBM_SHAPE_CIRCLE) {
            s.prm[1] = (U(g) - 0.5) * R; s.prm[2] = (U(g) - 0.5) * R; s.prm[3] = (U(g) - 0.5) * R;
        }
        if (s.shape == SDFIBM_SHAPE_ELLIPSE) { s.prm[3] = (U(g) - 0.5) * R; s.prm[4] = (U(g) - 0.5) * R; s.prm[5] = (U(g) - 0.5) * R; }
        const double off = std::pow(10.0, 3 * U(g) - 1) * (U(g) < 0.5 ? 1 : -1);
        for (int d = 0; d < 3; ++d) s.center[d] = off + (U(g) - 0.5) * 100 * h;
        double q[4], n2 = 0;
        std::normal_distribution<double> N(0, 1);
        for (int d = 0; d < 4; ++d) { q[d] = N(g); n2 += q[d] * q[d]; }
        const int axis_aligned = (r % 7 == 0);
        for (int d = 0; d < 4; ++d) s.q[d] = axis_aligned ? (d == 0 ? 1.0 : 0.0) : q[d] / std::sqrt(n2);
        const SolidK K = make_solid_k(s);
        // the row: a uniform x grid of n vertices around the solid, y/z near the solid
        const int n = 1 + (int)(63 * U(g));
        double xs[64];
        const double x0 = s.center[0] - (n / 2) * h + (U(g) - 0.5) * h;
        for (int v = 0; v < n; ++v) xs[v] = x0 + v * h;
        double y = s.center[1] + (U(g) - 0.5) * 2.6 * R, z = s.center[2] + (U(g) - 0.5) * 2.6 * R;
        if (r % 5 == 0) {
            // put a vertex on the surface: bisect along x for a sign change of the exact predicate
            int v = (int)(U(g) * n);
            double lo = xs[v], hi = xs[v] + 3 * R;
            bool in_lo = inside_world(K, d3{lo, y, z});
            for (int it = 0; it < 200; ++it) {
                double mid = 0.5 * (lo + hi);
                if (inside_world(K, d3{mid, y, z}) == in_lo) lo = mid; else hi = mid;
            }
            xs[v] = lo;
            for (int u = v + 1; u < n; ++u) xs[u] = xs[u - 1] + h;
        }
        const unsigned long long got = row_inside_bits(K, xs, n, y, z);
        unsigned long long want = 0;
        for (int v = 0; v < n; ++v) if (inside_world(K, d3{xs[v], y, z})) want |= 1ull << v;
        vert += n;
        if (got != want) { if (bad < 10) printf("mismatch row %lld shape %d got %llx want %llx\n", r, s.shape, got, want); ++bad; }
    }
    printf("rows %lld vertices %lld mismatches %lld\n", nrows, vert, bad);
    return bad != 0;
}
