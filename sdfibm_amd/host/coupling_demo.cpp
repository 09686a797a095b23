// coupling_demo.cpp -- a C++ host driving the coupling step through the C++
// mirror of SolidCloud (include/sdfibm_hip.hpp), the way the reference's
// solver loop calls it (src/main.cpp:41-104): interact, then U -= Fs*dt, then
// read the solids' fluid forces.  The fluid solve itself (OpenFOAM's PISO) is
// not part of this program: U is a prescribed uniform stream.
//
// Usage: coupling_demo [steps]
// Prints one line per step: step, sum of As*V, force on solid 0.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/sdfibm_hip.hpp"

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 3;
    const int nx = 64, ny = 64, nz = 1;
    std::vector<double> xs(nx + 1), ys(ny + 1), zs = {-0.5, 0.5};
    for (int i = 0; i <= nx; ++i) xs[i] = -2.0 + 4.0 * i / nx;
    for (int j = 0; j <= ny; ++j) ys[j] = -2.0 + 4.0 * j / ny;
    sdfibm_mesh_desc mesh{};
    mesh.kind = SDFIBM_MESH_RECTILINEAR;
    mesh.nx = nx;
    mesh.ny = ny;
    mesh.nz = nz;
    mesh.xs = xs.data();
    mesh.ys = ys.data();
    mesh.zs = zs.data();
    sdfibm_config cfg{};
    cfg.two_d = 1;
    cfg.precision = SDFIBM_FP64;
    cfg.rho_f = 1.0;
    try {
        sdfibm::hip::SolidCloud cloud(mesh, cfg);
        sdfibm_solid disc{};
        disc.shape = SDFIBM_SHAPE_CIRCLE;
        disc.prm[0] = 0.5;
        disc.q[0] = 1.0;
        cloud.setSolids({disc});
        const long long N = (long long)nx * ny * nz;
        std::vector<double> U(3 * N, 0.0), As, Fs;
        for (long long c = 0; c < N; ++c) U[3 * c] = 1.0;
        const double dt = 0.01, V = (4.0 / nx) * (4.0 / ny) * 1.0;
        for (int s = 0; s < steps; ++s) {
            cloud.setVelocity(U.data());
            cloud.interact(s * dt, dt);
            cloud.fields(&As, &Fs, nullptr, nullptr);
            double vol = 0.0;
            for (long long c = 0; c < N; ++c) {
                vol += As[c] * V;
                for (int d = 0; d < 3; ++d) U[3 * c + d] -= Fs[3 * c + d] * dt;  // main.cpp:70
            }
            const auto F = cloud.fluidForce(0);
            std::printf("%d %.12e %.12e %.12e\n", s, vol, F[0], F[1]);
        }
    } catch (const sdfibm::hip::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
