// Drives the C++ Stokes facade (include/fsd_stokes.hpp) the way the
// reference's integrator is driven: set-up calls in order, then two steps of
// 64 spheres on a cubic lattice under shear.  Exit 0 when both steps converge
// and the particles moved by a finite amount.  usage: stokes_facade_example TABLE
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "fsd_stokes.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: %s resistance_table.txt\n", argv[0]); return 2; }
    std::vector<float> dist, vals;
    std::ifstream in(argv[1]);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        float r;
        ss >> r;
        dist.push_back(r);
        for (int c = 0; c < 22; ++c) { float v; ss >> v; vals.push_back(v); }
    }
    if (dist.size() != 1000 || vals.size() != 22000) { fprintf(stderr, "bad table\n"); return 2; }
    const int n1 = 4, N = n1 * n1 * n1;
    const float L = 12.0f, a = L / n1;
    std::vector<float> pos(4 * N, 0.f), ori(3 * N, 0.f);
    for (int i = 0; i < N; ++i) {
        pos[4 * i] = -L / 2 + a * (i % n1) + 0.5f * a + 0.05f * (i % 3);
        pos[4 * i + 1] = -L / 2 + a * ((i / n1) % n1) + 0.5f * a;
        pos[4 * i + 2] = -L / 2 + a * (i / (n1 * n1)) + 0.5f * a;
        ori[3 * i] = 1.f;
    }
    float *d_pos, *d_ori, *d_gb, *d_vel;
    int32_t *d_img;
    if (hipMalloc(&d_pos, 16 * N) || hipMalloc(&d_ori, 12 * N) || hipMalloc(&d_gb, 16 * N) ||
        hipMalloc(&d_vel, 44 * N) || hipMalloc(&d_img, 12 * N)) { fprintf(stderr, "hipMalloc failed\n"); return 1; }
    hipMemcpy(d_pos, pos.data(), 16 * N, hipMemcpyHostToDevice);
    hipMemcpy(d_gb, pos.data(), 16 * N, hipMemcpyHostToDevice);
    hipMemcpy(d_ori, ori.data(), 12 * N, hipMemcpyHostToDevice);
    hipMemset(d_img, 0, 12 * N);
    try {
        fsd::Stokes st(N, L, L, L, 0.0f);
        st.setShear([](unsigned int) { return 1.0; }, 0.5f);
        st.setParams();
        st.setResistanceTable(dist.data(), vals.data());
        st.setFriction("NONE", 0.f, nullptr);
        st.setSparseMath();
        st.AllocateWorkSpaces();
        fsd::ParticleArrays arr;
        arr.pos4 = d_pos; arr.ori3 = d_ori; arr.pos_gb4 = d_gb; arr.image3 = d_img; arr.vel11 = d_vel;
        for (unsigned int t = 0; t < 2; ++t) {
            fsd_solve_stats s = st.integrateStepOne(t, arr);
            st.integrateStepTwo(t);
            printf("step %u: %d GMRES iterations, converged %d\n", t, s.iterations, (int)st.converged());
            if (!st.converged()) return 1;
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::vector<float> out(4 * N);
    hipMemcpy(out.data(), d_pos, 16 * N, hipMemcpyDeviceToHost);
    double moved = 0.0;
    for (int i = 0; i < N; ++i) moved = std::fmax(moved, std::fabs(out[4 * i + 1] - pos[4 * i + 1]) + std::fabs(out[4 * i] - pos[4 * i]));
    printf("max displacement %.3e\n", moved);
    return (std::isfinite(moved) && moved > 0.0 && moved < 1.0) ? 0 : 1;
}
