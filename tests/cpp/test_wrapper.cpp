// Drives the header-only C++ wrapper (include/pbat_vbd.hpp) through
// libpbat_vbd.so on the reference's 8-vertex cube
// (sim/vbd/Integrator.cpp:245-293, gpu/impl/vbd/Integrator.cu:384-432).
// "test_wrapper" checks the host-side Data construction only;
// "test_wrapper gpu" also steps the GPU integrator and checks the fall.
#include "pbat_vbd.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

using pbat_amd::Matrix;

static int fails = 0;
#define EXPECT(cond)                                                                                              \
    do                                                                                                            \
    {                                                                                                             \
        if (!(cond))                                                                                              \
        {                                                                                                         \
            std::printf("FAILED: %s (line %d)\n", #cond, __LINE__);                                              \
            ++fails;                                                                                              \
        }                                                                                                         \
    } while (0)

int main(int argc, char** argv)
{
    bool const gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    double const P[3][8] = {{0, 1, 0, 1, 0, 1, 0, 1}, {0, 0, 1, 1, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
    int64_t const T[4][5] = {{0, 3, 5, 6, 0}, {1, 2, 4, 7, 5}, {3, 0, 6, 5, 3}, {5, 6, 0, 3, 6}};
    Matrix<double> X(3, 8);
    Matrix<int64_t> E(4, 5);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 8; ++c)
            X(r, c) = P[r][c];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 5; ++c)
            E(r, c) = T[r][c];
    pbat_amd::sim::vbd::Data data;
    data.WithVolumeMesh(X, E).Construct();
    int64_t const expected[8] = {0, 3, 2, 1, 1, 2, 3, 0};
    for (int i = 0; i < 8; ++i)
        EXPECT(data.colors[i] == expected[i]);
    EXPECT(data.Pptr.size() == 5 && data.Padj.size() == 8);
    double total = 0.;
    for (double mi : data.m)
        total += mi;
    EXPECT(std::fabs(total - 1000. * 1.0000008) < 1e-9);
    EXPECT(std::fabs(data.aext(2, 0) + 9.81) < 1e-15);

    // collision mesh: the cube's 12 boundary triangles; areas sum to the surface area 6
    int64_t const Fc[3][12] = {{0, 0, 4, 4, 0, 0, 2, 2, 0, 0, 1, 1},
                               {2, 3, 5, 7, 1, 5, 6, 7, 4, 6, 3, 7},
                               {3, 1, 7, 6, 5, 4, 7, 3, 6, 2, 7, 5}};
    Matrix<int64_t> F(3, 12);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 12; ++c)
            F(r, c) = Fc[r][c];
    pbat_amd::sim::vbd::Data surf;
    surf.WithVolumeMesh(X, E).WithSurfaceMesh({0, 1, 2, 3, 4, 5, 6, 7}, F).WithContactParameters(1e5, 0.2, 1e-3);
    double sumFA = 0., sumXVA = 0.;
    for (double a : surf.FA)
        sumFA += a;
    for (double a : surf.XVA)
        sumXVA += a;
    EXPECT(std::fabs(sumFA - 6.) < 1e-12 && std::fabs(sumXVA - 6.) < 1e-12);
    EXPECT(std::fabs(surf.XVA[0] - 1.) < 1e-12); // corner 0 lies in 6 triangles of double area 1: 6 x 1/6
    EXPECT(surf.muC == 1e5 && surf.muF == 0.2);

    bool threw = false;
    try
    {
        Matrix<int64_t> Ebad = E;
        Ebad(0, 0) = 42;
        pbat_amd::sim::vbd::Data bad;
        bad.WithVolumeMesh(X, Ebad).Construct();
    }
    catch (std::invalid_argument const&)
    {
        threw = true;
    }
    EXPECT(threw);

    if (gpu)
    {
        pbat_amd::gpu::vbd::Integrator vbd(data);
        vbd.Step(1e-2f, 10, 1);
        auto const x = vbd.GetPositions();
        for (int i = 0; i < 8; ++i)
        {
            EXPECT(x(2, i) - P[2][i] < 0.f);
            EXPECT(std::fabs(x(0, i) - P[0][i]) < 1e-4f && std::fabs(x(1, i) - P[1][i]) < 1e-4f);
        }
        Matrix<float> bad(3, 7);
        threw = false;
        try
        {
            vbd.SetPositions(bad);
        }
        catch (std::invalid_argument const&)
        {
            threw = true;
        }
        EXPECT(threw);
    }
    std::printf("%s: %d failure(s)\n", gpu ? "wrapper+gpu" : "wrapper", fails);
    return fails == 0 ? 0 : 1;
}
