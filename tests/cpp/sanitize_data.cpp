// Drives the library's host precompute (pbat_vbd_data_construct in
// csrc/vbd_data.cpp: FP64 element geometry, lumped masses, vertex-element CSR,
// greedy colouring, Dirichlet removal) on a jittered Kuhn grid, for a build with
// AddressSanitizer and UndefinedBehaviorSanitizer that aborts on any report: every
// colouring strategy, with and without per-element materials, and the error paths
// (inverted element, index out of range, null output). Exits non-zero on a wrong
// status code.
#include "pbat_vbd.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace
{

struct Mesh
{
    int64_t nV = 0, nT = 0;
    std::vector<double> X; // nV x 3, row-major
    std::vector<int64_t> E; // nT x 4, row-major
};

double Volume(Mesh const& m, int64_t const* t)
{
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d)
    {
        a[d] = m.X[3 * t[1] + d] - m.X[3 * t[0] + d];
        b[d] = m.X[3 * t[2] + d] - m.X[3 * t[0] + d];
        c[d] = m.X[3 * t[3] + d] - m.X[3 * t[0] + d];
    }
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

Mesh KuhnGrid(int n)
{
    Mesh m;
    int const k1 = n + 1;
    m.nV = int64_t{k1} * k1 * k1;
    m.nT = 6 * int64_t{n} * n * n;
    m.X.resize(3 * m.nV);
    uint64_t s = 12345;
    auto uniform = [&s] {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<double>(s >> 11) / 4503599627370496.0 - 1.0;
    };
    for (int k = 0; k < k1; ++k)
        for (int j = 0; j < k1; ++j)
            for (int i = 0; i < k1; ++i)
            {
                int64_t const v = (int64_t{k} * k1 + j) * k1 + i;
                bool const inner = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                m.X[3 * v + 0] = (i + (inner ? 0.1 * uniform() : 0.)) / n;
                m.X[3 * v + 1] = (j + (inner ? 0.1 * uniform() : 0.)) / n;
                m.X[3 * v + 2] = (k + (inner ? 0.1 * uniform() : 0.)) / n;
            }
    int const perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (auto const& p : perm)
                {
                    int c[3] = {0, 0, 0};
                    int64_t t[4];
                    for (int q = 0; q < 4; ++q)
                    {
                        if (q > 0)
                            c[p[q - 1]] = 1;
                        t[q] = (int64_t{k + c[2]} * k1 + (j + c[1])) * k1 + (i + c[0]);
                    }
                    if (Volume(m, t) < 0.)
                        std::swap(t[2], t[3]);
                    m.E.insert(m.E.end(), t, t + 4);
                }
    return m;
}

struct Out
{
    std::vector<double> m, GP, wg, lame;
    std::vector<int64_t> GVGp, GVGe, GVGil, colors, Pptr, Padj;
    int64_t nPartitions = 0, nPadj = 0;
    explicit Out(Mesh const& mesh)
        : m(mesh.nV), GP(12 * mesh.nT), wg(mesh.nT), lame(2 * mesh.nT), GVGp(mesh.nV + 1), GVGe(4 * mesh.nT),
          GVGil(4 * mesh.nT), colors(mesh.nV), Pptr(mesh.nV + 1), Padj(mesh.nV)
    {
    }
};

int Construct(Mesh const& mesh, Out& o, double const* rho, double const* mu, double const* la,
              std::vector<int64_t> const& dbc, int ordering, int selection)
{
    return pbat_vbd_data_construct(mesh.nV, mesh.nT, mesh.X.data(), mesh.E.data(), rho, mu, la,
                                   dbc.empty() ? nullptr : dbc.data(), static_cast<int64_t>(dbc.size()), ordering,
                                   selection, o.m.data(), o.GP.data(), o.wg.data(), o.lame.data(), o.GVGp.data(),
                                   o.GVGe.data(), o.GVGil.data(), o.colors.data(), o.Pptr.data(), o.Padj.data(),
                                   &o.nPartitions, &o.nPadj);
}

int failures = 0;

void Expect(bool ok, char const* what)
{
    if (!ok)
    {
        std::fprintf(stderr, "FAILED: %s (%s)\n", what, pbat_vbd_last_error());
        ++failures;
    }
}

} // namespace

int main()
{
    Mesh const mesh = KuhnGrid(4);
    std::vector<double> rho(mesh.nT, 1000.), mu(mesh.nT, 3.4e5), la(mesh.nT, 3.1e6);
    std::vector<int64_t> dbc;
    for (int64_t v = 24; v >= 0; --v) // bottom face, unsorted
        dbc.push_back(v);
    double checksum = 0.;
    for (int ordering = 0; ordering <= 2; ++ordering)
        for (int selection = 0; selection <= 1; ++selection)
        {
            Out o(mesh);
            bool const materials = (ordering + selection) % 2 == 0;
            int const rc = Construct(mesh, o, materials ? rho.data() : nullptr, materials ? mu.data() : nullptr,
                                     materials ? la.data() : nullptr, dbc, ordering, selection);
            Expect(rc == PBAT_VBD_OK, "construct");
            Expect(o.nPartitions > 0 && o.nPadj == mesh.nV - static_cast<int64_t>(dbc.size()), "partition sizes");
            for (double x : o.m)
                checksum += x;
            checksum += o.GP[12 * mesh.nT - 1] + o.wg[mesh.nT - 1] + o.lame[0];
        }
    {
        Mesh bad = mesh;
        std::swap(bad.E[2], bad.E[3]); // first element inverted
        Out o(bad);
        Expect(Construct(bad, o, nullptr, nullptr, nullptr, {}, 2, 0) == PBAT_VBD_ERROR_SINGULAR_ELEMENT,
               "inverted element is rejected");
    }
    {
        Mesh bad = mesh;
        bad.E[5] = bad.nV; // vertex index out of range
        Out o(bad);
        Expect(Construct(bad, o, nullptr, nullptr, nullptr, {}, 2, 0) == PBAT_VBD_ERROR_INVALID_ARGUMENT,
               "out-of-range vertex index is rejected");
    }
    {
        Out o(mesh);
        std::vector<int64_t> const badDbc{mesh.nV};
        Expect(Construct(mesh, o, nullptr, nullptr, nullptr, badDbc, 2, 0) == PBAT_VBD_ERROR_INVALID_ARGUMENT,
               "out-of-range Dirichlet vertex is rejected");
        Expect(pbat_vbd_data_construct(mesh.nV, mesh.nT, mesh.X.data(), mesh.E.data(), nullptr, nullptr, nullptr,
                                       nullptr, 0, 2, 0, nullptr, o.GP.data(), o.wg.data(), o.lame.data(),
                                       o.GVGp.data(), o.GVGe.data(), o.GVGil.data(), o.colors.data(), o.Pptr.data(),
                                       o.Padj.data(), &o.nPartitions, &o.nPadj) == PBAT_VBD_ERROR_INVALID_ARGUMENT,
               "null output is rejected");
        Expect(Construct(mesh, o, nullptr, nullptr, nullptr, {}, 3, 0) == PBAT_VBD_ERROR_INVALID_ARGUMENT,
               "invalid colouring strategy is rejected");
    }
    std::printf("host precompute sanitizer run: %lld vertices, %lld tets, checksum %.6e, %d failures\n",
                static_cast<long long>(mesh.nV), static_cast<long long>(mesh.nT), checksum, failures);
    return failures == 0 && std::isfinite(checksum) ? 0 : 1;
}
