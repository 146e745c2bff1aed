// Host-side VBD precompute: the element, adjacency and partition data of
// pbat::sim::vbd::Data::Construct (sim/vbd/Data.cpp:179-244).
//
// FP64 throughout, like the reference. The element quantities use the closed
// form of the affine Tet1 element: with Dm = [x1-x0, x2-x0, x3-x0],
//   GP = GradN * Dm^{-1}     (fem/ShapeFunctions.h:266-299)
//   wg = 0.166667 * det Dm   (fem/MeshQuadrature.h:75-88, order-1 rule)
// and the lumped mass sums the order-2 consistent mass rows with the
// reference's truncated quadrature literals (fem/Mass.h:103-112, 838-874;
// math/SymmetricQuadratureRules.h:1719-1752).
#include "pbat_vbd.h"
#include "status.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace pbat_vbd {

namespace {

thread_local std::string gLastError;

struct Mat3
{
    // column-major a[3*c + r]
    std::array<double, 9> a{};
    double& operator()(int r, int c) { return a[3 * c + r]; }
    double operator()(int r, int c) const { return a[3 * c + r]; }
};

double Det(Mat3 const& A)
{
    return A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
           A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
}

Mat3 EdgeMatrix(double const* X, int64_t const* Ee)
{
    Mat3 D;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            D(r, c) = X[3 * Ee[c + 1] + r] - X[3 * Ee[0] + r];
    return D;
}

Mat3 Inverse(Mat3 const& A, double det)
{
    Mat3 B;
    double const s = 1.0 / det;
    // adjugate / det
    B(0, 0) = s * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1));
    B(0, 1) = s * (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2));
    B(0, 2) = s * (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1));
    B(1, 0) = s * (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2));
    B(1, 1) = s * (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0));
    B(1, 2) = s * (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2));
    B(2, 0) = s * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    B(2, 1) = s * (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1));
    B(2, 2) = s * (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0));
    return B;
}

// Vertex adjacency of the primal graph G G^T (every vertex is its own
// neighbour, graph/Mesh.h:116-123), as CSR with a per-vertex stamp for
// de-duplication.
void PrimalGraph(int64_t nV, int64_t const* E, std::vector<int64_t> const& vtp, std::vector<int64_t> const& vte,
                 std::vector<int64_t>& ptr, std::vector<int64_t>& adj)
{
    std::vector<int64_t> stamp(static_cast<size_t>(nV), -1);
    ptr.assign(static_cast<size_t>(nV) + 1, 0);
    for (int pass = 0; pass < 2; ++pass)
    {
        std::fill(stamp.begin(), stamp.end(), -1);
        for (int64_t u = 0; u < nV; ++u)
        {
            int64_t n = 0;
            int64_t const base = pass ? ptr[u] : 0;
            for (int64_t k = vtp[u]; k < vtp[u + 1]; ++k)
            {
                int64_t const* Ee = E + 4 * vte[k];
                for (int j = 0; j < 4; ++j)
                {
                    int64_t const w = Ee[j];
                    if (stamp[w] != u)
                    {
                        stamp[w] = u;
                        if (pass)
                            adj[base + n] = w;
                        ++n;
                    }
                }
            }
            if (!pass)
                ptr[u + 1] = n;
        }
        if (!pass)
        {
            std::partial_sum(ptr.begin(), ptr.end(), ptr.begin());
            adj.assign(static_cast<size_t>(ptr[nV]), 0);
        }
    }
}

// graph/Color.h:40-135: greedy colouring with a palette of at most 128 colours.
int64_t GreedyColor(int64_t nV, std::vector<int64_t> const& ptr, std::vector<int64_t> const& adj, int ordering,
                    int selection, int64_t* C)
{
    std::vector<int64_t> order(static_cast<size_t>(nV));
    std::iota(order.begin(), order.end(), int64_t{0});
    auto degree = [&](int64_t u) { return ptr[u + 1] - ptr[u]; };
    if (ordering == PBAT_VBD_ORDER_LARGEST_DEGREE)
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return degree(a) > degree(b); });
    else if (ordering == PBAT_VBD_ORDER_SMALLEST_DEGREE)
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return degree(a) < degree(b); });
    std::fill(C, C + nV, int64_t{-1});
    std::array<int64_t, PBAT_VBD_MAX_COLORS> used{};
    std::array<bool, PBAT_VBD_MAX_COLORS> usable{};
    int64_t nColors = 0;
    for (int64_t u : order)
    {
        std::fill(usable.begin(), usable.begin() + nColors, true);
        int64_t blocked = 0;
        for (int64_t k = ptr[u]; k < ptr[u + 1]; ++k)
        {
            int64_t const c = C[adj[k]];
            if (c >= 0 && usable[c])
            {
                usable[c] = false;
                ++blocked;
            }
        }
        if (blocked == nColors)
        {
            if (nColors == PBAT_VBD_MAX_COLORS)
                Fail(PBAT_VBD_ERROR_PALETTE, "vertex colouring needs more than 128 colours");
            C[u] = nColors;
            used[nColors++] = 1;
            continue;
        }
        // first colour that no other colour precedes under the selection order
        int64_t best = 0;
        for (int64_t c = 1; c < nColors; ++c)
        {
            bool precedes;
            if (usable[c] && usable[best])
                precedes = (selection == PBAT_VBD_SELECT_LEAST_USED) ? used[c] < used[best] : c < best;
            else
                precedes = usable[c];
            if (precedes)
                best = c;
        }
        C[u] = best;
        ++used[best];
    }
    return nColors;
}

} // namespace

void SetLastError(std::string const& what) { gLastError = what; }
void ClearLastError() { gLastError.clear(); }
char const* LastError() { return gLastError.c_str(); }

} // namespace pbat_vbd

extern "C" const char* pbat_vbd_last_error(void) { return pbat_vbd::LastError(); }

extern "C" const char* pbat_vbd_version(void) { return "pbat_vbd 0.1.0 (gfx950)"; }

extern "C" int pbat_vbd_data_construct(int64_t nV, int64_t nT, const double* X, const int64_t* E, const double* rhoe,
                                       const double* mue, const double* lambdae, const int64_t* dbc, int64_t nDbc,
                                       int ordering, int selection, double* m, double* GP, double* wg, double* lame,
                                       int64_t* GVGp, int64_t* GVGe, int64_t* GVGilocal, int64_t* colors,
                                       int64_t* Pptr, int64_t* Padj, int64_t* nPartitions, int64_t* nPadj)
{
    using namespace pbat_vbd;
    return Guard([&] {
        if (nV <= 0 || nT <= 0 || !X || !E)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "empty mesh: expected nV > 0, nT > 0 and non-null X, E");
        if (!m || !GP || !wg || !lame || !GVGp || !GVGe || !GVGilocal || !colors || !Pptr || !Padj || !nPartitions ||
            !nPadj)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null output array");
        if (nDbc < 0 || (nDbc > 0 && !dbc))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid Dirichlet vertex array");
        if (ordering < 0 || ordering > 2 || selection < 0 || selection > 1)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid colouring strategy");
        for (int64_t k = 0; k < 4 * nT; ++k)
            if (E[k] < 0 || E[k] >= nV)
                Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                     "element vertex index " + std::to_string(E[k]) + " out of range [0," + std::to_string(nV) +
                         ")");
        for (int64_t k = 0; k < nDbc; ++k)
            if (dbc[k] < 0 || dbc[k] >= nV)
                Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Dirichlet vertex index out of range");

        // Material defaults (sim/vbd/Data.cpp:199-209): LameCoefficients(1e6, 0.45), rho = 1e3
        double const Yd = 1e6, nud = 0.45;
        double const muDefault = Yd / (2. * (1. + nud));
        double const lambdaDefault = (Yd * nud) / ((1. + nud) * (1. - 2. * nud));
        for (int64_t e = 0; e < nT; ++e)
        {
            lame[2 * e] = mue ? mue[e] : muDefault;
            lame[2 * e + 1] = lambdae ? lambdae[e] : lambdaDefault;
        }

        // Element geometry
        static constexpr double kRule1Weight = 0.166667;
        static constexpr double kRule2Weight = 0.0416667;
        static constexpr double kRule2Points[4][4] = {{0.58541, 0.138197, 0.138197, 0.138197},
                                                      {0.138197, 0.58541, 0.138197, 0.138197},
                                                      {0.138197, 0.138197, 0.58541, 0.138197},
                                                      {0.138197, 0.138197, 0.138197, 0.58541}};
        // Tet1 shape functions at the order-2 points (reference coordinates = last 3 barycentrics)
        double N[4][4];
        for (int g = 0; g < 4; ++g)
        {
            double const* xi = &kRule2Points[g][1];
            N[g][0] = -xi[0] - xi[1] - xi[2] + 1.;
            N[g][1] = xi[0];
            N[g][2] = xi[1];
            N[g][3] = xi[2];
        }
        std::fill(m, m + nV, 0.);
        for (int64_t e = 0; e < nT; ++e)
        {
            int64_t const* Ee = E + 4 * e;
            Mat3 const D = EdgeMatrix(X, Ee);
            double const det = Det(D);
            if (!(det > 1e-10))
                Fail(PBAT_VBD_ERROR_SINGULAR_ELEMENT,
                     "Inverted or singular jacobian (element " + std::to_string(e) + ")");
            wg[e] = det * kRule1Weight;
            Mat3 const Di = Inverse(D, det);
            double* G = GP + 12 * e; // G(r, c) = G[4c + r]
            for (int c = 0; c < 3; ++c)
            {
                double s = 0.;
                for (int r = 1; r < 4; ++r)
                {
                    G[4 * c + r] = Di(r - 1, c);
                    s += Di(r - 1, c);
                }
                G[4 * c] = -s;
            }
            double const rho = rhoe ? rhoe[e] : 1e3;
            double const w = det * kRule2Weight;
            for (int g = 0; g < 4; ++g)
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j)
                        m[Ee[i]] += w * rho * N[g][i] * N[g][j];
        }

        // Vertex -> (element, local index) CSR, elements ascending per vertex
        std::fill(GVGp, GVGp + nV + 1, int64_t{0});
        for (int64_t k = 0; k < 4 * nT; ++k)
            ++GVGp[E[k] + 1];
        std::partial_sum(GVGp, GVGp + nV + 1, GVGp);
        {
            std::vector<int64_t> cursor(GVGp, GVGp + nV);
            for (int64_t e = 0; e < nT; ++e)
                for (int k = 0; k < 4; ++k)
                {
                    int64_t const i = E[4 * e + k];
                    GVGe[cursor[i]] = e;
                    GVGilocal[cursor[i]] = k;
                    ++cursor[i];
                }
        }

        // Colouring of the primal graph and partitions (sim/vbd/Data.cpp:228-231)
        std::vector<int64_t> vtp(GVGp, GVGp + nV + 1), vte(GVGe, GVGe + 4 * nT);
        std::vector<int64_t> gptr, gadj;
        PrimalGraph(nV, E, vtp, vte, gptr, gadj);
        int64_t const nColors = GreedyColor(nV, gptr, gadj, ordering, selection, colors);

        // MapToAdjacency (graph/Adjacency.h:172-181): vertices ascending within a colour
        std::fill(Pptr, Pptr + nColors + 1, int64_t{0});
        for (int64_t i = 0; i < nV; ++i)
            ++Pptr[colors[i] + 1];
        std::partial_sum(Pptr, Pptr + nColors + 1, Pptr);
        {
            std::vector<int64_t> cursor(Pptr, Pptr + nColors);
            for (int64_t i = 0; i < nV; ++i)
                Padj[cursor[colors[i]]++] = i;
        }
        // Dirichlet vertices leave the partitions (sim/vbd/Data.cpp:236-244, graph/Adjacency.h:307-342)
        std::vector<char> fixed(static_cast<size_t>(nV), 0);
        for (int64_t k = 0; k < nDbc; ++k)
            fixed[dbc[k]] = 1;
        int64_t w = 0, begin = 0;
        for (int64_t p = 0; p < nColors; ++p)
        {
            int64_t const end = Pptr[p + 1];
            for (int64_t k = begin; k < end; ++k)
                if (!fixed[Padj[k]])
                    Padj[w++] = Padj[k];
            begin = end;
            Pptr[p + 1] = w;
        }
        *nPartitions = nColors;
        *nPadj = w;
    });
}
