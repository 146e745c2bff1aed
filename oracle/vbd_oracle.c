/*
 * vbd_oracle.c -- FP64 CPU restatement of PBAT's Vertex Block Descent path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the MI355X product (see
 * vbd_oracle.h). Never linked into, or called by, the product library.
 *
 * Every routine cites the reference file:line it restates. The arithmetic is
 * written from the formulas those routines implement (the reference's SNH
 * derivatives are machine-generated expressions; here they are written from
 * the closed forms dPsi/dF = mu F + lambda (J - alpha) cof(F) and
 * d2Psi/dF2 = mu I + lambda vec(cof F) vec(cof F)^T + lambda (J - alpha) d2J/dF2).
 */
#define _POSIX_C_SOURCE 200809L
#include "vbd_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ===================================================================== */
/* small dense helpers                                                    */
/* ===================================================================== */

static inline void cross3(const double* a, const double* b, double* out)
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

/* Determinant of a 3x3 column-major matrix by cofactor expansion along the
 * first row (math/linalg/mini/Determinant.h:31-36). */
static inline double det3(const double* A)
{
#define A_(r, c) A[3 * (c) + (r)]
    return A_(0, 0) * (A_(1, 1) * A_(2, 2) - A_(2, 1) * A_(1, 2)) -
           A_(0, 1) * (A_(1, 0) * A_(2, 2) - A_(2, 0) * A_(1, 2)) +
           A_(0, 2) * (A_(1, 0) * A_(2, 1) - A_(2, 0) * A_(1, 1));
#undef A_
}

/* Inverse of a 3x3 column-major matrix through its adjugate
 * (math/linalg/mini/Inverse.h:21-63; textbook formula). */
static inline void inv3(const double* A, double* Ainv)
{
#define A_(r, c) A[3 * (c) + (r)]
#define I_(r, c) Ainv[3 * (c) + (r)]
    double const d = 1.0 / det3(A);
    I_(0, 0) = d * (A_(1, 1) * A_(2, 2) - A_(2, 1) * A_(1, 2));
    I_(0, 1) = d * (A_(2, 1) * A_(0, 2) - A_(0, 1) * A_(2, 2));
    I_(0, 2) = d * (A_(0, 1) * A_(1, 2) - A_(1, 1) * A_(0, 2));
    I_(1, 0) = d * (A_(2, 0) * A_(1, 2) - A_(1, 0) * A_(2, 2));
    I_(1, 1) = d * (A_(0, 0) * A_(2, 2) - A_(2, 0) * A_(0, 2));
    I_(1, 2) = d * (A_(1, 0) * A_(0, 2) - A_(0, 0) * A_(1, 2));
    I_(2, 0) = d * (A_(1, 0) * A_(2, 1) - A_(2, 0) * A_(1, 1));
    I_(2, 1) = d * (A_(2, 0) * A_(0, 1) - A_(0, 0) * A_(2, 1));
    I_(2, 2) = d * (A_(0, 0) * A_(1, 1) - A_(1, 0) * A_(0, 1));
#undef A_
#undef I_
}

/* ===================================================================== */
/* Stable Neo-Hookean                                                     */
/* ===================================================================== */

void ora_lame(double Y, double nu, double* mu, double* lambda)
{
    /* physics/HyperElasticity.cpp:6-11 */
    *mu = Y / (2. * (1. + nu));
    *lambda = (Y * nu) / ((1. + nu) * (1. - 2. * nu));
}

double ora_snh_eval(const double F[9], double mu, double lambda)
{
    /* Psi = mu/2 (|F|^2 - 3) + lambda/2 (det F - 1 - mu/lambda)^2
     * (physics/StableNeoHookeanEnergy.h:737-755) */
    double const J = det3(F);
    double const r = J - 1. - mu / lambda;
    double I2 = 0.;
    for (int k = 0; k < 9; ++k)
        I2 += F[k] * F[k];
    return 0.5 * lambda * (r * r) + 0.5 * mu * (I2 - 3.);
}

/* columns of cof(F) = dJ/dF: [f1 x f2, f2 x f0, f0 x f1] */
static inline void cofactor(const double F[9], double dJ[9])
{
    cross3(F + 3, F + 6, dJ + 0);
    cross3(F + 6, F + 0, dJ + 3);
    cross3(F + 0, F + 3, dJ + 6);
}

void ora_snh_grad(const double F[9], double mu, double lambda, double gF[9])
{
    /* physics/StableNeoHookeanEnergy.h:766-795 */
    double dJ[9];
    cofactor(F, dJ);
    double const J = F[0] * dJ[0] + F[1] * dJ[1] + F[2] * dJ[2];
    double const s = lambda * (J - 1. - mu / lambda);
    for (int k = 0; k < 9; ++k)
        gF[k] = mu * F[k] + s * dJ[k];
}

/* hat(a) b = a x b, written into the 3x3 block (rows 3*br.., cols 3*bc..) of a
 * 9x9 column-major matrix with the given sign. */
static inline void add_hat_block(double* H, int br, int bc, const double* a, double sgn)
{
#define H_(r, c) H[9 * (3 * bc + (c)) + 3 * br + (r)]
    H_(0, 1) += -sgn * a[2];
    H_(0, 2) += sgn * a[1];
    H_(1, 0) += sgn * a[2];
    H_(1, 2) += -sgn * a[0];
    H_(2, 0) += -sgn * a[1];
    H_(2, 1) += sgn * a[0];
#undef H_
}

void ora_snh_hessian(const double F[9], double mu, double lambda, double HF[81])
{
    /* physics/StableNeoHookeanEnergy.h:1164-1326 (exact Hessian, no projection) */
    double dJ[9];
    cofactor(F, dJ);
    double const J = F[0] * dJ[0] + F[1] * dJ[1] + F[2] * dJ[2];
    double const s = lambda * (J - 1. - mu / lambda);
    for (int c = 0; c < 9; ++c)
        for (int r = 0; r < 9; ++r)
            HF[9 * c + r] = lambda * dJ[r] * dJ[c] + (r == c ? mu : 0.);
    /* d2J/df_a df_b: (0,1) = -hat(f2), (0,2) = hat(f1), (1,0) = hat(f2),
     * (1,2) = -hat(f0), (2,0) = -hat(f1), (2,1) = hat(f0) */
    add_hat_block(HF, 0, 1, F + 6, -s);
    add_hat_block(HF, 0, 2, F + 3, s);
    add_hat_block(HF, 1, 0, F + 6, s);
    add_hat_block(HF, 1, 2, F + 0, -s);
    add_hat_block(HF, 2, 0, F + 3, -s);
    add_hat_block(HF, 2, 1, F + 0, s);
}

/* ===================================================================== */
/* FEM precompute                                                         */
/* ===================================================================== */

/* Jacobian of the affine Tet1 map: J = [x1-x0, x2-x0, x3-x0] (X * GradN,
 * fem/Tetrahedron.h:73-95, fem/Jacobian.h). Column-major. */
static inline void tet_jacobian(const double* X, const int64_t* Ee, double J[9])
{
    const double* x0 = X + 3 * Ee[0];
    for (int c = 0; c < 3; ++c)
    {
        const double* xc = X + 3 * Ee[c + 1];
        for (int r = 0; r < 3; ++r)
            J[3 * c + r] = xc[r] - x0[r];
    }
}

int ora_shape_function_gradients(int64_t nV, int64_t nT, const double* X, const int64_t* E, double* GP)
{
    (void)nV;
    /* GP = GradN * J^{-1} (fem/ShapeFunctions.h:266-299, order-1 quadrature point,
     * constant Jacobian). Rows 1..3 are the rows of J^{-1}; row 0 = -(sum of rows). */
    for (int64_t e = 0; e < nT; ++e)
    {
        double J[9], Ji[9];
        tet_jacobian(X, E + 4 * e, J);
        if (!(det3(J) > 1e-10))
            return ORA_ERR_SINGULAR_JACOBIAN;
        inv3(J, Ji);
        double* G = GP + 12 * e;
        for (int c = 0; c < 3; ++c)
        {
            double s = 0.;
            for (int r = 1; r < 4; ++r)
            {
                G[4 * c + r] = Ji[3 * c + (r - 1)];
                s += G[4 * c + r];
            }
            G[4 * c + 0] = -s;
        }
    }
    return ORA_OK;
}

/* Truncated literal quadrature constants of the reference
 * (math/SymmetricQuadratureRules.h:1719-1752). */
static const double kTetRule1Weight = 0.166667;
static const double kTetRule2Weight = 0.0416667;
static const double kTetRule2Points[16] = {0.58541,  0.138197, 0.138197, 0.138197, 0.138197, 0.58541,
                                           0.138197, 0.138197, 0.138197, 0.138197, 0.58541,  0.138197,
                                           0.138197, 0.138197, 0.138197, 0.58541};

int ora_quadrature_weights(int64_t nV, int64_t nT, const double* X, const int64_t* E, double* wg)
{
    (void)nV;
    /* w_e = weight * det(J_e), fem/MeshQuadrature.h:75-88; throws on det <= 1e-10 */
    for (int64_t e = 0; e < nT; ++e)
    {
        double J[9];
        tet_jacobian(X, E + 4 * e, J);
        double const d = det3(J);
        if (!(d > 1e-10))
            return ORA_ERR_SINGULAR_JACOBIAN;
        wg[e] = d * kTetRule1Weight;
    }
    return ORA_OK;
}

int ora_lumped_mass(int64_t nV, int64_t nT, const double* X, const int64_t* E, const double* rhoe, double* m)
{
    /* Shape functions at the 4 order-2 quadrature points (fem/ShapeFunctions.h:40-56):
     * xi = last 3 barycentric coordinates; N = (1 - xi0 - xi1 - xi2, xi0, xi1, xi2). */
    double N[4][4];
    for (int g = 0; g < 4; ++g)
    {
        double const* p = kTetRule2Points + 4 * g + 1;
        N[g][0] = -p[0] - p[1] - p[2] + 1.;
        N[g][1] = p[0];
        N[g][2] = p[1];
        N[g][3] = p[2];
    }
    for (int64_t i = 0; i < nV; ++i)
        m[i] = 0.;
    /* m_i += sum_j w_g rho_e N_i N_j for every (e, g), fem/Mass.h:103-112, 838-874 */
    for (int64_t e = 0; e < nT; ++e)
    {
        double J[9];
        tet_jacobian(X, E + 4 * e, J);
        double const d = det3(J);
        if (!(d > 1e-10))
            return ORA_ERR_SINGULAR_JACOBIAN;
        double const w = d * kTetRule2Weight;
        for (int g = 0; g < 4; ++g)
        {
            double const wr = w * rhoe[e];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    m[E[4 * e + i]] += wr * N[g][i] * N[g][j];
        }
    }
    return ORA_OK;
}

/* ===================================================================== */
/* graphs and colouring                                                   */
/* ===================================================================== */

void ora_vertex_tet_adjacency(int64_t nV, int64_t nT, const int64_t* E, int64_t* GVGp, int64_t* GVGe,
                              int64_t* GVGilocal)
{
    /* Transpose of the vertex x element incidence matrix weighted by ilocal:
     * per vertex, incident elements in ascending element order. */
    memset(GVGp, 0, sizeof(int64_t) * (size_t)(nV + 1));
    for (int64_t e = 0; e < nT; ++e)
        for (int k = 0; k < 4; ++k)
            ++GVGp[E[4 * e + k] + 1];
    for (int64_t i = 0; i < nV; ++i)
        GVGp[i + 1] += GVGp[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nV > 0 ? nV : 1));
    memcpy(fill, GVGp, sizeof(int64_t) * (size_t)nV);
    for (int64_t e = 0; e < nT; ++e)
        for (int k = 0; k < 4; ++k)
        {
            int64_t const i = E[4 * e + k];
            GVGe[fill[i]] = e;
            GVGilocal[fill[i]] = k;
            ++fill[i];
        }
    free(fill);
}

static int cmp_i64(const void* a, const void* b)
{
    int64_t const x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Vertex-vertex adjacency of G*G^T (self loops included), sorted per vertex. */
static int primal_graph(int64_t nV, int64_t nT, const int64_t* E, int64_t** pptr, int64_t** padj)
{
    int64_t* GVGp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nV + 1));
    int64_t* GVGe = (int64_t*)malloc(sizeof(int64_t) * (size_t)(4 * nT + 1));
    int64_t* GVGl = (int64_t*)malloc(sizeof(int64_t) * (size_t)(4 * nT + 1));
    int64_t* ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nV + 1));
    int64_t* adj = (int64_t*)malloc(sizeof(int64_t) * (size_t)(16 * nT + 1));
    if (!GVGp || !GVGe || !GVGl || !ptr || !adj)
    {
        free(GVGp);
        free(GVGe);
        free(GVGl);
        free(ptr);
        free(adj);
        return ORA_ERR_ALLOC;
    }
    ora_vertex_tet_adjacency(nV, nT, E, GVGp, GVGe, GVGl);
    int64_t n = 0;
    ptr[0] = 0;
    for (int64_t u = 0; u < nV; ++u)
    {
        int64_t const b = n;
        for (int64_t k = GVGp[u]; k < GVGp[u + 1]; ++k)
            for (int j = 0; j < 4; ++j)
                adj[n++] = E[4 * GVGe[k] + j];
        qsort(adj + b, (size_t)(n - b), sizeof(int64_t), cmp_i64);
        int64_t w = b;
        for (int64_t k = b; k < n; ++k)
            if (k == b || adj[k] != adj[w - 1])
                adj[w++] = adj[k];
        n = w;
        ptr[u + 1] = n;
    }
    free(GVGp);
    free(GVGe);
    free(GVGl);
    *pptr = ptr;
    *padj = adj;
    return ORA_OK;
}

/* Stable sort of vertex indices by degree (counting sort keeps index order
 * among equal degrees, as std::stable_sort does in common/ArgSort.h:31-39). */
static void order_by_degree(int64_t n, const int64_t* ptr, int descending, int64_t* order)
{
    int64_t maxd = 0;
    for (int64_t u = 0; u < n; ++u)
        if (ptr[u + 1] - ptr[u] > maxd)
            maxd = ptr[u + 1] - ptr[u];
    int64_t* cnt = (int64_t*)calloc((size_t)(maxd + 2), sizeof(int64_t));
    for (int64_t u = 0; u < n; ++u)
    {
        int64_t d = ptr[u + 1] - ptr[u];
        int64_t key = descending ? (maxd - d) : d;
        ++cnt[key + 1];
    }
    for (int64_t k = 0; k <= maxd; ++k)
        cnt[k + 1] += cnt[k];
    for (int64_t u = 0; u < n; ++u)
    {
        int64_t d = ptr[u + 1] - ptr[u];
        int64_t key = descending ? (maxd - d) : d;
        order[cnt[key]++] = u;
    }
    free(cnt);
}

#define ORA_PALETTE 128

int64_t ora_greedy_color(int64_t nV, int64_t nT, const int64_t* E, int ordering, int selection, int64_t* C)
{
    int64_t *ptr = NULL, *adj = NULL;
    int rc = primal_graph(nV, nT, E, &ptr, &adj);
    if (rc != ORA_OK)
        return rc;
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nV > 0 ? nV : 1));
    /* visiting order, graph/Color.h:59-78 */
    if (ordering == ORA_ORDER_LARGEST_DEGREE)
        order_by_degree(nV, ptr, 1, order);
    else if (ordering == ORA_ORDER_SMALLEST_DEGREE)
        order_by_degree(nV, ptr, 0, order);
    else
        for (int64_t u = 0; u < nV; ++u)
            order[u] = u;
    for (int64_t u = 0; u < nV; ++u)
        C[u] = -1;
    int64_t palette[ORA_PALETTE];
    int usable[ORA_PALETTE];
    int64_t nPalette = 0;
    int64_t result = 0;
    /* graph/Color.h:80-130 */
    for (int64_t t = 0; t < nV; ++t)
    {
        int64_t const u = order[t];
        for (int64_t c = 0; c < nPalette; ++c)
            usable[c] = 1;
        int64_t nUnusable = 0;
        for (int64_t k = ptr[u]; k < ptr[u + 1]; ++k)
        {
            int64_t const cv = C[adj[k]];
            if (cv >= 0 && usable[cv])
            {
                usable[cv] = 0;
                ++nUnusable;
            }
        }
        if (nUnusable == nPalette)
        {
            if (nPalette == ORA_PALETTE)
            {
                result = ORA_ERR_PALETTE_OVERFLOW;
                break;
            }
            C[u] = nPalette;
            palette[nPalette] = 1;
            ++nPalette;
        }
        else
        {
            /* std::min_element with the selection comparator: keep the first
             * element that no later element compares less than. */
            int64_t best = 0;
            for (int64_t c = 1; c < nPalette; ++c)
            {
                int less;
                if (selection == ORA_SELECT_LEAST_USED)
                    less = (usable[c] && usable[best]) ? (palette[c] < palette[best]) : usable[c];
                else
                    less = (usable[c] && usable[best]) ? (c < best) : usable[c];
                if (less)
                    best = c;
            }
            C[u] = best;
            ++palette[best];
        }
    }
    if (result == 0)
        result = nPalette;
    free(order);
    free(ptr);
    free(adj);
    return result;
}

int64_t ora_map_to_adjacency(int64_t n, const int64_t* colors, int64_t* Pptr, int64_t* Padj)
{
    /* graph/Adjacency.h:172-181: counts -> prefix sums; stable argsort by colour */
    int64_t nc = 0;
    for (int64_t i = 0; i < n; ++i)
        if (colors[i] + 1 > nc)
            nc = colors[i] + 1;
    for (int64_t c = 0; c <= nc; ++c)
        Pptr[c] = 0;
    for (int64_t i = 0; i < n; ++i)
        ++Pptr[colors[i] + 1];
    for (int64_t c = 0; c < nc; ++c)
        Pptr[c + 1] += Pptr[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    for (int64_t c = 0; c < nc; ++c)
        fill[c] = Pptr[c];
    for (int64_t i = 0; i < n; ++i)
        Padj[fill[colors[i]]++] = i;
    free(fill);
    return nc;
}

int64_t ora_remove_vertices(int64_t nPartitions, int64_t* Pptr, int64_t* Padj, int64_t nV, int64_t nRemove,
                            const int64_t* remove)
{
    /* graph/Adjacency.h:307-342 with the Dirichlet predicate of sim/vbd/Data.cpp:239-244 */
    char* del = (char*)calloc((size_t)(nV > 0 ? nV : 1), 1);
    for (int64_t k = 0; k < nRemove; ++k)
        if (remove[k] >= 0 && remove[k] < nV)
            del[remove[k]] = 1;
    int64_t kk = 0;
    int64_t begin = Pptr[0];
    for (int64_t p = 0; p < nPartitions; ++p)
    {
        int64_t const end = Pptr[p + 1];
        for (int64_t k = begin; k < end; ++k)
            if (!del[Padj[k]])
                Padj[kk++] = Padj[k];
        begin = end;
        Pptr[p + 1] = kk;
    }
    free(del);
    return kk;
}

/* ===================================================================== */
/* integrator                                                             */
/* ===================================================================== */

/* kernels::InitialPositionsForSolve (sim/vbd/Kernels.h:45-94) */
static void initial_position(const double* xt, const double* vtm1, const double* vt, const double* a, double dt,
                             double dt2, int strategy, double* x)
{
    if (strategy == ORA_INIT_POSITION)
    {
        for (int d = 0; d < 3; ++d)
            x[d] = xt[d];
        return;
    }
    if (strategy == ORA_INIT_INERTIA)
    {
        for (int d = 0; d < 3; ++d)
            x[d] = xt[d] + dt * vt[d];
        return;
    }
    if (strategy == ORA_INIT_KINETIC_ENERGY_MINIMUM)
    {
        for (int d = 0; d < 3; ++d)
            x[d] = xt[d] + dt * vt[d] + dt2 * a[d];
        return;
    }
    double const an2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double at = 0.;
    if (an2 != 0.)
    {
        if (strategy == ORA_INIT_ADAPTIVE_VBD)
        {
            double dot = 0.;
            for (int d = 0; d < 3; ++d)
                dot += ((vt[d] - vtm1[d]) / dt) * a[d];
            at = dot / an2;
            at = fmin(fmax(at, 0.), 1.);
        }
        if (strategy == ORA_INIT_ADAPTIVE_PBAT)
        {
            double const vn = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]) + DBL_MIN;
            double dot = 0.;
            for (int d = 0; d < 3; ++d)
                dot += (vt[d] / vn) * a[d];
            at = dot / an2;
            at = fmin(fabs(at), 1.);
        }
    }
    for (int d = 0; d < 3; ++d)
        x[d] = xt[d] + dt * vt[d] + dt2 * at * a[d];
}

static inline void element_F(const double* x, const int64_t* Ee, const double* G, double F[9])
{
    /* F = x_e * GP_e (3x4 times 4x3) */
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
        {
            double s = 0.;
            for (int k = 0; k < 4; ++k)
                s += x[3 * Ee[k] + r] * G[4 * c + k];
            F[3 * c + r] = s;
        }
}

void ora_vertex_elastic_block(const ora_vbd* s, int64_t i, double H[9], double g[3])
{
    for (int k = 0; k < 9; ++k)
        H[k] = 0.;
    g[0] = g[1] = g[2] = 0.;
    for (int64_t n = s->GVGp[i]; n < s->GVGp[i + 1]; ++n)
    {
        int64_t const e = s->GVGe[n];
        int64_t const il = s->GVGilocal[n];
        double const mu = s->lame[2 * e], lambda = s->lame[2 * e + 1];
        double const w = s->wg[e];
        const double* G = s->GP + 12 * e;
        double F[9], gF[9], HF[81];
        element_F(s->x, s->E + 4 * e, G, F);
        ora_snh_grad(F, mu, lambda, gF);
        ora_snh_hessian(F, mu, lambda, HF);
        /* kernels::AccumulateElasticHessian (sim/vbd/Kernels.h:131-151) */
        for (int kj = 0; kj < 3; ++kj)
            for (int ki = 0; ki < 3; ++ki)
            {
                double const c = w * G[4 * ki + il] * G[4 * kj + il];
                for (int cc = 0; cc < 3; ++cc)
                    for (int rr = 0; rr < 3; ++rr)
                        H[3 * cc + rr] += c * HF[9 * (3 * kj + cc) + 3 * ki + rr];
            }
        /* kernels::AccumulateElasticGradient (sim/vbd/Kernels.h:153-171) */
        for (int k = 0; k < 3; ++k)
        {
            double const c = w * G[4 * k + il];
            for (int d = 0; d < 3; ++d)
                g[d] += c * gF[3 * k + d];
        }
    }
}

static void sub3(const double* a, const double* b, double* r)
{
    r[0] = a[0] - b[0], r[1] = a[1] - b[1], r[2] = a[2] - b[2];
}

static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

double ora_vertex_triangle_contact(const double xtv[3], const double xv[3], const double xtf[9], const double xf[9],
                                   double dt, double muC, double muF, double epsv, double* g, double* H)
{
    const double *A = xf, *Bv = xf + 3, *C = xf + 6;
    double e0[3], e1[3], n[3];
    sub3(Bv, A, e0);
    sub3(C, A, e1);
    cross3(e0, e1, n);
    double const dbl = sqrt(dot3(n, n));
    if (dbl <= 1e-8) /* degenerate triangle */
        return 0.;
    for (int d = 0; d < 3; ++d)
        n[d] /= dbl;
    /* project xv on the triangle's plane; barycentric coordinates of the projection */
    double ap[3];
    sub3(xv, A, ap);
    double const t = dot3(n, ap);
    for (int d = 0; d < 3; ++d)
        ap[d] -= t * n[d];
    double const d00 = dot3(e0, e0), d01 = dot3(e0, e1), d11 = dot3(e1, e1);
    double const d20 = dot3(ap, e0), d21 = dot3(ap, e1);
    double const den = d00 * d11 - d01 * d01;
    double const bv = (d11 * d20 - d01 * d21) / den, bw = (d00 * d21 - d01 * d20) / den;
    double const bu = 1. - bv - bw;
    if (!(bu >= 0. && bu <= 1. && bv >= 0. && bv <= 1. && bw >= 0. && bw <= 1.))
        return 0.;
    /* normal penalty 1/2 muC d^2 on the penetration depth d = min(0, (xv - xb).n) */
    double xb[3], xtb[3];
    for (int d = 0; d < 3; ++d)
    {
        xb[d] = bu * A[d] + bv * Bv[d] + bw * C[d];
        xtb[d] = bu * xtf[d] + bv * xtf[3 + d] + bw * xtf[6 + d];
    }
    double r[3];
    sub3(xv, xb, r);
    double dpen = dot3(r, n);
    if (dpen > 0.)
        dpen = 0.;
    double const lambda = muC * dpen;
    double E = 0.5 * lambda * dpen;
    if (g)
        for (int d = 0; d < 3; ++d)
            g[d] += lambda * n[d];
    if (H)
        for (int c = 0; c < 3; ++c)
            for (int rr = 0; rr < 3; ++rr)
                H[3 * c + rr] += muC * n[rr] * n[c];
    /* lagged IPC friction in the tangent frame (e0, n x e0) on the relative
     * displacement of the contact point over the substep */
    double b1[3];
    cross3(n, e0, b1);
    double du[3];
    for (int d = 0; d < 3; ++d)
        du[d] = (xv[d] - xtv[d]) - (xb[d] - xtb[d]);
    double const u0 = dot3(e0, du), u1 = dot3(b1, du);
    double const unorm = sqrt(u0 * u0 + u1 * u1) + DBL_EPSILON;
    double const epsvh = epsv * dt;
    double const muFl = muF * fabs(lambda);
    double const y = unorm / epsvh;
    double const f1 = (y < 1.) ? 2. * y - y * y : 1.;
    double const c1 = muFl * f1 / unorm;
    if (g)
        for (int d = 0; d < 3; ++d)
            g[d] += c1 * (e0[d] * u0 + b1[d] * u1);
    if (H)
        for (int c = 0; c < 3; ++c)
            for (int rr = 0; rr < 3; ++rr)
                H[3 * c + rr] += c1 * (e0[rr] * e0[c] + b1[rr] * b1[c]);
    double const u2 = unorm * unorm;
    E += muFl * (epsvh / 3. + u2 / epsvh - u2 * unorm / (3. * epsvh * epsvh));
    return E;
}

double ora_point_triangle_sqdist(const double p[3], const double a[3], const double b[3], const double c[3])
{
    double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    double const d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    double u = 1., v = 0., w = 0.;
    sub3(p, b, bp);
    double const d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    sub3(p, c, cp);
    double const d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    double const vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    if (d1 <= 0. && d2 <= 0.)
        u = 1., v = 0., w = 0.; /* vertex region a */
    else if (d3 >= 0. && d4 <= d3)
        u = 0., v = 1., w = 0.; /* vertex region b */
    else if (vc <= 0. && d1 >= 0. && d3 <= 0.)
    {
        v = d1 / (d1 - d3); /* edge ab */
        u = 1. - v, w = 0.;
    }
    else if (d6 >= 0. && d5 <= d6)
        u = 0., v = 0., w = 1.; /* vertex region c */
    else if (vb <= 0. && d2 >= 0. && d6 <= 0.)
    {
        w = d2 / (d2 - d6); /* edge ac */
        u = 1. - w, v = 0.;
    }
    else if (va <= 0. && (d4 - d3) >= 0. && (d5 - d6) >= 0.)
    {
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6)); /* edge bc */
        u = 0., v = 1. - w;
    }
    else
    {
        double const den = 1. / (va + vb + vc); /* face */
        v = vb * den, w = vc * den, u = 1. - v - w;
    }
    for (int d = 0; d < 3; ++d)
        q[d] = p[d] - (u * a[d] + v * b[d] + w * c[d]);
    return dot3(q, q);
}

void ora_point_triangle_sqdist_many(const double p[3], const double *tri, int64_t n, double *out)
{
    for (int64_t k = 0; k < n; ++k)
        out[k] = ora_point_triangle_sqdist(p, tri + 9 * k, tri + 9 * k + 3, tri + 9 * k + 6);
}

/* ContactPenalty + AccumulateVertexTriangleContact over vertex i's contacts
 * (gpu/impl/vbd/Kernels.cuh:80-114, 204-223): the penalty muC is scaled by the
 * vertex area over the summed areas of its contacting triangles, times each
 * triangle's area. */
static void add_contacts(const ora_vbd* s, int64_t i, double sdt, double* g, double* H)
{
    const int64_t* f = s->fc + 8 * i;
    int nc = 0;
    for (int c = 0; c < 8; ++c)
        nc += f[c] >= 0;
    if (nc == 0)
        return;
    double sumfa = 0.;
    for (int c = 0; c < nc; ++c)
        sumfa += s->FA[f[c]];
    double const kC = s->XVA[i] * s->muC / sumfa;
    for (int c = 0; c < nc; ++c)
    {
        double xf[9], xtf[9];
        for (int k = 0; k < 3; ++k)
        {
            int64_t const j = s->F[3 * f[c] + k];
            for (int d = 0; d < 3; ++d)
            {
                xf[3 * k + d] = s->x[3 * j + d];
                xtf[3 * k + d] = s->xt[3 * j + d];
            }
        }
        ora_vertex_triangle_contact(s->xt + 3 * i, s->x + 3 * i, xtf, xf, sdt, kC * s->FA[f[c]], s->muF, s->epsv, g,
                                    H);
    }
}

/* Integrator::SolveVertex (sim/vbd/Integrator.cpp:98-136); the new position goes to out */
static void solve_vertex(ora_vbd* s, int64_t i, double sdt, double sdt2, double* out)
{
    double H[9], g[3];
    ora_vertex_elastic_block(s, i, H, g);
    double* xi = s->x + 3 * i;
    const double* xti = s->xt + 3 * i;
    const double* xtl = s->xtilde + 3 * i;
    /* AddDamping (sim/vbd/Kernels.h:173-191): g += D H (x - xt); H *= 1 + D */
    double const D = s->kD / sdt;
    double dx[3] = {xi[0] - xti[0], xi[1] - xti[1], xi[2] - xti[2]};
    for (int r = 0; r < 3; ++r)
    {
        double hx = 0.;
        for (int c = 0; c < 3; ++c)
            hx += H[3 * c + r] * dx[c];
        g[r] += D * hx;
    }
    for (int k = 0; k < 9; ++k)
        H[k] *= 1. + D;
    if (s->fc)
        add_contacts(s, i, sdt, g, H);
    /* AddInertiaDerivatives (sim/vbd/Kernels.h:304-322) */
    double const K = s->m[i] / sdt2;
    for (int d = 0; d < 3; ++d)
    {
        H[4 * d] += K;
        g[d] += K * (xi[d] - xtl[d]);
    }
    /* IntegratePositions (sim/vbd/Kernels.h:324-339): absolute determinant test */
    if (fabs(det3(H)) <= s->detHZero)
    {
        for (int d = 0; d < 3; ++d)
            out[d] = xi[d];
        return;
    }
    double Hi[9];
    inv3(H, Hi);
    double step[3];
    for (int r = 0; r < 3; ++r)
        step[r] = Hi[r] * g[0] + Hi[3 + r] * g[1] + Hi[6 + r] * g[2];
    for (int d = 0; d < 3; ++d)
        out[d] = xi[d] - step[d];
}

/* ---- thread pool over the vertices of one colour ----------------------- */

typedef struct worker_ctx {
    ora_vbd* s;
    int tid, nthreads;
    double sdt, sdt2;
    int64_t iterations, substeps;
    pthread_barrier_t* bar;
} worker_ctx;

static void range_split(int64_t b, int64_t e, int tid, int nt, int64_t* ob, int64_t* oe)
{
    int64_t const n = e - b;
    *ob = b + (n * tid) / nt;
    *oe = b + (n * (tid + 1)) / nt;
}

static void substep_begin(ora_vbd* s, int64_t vb, int64_t ve, double sdt, double sdt2)
{
    /* Integrator::Step (sim/vbd/Integrator.cpp:31-35) and InitializeSolve (:54-71) */
    for (int64_t i = vb; i < ve; ++i)
    {
        double vtm1[3];
        for (int d = 0; d < 3; ++d)
        {
            s->xt[3 * i + d] = s->x[3 * i + d];
            vtm1[d] = s->vt[3 * i + d];
            s->vt[3 * i + d] = s->v[3 * i + d];
            s->xtilde[3 * i + d] = s->xt[3 * i + d] + sdt * s->vt[3 * i + d] + sdt2 * s->aext[3 * i + d];
        }
        /* CPU: vtm1 == vt (both are the velocity at t); GPU semantics keep the
         * previous substep's velocity as vtm1. */
        const double* vprev = s->gpu_velocity_semantics ? vtm1 : s->vt + 3 * i;
        initial_position(s->xt + 3 * i, vprev, s->v + 3 * i, s->aext + 3 * i, sdt, sdt2, s->strategy,
                         s->x + 3 * i);
    }
}

static void substep_end(ora_vbd* s, int64_t vb, int64_t ve, double sdt)
{
    /* v = (x - xt) / sdt (sim/vbd/Integrator.cpp:39) */
    for (int64_t i = vb; i < ve; ++i)
        for (int d = 0; d < 3; ++d)
            s->v[3 * i + d] = (s->x[3 * i + d] - s->xt[3 * i + d]) / sdt;
}

double ora_chebyshev_omega(int64_t k, double rho2, double omega)
{
    if (k == 0)
        return 1.;
    if (k == 1)
        return 2. / (2. - rho2);
    return 4. / 4. - rho2 * omega;
}

/* sim/vbd/Kernels.h:104-118 over vertices [vb, ve): x = omega (x - xkm2) + xkm2
 * for k > 1, then xkm2 = xkm1, xkm1 = x. */
static void chebyshev_update(ora_vbd* s, int64_t vb, int64_t ve, int64_t k, double omega)
{
    for (int64_t i = 3 * vb; i < 3 * ve; ++i)
    {
        if (k > 1)
            s->x[i] = omega * (s->x[i] - s->xkm2[i]) + s->xkm2[i];
        s->xkm2[i] = s->xkm1[i];
        s->xkm1[i] = s->x[i];
    }
}

static void* worker(void* arg)
{
    worker_ctx* w = (worker_ctx*)arg;
    ora_vbd* s = w->s;
    int64_t vb, ve;
    range_split(0, s->nV, w->tid, w->nthreads, &vb, &ve);
    for (int64_t st = 0; st < w->substeps; ++st)
    {
        substep_begin(s, vb, ve, w->sdt, w->sdt2);
        pthread_barrier_wait(w->bar);
        double const rho2 = s->rho * s->rho;
        double omega = 0.;
        for (int64_t k = 0; k < w->iterations; ++k)
        {
            for (int64_t p = 0; p < s->nPartitions; ++p)
            {
                int64_t pb, pe;
                range_split(s->Pptr[p], s->Pptr[p + 1], w->tid, w->nthreads, &pb, &pe);
                for (int64_t q = pb; q < pe; ++q)
                    solve_vertex(s, s->Padj[q], w->sdt, w->sdt2, s->x + 3 * s->Padj[q]);
                pthread_barrier_wait(w->bar);
            }
            if (s->chebyshev)
            {
                omega = ora_chebyshev_omega(k, rho2, omega);
                chebyshev_update(s, vb, ve, k, omega);
                pthread_barrier_wait(w->bar);
            }
        }
        substep_end(s, vb, ve, w->sdt);
        pthread_barrier_wait(w->bar);
    }
    return NULL;
}

/* The pieces of one substep, for integrators that act between iterations
 * (sim/vbd/AndersonIntegrator.cpp:24-56): begin (x_t, v_t, x~, initial guess),
 * one VBD iteration over all colours, end (velocity). Single-threaded. */
void ora_substep_begin(ora_vbd* s, double sdt)
{
    substep_begin(s, 0, s->nV, sdt, sdt * sdt);
}

void ora_sweep(ora_vbd* s, double sdt)
{
    for (int64_t p = 0; p < s->nPartitions; ++p)
    {
        for (int64_t q = s->Pptr[p]; q < s->Pptr[p + 1]; ++q)
        {
            int64_t const i = s->Padj[q];
            solve_vertex(s, i, sdt, sdt * sdt, s->fc ? s->xb + 3 * i : s->x + 3 * i);
        }
        if (s->fc)
            for (int64_t q = s->Pptr[p]; q < s->Pptr[p + 1]; ++q)
                for (int d = 0; d < 3; ++d)
                    s->x[3 * s->Padj[q] + d] = s->xb[3 * s->Padj[q] + d];
    }
}

void ora_substep_end(ora_vbd* s, double sdt)
{
    substep_end(s, 0, s->nV, sdt);
}

int ora_step(ora_vbd* s, double dt, int64_t iterations, int64_t substeps)
{
    if (substeps < 1 || iterations < 0)
        return ORA_ERR_ARG;
    if (s->chebyshev && (!s->xkm1 || !s->xkm2))
        return ORA_ERR_ARG;
    if (s->fc) /* contact steps run through the substep pieces */
        return ORA_ERR_ARG;
    double const sdt = dt / (double)substeps;
    double const sdt2 = sdt * sdt;
    int nt = s->nthreads > 0 ? s->nthreads : 1;
    if (nt == 1)
    {
        for (int64_t st = 0; st < substeps; ++st)
        {
            substep_begin(s, 0, s->nV, sdt, sdt2);
            double const rho2 = s->rho * s->rho;
            double omega = 0.;
            for (int64_t k = 0; k < iterations; ++k)
            {
                for (int64_t p = 0; p < s->nPartitions; ++p)
                    for (int64_t q = s->Pptr[p]; q < s->Pptr[p + 1]; ++q)
                        solve_vertex(s, s->Padj[q], sdt, sdt2, s->x + 3 * s->Padj[q]);
                if (s->chebyshev)
                {
                    /* ChebyshevIntegrator::Solve (sim/vbd/ChebyshevIntegrator.cpp:15-32) */
                    omega = ora_chebyshev_omega(k, rho2, omega);
                    chebyshev_update(s, 0, s->nV, k, omega);
                }
            }
            substep_end(s, 0, s->nV, sdt);
        }
        return ORA_OK;
    }
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nt);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    worker_ctx* ctx = (worker_ctx*)malloc(sizeof(worker_ctx) * (size_t)nt);
    for (int t = 0; t < nt; ++t)
    {
        ctx[t].s = s;
        ctx[t].tid = t;
        ctx[t].nthreads = nt;
        ctx[t].sdt = sdt;
        ctx[t].sdt2 = sdt2;
        ctx[t].iterations = iterations;
        ctx[t].substeps = substeps;
        ctx[t].bar = &bar;
    }
    for (int t = 1; t < nt; ++t)
        pthread_create(&th[t], NULL, worker, &ctx[t]);
    worker(&ctx[0]);
    for (int t = 1; t < nt; ++t)
        pthread_join(th[t], NULL);
    pthread_barrier_destroy(&bar);
    free(th);
    free(ctx);
    return ORA_OK;
}

double ora_objective(const ora_vbd* s, const double* x, const double* xtilde, double dt)
{
    /* 1/2 (x - xtilde)^T M (x - xtilde) + dt^2 sum_e w_e Psi(F_e) (sim/vbd/Integrator.cpp:138-166) */
    double Ek = 0.;
    for (int64_t i = 0; i < s->nV; ++i)
        for (int d = 0; d < 3; ++d)
        {
            double const dx = x[3 * i + d] - xtilde[3 * i + d];
            Ek += dx * (s->m[i] * dx);
        }
    Ek *= 0.5;
    double Ep = 0.;
    for (int64_t e = 0; e < s->nT; ++e)
    {
        double F[9];
        element_F(x, s->E + 4 * e, s->GP + 12 * e, F);
        Ep += s->wg[e] * ora_snh_eval(F, s->lame[2 * e], s->lame[2 * e + 1]);
    }
    return Ek + dt * dt * Ep;
}

void ora_objective_gradient(const ora_vbd* s, const double* x, const double* xtilde, double dt, double* grad)
{
    /* M (x - xtilde) + dt^2 sum_e w_e dPsi/dx (sim/vbd/Integrator.cpp:168-200,
     * fem/DeformationGradient.h:126-141 for the dF/dx contraction) */
    double* gEp = (double*)calloc((size_t)(3 * s->nV), sizeof(double));
    for (int64_t e = 0; e < s->nT; ++e)
    {
        double F[9], gF[9];
        const double* G = s->GP + 12 * e;
        element_F(x, s->E + 4 * e, G, F);
        ora_snh_grad(F, s->lame[2 * e], s->lame[2 * e + 1], gF);
        for (int k = 0; k < 4; ++k)
        {
            int64_t const i = s->E[4 * e + k];
            for (int d = 0; d < 3; ++d)
            {
                double acc = 0.;
                for (int c = 0; c < 3; ++c)
                    acc += G[4 * c + k] * gF[3 * c + d];
                gEp[3 * i + d] += s->wg[e] * acc;
            }
        }
    }
    for (int64_t i = 0; i < s->nV; ++i)
        for (int d = 0; d < 3; ++d)
            grad[3 * i + d] = s->m[i] * (x[3 * i + d] - xtilde[3 * i + d]) + (dt * dt) * gEp[3 * i + d];
    free(gEp);
}

void ora_energy(const ora_vbd* s, double out[3])
{
    double K = 0., U = 0., P = 0.;
    for (int64_t i = 0; i < s->nV; ++i)
    {
        double v2 = 0., ax = 0.;
        for (int d = 0; d < 3; ++d)
        {
            v2 += s->v[3 * i + d] * s->v[3 * i + d];
            ax += s->aext[3 * i + d] * s->x[3 * i + d];
        }
        K += 0.5 * s->m[i] * v2;
        P -= s->m[i] * ax;
    }
    for (int64_t e = 0; e < s->nT; ++e)
    {
        double F[9];
        element_F(s->x, s->E + 4 * e, s->GP + 12 * e, F);
        U += s->wg[e] * ora_snh_eval(F, s->lame[2 * e], s->lame[2 * e + 1]);
    }
    out[0] = K;
    out[1] = U;
    out[2] = P;
}

void ora_element_derivatives(const ora_vbd* s, const double* x, double* ge, double* He, double* psi)
{
    for (int64_t e = 0; e < s->nT; ++e)
    {
        const double* G = s->GP + 12 * e;
        double const w = s->wg[e], mu = s->lame[2 * e], lambda = s->lame[2 * e + 1];
        double F[9];
        element_F(x, s->E + 4 * e, G, F);
        if (psi)
            psi[e] = w * ora_snh_eval(F, mu, lambda);
        if (ge)
        {
            double gF[9];
            ora_snh_grad(F, mu, lambda, gF);
            /* dF(r,c)/dx_k[d] = delta_rd GP(k,c) */
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d)
                {
                    double acc = 0.;
                    for (int c = 0; c < 3; ++c)
                        acc += G[4 * c + k] * gF[3 * c + d];
                    ge[12 * e + 3 * k + d] = w * acc;
                }
        }
        if (He)
        {
            double HF[81];
            ora_snh_hessian(F, mu, lambda, HF);
            for (int k = 0; k < 4; ++k)
                for (int a = 0; a < 3; ++a)
                    for (int l = 0; l < 4; ++l)
                        for (int b = 0; b < 3; ++b)
                        {
                            double acc = 0.;
                            for (int c = 0; c < 3; ++c)
                                for (int dd = 0; dd < 3; ++dd)
                                    acc += G[4 * c + k] * G[4 * dd + l] * HF[9 * (3 * dd + b) + 3 * c + a];
                            He[144 * e + 12 * (3 * k + a) + 3 * l + b] = w * acc;
                        }
        }
    }
}
