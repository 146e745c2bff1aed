/* Drives the oracle's C entry points on a small jittered Kuhn grid, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize): an
 * out-of-bounds access, a leak or an undefined operation anywhere in the checker
 * aborts the run with a report. Test infrastructure only; it checks memory and
 * arithmetic safety, not values (tests/test_oracle.py pins those). */
#include "vbd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void* alloc(size_t n)
{
    void* p = calloc(n ? n : 1, 1);
    if (!p)
    {
        fprintf(stderr, "allocation failed\n");
        exit(2);
    }
    return p;
}

static uint64_t lcg = 0x9e3779b97f4a7c15ull;
static double uniform(void) /* in [-1, 1) */
{
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg >> 11) / 4503599627370496.0 - 1.0;
}

static double volume(const double* X, const int64_t* t)
{
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d)
    {
        a[d] = X[3 * t[1] + d] - X[3 * t[0] + d];
        b[d] = X[3 * t[2] + d] - X[3 * t[0] + d];
        c[d] = X[3 * t[3] + d] - X[3 * t[0] + d];
    }
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

/* n^3 cells, 6 tets per cell along the main diagonal, interior vertices jittered by
 * up to 0.1 cell, every tet positively oriented. X is nV x 3, E nT x 4 (row-major). */
static void kuhn_grid(int n, double** Xo, int64_t** Eo, int64_t* nVo, int64_t* nTo)
{
    int const m = n + 1;
    int64_t const nV = (int64_t)m * m * m, nT = 6 * (int64_t)n * n * n;
    double* X = alloc(sizeof(double) * 3 * nV);
    int64_t* E = alloc(sizeof(int64_t) * 4 * nT);
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i)
            {
                int64_t const v = ((int64_t)k * m + j) * m + i;
                int const inner = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                X[3 * v + 0] = (i + (inner ? 0.1 * uniform() : 0.0)) / n;
                X[3 * v + 1] = (j + (inner ? 0.1 * uniform() : 0.0)) / n;
                X[3 * v + 2] = (k + (inner ? 0.1 * uniform() : 0.0)) / n;
            }
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int64_t e = 0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int p = 0; p < 6; ++p)
                {
                    int c[3] = {0, 0, 0};
                    int64_t* t = E + 4 * e;
                    for (int s = 0; s < 4; ++s)
                    {
                        if (s > 0)
                            c[perm[p][s - 1]] = 1;
                        t[s] = ((int64_t)(k + c[2]) * m + (j + c[1])) * m + (i + c[0]);
                    }
                    if (volume(X, t) < 0.0)
                    {
                        int64_t const w = t[2];
                        t[2] = t[3], t[3] = w;
                    }
                    ++e;
                }
    *Xo = X, *Eo = E, *nVo = nV, *nTo = nT;
}

int main(void)
{
    int const n = 3;
    double* X;
    int64_t *E, nV, nT;
    kuhn_grid(n, &X, &E, &nV, &nT);

    /* constitutive model */
    double mu, la, F[9], gF[9], HF[81];
    ora_lame(1e6, 0.45, &mu, &la);
    for (int k = 0; k < 9; ++k)
        F[k] = (k % 4 == 0 ? 1.0 : 0.0) + 0.1 * uniform();
    double const psi = ora_snh_eval(F, mu, la);
    ora_snh_grad(F, mu, la, gF);
    ora_snh_hessian(F, mu, la, HF);

    /* precompute */
    double* GP = alloc(sizeof(double) * 12 * nT);
    double* wg = alloc(sizeof(double) * nT);
    double* rho = alloc(sizeof(double) * nT);
    double* m = alloc(sizeof(double) * nV);
    for (int64_t e = 0; e < nT; ++e)
        rho[e] = 1000.0;
    if (ora_shape_function_gradients(nV, nT, X, E, GP) || ora_quadrature_weights(nV, nT, X, E, wg) ||
        ora_lumped_mass(nV, nT, X, E, rho, m))
    {
        fprintf(stderr, "precompute failed\n");
        return 1;
    }
    int64_t* GVGp = alloc(sizeof(int64_t) * (nV + 1));
    int64_t* GVGe = alloc(sizeof(int64_t) * 4 * nT);
    int64_t* GVGil = alloc(sizeof(int64_t) * 4 * nT);
    ora_vertex_tet_adjacency(nV, nT, E, GVGp, GVGe, GVGil);
    int64_t* colors = alloc(sizeof(int64_t) * nV);
    int64_t const nc = ora_greedy_color(nV, nT, E, ORA_ORDER_LARGEST_DEGREE, ORA_SELECT_LEAST_USED, colors);
    if (nc <= 0)
    {
        fprintf(stderr, "colouring failed\n");
        return 1;
    }
    int64_t* Pptr = alloc(sizeof(int64_t) * (nc + 1));
    int64_t* Padj = alloc(sizeof(int64_t) * nV);
    ora_map_to_adjacency(nV, colors, Pptr, Padj);
    int64_t const nb = (int64_t)(n + 1) * (n + 1); /* bottom face z = 0: the first (n+1)^2 vertices */
    int64_t* dbc = alloc(sizeof(int64_t) * nb);
    for (int64_t k = 0; k < nb; ++k)
        dbc[k] = nb - 1 - k; /* unsorted on purpose */
    int64_t const nPadj = ora_remove_vertices(nc, Pptr, Padj, nV, nb, dbc);

    /* integrator state */
    double* lame = alloc(sizeof(double) * 2 * nT);
    for (int64_t e = 0; e < nT; ++e)
        lame[2 * e] = mu * (1.0 + 0.5 * uniform()), lame[2 * e + 1] = la;
    double* x = alloc(sizeof(double) * 3 * nV);
    double* v = alloc(sizeof(double) * 3 * nV);
    double* aext = alloc(sizeof(double) * 3 * nV);
    double* xt = alloc(sizeof(double) * 3 * nV);
    double* vt = alloc(sizeof(double) * 3 * nV);
    double* xtilde = alloc(sizeof(double) * 3 * nV);
    double* xkm1 = alloc(sizeof(double) * 3 * nV);
    double* xkm2 = alloc(sizeof(double) * 3 * nV);
    double* xb = alloc(sizeof(double) * 3 * nV);
    memcpy(x, X, sizeof(double) * 3 * nV);
    for (int64_t i = nb; i < nV; ++i)
        for (int d = 0; d < 3; ++d)
            v[3 * i + d] = 0.3 * uniform(), aext[3 * i + d] = d == 1 ? -9.81 : 0.0;
    ora_vbd s;
    memset(&s, 0, sizeof(s));
    s.nV = nV, s.nT = nT, s.x = x, s.v = v, s.aext = aext, s.m = m, s.xt = xt, s.vt = vt, s.xtilde = xtilde;
    s.E = E, s.GP = GP, s.wg = wg, s.lame = lame, s.GVGp = GVGp, s.GVGe = GVGe, s.GVGilocal = GVGil;
    s.nPartitions = nc, s.Pptr = Pptr, s.Padj = Padj;
    s.detHZero = 1e-7, s.nthreads = 2, s.xkm1 = xkm1, s.xkm2 = xkm2, s.rho = 0.9;
    double const dt = 1.0 / 60.0;
    for (int strategy = ORA_INIT_POSITION; strategy <= ORA_INIT_ADAPTIVE_PBAT; ++strategy)
    {
        s.strategy = strategy;
        s.gpu_velocity_semantics = strategy % 2;
        s.kD = strategy == ORA_INIT_INERTIA ? 0.01 : 0.0;
        if (ora_step(&s, dt, 5, 2) != ORA_OK)
        {
            fprintf(stderr, "step failed\n");
            return 1;
        }
    }
    s.chebyshev = 1;
    memcpy(xkm1, x, sizeof(double) * 3 * nV);
    memcpy(xkm2, x, sizeof(double) * 3 * nV);
    ora_step(&s, dt, 6, 1);
    s.chebyshev = 0;
    double en[3], H[9], g[3], obj;
    ora_energy(&s, en);
    obj = ora_objective(&s, x, xtilde, dt);
    double* grad = alloc(sizeof(double) * 3 * nV);
    ora_objective_gradient(&s, x, xtilde, dt, grad);
    ora_vertex_elastic_block(&s, nV - 1, H, g);
    double* ge = alloc(sizeof(double) * 12 * nT);
    double* He = alloc(sizeof(double) * 144 * nT);
    double* pe = alloc(sizeof(double) * nT);
    ora_element_derivatives(&s, x, ge, He, pe);
    double const omega = ora_chebyshev_omega(3, 0.81, ora_chebyshev_omega(2, 0.81, 1.0));

    /* contact: the boundary faces of the first tets as triangles, each vertex of the
     * top layer touching up to 8 of them */
    int64_t const nF = 4 * nT;
    int64_t* Ftri = alloc(sizeof(int64_t) * 3 * nF);
    double* FA = alloc(sizeof(double) * nF);
    double* XVA = alloc(sizeof(double) * nV);
    int64_t* fc = alloc(sizeof(int64_t) * 8 * nV);
    for (int64_t e = 0; e < nT; ++e)
        for (int f = 0; f < 4; ++f)
        {
            int64_t* t = Ftri + 3 * (4 * e + f);
            for (int k = 0, c = 0; k < 4; ++k)
                if (k != f)
                    t[c++] = E[4 * e + k];
            FA[4 * e + f] = 0.01;
        }
    for (int64_t i = 0; i < nV; ++i)
    {
        XVA[i] = 0.01;
        for (int k = 0; k < 8; ++k)
            fc[8 * i + k] = -1;
        if (i >= nV - nb)
            for (int k = 0; k < 1 + (int)(i % 8); ++k)
                fc[8 * i + k] = (i * 7 + k * 13) % (nF / 2);
    }
    s.fc = fc, s.F = Ftri, s.XVA = XVA, s.FA = FA, s.muC = 1e6, s.muF = 0.3, s.epsv = 1e-3, s.xb = xb;
    ora_step(&s, dt, 4, 1);
    double xf[9], xtf[9], Hc[9], gc[3];
    for (int k = 0; k < 9; ++k)
        xf[k] = uniform(), xtf[k] = xf[k] + 0.01 * uniform();
    double const p[3] = {0.1, 0.2, 0.05}, pt[3] = {0.1, 0.2, 0.1};
    double const ec = ora_vertex_triangle_contact(pt, p, xtf, xf, dt, 1e6, 0.3, 1e-3, gc, Hc);
    double dmany[16], tris[9 * 16];
    for (int k = 0; k < 9 * 16; ++k)
        tris[k] = uniform();
    ora_point_triangle_sqdist_many(p, tris, 16, dmany);
    double const d0 = ora_point_triangle_sqdist(p, tris, tris + 3, tris + 6);

    double sum = psi + gF[0] + HF[80] + en[0] + en[1] + en[2] + obj + grad[3 * nV - 1] + H[8] + g[2] + ge[0] + He[0] +
                 pe[0] + omega + ec + gc[0] + Hc[0] + dmany[15] + d0;
    for (int64_t i = 0; i < 3 * nV; ++i)
        sum += x[i];
    printf("oracle sanitizer run: %lld vertices, %lld tets, %lld colours, %lld updated, checksum %.6e\n",
           (long long)nV, (long long)nT, (long long)nc, (long long)nPadj, sum);
    int const ok = isfinite(sum);

    free(X), free(E), free(GP), free(wg), free(rho), free(m), free(GVGp), free(GVGe), free(GVGil), free(colors);
    free(Pptr), free(Padj), free(dbc), free(lame), free(x), free(v), free(aext), free(xt), free(vt), free(xtilde);
    free(xkm1), free(xkm2), free(xb), free(grad), free(ge), free(He), free(pe), free(Ftri), free(FA), free(XVA);
    free(fc);
    return ok ? 0 : 1;
}
