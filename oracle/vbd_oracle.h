/*
 * vbd_oracle.h -- FP64 CPU restatement of PBAT's Vertex Block Descent path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the MI355X product is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Each function cites the reference source it restates (paths relative to the
 * reference repository root). All arithmetic is double precision, all indices
 * int64, matching pbat::Scalar / pbat::Index (source/pbat/Aliases.h:17-18).
 *
 * Memory layouts follow the reference's Eigen column-major matrices:
 *   X, x, v, aext, ...  3 x nV  -> x[3*i + d]
 *   E                   4 x nT  -> E[4*e + k]
 *   GP                  4 x 3nT -> element e block GP_e(r,c) = GP[12*e + 4*c + r]
 *   lame                2 x nT  -> mu = lame[2*e], lambda = lame[2*e+1]
 *   F (3x3) vectorised column-major: F[3*c + r] = F(r,c)
 */
#ifndef PBAT_VBD_ORACLE_H
#define PBAT_VBD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sim/vbd/Enums.h:12-18 */
enum {
    ORA_INIT_POSITION = 0,
    ORA_INIT_INERTIA = 1,
    ORA_INIT_KINETIC_ENERGY_MINIMUM = 2,
    ORA_INIT_ADAPTIVE_VBD = 3,
    ORA_INIT_ADAPTIVE_PBAT = 4
};
/* graph/Enums.h */
enum { ORA_SELECT_LEAST_USED = 0, ORA_SELECT_FIRST_AVAILABLE = 1 };
enum { ORA_ORDER_NATURAL = 0, ORA_ORDER_SMALLEST_DEGREE = 1, ORA_ORDER_LARGEST_DEGREE = 2 };

/* Error codes */
enum {
    ORA_OK = 0,
    ORA_ERR_SINGULAR_JACOBIAN = -1, /* fem/Jacobian.h:74-77 */
    ORA_ERR_PALETTE_OVERFLOW = -2,  /* graph/Color.h NC=128 */
    ORA_ERR_ALLOC = -3,
    ORA_ERR_ARG = -4
};

/* ---- constitutive model ---------------------------------------------- */
/* physics/HyperElasticity.cpp:6-11 */
void ora_lame(double Y, double nu, double* mu, double* lambda);
/* physics/StableNeoHookeanEnergy.h:737-755 */
double ora_snh_eval(const double F[9], double mu, double lambda);
/* physics/StableNeoHookeanEnergy.h:766-795 */
void ora_snh_grad(const double F[9], double mu, double lambda, double gF[9]);
/* physics/StableNeoHookeanEnergy.h:1164-1326 (exact, no PSD projection). HF column-major. */
void ora_snh_hessian(const double F[9], double mu, double lambda, double HF[81]);

/* ---- FEM precompute ---------------------------------------------------- */
/* fem/ShapeFunctions.h:266-350 with fem/Tetrahedron.h:43-97 (Tet1 gradients). */
int ora_shape_function_gradients(int64_t nV, int64_t nT, const double* X, const int64_t* E, double* GP);
/* fem/MeshQuadrature.h:75-88, fem/Jacobian.h:67-140, rule <3,1> weight 0.166667 */
int ora_quadrature_weights(int64_t nV, int64_t nT, const double* X, const int64_t* E, double* wg);
/* sim/vbd/Data.cpp:210-219 via fem/Mass.h:103-112,838-874, rule <3,2> */
int ora_lumped_mass(int64_t nV, int64_t nT, const double* X, const int64_t* E, const double* rhoe, double* m);

/* ---- graphs ------------------------------------------------------------ */
/* sim/vbd/Data.cpp:223-226 with graph/Mesh.h:42-85: CSR vertex -> (tet, ilocal), tets ascending */
void ora_vertex_tet_adjacency(int64_t nV, int64_t nT, const int64_t* E, int64_t* GVGp, int64_t* GVGe,
                              int64_t* GVGilocal);
/* graph/Mesh.h:116-123 (G*G^T incl. self loops) + graph/Color.h:40-135 + common/ArgSort.h:31-39.
 * Returns the number of colours (>0) or an error code. */
int64_t ora_greedy_color(int64_t nV, int64_t nT, const int64_t* E, int ordering, int selection, int64_t* colors);
/* graph/Adjacency.h:172-181. Pptr has nColours+1 entries, Padj nV. Returns nColours. */
int64_t ora_map_to_adjacency(int64_t n, const int64_t* colors, int64_t* Pptr, int64_t* Padj);
/* graph/Adjacency.h:307-342 applied as in sim/vbd/Data.cpp:236-244. dbc need not be sorted.
 * Compresses Pptr/Padj in place and returns the new Padj size. */
int64_t ora_remove_vertices(int64_t nPartitions, int64_t* Pptr, int64_t* Padj, int64_t nV, int64_t nRemove,
                            const int64_t* remove);

/* ---- integrator -------------------------------------------------------- */
typedef struct ora_vbd {
    int64_t nV, nT;
    double* x;      /* 3*nV */
    double* v;      /* 3*nV */
    double* aext;   /* 3*nV */
    const double* m;/* nV */
    double* xt;     /* 3*nV */
    double* vt;     /* 3*nV */
    double* xtilde; /* 3*nV */
    const int64_t* E;
    const double* GP;
    const double* wg;
    const double* lame;
    const int64_t* GVGp;
    const int64_t* GVGe;
    const int64_t* GVGilocal;
    int64_t nPartitions;
    const int64_t* Pptr;
    const int64_t* Padj;
    int strategy;
    double kD;
    double detHZero;
    int nthreads;
    /* 0: CPU integrator semantics (vt = v before initialisation, sim/vbd/Integrator.cpp:31-32).
     * 1: GPU integrator semantics (vt keeps the previous substep's velocity,
     *    gpu/impl/vbd/Integrator.cu:329-347). Only AdaptiveVbd initialisation sees the difference. */
    int gpu_velocity_semantics;
    /* Chebyshev semi-iterative acceleration (sim/vbd/ChebyshevIntegrator.cpp:15-32):
     * on when chebyshev != 0, with estimated spectral radius rho and the previous
     * two iterates xkm1, xkm2 (3*nV each, caller-owned). */
    int chebyshev;
    double rho;
    double* xkm1;
    double* xkm2;
    /* Vertex-triangle contact of the GPU integrator (gpu/impl/vbd/Kernels.cuh:80-114,
     * 204-223): off when fc is NULL. fc holds up to 8 contacting triangles per vertex
     * (8*nV, packed, -1 padded), F the triangles (3 vertex indices each), XVA the
     * vertex areas (nV), FA the triangle areas. With contact on, ora_sweep writes
     * each colour's new positions to xb (3*nV) and copies them back after the colour,
     * so contact terms read the positions from before the colour pass, as the GPU
     * integrator's per-colour write buffer does (gpu/impl/vbd/Integrator.cu:303-326). */
    const int64_t* fc;
    const int64_t* F;
    const double* XVA;
    const double* FA;
    double muC;
    double muF;
    double epsv;
    double* xb;
} ora_vbd;

/* AccumulateVertexTriangleContact (sim/vbd/Kernels.h:215-302) for vertex position xv
 * (xtv at the start of the substep) against the triangle with vertex positions xf
 * (xtf), each 3x3 column-major (column k = vertex k): adds the penalty and lagged
 * IPC friction gradient to g and Hessian approximation to H (column-major; either may
 * be NULL) and returns the contact energy. */
double ora_vertex_triangle_contact(const double xtv[3], const double xv[3], const double xtf[9], const double xf[9],
                                   double dt, double muC, double muF, double epsv, double* g, double* H);

/* Squared distance from p to the triangle (a, b, c): closest point by Voronoi regions
 * (Ericson, Real-Time Collision Detection, 5.1.5), as geometry/DistanceQueries.h:217-222. */
double ora_point_triangle_sqdist(const double p[3], const double a[3], const double b[3], const double c[3]);

/* ora_point_triangle_sqdist from p to each of n triangles, tri holding (a, b, c) of
 * triangle k at tri[9k .. 9k+8]. */
void ora_point_triangle_sqdist_many(const double p[3], const double *tri, int64_t n, double *out);

/* sim/vbd/Kernels.h:97-102, including its operator precedence: for k >= 2 the
 * result is 4/4 - rho2*omega = 1 - rho2*omega. */
double ora_chebyshev_omega(int64_t k, double rho2, double omega);

/* sim/vbd/Integrator.cpp:22-45 */
int ora_step(ora_vbd* s, double dt, int64_t iterations, int64_t substeps);
/* One substep in pieces (single-threaded): Integrator::Step's preamble
 * (sim/vbd/Integrator.cpp:31-35, 54-71), one RunVbdIteration (:73-87), and the
 * velocity update (:39). sdt is the substep length. */
void ora_substep_begin(ora_vbd* s, double sdt);
void ora_sweep(ora_vbd* s, double sdt);
void ora_substep_end(ora_vbd* s, double sdt);
/* sim/vbd/Integrator.cpp:98-136, elastic part only (before damping and inertia):
 * H (3x3 column-major) and g accumulated through the full 9x9 SNH Hessian. */
void ora_vertex_elastic_block(const ora_vbd* s, int64_t i, double H[9], double g[3]);
/* sim/vbd/Integrator.cpp:138-166 */
double ora_objective(const ora_vbd* s, const double* x, const double* xtilde, double dt);
/* sim/vbd/Integrator.cpp:168-200 */
void ora_objective_gradient(const ora_vbd* s, const double* x, const double* xtilde, double dt, double* grad);
/* Total mechanical energy terms: out = {kinetic 1/2 sum m|v|^2, elastic sum w Psi(F), external -sum m aext.x} */
void ora_energy(const ora_vbd* s, double out[3]);
/* Per-element derivatives of w_e Psi_e at positions x (3*nV):
 *   ge[12 e + 3 k + d]            = w_e dPsi_e / dx_k[d]   (fem/DeformationGradient.h:126-141)
 *   He[144 e + 12 (3k+a) + 3l+b]  = w_e d2Psi_e / dx_k[a] dx_l[b]  (fem/DeformationGradient.h:231-253)
 *   psi[e]                        = w_e Psi_e
 * Any output may be NULL. */
void ora_element_derivatives(const ora_vbd* s, const double* x, double* ge, double* He, double* psi);

#ifdef __cplusplus
}
#endif

#endif
