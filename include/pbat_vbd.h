/*
 * pbat_vbd.h -- C ABI of the MI355X-native Vertex Block Descent library
 * (libpbat_vbd.so).
 *
 * Drop-in for PBAT's GPU VBD path: pbat::gpu::vbd::Integrator
 * (source/pbat/gpu/vbd/Integrator.h:33-150) constructed from a
 * pbat::sim::vbd::Data (source/pbat/sim/vbd/Data.h:14-247). Each entry point
 * names the reference member it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only. Host arrays are caller-owned and are copied
 *    in or out during the call; nothing is retained across the boundary.
 *  - Per-vertex 3-vectors are passed as 3 x n column-major matrices, i.e. the
 *    layout of Eigen's Matrix<Scalar, 3, Dynamic>: value (d, i) at [3*i + d].
 *    Elements are 4 x nT column-major: vertex k of element e at [4*e + k].
 *  - Every function returns PBAT_VBD_OK (0) or a negative status; the message
 *    of the last failure on the calling thread is pbat_vbd_last_error().
 *    The reference raises std::invalid_argument / std::runtime_error in the
 *    same situations (sim/vbd/Data.cpp:245-305, gpu/impl/common/Eigen.cuh).
 *  - An integrator handle is not thread-safe. Its work runs on one HIP stream
 *    owned by the handle; setters and getters synchronise that stream, so the
 *    observable behaviour is host-synchronous as in the reference.
 */
#ifndef PBAT_VBD_H
#define PBAT_VBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBAT_VBD_OK 0
#define PBAT_VBD_ERROR_INVALID_ARGUMENT (-1) /* std::invalid_argument in the reference */
#define PBAT_VBD_ERROR_HIP (-2)              /* HIP runtime failure */
#define PBAT_VBD_ERROR_OUT_OF_MEMORY (-3)
#define PBAT_VBD_ERROR_SINGULAR_ELEMENT (-4) /* fem/Jacobian.h:74-77 "Inverted or singular jacobian" */
#define PBAT_VBD_ERROR_PALETTE (-5)          /* more than 128 colours (graph/Color.h NC) */
#define PBAT_VBD_ERROR_NONFINITE (-6)        /* NaN/Inf positions detected on the device */

/* sim/vbd/Enums.h:12-18 */
#define PBAT_VBD_INIT_POSITION 0
#define PBAT_VBD_INIT_INERTIA 1
#define PBAT_VBD_INIT_KINETIC_ENERGY_MINIMUM 2
#define PBAT_VBD_INIT_ADAPTIVE_VBD 3
#define PBAT_VBD_INIT_ADAPTIVE_PBAT 4

/* graph/Enums.h: EGreedyColorOrderingStrategy / EGreedyColorSelectionStrategy */
#define PBAT_VBD_ORDER_NATURAL 0
#define PBAT_VBD_ORDER_SMALLEST_DEGREE 1
#define PBAT_VBD_ORDER_LARGEST_DEGREE 2
#define PBAT_VBD_SELECT_LEAST_USED 0
#define PBAT_VBD_SELECT_FIRST_AVAILABLE 1

#define PBAT_VBD_MAX_COLORS 128

/* sim/vbd/Enums.h:21-28 EAccelerationStrategy; gpu::vbd::Integrator's factory
 * (gpu/vbd/Integrator.cu:15-33) picks the integrator from it. None, Chebyshev
 * (gpu/impl/vbd/ChebyshevIntegrator.cu), Anderson (gpu/impl/vbd/AndersonIntegrator.cu)
 * and trust region (gpu/impl/vbd/TrustRegionIntegrator.cu) are available; Anderson
 * and trust region couple every vertex of a handle and take single meshes, not
 * batches. */
#define PBAT_VBD_ACCELERATION_NONE 0
#define PBAT_VBD_ACCELERATION_CHEBYSHEV 1
#define PBAT_VBD_ACCELERATION_ANDERSON 2
#define PBAT_VBD_ACCELERATION_TRUST_REGION 3
#define PBAT_VBD_MAX_ANDERSON_WINDOW 16

/* Message of the last failure on the calling thread ("" if none). */
const char* pbat_vbd_last_error(void);

/* Library version string. */
const char* pbat_vbd_version(void);

/* ------------------------------------------------------------------------- */
/* Data construction                                                         */
/* ------------------------------------------------------------------------- */

/*
 * Replaces pbat::sim::vbd::Data::Construct (sim/vbd/Data.cpp:179-244) for the
 * element, adjacency and partition data; host-only, touches no GPU.
 *
 * Inputs: rest positions X (3 x nV), tetrahedra E (4 x nT), per-element mass
 * density rhoe, Lame coefficients mue, lambdae (each nT; pass NULL for the
 * reference defaults rho = 1e3, LameCoefficients(1e6, 0.45)), Dirichlet
 * vertices dbc (nDbc, any order; NULL if none) and the colouring strategy.
 *
 * Outputs (caller-allocated, sizes in brackets):
 *   m [nV] lumped masses; GP [12 nT] shape function gradients as the 4 x 3nT
 *   matrix of the reference; wg [nT] quadrature weights; lame [2 nT];
 *   GVGp [nV+1], GVGe [4 nT], GVGilocal [4 nT] vertex -> (element, local index)
 *   CSR; colors [nV]; Pptr [PBAT_VBD_MAX_COLORS + 1]; Padj [nV].
 *   *nPartitions receives the number of colours, *nPadj the number of
 *   non-Dirichlet vertices listed in Padj.
 */
int pbat_vbd_data_construct(int64_t nV, int64_t nT, const double* X, const int64_t* E, const double* rhoe,
                            const double* mue, const double* lambdae, const int64_t* dbc, int64_t nDbc,
                            int ordering, int selection, double* m, double* GP, double* wg, double* lame,
                            int64_t* GVGp, int64_t* GVGe, int64_t* GVGilocal, int64_t* colors, int64_t* Pptr,
                            int64_t* Padj, int64_t* nPartitions, int64_t* nPadj);

/* ------------------------------------------------------------------------- */
/* GPU integrator                                                            */
/* ------------------------------------------------------------------------- */

/* Fields of pbat::sim::vbd::Data consumed by gpu::impl::vbd::Integrator's
 * constructor (gpu/impl/vbd/Integrator.cu:22-80). All FP64/int64 host arrays,
 * converted to the device's FP32/int32 layout on upload. */
typedef struct pbat_vbd_data {
    int64_t nV;
    int64_t nT;
    const double* x;         /* 3 x nV positions */
    const double* v;         /* 3 x nV velocities */
    const double* aext;      /* 3 x nV external accelerations */
    const double* m;         /* nV lumped masses */
    const int64_t* E;        /* 4 x nT */
    const double* GP;        /* 4 x 3nT */
    const double* wg;        /* nT */
    const double* lame;      /* 2 x nT */
    const int64_t* GVGp;     /* nV + 1 */
    const int64_t* GVGe;     /* GVGp[nV] */
    const int64_t* GVGilocal;/* GVGp[nV] */
    int64_t nPartitions;
    const int64_t* Pptr;     /* nPartitions + 1 */
    const int64_t* Padj;     /* Pptr[nPartitions] */
    int32_t strategy;        /* PBAT_VBD_INIT_* */
    double kD;               /* Rayleigh damping coefficient */
    double detHZero;         /* numerical zero of |det H| */
    int32_t acceleration;    /* PBAT_VBD_ACCELERATION_* (sim/vbd/Enums.h:21-28) */
    double rho;              /* Chebyshev estimated spectral radius, 0 < rho < 1 (sim/vbd/Data.h:233-234) */
    int32_t window;          /* Anderson window size, 1..PBAT_VBD_MAX_ANDERSON_WINDOW (sim/vbd/Data.h:236-237) */
    double eta;              /* trust region: energy reduction accuracy threshold, >= 0 (sim/vbd/Data.h:244) */
    double tau;              /* trust region: radius increase factor, > 1 (sim/vbd/Data.h:245) */
    int32_t curved;          /* trust region: curved (1) or linear (0) accelerated path (sim/vbd/Data.h:246) */
    /* Collision mesh and contact model (sim/vbd/Data.h:165-178, 225-229; Data.cpp:34-60,
     * 121-133). Contact runs when two bodies share the collision mesh; zeros/NULL
     * leave it off. */
    int64_t nCollisionVertices;
    const int64_t* V;        /* collision vertices (indices into x) */
    int64_t nCollisionTriangles;
    const int64_t* F;        /* 3 x nCollisionTriangles, column-major (F[3f + k]) */
    const int64_t* B;        /* nV body index of every vertex (NULL: one body) */
    const double* XVA;       /* nV vertex areas */
    const double* FA;        /* nCollisionTriangles triangle areas */
    double muC;              /* collision penalty */
    double muF;              /* friction coefficient */
    double epsv;             /* IPC friction's relative velocity threshold */
    int32_t activeSetUpdateFrequency; /* substeps between contact updates, >= 1 */
} pbat_vbd_data;

typedef struct pbat_vbd_integrator pbat_vbd_integrator;

/* gpu::vbd::Integrator::Integrator(Data const&) (gpu/vbd/Integrator.cu:15-33).
 * Uses the current HIP device. */
int pbat_vbd_integrator_create(const pbat_vbd_data* data, pbat_vbd_integrator** out);
/* gpu::vbd::Integrator::~Integrator (gpu/vbd/Integrator.cu:52-56) */
int pbat_vbd_integrator_destroy(pbat_vbd_integrator* h);

/* gpu::vbd::Integrator::Step (gpu/vbd/Integrator.h:72; gpu/impl/vbd/Integrator.cu:82-103).
 * Returns PBAT_VBD_ERROR_NONFINITE when a substep ended with a non-finite
 * position or velocity at any vertex (the reference has no such check and
 * carries NaN on); the state is left as computed. */
int pbat_vbd_integrator_step(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t substeps);
/* gpu::vbd::Integrator::TracedStep (gpu/vbd/Integrator.h:85-90; gpu/impl/vbd/Integrator.cu:105-148).
 * Writes MatrixMarket files named {variable}.t.{t}.s.{s}[.k.{k}].mtx into dir. */
int pbat_vbd_integrator_traced_step(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t substeps,
                                    int32_t t, const char* dir);

/* gpu::vbd::Integrator::SetPositions / SetVelocities / SetExternalAcceleration
 * (gpu/vbd/Integrator.cu:73-91): 3 x n column-major FP32, n must equal nV. */
int pbat_vbd_integrator_set_positions(pbat_vbd_integrator* h, const float* x, int64_t n);
int pbat_vbd_integrator_set_velocities(pbat_vbd_integrator* h, const float* v, int64_t n);
int pbat_vbd_integrator_set_external_acceleration(pbat_vbd_integrator* h, const float* a, int64_t n);
/* gpu::vbd::Integrator::GetPositions / GetVelocities (gpu/vbd/Integrator.cu:117-125) */
int pbat_vbd_integrator_get_positions(pbat_vbd_integrator* h, float* x, int64_t n);
int pbat_vbd_integrator_get_velocities(pbat_vbd_integrator* h, float* v, int64_t n);

/* gpu::vbd::Integrator::SetNumericalZeroForHessianDeterminant (gpu/vbd/Integrator.cu:93-96) */
int pbat_vbd_integrator_set_detH_zero(pbat_vbd_integrator* h, float zero);
/* gpu::vbd::Integrator::SetRayleighDampingCoefficient (gpu/vbd/Integrator.cu:98-101) */
int pbat_vbd_integrator_set_rayleigh_damping(pbat_vbd_integrator* h, float kD);
/* gpu::vbd::Integrator::SetInitializationStrategy (gpu/vbd/Integrator.cu:103-106) */
int pbat_vbd_integrator_set_initialization_strategy(pbat_vbd_integrator* h, int32_t strategy);
/* gpu::vbd::Integrator::SetBlockSize (gpu/vbd/Integrator.cu:108-111; clamp as gpu/impl/vbd/Integrator.cu:158-161).
 * Threads per workgroup of the sweep kernel; clamped to [64, 256] and rounded down to a multiple of 64. */
int pbat_vbd_integrator_set_block_size(pbat_vbd_integrator* h, int32_t block_size);
/* gpu::vbd::Integrator::SetSceneBoundingBox (gpu/vbd/Integrator.cu:113-116). Stored only: in
 * the reference the box bounds the Morton codes of the contact detector's BVH; the
 * exhaustive detector here needs no such bounds. */
int pbat_vbd_integrator_set_scene_bounding_box(pbat_vbd_integrator* h, const float min[3], const float max[3]);

/* Lanes of a wavefront that cooperate on one vertex (1, 2, 4, 8, 16, 32 or 64;
 * 0 = choose from the mesh's maximum vertex degree). No reference counterpart:
 * the reference always uses one thread block per vertex. */
int pbat_vbd_integrator_set_lanes_per_vertex(pbat_vbd_integrator* h, int32_t lanes);
/* Shape of the sweep kernel. No reference counterpart. */
typedef struct pbat_vbd_sweep_config {
    int32_t lanes_per_vertex;    /* power of two in [1, 64]; 0 = automatic */
    int32_t incidences_per_lane; /* incidences a lane prefetches before accumulating, 1..4; 0 = automatic */
    int32_t xcd_remap;           /* reorder workgroups so each XCD's L2 serves one range of vertices */
    int32_t nontemporal;         /* non-temporal loads of the incidence stream and tet constants */
    int32_t fixed_stride;        /* incidences in fixed-stride (ELL) rows instead of CSR rows */
} pbat_vbd_sweep_config;

/* Set / read the sweep shape. get_ reports the values in effect (automatic
 * choices resolved). fixed_stride = 1 fails if the mesh's degree spread made
 * the library skip the fixed-stride copy at construction. */
int pbat_vbd_integrator_set_sweep_config(pbat_vbd_integrator* h, const pbat_vbd_sweep_config* config);
int pbat_vbd_integrator_get_sweep_config(pbat_vbd_integrator* h, pbat_vbd_sweep_config* config);
/* Autotuner: times `sweeps` full colour sweeps at substep dt for every
 * candidate sweep shape (lanes with their natural prefetch depth, XCD order
 * on/off, fixed-stride/CSR rows) and keeps the fastest; the simulation state
 * is restored afterwards. timings (nullable, >= 192 doubles) receives up to 32
 * rows of (lanes, incidences per lane, xcd remap, non-temporal, fixed stride,
 * microseconds per sweep); *n_candidates the number of candidates timed. */
int pbat_vbd_integrator_autotune(pbat_vbd_integrator* h, float dt, int32_t sweeps, double* timings,
                                 int32_t* n_candidates);
/* Capture each Step's launch sequence in a hipGraph and replay it (default on). */
int pbat_vbd_integrator_set_graph_capture(pbat_vbd_integrator* h, int32_t enabled);

/* ------------------------------------------------------------------------- */
/* Batched integrator (no reference counterpart)                             */
/* ------------------------------------------------------------------------- */

/* One integrator for `count` independent meshes (each a constructed Data, all
 * sharing strategy, kD and detHZero). The meshes become one mesh with
 * disconnected components: global colour c is the union of every mesh's
 * colour c, so each mesh is swept in its own order and one launch per colour
 * serves the batch. Every pbat_vbd_integrator_* entry point applies to the
 * returned handle; positions and velocities are the meshes' 3 x nV_b matrices
 * concatenated in batch order. */
int pbat_vbd_batch_create(const pbat_vbd_data* meshes, int32_t count, pbat_vbd_integrator** out);
/* First vertex of each mesh in the concatenated arrays: offsets[count + 1]
 * (nullable to query *count first). A plain integrator reports one mesh. */
int pbat_vbd_batch_vertex_offsets(pbat_vbd_integrator* h, int64_t* offsets, int32_t* count);

/* Total mechanical energy terms reduced in FP64 on the device:
 * out = {1/2 sum m |v|^2, sum_e w_e Psi(F_e), -sum m aext . x}.
 * Counterpart of the FP64 objective reductions of
 * gpu/impl/vbd/TrustRegionIntegrator.cu:265-364 and sim/vbd/Integrator.cpp:138-166. */
int pbat_vbd_integrator_energy(pbat_vbd_integrator* h, double out[3]);

/* sim::vbd::Integrator::ObjectiveFunction and ObjectiveFunctionGradient
 * (sim/vbd/Integrator.h:41-49, Integrator.cpp:138-200) on the device, reduced in
 * FP64 in a fixed order: *f = 1/2 sum m |x - x~|^2 + dt^2 sum_e w_e Psi_e at
 * positions x and inertial target xtilde (each 3 x nV FP32; NULL = the handle's
 * current positions / the inertial target of its last substep); grad (3 x nV
 * FP64, vertex-major as the reference's reshaped gradient, Dirichlet vertices
 * included) its gradient; *residual2 the squared gradient norm over the
 * non-Dirichlet vertices, the quantity the VBD iteration drives to zero. Each
 * output pointer may be NULL. */
int pbat_vbd_integrator_objective(pbat_vbd_integrator* h, const float* x, const float* xtilde, float dt, double* f,
                                  double* grad, double* residual2);
/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code between other work on a stream it must not disturb, serving loops,
 * harnesses that forbid hipMalloc). The same value, bit for bit, as
 * pbat_vbd_integrator_objective(h, NULL, NULL, dt, f, NULL, NULL): the objective at
 * the handle's current positions and inertial target, reduced in FP64 in a fixed
 * order, with its partial sums in a file-scope device array of 16 KiB per GPU
 * (looked up once per GPU, kept until the process ends) instead of buffers the
 * handle allocates. Calls no hipMalloc or hipFree. The scratch does not grow with
 * the mesh (a fixed 512-workgroup reduction), so the only size limit is the
 * kernels' 32-bit indexing: nV and nT below 2^31, larger meshes are rejected with
 * PBAT_VBD_ERROR_INVALID_ARGUMENT. Every call on a GPU overwrites that GPU's
 * scratch: calls that overlap on one GPU (several host threads) race. */
int pbat_vbd_integrator_objective_noalloc(pbat_vbd_integrator* h, float dt, double* f);
/* sim::vbd::Integrator::TraceNextStep followed by Step (sim/vbd/Integrator.cpp:22-52,
 * 73-87, 202-235): one step whose iterates are traced. Before each VBD iteration and
 * after the last, the FP64 objective, its gradient and the positions are recorded;
 * per substep s they are written to {dir}/{t}.{s}.f.mtx ((iterations + 1) x 1),
 * {t}.{s}.grad.mtx and {t}.{s}.x.mtx (3 nV x (iterations + 1), columns vertex-major),
 * MatrixMarket dense arrays as the reference's Eigen::saveMarketDense writes them.
 * The step itself is the handle's Step, run eagerly (the linear trust-region path
 * runs plain VBD iterations here). */
int pbat_vbd_integrator_trace_objective_step(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t substeps,
                                             int32_t t, const char* dir);

/* ------------------------------------------------------------------------- */
/* Differentiable rollouts (no reference counterpart)                        */
/* ------------------------------------------------------------------------- */
/* The reference has no gradients of its VBD step. These entry points give
 * d L / d (per-tet log Young's modulus, v_0, x_0) of a loss L(x_T, v_T) of a
 * rollout by the implicit-function theorem on each step's incremental
 * potential (sim/vbd/Integrator.cpp:138-166: its gradient is zero at the
 * step's minimiser). Per step, backwards:
 *   solve (M / h^2 + d2U/dx2(x_{t+1})) y = xbar + vbar / h   (free vertices)
 *   dL/dlogY_e -= y . (w_e dPsi_e/dx)
 *   xbar <- M y / h^2 - vbar / h,  vbar <- M y / h
 * Assumes one substep, no Rayleigh damping and a converged forward: with a
 * fixed iteration count the adjoint is that of the exact minimiser. */

enum pbat_vbd_adjoint_solver {
    PBAT_VBD_ADJOINT_PCG = 0,          /* CG preconditioned by symmetric coloured block Gauss-Seidel */
    PBAT_VBD_ADJOINT_GAUSS_SEIDEL = 1, /* coloured block Gauss-Seidel sweeps */
    PBAT_VBD_ADJOINT_PCG_JACOBI = 2,   /* CG preconditioned by the inverse 3x3 diagonal blocks */
    PBAT_VBD_ADJOINT_PCG_TWO_LEVEL = 3 /* block Jacobi plus a coarse correction over the rigid motions
                                          (translations, rotations) of vertex aggregates (coarse
                                          operator from the rollout's last state, inverted once per
                                          backward pass) */
};

/* Each mesh of a batch is solved with its own CG scalars and stopping test
 * (a converged mesh is frozen), as if it were solved alone. */
typedef struct pbat_vbd_adjoint_config {
    int32_t solver;         /* pbat_vbd_adjoint_solver */
    int32_t max_iterations; /* per linear solve */
    float tolerance;        /* on |b - H y| / |b| per mesh */
    int32_t check_every;    /* iterations between host polls of the live-mesh count */
    int32_t history;        /* 0..8: the block Jacobi and two-level solvers start each solve of a
                               backward pass from the projection (H-norm) of its solution onto the
                               span of the last `history` solutions; 0: from the last solution */
} pbat_vbd_adjoint_config;

typedef struct pbat_vbd_adjoint_stats {
    int32_t solves;
    int32_t total_iterations; /* sum over solves of the largest per-mesh iteration count */
    int32_t max_iterations;
    float max_relative_residual; /* largest per-mesh |b - H y| / |b| */
} pbat_vbd_adjoint_stats;

/* Per-tet Young's modulus Y_e = exp(logY[e]) with each tet's Poisson ratio
 * kept: (mu_e, lambda_e) are rescaled by Y_e / Y_e(construction). nT values. */
int pbat_vbd_integrator_set_log_young_modulus(pbat_vbd_integrator* h, const float* logY, int64_t nT);
int pbat_vbd_integrator_get_log_young_modulus(pbat_vbd_integrator* h, float* logY, int64_t nT);
/* Step `steps` times (one substep each) from the current state, keeping the
 * state every `checkpoint_every` steps for the backward pass. Setting
 * positions, velocities, external accelerations, moduli, solver settings or
 * the sweep shape (or autotuning) afterwards discards the rollout. */
int pbat_vbd_rollout_forward(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t steps,
                             int32_t checkpoint_every);
/* Gradients of L given dL/dx_T (3 x nV, required) and dL/dv_T (3 x nV,
 * nullable = 0), original vertex order. Outputs (each nullable): dL/dlogY
 * (nT), dL/dv_0 and dL/dx_0 (3 x nV; zero on Dirichlet vertices). Segments
 * between checkpoints are recomputed; the integrator's state at the end of
 * the rollout is restored afterwards. config nullable (PCG, 200, 1e-6, 5, history 0). */
int pbat_vbd_rollout_backward(pbat_vbd_integrator* h, const float* dL_dx, const float* dL_dv, float* dL_dlogY,
                              float* dL_dv0, float* dL_dx0, const pbat_vbd_adjoint_config* config,
                              pbat_vbd_adjoint_stats* stats);
/* Identity of the rollout that pbat_vbd_rollout_backward would differentiate:
 * *generation changes whenever a rollout is started (rollout_forward) or
 * discarded (any of the setters listed above); *valid is 1 while one is held.
 * A caller that keeps a rollout across other work (an autograd graph, say)
 * compares the generation before calling rollout_backward. Either pointer may
 * be NULL. */
int pbat_vbd_rollout_generation(pbat_vbd_integrator* h, int64_t* generation, int32_t* valid);
/* out = (M / h^2 + d2U/dx2(x)) p on free vertices (0 on Dirichlet ones);
 * x, p, out: 3 x nV. */
int pbat_vbd_hessian_vector(pbat_vbd_integrator* h, const float* x, float dt, const float* p, float* out);
/* Solve (M / h^2 + d2U/dx2(x)) y = b on free vertices with the adjoint's
 * solver (b's Dirichlet rows ignored). x, b, y: 3 x nV. */
int pbat_vbd_adjoint_solve(pbat_vbd_integrator* h, const float* x, float dt, const float* b, float* y,
                           const pbat_vbd_adjoint_config* config, pbat_vbd_adjoint_stats* stats);

/* Device memory held by the library's arrays in this process (all handles),
 * in bytes, and its peak since load or the last reset_peak != 0. No reference
 * counterpart. */
int pbat_vbd_device_memory(int64_t* current, int64_t* peak, int32_t reset_peak);

/* Number of 3x3 block solves of the sweep that took the FP64 path since the
 * handle was created or last reset (reset != 0 clears it after reading). The
 * sweep solves a vertex's block in FP32 by its adjugate, and in FP64 from the
 * same FP32 entries when |det H| < 1e-4 |H00 H11 H22| (ill-conditioned; the
 * reference solves in its build's scalar type, sim/vbd/Kernels.h:324-339). No
 * reference counterpart. */
int pbat_vbd_integrator_fp64_solves(pbat_vbd_integrator* h, int64_t* count, int32_t reset);

/* Wait for all work queued on the handle's stream. */
int pbat_vbd_integrator_synchronize(pbat_vbd_integrator* h);

/* Number of colours (vertex partitions) and the number of vertices updated per sweep. */
int pbat_vbd_integrator_counts(pbat_vbd_integrator* h, int64_t* nPartitions, int64_t* nUpdated,
                               int64_t* nIncidences);

/* Contact state after the last Step (the reference keeps it in
 * VertexTriangleMixedCcdDcd::active and Integrator::fc,
 * gpu/impl/contact/VertexTriangleMixedCcdDcd.cuh:151-160, gpu/impl/vbd/Integrator.cuh:165):
 * active[c] for each collision vertex c in the order of pbat_vbd_data::V, and up to
 * 8 contacting triangles per vertex (fc[8 i + k], -1 padded, vertices in original
 * order). Either may be NULL; fails when contact is off. */
int pbat_vbd_integrator_contacts(pbat_vbd_integrator* h, int32_t* active, int32_t* fc);

/* Decisions of the linear trust-region path in the last Step, one entry per
 * iteration k >= 2 of each substep, in order: the step length t along the VBD
 * step and whether the step was kept. The structured form of the reference's
 * profiling log lines "t={},t*={}" and "Accepted/Rejected TR step"
 * (gpu/impl/vbd/TrustRegionIntegrator.cu:111,133,146). Writes at most
 * capacity entries; *count receives the number recorded. The log is kept in
 * device memory during the step; this call synchronises the handle's stream. */
int pbat_vbd_integrator_trust_region_log(pbat_vbd_integrator* h, double* t, int32_t* accepted, int64_t capacity,
                                         int64_t* count);

/* Kernel-time instrumentation: when enabled, HIP events on the handle's stream
 * bracket each replay of the captured step graph (which also holds the substep
 * begin/end kernels), or, with graph capture off, each substep's back-to-back
 * sweep launches. pbat_vbd_integrator_sweep_timing returns (summed event ms,
 * number of sweep launches inside the brackets) since the last reset. No
 * reference counterpart. */
int pbat_vbd_integrator_set_timing(pbat_vbd_integrator* h, int32_t enabled);
int pbat_vbd_integrator_sweep_timing(pbat_vbd_integrator* h, double* total_ms, int64_t* launches, int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* PBAT_VBD_H */
