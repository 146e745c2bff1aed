// Internal declarations shared by the translation units of libpbat_vbd.so:
// device arrays, the integrator handle and the host helpers that the
// forward integrator (vbd_integrator.hip) and the adjoint (vbd_adjoint.hip)
// both use. Not part of the C ABI.
#pragma once

#include "coarse_space.hpp"
#include "pbat_vbd.h"
#include "status.hpp"
#include "vbd_kernels.hpp"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

namespace pbat_vbd {

// Device bytes held by the library's arrays, process-wide, and their peak.
inline std::atomic<int64_t> gDeviceBytes{0};
inline std::atomic<int64_t> gDevicePeak{0};

inline void CountDeviceBytes(int64_t delta)
{
    int64_t const now = gDeviceBytes.fetch_add(delta) + delta;
    int64_t peak = gDevicePeak.load();
    while (now > peak && !gDevicePeak.compare_exchange_weak(peak, now))
    {
    }
}

#define PBAT_HIP(call)                                                                                            \
    do                                                                                                            \
    {                                                                                                             \
        hipError_t const _e = (call);                                                                             \
        if (_e != hipSuccess)                                                                                     \
            Fail(PBAT_VBD_ERROR_HIP, std::string(#call) + ": " + hipGetErrorString(_e));                          \
    } while (0)

template <class T>
struct DeviceArray
{
    T* ptr = nullptr;
    size_t n = 0;
    DeviceArray() = default;
    DeviceArray(DeviceArray const&) = delete;
    DeviceArray& operator=(DeviceArray const&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr, o.n = 0; }
    DeviceArray& operator=(DeviceArray&& o) noexcept
    {
        if (this != &o)
        {
            reset();
            ptr = o.ptr, n = o.n;
            o.ptr = nullptr, o.n = 0;
        }
        return *this;
    }
    ~DeviceArray() { reset(); }
    void reset()
    {
        if (ptr)
        {
            (void)hipFree(ptr);
            CountDeviceBytes(-static_cast<int64_t>(n * sizeof(T)));
        }
        ptr = nullptr;
        n = 0;
    }
    void allocate(size_t count)
    {
        reset();
        if (count == 0)
            return;
        hipError_t const e = hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T));
        if (e != hipSuccess)
        {
            ptr = nullptr;
            Fail(PBAT_VBD_ERROR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        n = count;
        CountDeviceBytes(static_cast<int64_t>(count * sizeof(T)));
    }
    void upload(std::vector<T> const& h, hipStream_t s)
    {
        allocate(h.size());
        if (!h.empty())
            PBAT_HIP(hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// Incidence records (structure of arrays, see SweepArgs): the sweep reads sids
// and sgeo (32 B); the adjoint's vertex kernel reads ids, g0, g1, g2 (the three
// GP rows, 56 B); shape holds the rest-shape constants the sweep's records are
// rebuilt from when the moduli change.
struct IncidenceRecords
{
    DeviceArray<int4> ids, sids;
    DeviceArray<float4> g0, g1, sgeo, shape;
    DeviceArray<float2> g2;
    void reset() { ids.reset(), sids.reset(), g0.reset(), g1.reset(), sgeo.reset(), shape.reset(), g2.reset(); }
};

struct HostIncidenceRecords
{
    std::vector<int4> ids, sids;
    std::vector<float4> g0, g1, sgeo, shape;
    std::vector<float2> g2;
    void resize(size_t n)
    {
        ids.assign(n, make_int4(-1, 0, 0, 0)); // a = -1: unused fixed-stride slot
        sids = ids;
        g0.assign(n, make_float4(0.f, 0.f, 0.f, 0.f));
        g1 = g0;
        sgeo = g0;
        shape = g0;
        g2.assign(n, make_float2(0.f, 0.f));
    }
    void upload(IncidenceRecords& d, hipStream_t s) const
    {
        d.ids.upload(ids, s);
        d.sids.upload(sids, s);
        d.g0.upload(g0, s);
        d.g1.upload(g1, s);
        d.g2.upload(g2, s);
        d.sgeo.upload(sgeo, s);
        d.shape.upload(shape, s);
    }
};

// Rollout checkpoints and adjoint work vectors (vbd_adjoint.hip).
struct RolloutState
{
    bool valid = false;
    // changes whenever a rollout is started or discarded, so a caller holding a
    // rollout can tell whether rollout_backward would still differentiate it
    int64_t generation = 0;
    void Discard()
    {
        valid = false;
        ++generation;
    }
    float dt = 0.f;
    int iterations = 0, steps = 0, every = 1;
    std::vector<DeviceArray<float4>> ckptX, ckptV, ckptVprev; // every `every` steps
    std::vector<DeviceArray<float4>> segX;                    // minimisers of one recomputed segment
    DeviceArray<float4> lastX;                                // positions after the last step
    DeviceArray<float4> y, r, z, p, q, rhs, xbar, vbar, hold;
    DeviceArray<float4> yS, bS; // the fused solvers' solution and right-hand side in row order
    // previous solutions of the backward pass (a ring, newest at histHead - 1),
    // the projection basis, its products with H, the per-aggregate partial sums
    // and the per-mesh coefficients of the projected initial guess
    DeviceArray<float4> hist[8], projBasis[8], projHb[8];
    DeviceArray<double> projPart, projCoef;
    DeviceArray<double> solveStats; // per solve of a backward pass: iterations, relative residual
    int histHead = 0, histCount = 0;
    DeviceArray<float4> dinv;                                 // inverse diagonal blocks, 2 per vertex
    DeviceArray<double> coarse[2];                            // coarse operator / its inverse (ping-pong)
    DeviceArray<int> coarseFinal;                             // per mesh: buffer holding its inverse
    DeviceArray<float> coarseF;                               // the inverse, symmetrised, FP32
    DeviceArray<float4> coarseD;                              // per solver row: rotation offset of the coarse space
    DeviceArray<double> pivotRows;                            // pivot rows of a Gauss-Jordan step, per mesh
    DeviceArray<int4> coarseInfo;                             // per aggregate: (b0, n, rows begin, rows end)
    DeviceArray<long long> coarseRowOffset;                   // per aggregate: its first row in coarseF
    int coarseMaxN = 0;                                       // rows of the largest mesh's coarse operator
    DeviceArray<double> rc, zc;                               // coarse residual and correction
    DeviceArray<double> scal;                                 // per-mesh solver scalars (kScalarCount x nMeshes)
    DeviceArray<int> flags;                                   // per-mesh done flags, iteration counts, live count
    DeviceArray<double> grad;                                 // d L / d log Y per tet
    DeviceArray<float> logY;                                  // parameter upload buffer
    // step Hessian assembled per solve as 3x3 blocks over the updated vertices:
    // the pattern and assembly sources (built once per handle: per row its
    // incidences and their tets' row positions, or, for rows too long for the
    // row-major assembly, per block its sources), the values, the longest row,
    // and the fused iteration's second direction buffer
    DeviceArray<int> bsrRowPtr, bsrCol, bsrBlockRow, bsrSrcPtr, bsrSrc;
    // row order of the pattern (-1: none yet, 0: the vertices' own, 1: aggregate-
    // major), the vertex of each row, and each row's mesh and aggregate, and each
    // aggregate's rows (CSR by aggVertPtr)
    int bsrOrder = -1;
    DeviceArray<int> rowVertex, meshOfRow, aggOfRow, aggRows;
    DeviceArray<int> incPtr, incCode;
    DeviceArray<int4> incSlot;
    DeviceArray<float4> bsrA, bsrB;
    DeviceArray<float> bsrC;
    int bsrMaxRow = 0;
    DeviceArray<float4> p2;
    // row chunks of the fused iteration (begin, end, mesh, 0), mesh-major, and
    // for each aggregate the chunks of its mesh (begin, end); chunks hold at
    // most bsrChunkRows rows
    DeviceArray<int4> bsrChunks;
    DeviceArray<int2> aggChunks;
    // per aggregate: its first kAggFirst vertices (-1 past its size) and its
    // mesh's done flag as of the last finish kernel
    DeviceArray<int> aggFirst, aggDone;
    int bsrChunkRows = 0;
    // per-chunk p . H p, per-aggregate r . z, r . r, b . b, per-coarse-row rc . zc
    DeviceArray<double> pqPart, rzPart, rrPart, bbPart, rczcPart;
    // Galerkin coarse operator from the blocks: coarse block k couples aggregates
    // (cbA[k], cbB[k]) and sums the fine blocks cbSrc[cbSrcPtr[k] .. cbSrcPtr[k + 1])
    DeviceArray<int> cbA, cbB, cbSrcPtr, cbSrc;
    // captured chunk of CG iterations and the settings it was captured for
    hipGraph_t cgGraph = nullptr;
    hipGraphExec_t cgExec = nullptr;
    // host-mapped copy of the per-mesh done flags the CG's finish kernel writes
    // (doneDev its device address), and the events that mark the ends of the
    // two most recent chunks of iterations
    int* doneHost = nullptr;
    int* doneDev = nullptr;
    int doneHostMeshes = 0;
    hipEvent_t chunkEvent[2] = {nullptr, nullptr};
    std::vector<double> cgKey;
    std::vector<void const*> cgPtrs; // device addresses the captured kernels take
    void ResetGraph()
    {
        if (cgExec)
            (void)hipGraphExecDestroy(cgExec);
        if (cgGraph)
            (void)hipGraphDestroy(cgGraph);
        cgExec = nullptr;
        cgGraph = nullptr;
        cgKey.clear();
        cgPtrs.clear();
    }
};

struct GraphKey
{
    float dt = 0.f;
    int iterations = -1, substeps = -1, strategy = -1, lanes = -1, block = -1, perLane = -1;
    bool xcd = false, nt = false, ell = false;
    float kD = 0.f, detHZero = 0.f;
    int acceleration = -1, window = 0;
    float rho = 0.f;
    bool operator==(GraphKey const& o) const
    {
        return dt == o.dt && iterations == o.iterations && substeps == o.substeps && strategy == o.strategy &&
               lanes == o.lanes && block == o.block && perLane == o.perLane && xcd == o.xcd && nt == o.nt &&
               ell == o.ell && kD == o.kD && detHZero == o.detHZero && acceleration == o.acceleration &&
               rho == o.rho && window == o.window;
    }
};

} // namespace pbat_vbd

struct pbat_vbd_integrator
{
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nV = 0, nT = 0, nIncidences = 0, nUpdated = 0, maxDegree = 0;
    std::vector<int> colorBegin; // internal vertex ranges per colour (size nColours + 1)
    std::vector<int> perm;       // internal -> original vertex index
    std::vector<int> iperm;      // original -> internal
    std::vector<float> massOrig; // lumped masses in original order (FP32)
    std::vector<int64_t> meshOffsets; // batch: first vertex of each mesh (count + 1), empty otherwise
    int nMeshes = 1;
    pbat_vbd::DeviceArray<int> meshOf; // mesh of each internal vertex (0 for a plain integrator)
    // Aggregates of updated vertices for the adjoint's two-level preconditioner:
    // ids contiguous per mesh; each mesh's coarse operator is a dense block of
    // (3 x its aggregates)^2 doubles.
    int nAggregates = 0;
    std::vector<int> aggBegin;          // per mesh: first aggregate id (nMeshes + 1)
    std::vector<int64_t> coarseOffset;  // per mesh: first double of its dense block (nMeshes + 1)
    pbat_vbd::DeviceArray<int> aggOf;   // aggregate of each internal vertex (-1: not updated)
    pbat_vbd::DeviceArray<int> aggMesh; // mesh of each aggregate
    pbat_vbd::DeviceArray<int> aggVertPtr, aggVerts; // vertices of each aggregate (CSR)
    pbat_vbd::DeviceArray<int> meshAggBegin;          // device copy of aggBegin
    pbat_vbd::DeviceArray<long long> meshCoarseOffset; // device copy of coarseOffset
    std::vector<float> wgHost;   // quadrature weights (traces only)
    std::vector<float> lameHost; // (mu, lambda) per element (traces only)

    pbat_vbd::DeviceArray<float4> xm, xtil, xt, vel, vprev, aext;
    pbat_vbd::DeviceArray<float4> tgp; // 3 per tet
    pbat_vbd::DeviceArray<int4> tetv;
    pbat_vbd::IncidenceRecords rows, ell; // CSR rows, fixed-stride rows
    int ellStride = 0;   // > 0 when the fixed-stride rows exist
    bool useEll = false; // sweep reads the fixed-stride rows
    pbat_vbd::DeviceArray<int> vtp;
    pbat_vbd::DeviceArray<double> energy;
    // record slot of (tet, local vertex) in the CSR and fixed-stride rows (-1: not swept)
    pbat_vbd::DeviceArray<int4> slotsRows, slotsEll;
    pbat_vbd::DeviceArray<float2> matBase; // (w mu / Y, w lambda / Y) per tet
    std::vector<float> logY0;              // log Young's modulus per tet at construction
    std::vector<float> logYCurrent;        // last value set (log Young's modulus per tet)
    pbat_vbd::RolloutState rollout;

    int strategy = PBAT_VBD_INIT_ADAPTIVE_PBAT;
    int acceleration = PBAT_VBD_ACCELERATION_NONE;
    float rho = 1.f;                            // Chebyshev estimated spectral radius
    int window = 5;                             // Anderson window size
    pbat_vbd::DeviceArray<float4> xkm1, xkm2;   // Chebyshev: previous two iterates; Anderson: x_{k-1}
    pbat_vbd::DeviceArray<float4> gkm1, fkm1, dgk, dfk; // Anderson: G_{k-1}, F_{k-1}, window columns
    pbat_vbd::DeviceArray<double> andersonAcc, andersonGram; // Anderson: dot products, Gram matrices
    double eta = 0.2, tau = 2.;               // trust region: acceptance threshold, radius factor
    bool curved = true;                       // trust region: curved (plain VBD iterates) or linear path
    pbat_vbd::DeviceArray<double> trPartials; // trust region: per-workgroup reduction partials + sums
    pbat_vbd::DeviceArray<double> trState;    // trust region: the model's FP64 scalars (device::TrSlot)
    pbat_vbd::DeviceArray<double> trLogT;     // trust region: t of each iteration k >= 2 of the last Step
    pbat_vbd::DeviceArray<int> trLogAccepted; // trust region: whether that step was kept
    int64_t trLogCount = 0;                   // trust region: entries of the last Step's log
    // Vertex-triangle contact (on when two bodies share the collision mesh)
    bool contact = false;
    int nCV = 0, nF = 0, activeSetFrequency = 1;
    float muC = 1e6f, muF = 0.3f, epsv = 1e-3f;
    pbat_vbd::DeviceArray<int> cvIdx, cvBody, fc, cvActive, cvInAv;
    pbat_vbd::DeviceArray<int4> tri;
    // the triangles in the detection kernels' storage order (grouped by body,
    // then along a Morton curve of their rest centroids) and each one's index
    pbat_vbd::DeviceArray<int4> triS;
    pbat_vbd::DeviceArray<int> triId;
    pbat_vbd::DeviceArray<float> triArea, xva, dupper;
    pbat_vbd::DeviceArray<float4> triLo, triHi, chunkLo, chunkHi, superLo, superHi, xfull, xb;
    pbat_vbd::DeviceArray<float4> triA, triB, triC; // triangle positions of the nearest pass
    pbat_vbd::DeviceArray<int4> crec;               // per-slot contact records of the sweep
    std::vector<int64_t> cvOrig; // original index of each collision vertex
    int pendBeg = 0, pendEnd = 0; // colour range whose new positions are still in xb
    float kD = 0.f;
    float detHZero = 1e-7f;
    // Step status the kernels report: statusHost[0] = 1 when a substep ended with a
    // non-finite position or velocity. Pinned, mapped, coherent host memory
    // (statusDev is its device address), so the host reads it after its stream
    // synchronisation with no copy; written by the device only in that case.
    int* statusHost = nullptr;
    int* statusDev = nullptr;
    pbat_vbd::DeviceArray<unsigned int> fp64Solves; // 3x3 block solves that took the FP64 path
    // FP64 objective and its gradient (allocated on first use): every vertex's
    // incident (tet, local index) as 4 e + l (CSR), staged x / x~, the gradient
    // and the reductions' per-workgroup partials
    pbat_vbd::DeviceArray<int> objVtp, objInc;
    pbat_vbd::DeviceArray<float4> objX, objXtil;
    pbat_vbd::DeviceArray<double> objGrad, objPart;
    int blockSize = 256;
    int lanes = 0;         // 0 = automatic
    int perLane = 0;       // incidences prefetched per lane, 0 = automatic
    bool xcdRemap = true;  // XCD-aware block order
    bool ntStream = false; // non-temporal loads of the incidence stream
    bool tuneNontemporal = true; // autotune also tries non-temporal stream loads
    bool useGraph = true;
    float worldMin[3] = {0.f, 0.f, 0.f}, worldMax[3] = {0.f, 0.f, 0.f};

    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    pbat_vbd::GraphKey graphKey;

    bool timing = false;
    int graphSweeps = 0; // sweep launches in the captured graph
    std::vector<hipEvent_t> events; // pairs (start, stop)
    std::vector<int> pendingLaunches; // sweep launches between each pair
    size_t eventsUsed = 0;
    double timedMs = 0.;
    int64_t timedLaunches = 0;

    ~pbat_vbd_integrator()
    {
        if (graphExec)
            (void)hipGraphExecDestroy(graphExec);
        if (graph)
            (void)hipGraphDestroy(graph);
        rollout.ResetGraph();
        for (hipEvent_t e : events)
            (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamSynchronize(stream);
        xm.reset(), xtil.reset(), xt.reset(), vel.reset(), vprev.reset(), aext.reset();
        tgp.reset(), tetv.reset(), rows.reset(), ell.reset(), vtp.reset(), energy.reset();
        slotsRows.reset(), slotsEll.reset(), matBase.reset(), meshOf.reset(), xkm1.reset(), xkm2.reset();
        aggOf.reset(), aggMesh.reset(), meshAggBegin.reset(), meshCoarseOffset.reset();
        aggVertPtr.reset(), aggVerts.reset();
        gkm1.reset(), fkm1.reset(), dgk.reset(), dfk.reset(), andersonAcc.reset(), andersonGram.reset();
        trPartials.reset(), trState.reset(), trLogT.reset(), trLogAccepted.reset();
        cvIdx.reset(), cvBody.reset(), fc.reset(), cvActive.reset(), cvInAv.reset();
        tri.reset(), triS.reset(), triId.reset(), triArea.reset(), xva.reset(), dupper.reset();
        triLo.reset(), triHi.reset(), chunkLo.reset(), chunkHi.reset(), superLo.reset(), superHi.reset();
        xfull.reset(), xb.reset();
        triA.reset(), triB.reset(), triC.reset(), crec.reset();
        rollout.ckptX.clear(), rollout.ckptV.clear(), rollout.ckptVprev.clear(), rollout.segX.clear();
        rollout.lastX.reset();
        fp64Solves.reset();
        objVtp.reset(), objInc.reset(), objX.reset(), objXtil.reset(), objGrad.reset(), objPart.reset();
        if (statusHost)
            (void)hipHostFree(statusHost);
        if (rollout.doneHost)
            (void)hipHostFree(rollout.doneHost);
        for (hipEvent_t e : rollout.chunkEvent)
            if (e)
                (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace pbat_vbd {

// Sweep shape in effect (automatic choices resolved).
int EffectiveLanes(pbat_vbd_integrator const& h);
int EffectivePerLane(pbat_vbd_integrator const& h);
// Launch one colour's sweep over nVerts vertices with the current shape.
void LaunchSweep(pbat_vbd_integrator& h, device::SweepArgs const& a, int nVerts);
device::SweepArgs MakeSweepArgs(pbat_vbd_integrator& h, float sdt, float sdt2);
// One forward Step (gpu/impl/vbd/Integrator.cu:82-103): graph replay or eager launches.
// sync = false (rollouts): the step is only queued; the caller clears the status
// before its steps and synchronises and checks it after them.
void Step(pbat_vbd_integrator& h, float dt, int iterations, int substeps, bool sync = true);
// Status reported by the kernels of the work queued since the last ClearStatus,
// read after a synchronisation of the handle's stream: fails with
// PBAT_VBD_ERROR_NONFINITE when a substep ended with non-finite positions.
void ClearStatus(pbat_vbd_integrator& h);
void CheckStatus(pbat_vbd_integrator& h);
void InvalidateGraph(pbat_vbd_integrator& h);
// Host <-> device copies of 3 x nV column-major matrices (original vertex order)
// into / out of float4 arrays (internal order); w components from wsrc (or 0).
void UploadVec(pbat_vbd_integrator& h, float4* dst, float const* src, float const* wsrc);
void DownloadVec(pbat_vbd_integrator& h, float4 const* src, float* dst);
void CheckVertexArray(pbat_vbd_integrator const& h, void const* p, int64_t n, char const* what);
pbat_vbd_integrator& Handle(pbat_vbd_integrator* h);
// Vertex-triangle contact (vbd_contact.hip): set-up from the data, the step's active
// set, contacts (finalize = 0) or the end-of-step active flags (finalize = 1), and
// the copy of the last contact-swept colour back from the write buffer.
bool CollisionMeshSpansBodies(pbat_vbd_data const& d);
void SetUpContact(pbat_vbd_integrator& h, pbat_vbd_data const* d);
void ContactBeginStep(pbat_vbd_integrator& h, float dt);
void ContactNearestPass(pbat_vbd_integrator& h, int finalize);
void FlushContactSweep(pbat_vbd_integrator& h);

} // namespace pbat_vbd
