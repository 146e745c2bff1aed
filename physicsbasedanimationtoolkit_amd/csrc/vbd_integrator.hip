// MI355X VBD integrator: host orchestration behind the C ABI of pbat_vbd.h.
//
// Counterpart of pbat::gpu::impl::vbd::Integrator (gpu/impl/vbd/Integrator.cu:22-376)
// with a different device design:
//  - vertices are renumbered colour-major, so each colour is one contiguous
//    range and a sweep over it reads per-vertex data coalesced; within a colour
//    vertices are sorted by degree (stable) so lanes of a wave do equal work;
//  - per-tet data is one 64-byte record (vertex ids, 9 shape-function
//    gradients, w*mu, w*lambda, alpha), gathered through a 4-byte per-incidence
//    CSR reference (tet << 2 | ilocal);
//  - positions carry the lumped mass in their 4th component, so one 16-byte
//    gather serves both;
//  - updates are written in place (no xb write buffer + copy-back kernel):
//    vertices of one colour share no tet, so no vertex reads what another
//    vertex of its colour writes;
//  - the substep-begin/-end work is two fused elementwise kernels, and the
//    whole launch sequence of a Step is captured once in a hipGraph and replayed.
#include "pbat_vbd.h"
#include "status.hpp"
#include "vbd_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace pbat_vbd {

namespace {

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
    ~DeviceArray() { reset(); }
    void reset()
    {
        if (ptr)
            (void)hipFree(ptr);
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
    }
    void upload(std::vector<T> const& h, hipStream_t s)
    {
        allocate(h.size());
        if (!h.empty())
            PBAT_HIP(hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct GraphKey
{
    float dt = 0.f;
    int iterations = -1, substeps = -1, strategy = -1, lanes = -1, block = -1;
    float kD = 0.f, detHZero = 0.f;
    bool operator==(GraphKey const& o) const
    {
        return dt == o.dt && iterations == o.iterations && substeps == o.substeps && strategy == o.strategy &&
               lanes == o.lanes && block == o.block && kD == o.kD && detHZero == o.detHZero;
    }
};

} // namespace
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

    pbat_vbd::DeviceArray<float4> xm, xtil, xt, vel, vprev, aext;
    pbat_vbd::DeviceArray<pbat_vbd::device::TetRecord> tets;
    pbat_vbd::DeviceArray<int> vtp, vtref;
    pbat_vbd::DeviceArray<double> energy;

    int strategy = PBAT_VBD_INIT_ADAPTIVE_PBAT;
    float kD = 0.f;
    float detHZero = 1e-7f;
    int blockSize = 256;
    int lanes = 0; // 0 = automatic
    bool useGraph = true;
    float worldMin[3] = {0.f, 0.f, 0.f}, worldMax[3] = {0.f, 0.f, 0.f};

    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    pbat_vbd::GraphKey graphKey;

    bool timing = false;
    std::vector<hipEvent_t> events; // pairs (start, stop)
    size_t eventsUsed = 0;
    double timedMs = 0.;
    int64_t timedLaunches = 0;

    ~pbat_vbd_integrator()
    {
        if (graphExec)
            (void)hipGraphExecDestroy(graphExec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (hipEvent_t e : events)
            (void)hipEventDestroy(e);
        if (stream)
            (void)hipStreamSynchronize(stream);
        xm.reset(), xtil.reset(), xt.reset(), vel.reset(), vprev.reset(), aext.reset();
        tets.reset(), vtp.reset(), vtref.reset(), energy.reset();
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace pbat_vbd {
namespace {

int EffectiveLanes(pbat_vbd_integrator const& h)
{
    if (h.lanes > 0)
        return h.lanes;
    // enough lanes that each handles about 3 incidences of the busiest vertex
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    int L = 1;
    while (L < 64 && (d + L - 1) / L > 3)
        L <<= 1;
    return L;
}

template <int L>
void LaunchSweepL(pbat_vbd_integrator& h, device::SweepArgs const& a, int nVerts)
{
    int const block = h.blockSize;
    int64_t const threads = static_cast<int64_t>(nVerts) * L;
    int64_t const grid64 = (threads + block - 1) / block;
    // the kernel indexes threads with 32-bit ints
    if (grid64 * block >= (int64_t{1} << 31))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "colour too large for the sweep launch: reduce lanes per vertex");
    int const grid = static_cast<int>(grid64);
    hipLaunchKernelGGL(device::VbdSweep<L>, dim3(grid), dim3(block), 0, h.stream, a);
}

void LaunchSweep(pbat_vbd_integrator& h, device::SweepArgs const& a, int nVerts)
{
    switch (EffectiveLanes(h))
    {
        case 1: LaunchSweepL<1>(h, a, nVerts); break;
        case 2: LaunchSweepL<2>(h, a, nVerts); break;
        case 4: LaunchSweepL<4>(h, a, nVerts); break;
        case 8: LaunchSweepL<8>(h, a, nVerts); break;
        case 16: LaunchSweepL<16>(h, a, nVerts); break;
        case 32: LaunchSweepL<32>(h, a, nVerts); break;
        default: LaunchSweepL<64>(h, a, nVerts); break;
    }
}

hipEvent_t NextEvent(pbat_vbd_integrator& h)
{
    if (h.eventsUsed == h.events.size())
    {
        hipEvent_t e;
        PBAT_HIP(hipEventCreate(&e));
        h.events.push_back(e);
    }
    return h.events[h.eventsUsed++];
}

void CollectTiming(pbat_vbd_integrator& h)
{
    if (h.eventsUsed == 0)
        return;
    PBAT_HIP(hipStreamSynchronize(h.stream));
    for (size_t k = 0; k + 1 < h.eventsUsed; k += 2)
    {
        float ms = 0.f;
        PBAT_HIP(hipEventElapsedTime(&ms, h.events[k], h.events[k + 1]));
        h.timedMs += ms;
        ++h.timedLaunches;
    }
    h.eventsUsed = 0;
}

// Enqueue the launches of one Step (gpu/impl/vbd/Integrator.cu:82-103).
void EnqueueStep(pbat_vbd_integrator& h, float dt, int iterations, int substeps)
{
    float const sdt = dt / static_cast<float>(substeps);
    float const sdt2 = sdt * sdt;
    int const n = static_cast<int>(h.nV);
    int const eb = 256, eg = (n + eb - 1) / eb;
    device::SweepArgs a{};
    a.vtp = h.vtp.ptr;
    a.vtref = h.vtref.ptr;
    a.tets = h.tets.ptr;
    a.xm = h.xm.ptr;
    a.xtil = h.xtil.ptr;
    a.xt = h.xt.ptr;
    a.dt = sdt;
    a.dt2 = sdt2;
    a.kD = h.kD;
    a.detHZero = h.detHZero;
    int const nColors = static_cast<int>(h.colorBegin.size()) - 1;
    for (int s = 0; s < substeps; ++s)
    {
        hipLaunchKernelGGL(device::SubstepBegin, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xtil.ptr, h.xt.ptr,
                           h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, sdt, sdt2, h.strategy);
        for (int k = 0; k < iterations; ++k)
            for (int c = 0; c < nColors; ++c)
            {
                a.vbeg = h.colorBegin[c];
                a.vend = h.colorBegin[c + 1];
                int const nc = a.vend - a.vbeg;
                if (nc <= 0)
                    continue;
                if (h.timing)
                    PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
                LaunchSweep(h, a, nc);
                if (h.timing)
                    PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
            }
        hipLaunchKernelGGL(device::SubstepEnd, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xt.ptr, h.vel.ptr,
                           h.vprev.ptr, n, sdt);
    }
    PBAT_HIP(hipGetLastError());
}

void InvalidateGraph(pbat_vbd_integrator& h)
{
    if (h.graphExec)
        (void)hipGraphExecDestroy(h.graphExec);
    if (h.graph)
        (void)hipGraphDestroy(h.graph);
    h.graphExec = nullptr;
    h.graph = nullptr;
    h.graphKey = GraphKey{};
}

void Step(pbat_vbd_integrator& h, float dt, int iterations, int substeps)
{
    if (!(dt > 0.f) || !std::isfinite(dt))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (iterations < 0 || substeps < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "expected iterations >= 0 and substeps >= 1");
    PBAT_HIP(hipSetDevice(h.device));
    if (h.timing || !h.useGraph)
    {
        EnqueueStep(h, dt, iterations, substeps);
        if (h.timing)
            CollectTiming(h);
        PBAT_HIP(hipStreamSynchronize(h.stream));
        return;
    }
    GraphKey key;
    key.dt = dt, key.iterations = iterations, key.substeps = substeps, key.strategy = h.strategy;
    key.lanes = EffectiveLanes(h), key.block = h.blockSize, key.kD = h.kD, key.detHZero = h.detHZero;
    if (!(h.graphExec && h.graphKey == key))
    {
        InvalidateGraph(h);
        PBAT_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
        try
        {
            EnqueueStep(h, dt, iterations, substeps);
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(h.stream, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        PBAT_HIP(hipStreamEndCapture(h.stream, &h.graph));
        PBAT_HIP(hipGraphInstantiate(&h.graphExec, h.graph, nullptr, nullptr, 0));
        h.graphKey = key;
    }
    PBAT_HIP(hipGraphLaunch(h.graphExec, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

void CheckVertexArray(pbat_vbd_integrator const& h, void const* p, int64_t n, char const* what)
{
    if (!p)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, std::string(what) + " is null");
    if (n != h.nV)
    {
        std::ostringstream ss;
        ss << "Expected input dimensions 3x" << h.nV << ", but got 3x" << n << " for " << what;
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, ss.str());
    }
}

// Copy a 3 x nV host matrix (original numbering) into a float4 device array
// (internal numbering); w components come from wsrc (nullptr -> 0).
void UploadVec(pbat_vbd_integrator& h, float4* dst, float const* src, float const* wsrc)
{
    std::vector<float4> buf(static_cast<size_t>(h.nV));
    for (int64_t v = 0; v < h.nV; ++v)
    {
        int64_t const i = h.perm[v];
        buf[v] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], wsrc ? wsrc[i] : 0.f);
    }
    PBAT_HIP(hipMemcpyAsync(dst, buf.data(), buf.size() * sizeof(float4), hipMemcpyHostToDevice, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

void DownloadVec(pbat_vbd_integrator& h, float4 const* src, float* dst)
{
    std::vector<float4> buf(static_cast<size_t>(h.nV));
    PBAT_HIP(hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(float4), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    for (int64_t v = 0; v < h.nV; ++v)
    {
        int64_t const i = h.perm[v];
        dst[3 * i] = buf[v].x;
        dst[3 * i + 1] = buf[v].y;
        dst[3 * i + 2] = buf[v].z;
    }
}

void Create(pbat_vbd_data const* d, pbat_vbd_integrator** out)
{
    if (!d || !out)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (d->nV <= 0 || d->nT <= 0)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "empty mesh");
    if (d->nV > (int64_t{1} << 30) || d->nT >= (int64_t{1} << 29))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "mesh too large for 32-bit device indexing");
    if (!d->x || !d->v || !d->aext || !d->m || !d->E || !d->GP || !d->wg || !d->lame || !d->GVGp || !d->GVGe ||
        !d->GVGilocal || !d->Pptr || (d->nPartitions > 0 && !d->Padj))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null data array");
    if (d->nPartitions < 0 || d->nPartitions > PBAT_VBD_MAX_COLORS)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid number of partitions");
    if (d->strategy < 0 || d->strategy > 4)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid initialization strategy");
    int64_t const nV = d->nV, nT = d->nT;
    if (d->GVGp[0] != 0)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGp must start at 0");
    for (int64_t i = 0; i < nV; ++i)
        if (d->GVGp[i + 1] < d->GVGp[i])
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGp must be non-decreasing");
    int64_t const nInc = d->GVGp[nV];
    for (int64_t k = 0; k < nInc; ++k)
        if (d->GVGe[k] < 0 || d->GVGe[k] >= nT || d->GVGilocal[k] < 0 || d->GVGilocal[k] > 3)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGe/GVGilocal entry out of range");
    for (int64_t k = 0; k < 4 * nT; ++k)
        if (d->E[k] < 0 || d->E[k] >= nV)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "element vertex index out of range");
    if (d->Pptr[0] != 0)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Pptr must start at 0");
    for (int64_t p = 0; p < d->nPartitions; ++p)
        if (d->Pptr[p + 1] < d->Pptr[p])
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Pptr must be non-decreasing");
    int64_t const nUpd = d->Pptr[d->nPartitions];
    if (nUpd > nV)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "more partition entries than vertices");

    auto h = std::make_unique<pbat_vbd_integrator>();
    PBAT_HIP(hipGetDevice(&h->device));
    PBAT_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->nV = nV, h->nT = nT, h->nIncidences = nInc, h->nUpdated = nUpd;
    h->strategy = d->strategy;
    h->kD = static_cast<float>(d->kD);
    h->detHZero = static_cast<float>(d->detHZero);

    // Colour-major internal numbering; degree-sorted (stable) within a colour.
    std::vector<char> seen(static_cast<size_t>(nV), 0);
    h->perm.reserve(static_cast<size_t>(nV));
    h->colorBegin.push_back(0);
    for (int64_t p = 0; p < d->nPartitions; ++p)
    {
        std::vector<int> part;
        for (int64_t k = d->Pptr[p]; k < d->Pptr[p + 1]; ++k)
        {
            int64_t const i = d->Padj[k];
            if (i < 0 || i >= nV || seen[i])
                Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Padj entry out of range or repeated");
            seen[i] = 1;
            part.push_back(static_cast<int>(i));
        }
        std::stable_sort(part.begin(), part.end(), [&](int a, int b) {
            return (d->GVGp[a + 1] - d->GVGp[a]) > (d->GVGp[b + 1] - d->GVGp[b]);
        });
        h->perm.insert(h->perm.end(), part.begin(), part.end());
        h->colorBegin.push_back(static_cast<int>(h->perm.size()));
    }
    for (int64_t i = 0; i < nV; ++i)
        if (!seen[i])
            h->perm.push_back(static_cast<int>(i));
    h->iperm.assign(static_cast<size_t>(nV), 0);
    for (int64_t v = 0; v < nV; ++v)
        h->iperm[h->perm[v]] = static_cast<int>(v);

    // Validate the colouring: no element may hold two vertices of one colour.
    {
        std::vector<int> colorOf(static_cast<size_t>(nV), -1);
        for (size_t c = 0; c + 1 < h->colorBegin.size(); ++c)
            for (int v = h->colorBegin[c]; v < h->colorBegin[c + 1]; ++v)
                colorOf[h->perm[v]] = static_cast<int>(c);
        for (int64_t e = 0; e < nT; ++e)
            for (int a = 0; a < 4; ++a)
                for (int b = a + 1; b < 4; ++b)
                {
                    int const ca = colorOf[d->E[4 * e + a]], cb = colorOf[d->E[4 * e + b]];
                    if (ca >= 0 && ca == cb)
                        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                             "partitions are not independent sets: element " + std::to_string(e) +
                                 " has two vertices in partition " + std::to_string(ca));
                }
    }

    // Tet records
    std::vector<device::TetRecord> T(static_cast<size_t>(nT));
    for (int64_t e = 0; e < nT; ++e)
    {
        double const* G = d->GP + 12 * e; // G(r, c) = G[4c + r]
        auto gp = [&](int r, int c) { return static_cast<float>(G[4 * c + r]); };
        double const w = d->wg[e], mu = d->lame[2 * e], lam = d->lame[2 * e + 1];
        device::TetRecord& R = T[e];
        R.v = make_int4(h->iperm[d->E[4 * e]], h->iperm[d->E[4 * e + 1]], h->iperm[d->E[4 * e + 2]],
                        h->iperm[d->E[4 * e + 3]]);
        R.q0 = make_float4(gp(1, 0), gp(1, 1), gp(1, 2), gp(2, 0));
        R.q1 = make_float4(gp(2, 1), gp(2, 2), gp(3, 0), gp(3, 1));
        R.q2 = make_float4(gp(3, 2), static_cast<float>(w * mu), static_cast<float>(w * lam),
                           static_cast<float>(1. + mu / lam));
    }
    // CSR in internal numbering
    std::vector<int> P(static_cast<size_t>(nV) + 1, 0), R;
    R.reserve(static_cast<size_t>(nInc));
    for (int64_t v = 0; v < nV; ++v)
    {
        int64_t const i = h->perm[v];
        h->maxDegree = std::max(h->maxDegree, d->GVGp[i + 1] - d->GVGp[i]);
        for (int64_t k = d->GVGp[i]; k < d->GVGp[i + 1]; ++k)
            R.push_back(static_cast<int>((d->GVGe[k] << 2) | d->GVGilocal[k]));
        P[v + 1] = static_cast<int>(R.size());
    }
    // maximum degree over updated vertices decides the default lanes per vertex
    h->maxDegree = 0;
    for (int64_t v = 0; v < nUpd; ++v)
        h->maxDegree = std::max<int64_t>(h->maxDegree, P[v + 1] - P[v]);

    h->tets.upload(T, h->stream);
    h->vtp.upload(P, h->stream);
    h->vtref.upload(R, h->stream);
    for (auto* b : {&h->xm, &h->xtil, &h->xt, &h->vel, &h->vprev, &h->aext})
        b->allocate(static_cast<size_t>(nV));
    h->energy.allocate(3);
    PBAT_HIP(hipMemsetAsync(h->xtil.ptr, 0, sizeof(float4) * nV, h->stream));
    PBAT_HIP(hipMemsetAsync(h->vprev.ptr, 0, sizeof(float4) * nV, h->stream));

    std::vector<float> x32(static_cast<size_t>(3 * nV)), v32(x32.size()), a32(x32.size());
    h->massOrig.resize(static_cast<size_t>(nV));
    for (int64_t k = 0; k < 3 * nV; ++k)
    {
        x32[k] = static_cast<float>(d->x[k]);
        v32[k] = static_cast<float>(d->v[k]);
        a32[k] = static_cast<float>(d->aext[k]);
    }
    for (int64_t i = 0; i < nV; ++i)
        h->massOrig[i] = static_cast<float>(d->m[i]);
    UploadVec(*h, h->xm.ptr, x32.data(), h->massOrig.data());
    UploadVec(*h, h->xt.ptr, x32.data(), nullptr);
    UploadVec(*h, h->vel.ptr, v32.data(), nullptr);
    UploadVec(*h, h->aext.ptr, a32.data(), nullptr);
    PBAT_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
}

void Energy(pbat_vbd_integrator& h, double out[3])
{
    PBAT_HIP(hipSetDevice(h.device));
    hipLaunchKernelGGL(device::ZeroDoubles, dim3(1), dim3(64), 0, h.stream, h.energy.ptr, 3);
    int const n = static_cast<int>(std::max(h.nV, h.nT));
    int const grid = std::min(4096, (n + 255) / 256);
    hipLaunchKernelGGL(device::EnergyTerms, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr, h.vel.ptr, h.aext.ptr,
                       h.tets.ptr, static_cast<int>(h.nV), static_cast<int>(h.nT), h.energy.ptr);
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipMemcpyAsync(out, h.energy.ptr, 3 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

// ---- MatrixMarket dense export (gpu/impl/vbd/Integrator.cu:105-148, 284-301) ----
void WriteDense(std::string const& path, int64_t rows, int64_t cols, std::vector<double> const& colMajor)
{
    std::ofstream f(path);
    if (!f)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "cannot open " + path + " for writing");
    f << "%%MatrixMarket matrix array real general\n" << rows << " " << cols << "\n";
    f << std::setprecision(9);
    for (double v : colMajor)
        f << v << "\n";
}

// nV x 3 matrix of a float4 device array (original vertex order)
void WriteVertexMatrix(pbat_vbd_integrator& h, float4 const* src, std::string const& path)
{
    std::vector<float> buf(static_cast<size_t>(3 * h.nV));
    DownloadVec(h, src, buf.data());
    std::vector<double> M(static_cast<size_t>(3 * h.nV));
    for (int64_t i = 0; i < h.nV; ++i)
        for (int d = 0; d < 3; ++d)
            M[d * h.nV + i] = buf[3 * i + d];
    WriteDense(path, h.nV, 3, M);
}

std::string Join(char const* dir, std::string const& name)
{
    std::string d = (dir && *dir) ? dir : ".";
    return d + "/" + name;
}

void TracedStep(pbat_vbd_integrator& h, float dt, int iterations, int substeps, int t, char const* dir)
{
    if (!(dt > 0.f) || iterations < 0 || substeps < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid step arguments");
    PBAT_HIP(hipSetDevice(h.device));
    auto name = [](char const* var, int tt, int s, int k) {
        std::ostringstream ss;
        ss << var << ".t." << tt << ".s." << s;
        if (k >= 0)
            ss << ".k." << k;
        ss << ".mtx";
        return ss.str();
    };
    if (t == 0)
    {
        // Element and vertex data of the first frame, in original numbering
        std::vector<device::TetRecord> T(static_cast<size_t>(h.nT));
        PBAT_HIP(hipMemcpy(T.data(), h.tets.ptr, T.size() * sizeof(device::TetRecord), hipMemcpyDeviceToHost));
        std::vector<double> Tm(static_cast<size_t>(4 * h.nT)), M(h.massOrig.begin(), h.massOrig.end());
        for (int64_t e = 0; e < h.nT; ++e)
        {
            int const ids[4] = {T[e].v.x, T[e].v.y, T[e].v.z, T[e].v.w};
            for (int k = 0; k < 4; ++k)
                Tm[k * h.nT + e] = h.perm[ids[k]];
        }
        WriteDense(Join(dir, "T.mtx"), h.nT, 4, Tm);
        WriteDense(Join(dir, "M.mtx"), h.nV, 1, M);
    }
    float const sdt = dt / static_cast<float>(substeps);
    float const sdt2 = sdt * sdt;
    int const n = static_cast<int>(h.nV);
    int const eb = 256, eg = (n + eb - 1) / eb;
    int const nColors = static_cast<int>(h.colorBegin.size()) - 1;
    device::SweepArgs a{};
    a.vtp = h.vtp.ptr, a.vtref = h.vtref.ptr, a.tets = h.tets.ptr, a.xm = h.xm.ptr, a.xtil = h.xtil.ptr;
    a.xt = h.xt.ptr, a.dt = sdt, a.dt2 = sdt2, a.kD = h.kD, a.detHZero = h.detHZero;
    for (int s = 0; s < substeps; ++s)
    {
        hipLaunchKernelGGL(device::SubstepBegin, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xtil.ptr, h.xt.ptr,
                           h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, sdt, sdt2, h.strategy);
        WriteVertexMatrix(h, h.xtil.ptr, Join(dir, name("xtilde", t, s, -1)));
        for (int k = 0; k < iterations; ++k)
        {
            WriteVertexMatrix(h, h.xm.ptr, Join(dir, name("x", t, s, k)));
            for (int c = 0; c < nColors; ++c)
            {
                a.vbeg = h.colorBegin[c];
                a.vend = h.colorBegin[c + 1];
                if (a.vend > a.vbeg)
                    LaunchSweep(h, a, a.vend - a.vbeg);
            }
        }
        WriteVertexMatrix(h, h.xm.ptr, Join(dir, name("x", t, s, iterations)));
        hipLaunchKernelGGL(device::SubstepEnd, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xt.ptr, h.vel.ptr,
                           h.vprev.ptr, n, sdt);
    }
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

pbat_vbd_integrator& Handle(pbat_vbd_integrator* h)
{
    if (!h)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null integrator handle");
    return *h;
}

} // namespace
} // namespace pbat_vbd

using namespace pbat_vbd;

extern "C" int pbat_vbd_integrator_create(const pbat_vbd_data* data, pbat_vbd_integrator** out)
{
    return Guard([&] { Create(data, out); });
}

extern "C" int pbat_vbd_integrator_destroy(pbat_vbd_integrator* h)
{
    return Guard([&] {
        if (h)
        {
            (void)hipSetDevice(h->device);
            delete h;
        }
    });
}

extern "C" int pbat_vbd_integrator_step(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t substeps)
{
    return Guard([&] { Step(Handle(h), dt, iterations, substeps); });
}

extern "C" int pbat_vbd_integrator_traced_step(pbat_vbd_integrator* h, float dt, int32_t iterations,
                                               int32_t substeps, int32_t t, const char* dir)
{
    return Guard([&] { TracedStep(Handle(h), dt, iterations, substeps, t, dir); });
}

extern "C" int pbat_vbd_integrator_set_positions(pbat_vbd_integrator* h, const float* x, int64_t n)
{
    return Guard([&] {
        auto& H = Handle(h);
        CheckVertexArray(H, x, n, "positions");
        PBAT_HIP(hipSetDevice(H.device));
        UploadVec(H, H.xm.ptr, x, H.massOrig.data());
        UploadVec(H, H.xt.ptr, x, nullptr);
    });
}

extern "C" int pbat_vbd_integrator_set_velocities(pbat_vbd_integrator* h, const float* v, int64_t n)
{
    return Guard([&] {
        auto& H = Handle(h);
        CheckVertexArray(H, v, n, "velocities");
        PBAT_HIP(hipSetDevice(H.device));
        UploadVec(H, H.vel.ptr, v, nullptr);
        UploadVec(H, H.vprev.ptr, v, nullptr);
    });
}

extern "C" int pbat_vbd_integrator_set_external_acceleration(pbat_vbd_integrator* h, const float* a, int64_t n)
{
    return Guard([&] {
        auto& H = Handle(h);
        CheckVertexArray(H, a, n, "external acceleration");
        PBAT_HIP(hipSetDevice(H.device));
        UploadVec(H, H.aext.ptr, a, nullptr);
    });
}

extern "C" int pbat_vbd_integrator_get_positions(pbat_vbd_integrator* h, float* x, int64_t n)
{
    return Guard([&] {
        auto& H = Handle(h);
        CheckVertexArray(H, x, n, "positions output");
        PBAT_HIP(hipSetDevice(H.device));
        DownloadVec(H, H.xm.ptr, x);
    });
}

extern "C" int pbat_vbd_integrator_get_velocities(pbat_vbd_integrator* h, float* v, int64_t n)
{
    return Guard([&] {
        auto& H = Handle(h);
        CheckVertexArray(H, v, n, "velocities output");
        PBAT_HIP(hipSetDevice(H.device));
        DownloadVec(H, H.vel.ptr, v);
    });
}

extern "C" int pbat_vbd_integrator_set_detH_zero(pbat_vbd_integrator* h, float zero)
{
    return Guard([&] { Handle(h).detHZero = zero; });
}

extern "C" int pbat_vbd_integrator_set_rayleigh_damping(pbat_vbd_integrator* h, float kD)
{
    return Guard([&] { Handle(h).kD = kD; });
}

extern "C" int pbat_vbd_integrator_set_initialization_strategy(pbat_vbd_integrator* h, int32_t strategy)
{
    return Guard([&] {
        if (strategy < 0 || strategy > 4)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid initialization strategy");
        Handle(h).strategy = strategy;
    });
}

extern "C" int pbat_vbd_integrator_set_block_size(pbat_vbd_integrator* h, int32_t block_size)
{
    return Guard([&] {
        int b = std::clamp<int>(block_size, 64, 256);
        Handle(h).blockSize = (b / 64) * 64;
    });
}

extern "C" int pbat_vbd_integrator_set_scene_bounding_box(pbat_vbd_integrator* h, const float min[3],
                                                          const float max[3])
{
    return Guard([&] {
        auto& H = Handle(h);
        if (!min || !max)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null bounding box");
        for (int d = 0; d < 3; ++d)
            H.worldMin[d] = min[d], H.worldMax[d] = max[d];
    });
}

extern "C" int pbat_vbd_integrator_set_lanes_per_vertex(pbat_vbd_integrator* h, int32_t lanes)
{
    return Guard([&] {
        if (lanes != 0 && (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "lanes per vertex must be 0 or a power of two in [1, 64]");
        Handle(h).lanes = lanes;
    });
}

extern "C" int pbat_vbd_integrator_set_graph_capture(pbat_vbd_integrator* h, int32_t enabled)
{
    return Guard([&] {
        auto& H = Handle(h);
        H.useGraph = enabled != 0;
        if (!H.useGraph)
            InvalidateGraph(H);
    });
}

extern "C" int pbat_vbd_integrator_energy(pbat_vbd_integrator* h, double out[3])
{
    return Guard([&] {
        if (!out)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null output");
        Energy(Handle(h), out);
    });
}

extern "C" int pbat_vbd_integrator_synchronize(pbat_vbd_integrator* h)
{
    return Guard([&] { PBAT_HIP(hipStreamSynchronize(Handle(h).stream)); });
}

extern "C" int pbat_vbd_integrator_counts(pbat_vbd_integrator* h, int64_t* nPartitions, int64_t* nUpdated,
                                          int64_t* nIncidences)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (nPartitions)
            *nPartitions = static_cast<int64_t>(H.colorBegin.size()) - 1;
        if (nUpdated)
            *nUpdated = H.nUpdated;
        if (nIncidences)
        {
            int64_t s = 0;
            std::vector<int> P(static_cast<size_t>(H.nUpdated) + 1);
            PBAT_HIP(hipMemcpy(P.data(), H.vtp.ptr, P.size() * sizeof(int), hipMemcpyDeviceToHost));
            s = P.back() - P.front();
            *nIncidences = s;
        }
    });
}

extern "C" int pbat_vbd_integrator_set_timing(pbat_vbd_integrator* h, int32_t enabled)
{
    return Guard([&] { Handle(h).timing = enabled != 0; });
}

extern "C" int pbat_vbd_integrator_sweep_timing(pbat_vbd_integrator* h, double* total_ms, int64_t* launches,
                                                int32_t reset)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (total_ms)
            *total_ms = H.timedMs;
        if (launches)
            *launches = H.timedLaunches;
        if (reset)
            H.timedMs = 0., H.timedLaunches = 0;
    });
}
