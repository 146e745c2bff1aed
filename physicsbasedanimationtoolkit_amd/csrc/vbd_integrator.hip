// MI355X VBD integrator: host orchestration behind the C ABI of pbat_vbd.h.
//
// Counterpart of pbat::gpu::impl::vbd::Integrator (gpu/impl/vbd/Integrator.cu:22-376)
// with a different device design:
//  - vertices are renumbered colour-major, so each colour is one contiguous
//    range and a sweep over it reads per-vertex data coalesced; within a colour
//    vertices are grouped by mesh (batches) and sorted by degree (stable) so
//    lanes of a wave do equal work;
//  - each vertex-tet incidence is a 56-byte record stored in sweep order: the
//    tet's other three vertex ids, their three shape-function-gradient rows
//    and the tet's w*mu, w*lambda (alpha = 1 + mu/lambda is derived in the
//    kernel). A colour pass streams its records
//    (instead of gathering per-tet data that straddles cache lines), and only
//    the neighbour positions are gathered, one dependent load after the record;
//  - positions carry the lumped mass in their 4th component, so one 16-byte
//    gather serves both;
//  - workgroups are reordered so each XCD's L2 serves one contiguous range of
//    vertices; the sweep shape (lanes per vertex, incidences prefetched per
//    lane, XCD order, non-temporal stream loads) is chosen by an autotuner;
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
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "integrator.hpp"

namespace pbat_vbd {

int EffectiveLanes(pbat_vbd_integrator const& h)
{
    if (h.lanes > 0)
        return h.lanes;
    // enough lanes that each handles about 2 incidences of the busiest vertex
    // (degree 24: 16 lanes x 2, the shape the autotuner picks on the 64^3 grid)
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    int L = 1;
    while (L < 64 && (d + L - 1) / L > 2)
        L <<= 1;
    return L;
}

// Incidences each lane prefetches before accumulating (1..4): enough to cover
// the busiest updated vertex, capped to bound register use.
int EffectivePerLane(pbat_vbd_integrator const& h)
{
    if (h.perLane > 0)
        return h.perLane;
    int const L = EffectiveLanes(h);
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    return std::clamp((d + L - 1) / L, 1, 4);
}

template <int L, int K>
void LaunchSweepLK(pbat_vbd_integrator& h, device::SweepArgs a, int nVerts)
{
    int const vbeg = a.vbeg;
    int const block = h.blockSize;
    int64_t const threads = static_cast<int64_t>(nVerts) * L;
    int64_t grid64 = (threads + block - 1) / block;
    a.xcdChunk = 0;
    if (h.xcdRemap && grid64 >= 16)
    {
        // pad to a multiple of the 8 XCDs; padded blocks hold no vertex
        grid64 = (grid64 + 7) / 8 * 8;
        a.xcdChunk = static_cast<int>(grid64 / 8);
    }
    // the kernel indexes threads with 32-bit ints
    if (grid64 * block >= (int64_t{1} << 31))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "colour too large for the sweep launch: reduce lanes per vertex");
    if (!h.contact)
    {
        hipLaunchKernelGGL((device::VbdSweep<L, K>), dim3(static_cast<int>(grid64)), dim3(block), 0, h.stream, a);
        return;
    }
    // gpu/impl/vbd/Integrator.cu:303-326: the colour writes a buffer; the next colour's
    // launch reads the pending range there and copies it back
    // the same colour again (a single colour), or a pending range larger than this
    // launch can copy: copy it back first
    if ((h.pendBeg == vbeg && h.pendEnd > h.pendBeg) || static_cast<int64_t>(h.pendEnd - h.pendBeg) > grid64 * block)
        FlushContactSweep(h);
    a.contact.rec = h.crec.ptr, a.contact.xb = h.xb.ptr, a.contact.muF = h.muF, a.contact.epsv = h.epsv;
    a.contact.pendBeg = h.pendBeg, a.contact.pendEnd = h.pendEnd;
    hipLaunchKernelGGL((device::VbdSweep<L, K, true>), dim3(static_cast<int>(grid64)), dim3(block), 0, h.stream, a);
    h.pendBeg = vbeg, h.pendEnd = vbeg + nVerts;
}

template <int L>
void LaunchSweepL(pbat_vbd_integrator& h, device::SweepArgs const& a, int nVerts)
{
    switch (EffectivePerLane(h))
    {
        case 1: LaunchSweepLK<L, 1>(h, a, nVerts); break;
        case 2: LaunchSweepLK<L, 2>(h, a, nVerts); break;
        case 3: LaunchSweepLK<L, 3>(h, a, nVerts); break;
        default: LaunchSweepLK<L, 4>(h, a, nVerts); break;
    }
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

device::SweepArgs MakeSweepArgs(pbat_vbd_integrator& h, float sdt, float sdt2)
{
    device::SweepArgs a{};
    pbat_vbd::IncidenceRecords const& r = (h.useEll && h.ellStride > 0) ? h.ell : h.rows;
    a.vtp = h.vtp.ptr;
    a.ids = r.sids.ptr;
    a.geo = r.sgeo.ptr;
    a.xm = h.xm.ptr;
    a.xtil = h.xtil.ptr;
    a.xt = h.xt.ptr;
    a.c.dt = sdt;
    a.c.dt2 = sdt2;
    a.c.kD = h.kD;
    a.c.detHZero = h.detHZero;
    a.c.fp64Solves = h.fp64Solves.ptr;
    a.ntStream = h.ntStream ? 1 : 0;
    if (h.useEll && h.ellStride > 0)
        a.ellStride = h.ellStride;
    return a;
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

// Sum the event pairs recorded by the last Step and release them.
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
        h.timedLaunches += h.pendingLaunches[k / 2];
    }
    h.eventsUsed = 0;
    h.pendingLaunches.clear();
}

// Enqueue the launches of one Step (gpu/impl/vbd/Integrator.cu:82-103) and
// return the number of sweep launches. With eventBrackets, one event pair
// brackets each substep's back-to-back sweep launches.
// sim/vbd/Kernels.h:97-102 in FP32 as gpu/impl/vbd/ChebyshevIntegrator.cu:28-33
// evaluates it, operator precedence included: for k >= 2, 4/4 - rho2 * omega.
float ChebyshevOmega(int k, float rho2, float omega)
{
    return (k == 0) ? 1.f : (k == 1) ? 2.f / (2.f - rho2) : 4.f / 4.f - rho2 * omega;
}

// Before iteration 0 of a substep (Anderson: remember x).
void BeginAcceleration(pbat_vbd_integrator& h)
{
    if (h.acceleration != PBAT_VBD_ACCELERATION_ANDERSON)
        return;
    int const n = static_cast<int>(h.nV);
    hipLaunchKernelGGL(device::AndersonStart, dim3((n + 255) / 256), dim3(256), 0, h.stream, h.xm.ptr, h.xkm1.ptr,
                       n, h.andersonAcc.ptr);
}

template <int M>
void LaunchAndersonWindow(pbat_vbd_integrator& h, int n, int col, int mk, double* acc)
{
    hipLaunchKernelGGL((device::AndersonWindow<M>), dim3((n + 255) / 256), dim3(256), 0, h.stream, h.xm.ptr,
                       h.xkm1.ptr, h.gkm1.ptr, h.fkm1.ptr, h.dgk.ptr, h.dfk.ptr, n, col, mk, acc);
}

// After iteration k of a substep: the Chebyshev update
// (gpu/impl/vbd/ChebyshevIntegrator.cu:24-35) or the Anderson window update
// and accelerated step (gpu/impl/vbd/AndersonIntegrator.cu:36-52); no-op
// without acceleration.
void LaunchAcceleration(pbat_vbd_integrator& h, int k, float& omega)
{
    int const n = static_cast<int>(h.nV);
    int const grid = (n + 255) / 256;
    if (h.acceleration == PBAT_VBD_ACCELERATION_CHEBYSHEV)
    {
        omega = ChebyshevOmega(k, h.rho * h.rho, omega);
        hipLaunchKernelGGL(device::ChebyshevUpdate, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr, h.xkm1.ptr,
                           h.xkm2.ptr, n, k, omega);
        return;
    }
    if (h.acceleration != PBAT_VBD_ACCELERATION_ANDERSON)
        return;
    if (k == 0)
    {
        hipLaunchKernelGGL(device::AndersonFirst, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr, h.xkm1.ptr,
                           h.gkm1.ptr, h.fkm1.ptr, n);
        return;
    }
    int const m = h.window;
    int const col = (k - 1) % m, mk = std::min(m, k);
    size_t const setDoubles = static_cast<size_t>(device::AndersonAccDoubles());
    double* acc = h.andersonAcc.ptr + (k & 1) * setDoubles;
    double* accNext = h.andersonAcc.ptr + ((k + 1) & 1) * setDoubles;
    if (m <= 2)
        LaunchAndersonWindow<2>(h, n, col, mk, acc);
    else if (m <= 4)
        LaunchAndersonWindow<4>(h, n, col, mk, acc);
    else if (m <= 8)
        LaunchAndersonWindow<8>(h, n, col, mk, acc);
    else
        LaunchAndersonWindow<16>(h, n, col, mk, acc);
    size_t const gram = static_cast<size_t>(m) * m;
    hipLaunchKernelGGL(device::AndersonApply, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr, h.xkm1.ptr,
                       static_cast<float4 const*>(h.gkm1.ptr), static_cast<float4 const*>(h.dgk.ptr), n, m, col, mk,
                       static_cast<double const*>(acc), accNext,
                       static_cast<double const*>(h.andersonGram.ptr + ((k - 1) & 1) * gram),
                       h.andersonGram.ptr + (k & 1) * gram);
}

// Sweep every colour once (one VBD iteration); returns the number of launches.
// One sweep over every colour. With contact, the last colour's new positions are
// left in the write buffer when flush is false: the next sweep's first launch
// reads them there and copies them back, as between colours.
int SweepAllColours(pbat_vbd_integrator& h, device::SweepArgs& a, bool flush = true)
{
    int const nColors = static_cast<int>(h.colorBegin.size()) - 1;
    int launched = 0;
    for (int c = 0; c < nColors; ++c)
    {
        a.vbeg = h.colorBegin[c];
        a.vend = h.colorBegin[c + 1];
        int const nc = a.vend - a.vbeg;
        if (nc <= 0)
            continue;
        LaunchSweep(h, a, nc);
        ++launched;
    }
    if (flush)
        FlushContactSweep(h);
    return launched;
}

constexpr int kObjectiveBlocks = 512;

// (sum m |x - x~|^2, sum w Psi(F), sum |x - xprev|^2) at the current iterate into
// the device sums, in FP64 with a fixed reduction order; skipped when gate points
// at 0. The trust-region update e runs after the sums (or in their place).
void ObjectiveTermsOnDevice(pbat_vbd_integrator& h, float4 const* xprev, double const* gate,
                            device::TrEpilogue const& e)
{
    double* part = h.trPartials.ptr;
    double* sums = part + 3 * kObjectiveBlocks;
    hipLaunchKernelGGL(device::ObjectiveTerms, dim3(kObjectiveBlocks), dim3(256), 0, h.stream,
                       static_cast<float4 const*>(h.xm.ptr), static_cast<float4 const*>(h.xtil.ptr), xprev,
                       static_cast<int4 const*>(h.tetv.ptr), static_cast<float4 const*>(h.tgp.ptr),
                       static_cast<int>(h.nV), static_cast<int>(h.nT), part, gate);
    hipLaunchKernelGGL(device::SumPartials3, dim3(1), dim3(256), 0, h.stream, static_cast<double const*>(part),
                       kObjectiveBlocks, sums, gate, e);
}

// Trust-region log slots of one step: one per iteration k >= 2 of each substep.
int64_t TrLogSlots(int iterations, int substeps)
{
    return static_cast<int64_t>(std::max(iterations - 2, 0)) * std::max(substeps, 0);
}

// Device buffers of the linear trust-region path for a step of this shape
// (allocated before any capture).
void EnsureTrustRegionBuffers(pbat_vbd_integrator& h, int iterations, int substeps)
{
    if (!h.trState.ptr)
        h.trState.allocate(device::kTrSlots);
    size_t const slots = static_cast<size_t>(std::max<int64_t>(TrLogSlots(iterations, substeps), 1));
    if (h.trLogT.n < slots)
        h.trLogT.allocate(slots), h.trLogAccepted.allocate(slots);
}

// TrustRegionIntegrator::SolveWithLinearAcceleratedPath
// (gpu/impl/vbd/TrustRegionIntegrator.cu:48-165). f = 1/2 |x - x~|_M^2 + h^2 sum w Psi
// (:265-364 without the contact term: no active contacts on a single body). The
// first two iterations are plain VBD iterations; from k = 2, m(t) = a t^2 + b t + c
// through the last three (t, f) pairs, shifted so the current iterate sits at
// t = 0 (:452-468), picks t in [1, R / |dx|] along x(t) = x_k + t dx, dx the VBD
// step (:400-430); the step is kept if (f_k - f(x(t))) / (f_k - m(t)) > eta, else
// rolled back to the VBD iterate; R grows or shrinks by tau. The scalars are FP64
// on the device (the reference keeps them in FP32 and reads each reduction back);
// the objective after a step that is not tried, or not rolled back, is skipped.
// Logs (t, accepted) of substep `substep` from slot substep * (iterations - 2).
// Returns the number of sweep launches.
int TrustRegionSolve(pbat_vbd_integrator& h, device::SweepArgs& a, int iterations, float sdt, int substep)
{
    device::TrConstants const c{double(sdt) * double(sdt), h.tau * h.tau, h.eta, double(1e-6f),
                                double(std::numeric_limits<float>::min()), double(std::numeric_limits<float>::max())};
    double* st = h.trState.ptr;
    int const n = static_cast<int>(h.nV);
    int const grid = (n + 255) / 256;
    size_t const bytes = sizeof(float4) * static_cast<size_t>(h.nV);
    auto update = [&](int op) {
        device::TrEpilogue e{};
        e.op = op, e.s = st, e.c = c, e.logT = h.trLogT.ptr, e.logAccepted = h.trLogAccepted.ptr;
        return e;
    };
    int launched = 0;
    ObjectiveTermsOnDevice(h, nullptr, nullptr, update(device::kTrOpBegin));
    for (int k = 0; k < iterations; ++k)
    {
        if (k < 2)
        {
            launched += SweepAllColours(h, a);
            device::TrEpilogue e = update(device::kTrOpPlain);
            e.k = k;
            ObjectiveTermsOnDevice(h, nullptr, nullptr, e);
            continue;
        }
        hipLaunchKernelGGL(device::TrModel, dim3(1), dim3(64), 0, h.stream, st);
        PBAT_HIP(hipMemcpyAsync(h.xkm1.ptr, h.xm.ptr, bytes, hipMemcpyDeviceToDevice, h.stream));
        launched += SweepAllColours(h, a);
        ObjectiveTermsOnDevice(h, h.xkm1.ptr, nullptr, update(device::kTrOpDecide));
        hipLaunchKernelGGL(device::LinearPathStep, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr,
                           static_cast<float4 const*>(h.xkm1.ptr), n, static_cast<double const*>(st), 0);
        device::TrEpilogue accept = update(device::kTrOpAccept);
        accept.slot = substep * (iterations - 2) + (k - 2);
        ObjectiveTermsOnDevice(h, nullptr, st + device::kTrAttempt, accept);
        hipLaunchKernelGGL(device::LinearPathStep, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr,
                           static_cast<float4 const*>(h.xkm1.ptr), n, static_cast<double const*>(st), 1);
        ObjectiveTermsOnDevice(h, nullptr, st + device::kTrRevert, update(device::kTrOpRevert));
    }
    return launched;
}

bool LinearTrustRegion(pbat_vbd_integrator const& h)
{
    return h.acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION && !h.curved;
}

int EnqueueStep(pbat_vbd_integrator& h, float dt, int iterations, int substeps, bool eventBrackets)
{
    float const sdt = dt / static_cast<float>(substeps);
    float const sdt2 = sdt * sdt;
    int const n = static_cast<int>(h.nV);
    int const eb = 256, eg = (n + eb - 1) / eb;
    device::SweepArgs a = MakeSweepArgs(h, sdt, sdt2);
    h.trLogCount = LinearTrustRegion(h) ? TrLogSlots(iterations, substeps) : 0;
    int total = 0;
    if (h.contact)
        ContactBeginStep(h, dt);
    for (int s = 0; s < substeps; ++s)
    {
        hipLaunchKernelGGL(device::SubstepBegin, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xtil.ptr, h.xt.ptr,
                           h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, sdt, sdt2, h.strategy);
        if (h.contact && s % h.activeSetFrequency == 0)
            ContactNearestPass(h, 0);
        // Per-launch events would add their own overhead to every kernel;
        // the gaps between queued launches are small.
        int launched = 0;
        if (eventBrackets)
            PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
        float omega = 0.f;
        BeginAcceleration(h);
        if (LinearTrustRegion(h))
            launched += TrustRegionSolve(h, a, iterations, sdt, s);
        else
            for (int k = 0; k < iterations; ++k)
            {
                // plain VBD reads nothing between sweeps: the write buffer is
                // copied back after the last one
                bool const keep = h.acceleration == PBAT_VBD_ACCELERATION_NONE && k + 1 < iterations;
                launched += SweepAllColours(h, a, !keep);
                LaunchAcceleration(h, k, omega);
            }
        if (eventBrackets)
        {
            PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
            h.pendingLaunches.push_back(launched);
        }
        total += launched;
        hipLaunchKernelGGL(device::SubstepEnd, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xt.ptr, h.vel.ptr,
                           h.vprev.ptr, n, sdt, h.statusDev);
    }
    if (h.contact)
        ContactNearestPass(h, 1);
    PBAT_HIP(hipGetLastError());
    return total;
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
    h.graphSweeps = 0;
}

void ClearStatus(pbat_vbd_integrator& h) { __atomic_store_n(h.statusHost, 0, __ATOMIC_RELEASE); }

void CheckStatus(pbat_vbd_integrator& h)
{
    if (__atomic_load_n(h.statusHost, __ATOMIC_ACQUIRE) != 0)
        Fail(PBAT_VBD_ERROR_NONFINITE,
             "non-finite positions or velocities after the step (a diverged solve or a non-finite input state); "
             "the state is left as computed");
}

void Step(pbat_vbd_integrator& h, float dt, int iterations, int substeps, bool sync)
{
    sync = sync || h.timing; // timed launches are collected after a synchronisation
    if (!(dt > 0.f) || !std::isfinite(dt))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (iterations < 0 || substeps < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "expected iterations >= 0 and substeps >= 1");
    PBAT_HIP(hipSetDevice(h.device));
    if (LinearTrustRegion(h))
        EnsureTrustRegionBuffers(h, iterations, substeps);
    if (sync)
        ClearStatus(h);
    if (!h.useGraph)
    {
        InvalidateGraph(h);
        EnqueueStep(h, dt, iterations, substeps, h.timing);
        if (!sync)
            return;
        if (h.timing)
            CollectTiming(h);
        PBAT_HIP(hipStreamSynchronize(h.stream));
        CheckStatus(h);
        return;
    }
    GraphKey key;
    key.dt = dt, key.iterations = iterations, key.substeps = substeps, key.strategy = h.strategy;
    key.lanes = EffectiveLanes(h), key.block = h.blockSize, key.kD = h.kD, key.detHZero = h.detHZero;
    key.perLane = EffectivePerLane(h), key.xcd = h.xcdRemap, key.nt = h.ntStream, key.ell = h.useEll;
    key.acceleration = h.acceleration, key.rho = h.rho, key.window = h.window;
    if (!(h.graphExec && h.graphKey == key))
    {
        InvalidateGraph(h);
        PBAT_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
        try
        {
            h.graphSweeps = EnqueueStep(h, dt, iterations, substeps, false);
            PBAT_HIP(hipStreamEndCapture(h.stream, &h.graph));
            PBAT_HIP(hipGraphInstantiate(&h.graphExec, h.graph, nullptr, nullptr, 0));
        }
        catch (...)
        {
            hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(h.stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone)
            {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(h.stream, &g);
                if (g)
                    (void)hipGraphDestroy(g);
            }
            InvalidateGraph(h);
            throw;
        }
        h.graphKey = key;
    }
    // Timing a replay: one event pair around the graph launch. The bracket also
    // holds the substep begin/end kernels (two elementwise passes per substep).
    if (h.timing)
        PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
    PBAT_HIP(hipGraphLaunch(h.graphExec, h.stream));
    if (h.timing)
    {
        PBAT_HIP(hipEventRecord(NextEvent(h), h.stream));
        h.pendingLaunches.push_back(h.graphSweeps);
    }
    if (!sync)
        return;
    PBAT_HIP(hipStreamSynchronize(h.stream));
    if (h.timing)
        CollectTiming(h);
    CheckStatus(h);
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

// Aggregates of the adjoint's two-level preconditioner: each mesh's updated
// vertices binned by rest position into cubes of four mean edge lengths
// (enlarged so that the cubes spanning a mesh's box number about
// kMaxAggregatesPerMesh: its dense coarse inverse has 6 rows per aggregate).
constexpr int kMaxAggregatesPerMesh = 512;
constexpr double kAggregateEdges = 4.;

void BuildAggregates(pbat_vbd_integrator& h, pbat_vbd_data const* d, std::vector<int> const& meshOfOrig)
{
    int const nM = h.nMeshes;
    int64_t const nV = h.nV, nT = h.nT, nUpd = h.nUpdated;
    std::vector<double> lo(3 * nM, 1e300), hi(3 * nM, -1e300), edgeSum(nM, 0.);
    std::vector<int64_t> edgeCount(nM, 0), count(nM, 0);
    for (int64_t i = 0; i < nV; ++i)
    {
        int const m = meshOfOrig[i];
        for (int k = 0; k < 3; ++k)
        {
            lo[3 * m + k] = std::min(lo[3 * m + k], d->x[3 * i + k]);
            hi[3 * m + k] = std::max(hi[3 * m + k], d->x[3 * i + k]);
        }
    }
    for (int64_t e = 0; e < nT; ++e)
    {
        int const m = meshOfOrig[d->E[4 * e]];
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b)
            {
                double const* p = d->x + 3 * d->E[4 * e + a];
                double const* q = d->x + 3 * d->E[4 * e + b];
                edgeSum[m] += std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                                        (p[2] - q[2]) * (p[2] - q[2]));
                ++edgeCount[m];
            }
    }
    std::vector<double> cell(nM);
    for (int m = 0; m < nM; ++m)
    {
        double const mean = edgeCount[m] ? edgeSum[m] / double(edgeCount[m]) : 1.;
        double vol = 1.;
        for (int k = 0; k < 3; ++k)
            vol *= std::max(hi[3 * m + k] - lo[3 * m + k], mean);
        cell[m] = std::max(kAggregateEdges * mean, std::cbrt(vol / kMaxAggregatesPerMesh));
    }
    // bin updated vertices (internal order), ids contiguous per mesh
    std::vector<std::vector<std::pair<std::array<int64_t, 3>, int>>> keys(nM);
    for (int64_t v = 0; v < nUpd; ++v)
    {
        int64_t const i = h.perm[v];
        int const m = meshOfOrig[i];
        std::array<int64_t, 3> k;
        for (int a = 0; a < 3; ++a)
            k[a] = static_cast<int64_t>(std::floor((d->x[3 * i + a] - lo[3 * m + a]) / cell[m]));
        keys[m].push_back({k, static_cast<int>(v)});
    }
    std::vector<int> aggOf(static_cast<size_t>(nV), -1), aggMesh;
    h.aggBegin.assign(static_cast<size_t>(nM) + 1, 0);
    h.coarseOffset.assign(static_cast<size_t>(nM) + 1, 0);
    int next = 0;
    for (int m = 0; m < nM; ++m)
    {
        auto& km = keys[m];
        std::sort(km.begin(), km.end());
        h.aggBegin[m] = next;
        for (size_t j = 0; j < km.size(); ++j)
        {
            if (j == 0 || km[j].first != km[j - 1].first)
            {
                aggMesh.push_back(m);
                ++next;
            }
            aggOf[km[j].second] = next - 1;
        }
        int64_t const nc = device::kCoarseDofs * int64_t(next - h.aggBegin[m]);
        h.coarseOffset[m + 1] = h.coarseOffset[m] + nc * nc;
    }
    h.aggBegin[nM] = next;
    h.nAggregates = next;
    std::vector<int> ptr(static_cast<size_t>(next) + 1, 0), verts(static_cast<size_t>(nUpd));
    for (int64_t v = 0; v < nUpd; ++v)
        ++ptr[aggOf[v] + 1];
    for (int g = 0; g < next; ++g)
        ptr[g + 1] += ptr[g];
    {
        std::vector<int> fillp(ptr.begin(), ptr.end() - 1);
        for (int64_t v = 0; v < nUpd; ++v)
            verts[fillp[aggOf[v]]++] = static_cast<int>(v);
    }
    h.aggVertPtr.upload(ptr, h.stream);
    h.aggVerts.upload(verts, h.stream);
    std::vector<long long> off(h.coarseOffset.begin(), h.coarseOffset.end());
    h.aggOf.upload(aggOf, h.stream);
    h.aggMesh.upload(aggMesh, h.stream);
    h.meshAggBegin.upload(h.aggBegin, h.stream);
    h.meshCoarseOffset.upload(off, h.stream);
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

// vertexOffsets (batches): first vertex of each mesh in d's numbering, count + 1
// entries; vertices of one colour are then grouped by mesh.
void Create(pbat_vbd_data const* d, pbat_vbd_integrator** out, std::vector<int64_t> const* vertexOffsets = nullptr)
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
    if (d->acceleration != PBAT_VBD_ACCELERATION_NONE && d->acceleration != PBAT_VBD_ACCELERATION_CHEBYSHEV &&
        d->acceleration != PBAT_VBD_ACCELERATION_ANDERSON && d->acceleration != PBAT_VBD_ACCELERATION_TRUST_REGION)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
             "only the base, Chebyshev-, Anderson- and trust-region-accelerated VBD integrators are available");
    // sim/vbd/Data.cpp:294-302
    if (d->acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION && !(d->eta >= 0.))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Expected eta >= 0");
    if (d->acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION && !(d->tau > 1.))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Expected tau > 1");
    if (d->acceleration == PBAT_VBD_ACCELERATION_CHEBYSHEV && !(d->rho > 0. && d->rho < 1.))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Expected 0 < rho < 1");
    if (d->acceleration == PBAT_VBD_ACCELERATION_ANDERSON &&
        (d->window < 1 || d->window > PBAT_VBD_MAX_ANDERSON_WINDOW))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
             "Anderson window size must be in [1, " + std::to_string(PBAT_VBD_MAX_ANDERSON_WINDOW) + "]");
    int64_t const nV = d->nV, nT = d->nT;
    if (d->GVGp[0] != 0)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGp must start at 0");
    for (int64_t i = 0; i < nV; ++i)
        if (d->GVGp[i + 1] < d->GVGp[i])
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGp must be non-decreasing");
    int64_t const nInc = d->GVGp[nV];
    // device kernels index vertices, incidence records and 3 or 4 entries per tet with
    // 32-bit integers
    if (nV >= (int64_t{1} << 31) || 4 * nT >= (int64_t{1} << 31) || nInc >= (int64_t{1} << 31))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "mesh too large for the device's 32-bit indices");
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
    h->acceleration = d->acceleration;
    h->rho = static_cast<float>(d->rho);
    h->window = d->acceleration == PBAT_VBD_ACCELERATION_ANDERSON ? d->window : 0;
    if (d->acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION)
        h->eta = d->eta, h->tau = d->tau, h->curved = d->curved != 0;
    h->kD = static_cast<float>(d->kD);
    h->detHZero = static_cast<float>(d->detHZero);

    // Mesh of each vertex (batches; one mesh otherwise)
    std::vector<int> meshOfOrig(static_cast<size_t>(nV), 0);
    if (vertexOffsets)
    {
        auto const& o = *vertexOffsets;
        if (o.size() < 2 || o.front() != 0 || o.back() != nV)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid mesh vertex offsets");
        for (size_t b = 0; b + 1 < o.size(); ++b)
            for (int64_t i = o[b]; i < o[b + 1]; ++i)
                meshOfOrig[i] = static_cast<int>(b);
        h->meshOffsets = o;
    }
    h->nMeshes = vertexOffsets ? static_cast<int>(vertexOffsets->size()) - 1 : 1;

    // Colour-major internal numbering; within a colour grouped by mesh, then
    // sorted by degree (stable).
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
            if (meshOfOrig[a] != meshOfOrig[b])
                return meshOfOrig[a] < meshOfOrig[b];
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
    {
        std::vector<int> meshOf(static_cast<size_t>(nV));
        for (int64_t v = 0; v < nV; ++v)
            meshOf[v] = meshOfOrig[h->perm[v]];
        h->meshOf.upload(meshOf, h->stream);
        PBAT_HIP(hipStreamSynchronize(h->stream));
    }

    BuildAggregates(*h, d, meshOfOrig);

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

    // Per-tet constants (3 float4) and vertex ids
    std::vector<float4> G4(static_cast<size_t>(3 * nT));
    std::vector<int4> TV(static_cast<size_t>(nT));
    for (int64_t e = 0; e < nT; ++e)
    {
        double const* G = d->GP + 12 * e; // G(r, c) = G[4c + r]
        auto gp = [&](int r, int c) { return static_cast<float>(G[4 * c + r]); };
        double const w = d->wg[e], mu = d->lame[2 * e], lam = d->lame[2 * e + 1];
        if (!(lam != 0.))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "lambda must be non-zero (element " + std::to_string(e) + ")");
        TV[e] = make_int4(h->iperm[d->E[4 * e]], h->iperm[d->E[4 * e + 1]], h->iperm[d->E[4 * e + 2]],
                          h->iperm[d->E[4 * e + 3]]);
        G4[3 * e] = make_float4(gp(1, 0), gp(1, 1), gp(1, 2), gp(2, 0));
        G4[3 * e + 1] = make_float4(gp(2, 1), gp(2, 2), gp(3, 0), gp(3, 1));
        G4[3 * e + 2] = make_float4(gp(3, 2), static_cast<float>(w * mu), static_cast<float>(w * lam),
                                    static_cast<float>(1. + mu / lam));
    }
    // Incidence records in internal numbering and sweep order: for each
    // incidence the tet's other three vertices (local order) with their GP
    // rows, and the tet's w*mu (in the id word) and w*lambda. Only updated
    // vertices are swept, so only their rows are stored.
    std::vector<int> P(static_cast<size_t>(nV) + 1, 0);
    for (int64_t v = 0; v < nV; ++v)
    {
        int64_t const i = h->perm[v];
        for (int64_t k = d->GVGp[i]; k < d->GVGp[i + 1]; ++k)
        {
            int64_t const e = d->GVGe[k];
            if (d->E[4 * e + d->GVGilocal[k]] != i)
                Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "GVGe/GVGilocal inconsistent with E at vertex " +
                                                          std::to_string(i));
        }
        P[v + 1] = P[v] + static_cast<int>(d->GVGp[i + 1] - d->GVGp[i]);
    }
    h->maxDegree = 0;
    for (int64_t v = 0; v < nUpd; ++v)
        h->maxDegree = std::max<int64_t>(h->maxDegree, P[v + 1] - P[v]);
    int64_t const nIncUpd = P[nUpd];
    std::vector<int4> slots(static_cast<size_t>(nT), make_int4(-1, -1, -1, -1));
    auto setSlot = [](int4& s, int il, int slot) {
        (il == 0 ? s.x : il == 1 ? s.y : il == 2 ? s.z : s.w) = slot;
    };
    auto fill = [&](HostIncidenceRecords& R, int64_t slot, int64_t k) {
        int64_t const e = d->GVGe[k];
        int const il = static_cast<int>(d->GVGilocal[k]);
        setSlot(slots[e], il, static_cast<int>(slot));
        double const* G = d->GP + 12 * e; // G(r, c) = G[4c + r]
        double const w = d->wg[e], mu = d->lame[2 * e], lam = d->lame[2 * e + 1];
        int ids[3];
        float g[3][3];
        double gd[3][3], own[3];
        int n = 0;
        for (int c = 0; c < 3; ++c)
            own[c] = G[4 * c + il];
        for (int j = 0; j < 4; ++j)
        {
            if (j == il)
                continue;
            ids[n] = h->iperm[d->E[4 * e + j]];
            for (int c = 0; c < 3; ++c)
            {
                gd[n][c] = G[4 * c + j];
                g[n][c] = static_cast<float>(gd[n][c]);
            }
            ++n;
        }
        float4 const m = G4[3 * e + 2]; // (., w*mu, w*lambda, alpha)
        int wmuBits;
        std::memcpy(&wmuBits, &m.y, sizeof(int));
        R.ids[slot] = make_int4(ids[0], ids[1], ids[2], wmuBits);
        R.g0[slot] = make_float4(g[0][0], g[0][1], g[0][2], g[1][0]);
        R.g1[slot] = make_float4(g[1][1], g[1][2], g[2][0], g[2][1]);
        R.g2[slot] = make_float2(g[2][2], m.z);
        // sweep record (FP64 from the FP64 GP and material): with gamma_k = g_k . g
        // (the other rows with the own row g) and kappa = det of the other three
        // rows, sids.w = bits of w lambda kappa^2 and sgeo = (w mu gamma_a,
        // w mu gamma_b, w mu gamma_c, kappa (w lambda + w mu)); shape keeps
        // (gamma_a, gamma_b, gamma_c, kappa) for later changes of the moduli
        auto dot = [](double const* a, double const* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        double const kappa = gd[0][0] * (gd[1][1] * gd[2][2] - gd[1][2] * gd[2][1]) -
                             gd[0][1] * (gd[1][0] * gd[2][2] - gd[1][2] * gd[2][0]) +
                             gd[0][2] * (gd[1][0] * gd[2][1] - gd[1][1] * gd[2][0]);
        double const ga = dot(gd[0], own), gb = dot(gd[1], own), gc = dot(gd[2], own);
        double const wmu = w * mu, wlam = w * lam;
        float const A = static_cast<float>(wlam * kappa * kappa);
        int Abits;
        std::memcpy(&Abits, &A, sizeof(int));
        R.sids[slot] = make_int4(ids[0], ids[1], ids[2], Abits);
        R.sgeo[slot] = make_float4(static_cast<float>(wmu * ga), static_cast<float>(wmu * gb),
                                   static_cast<float>(wmu * gc), static_cast<float>(kappa * (wlam + wmu)));
        R.shape[slot] = make_float4(static_cast<float>(ga), static_cast<float>(gb), static_cast<float>(gc),
                                    static_cast<float>(kappa));
    };
    {
        HostIncidenceRecords R;
        R.resize(static_cast<size_t>(nIncUpd));
        for (int64_t v = 0; v < nUpd; ++v)
        {
            int64_t const i = h->perm[v];
            for (int64_t k = d->GVGp[i]; k < d->GVGp[i + 1]; ++k)
                fill(R, P[v] + (k - d->GVGp[i]), k);
        }
        R.upload(h->rows, h->stream);
        h->slotsRows.upload(slots, h->stream);
        PBAT_HIP(hipStreamSynchronize(h->stream));
    }
    // Fixed-stride copy when padding every row to the maximum degree costs at
    // most 35% more records than the CSR rows.
    if (nUpd > 0 && h->maxDegree > 0 && h->maxDegree * nUpd * 100 <= 135 * nIncUpd &&
        h->maxDegree * nUpd < (int64_t{1} << 31))
    {
        int64_t const S = h->maxDegree;
        HostIncidenceRecords R;
        R.resize(static_cast<size_t>(nUpd * S));
        std::fill(slots.begin(), slots.end(), make_int4(-1, -1, -1, -1));
        for (int64_t v = 0; v < nUpd; ++v)
        {
            int64_t const i = h->perm[v];
            for (int64_t k = d->GVGp[i]; k < d->GVGp[i + 1]; ++k)
                fill(R, v * S + (k - d->GVGp[i]), k);
        }
        R.upload(h->ell, h->stream);
        h->slotsEll.upload(slots, h->stream);
        PBAT_HIP(hipStreamSynchronize(h->stream));
        h->ellStride = static_cast<int>(S);
    }

    h->wgHost.resize(static_cast<size_t>(nT));
    h->lameHost.resize(static_cast<size_t>(2 * nT));
    h->logY0.resize(static_cast<size_t>(nT));
    std::vector<float2> base(static_cast<size_t>(nT));
    for (int64_t e = 0; e < nT; ++e)
    {
        double const w = d->wg[e], mu = d->lame[2 * e], lam = d->lame[2 * e + 1];
        h->wgHost[e] = static_cast<float>(w);
        h->lameHost[2 * e] = static_cast<float>(mu);
        h->lameHost[2 * e + 1] = static_cast<float>(lam);
        // Young's modulus of the tet: Y = mu (3 lambda + 2 mu) / (lambda + mu)
        double const Y = mu * (3. * lam + 2. * mu) / (lam + mu);
        h->logY0[e] = static_cast<float>(std::log(Y));
        base[e] = make_float2(static_cast<float>(w * mu / Y), static_cast<float>(w * lam / Y));
    }
    h->logYCurrent = h->logY0;
    h->matBase.upload(base, h->stream);
    h->tgp.upload(G4, h->stream);
    h->tetv.upload(TV, h->stream);
    h->vtp.upload(P, h->stream);
    for (auto* b : {&h->xm, &h->xtil, &h->xt, &h->vel, &h->vprev, &h->aext})
        b->allocate(static_cast<size_t>(nV));
    if (h->acceleration == PBAT_VBD_ACCELERATION_CHEBYSHEV)
    {
        h->xkm1.allocate(static_cast<size_t>(nV));
        h->xkm2.allocate(static_cast<size_t>(nV));
    }
    if (h->acceleration == PBAT_VBD_ACCELERATION_ANDERSON)
    {
        size_t const n = static_cast<size_t>(nV), m = static_cast<size_t>(h->window);
        h->xkm1.allocate(n), h->gkm1.allocate(n), h->fkm1.allocate(n);
        h->dgk.allocate(m * n), h->dfk.allocate(m * n);
        h->andersonAcc.allocate(2 * static_cast<size_t>(device::AndersonAccDoubles()));
        h->andersonGram.allocate(2 * m * m);
    }
    if (LinearTrustRegion(*h))
    {
        h->xkm1.allocate(static_cast<size_t>(nV));
        h->trPartials.allocate(3 * static_cast<size_t>(kObjectiveBlocks) + 3);
    }
    h->energy.allocate(3);
    {
        void* p = nullptr;
        if (hipHostMalloc(&p, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess || !p)
            Fail(PBAT_VBD_ERROR_OUT_OF_MEMORY, "hipHostMalloc of the step status failed");
        h->statusHost = static_cast<int*>(p);
        std::memset(p, 0, 64);
        void* dp = nullptr;
        PBAT_HIP(hipHostGetDevicePointer(&dp, p, 0));
        h->statusDev = static_cast<int*>(dp);
        h->fp64Solves.allocate(1);
        unsigned int const zero = 0;
        PBAT_HIP(hipMemcpyAsync(h->fp64Solves.ptr, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
        PBAT_HIP(hipStreamSynchronize(h->stream));
    }
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
    SetUpContact(*h, d);
    PBAT_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
}

// MatrixMarket dense array file (column-major values), `digits` significant digits.
void WriteDense(std::string const& path, int64_t rows, int64_t cols, std::vector<double> const& colMajor,
                int digits = 9);

// Buffers of the FP64 objective (allocated on first use). The vertex -> (tet,
// local index) CSR over every vertex, Dirichlet ones included, comes from the
// device's element table, in ascending tet order per vertex.
void EnsureObjectiveBuffers(pbat_vbd_integrator& h)
{
    size_t const nV = static_cast<size_t>(h.nV);
    if (h.objVtp.n != nV + 1)
    {
        std::vector<int4> TV(static_cast<size_t>(h.nT));
        PBAT_HIP(hipMemcpy(TV.data(), h.tetv.ptr, TV.size() * sizeof(int4), hipMemcpyDeviceToHost));
        std::vector<int> ptr(nV + 1, 0), inc(4 * TV.size());
        for (int4 const& t : TV)
            for (int id : {t.x, t.y, t.z, t.w})
                ++ptr[id + 1];
        for (size_t v = 0; v < nV; ++v)
            ptr[v + 1] += ptr[v];
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (size_t e = 0; e < TV.size(); ++e)
        {
            int const ids[4] = {TV[e].x, TV[e].y, TV[e].z, TV[e].w};
            for (int l = 0; l < 4; ++l)
                inc[fill[ids[l]]++] = static_cast<int>(4 * e + l);
        }
        h.objVtp.upload(ptr, h.stream);
        h.objInc.upload(inc, h.stream);
        PBAT_HIP(hipStreamSynchronize(h.stream));
    }
    if (h.objX.n != nV)
    {
        h.objX.allocate(nV), h.objXtil.allocate(nV);
        h.objGrad.allocate(3 * nV);
    }
    size_t const nb = static_cast<size_t>(std::max<int64_t>((h.nV + 255) / 256, kObjectiveBlocks));
    if (h.objPart.n != 6 * nb + 6)
        h.objPart.allocate(6 * nb + 6);
}

// sim::vbd::Integrator::ObjectiveFunction and ObjectiveFunctionGradient
// (sim/vbd/Integrator.cpp:138-200) at device positions x and inertial target xtil
// (float4, internal order), in FP64: f = 1/2 sum m |x - x~|^2 + dt^2 sum_e w_e Psi_e,
// its gradient (3 x nV, original vertex order, nullable) and the squared gradient
// norm over the non-Dirichlet vertices (nullable).
void ObjectiveOnDevice(pbat_vbd_integrator& h, float4 const* x, float4 const* xtil, float dt, double* f,
                       double* gradOrig, double* r2)
{
    EnsureObjectiveBuffers(h);
    int const nV = static_cast<int>(h.nV), nT = static_cast<int>(h.nT);
    int const gridG = (nV + 255) / 256;
    size_t const nb = static_cast<size_t>(std::max(gridG, kObjectiveBlocks));
    double* partF = h.objPart.ptr;
    double* partG = h.objPart.ptr + 3 * nb;
    double* sums = h.objPart.ptr + 6 * nb; // (K, U, .) then (|g|^2, ., .)
    device::TrEpilogue none{};
    hipLaunchKernelGGL(device::ObjectiveTerms, dim3(kObjectiveBlocks), dim3(256), 0, h.stream, x, xtil,
                       static_cast<float4 const*>(nullptr), static_cast<int4 const*>(h.tetv.ptr),
                       static_cast<float4 const*>(h.tgp.ptr), nV, nT, partF, static_cast<double const*>(nullptr));
    hipLaunchKernelGGL(device::SumPartials3, dim3(1), dim3(256), 0, h.stream, static_cast<double const*>(partF),
                       kObjectiveBlocks, sums, static_cast<double const*>(nullptr), none);
    double const dt2 = double(dt) * double(dt);
    hipLaunchKernelGGL(device::ObjectiveGradient, dim3(std::max(gridG, 1)), dim3(256), 0, h.stream, x, xtil,
                       static_cast<int const*>(h.objVtp.ptr), static_cast<int const*>(h.objInc.ptr),
                       static_cast<int4 const*>(h.tetv.ptr), static_cast<float4 const*>(h.tgp.ptr), nV,
                       static_cast<int>(h.nUpdated), dt2, h.objGrad.ptr, partG);
    hipLaunchKernelGGL(device::SumPartials3, dim3(1), dim3(256), 0, h.stream, static_cast<double const*>(partG),
                       std::max(gridG, 1), sums + 3, static_cast<double const*>(nullptr), none);
    PBAT_HIP(hipGetLastError());
    double s[6];
    PBAT_HIP(hipMemcpyAsync(s, sums, sizeof(s), hipMemcpyDeviceToHost, h.stream));
    std::vector<double> g;
    if (gradOrig)
    {
        g.resize(3 * static_cast<size_t>(nV));
        PBAT_HIP(hipMemcpyAsync(g.data(), h.objGrad.ptr, g.size() * sizeof(double), hipMemcpyDeviceToHost,
                                h.stream));
    }
    PBAT_HIP(hipStreamSynchronize(h.stream));
    if (f)
        *f = 0.5 * s[0] + dt2 * s[1];
    if (r2)
        *r2 = s[3];
    if (gradOrig)
        for (int v = 0; v < nV; ++v)
            for (int d = 0; d < 3; ++d)
                gradOrig[3 * static_cast<size_t>(h.perm[v]) + d] = g[3 * static_cast<size_t>(v) + d];
}

void Objective(pbat_vbd_integrator& h, float const* x, float const* xtilde, float dt, double* f, double* grad,
               double* r2)
{
    if (!(dt > 0.f) || !std::isfinite(dt))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "dt must be positive and finite");
    PBAT_HIP(hipSetDevice(h.device));
    EnsureObjectiveBuffers(h);
    float4 const* xd = h.xm.ptr;
    float4 const* xl = h.xtil.ptr;
    if (x)
    {
        UploadVec(h, h.objX.ptr, x, h.massOrig.data());
        xd = h.objX.ptr;
    }
    if (xtilde)
    {
        UploadVec(h, h.objXtil.ptr, xtilde, nullptr);
        xl = h.objXtil.ptr;
    }
    ObjectiveOnDevice(h, xd, xl, dt, f, grad, r2);
}

// ---- objective without device allocation -----------------------------------
// Scratch of ObjectiveNoAlloc: the per-workgroup partials of ObjectiveTerms
// (3 doubles for each of its kObjectiveBlocks workgroups) and their 3 sums,
// 16 KiB. Each GPU holds its own copy until the process ends; every call on a
// GPU overwrites it, so calls that overlap on one GPU race.
constexpr int kNoAllocScratchDoubles = 2048;
static_assert(3 * kObjectiveBlocks + 3 <= kNoAllocScratchDoubles, "objective partials exceed the scratch");
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

// Device address of the scratch on each GPU, looked up once per device (with
// that device current); the host-side name is not a device address.
constexpr int kMaxScratchDevices = 64;
std::mutex gScratchMutex;
double* gScratchPtr[kMaxScratchDevices] = {};

double* ScratchOnDevice(int device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "device index beyond the no-allocation entry point's table");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[device])
    {
        void* ptr = nullptr;
        PBAT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr)
            Fail(PBAT_VBD_ERROR_HIP, "hipGetSymbolAddress returned no address for the objective scratch");
        gScratchPtr[device] = static_cast<double*>(ptr);
    }
    return gScratchPtr[device];
}

// The objective of ObjectiveOnDevice at the handle's current positions and
// inertial target (f only), with its partials in the static scratch instead of
// the handle's lazily allocated buffers: the same kernels and partial layout,
// so the same bits. Allocates no device memory.
void ObjectiveNoAlloc(pbat_vbd_integrator& h, float dt, double* f)
{
    if (!(dt > 0.f) || !std::isfinite(dt))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (!f)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null objective output");
    if (h.nV > std::numeric_limits<int>::max() || h.nT > std::numeric_limits<int>::max())
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "mesh too large for the objective reduction (2^31 vertices or tets)");
    PBAT_HIP(hipSetDevice(h.device));
    double* const part = ScratchOnDevice(h.device);
    double* const sums = part + 3 * kObjectiveBlocks;
    device::TrEpilogue none{};
    hipLaunchKernelGGL(device::ObjectiveTerms, dim3(kObjectiveBlocks), dim3(256), 0, h.stream,
                       static_cast<float4 const*>(h.xm.ptr), static_cast<float4 const*>(h.xtil.ptr),
                       static_cast<float4 const*>(nullptr), static_cast<int4 const*>(h.tetv.ptr),
                       static_cast<float4 const*>(h.tgp.ptr), static_cast<int>(h.nV), static_cast<int>(h.nT), part,
                       static_cast<double const*>(nullptr));
    hipLaunchKernelGGL(device::SumPartials3, dim3(1), dim3(256), 0, h.stream, static_cast<double const*>(part),
                       kObjectiveBlocks, sums, static_cast<double const*>(nullptr), none);
    PBAT_HIP(hipGetLastError());
    double s[2];
    PBAT_HIP(hipMemcpyAsync(s, sums, sizeof(s), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    *f = 0.5 * s[0] + double(dt) * double(dt) * s[1];
}

// sim::vbd::Integrator::TraceNextStep followed by Step (sim/vbd/Integrator.cpp:22-52,
// 73-87, 202-235): one step whose iterates are traced. Before every VBD iteration
// and after the last, the objective, its gradient and the positions at that iterate
// are recorded (FP64 reductions on the device); per substep s they go to
// {dir}/{t}.{s}.f.mtx ((iterations + 1) x 1), {t}.{s}.grad.mtx and {t}.{s}.x.mtx
// (3 nV x (iterations + 1), each column vertex-major as the reference's
// MatrixX::reshaped()). Accelerated integrators trace between their sweeps as the
// reference's do (RunVbdIteration records before each sweep); the linear
// trust-region path traces plain VBD iterations. Runs eagerly, no graph.
void TraceObjectiveStep(pbat_vbd_integrator& h, float dt, int iterations, int substeps, int t, char const* dir)
{
    if (!(dt > 0.f) || !std::isfinite(dt) || iterations < 0 || substeps < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid step arguments");
    PBAT_HIP(hipSetDevice(h.device));
    EnsureObjectiveBuffers(h);
    ClearStatus(h);
    float const sdt = dt / static_cast<float>(substeps);
    float const sdt2 = sdt * sdt;
    int const n = static_cast<int>(h.nV);
    int const eb = 256, eg = (n + eb - 1) / eb;
    size_t const nd = 3 * static_cast<size_t>(n);
    device::SweepArgs a = MakeSweepArgs(h, sdt, sdt2);
    if (h.contact)
        ContactBeginStep(h, dt);
    std::string const d = (dir && *dir) ? dir : ".";
    for (int s = 0; s < substeps; ++s)
    {
        hipLaunchKernelGGL(device::SubstepBegin, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xtil.ptr, h.xt.ptr,
                           h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, sdt, sdt2, h.strategy);
        if (h.contact && s % h.activeSetFrequency == 0)
            ContactNearestPass(h, 0);
        std::vector<double> F, G, X;
        std::vector<float> xbuf(nd);
        auto record = [&] {
            double fk = 0.;
            G.resize(G.size() + nd);
            ObjectiveOnDevice(h, h.xm.ptr, h.xtil.ptr, sdt, &fk, G.data() + G.size() - nd, nullptr);
            F.push_back(fk);
            DownloadVec(h, h.xm.ptr, xbuf.data());
            X.insert(X.end(), xbuf.begin(), xbuf.end());
        };
        float omega = 0.f;
        BeginAcceleration(h);
        for (int k = 0; k < iterations; ++k)
        {
            record();
            SweepAllColours(h, a);
            LaunchAcceleration(h, k, omega);
        }
        record();
        hipLaunchKernelGGL(device::SubstepEnd, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xt.ptr, h.vel.ptr,
                           h.vprev.ptr, n, sdt, h.statusDev);
        std::string const stem = d + "/" + std::to_string(t) + "." + std::to_string(s) + ".";
        int64_t const cols = static_cast<int64_t>(F.size());
        WriteDense(stem + "f.mtx", cols, 1, F, 17);
        WriteDense(stem + "grad.mtx", static_cast<int64_t>(nd), cols, G, 17);
        WriteDense(stem + "x.mtx", static_cast<int64_t>(nd), cols, X, 17);
    }
    if (h.contact)
        ContactNearestPass(h, 1);
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipStreamSynchronize(h.stream));
    CheckStatus(h);
}

void Energy(pbat_vbd_integrator& h, double out[3])
{
    PBAT_HIP(hipSetDevice(h.device));
    hipLaunchKernelGGL(device::ZeroDoubles, dim3(1), dim3(64), 0, h.stream, h.energy.ptr, 3);
    int const n = static_cast<int>(std::max(h.nV, h.nT));
    int const grid = std::min(4096, (n + 255) / 256);
    hipLaunchKernelGGL(device::EnergyTerms, dim3(grid), dim3(256), 0, h.stream, h.xm.ptr, h.vel.ptr, h.aext.ptr,
                       h.tetv.ptr, h.tgp.ptr, static_cast<int>(h.nV), static_cast<int>(h.nT), h.energy.ptr);
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipMemcpyAsync(out, h.energy.ptr, 3 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
}

// ---- MatrixMarket dense export (gpu/impl/vbd/Integrator.cu:105-148, 284-301) ----
void WriteDense(std::string const& path, int64_t rows, int64_t cols, std::vector<double> const& colMajor,
                int digits)
{
    std::ofstream f(path);
    if (!f)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "cannot open " + path + " for writing");
    f << "%%MatrixMarket matrix array real general\n" << rows << " " << cols << "\n";
    f << std::setprecision(digits);
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
    ClearStatus(h);
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
        // (T |#elements|x4, M |#vertices|x1, wg |#elements|x1, GP 4x3|#elements|, lame |#elements|x2)
        int64_t const nT = h.nT;
        std::vector<int4> TV(static_cast<size_t>(nT));
        std::vector<float4> G4(static_cast<size_t>(3 * nT));
        PBAT_HIP(hipMemcpy(TV.data(), h.tetv.ptr, TV.size() * sizeof(int4), hipMemcpyDeviceToHost));
        PBAT_HIP(hipMemcpy(G4.data(), h.tgp.ptr, G4.size() * sizeof(float4), hipMemcpyDeviceToHost));
        std::vector<double> Tm(static_cast<size_t>(4 * nT)), M(h.massOrig.begin(), h.massOrig.end());
        std::vector<double> W(static_cast<size_t>(nT)), GP(static_cast<size_t>(12 * nT)), La(2 * W.size());
        for (int64_t e = 0; e < nT; ++e)
        {
            int const ids[4] = {TV[e].x, TV[e].y, TV[e].z, TV[e].w};
            for (int k = 0; k < 4; ++k)
                Tm[k * nT + e] = h.perm[ids[k]];
            W[e] = h.wgHost[e];
            La[e] = h.lameHost[2 * e];
            La[nT + e] = h.lameHost[2 * e + 1];
            float4 const q0 = G4[3 * e], q1 = G4[3 * e + 1], q2 = G4[3 * e + 2];
            float const rows[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
            for (int c = 0; c < 3; ++c)
            {
                // column 3e + c of the 4 x 3nT matrix
                double* col = GP.data() + 4 * (3 * e + c);
                col[0] = -(rows[0][c] + rows[1][c] + rows[2][c]);
                for (int r = 0; r < 3; ++r)
                    col[r + 1] = rows[r][c];
            }
        }
        WriteDense(Join(dir, "T.mtx"), nT, 4, Tm);
        WriteDense(Join(dir, "M.mtx"), h.nV, 1, M);
        WriteDense(Join(dir, "wg.mtx"), nT, 1, W);
        WriteDense(Join(dir, "GP.mtx"), 4, 3 * nT, GP);
        WriteDense(Join(dir, "lame.mtx"), nT, 2, La);
    }
    float const sdt = dt / static_cast<float>(substeps);
    float const sdt2 = sdt * sdt;
    int const n = static_cast<int>(h.nV);
    int const eb = 256, eg = (n + eb - 1) / eb;
    device::SweepArgs a = MakeSweepArgs(h, sdt, sdt2);
    if (h.contact)
        ContactBeginStep(h, dt);
    for (int s = 0; s < substeps; ++s)
    {
        hipLaunchKernelGGL(device::SubstepBegin, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xtil.ptr, h.xt.ptr,
                           h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, sdt, sdt2, h.strategy);
        WriteVertexMatrix(h, h.xtil.ptr, Join(dir, name("xtilde", t, s, -1)));
        if (h.contact && s % h.activeSetFrequency == 0)
            ContactNearestPass(h, 0);
        // TracedSolve (gpu/impl/vbd/Integrator.cu:284-301) is not overridden by the
        // accelerated integrators: traced iterations are plain VBD iterations.
        for (int k = 0; k < iterations; ++k)
        {
            WriteVertexMatrix(h, h.xm.ptr, Join(dir, name("x", t, s, k)));
            SweepAllColours(h, a);
        }
        WriteVertexMatrix(h, h.xm.ptr, Join(dir, name("x", t, s, iterations)));
        hipLaunchKernelGGL(device::SubstepEnd, dim3(eg), dim3(eb), 0, h.stream, h.xm.ptr, h.xt.ptr, h.vel.ptr,
                           h.vprev.ptr, n, sdt, h.statusDev);
    }
    if (h.contact)
        ContactNearestPass(h, 1);
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipStreamSynchronize(h.stream));
    CheckStatus(h);
}

// Time full sweeps (all colours) for each candidate sweep shape and keep the
// fastest. Device state is saved before and restored after, so the simulation
// is unchanged. Candidates: lanes L in {2..64} with the natural prefetch depth
// K = ceil(maxDegree / L) (when K <= 4), each with and without the XCD order
// and with and without non-temporal loads of the incidence stream.
constexpr int kMaxTuneRows = 32;

void Autotune(pbat_vbd_integrator& h, float dt, int sweeps, double* timingsUs, int32_t* nCandidates)
{
    if (!(dt > 0.f) || sweeps < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "autotune needs dt > 0 and sweeps >= 1");
    PBAT_HIP(hipSetDevice(h.device));
    size_t const bytes = sizeof(float4) * static_cast<size_t>(h.nV);
    DeviceArray<float4> saved;
    saved.allocate(5 * static_cast<size_t>(h.nV));
    float4* bufs[5] = {h.xm.ptr, h.xtil.ptr, h.xt.ptr, h.vel.ptr, h.vprev.ptr};
    for (int b = 0; b < 5; ++b)
        PBAT_HIP(hipMemcpyAsync(saved.ptr + b * h.nV, bufs[b], bytes, hipMemcpyDeviceToDevice, h.stream));
    int const n = static_cast<int>(h.nV);
    hipLaunchKernelGGL(device::SubstepBegin, dim3((n + 255) / 256), dim3(256), 0, h.stream, h.xm.ptr, h.xtil.ptr,
                       h.xt.ptr, h.vel.ptr, h.vprev.ptr, h.aext.ptr, n, dt, dt * dt, h.strategy);
    int const oldL = h.lanes, oldK = h.perLane;
    bool const oldX = h.xcdRemap, oldN = h.ntStream, oldTiming = h.timing;
    h.timing = false;
    hipEvent_t e0, e1;
    PBAT_HIP(hipEventCreate(&e0));
    PBAT_HIP(hipEventCreate(&e1));
    device::SweepArgs a = MakeSweepArgs(h, dt, dt * dt);
    int const nColors = static_cast<int>(h.colorBegin.size()) - 1;
    auto sweep = [&] {
        for (int c = 0; c < nColors; ++c)
        {
            a.vbeg = h.colorBegin[c];
            a.vend = h.colorBegin[c + 1];
            if (a.vend > a.vbeg)
                LaunchSweep(h, a, a.vend - a.vbeg);
        }
        FlushContactSweep(h);
    };
    struct Candidate
    {
        int L, K;
        bool xcd, nt, ell;
    };
    std::vector<Candidate> cands;
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    for (int L = 2; L <= 64; L <<= 1)
    {
        int const K = (d + L - 1) / L;
        if (K > 4)
            continue;
        for (int ell = (h.ellStride > 0 ? 1 : 0); ell >= 0; --ell)
            for (int x = 1; x >= 0; --x)
                for (int nt = 0; nt <= (h.tuneNontemporal ? 1 : 0); ++nt)
                    cands.push_back({L, K, x != 0, nt != 0, ell != 0});
    }
    bool const oldE = h.useEll;
    double best = 1e300;
    Candidate bestC{EffectiveLanes(h), EffectivePerLane(h), h.xcdRemap, h.ntStream, h.useEll};
    int count = 0;
    try
    {
        for (Candidate const& c : cands)
        {
            h.lanes = c.L, h.perLane = c.K, h.xcdRemap = c.xcd, h.ntStream = c.nt, h.useEll = c.ell;
            a = MakeSweepArgs(h, dt, dt * dt);
            sweep(); // warm-up
            PBAT_HIP(hipEventRecord(e0, h.stream));
            for (int s = 0; s < sweeps; ++s)
                sweep();
            PBAT_HIP(hipEventRecord(e1, h.stream));
            PBAT_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PBAT_HIP(hipEventElapsedTime(&ms, e0, e1));
            double const us = 1e3 * ms / sweeps;
            if (timingsUs && count < kMaxTuneRows)
            {
                double* row = timingsUs + 6 * count;
                row[0] = c.L, row[1] = c.K, row[2] = c.xcd, row[3] = c.nt, row[4] = c.ell, row[5] = us;
            }
            ++count;
            if (us < best)
                best = us, bestC = c;
        }
    }
    catch (...)
    {
        h.lanes = oldL, h.perLane = oldK, h.xcdRemap = oldX, h.ntStream = oldN, h.useEll = oldE;
        h.timing = oldTiming;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h.lanes = bestC.L, h.perLane = bestC.K, h.xcdRemap = bestC.xcd, h.ntStream = bestC.nt, h.useEll = bestC.ell;
    h.timing = oldTiming;
    for (int b = 0; b < 5; ++b)
        PBAT_HIP(hipMemcpyAsync(bufs[b], saved.ptr + b * h.nV, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    InvalidateGraph(h);
    if (nCandidates)
        *nCandidates = count;
}

// A batch of independent meshes is one mesh with disconnected components:
// vertex, element and incidence indices are offset per mesh, and global
// colour c is the union of every mesh's colour c. Each mesh is therefore
// swept in exactly the order its own integrator would use, and one launch
// per colour serves the whole batch.
struct UnionData
{
    std::vector<double> x, v, aext, m, GP, wg, lame;
    std::vector<int64_t> E, GVGp, GVGe, GVGilocal, Pptr, Padj;
    std::vector<int64_t> vertexOffsets;
    pbat_vbd_data d{};
};

void BuildUnion(pbat_vbd_data const* meshes, int count, UnionData& U)
{
    if (!meshes || count < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "a batch needs at least one mesh");
    int64_t nV = 0, nT = 0, nInc = 0, nColors = 0;
    for (int b = 0; b < count; ++b)
    {
        pbat_vbd_data const& m = meshes[b];
        if (m.nV <= 0 || m.nT <= 0 || !m.GVGp || !m.Pptr || m.nPartitions < 0 ||
            m.nPartitions > PBAT_VBD_MAX_COLORS)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid mesh " + std::to_string(b) + " in batch");
        if (m.strategy != meshes[0].strategy || m.kD != meshes[0].kD || m.detHZero != meshes[0].detHZero ||
            m.acceleration != meshes[0].acceleration ||
            (m.acceleration == PBAT_VBD_ACCELERATION_CHEBYSHEV && m.rho != meshes[0].rho))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                 "meshes of a batch must share the initialization strategy, kD, detHZero and acceleration");
        if (m.acceleration == PBAT_VBD_ACCELERATION_ANDERSON)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                 "Anderson acceleration couples every vertex of a handle; batches do not support it");
        if (m.acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                 "trust-region acceleration couples every vertex of a handle; batches do not support it");
        if (CollisionMeshSpansBodies(m))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                 "contact couples bodies within a mesh's data; batches of independent meshes do not run it");
        nV += m.nV;
        nT += m.nT;
        nInc += m.GVGp[m.nV];
        nColors = std::max(nColors, m.nPartitions);
    }
    U.vertexOffsets.assign(static_cast<size_t>(count) + 1, 0);
    U.GVGp.assign(1, 0);
    U.x.reserve(3 * nV), U.v.reserve(3 * nV), U.aext.reserve(3 * nV), U.m.reserve(nV);
    U.E.reserve(4 * nT), U.GP.reserve(12 * nT), U.wg.reserve(nT), U.lame.reserve(2 * nT);
    U.GVGp.reserve(nV + 1), U.GVGe.reserve(nInc), U.GVGilocal.reserve(nInc);
    int64_t vOff = 0, tOff = 0;
    for (int b = 0; b < count; ++b)
    {
        pbat_vbd_data const& m = meshes[b];
        if (!m.x || !m.v || !m.aext || !m.m || !m.E || !m.GP || !m.wg || !m.lame || !m.GVGe || !m.GVGilocal ||
            (m.nPartitions > 0 && !m.Padj))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null array in mesh " + std::to_string(b) + " of batch");
        U.x.insert(U.x.end(), m.x, m.x + 3 * m.nV);
        U.v.insert(U.v.end(), m.v, m.v + 3 * m.nV);
        U.aext.insert(U.aext.end(), m.aext, m.aext + 3 * m.nV);
        U.m.insert(U.m.end(), m.m, m.m + m.nV);
        for (int64_t k = 0; k < 4 * m.nT; ++k)
            U.E.push_back(m.E[k] + vOff);
        U.GP.insert(U.GP.end(), m.GP, m.GP + 12 * m.nT);
        U.wg.insert(U.wg.end(), m.wg, m.wg + m.nT);
        U.lame.insert(U.lame.end(), m.lame, m.lame + 2 * m.nT);
        int64_t const incBase = U.GVGp.back();
        for (int64_t i = 0; i < m.nV; ++i)
            U.GVGp.push_back(incBase + m.GVGp[i + 1]);
        for (int64_t k = 0; k < m.GVGp[m.nV]; ++k)
        {
            U.GVGe.push_back(m.GVGe[k] + tOff);
            U.GVGilocal.push_back(m.GVGilocal[k]);
        }
        vOff += m.nV;
        tOff += m.nT;
        U.vertexOffsets[b + 1] = vOff;
    }
    U.Pptr.assign(1, 0);
    for (int64_t c = 0; c < nColors; ++c)
    {
        for (int b = 0; b < count; ++b)
        {
            pbat_vbd_data const& m = meshes[b];
            if (c >= m.nPartitions)
                continue;
            for (int64_t k = m.Pptr[c]; k < m.Pptr[c + 1]; ++k)
            {
                if (m.Padj[k] < 0 || m.Padj[k] >= m.nV)
                    Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "Padj entry out of range in mesh " + std::to_string(b));
                U.Padj.push_back(m.Padj[k] + U.vertexOffsets[b]);
            }
        }
        U.Pptr.push_back(static_cast<int64_t>(U.Padj.size()));
    }
    pbat_vbd_data& d = U.d;
    d.nV = nV, d.nT = nT;
    d.x = U.x.data(), d.v = U.v.data(), d.aext = U.aext.data(), d.m = U.m.data();
    d.E = U.E.data(), d.GP = U.GP.data(), d.wg = U.wg.data(), d.lame = U.lame.data();
    d.GVGp = U.GVGp.data(), d.GVGe = U.GVGe.data(), d.GVGilocal = U.GVGilocal.data();
    // Padj may be empty (every vertex fixed); any non-null pointer serves then
    d.nPartitions = nColors, d.Pptr = U.Pptr.data(), d.Padj = U.Padj.empty() ? U.Pptr.data() : U.Padj.data();
    d.strategy = meshes[0].strategy, d.kD = meshes[0].kD, d.detHZero = meshes[0].detHZero;
    d.acceleration = meshes[0].acceleration, d.rho = meshes[0].rho, d.window = meshes[0].window;
}

pbat_vbd_integrator& Handle(pbat_vbd_integrator* h)
{
    if (!h)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null integrator handle");
    return *h;
}

} // namespace pbat_vbd

using namespace pbat_vbd;

extern "C" int pbat_vbd_integrator_create(const pbat_vbd_data* data, pbat_vbd_integrator** out)
{
    return Guard([&] { Create(data, out); });
}

extern "C" int pbat_vbd_batch_create(const pbat_vbd_data* meshes, int32_t count, pbat_vbd_integrator** out)
{
    return Guard([&] {
        if (!out)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null output handle");
        UnionData U;
        BuildUnion(meshes, count, U);
        Create(&U.d, out, &U.vertexOffsets);
    });
}

extern "C" int pbat_vbd_batch_vertex_offsets(pbat_vbd_integrator* h, int64_t* offsets, int32_t* count)
{
    return Guard([&] {
        auto& H = Handle(h);
        int const n = H.meshOffsets.empty() ? 1 : static_cast<int>(H.meshOffsets.size()) - 1;
        if (count)
            *count = n;
        if (offsets)
        {
            if (H.meshOffsets.empty())
                offsets[0] = 0, offsets[1] = H.nV;
            else
                std::copy(H.meshOffsets.begin(), H.meshOffsets.end(), offsets);
        }
    });
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
        H.rollout.Discard();
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
        H.rollout.Discard();
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
        H.rollout.Discard();
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
    return Guard([&] {
        Handle(h).detHZero = zero;
        Handle(h).rollout.Discard();
    });
}

extern "C" int pbat_vbd_integrator_set_rayleigh_damping(pbat_vbd_integrator* h, float kD)
{
    return Guard([&] {
        Handle(h).kD = kD;
        Handle(h).rollout.Discard();
    });
}

extern "C" int pbat_vbd_integrator_set_initialization_strategy(pbat_vbd_integrator* h, int32_t strategy)
{
    return Guard([&] {
        if (strategy < 0 || strategy > 4)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid initialization strategy");
        Handle(h).strategy = strategy;
        Handle(h).rollout.Discard();
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
        Handle(h).rollout.Discard();
    });
}

extern "C" int pbat_vbd_integrator_set_sweep_config(pbat_vbd_integrator* h, const pbat_vbd_sweep_config* config)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (!config)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null sweep config");
        int const lanes = config->lanes_per_vertex, perLane = config->incidences_per_lane;
        if (lanes != 0 && (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "lanes per vertex must be 0 or a power of two in [1, 64]");
        if (perLane < 0 || perLane > 4)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "incidences per lane must be in [0, 4]");
        if (config->fixed_stride && H.ellStride == 0)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
                 "fixed-stride incidence rows are not available for this mesh (degree spread too large)");
        H.lanes = lanes;
        H.perLane = perLane;
        H.xcdRemap = config->xcd_remap != 0;
        H.ntStream = config->nontemporal != 0;
        H.useEll = config->fixed_stride != 0;
        H.rollout.Discard(); // the recomputed steps must repeat the forward bit for bit
    });
}

extern "C" int pbat_vbd_integrator_get_sweep_config(pbat_vbd_integrator* h, pbat_vbd_sweep_config* config)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (!config)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null sweep config");
        config->lanes_per_vertex = EffectiveLanes(H);
        config->incidences_per_lane = EffectivePerLane(H);
        config->xcd_remap = H.xcdRemap ? 1 : 0;
        config->nontemporal = H.ntStream ? 1 : 0;
        config->fixed_stride = H.useEll ? 1 : 0;
    });
}

extern "C" int pbat_vbd_integrator_autotune(pbat_vbd_integrator* h, float dt, int32_t sweeps, double* timings,
                                            int32_t* n_candidates)
{
    return Guard([&] {
        Autotune(Handle(h), dt, sweeps, timings, n_candidates);
        Handle(h).rollout.Discard();
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

extern "C" int pbat_vbd_integrator_fp64_solves(pbat_vbd_integrator* h, int64_t* count, int32_t reset)
{
    return Guard([&] {
        auto& H = Handle(h);
        PBAT_HIP(hipSetDevice(H.device));
        unsigned int c = 0;
        PBAT_HIP(hipMemcpyAsync(&c, H.fp64Solves.ptr, sizeof(c), hipMemcpyDeviceToHost, H.stream));
        PBAT_HIP(hipStreamSynchronize(H.stream));
        if (count)
            *count = static_cast<int64_t>(c);
        if (reset)
        {
            unsigned int const zero = 0;
            PBAT_HIP(hipMemcpyAsync(H.fp64Solves.ptr, &zero, sizeof(zero), hipMemcpyHostToDevice, H.stream));
            PBAT_HIP(hipStreamSynchronize(H.stream));
        }
    });
}

extern "C" int pbat_vbd_integrator_objective(pbat_vbd_integrator* h, const float* x, const float* xtilde, float dt,
                                             double* f, double* grad, double* residual2)
{
    return Guard([&] { Objective(Handle(h), x, xtilde, dt, f, grad, residual2); });
}

extern "C" int pbat_vbd_integrator_objective_noalloc(pbat_vbd_integrator* h, float dt, double* f)
{
    return Guard([&] { ObjectiveNoAlloc(Handle(h), dt, f); });
}

extern "C" int pbat_vbd_integrator_trace_objective_step(pbat_vbd_integrator* h, float dt, int32_t iterations,
                                                        int32_t substeps, int32_t t, const char* dir)
{
    return Guard([&] { TraceObjectiveStep(Handle(h), dt, iterations, substeps, t, dir); });
}

extern "C" int pbat_vbd_integrator_synchronize(pbat_vbd_integrator* h)
{
    return Guard([&] { PBAT_HIP(hipStreamSynchronize(Handle(h).stream)); });
}

extern "C" int pbat_vbd_device_memory(int64_t* current, int64_t* peak, int32_t reset_peak)
{
    return Guard([&] {
        if (current)
            *current = gDeviceBytes.load();
        if (peak)
            *peak = gDevicePeak.load();
        if (reset_peak)
            gDevicePeak.store(gDeviceBytes.load());
    });
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

extern "C" int pbat_vbd_integrator_trust_region_log(pbat_vbd_integrator* h, double* t, int32_t* accepted,
                                                    int64_t capacity, int64_t* count)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (capacity < 0 || (capacity > 0 && (!t || !accepted)) || !count)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid trust-region log buffers");
        int64_t const n = H.trLogCount, m = std::min(n, capacity);
        if (m > 0)
        {
            PBAT_HIP(hipSetDevice(H.device));
            PBAT_HIP(hipStreamSynchronize(H.stream));
            std::vector<int> acc(static_cast<size_t>(m));
            PBAT_HIP(hipMemcpy(t, H.trLogT.ptr, sizeof(double) * m, hipMemcpyDeviceToHost));
            PBAT_HIP(hipMemcpy(acc.data(), H.trLogAccepted.ptr, sizeof(int) * m, hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < m; ++k)
                accepted[k] = acc[k];
        }
        *count = n;
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
