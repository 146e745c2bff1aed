// Differentiable VBD: gradients of a rollout's loss with respect to per-tet
// log Young's modulus, initial velocities and initial positions, by the
// implicit-function theorem applied to each step's incremental potential.
//
// Forward (rollout_forward): steps the integrator, keeping (x, v, v_prev)
// every k steps. Backward (rollout_backward): for each segment, from the last,
// restore its checkpoint and recompute its steps (bitwise identical to the
// forward), then walk the segment backwards. Per step, with H the Hessian of
// the step's incremental potential at its minimiser x* in the forward's
// scaling (M / h^2 + sum_e w_e d2Psi_e/dx2, updated vertices only):
//   solve  H y = xbar + vbar / h
//   d L / d log Y_e -= (w_e P_e(x*)) : dF_e(y)      (mu, lambda proportional to Y)
//   xt~bar = M y / h^2,  xbar <- xt~bar - vbar / h,  vbar <- h xt~bar
// The solve is coloured block Gauss-Seidel (the forward's colouring and 3x3
// diagonal blocks) either on its own or as the symmetric preconditioner of a
// conjugate-gradient iteration driven by the Hessian-vector kernel.
#include "integrator.hpp"
#include "vbd_adjoint_kernels.hpp"
#include "vbd_hessian_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace pbat_vbd {
namespace {

int Blocks(int64_t n, int block = 256) { return static_cast<int>((std::max<int64_t>(n, 1) + block - 1) / block); }

using device::AdjMode;

device::AdjointArgs MakeAdjointArgs(pbat_vbd_integrator& h, float4 const* xs, float dt)
{
    device::AdjointArgs a{};
    IncidenceRecords const& r = (h.useEll && h.ellStride > 0) ? h.ell : h.rows;
    a.vtp = h.vtp.ptr;
    a.ids = r.ids.ptr;
    a.g0 = r.g0.ptr;
    a.g1 = r.g1.ptr;
    a.g2 = r.g2.ptr;
    a.xs = xs;
    a.meshOf = h.meshOf.ptr;
    a.ellStride = (h.useEll && h.ellStride > 0) ? h.ellStride : 0;
    a.invDt2 = 1.f / (dt * dt);
    return a;
}

template <int L, int K, AdjMode M>
void LaunchAdjointLK(pbat_vbd_integrator& h, device::AdjointArgs a)
{
    if (a.vend <= a.vbeg)
        return;  // an empty colour (or no updated vertex at all)
    int const block = h.blockSize;
    int64_t const threads = static_cast<int64_t>(a.vend - a.vbeg) * L;
    int64_t grid64 = (threads + block - 1) / block;
    a.xcdChunk = 0;
    if (h.xcdRemap && grid64 >= 16)
    {
        grid64 = (grid64 + 7) / 8 * 8;
        a.xcdChunk = static_cast<int>(grid64 / 8);
    }
    if (grid64 * block >= (int64_t{1} << 31))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "vertex range too large for the adjoint launch");
    hipLaunchKernelGGL((device::AdjointVertexKernel<L, K, M>), dim3(static_cast<int>(grid64)), dim3(block), 0,
                       h.stream, a);
}

// Shape of the adjoint's vertex kernel: lanes per vertex = the busiest updated
// vertex's degree rounded up to a power of two (at most 64), so each lane takes
// one incidence; the remainder, if any, goes to the prefetch depth. Measured on
// the 32^3 workload (degree 24): 32 lanes beat 16 x 2, 8 x 3 and 64 x 1 by 4-15%.
// Unlike the forward sweep's shape it is not autotuned, so every handle of a
// given mesh runs the same summation order.
int AdjointLanes(pbat_vbd_integrator const& h)
{
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    int L = 1;
    while (L < 64 && L < d)
        L <<= 1;
    return L;
}

int AdjointPerLane(pbat_vbd_integrator const& h)
{
    int const d = static_cast<int>(std::max<int64_t>(h.maxDegree, 1));
    int const L = AdjointLanes(h);
    return std::clamp((d + L - 1) / L, 1, 4);
}

template <int L, AdjMode M>
void LaunchAdjointL(pbat_vbd_integrator& h, device::AdjointArgs const& a)
{
    switch (AdjointPerLane(h))
    {
        case 1: LaunchAdjointLK<L, 1, M>(h, a); break;
        case 2: LaunchAdjointLK<L, 2, M>(h, a); break;
        case 3: LaunchAdjointLK<L, 3, M>(h, a); break;
        default: LaunchAdjointLK<L, 4, M>(h, a); break;
    }
}

template <AdjMode M>
void LaunchAdjoint(pbat_vbd_integrator& h, device::AdjointArgs const& a)
{
    if (a.vend <= a.vbeg)
        return;
    switch (AdjointLanes(h))
    {
        case 1: LaunchAdjointL<1, M>(h, a); break;
        case 2: LaunchAdjointL<2, M>(h, a); break;
        case 4: LaunchAdjointL<4, M>(h, a); break;
        case 8: LaunchAdjointL<8, M>(h, a); break;
        case 16: LaunchAdjointL<16, M>(h, a); break;
        case 32: LaunchAdjointL<32, M>(h, a); break;
        default: LaunchAdjointL<64, M>(h, a); break;
    }
}

// Solver scalars: plain per-mesh values (padded to 128 B), then the accumulators.
size_t PlainDoubles(pbat_vbd_integrator const& h)
{
    return (static_cast<size_t>(device::kPlainCount) * h.nMeshes + 15) / 16 * 16;
}
size_t ScalarDoubles(pbat_vbd_integrator const& h) { return PlainDoubles(h) + device::AccDoubles(h.nMeshes); }

void EnsureBuffers(pbat_vbd_integrator& h)
{
    RolloutState& R = h.rollout;
    size_t const n = static_cast<size_t>(h.nV);
    for (auto* b : {&R.y, &R.yS, &R.bS, &R.r, &R.z, &R.p, &R.p2, &R.q, &R.rhs, &R.xbar, &R.vbar, &R.hold})
        if (b->n != n)
            b->allocate(n);
    static_assert(device::kProjMax == sizeof(R.hist) / sizeof(R.hist[0]), "history ring size");
    for (int j = 0; j < device::kProjMax; ++j)
        for (auto* b : {&R.hist[j], &R.projBasis[j], &R.projHb[j]})
            if (b->n != n)
                b->allocate(n);
    size_t const nProjPart = static_cast<size_t>(device::kProjTerms) * std::max(h.nAggregates, 1);
    if (R.projPart.n != nProjPart)
        R.projPart.allocate(nProjPart);
    if (R.projCoef.n != static_cast<size_t>(device::kProjMax) * h.nMeshes)
        R.projCoef.allocate(static_cast<size_t>(device::kProjMax) * h.nMeshes);
    if (R.dinv.n != 2 * n)
        R.dinv.allocate(2 * n);
    size_t const nM = static_cast<size_t>(h.nMeshes);
    if (R.scal.n != ScalarDoubles(h))
        R.scal.allocate(ScalarDoubles(h));
    if (R.flags.n != 2 * nM + 1)
        R.flags.allocate(2 * nM + 1);
    if (R.grad.n != static_cast<size_t>(h.nT))
        R.grad.allocate(static_cast<size_t>(h.nT));
}

device::CoarseArgs MakeCoarseArgs(pbat_vbd_integrator& h)
{
    device::CoarseArgs c{};
    c.aggOf = h.aggOf.ptr;
    c.aggMesh = h.aggMesh.ptr;
    c.meshAggBegin = h.meshAggBegin.ptr;
    c.meshOffset = h.meshCoarseOffset.ptr;
    c.nAgg = h.nAggregates;
    return c;
}

void AssembleBsr(pbat_vbd_integrator& h, float4 const* xs, float dt, bool aggregateOrder);

// The coarse operator Z^T H Z at xs (positions with masses): the step Hessian
// assembled there, its blocks summed per pair of aggregates in a fixed order,
// then each mesh's dense block inverted in place.
void BuildCoarse(pbat_vbd_integrator& h, float4 const* xs, float dt)
{
    RolloutState& R = h.rollout;
    size_t const total = static_cast<size_t>(h.coarseOffset.back());
    int const nM = h.nMeshes;
    if (R.coarse[0].n != total)
    {
        R.coarse[0].allocate(total);
        R.coarse[1].allocate(total);
    }
    size_t const nCoarse = static_cast<size_t>(device::kCoarseDofs) * h.nAggregates;
    if (R.rc.n != nCoarse)
    {
        R.rc.allocate(nCoarse);
        R.zc.allocate(nCoarse);
    }
    if (total == 0 || h.nAggregates == 0)
        return;
    device::CoarseArgs const c = MakeCoarseArgs(h);
    hipLaunchKernelGGL(device::ZeroDoublesLong, dim3(1024), dim3(256), 0, h.stream, R.coarse[0].ptr,
                       static_cast<long long>(total));
    hipLaunchKernelGGL(device::ZeroDoublesLong, dim3(64), dim3(256), 0, h.stream, R.rc.ptr,
                       static_cast<long long>(R.rc.n));
    AssembleBsr(h, xs, dt, true); // the two-level solver is a fused one: aggregate-major rows
    // the rotation offsets d of the coarse space at xs, per solver row
    if (R.coarseD.n != static_cast<size_t>(h.nV))
        R.coarseD.allocate(static_cast<size_t>(h.nV));
    hipLaunchKernelGGL(device::CoarseBasis, dim3(h.nAggregates), dim3(128), 0, h.stream,
                       static_cast<int const*>(h.aggVertPtr.ptr), static_cast<int const*>(R.aggRows.ptr),
                       static_cast<int const*>(R.rowVertex.ptr), xs, R.coarseD.ptr);
    int const nCB = static_cast<int>(R.cbA.n);
    if (nCB > 0)
    {
        device::CoarseFromBsrArgs g{};
        g.cbA = R.cbA.ptr, g.cbB = R.cbB.ptr, g.cbSrcPtr = R.cbSrcPtr.ptr, g.cbSrc = R.cbSrc.ptr;
        g.nCoarseBlocks = nCB;
        g.va = R.bsrA.ptr, g.vb = R.bsrB.ptr, g.vc = R.bsrC.ptr;
        g.blockRow = R.bsrBlockRow.ptr, g.col = R.bsrCol.ptr, g.coarseD = R.coarseD.ptr;
        g.c = c, g.Ac = R.coarse[0].ptr;
        hipLaunchKernelGGL(device::CoarseFromBsr, dim3((nCB + 3) / 4), dim3(256), 0, h.stream, g);
    }
    int maxN = 0;
    std::vector<int> finalBuf(static_cast<size_t>(nM));
    for (int m = 0; m < nM; ++m)
    {
        int const n = device::kCoarseDofs * (h.aggBegin[m + 1] - h.aggBegin[m]);
        maxN = std::max(maxN, n);
        finalBuf[m] = ((n + device::kPivot - 1) / device::kPivot) % 2; // parity of its step count
    }
    R.coarseMaxN = maxN;
    int64_t const maxEntries = int64_t(maxN) * maxN;
    dim3 const grid(static_cast<unsigned>((maxEntries + 255) / 256), static_cast<unsigned>(nM));
    // block Gauss-Jordan: per step the pivot rows, then the rest in 64 x 64 tiles
    int const tilesPerRow = (maxN + 63) / 64;
    if (R.pivotRows.n != static_cast<size_t>(nM) * device::kPivot * maxN)
        R.pivotRows.allocate(static_cast<size_t>(nM) * device::kPivot * maxN);
    for (int s = 0, k = 0; k < maxN; ++s, k += device::kPivot)
    {
        double const* Ak = R.coarse[s & 1].ptr;
        hipLaunchKernelGGL(device::CoarsePivotRows, dim3(tilesPerRow, nM), dim3(256), 0, h.stream, Ak,
                           R.pivotRows.ptr, k, maxN, c);
        hipLaunchKernelGGL(device::CoarsePivotUpdate, dim3(tilesPerRow * tilesPerRow, nM), dim3(256), 0, h.stream,
                           Ak, R.coarse[(s + 1) & 1].ptr, static_cast<double const*>(R.pivotRows.ptr), k, maxN,
                           tilesPerRow, c);
    }
    // allocated once per size: the captured CG chunk graph holds this pointer
    if (R.coarseFinal.n != finalBuf.size())
        R.coarseFinal.allocate(finalBuf.size());
    PBAT_HIP(hipMemcpyAsync(R.coarseFinal.ptr, finalBuf.data(), finalBuf.size() * sizeof(int),
                            hipMemcpyHostToDevice, h.stream));
    // per aggregate: its mesh's coarse rows (b0, n), its fine rows' range and
    // the offset of its first row of the inverse (allocated once per size: the
    // captured CG chunk graph holds these pointers)
    {
        std::vector<int4> info(static_cast<size_t>(h.nAggregates));
        std::vector<long long> rowOff(static_cast<size_t>(h.nAggregates));
        std::vector<int> aggPtr(static_cast<size_t>(h.nAggregates) + 1);
        PBAT_HIP(hipMemcpy(aggPtr.data(), h.aggVertPtr.ptr, aggPtr.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int m = 0; m < nM; ++m)
        {
            int const b0 = device::kCoarseDofs * h.aggBegin[m];
            int const n = device::kCoarseDofs * (h.aggBegin[m + 1] - h.aggBegin[m]);
            for (int g = h.aggBegin[m]; g < h.aggBegin[m + 1]; ++g)
            {
                info[g] = make_int4(b0, n, aggPtr[g], aggPtr[g + 1]);
                rowOff[g] = static_cast<long long>(h.coarseOffset[m]) +
                            static_cast<long long>(device::kCoarseDofs * g - b0) * n;
            }
        }
        if (R.coarseInfo.n != info.size())
        {
            R.coarseInfo.allocate(info.size());
            R.coarseRowOffset.allocate(rowOff.size());
        }
        PBAT_HIP(hipMemcpyAsync(R.coarseInfo.ptr, info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice,
                                h.stream));
        PBAT_HIP(hipMemcpyAsync(R.coarseRowOffset.ptr, rowOff.data(), rowOff.size() * sizeof(long long),
                                hipMemcpyHostToDevice, h.stream));
        PBAT_HIP(hipStreamSynchronize(h.stream)); // host temporaries
    }
    PBAT_HIP(hipStreamSynchronize(h.stream)); // finalBuf is a host temporary
    hipLaunchKernelGGL(device::CoarseGuard, dim3(nM), dim3(256), 0, h.stream, R.coarse[0].ptr, R.coarse[1].ptr,
                       static_cast<int const*>(R.coarseFinal.ptr), c);
    if (R.coarseF.n != total)
        R.coarseF.allocate(total);
    hipLaunchKernelGGL(device::CoarseToFloat, grid, dim3(256), 0, h.stream,
                       static_cast<double const*>(R.coarse[0].ptr), static_cast<double const*>(R.coarse[1].ptr),
                       static_cast<int const*>(R.coarseFinal.ptr), c, R.coarseF.ptr);
    PBAT_HIP(hipGetLastError());
}

void Zero4(pbat_vbd_integrator& h, float4* p, int n)
{
    if (n > 0)
        hipLaunchKernelGGL(device::Fill4, dim3(Blocks(n)), dim3(256), 0, h.stream, p, n);
}

// Vertex-kernel pass of mode M over [vbeg, vend): y, b, out, acc as the mode uses them.
template <AdjMode M>
void VertexPass(pbat_vbd_integrator& h, float4 const* xs, float dt, float4* y, float4 const* b, float4* out,
                double* acc, int accArray, int vbeg, int vend)
{
    device::AdjointArgs a = MakeAdjointArgs(h, xs, dt);
    a.y = y;
    a.b = b;
    a.out = out;
    a.acc = acc;
    a.accArray = accArray;
    a.nM = h.nMeshes;
    a.vbeg = vbeg;
    a.vend = vend;
    LaunchAdjoint<M>(h, a);
}

// ---- assembled step Hessian (vbd_hessian_kernels.hpp) ------------------------

int BsrLanes(pbat_vbd_integrator const& h);

// Longest block row the row-major assembly takes (4 blocks per lane of 32).
constexpr int kRowAssembleMaxRow = 128;

// Host threads for the pattern build over n independent items: up to 16, one
// per 8,192 items.
int HostThreads(int n)
{
    int const hw = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    return std::max(1, std::min({16, hw, n / 8192}));
}

// f(begin, end, part) over [0, n) cut into nParts contiguous ranges, the same
// ranges for the same (n, nParts), run on nParts host threads; the first
// exception a part throws is rethrown after every thread has joined.
template <class F>
void HostParallel(int n, int nParts, F&& f)
{
    auto range = [&](int t) { return static_cast<int>(int64_t(n) * t / nParts); };
    if (nParts <= 1)
    {
        f(0, n, 0);
        return;
    }
    std::vector<std::exception_ptr> err(static_cast<size_t>(nParts));
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(nParts));
    auto run = [&](int t) {
        try
        {
            f(range(t), range(t + 1), t);
        }
        catch (...)
        {
            err[t] = std::current_exception();
        }
    };
    try
    {
        for (int t = 0; t < nParts; ++t)
            pool.emplace_back(run, t);
    }
    catch (...)
    {
        for (auto& th : pool)
            th.join();
        throw;
    }
    for (auto& th : pool)
        th.join();
    for (auto& e : err)
        if (e)
            std::rethrow_exception(e);
}

// Block pattern of the step Hessian over the updated vertices and each block's
// assembly sources, from the element table; built once per handle and row
// order. Row s stands for updated vertex rowVertex[s]: the vertices in their
// own (colour-major) order for the Gauss-Seidel solvers, or aggregate-major for
// the fused CG, so that an aggregate's rows are consecutive and its updates and
// sums read consecutive entries. Row s lists the rows of the updated vertices
// sharing a tet with rowVertex[s], itself included, ascending.
void BuildBsrPattern(pbat_vbd_integrator& h, bool aggregateOrder)
{
    RolloutState& R = h.rollout;
    int const nUpd = static_cast<int>(h.nUpdated);
    int const order = aggregateOrder ? 1 : 0;
    if (R.bsrOrder == order && R.bsrRowPtr.n == static_cast<size_t>(nUpd) + 1)
        return;
    R.ResetGraph(); // a captured CG chunk holds the previous pattern's buffers
    int64_t const nT = h.nT;
    if (nT >= (int64_t{1} << 27))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "mesh too large for the adjoint's assembled Hessian (2^27 tets)");
    std::vector<int4> TV(static_cast<size_t>(nT));
    PBAT_HIP(hipMemcpy(TV.data(), h.tetv.ptr, TV.size() * sizeof(int4), hipMemcpyDeviceToHost));
    auto vid = [&](int64_t e, int l) {
        int4 const& t = TV[e];
        return l == 0 ? t.x : l == 1 ? t.y : l == 2 ? t.z : t.w;
    };
    // updated vertex -> 4 e + l, ascending e
    std::vector<int> vp(static_cast<size_t>(nUpd) + 1, 0);
    for (int64_t e = 0; e < nT; ++e)
        for (int l = 0; l < 4; ++l)
            if (vid(e, l) < nUpd)
                ++vp[vid(e, l) + 1];
    for (int i = 0; i < nUpd; ++i)
        vp[i + 1] += vp[i];
    std::vector<int> vinc(static_cast<size_t>(vp[nUpd]));
    {
        std::vector<int> fill(vp.begin(), vp.end() - 1);
        for (int64_t e = 0; e < nT; ++e)
            for (int l = 0; l < 4; ++l)
                if (vid(e, l) < nUpd)
                    vinc[fill[vid(e, l)]++] = static_cast<int>(4 * e + l);
    }
    // mesh and aggregate of every updated vertex, and the aggregates' vertices
    std::vector<int> meshOf(static_cast<size_t>(nUpd)), aggOf(static_cast<size_t>(nUpd));
    std::vector<int> aggPtr(static_cast<size_t>(h.nAggregates) + 1, 0), aggVerts(static_cast<size_t>(nUpd));
    if (nUpd > 0)
    {
        PBAT_HIP(hipMemcpy(meshOf.data(), h.meshOf.ptr, sizeof(int) * nUpd, hipMemcpyDeviceToHost));
        PBAT_HIP(hipMemcpy(aggOf.data(), h.aggOf.ptr, sizeof(int) * nUpd, hipMemcpyDeviceToHost));
    }
    if (h.nAggregates > 0)
    {
        PBAT_HIP(hipMemcpy(aggPtr.data(), h.aggVertPtr.ptr, aggPtr.size() * sizeof(int), hipMemcpyDeviceToHost));
        if (nUpd > 0)
            PBAT_HIP(hipMemcpy(aggVerts.data(), h.aggVerts.ptr, sizeof(int) * nUpd, hipMemcpyDeviceToHost));
    }
    // row order: rowVertex[s] = vertex of row s, rowOf its inverse
    std::vector<int> rowVertex(static_cast<size_t>(nUpd)), rowOf(static_cast<size_t>(nUpd));
    for (int i = 0; i < nUpd; ++i)
        rowVertex[i] = aggregateOrder ? aggVerts[i] : i;
    for (int s = 0; s < nUpd; ++s)
        rowOf[rowVertex[s]] = s;
    // rows, and per incidence of a row (in row order) the row positions of its
    // tet's vertices (-1: not updated). Rows are independent: each range of rows
    // is built on its own thread with its own dense row -> position marker
    // (reset after each row), then the ranges' columns are concatenated.
    std::vector<int> rowPtr(static_cast<size_t>(nUpd) + 1, 0), incPtr(static_cast<size_t>(nUpd) + 1, 0);
    std::vector<int> incCode(vinc.size());
    std::vector<int4> incSlot(vinc.size());
    for (int s = 0; s < nUpd; ++s)
        incPtr[s + 1] = incPtr[s] + vp[rowVertex[s] + 1] - vp[rowVertex[s]];
    int const nParts = HostThreads(nUpd);
    std::vector<std::vector<int>> colPart(static_cast<size_t>(nParts));
    std::vector<int> maxPart(static_cast<size_t>(nParts), 0);
    HostParallel(nUpd, nParts, [&](int r0, int r1, int part) {
        std::vector<int> slotOf(static_cast<size_t>(nUpd), -1), cols;
        std::vector<int>& out = colPart[part];
        out.reserve(16 * static_cast<size_t>(r1 - r0));
        for (int s = r0; s < r1; ++s)
        {
            int const v = rowVertex[s];
            // the row's distinct columns (slotOf marks those seen), then sorted
            cols.clear();
            for (int k = vp[v]; k < vp[v + 1]; ++k)
                for (int l = 0; l < 4; ++l)
                {
                    int const j = vid(vinc[k] >> 2, l);
                    if (j < nUpd && slotOf[rowOf[j]] < 0)
                    {
                        slotOf[rowOf[j]] = 0;
                        cols.push_back(rowOf[j]);
                    }
                }
            std::sort(cols.begin(), cols.end());
            for (size_t c = 0; c < cols.size(); ++c)
                slotOf[cols[c]] = static_cast<int>(c);
            for (int k = vp[v], t = incPtr[s]; k < vp[v + 1]; ++k, ++t)
            {
                int s4[4];
                for (int l = 0; l < 4; ++l)
                {
                    int const j = vid(vinc[k] >> 2, l);
                    s4[l] = j < nUpd ? slotOf[rowOf[j]] : -1;
                }
                incCode[t] = vinc[k];
                incSlot[t] = make_int4(s4[0], s4[1], s4[2], s4[3]);
            }
            for (int j : cols)
                slotOf[j] = -1;
            out.insert(out.end(), cols.begin(), cols.end());
            rowPtr[s + 1] = static_cast<int>(cols.size()); // the row's length until the scan below
            maxPart[part] = std::max(maxPart[part], static_cast<int>(cols.size()));
        }
    });
    for (int s = 0; s < nUpd; ++s)
        rowPtr[s + 1] += rowPtr[s];
    int const maxRow = *std::max_element(maxPart.begin(), maxPart.end());
    std::vector<int> col(static_cast<size_t>(rowPtr[nUpd])), blockRow(col.size());
    HostParallel(nUpd, nParts, [&](int r0, int r1, int part) {
        std::copy(colPart[part].begin(), colPart[part].end(), col.begin() + rowPtr[r0]);
        for (int s = r0; s < r1; ++s)
            std::fill(blockRow.begin() + rowPtr[s], blockRow.begin() + rowPtr[s + 1], s);
    });
    colPart.clear();
    size_t const nBlocks = col.size();
    auto forEachSource = [&](auto&& f) {
        for (int s = 0; s < nUpd; ++s)
            for (int t = incPtr[s]; t < incPtr[s + 1]; ++t)
            {
                int const e = incCode[t] >> 2, la = incCode[t] & 3;
                int const s4[4] = {incSlot[t].x, incSlot[t].y, incSlot[t].z, incSlot[t].w};
                for (int lb = 0; lb < 4; ++lb)
                    if (s4[lb] >= 0)
                        f(rowPtr[s] + s4[lb], 16 * e + 4 * la + lb);
            }
    };
    R.bsrRowPtr.upload(rowPtr, h.stream);
    R.bsrCol.upload(col, h.stream);
    R.bsrBlockRow.upload(blockRow, h.stream);
    R.bsrA.allocate(nBlocks), R.bsrB.allocate(nBlocks), R.bsrC.allocate(nBlocks);
    R.rowVertex.upload(rowVertex, h.stream);
    if (maxRow <= kRowAssembleMaxRow)
    {
        // row-major assembly: the rows' incidences and their tets' row positions
        R.incPtr.upload(incPtr, h.stream);
        R.incCode.upload(incCode, h.stream);
        R.incSlot.upload(incSlot, h.stream);
    }
    else
    {
        // rows too long for it: one thread per block over the block's sources
        std::vector<int> srcPtr(nBlocks + 1, 0);
        forEachSource([&](int slot, int) { ++srcPtr[slot + 1]; });
        for (size_t k = 0; k < nBlocks; ++k)
            srcPtr[k + 1] += srcPtr[k];
        std::vector<int> src(static_cast<size_t>(srcPtr[nBlocks]));
        std::vector<int> fill(srcPtr.begin(), srcPtr.end() - 1);
        forEachSource([&](int slot, int code) { src[fill[slot]++] = code; });
        R.bsrSrcPtr.upload(srcPtr, h.stream);
        R.bsrSrc.upload(src, h.stream);
    }
    R.bsrMaxRow = maxRow;

    // mesh and aggregate of every row, each aggregate's rows
    std::vector<int> meshOfRow(static_cast<size_t>(nUpd)), aggOfRow(static_cast<size_t>(nUpd));
    std::vector<int> aggRows(static_cast<size_t>(nUpd));
    for (int s = 0; s < nUpd; ++s)
    {
        meshOfRow[s] = meshOf[rowVertex[s]];
        aggOfRow[s] = aggOf[rowVertex[s]];
    }
    for (int k = 0; k < nUpd; ++k)
        aggRows[k] = rowOf[aggVerts[k]]; // aggregate-major rows: k itself
    R.meshOfRow.upload(meshOfRow, h.stream);
    R.aggOfRow.upload(aggOfRow, h.stream);
    R.aggRows.upload(aggRows, h.stream);
    // aggregate of each block's column (the coarse pattern below)
    std::vector<int> colAgg(nBlocks);
    for (size_t k = 0; k < nBlocks; ++k)
        colAgg[k] = aggOfRow[col[k]];
    // row chunks: every run of rows of one mesh (and, in the vertices' own order,
    // one colour) cut into chunks of at most 256 / L rows, then ordered
    // mesh-major (stable: colour, then row order)
    int const nM = h.nMeshes;
    int const chunkRows = 256 / BsrLanes(h);
    struct Chunk
    {
        int mesh, begin, end;
    };
    std::vector<Chunk> chunks;
    std::vector<int> bounds;
    if (aggregateOrder)
        bounds = {0, nUpd};
    else
        for (int b : h.colorBegin)
            bounds.push_back(std::min(b, nUpd));
    for (size_t c = 0; c + 1 < bounds.size(); ++c)
    {
        int r = bounds[c];
        int const re = bounds[c + 1];
        while (r < re)
        {
            int runEnd = r + 1;
            while (runEnd < re && meshOfRow[runEnd] == meshOfRow[r])
                ++runEnd;
            for (int b = r; b < runEnd; b += chunkRows)
                chunks.push_back({meshOfRow[r], b, std::min(b + chunkRows, runEnd)});
            r = runEnd;
        }
    }
    std::stable_sort(chunks.begin(), chunks.end(), [](Chunk const& a, Chunk const& b) { return a.mesh < b.mesh; });
    std::vector<int4> chunkRange(chunks.size());
    std::vector<int> meshChunk(static_cast<size_t>(nM) + 1, 0);
    for (size_t k = 0; k < chunks.size(); ++k)
    {
        chunkRange[k] = make_int4(chunks[k].begin, chunks[k].end, chunks[k].mesh, 0);
        ++meshChunk[chunks[k].mesh + 1];
    }
    for (int m = 0; m < nM; ++m)
        meshChunk[m + 1] += meshChunk[m];
    R.bsrChunks.upload(chunkRange, h.stream);
    std::vector<int2> aggChunk(static_cast<size_t>(std::max(h.nAggregates, 1)), make_int2(0, 0));
    for (int m = 0; m < nM; ++m)
        for (int g = h.aggBegin[m]; g < h.aggBegin[m + 1]; ++g)
            aggChunk[g] = make_int2(meshChunk[m], meshChunk[m + 1]);
    R.aggChunks.upload(aggChunk, h.stream);
    // each aggregate's first rows, one per AggUpdate thread
    {
        std::vector<int> first(static_cast<size_t>(std::max(h.nAggregates, 1)) * device::kAggFirst, -1);
        for (int g = 0; g < h.nAggregates; ++g)
            for (int k = aggPtr[g]; k < std::min(aggPtr[g + 1], aggPtr[g] + device::kAggFirst); ++k)
                first[static_cast<size_t>(g) * device::kAggFirst + (k - aggPtr[g])] = aggRows[k];
        R.aggFirst.upload(first, h.stream);
        R.aggDone.allocate(static_cast<size_t>(std::max(h.nAggregates, 1)));
    }
    R.bsrChunkRows = chunkRows;
    R.pqPart.allocate(std::max<size_t>(chunks.size(), 1));
    size_t const nAggAlloc = static_cast<size_t>(std::max(h.nAggregates, 1));
    R.rzPart.allocate(nAggAlloc), R.rrPart.allocate(nAggAlloc), R.bbPart.allocate(nAggAlloc);
    R.rczcPart.allocate(static_cast<size_t>(device::kCoarseDofs) * nAggAlloc);

    // coarse blocks: the fine blocks grouped by (aggregate of the row, aggregate
    // of the column), ascending within each group. Built per row aggregate, in
    // parallel ranges of aggregates: its rows ascending and their blocks, binned
    // by the column's aggregate (ascending) in that order.
    int const nAgg = h.nAggregates;
    std::vector<int> aggRowPtr(static_cast<size_t>(nAgg) + 1, 0), aggRowList(static_cast<size_t>(nUpd));
    for (int s = 0; s < nUpd; ++s)
        ++aggRowPtr[aggOfRow[s] + 1];
    for (int a = 0; a < nAgg; ++a)
        aggRowPtr[a + 1] += aggRowPtr[a];
    {
        std::vector<int> fill(aggRowPtr.begin(), aggRowPtr.end() - 1);
        for (int s = 0; s < nUpd; ++s)
            aggRowList[fill[aggOfRow[s]]++] = s;
    }
    struct CoarsePart
    {
        std::vector<int> a, b, len, src;
    };
    int const nAggParts = std::max(1, std::min(nParts, nAgg / 4));
    std::vector<CoarsePart> cparts(static_cast<size_t>(nAggParts));
    HostParallel(nAgg, nAggParts, [&](int a0, int a1, int part) {
        CoarsePart& P = cparts[part];
        std::vector<int> groupOf(static_cast<size_t>(nAgg), -1), keys, start;
        for (int a = a0; a < a1; ++a)
        {
            // distinct column aggregates, ascending
            keys.clear();
            size_t total = 0;
            for (int q = aggRowPtr[a]; q < aggRowPtr[a + 1]; ++q)
            {
                int const s = aggRowList[q];
                total += static_cast<size_t>(rowPtr[s + 1] - rowPtr[s]);
                for (int k = rowPtr[s]; k < rowPtr[s + 1]; ++k)
                    if (groupOf[colAgg[k]] < 0)
                    {
                        groupOf[colAgg[k]] = 0;
                        keys.push_back(colAgg[k]);
                    }
            }
            std::sort(keys.begin(), keys.end());
            for (size_t g = 0; g < keys.size(); ++g)
                groupOf[keys[g]] = static_cast<int>(g);
            // group sizes, then the blocks placed in order
            start.assign(keys.size() + 1, 0);
            for (int q = aggRowPtr[a]; q < aggRowPtr[a + 1]; ++q)
            {
                int const s = aggRowList[q];
                for (int k = rowPtr[s]; k < rowPtr[s + 1]; ++k)
                    ++start[groupOf[colAgg[k]] + 1];
            }
            size_t const base = P.src.size();
            for (size_t g = 0; g < keys.size(); ++g)
            {
                P.a.push_back(a), P.b.push_back(keys[g]), P.len.push_back(start[g + 1]);
                start[g + 1] += start[g];
            }
            P.src.resize(base + total);
            for (int q = aggRowPtr[a]; q < aggRowPtr[a + 1]; ++q)
            {
                int const s = aggRowList[q];
                for (int k = rowPtr[s]; k < rowPtr[s + 1]; ++k)
                    P.src[base + static_cast<size_t>(start[groupOf[colAgg[k]]]++)] = k;
            }
            for (int key : keys)
                groupOf[key] = -1;
        }
    });
    std::vector<int> cbA, cbB, cbPtr{0}, cbSrc;
    cbSrc.reserve(nBlocks);
    for (CoarsePart const& P : cparts)
    {
        cbA.insert(cbA.end(), P.a.begin(), P.a.end());
        cbB.insert(cbB.end(), P.b.begin(), P.b.end());
        for (int len : P.len)
            cbPtr.push_back(cbPtr.back() + len);
        cbSrc.insert(cbSrc.end(), P.src.begin(), P.src.end());
    }
    R.cbA.upload(cbA, h.stream);
    R.cbB.upload(cbB, h.stream);
    R.cbSrcPtr.upload(cbPtr, h.stream);
    R.cbSrc.upload(cbSrc, h.stream);
    PBAT_HIP(hipStreamSynchronize(h.stream));
    R.bsrOrder = order;
}

// H = M / h^2 + sum_e w_e d2Psi_e/dx2 at xs into the block values, and the
// inverse diagonal blocks into dinv.
void AssembleBsr(pbat_vbd_integrator& h, float4 const* xs, float dt, bool aggregateOrder)
{
    BuildBsrPattern(h, aggregateOrder);
    RolloutState& R = h.rollout;
    device::BsrPattern p{};
    p.rowPtr = R.bsrRowPtr.ptr, p.col = R.bsrCol.ptr, p.blockRow = R.bsrBlockRow.ptr;
    p.srcPtr = R.bsrSrcPtr.ptr, p.src = R.bsrSrc.ptr;
    p.nRows = static_cast<int>(h.nUpdated), p.nBlocks = static_cast<int>(R.bsrCol.n);
    if (p.nBlocks == 0)
        return;
    double const invDt2 = 1. / (double(dt) * double(dt));
    device::BsrValues const v{R.bsrA.ptr, R.bsrB.ptr, R.bsrC.ptr};
    if (R.bsrMaxRow > kRowAssembleMaxRow)
    {
        hipLaunchKernelGGL(device::BsrAssemble, dim3(Blocks(p.nBlocks)), dim3(256), 0, h.stream, xs,
                           static_cast<int4 const*>(h.tetv.ptr), static_cast<float4 const*>(h.tgp.ptr), p, invDt2,
                           v, R.dinv.ptr, static_cast<int const*>(R.rowVertex.ptr));
        return;
    }
    device::RowAssembleArgs a{};
    a.xs = xs, a.tetv = h.tetv.ptr, a.tgp = h.tgp.ptr;
    a.incPtr = R.incPtr.ptr, a.incCode = R.incCode.ptr, a.incSlot = R.incSlot.ptr;
    a.rowPtr = R.bsrRowPtr.ptr, a.col = R.bsrCol.ptr, a.nRows = p.nRows, a.invDt2 = invDt2, a.v = v;
    a.dinv = R.dinv.ptr, a.rowVertex = R.rowVertex.ptr;
    dim3 const grid((p.nRows + 1) / 2);
    int const nb = (R.bsrMaxRow + 31) / 32;
    if (nb <= 1)
        hipLaunchKernelGGL(device::RowAssemble<1>, grid, dim3(64), 0, h.stream, a);
    else if (nb == 2)
        hipLaunchKernelGGL(device::RowAssemble<2>, grid, dim3(64), 0, h.stream, a);
    else if (nb == 3)
        hipLaunchKernelGGL(device::RowAssemble<3>, grid, dim3(64), 0, h.stream, a);
    else
        hipLaunchKernelGGL(device::RowAssemble<4>, grid, dim3(64), 0, h.stream, a);
}

// Lanes per row of the block products: the longest row rounded up to a power of
// two (at most 64), one block per lane; longer rows take more per lane.
int BsrLanes(pbat_vbd_integrator const& h)
{
    int const d = std::max(h.rollout.bsrMaxRow, 1);
    int L = 1;
    while (L < 64 && L < d)
        L <<= 1;
    return L;
}

int BsrPerLane(pbat_vbd_integrator const& h)
{
    int const d = std::max(h.rollout.bsrMaxRow, 1);
    return std::clamp((d + BsrLanes(h) - 1) / BsrLanes(h), 1, 4);
}

using device::BsrMode;

template <int L, int K, BsrMode M>
void LaunchBsrLK(pbat_vbd_integrator& h, device::BsrArgs a)
{
    int const block = 256;
    int64_t const threads = static_cast<int64_t>(a.vend - a.vbeg) * L;
    int64_t grid64 = (threads + block - 1) / block;
    a.xcdChunk = 0;
    if (h.xcdRemap && grid64 >= 16)
    {
        grid64 = (grid64 + 7) / 8 * 8;
        a.xcdChunk = static_cast<int>(grid64 / 8);
    }
    if (grid64 * block >= (int64_t{1} << 31))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "row range too large for the block product launch");
    hipLaunchKernelGGL((device::BsrKernel<L, K, M>), dim3(static_cast<int>(grid64)), dim3(block), 0, h.stream, a);
}

template <int L, BsrMode M>
void LaunchBsrL(pbat_vbd_integrator& h, device::BsrArgs const& a)
{
    switch (BsrPerLane(h))
    {
        case 1: LaunchBsrLK<L, 1, M>(h, a); break;
        case 2: LaunchBsrLK<L, 2, M>(h, a); break;
        case 3: LaunchBsrLK<L, 3, M>(h, a); break;
        default: LaunchBsrLK<L, 4, M>(h, a); break;
    }
}

// Block product of mode M over rows [vbeg, vend) (nothing for an empty range).
template <BsrMode M>
void BsrPass(pbat_vbd_integrator& h, device::BsrArgs a, int vbeg, int vend)
{
    if (vend <= vbeg)
        return;
    RolloutState& R = h.rollout;
    a.rowPtr = R.bsrRowPtr.ptr, a.col = R.bsrCol.ptr;
    a.va = R.bsrA.ptr, a.vb = R.bsrB.ptr, a.vc = R.bsrC.ptr;
    a.meshOf = R.meshOfRow.ptr, a.dinv = R.dinv.ptr;
    a.vbeg = vbeg, a.vend = vend;
    switch (BsrLanes(h))
    {
        case 1: LaunchBsrL<1, M>(h, a); break;
        case 2: LaunchBsrL<2, M>(h, a); break;
        case 4: LaunchBsrL<4, M>(h, a); break;
        case 8: LaunchBsrL<8, M>(h, a); break;
        case 16: LaunchBsrL<16, M>(h, a); break;
        case 32: LaunchBsrL<32, M>(h, a); break;
        default: LaunchBsrL<64, M>(h, a); break;
    }
}

template <int L, int K>
void LaunchDirectionLK(pbat_vbd_integrator& h, device::BsrArgs a)
{
    RolloutState& R = h.rollout;
    int const nChunks = static_cast<int>(R.bsrChunks.n);
    int64_t grid64 = std::max(nChunks, 1);
    a.xcdChunk = 0;
    if (h.xcdRemap && grid64 >= 16)
    {
        grid64 = (grid64 + 7) / 8 * 8;
        a.xcdChunk = static_cast<int>(grid64 / 8);
    }
    hipLaunchKernelGGL((device::BsrDirection<L, K>), dim3(static_cast<int>(grid64)), dim3(256), 0, h.stream, a,
                       static_cast<int4 const*>(R.bsrChunks.ptr), nChunks, R.pqPart.ptr);
}

template <int L>
void LaunchDirectionL(pbat_vbd_integrator& h, device::BsrArgs const& a)
{
    switch (BsrPerLane(h))
    {
        case 1: LaunchDirectionLK<L, 1>(h, a); break;
        case 2: LaunchDirectionLK<L, 2>(h, a); break;
        case 3: LaunchDirectionLK<L, 3>(h, a); break;
        default: LaunchDirectionLK<L, 4>(h, a); break;
    }
}

// The fused iteration's direction product over the row chunks (BsrDirection):
// a.y = p (written), a.out = H p, a.z, a.pOld as BsrArgs describes them.
void DirectionPass(pbat_vbd_integrator& h, device::BsrArgs a)
{
    RolloutState& R = h.rollout;
    if (R.bsrChunks.n == 0)
        return;
    if (256 / BsrLanes(h) != R.bsrChunkRows)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "row chunks built for another lane count");
    a.rowPtr = R.bsrRowPtr.ptr, a.col = R.bsrCol.ptr;
    a.va = R.bsrA.ptr, a.vb = R.bsrB.ptr, a.vc = R.bsrC.ptr;
    a.meshOf = R.meshOfRow.ptr;
    switch (BsrLanes(h))
    {
        case 1: LaunchDirectionL<1>(h, a); break;
        case 2: LaunchDirectionL<2>(h, a); break;
        case 4: LaunchDirectionL<4>(h, a); break;
        case 8: LaunchDirectionL<8>(h, a); break;
        case 16: LaunchDirectionL<16>(h, a); break;
        case 32: LaunchDirectionL<32>(h, a); break;
        default: LaunchDirectionL<64>(h, a); break;
    }
}

// One Gauss-Seidel pass over the colours (forward or reverse order) on H y = b
// with the assembled blocks. With acc (a reverse pass after a forward pass,
// which updates each vertex last), the accumulator r.z of each mesh receives
// b . y of the final values.
void GaussSeidelPass(pbat_vbd_integrator& h, device::CgState const& cg, float4* y, float4 const* b, bool reverse,
                     double* acc = nullptr)
{
    int const nC = static_cast<int>(h.colorBegin.size()) - 1;
    device::BsrArgs a{};
    a.y = y, a.b = b, a.cg = cg, a.acc = acc, a.accArray = device::kAccRzNew;
    for (int k = 0; k < nC; ++k)
    {
        int const c = reverse ? nC - 1 - k : k;
        if (acc)
            BsrPass<BsrMode::kGaussSeidelDot>(h, a, h.colorBegin[c], h.colorBegin[c + 1]);
        else
            BsrPass<BsrMode::kGaussSeidel>(h, a, h.colorBegin[c], h.colorBegin[c + 1]);
    }
}

struct SolveResult
{
    int iterations = 0;         // largest per-mesh iteration count
    double relResidual = 0.;    // largest per-mesh |r| / |b| at its last iteration
};

// Meshes whose solve has not converged yet (their done flags).
int ReadLive(pbat_vbd_integrator& h)
{
    int const nM = h.nMeshes;
    std::vector<int> done(static_cast<size_t>(nM));
    PBAT_HIP(hipMemcpyAsync(done.data(), h.rollout.flags.ptr + device::kDone * nM, sizeof(int) * nM,
                            hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    return static_cast<int>(std::count(done.begin(), done.end(), 0));
}

// Host-mapped copy of the per-mesh done flags (written by CgFinish) and the
// chunk-end events, allocated on first use.
void EnsureDoneHost(pbat_vbd_integrator& h)
{
    RolloutState& R = h.rollout;
    int const nM = h.nMeshes;
    if (R.doneHostMeshes < nM)
    {
        if (R.doneHost)
            PBAT_HIP(hipHostFree(R.doneHost));
        R.doneHost = nullptr, R.doneDev = nullptr, R.doneHostMeshes = 0;
        void* p = nullptr;
        if (hipHostMalloc(&p, sizeof(int) * static_cast<size_t>(nM), hipHostMallocMapped | hipHostMallocCoherent) !=
                hipSuccess ||
            !p)
            Fail(PBAT_VBD_ERROR_OUT_OF_MEMORY, "hipHostMalloc of the live-mesh flags failed");
        R.doneHost = static_cast<int*>(p), R.doneHostMeshes = nM;
        void* dp = nullptr;
        PBAT_HIP(hipHostGetDevicePointer(&dp, p, 0));
        R.doneDev = static_cast<int*>(dp);
    }
    for (hipEvent_t& e : R.chunkEvent)
        if (!e)
            PBAT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// Wait for the chunk that event `slot` ends and count the meshes still
// iterating. The flags may already hold a later chunk's values (a mesh stays
// done once done within a solve).
int LiveAfterChunk(pbat_vbd_integrator& h, int slot)
{
    RolloutState& R = h.rollout;
    PBAT_HIP(hipEventSynchronize(R.chunkEvent[slot]));
    int live = 0;
    for (int m = 0; m < h.nMeshes; ++m)
        live += __atomic_load_n(R.doneHost + m, __ATOMIC_ACQUIRE) == 0;
    return live;
}

SolveResult ReadResult(pbat_vbd_integrator& h)
{
    int const nM = h.nMeshes;
    RolloutState& R = h.rollout;
    std::vector<double> S(static_cast<size_t>(device::kPlainCount * nM));
    std::vector<int> F(static_cast<size_t>(2 * nM + 1));
    PBAT_HIP(hipMemcpyAsync(S.data(), R.scal.ptr, S.size() * sizeof(double), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipMemcpyAsync(F.data(), R.flags.ptr, F.size() * sizeof(int), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    SolveResult out;
    for (int m = 0; m < nM; ++m)
    {
        out.iterations = std::max(out.iterations, F[device::kIters * nM + m]);
        double const bb = S[device::kBb * nM + m];
        if (bb > 0.)
            out.relResidual = std::max(out.relResidual, std::sqrt(S[device::kRrLast * nM + m] / bb));
    }
    return out;
}

// Solve H y = b (y holds the initial guess) for every mesh of the handle, each
// with its own CG scalars and stopping test |b - H y| <= tol |b|. H is assembled
// once per solve (AssembleBsr); the host polls the number of live meshes every
// cfg.check_every iterations, and the iteration kernels do nothing once every
// mesh has converged.
//
// Block Jacobi and two-level preconditioners (the fused iteration), per iteration:
//   (1) per row chunk: p = z + beta p_old formed on the fly, q = H p, the
//       chunk's p . q;
//   (2) one workgroup per aggregate: alpha = rz / pq from its mesh's chunk sums,
//       y += alpha p, r -= alpha q, z = D^-1 r, the aggregate's r . z, r . r and
//       restriction rc = Z^T r;
//   (3) two-level, one workgroup per aggregate: its rows of zc = (Z^T H Z)^-1 rc,
//       each coarse row's rc . zc, and z += Z zc on its rows;
//   (4) one workgroup: per mesh, r . z = r . D^-1 r + rc . zc and r . r summed
//       over its partials, stopping test, beta.
// No atomics: every sum is taken in a fixed order, so a solve gives the same
// bits on every run, and a mesh gives the same bits alone or in a batch (its
// row chunks, aggregates and coarse rows are the same).
// The symmetric Gauss-Seidel preconditioner keeps the four-launch iteration
//   q = H p and p . q; y, r update with z = 0; a forward and a reverse coloured
//   Gauss-Seidel pass for z with r . z; one workgroup: scalars; p = z + beta p.
//
// With graph capture on (the integrator's setting), each chunk of iterations is
// captured once into a hipGraph and replayed; the graph's kernels take the
// handle's fixed work vectors (x = hold, y, b = rhs), so it stays valid across
// solves with the same settings.
// y = the H-norm projection of the solution onto the span of the nHist latest
// entries of the solution history (vbd_hessian_kernels.hpp: ProjBasis .. ProjCombine),
// H assembled.
void ProjectInitialGuess(pbat_vbd_integrator& h, float4* y, float4 const* b, int nHist)
{
    RolloutState& R = h.rollout;
    int const n = static_cast<int>(h.nUpdated), nM = h.nMeshes;
    device::ProjArgs a{};
    for (int j = 0; j < device::kProjMax; ++j)
    {
        int const slot = ((R.histHead - 1 - j) % device::kProjMax + device::kProjMax) % device::kProjMax;
        a.hist[j] = R.hist[slot].ptr;
        a.basis[j] = R.projBasis[j].ptr;
        a.hb[j] = R.projHb[j].ptr;
    }
    a.rhs = b, a.y = y;
    a.aggPtr = h.aggVertPtr.ptr, a.aggVerts = R.aggRows.ptr, a.meshAggBegin = h.meshAggBegin.ptr;
    a.meshOf = R.meshOfRow.ptr;
    a.part = R.projPart.ptr, a.coef = R.projCoef.ptr;
    a.k = nHist, a.n = n, a.nM = nM, a.nAgg = h.nAggregates;
    hipLaunchKernelGGL(device::ProjBasis, dim3(Blocks(n)), dim3(256), 0, h.stream, a);
    hipLaunchKernelGGL(device::ProjProducts, dim3(Blocks(static_cast<int64_t>(n) * 16)), dim3(256), 0, h.stream, a,
                       static_cast<int const*>(R.bsrRowPtr.ptr), static_cast<int const*>(R.bsrCol.ptr),
                       static_cast<float4 const*>(R.bsrA.ptr), static_cast<float4 const*>(R.bsrB.ptr),
                       static_cast<float const*>(R.bsrC.ptr));
    hipLaunchKernelGGL(device::ProjGram, dim3(h.nAggregates), dim3(128), 0, h.stream, a);
    hipLaunchKernelGGL(device::ProjSolve, dim3(nM), dim3(256), 0, h.stream, a);
    hipLaunchKernelGGL(device::ProjCombine, dim3(Blocks(n)), dim3(256), 0, h.stream, a);
}

SolveResult Solve(pbat_vbd_integrator& h, float4 const* xs, float dt, float4* y, float4 const* b,
                  pbat_vbd_adjoint_config const& cfg, int nHist, double* statsOut)
{
    RolloutState& R = h.rollout;
    int const n = static_cast<int>(h.nUpdated), nV = static_cast<int>(h.nV), nM = h.nMeshes;
    device::CgState cg{};
    cg.P = R.scal.ptr;
    cg.A = R.scal.ptr + PlainDoubles(h);
    cg.F = R.flags.ptr;
    cg.nM = nM;
    cg.tol2 = static_cast<double>(cfg.tolerance) * cfg.tolerance;
    int const gridN = Blocks(n);
    bool const twoLevel = cfg.solver == PBAT_VBD_ADJOINT_PCG_TWO_LEVEL && h.nAggregates > 0;
    bool const fused = cfg.solver == PBAT_VBD_ADJOINT_PCG_JACOBI || cfg.solver == PBAT_VBD_ADJOINT_PCG_TWO_LEVEL;
    // fused chunks hold an even number of iterations: the two direction buffers swap
    int const check = fused ? (std::max(1, cfg.check_every) + 1) / 2 * 2 : std::max(1, cfg.check_every);
    AssembleBsr(h, xs, dt, fused);
    // the fused solvers work on aggregate-major rows: y and b gathered into that
    // order here, y scattered back at the end
    float4* const yIn = y;
    bool const fixedBuffers = xs == R.hold.ptr && y == R.y.ptr && b == R.rhs.ptr;
    if (fused && n > 0)
    {
        hipLaunchKernelGGL(device::GatherRows, dim3(gridN), dim3(256), 0, h.stream, R.yS.ptr,
                           static_cast<float4 const*>(y), static_cast<int const*>(R.rowVertex.ptr), n);
        hipLaunchKernelGGL(device::GatherRows, dim3(gridN), dim3(256), 0, h.stream, R.bS.ptr, b,
                           static_cast<int const*>(R.rowVertex.ptr), n);
        y = R.yS.ptr;
        b = R.bS.ptr;
    }
    if (fused && nHist > 0 && n > 0 && h.nAggregates > 0)
        ProjectInitialGuess(h, y, b, nHist);
    auto zeroScalars = [&] {
        int const nd = static_cast<int>(ScalarDoubles(h));
        hipLaunchKernelGGL(device::ZeroScalars, dim3(std::min(1024, Blocks(nd + 2 * nM + 1))), dim3(256), 0,
                           h.stream, cg.P, nd, cg.F, 2 * nM + 1);
    };
    auto product = [&](float4* v, float4* out, double* acc) {
        device::BsrArgs a{};
        a.y = v, a.out = out, a.cg = cg, a.acc = acc, a.accArray = device::kAccPq;
        if (acc)
            BsrPass<BsrMode::kProductDot>(h, a, 0, n);
        else
            BsrPass<BsrMode::kProduct>(h, a, 0, n);
    };
    zeroScalars();
    Zero4(h, R.z.ptr, nV);
    if (cfg.solver == PBAT_VBD_ADJOINT_GAUSS_SEIDEL)
    {
        int it = 0;
        while (it < cfg.max_iterations)
        {
            int const chunk = std::min(check, cfg.max_iterations - it);
            for (int k = 0; k < chunk; ++k)
                GaussSeidelPass(h, cg, y, b, false);
            it += chunk;
            zeroScalars();
            product(y, R.q.ptr, nullptr);
            hipLaunchKernelGGL(device::CgInitResidual, dim3(gridN), dim3(256), 0, h.stream, b, R.q.ptr, R.r.ptr,
                               R.z.ptr, static_cast<float4 const*>(nullptr), h.meshOf.ptr, n, nM, cg.A);
            hipLaunchKernelGGL(device::CgScalars, dim3(1), dim3(256), 0, h.stream, cg, 1);
            if (ReadLive(h) == 0)
                break;
        }
        SolveResult out = ReadResult(h);
        out.iterations = it;
        PBAT_HIP(hipGetLastError());
        return out;
    }
    auto aggUpdate = [&](int init, float4 const* p) {
        device::AggUpdateArgs u{};
        u.aggPtr = h.aggVertPtr.ptr, u.aggVerts = R.aggRows.ptr, u.aggMesh = h.aggMesh.ptr;
        u.y = y, u.r = R.r.ptr, u.z = R.z.ptr, u.p = p, u.q = R.q.ptr, u.b = b, u.dinv = R.dinv.ptr;
        u.rc = twoLevel ? R.rc.ptr : nullptr;
        u.coarseD = R.coarseD.ptr;
        u.pqPart = R.pqPart.ptr, u.aggChunks = R.aggChunks.ptr;
        u.aggFirst = R.aggFirst.ptr, u.aggDone = R.aggDone.ptr;
        u.rzPart = R.rzPart.ptr, u.rrPart = R.rrPart.ptr, u.bbPart = R.bbPart.ptr;
        u.cg = cg, u.init = init;
        if (h.nAggregates > 0)
            hipLaunchKernelGGL(device::AggUpdate, dim3(h.nAggregates), dim3(128), 0, h.stream, u);
    };
    auto coarseRows = [&](int first) {
        device::CoarseApplyArgs ca{};
        ca.coarseF = R.coarseF.ptr, ca.rc = R.rc.ptr, ca.zc = R.zc.ptr, ca.rczcPart = R.rczcPart.ptr;
        ca.info = R.coarseInfo.ptr, ca.rowOffset = R.coarseRowOffset.ptr, ca.aggDone = R.aggDone.ptr;
        ca.coarseD = R.coarseD.ptr, ca.z = R.z.ptr;
        // 64-column chunks per wave of the largest mesh, rounded up to a kernel's
        int const perWave = ((R.coarseMaxN + 63) / 64 + device::kCoarseWaves - 1) / device::kCoarseWaves;
        dim3 const grid(h.nAggregates), block(64 * device::kCoarseWaves);
        if (perWave <= 1)
            hipLaunchKernelGGL(device::CoarseApplyRows<1>, grid, block, 0, h.stream, ca, first);
        else if (perWave <= 2)
            hipLaunchKernelGGL(device::CoarseApplyRows<2>, grid, block, 0, h.stream, ca, first);
        else if (perWave <= 3)
            hipLaunchKernelGGL(device::CoarseApplyRows<3>, grid, block, 0, h.stream, ca, first);
        else if (perWave <= 4)
            hipLaunchKernelGGL(device::CoarseApplyRows<4>, grid, block, 0, h.stream, ca, first);
        else if (perWave <= 6)
            hipLaunchKernelGGL(device::CoarseApplyRows<6>, grid, block, 0, h.stream, ca, first);
        else if (perWave <= 9)
            hipLaunchKernelGGL(device::CoarseApplyRows<9>, grid, block, 0, h.stream, ca, first);
        else // several passes beyond 3,072 rows
            hipLaunchKernelGGL(device::CoarseApplyRows<12>, grid, block, 0, h.stream, ca, first);
    };
    auto finish = [&](int first) {
        device::CgFinishArgs f{};
        f.cg = cg, f.meshAggBegin = h.meshAggBegin.ptr;
        f.rzPart = R.rzPart.ptr, f.rrPart = R.rrPart.ptr, f.bbPart = R.bbPart.ptr;
        f.rczcPart = twoLevel ? R.rczcPart.ptr : nullptr;
        f.doneOut = R.doneDev, f.aggDone = R.aggDone.ptr;
        hipLaunchKernelGGL(device::CgFinish, dim3(nM), dim3(256), 0, h.stream, f, first);
    };
    std::function<void(int)> iteration;
    if (fused)
    {
        EnsureDoneHost(h);
        // r = b - H y, z = D^-1 r, (rc, zc); p_old = 0 for the first direction
        product(y, R.q.ptr, nullptr);
        aggUpdate(1, nullptr);
        if (twoLevel && h.nAggregates > 0)
            coarseRows(1);
        finish(1);
        Zero4(h, R.p.ptr, nV);
        iteration = [&](int k) {
            float4* const pOld = (k & 1) ? R.p2.ptr : R.p.ptr;
            float4* const pNew = (k & 1) ? R.p.ptr : R.p2.ptr;
            device::BsrArgs a{};
            a.y = pNew, a.out = R.q.ptr, a.z = R.z.ptr, a.pOld = pOld;
            a.cg = cg;
            DirectionPass(h, a);
            aggUpdate(0, pNew);
            if (twoLevel && h.nAggregates > 0)
                coarseRows(0);
            finish(0);
        };
    }
    else
    {
        product(y, R.q.ptr, nullptr);
        hipLaunchKernelGGL(device::CgInitResidual, dim3(gridN), dim3(256), 0, h.stream, b, R.q.ptr, R.r.ptr, R.z.ptr,
                           static_cast<float4 const*>(nullptr), h.meshOf.ptr, n, nM, cg.A);
        GaussSeidelPass(h, cg, R.z.ptr, R.r.ptr, false);
        GaussSeidelPass(h, cg, R.z.ptr, R.r.ptr, true, cg.A);
        hipLaunchKernelGGL(device::CgScalars, dim3(1), dim3(256), 0, h.stream, cg, 1);
        hipLaunchKernelGGL(device::Copy4, dim3(Blocks(nV)), dim3(256), 0, h.stream, R.p.ptr, R.z.ptr, nV);
        iteration = [&](int) {
            product(R.p.ptr, R.q.ptr, cg.A);
            hipLaunchKernelGGL(device::CgUpdate, dim3(gridN), dim3(256), 0, h.stream, y, R.r.ptr, R.z.ptr,
                               static_cast<float4 const*>(R.p.ptr), static_cast<float4 const*>(R.q.ptr),
                               static_cast<float4 const*>(nullptr), h.meshOf.ptr, n, cg);
            GaussSeidelPass(h, cg, R.z.ptr, R.r.ptr, false);
            GaussSeidelPass(h, cg, R.z.ptr, R.r.ptr, true, cg.A);
            hipLaunchKernelGGL(device::CgScalars, dim3(1), dim3(256), 0, h.stream, cg, 0);
            hipLaunchKernelGGL(device::CgDirection, dim3(gridN), dim3(256), 0, h.stream, R.p.ptr,
                               static_cast<float4 const*>(R.z.ptr), h.meshOf.ptr, n, cg);
        };
    }
    bool const graph = h.useGraph && fixedBuffers;
    if (graph)
    {
        std::vector<double> const key{double(dt),        double(cfg.solver), double(check),       cg.tol2,
                                      double(BsrLanes(h)), double(BsrPerLane(h)), double(h.xcdRemap),
                                      double(R.coarseMaxN)};
        // every device address the captured kernels take: a buffer reallocated at
        // another address forces a new capture
        std::vector<void const*> const ptrs{
            R.y.ptr,           R.r.ptr,         R.z.ptr,       R.p.ptr,          R.p2.ptr,        R.q.ptr,
            R.rhs.ptr,         R.hold.ptr,      R.dinv.ptr,    R.scal.ptr,       R.flags.ptr,     R.coarse[0].ptr,
            R.coarse[1].ptr,   R.coarseF.ptr,     R.rc.ptr,    R.zc.ptr,         R.bsrRowPtr.ptr, R.bsrCol.ptr,
            R.bsrA.ptr,        R.bsrB.ptr,      R.bsrC.ptr,    h.meshOf.ptr,     h.aggOf.ptr,     h.aggVertPtr.ptr,
            h.aggVerts.ptr,    h.aggMesh.ptr,   h.meshAggBegin.ptr, R.bsrChunks.ptr, R.aggChunks.ptr,
            R.pqPart.ptr,      R.rzPart.ptr,    R.rrPart.ptr,  R.bbPart.ptr,     R.rczcPart.ptr,
            R.doneDev,       R.aggFirst.ptr, R.aggDone.ptr,    R.yS.ptr,        R.bS.ptr,
            R.aggRows.ptr,     R.aggOfRow.ptr,  R.meshOfRow.ptr, R.coarseD.ptr, R.coarseInfo.ptr,
            R.coarseRowOffset.ptr};
        if (!R.cgExec || R.cgKey != key || R.cgPtrs != ptrs)
        {
            R.ResetGraph();
            PBAT_HIP(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < check; ++k)
                iteration(k);
            hipGraph_t g = nullptr;
            PBAT_HIP(hipStreamEndCapture(h.stream, &g));
            R.cgGraph = g;
            PBAT_HIP(hipGraphInstantiate(&R.cgExec, g, nullptr, nullptr, 0));
            R.cgKey = key;
            R.cgPtrs = ptrs;
        }
    }
    // Fused iterations: the host polls the live-mesh count one chunk behind.
    // Chunk c + 1 is queued before the host waits for the end of chunk c and
    // reads the done flags the finish kernel mirrors into host memory, so the
    // device does not idle during the poll. A chunk queued after every mesh has
    // converged (or that starts with every mesh converged) changes nothing: each
    // of its kernels returns at once for a mesh whose done flag is set.
    auto launchChunk = [&](int it) {
        int const chunk = std::min(check, cfg.max_iterations - it);
        if (graph && chunk == check)
            PBAT_HIP(hipGraphLaunch(R.cgExec, h.stream));
        else
            for (int k = 0; k < chunk; ++k)
                iteration(it + k);
        return it + chunk;
    };
    int it = 0;
    if (!fused)
    {
        // the matrix-free iteration keeps its flags on the device only: poll
        // before every chunk
        while (it < cfg.max_iterations && ReadLive(h) > 0)
            it = launchChunk(it);
    }
    else
    {
        it = launchChunk(0);
        for (int slot = 0; it < cfg.max_iterations; slot ^= 1)
        {
            PBAT_HIP(hipEventRecord(R.chunkEvent[slot], h.stream));
            it = launchChunk(it);
            if (LiveAfterChunk(h, slot) == 0)
                break;
        }
    }
    if (fused && n > 0)
        hipLaunchKernelGGL(device::ScatterRows, dim3(gridN), dim3(256), 0, h.stream, yIn,
                           static_cast<float4 const*>(y), static_cast<int const*>(R.rowVertex.ptr), n);
    PBAT_HIP(hipGetLastError());
    if (statsOut)
    {
        // the fused solvers' statistics stay on the device: no host round trip
        hipLaunchKernelGGL(device::SolveStats, dim3(1), dim3(64), 0, h.stream, cg, statsOut);
        return SolveResult{};
    }
    return ReadResult(h);
}
void CheckConfig(pbat_vbd_adjoint_config const& cfg)
{
    if (cfg.max_iterations < 1 || !(cfg.tolerance >= 0.f) || cfg.history < 0 || cfg.history > device::kProjMax ||
        (cfg.solver != PBAT_VBD_ADJOINT_PCG && cfg.solver != PBAT_VBD_ADJOINT_GAUSS_SEIDEL &&
         cfg.solver != PBAT_VBD_ADJOINT_PCG_JACOBI && cfg.solver != PBAT_VBD_ADJOINT_PCG_TWO_LEVEL))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid adjoint solver configuration");
}

void SaveState(pbat_vbd_integrator& h, float4* x, float4* v, float4* vp)
{
    size_t const bytes = sizeof(float4) * static_cast<size_t>(h.nV);
    PBAT_HIP(hipMemcpyAsync(x, h.xm.ptr, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipMemcpyAsync(v, h.vel.ptr, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipMemcpyAsync(vp, h.vprev.ptr, bytes, hipMemcpyDeviceToDevice, h.stream));
}

// Restore (x, v, v_prev); x_t equals x at a step boundary.
void RestoreState(pbat_vbd_integrator& h, float4 const* x, float4 const* v, float4 const* vp)
{
    size_t const bytes = sizeof(float4) * static_cast<size_t>(h.nV);
    PBAT_HIP(hipMemcpyAsync(h.xm.ptr, x, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipMemcpyAsync(h.xt.ptr, x, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipMemcpyAsync(h.vel.ptr, v, bytes, hipMemcpyDeviceToDevice, h.stream));
    PBAT_HIP(hipMemcpyAsync(h.vprev.ptr, vp, bytes, hipMemcpyDeviceToDevice, h.stream));
}

// The adjoint differentiates the elastic step only.
void NoContact(pbat_vbd_integrator const& h)
{
    if (h.contact)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "differentiable rollouts do not support contact");
}

void RolloutForward(pbat_vbd_integrator& h, float dt, int iterations, int steps, int every)
{
    if (!(dt > 0.f) || iterations < 0 || steps < 1 || every < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "rollout needs dt > 0, iterations >= 0, steps >= 1, every >= 1");
    NoContact(h);
    PBAT_HIP(hipSetDevice(h.device));
    RolloutState& R = h.rollout;
    R.Discard();
    int const nCkpt = (steps + every - 1) / every;
    size_t const n = static_cast<size_t>(h.nV);
    R.ckptX.resize(static_cast<size_t>(nCkpt));
    R.ckptV.resize(static_cast<size_t>(nCkpt));
    R.ckptVprev.resize(static_cast<size_t>(nCkpt));
    for (int c = 0; c < nCkpt; ++c)
    {
        if (R.ckptX[c].n != n)
            R.ckptX[c].allocate(n), R.ckptV[c].allocate(n), R.ckptVprev[c].allocate(n);
    }
    ClearStatus(h);
    for (int t = 0; t < steps; ++t)
    {
        if (t % every == 0)
        {
            int const c = t / every;
            SaveState(h, R.ckptX[c].ptr, R.ckptV[c].ptr, R.ckptVprev[c].ptr);
        }
        Step(h, dt, iterations, 1, false);
    }
    PBAT_HIP(hipStreamSynchronize(h.stream));
    CheckStatus(h);
    if (R.lastX.n != n)
        R.lastX.allocate(n);
    PBAT_HIP(hipMemcpyAsync(R.lastX.ptr, h.xm.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, h.stream));
    R.dt = dt, R.iterations = iterations, R.steps = steps, R.every = every;
    R.valid = true;
}

void RolloutBackward(pbat_vbd_integrator& h, float const* dLdx, float const* dLdv, float* dLogY, float* dv0,
                     float* dx0, pbat_vbd_adjoint_config const* cfgIn, pbat_vbd_adjoint_stats* stats)
{
    RolloutState& R = h.rollout;
    if (!R.valid)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "rollout_backward needs a preceding rollout_forward");
    if (h.kD != 0.f)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "the adjoint assumes no Rayleigh damping (kD = 0)");
    if (!dLdx)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "null dL/dx");
    pbat_vbd_adjoint_config cfg{PBAT_VBD_ADJOINT_PCG, 200, 1e-6f, 5, 0};
    if (cfgIn)
        cfg = *cfgIn;
    CheckConfig(cfg);
    PBAT_HIP(hipSetDevice(h.device));
    EnsureBuffers(h);
    int const nV = static_cast<int>(h.nV), nUpd = static_cast<int>(h.nUpdated), nT = static_cast<int>(h.nT);
    size_t const n = static_cast<size_t>(nV);
    float const dt = R.dt;
    // end state, restored at the end
    DeviceArray<float4> endX, endV, endVp;
    endX.allocate(n), endV.allocate(n), endVp.allocate(n);
    SaveState(h, endX.ptr, endV.ptr, endVp.ptr);
    // loss gradients of the final state (updated vertices only)
    UploadVec(h, R.xbar.ptr, dLdx, nullptr);
    if (dLdv)
        UploadVec(h, R.vbar.ptr, dLdv, nullptr);
    else
        Zero4(h, R.vbar.ptr, nV);
    if (nV > nUpd)
    {
        Zero4(h, R.xbar.ptr + nUpd, nV - nUpd);
        Zero4(h, R.vbar.ptr + nUpd, nV - nUpd);
    }
    Zero4(h, R.y.ptr, nV);
    R.histHead = R.histCount = 0;
    bool const fused = cfg.solver == PBAT_VBD_ADJOINT_PCG_JACOBI || cfg.solver == PBAT_VBD_ADJOINT_PCG_TWO_LEVEL;
    if (fused && R.solveStats.n < 2 * static_cast<size_t>(R.steps))
        R.solveStats.allocate(2 * static_cast<size_t>(R.steps));
    hipLaunchKernelGGL(device::ZeroDoubles, dim3(Blocks(std::max(nT, 1))), dim3(256), 0, h.stream, R.grad.ptr, nT);
    if (R.segX.size() != static_cast<size_t>(R.every))
        R.segX.resize(static_cast<size_t>(R.every));
    for (auto& s : R.segX)
        if (s.n != n)
            s.allocate(n);
    if (cfg.solver == PBAT_VBD_ADJOINT_PCG_TWO_LEVEL)
        BuildCoarse(h, h.xm.ptr, dt); // at x_T; reused by every step's solve
    pbat_vbd_adjoint_stats st{0, 0, 0, 0.f};
    int const nSeg = static_cast<int>(R.ckptX.size());
    for (int sgm = nSeg - 1; sgm >= 0; --sgm)
    {
        // the segment's minimisers: all but the last recomputed from its checkpoint
        // (bitwise equal to the forward), the last one being the positions the next
        // checkpoint (or the end of the rollout) saved
        int const t0 = sgm * R.every, t1 = std::min(R.steps, t0 + R.every);
        if (t1 - t0 > 1)
            RestoreState(h, R.ckptX[sgm].ptr, R.ckptV[sgm].ptr, R.ckptVprev[sgm].ptr);
        // the recomputed steps repeat the forward's bit for bit: no status checks
        for (int t = t0; t < t1 - 1; ++t)
        {
            Step(h, dt, R.iterations, 1, false);
            PBAT_HIP(hipMemcpyAsync(R.segX[t - t0].ptr, h.xm.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                    h.stream));
        }
        float4 const* last = (sgm + 1 < nSeg) ? R.ckptX[sgm + 1].ptr : R.lastX.ptr;
        PBAT_HIP(hipMemcpyAsync(R.segX[t1 - 1 - t0].ptr, last, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                h.stream));
        for (int t = t1 - 1; t >= t0; --t)
        {
            // the step's minimiser, in the solver's fixed buffer
            PBAT_HIP(hipMemcpyAsync(R.hold.ptr, R.segX[t - t0].ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                    h.stream));
            float4 const* xs = R.hold.ptr;
            hipLaunchKernelGGL(device::AdjointRhs, dim3(Blocks(nV)), dim3(256), 0, h.stream, R.xbar.ptr, R.vbar.ptr,
                               R.rhs.ptr, nUpd, nV, 1.f / dt);
            // initial guess: the previous solution, or (fused solvers) the
            // projection onto the span of the last kProjMax solutions
            double* const statsOut = fused ? R.solveStats.ptr + 2 * static_cast<size_t>(st.solves) : nullptr;
            SolveResult const sr =
                Solve(h, xs, dt, R.y.ptr, R.rhs.ptr, cfg, std::min(cfg.history, R.histCount), statsOut);
            if (cfg.history > 0)
            {
                hipLaunchKernelGGL(device::Copy4, dim3(Blocks(nV)), dim3(256), 0, h.stream,
                                   R.hist[R.histHead].ptr, static_cast<float4 const*>(fused ? R.yS.ptr : R.y.ptr),
                                   nV);
                R.histHead = (R.histHead + 1) % device::kProjMax;
                R.histCount = std::min(R.histCount + 1, device::kProjMax);
            }
            ++st.solves;
            if (!fused)
            {
                st.total_iterations += sr.iterations;
                st.max_iterations = std::max(st.max_iterations, sr.iterations);
                st.max_relative_residual = std::max(st.max_relative_residual, static_cast<float>(sr.relResidual));
            }
            hipLaunchKernelGGL(device::LogYoungGradient, dim3(Blocks(std::max(nT, 1))), dim3(256), 0, h.stream, xs,
                               R.y.ptr, h.tetv.ptr, h.tgp.ptr, nT, R.grad.ptr);
            hipLaunchKernelGGL(device::AdjointStepBack, dim3(Blocks(nV)), dim3(256), 0, h.stream, xs, R.y.ptr,
                               R.xbar.ptr, R.vbar.ptr, nUpd, nV, dt);
        }
    }
    PBAT_HIP(hipGetLastError());
    if (fused && st.solves > 0)
    {
        std::vector<double> S(2 * static_cast<size_t>(st.solves));
        PBAT_HIP(hipMemcpyAsync(S.data(), R.solveStats.ptr, S.size() * sizeof(double), hipMemcpyDeviceToHost,
                                h.stream));
        PBAT_HIP(hipStreamSynchronize(h.stream));
        for (int k = 0; k < st.solves; ++k)
        {
            int const it = static_cast<int>(S[2 * k]);
            st.total_iterations += it;
            st.max_iterations = std::max(st.max_iterations, it);
            st.max_relative_residual = std::max(st.max_relative_residual, static_cast<float>(S[2 * k + 1]));
        }
    }
    RestoreState(h, endX.ptr, endV.ptr, endVp.ptr);
    if (dv0)
        DownloadVec(h, R.vbar.ptr, dv0);
    if (dx0)
        DownloadVec(h, R.xbar.ptr, dx0);
    if (dLogY)
    {
        std::vector<double> g(static_cast<size_t>(nT));
        PBAT_HIP(hipMemcpyAsync(g.data(), R.grad.ptr, sizeof(double) * nT, hipMemcpyDeviceToHost, h.stream));
        PBAT_HIP(hipStreamSynchronize(h.stream));
        for (int e = 0; e < nT; ++e)
            dLogY[e] = static_cast<float>(g[e]);
    }
    PBAT_HIP(hipStreamSynchronize(h.stream));
    if (stats)
        *stats = st;
}

void SetLogYoung(pbat_vbd_integrator& h, float const* logY, int64_t nT)
{
    if (!logY || nT != h.nT)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "expected |#elements| = " + std::to_string(h.nT) + " values");
    for (int64_t e = 0; e < nT; ++e)
        if (!std::isfinite(logY[e]))
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "non-finite log Young's modulus at element " + std::to_string(e));
    PBAT_HIP(hipSetDevice(h.device));
    RolloutState& R = h.rollout;
    std::vector<float> v(logY, logY + nT);
    R.logY.upload(v, h.stream);
    bool const ell = h.ellStride > 0;
    auto set = [](int4 const* slots, IncidenceRecords& r) {
        return device::RecordSet{slots, r.ids.ptr, r.g2.ptr, r.sids.ptr, r.sgeo.ptr, r.shape.ptr};
    };
    device::RecordSet const rows = set(h.slotsRows.ptr, h.rows);
    device::RecordSet const ellRows = ell ? set(h.slotsEll.ptr, h.ell) : device::RecordSet{};
    hipLaunchKernelGGL(device::SetLogYoung, dim3(Blocks(nT)), dim3(256), 0, h.stream, R.logY.ptr, h.matBase.ptr,
                       h.tgp.ptr, rows, ellRows, static_cast<int>(nT));
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipStreamSynchronize(h.stream));
    for (int64_t e = 0; e < nT; ++e)
    {
        // (mu, lambda) of the traces follow the new moduli
        float const scale = std::exp(v[e] - h.logYCurrent[e]);
        h.lameHost[2 * e] *= scale;
        h.lameHost[2 * e + 1] *= scale;
    }
    h.logYCurrent = v;
    R.Discard();
}

void HessianVectorProduct(pbat_vbd_integrator& h, float const* x, float dt, float const* p, float* out)
{
    if (!x || !p || !out || !(dt > 0.f))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid Hessian-vector arguments");
    NoContact(h);
    PBAT_HIP(hipSetDevice(h.device));
    EnsureBuffers(h);
    RolloutState& R = h.rollout;
    UploadVec(h, R.hold.ptr, x, h.massOrig.data());
    UploadVec(h, R.p.ptr, p, nullptr);
    int const nV = static_cast<int>(h.nV), nUpd = static_cast<int>(h.nUpdated);
    if (nV > nUpd)
        Zero4(h, R.p.ptr + nUpd, nV - nUpd);
    Zero4(h, R.q.ptr, nV);
    VertexPass<AdjMode::kHessVec>(h, R.hold.ptr, dt, R.p.ptr, nullptr, R.q.ptr, nullptr, 0, 0,
                                  static_cast<int>(h.nUpdated));
    PBAT_HIP(hipGetLastError());
    DownloadVec(h, R.q.ptr, out);
}

void AdjointSolve(pbat_vbd_integrator& h, float const* x, float dt, float const* b, float* y,
                  pbat_vbd_adjoint_config const* cfgIn, pbat_vbd_adjoint_stats* stats)
{
    if (!x || !b || !y || !(dt > 0.f))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid adjoint solve arguments");
    NoContact(h);
    pbat_vbd_adjoint_config cfg{PBAT_VBD_ADJOINT_PCG, 200, 1e-6f, 5, 0};
    if (cfgIn)
        cfg = *cfgIn;
    CheckConfig(cfg);
    PBAT_HIP(hipSetDevice(h.device));
    EnsureBuffers(h);
    RolloutState& R = h.rollout;
    int const nV = static_cast<int>(h.nV), nUpd = static_cast<int>(h.nUpdated);
    UploadVec(h, R.hold.ptr, x, h.massOrig.data());
    UploadVec(h, R.rhs.ptr, b, nullptr);
    if (nV > nUpd)
        Zero4(h, R.rhs.ptr + nUpd, nV - nUpd);
    Zero4(h, R.y.ptr, nV);
    if (cfg.solver == PBAT_VBD_ADJOINT_PCG_TWO_LEVEL)
        BuildCoarse(h, R.hold.ptr, dt);
    SolveResult const sr = Solve(h, R.hold.ptr, dt, R.y.ptr, R.rhs.ptr, cfg, 0, nullptr);
    DownloadVec(h, R.y.ptr, y);
    if (stats)
        *stats = pbat_vbd_adjoint_stats{1, sr.iterations, sr.iterations, static_cast<float>(sr.relResidual)};
}

} // namespace
} // namespace pbat_vbd

using namespace pbat_vbd;

extern "C" int pbat_vbd_integrator_set_log_young_modulus(pbat_vbd_integrator* h, const float* logY, int64_t nT)
{
    return Guard([&] { SetLogYoung(Handle(h), logY, nT); });
}

extern "C" int pbat_vbd_integrator_get_log_young_modulus(pbat_vbd_integrator* h, float* logY, int64_t nT)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (!logY || nT != H.nT)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "expected |#elements| outputs");
        std::copy(H.logYCurrent.begin(), H.logYCurrent.end(), logY);
    });
}

extern "C" int pbat_vbd_rollout_forward(pbat_vbd_integrator* h, float dt, int32_t iterations, int32_t steps,
                                        int32_t checkpoint_every)
{
    return Guard([&] { RolloutForward(Handle(h), dt, iterations, steps, checkpoint_every); });
}

extern "C" int pbat_vbd_rollout_backward(pbat_vbd_integrator* h, const float* dL_dx, const float* dL_dv,
                                         float* dL_dlogY, float* dL_dv0, float* dL_dx0,
                                         const pbat_vbd_adjoint_config* config, pbat_vbd_adjoint_stats* stats)
{
    return Guard([&] { RolloutBackward(Handle(h), dL_dx, dL_dv, dL_dlogY, dL_dv0, dL_dx0, config, stats); });
}

extern "C" int pbat_vbd_rollout_generation(pbat_vbd_integrator* h, int64_t* generation, int32_t* valid)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (generation)
            *generation = H.rollout.generation;
        if (valid)
            *valid = H.rollout.valid ? 1 : 0;
    });
}

extern "C" int pbat_vbd_hessian_vector(pbat_vbd_integrator* h, const float* x, float dt, const float* p, float* out)
{
    return Guard([&] { HessianVectorProduct(Handle(h), x, dt, p, out); });
}

extern "C" int pbat_vbd_adjoint_solve(pbat_vbd_integrator* h, const float* x, float dt, const float* b, float* y,
                                      const pbat_vbd_adjoint_config* config, pbat_vbd_adjoint_stats* stats)
{
    return Guard([&] { AdjointSolve(Handle(h), x, dt, b, y, config, stats); });
}
