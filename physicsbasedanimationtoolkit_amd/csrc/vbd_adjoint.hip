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

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace pbat_vbd {
namespace {

enum ScalarSlot { kRz0 = 0, kRz1 = 1, kPq = 2, kRr = 3, kBb = 4 };

int Blocks(int64_t n, int block = 256) { return static_cast<int>((n + block - 1) / block); }

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
    a.ellStride = (h.useEll && h.ellStride > 0) ? h.ellStride : 0;
    a.invDt2 = 1.f / (dt * dt);
    return a;
}

template <int L, int K, bool kGS>
void LaunchAdjointLK(pbat_vbd_integrator& h, device::AdjointArgs a)
{
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
    hipLaunchKernelGGL((device::AdjointVertexKernel<L, K, kGS>), dim3(static_cast<int>(grid64)), dim3(block), 0,
                       h.stream, a);
}

template <int L, bool kGS>
void LaunchAdjointL(pbat_vbd_integrator& h, device::AdjointArgs const& a)
{
    switch (EffectivePerLane(h))
    {
        case 1: LaunchAdjointLK<L, 1, kGS>(h, a); break;
        case 2: LaunchAdjointLK<L, 2, kGS>(h, a); break;
        case 3: LaunchAdjointLK<L, 3, kGS>(h, a); break;
        default: LaunchAdjointLK<L, 4, kGS>(h, a); break;
    }
}

template <bool kGS>
void LaunchAdjoint(pbat_vbd_integrator& h, device::AdjointArgs const& a)
{
    if (a.vend <= a.vbeg)
        return;
    switch (EffectiveLanes(h))
    {
        case 1: LaunchAdjointL<1, kGS>(h, a); break;
        case 2: LaunchAdjointL<2, kGS>(h, a); break;
        case 4: LaunchAdjointL<4, kGS>(h, a); break;
        case 8: LaunchAdjointL<8, kGS>(h, a); break;
        case 16: LaunchAdjointL<16, kGS>(h, a); break;
        case 32: LaunchAdjointL<32, kGS>(h, a); break;
        default: LaunchAdjointL<64, kGS>(h, a); break;
    }
}

void EnsureBuffers(pbat_vbd_integrator& h)
{
    RolloutState& R = h.rollout;
    size_t const n = static_cast<size_t>(h.nV);
    for (auto* b : {&R.y, &R.r, &R.z, &R.p, &R.q, &R.rhs, &R.xbar, &R.vbar, &R.hold})
        if (b->n != n)
            b->allocate(n);
    if (R.scal.n != 8)
        R.scal.allocate(8);
    if (R.grad.n != static_cast<size_t>(h.nT))
        R.grad.allocate(static_cast<size_t>(h.nT));
}

void Zero4(pbat_vbd_integrator& h, float4* p, int n)
{
    if (n > 0)
        hipLaunchKernelGGL(device::Fill4, dim3(Blocks(n)), dim3(256), 0, h.stream, p, n);
}

// acc = a . b over the updated vertices (FP64, device scalar)
void Dot(pbat_vbd_integrator& h, float4 const* a, float4 const* b, double* acc)
{
    int const n = static_cast<int>(h.nUpdated);
    hipLaunchKernelGGL(device::ZeroDoubles, dim3(1), dim3(64), 0, h.stream, acc, 1);
    if (n > 0)
        hipLaunchKernelGGL(device::Dot3, dim3(std::min(1024, Blocks(n))), dim3(256), 0, h.stream, a, b, n, acc);
}

double ReadScalar(pbat_vbd_integrator& h, double const* p)
{
    double v = 0.;
    PBAT_HIP(hipMemcpyAsync(&v, p, sizeof(double), hipMemcpyDeviceToHost, h.stream));
    PBAT_HIP(hipStreamSynchronize(h.stream));
    return v;
}

// out = H p over the updated vertices
void HessVec(pbat_vbd_integrator& h, float4 const* xs, float dt, float4* p, float4* out)
{
    device::AdjointArgs a = MakeAdjointArgs(h, xs, dt);
    a.y = p;
    a.out = out;
    a.vbeg = 0;
    a.vend = static_cast<int>(h.nUpdated);
    LaunchAdjoint<false>(h, a);
}

// One Gauss-Seidel pass over the colours (forward or reverse order) on H y = b.
void GaussSeidelPass(pbat_vbd_integrator& h, float4 const* xs, float dt, float4* y, float4 const* b, bool reverse)
{
    device::AdjointArgs a = MakeAdjointArgs(h, xs, dt);
    a.y = y;
    a.b = b;
    int const nC = static_cast<int>(h.colorBegin.size()) - 1;
    for (int k = 0; k < nC; ++k)
    {
        int const c = reverse ? nC - 1 - k : k;
        a.vbeg = h.colorBegin[c];
        a.vend = h.colorBegin[c + 1];
        LaunchAdjoint<true>(h, a);
    }
}

struct SolveResult
{
    int iterations = 0;
    double relResidual = 0.;
};

// Solve H y = b (y holds the initial guess). Residuals are checked on the host
// every cfg.check_every iterations.
SolveResult Solve(pbat_vbd_integrator& h, float4 const* xs, float dt, float4* y, float4 const* b,
                  pbat_vbd_adjoint_config const& cfg)
{
    RolloutState& R = h.rollout;
    int const n = static_cast<int>(h.nUpdated);
    double* S = R.scal.ptr;
    Dot(h, b, b, S + kBb);
    double const bb = ReadScalar(h, S + kBb);
    SolveResult out;
    if (!(bb > 0.))
    {
        Zero4(h, y, static_cast<int>(h.nV));
        return out;
    }
    int const check = std::max(1, cfg.check_every);
    auto residual = [&]() {
        HessVec(h, xs, dt, y, R.q.ptr);
        hipLaunchKernelGGL(device::Sub4, dim3(Blocks(std::max(n, 1))), dim3(256), 0, h.stream, R.r.ptr, b, R.q.ptr,
                           n);
        Dot(h, R.r.ptr, R.r.ptr, S + kRr);
        return std::sqrt(ReadScalar(h, S + kRr) / bb);
    };
    if (cfg.solver == PBAT_VBD_ADJOINT_GAUSS_SEIDEL)
    {
        for (int it = 1; it <= cfg.max_iterations; ++it)
        {
            GaussSeidelPass(h, xs, dt, y, b, false);
            out.iterations = it;
            if (it % check == 0 || it == cfg.max_iterations)
            {
                out.relResidual = residual();
                if (out.relResidual <= cfg.tolerance)
                    break;
            }
        }
        PBAT_HIP(hipGetLastError());
        return out;
    }
    // Conjugate gradients preconditioned by a symmetric Gauss-Seidel pass
    // (forward then reverse colour order, from zero), all scalars on the device.
    out.relResidual = residual(); // r = b - H y
    if (out.relResidual <= cfg.tolerance)
        return out;
    auto precondition = [&]() {
        Zero4(h, R.z.ptr, static_cast<int>(h.nV));
        GaussSeidelPass(h, xs, dt, R.z.ptr, R.r.ptr, false);
        GaussSeidelPass(h, xs, dt, R.z.ptr, R.r.ptr, true);
    };
    precondition();
    PBAT_HIP(hipMemcpyAsync(R.p.ptr, R.z.ptr, sizeof(float4) * h.nV, hipMemcpyDeviceToDevice, h.stream));
    Dot(h, R.r.ptr, R.z.ptr, S + kRz0);
    int cur = kRz0;
    for (int it = 1; it <= cfg.max_iterations; ++it)
    {
        HessVec(h, xs, dt, R.p.ptr, R.q.ptr);
        Dot(h, R.p.ptr, R.q.ptr, S + kPq);
        hipLaunchKernelGGL(device::CgUpdate, dim3(Blocks(std::max(n, 1))), dim3(256), 0, h.stream, y, R.r.ptr,
                           R.p.ptr, R.q.ptr, n, S + cur, S + kPq);
        out.iterations = it;
        if (it % check == 0 || it == cfg.max_iterations)
        {
            Dot(h, R.r.ptr, R.r.ptr, S + kRr);
            out.relResidual = std::sqrt(ReadScalar(h, S + kRr) / bb);
            if (out.relResidual <= cfg.tolerance)
                break;
        }
        precondition();
        int const next = (cur == kRz0) ? kRz1 : kRz0;
        Dot(h, R.r.ptr, R.z.ptr, S + next);
        hipLaunchKernelGGL(device::CgDirection, dim3(Blocks(std::max(n, 1))), dim3(256), 0, h.stream, R.p.ptr,
                           R.z.ptr, n, S + next, S + cur);
        cur = next;
    }
    PBAT_HIP(hipGetLastError());
    return out;
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

void RolloutForward(pbat_vbd_integrator& h, float dt, int iterations, int steps, int every)
{
    if (!(dt > 0.f) || iterations < 0 || steps < 1 || every < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "rollout needs dt > 0, iterations >= 0, steps >= 1, every >= 1");
    PBAT_HIP(hipSetDevice(h.device));
    RolloutState& R = h.rollout;
    R.valid = false;
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
    for (int t = 0; t < steps; ++t)
    {
        if (t % every == 0)
        {
            int const c = t / every;
            SaveState(h, R.ckptX[c].ptr, R.ckptV[c].ptr, R.ckptVprev[c].ptr);
        }
        Step(h, dt, iterations, 1);
    }
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
    pbat_vbd_adjoint_config cfg{PBAT_VBD_ADJOINT_PCG, 200, 1e-6f, 5};
    if (cfgIn)
        cfg = *cfgIn;
    if (cfg.max_iterations < 1 || (cfg.solver != PBAT_VBD_ADJOINT_PCG && cfg.solver != PBAT_VBD_ADJOINT_GAUSS_SEIDEL))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid adjoint solver configuration");
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
    hipLaunchKernelGGL(device::ZeroDoubles, dim3(Blocks(std::max(nT, 1))), dim3(256), 0, h.stream, R.grad.ptr, nT);
    if (R.segX.size() != static_cast<size_t>(R.every))
        R.segX.resize(static_cast<size_t>(R.every));
    for (auto& s : R.segX)
        if (s.n != n)
            s.allocate(n);
    pbat_vbd_adjoint_stats st{0, 0, 0, 0.f};
    int const nSeg = static_cast<int>(R.ckptX.size());
    for (int sgm = nSeg - 1; sgm >= 0; --sgm)
    {
        RestoreState(h, R.ckptX[sgm].ptr, R.ckptV[sgm].ptr, R.ckptVprev[sgm].ptr);
        int const t0 = sgm * R.every, t1 = std::min(R.steps, t0 + R.every);
        for (int t = t0; t < t1; ++t)
        {
            Step(h, dt, R.iterations, 1);
            PBAT_HIP(hipMemcpyAsync(R.segX[t - t0].ptr, h.xm.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                    h.stream));
        }
        for (int t = t1 - 1; t >= t0; --t)
        {
            float4 const* xs = R.segX[t - t0].ptr;
            hipLaunchKernelGGL(device::AdjointRhs, dim3(Blocks(nV)), dim3(256), 0, h.stream, R.xbar.ptr, R.vbar.ptr,
                               R.rhs.ptr, nUpd, nV, 1.f / dt);
            SolveResult const sr = Solve(h, xs, dt, R.y.ptr, R.rhs.ptr, cfg);
            ++st.solves;
            st.total_iterations += sr.iterations;
            st.max_iterations = std::max(st.max_iterations, sr.iterations);
            st.max_relative_residual = std::max(st.max_relative_residual, static_cast<float>(sr.relResidual));
            hipLaunchKernelGGL(device::LogYoungGradient, dim3(Blocks(std::max(nT, 1))), dim3(256), 0, h.stream, xs,
                               R.y.ptr, h.tetv.ptr, h.tgp.ptr, nT, R.grad.ptr);
            hipLaunchKernelGGL(device::AdjointStepBack, dim3(Blocks(nV)), dim3(256), 0, h.stream, xs, R.y.ptr,
                               R.xbar.ptr, R.vbar.ptr, nUpd, nV, dt);
        }
    }
    PBAT_HIP(hipGetLastError());
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
    hipLaunchKernelGGL(device::SetLogYoung, dim3(Blocks(nT)), dim3(256), 0, h.stream, R.logY.ptr, h.matBase.ptr,
                       h.tgp.ptr, h.slotsRows.ptr, h.rows.g2.ptr, h.ellStride > 0 ? h.slotsEll.ptr : nullptr,
                       h.ellStride > 0 ? h.ell.g2.ptr : nullptr, static_cast<int>(nT));
    PBAT_HIP(hipGetLastError());
    PBAT_HIP(hipStreamSynchronize(h.stream));
    h.logYCurrent = v;
    R.valid = false;
}

void HessianVectorProduct(pbat_vbd_integrator& h, float const* x, float dt, float const* p, float* out)
{
    if (!x || !p || !out || !(dt > 0.f))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid Hessian-vector arguments");
    PBAT_HIP(hipSetDevice(h.device));
    EnsureBuffers(h);
    RolloutState& R = h.rollout;
    UploadVec(h, R.hold.ptr, x, h.massOrig.data());
    UploadVec(h, R.p.ptr, p, nullptr);
    int const nV = static_cast<int>(h.nV), nUpd = static_cast<int>(h.nUpdated);
    if (nV > nUpd)
        Zero4(h, R.p.ptr + nUpd, nV - nUpd);
    Zero4(h, R.q.ptr, nV);
    HessVec(h, R.hold.ptr, dt, R.p.ptr, R.q.ptr);
    PBAT_HIP(hipGetLastError());
    DownloadVec(h, R.q.ptr, out);
}

void AdjointSolve(pbat_vbd_integrator& h, float const* x, float dt, float const* b, float* y,
                  pbat_vbd_adjoint_config const* cfgIn, pbat_vbd_adjoint_stats* stats)
{
    if (!x || !b || !y || !(dt > 0.f))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid adjoint solve arguments");
    pbat_vbd_adjoint_config cfg{PBAT_VBD_ADJOINT_PCG, 200, 1e-6f, 5};
    if (cfgIn)
        cfg = *cfgIn;
    PBAT_HIP(hipSetDevice(h.device));
    EnsureBuffers(h);
    RolloutState& R = h.rollout;
    int const nV = static_cast<int>(h.nV), nUpd = static_cast<int>(h.nUpdated);
    UploadVec(h, R.hold.ptr, x, h.massOrig.data());
    UploadVec(h, R.rhs.ptr, b, nullptr);
    if (nV > nUpd)
        Zero4(h, R.rhs.ptr + nUpd, nV - nUpd);
    Zero4(h, R.y.ptr, nV);
    SolveResult const sr = Solve(h, R.hold.ptr, dt, R.y.ptr, R.rhs.ptr, cfg);
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

extern "C" int pbat_vbd_hessian_vector(pbat_vbd_integrator* h, const float* x, float dt, const float* p, float* out)
{
    return Guard([&] { HessianVectorProduct(Handle(h), x, dt, p, out); });
}

extern "C" int pbat_vbd_adjoint_solve(pbat_vbd_integrator* h, const float* x, float dt, const float* b, float* y,
                                      const pbat_vbd_adjoint_config* config, pbat_vbd_adjoint_stats* stats)
{
    return Guard([&] { AdjointSolve(Handle(h), x, dt, b, y, config, stats); });
}
