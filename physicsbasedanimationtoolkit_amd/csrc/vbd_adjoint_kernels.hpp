// Device kernels of the VBD adjoint (implicit-function-theorem gradient of a
// rollout). The linear systems are H y = b with
//   H = M / h^2 + sum_e w_e d2Psi_e/dx2   (the forward's scaling)
// evaluated at a step's minimiser x*, over the updated (non-Dirichlet)
// vertices, which are the internal range [0, nUpdated). Vectors are float4
// arrays in internal vertex order with zeros on Dirichlet vertices.
#pragma once

#include "coarse_space.hpp"
#include "vbd_kernels.hpp"

namespace pbat_vbd::device {

// Directional derivative of one incidence's elastic gradient: with
// dF = sum_k (p_k - p_i) g_k^T (the own frame: F = sum_k (x_k - x_i) g_k^T),
//   (H_e p)_i = [w mu dF + w lambda (cof F : dF) cof F + w lambda (J - alpha) dcof(F)[dF]] g
// where dcof's columns are df1 x f2 + f1 x df2, df2 x f0 + f2 x df0,
// df0 x f1 + f0 x df1. Adds the block-diagonal part w (mu |g|^2 I + lambda c c^T)
// into D when kDiag.
template <bool kDiag>
__device__ __forceinline__ void AccumulateHessVec(float4 const& q0, float4 const& q1, float2 const& q2, float wmu,
                                                  float4 const& own, float4 const& pa, float4 const& pb,
                                                  float4 const& pc, float4 const& yo, float4 const& ya,
                                                  float4 const& yb, float4 const& yc, float3& hv, Block& D)
{
    float const gax = q0.x, gay = q0.y, gaz = q0.z;
    float const gbx = q0.w, gby = q1.x, gbz = q1.y;
    float const gcx = q1.z, gcy = q1.w, gcz = q2.x;
    float const wlam = q2.y, alpha = 1.f + wmu / wlam;
    // F
    float const dax = pa.x - own.x, day = pa.y - own.y, daz = pa.z - own.z;
    float const dbx = pb.x - own.x, dby = pb.y - own.y, dbz = pb.z - own.z;
    float const dcx = pc.x - own.x, dcy = pc.y - own.y, dcz = pc.z - own.z;
    float const f0x = dax * gax + dbx * gbx + dcx * gcx, f0y = day * gax + dby * gbx + dcy * gcx,
                f0z = daz * gax + dbz * gbx + dcz * gcx;
    float const f1x = dax * gay + dbx * gby + dcx * gcy, f1y = day * gay + dby * gby + dcy * gcy,
                f1z = daz * gay + dbz * gby + dcz * gcy;
    float const f2x = dax * gaz + dbx * gbz + dcx * gcz, f2y = day * gaz + dby * gbz + dcy * gcz,
                f2z = daz * gaz + dbz * gbz + dcz * gcz;
    // dF
    float const eax = ya.x - yo.x, eay = ya.y - yo.y, eaz = ya.z - yo.z;
    float const ebx = yb.x - yo.x, eby = yb.y - yo.y, ebz = yb.z - yo.z;
    float const ecx = yc.x - yo.x, ecy = yc.y - yo.y, ecz = yc.z - yo.z;
    float const h0x = eax * gax + ebx * gbx + ecx * gcx, h0y = eay * gax + eby * gbx + ecy * gcx,
                h0z = eaz * gax + ebz * gbx + ecz * gcx;
    float const h1x = eax * gay + ebx * gby + ecx * gcy, h1y = eay * gay + eby * gby + ecy * gcy,
                h1z = eaz * gay + ebz * gby + ecz * gcy;
    float const h2x = eax * gaz + ebx * gbz + ecx * gcz, h2y = eay * gaz + eby * gbz + ecy * gcz,
                h2z = eaz * gaz + ebz * gbz + ecz * gcz;
    // cofactor columns A = f1 x f2, B = f2 x f0, C = f0 x f1
    float const Ax = f1y * f2z - f1z * f2y, Ay = f1z * f2x - f1x * f2z, Az = f1x * f2y - f1y * f2x;
    float const Bx = f2y * f0z - f2z * f0y, By = f2z * f0x - f2x * f0z, Bz = f2x * f0y - f2y * f0x;
    float const Cx = f0y * f1z - f0z * f1y, Cy = f0z * f1x - f0x * f1z, Cz = f0x * f1y - f0y * f1x;
    float const J = f0x * Ax + f0y * Ay + f0z * Az;
    float const dJ = Ax * h0x + Ay * h0y + Az * h0z + Bx * h1x + By * h1y + Bz * h1z + Cx * h2x + Cy * h2y +
                     Cz * h2z;
    // d(cof) columns
    float const dAx = (h1y * f2z - h1z * f2y) + (f1y * h2z - f1z * h2y);
    float const dAy = (h1z * f2x - h1x * f2z) + (f1z * h2x - f1x * h2z);
    float const dAz = (h1x * f2y - h1y * f2x) + (f1x * h2y - f1y * h2x);
    float const dBx = (h2y * f0z - h2z * f0y) + (f2y * h0z - f2z * h0y);
    float const dBy = (h2z * f0x - h2x * f0z) + (f2z * h0x - f2x * h0z);
    float const dBz = (h2x * f0y - h2y * f0x) + (f2x * h0y - f2y * h0x);
    float const dCx = (h0y * f1z - h0z * f1y) + (f0y * h1z - f0z * h1y);
    float const dCy = (h0z * f1x - h0x * f1z) + (f0z * h1x - f0x * h1z);
    float const dCz = (h0x * f1y - h0y * f1x) + (f0x * h1y - f0y * h1x);
    float const gx = -(gax + gbx + gcx), gy = -(gay + gby + gcy), gz = -(gaz + gbz + gcz);
    // cof(F) g, dcof g, dF g
    float const kx = gx * Ax + gy * Bx + gz * Cx, ky = gx * Ay + gy * By + gz * Cy, kz = gx * Az + gy * Bz + gz * Cz;
    float const lx = gx * dAx + gy * dBx + gz * dCx, ly = gx * dAy + gy * dBy + gz * dCy,
                lz = gx * dAz + gy * dBz + gz * dCz;
    float const mx = h0x * gx + h1x * gy + h2x * gz, my = h0y * gx + h1y * gy + h2y * gz,
                mz = h0z * gx + h1z * gy + h2z * gz;
    float const s1 = wlam * dJ, s2 = wlam * (J - alpha);
    hv.x += wmu * mx + s1 * kx + s2 * lx;
    hv.y += wmu * my + s1 * ky + s2 * ly;
    hv.z += wmu * mz + s1 * kz + s2 * lz;
    if constexpr (kDiag)
    {
        float const dg = wmu * (gx * gx + gy * gy + gz * gz);
        float const lkx = wlam * kx, lky = wlam * ky, lkz = wlam * kz;
        D.h00 += dg + lkx * kx;
        D.h11 += dg + lky * ky;
        D.h22 += dg + lkz * kz;
        D.h01 += lkx * ky;
        D.h02 += lkx * kz;
        D.h12 += lky * kz;
    }
}

// Per-mesh solver scalars. Plain values: kPlainCount arrays of nM doubles.
enum PlainArray : int
{
    kRz = 0,     // r . z of the current iteration
    kBeta = 1,   // CG direction coefficient
    kRrLast = 2, // r . r at the last completed iteration
    kBb = 3,     // b . b
    kPlainCount = 4
};
// Accumulated dot products: each (array, mesh) is spread over kSpread atomic
// targets kSlotStride doubles (128 B) apart, so that thousands of workgroups
// adding into one mesh's sum do not serialise on one address (same-address
// FP64 atomics from many workgroups cost ~20 ns each, serialised).
enum AccArray : int
{
    kAccRzNew = 0, // r . z being accumulated
    kAccPq = 1,    // p . H p
    kAccRr = 2,    // r . r
    kAccBb = 3,    // b . b
    kAccCount = 4
};
constexpr int kSpread = 16;
constexpr int kSlotStride = 16;

__host__ __device__ constexpr size_t AccDoubles(int nM)
{
    return static_cast<size_t>(kAccCount) * nM * kSpread * kSlotStride;
}

__device__ __forceinline__ double* AccSlot(double* A, int arr, int nM, int m, int slot)
{
    return A + ((static_cast<size_t>(arr) * nM + m) * kSpread + slot) * kSlotStride;
}

// Sum of a mesh's targets (read by a later launch than the one that added).
__device__ __forceinline__ double AccSum(double const* A, int arr, int nM, int m)
{
    double s = 0.;
#pragma unroll
    for (int k = 0; k < kSpread; ++k)
        s += A[((static_cast<size_t>(arr) * nM + m) * kSpread + k) * kSlotStride];
    return s;
}

// Per-mesh flags: done[nM], iterations[nM], then one count of live meshes.
enum FlagArray : int
{
    kDone = 0,
    kIters = 1
};

// Device state of the per-mesh conjugate-gradient solver.
struct CgState
{
    double* P; // plain values (kPlainCount x nM)
    double* A; // accumulators (AccDoubles(nM))
    int* F;    // flags (2 nM + 1)
    int nM;
    double tol2; // squared relative tolerance
};

// End of one CG iteration of mesh m (one thread) from its sums r . r, r . z
// and, at the initialisation (first), b . b: the stopping test |r| <= tol |b|,
// beta = rz / rz_old, rz rotated, the iteration counted. Returns whether the
// mesh is still live.
__device__ __forceinline__ bool FinishMesh(CgState const& c, int m, double rr, double rz, double bb, int first)
{
    int const nM = c.nM;
    if (first)
        c.P[kBb * nM + m] = bb;
    int done = c.F[kDone * nM + m];
    double const b2 = c.P[kBb * nM + m];
    if (!done)
    {
        if (!first)
            c.F[kIters * nM + m] += 1;
        c.P[kRrLast * nM + m] = rr;
        if (!(b2 > 0.) || rr <= c.tol2 * b2)
            done = 1;
    }
    double const rzOld = c.P[kRz * nM + m];
    c.P[kBeta * nM + m] = (!done && rzOld != 0.) ? rz / rzOld : 0.;
    c.P[kRz * nM + m] = rz;
    c.F[kDone * nM + m] = done;
    return !done;
}

// Per mesh, after an iteration's accumulations (one workgroup, blockDim a
// multiple of 64): each mesh's kSpread targets are summed by one half-wave in
// a fixed order, every accumulator of the mesh cleared, then FinishMesh and
// the live meshes counted. first: the initialisation (takes b.b, counts no
// iteration).
__device__ __forceinline__ void FinishIteration(CgState const& c, int first)
{
    static_assert(kSpread <= 32, "one half-wave per mesh");
    __shared__ int live;
    if (threadIdx.x == 0)
        live = 0;
    __syncthreads();
    int const nM = c.nM;
    int const half = static_cast<int>(threadIdx.x >> 5), hl = static_cast<int>(threadIdx.x & 31);
    int const nHalves = static_cast<int>(blockDim.x >> 5);
    int count = 0;
    for (int m = half; m < nM; m += nHalves)
    {
        double rr = 0., rz = 0., bb = 0.;
        if (hl < kSpread)
        {
            rr = *AccSlot(c.A, kAccRr, nM, m, hl);
            rz = *AccSlot(c.A, kAccRzNew, nM, m, hl);
            bb = *AccSlot(c.A, kAccBb, nM, m, hl);
            for (int arr = 0; arr < kAccCount; ++arr)
                *AccSlot(c.A, arr, nM, m, hl) = 0.;
        }
        for (int off = 16; off > 0; off >>= 1)
        {
            rr += __shfl_xor(rr, off);
            rz += __shfl_xor(rz, off);
            bb += __shfl_xor(bb, off);
        }
        if (hl == 0)
            count += FinishMesh(c, m, rr, rz, bb, first) ? 1 : 0;
    }
    if (count)
        atomicAdd(&live, count);
    __syncthreads();
    if (threadIdx.x == 0)
        c.F[2 * nM] = live;
}

// Sum val into accumulated array arr of mesh m over the calling workgroup;
// every thread of the workgroup must call it (threads without a contribution
// pass m < 0); blockDim <= 256. Waves whose contributors share one mesh (the
// vertices of a colour are grouped by mesh) reduce in registers, the
// workgroup's waves then combine in LDS, and one FP64 atomic per mesh per
// workgroup lands on one of the mesh's kSpread targets.
__device__ __forceinline__ void BlockAccumulate(double val, int m, double* __restrict__ A, int arr, int nM)
{
    __shared__ double sv[4];
    __shared__ int sm[4];
    int const lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
    int const slot = static_cast<int>(blockIdx.x) & (kSpread - 1);
    unsigned long long const live = __ballot(m >= 0);
    int mref = -1;
    double v = 0.;
    if (live != 0ull)
    {
        int const first = __ffsll(static_cast<long long>(live)) - 1;
        int const mf = __shfl(m, first);
        if (__all(m < 0 || m == mf))
        {
            v = (m >= 0) ? val : 0.;
            for (int off = 32; off > 0; off >>= 1)
                v += __shfl_xor(v, off);
            mref = mf;
        }
        else if (m >= 0)
        {
            atomicAdd(AccSlot(A, arr, nM, m, slot), val);
        }
    }
    if (lane == 0)
    {
        sv[wave] = v;
        sm[wave] = mref;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int const nw = static_cast<int>((blockDim.x + 63) / 64);
        for (int w = 0; w < nw; ++w)
        {
            int const mw = sm[w];
            if (mw < 0)
                continue;
            double sum = sv[w];
            for (int u = w + 1; u < nw; ++u)
                if (sm[u] == mw)
                {
                    sum += sv[u];
                    sm[u] = -1;
                }
            atomicAdd(AccSlot(A, arr, nM, mw, slot), sum);
        }
    }
    __syncthreads();
}

enum class AdjMode : int
{
    kGaussSeidel = 0,    // one colour of block Gauss-Seidel on H y = b
    kGaussSeidelDot = 1, // same, then r.z += b . y (final pass of symmetric GS)
    kHessVec = 2,        // out = H y
    kHessVecDot = 3,     // out = H y, acc[mesh] += y . out
    kDiagInverse = 4     // dinv = inverse of the 3x3 diagonal block of H
};

struct AdjointArgs
{
    int const* __restrict__ vtp;
    int4 const* __restrict__ ids;
    float4 const* __restrict__ g0;
    float4 const* __restrict__ g1;
    float2 const* __restrict__ g2;
    float4 const* __restrict__ xs;  // minimiser positions (with lumped mass in .w)
    float4* __restrict__ y;         // unknowns (Gauss-Seidel: updated in place) / vector multiplied
    float4 const* __restrict__ b;   // right-hand side (Gauss-Seidel)
    float4* __restrict__ out;       // H y (Hessian-vector product) / inverse blocks (2 per vertex)
    int const* __restrict__ meshOf; // mesh of each vertex
    double* __restrict__ acc;       // accumulators (Dot modes): array accArray of nM meshes
    int accArray;
    int nM;
    int vbeg;
    int vend;
    int xcdChunk;
    int ellStride;
    float invDt2; // 1 / h^2
};

// The 3x3 symmetric block solve / inverse in FP64 (adjugate / determinant).
struct Inverse3
{
    double c00, c01, c02, c11, c12, c22, inv;
    __device__ __forceinline__ Inverse3(double d00, double d01, double d02, double d11, double d12, double d22)
    {
        c00 = d11 * d22 - d12 * d12, c01 = d02 * d12 - d01 * d22, c02 = d01 * d12 - d02 * d11;
        c11 = d00 * d22 - d02 * d02, c12 = d01 * d02 - d00 * d12, c22 = d00 * d11 - d01 * d01;
        double const det = d00 * c00 + d01 * c01 + d02 * c02;
        inv = (fabs(det) > 0.) ? 1. / det : 0.;
    }
};

// Per vertex i of [vbeg, vend), with D_i = m_i / h^2 I + sum_e w_e (mu |g|^2 I + lambda c c^T)
// and (H y)_i = m_i / h^2 y_i + sum_e (H_e y)_i:
//   kGaussSeidel:  y_i += D_i^{-1} (b_i - (H y)_i)
//   kHessVec:      out_i = (H y)_i
//   kDiagInverse:  out[2i], out[2i+1] = D_i^{-1} as (00, 01, 02, 11), (12, 22, ., .)
template <int L, int K, AdjMode kMode>
__global__ __launch_bounds__(256) void AdjointVertexKernel(AdjointArgs a)
{
    constexpr bool kGS = kMode == AdjMode::kGaussSeidel || kMode == AdjMode::kGaussSeidelDot;
    constexpr bool kDiag = kGS || kMode == AdjMode::kDiagInverse;
    constexpr bool kDot = kMode == AdjMode::kGaussSeidelDot || kMode == AdjMode::kHessVecDot;
    constexpr bool kProduct = kMode == AdjMode::kHessVec || kMode == AdjMode::kHessVecDot;
    int block = static_cast<int>(blockIdx.x);
    if (a.xcdChunk > 0)
        block = (block & 7) * a.xcdChunk + (block >> 3);
    int const tid = block * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    int const lane = tid & (L - 1);
    int const v = a.vbeg + tid / L;
    bool const active = v < a.vend;
    int kb = 0, ke = 0;
    float4 own = make_float4(0.f, 0.f, 0.f, 0.f), yo = own, bo = own;
    auto vec = [&](int k) { return a.y[k]; };
    if (active)
    {
        if (a.ellStride > 0)
        {
            kb = v * a.ellStride;
            ke = kb + a.ellStride;
        }
        else
        {
            kb = a.vtp[v];
            ke = a.vtp[v + 1];
        }
        own = a.xs[v];
        yo = vec(v);
        if (kGS && lane == 0)
            bo = a.b[v];
    }
    float3 hv = make_float3(0.f, 0.f, 0.f);
    Block D{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int4 I[K];
    float4 q0[K], q1[K];
    float2 q2[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
    {
        int const k = kb + lane + j * L;
        if (k < ke)
        {
            I[j] = a.ids[k];
            q0[j] = a.g0[k];
            q1[j] = a.g1[k];
            q2[j] = a.g2[k];
        }
        else
            I[j] = make_int4(-1, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
    {
        if (I[j].x >= 0)
            AccumulateHessVec<kDiag>(q0[j], q1[j], q2[j], __int_as_float(I[j].w), own, a.xs[I[j].x], a.xs[I[j].y],
                                     a.xs[I[j].z], yo, vec(I[j].x), vec(I[j].y), vec(I[j].z), hv, D);
    }
    for (int k = kb + lane + K * L; k < ke; k += L)
    {
        int4 const J = a.ids[k];
        if (J.x < 0)
            continue;
        AccumulateHessVec<kDiag>(a.g0[k], a.g1[k], a.g2[k], __int_as_float(J.w), own, a.xs[J.x], a.xs[J.y],
                                 a.xs[J.z], yo, vec(J.x), vec(J.y), vec(J.z), hv, D);
    }
    hv.x = GroupSum<L>(hv.x);
    hv.y = GroupSum<L>(hv.y);
    hv.z = GroupSum<L>(hv.z);
    if constexpr (kDiag)
        GroupSum<L>(D);
    bool const leader = active && lane == 0;
    float const K0 = own.w * a.invDt2;
    float const Hx = hv.x + K0 * yo.x, Hy = hv.y + K0 * yo.y, Hz = hv.z + K0 * yo.z;
    double dot = 0.;
    if (leader)
    {
        if constexpr (kProduct)
        {
            a.out[v] = make_float4(Hx, Hy, Hz, 0.f);
            dot = double(yo.x) * Hx + double(yo.y) * Hy + double(yo.z) * Hz;
        }
        else if constexpr (kMode == AdjMode::kDiagInverse)
        {
            Inverse3 const C(double(D.h00) + K0, D.h01, D.h02, double(D.h11) + K0, D.h12, double(D.h22) + K0);
            a.out[2 * v] = make_float4(static_cast<float>(C.c00 * C.inv), static_cast<float>(C.c01 * C.inv),
                                       static_cast<float>(C.c02 * C.inv), static_cast<float>(C.c11 * C.inv));
            a.out[2 * v + 1] =
                make_float4(static_cast<float>(C.c12 * C.inv), static_cast<float>(C.c22 * C.inv), 0.f, 0.f);
        }
        else
        {
            float const rx = bo.x - Hx, ry = bo.y - Hy, rz = bo.z - Hz;
            Inverse3 const C(double(D.h00) + K0, D.h01, D.h02, double(D.h11) + K0, D.h12, double(D.h22) + K0);
            float4 const yn =
                make_float4(yo.x + static_cast<float>((C.c00 * rx + C.c01 * ry + C.c02 * rz) * C.inv),
                            yo.y + static_cast<float>((C.c01 * rx + C.c11 * ry + C.c12 * rz) * C.inv),
                            yo.z + static_cast<float>((C.c02 * rx + C.c12 * ry + C.c22 * rz) * C.inv), 0.f);
            a.y[v] = yn;
            dot = double(bo.x) * yn.x + double(bo.y) * yn.y + double(bo.z) * yn.z;
        }
    }
    if constexpr (kDot)
        BlockAccumulate(dot, leader ? a.meshOf[v] : -1, a.acc, a.accArray, a.nM);
}

// d L / d log Y_e -= (w_e P_e) : dF_e(y) at the minimiser, for every tet
// (mu, lambda proportional to Y, so d(grad U)/d log Y_e = w_e dPsi_e/dx).
// Evaluated in FP64 from the FP32 positions: near rest the stress
// P = mu F + lambda (J - alpha) cof F is a small difference of O(Y) terms
// (P(I) = 0), which FP32 would resolve only to ~1e-7 / strain.
__global__ __launch_bounds__(256) void LogYoungGradient(float4 const* __restrict__ xs, float4 const* __restrict__ y,
                                                        int4 const* __restrict__ tetv,
                                                        float4 const* __restrict__ tgp, int nT,
                                                        double* __restrict__ grad)
{
    int const e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= nT)
        return;
    int4 const t = tetv[e];
    float4 const q0 = tgp[3 * e], q1 = tgp[3 * e + 1], q2 = tgp[3 * e + 2];
    float4 const x0 = xs[t.x], x1 = xs[t.y], x2 = xs[t.z], x3 = xs[t.w];
    float4 const y0 = y[t.x], y1 = y[t.y], y2 = y[t.z], y3 = y[t.w];
    double const r[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    double const d[3][3] = {{double(x1.x) - x0.x, double(x1.y) - x0.y, double(x1.z) - x0.z},
                            {double(x2.x) - x0.x, double(x2.y) - x0.y, double(x2.z) - x0.z},
                            {double(x3.x) - x0.x, double(x3.y) - x0.y, double(x3.z) - x0.z}};
    double const u[3][3] = {{double(y1.x) - y0.x, double(y1.y) - y0.y, double(y1.z) - y0.z},
                            {double(y2.x) - y0.x, double(y2.y) - y0.y, double(y2.z) - y0.z},
                            {double(y3.x) - y0.x, double(y3.y) - y0.y, double(y3.z) - y0.z}};
    double F[3][3], dF[3][3]; // [row][col]
    for (int rr = 0; rr < 3; ++rr)
        for (int c = 0; c < 3; ++c)
        {
            F[rr][c] = d[0][rr] * r[0][c] + d[1][rr] * r[1][c] + d[2][rr] * r[2][c];
            dF[rr][c] = u[0][rr] * r[0][c] + u[1][rr] * r[1][c] + u[2][rr] * r[2][c];
        }
    // cofactor matrix (column c = f_{c+1} x f_{c+2})
    double cof[3][3];
    for (int c = 0; c < 3; ++c)
    {
        int const c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        cof[0][c] = F[1][c1] * F[2][c2] - F[2][c1] * F[1][c2];
        cof[1][c] = F[2][c1] * F[0][c2] - F[0][c1] * F[2][c2];
        cof[2][c] = F[0][c1] * F[1][c2] - F[1][c1] * F[0][c2];
    }
    double const J = F[0][0] * cof[0][0] + F[1][0] * cof[1][0] + F[2][0] * cof[2][0];
    double const wmu = q2.y, s = double(q2.z) * (J - double(q2.w));
    double acc = 0.;
    for (int rr = 0; rr < 3; ++rr)
        for (int c = 0; c < 3; ++c)
            acc += (wmu * F[rr][c] + s * cof[rr][c]) * dF[rr][c];
    grad[e] -= acc;
}

// rhs = xbar + vbar / h on updated vertices, 0 elsewhere
__global__ __launch_bounds__(256) void AdjointRhs(float4 const* __restrict__ xbar, float4 const* __restrict__ vbar,
                                                  float4* __restrict__ rhs, int nUpd, int nV, float invDt)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nV)
        return;
    float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < nUpd)
    {
        float4 const x = xbar[i], v = vbar[i];
        r = make_float4(x.x + v.x * invDt, x.y + v.y * invDt, x.z + v.z * invDt, 0.f);
    }
    rhs[i] = r;
}

// xt~bar = m y / h^2; xbar <- xt~bar - vbar / h; vbar <- h xt~bar (updated vertices; 0 elsewhere)
__global__ __launch_bounds__(256) void AdjointStepBack(float4 const* __restrict__ xs, float4 const* __restrict__ y,
                                                       float4* __restrict__ xbar, float4* __restrict__ vbar, int nUpd,
                                                       int nV, float dt)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nV)
        return;
    float4 xb = make_float4(0.f, 0.f, 0.f, 0.f), vb = xb;
    if (i < nUpd)
    {
        float const s = xs[i].w / (dt * dt);
        float4 const yi = y[i], vo = vbar[i];
        float const tx = s * yi.x, ty = s * yi.y, tz = s * yi.z;
        xb = make_float4(tx - vo.x / dt, ty - vo.y / dt, tz - vo.z / dt, 0.f);
        vb = make_float4(dt * tx, dt * ty, dt * tz, 0.f);
    }
    xbar[i] = xb;
    vbar[i] = vb;
}

// ---- conjugate-gradient kernels with per-mesh scalars kept on the device ----
// Every mesh of a batch runs its own CG (its own alpha, beta and stopping
// test); a mesh that has converged is frozen. Vertices [0, n) are the updated
// ones; P points at kPlainCount x nM doubles, A at the accumulators
// (AccDoubles(nM)), F at 2 nM + 1 ints.

__device__ __forceinline__ double Dot3(float4 const& u, float4 const& w)
{
    return double(u.x) * w.x + double(u.y) * w.y + double(u.z) * w.z;
}

__device__ __forceinline__ float4 ApplyInverse(float4 const* __restrict__ dinv, int i, float4 const& r)
{
    float4 const A = dinv[2 * i], B = dinv[2 * i + 1]; // (00, 01, 02, 11), (12, 22)
    return make_float4(A.x * r.x + A.y * r.y + A.z * r.z, A.y * r.x + A.w * r.y + B.x * r.z,
                       A.z * r.x + B.x * r.y + B.y * r.z, 0.f);
}

// r = b - q (q = H y); z = D^-1 r when dinv; accumulates b.b, r.r (and r.z)
__global__ __launch_bounds__(256) void CgInitResidual(float4 const* __restrict__ b, float4 const* __restrict__ q,
                                                      float4* __restrict__ r, float4* __restrict__ z,
                                                      float4 const* __restrict__ dinv, int const* __restrict__ meshOf,
                                                      int n, int nM, double* __restrict__ A)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    double bb = 0., rr = 0., rz = 0.;
    int m = -1;
    if (i < n)
    {
        m = meshOf[i];
        float4 const bi = b[i], qi = q[i];
        float4 const ri = make_float4(bi.x - qi.x, bi.y - qi.y, bi.z - qi.z, 0.f);
        r[i] = ri;
        bb = Dot3(bi, bi);
        rr = Dot3(ri, ri);
        if (dinv)
        {
            float4 const zi = ApplyInverse(dinv, i, ri);
            z[i] = zi;
            rz = Dot3(ri, zi);
        }
    }
    BlockAccumulate(bb, m, A, kAccBb, nM);
    BlockAccumulate(rr, m, A, kAccRr, nM);
    if (dinv)
        BlockAccumulate(rz, m, A, kAccRzNew, nM);
}

// y += alpha p, r -= alpha q with alpha = rz / pq of the vertex's mesh (0 when
// it is done); r.r accumulated; with dinv: z = D^-1 r and r.z accumulated;
// without: z = 0 (a Gauss-Seidel preconditioner pass follows).
__global__ __launch_bounds__(256) void CgUpdate(float4* __restrict__ y, float4* __restrict__ r,
                                                float4* __restrict__ z, float4 const* __restrict__ p,
                                                float4 const* __restrict__ q, float4 const* __restrict__ dinv,
                                                int const* __restrict__ meshOf, int n, CgState c)
{
    int const nM = c.nM;
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    double rr = 0., rz = 0.;
    int m = -1;
    if (i < n)
    {
        m = meshOf[i];
        double const pq = AccSum(c.A, kAccPq, nM, m);
        bool const done = c.F[kDone * nM + m] != 0;
        float const alpha = (!done && pq != 0.) ? static_cast<float>(c.P[kRz * nM + m] / pq) : 0.f;
        float4 const pi = p[i], qi = q[i];
        float4 yi = y[i], ri = r[i];
        yi.x += alpha * pi.x, yi.y += alpha * pi.y, yi.z += alpha * pi.z;
        ri.x -= alpha * qi.x, ri.y -= alpha * qi.y, ri.z -= alpha * qi.z;
        y[i] = yi;
        r[i] = ri;
        rr = Dot3(ri, ri);
        if (dinv)
        {
            float4 const zi = ApplyInverse(dinv, i, ri);
            z[i] = zi;
            rz = Dot3(ri, zi);
        }
        else
        {
            z[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    BlockAccumulate(rr, m, c.A, kAccRr, nM);
    if (dinv)
        BlockAccumulate(rz, m, c.A, kAccRzNew, nM);
}

// End of a CG iteration (or of the initialisation when first): one workgroup.
__global__ __launch_bounds__(256) void CgScalars(CgState c, int first)
{
    FinishIteration(c, first);
}

// p = z + beta p (beta of the vertex's mesh; finished meshes keep p)
__global__ __launch_bounds__(256) void CgDirection(float4* __restrict__ p, float4 const* __restrict__ z,
                                                   int const* __restrict__ meshOf, int n, CgState c)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    int const m = meshOf[i];
    if (c.F[kDone * c.nM + m])
        return;
    float const beta = static_cast<float>(c.P[kBeta * c.nM + m]);
    float4 const zi = z[i];
    float4 const pi = p[i];
    p[i] = make_float4(zi.x + beta * pi.x, zi.y + beta * pi.y, zi.z + beta * pi.z, 0.f);
}

// ---- two-level preconditioner: rigid motions of vertex aggregates ----------
// Coarse space Z: one 6-column block per aggregate, the unit translations of
// its vertices and the infinitesimal rotations omega x d_i about its centroid
// (d_i = (x_i - c) / s, s the aggregate's RMS radius; d = 0 for an aggregate
// whose vertices are (nearly) collinear, whose rotation columns are then zero
// and carry an identity diagonal in the coarse operator). Z_i = [I | -[d_i]_x].
// The coarse operator Z^T H Z is formed per mesh as a dense block from the
// assembled step Hessian (CoarseFromBsr), inverted in place by Gauss-Jordan
// steps (SPD: no pivoting), and the preconditioner adds Z (Z^T H Z)^-1 Z^T r
// to the block-Jacobi one.

struct CoarseArgs
{
    int const* __restrict__ aggOf;          // aggregate of each internal vertex (-1: not updated)
    int const* __restrict__ aggMesh;        // mesh of each aggregate
    int const* __restrict__ meshAggBegin;   // first aggregate of each mesh (nM + 1)
    long long const* __restrict__ meshOffset; // first double of each mesh's dense block
    int nAgg;
};

// Block Gauss-Jordan step on every mesh block (blockIdx.y = mesh): with the
// pivot block P = A[K, K] of the rows and columns K = [k, k + b) (b = min(kPivot,
// n - k)), Bk = step(Ak) is
//   B[K, K] = P^-1,   B[K, j] = P^-1 A[K, j],   B[i, K] = -A[i, K] P^-1,
//   B[i, j] = A[i, j] - A[i, K] P^-1 A[K, j]     (i, j outside K),
// the same as b scalar steps; after ceil(n / kPivot) steps an n x n SPD block
// holds its inverse. A step is two launches: CoarsePivotRows forms the b rows
// W = B[K, :] (P^-1 in the K columns), CoarsePivotUpdate every other row from
// Ak and W: B[i, j] = (j in K ? 0 : A[i, j]) - A[i, K] W[:, j].
constexpr int kPivot = 32;

// P inverted in LDS (SPD, no pivoting, Gauss-Jordan one entry per thread step)
// by every workgroup, then W[:, j] for its 64 columns, 4 threads per column
// each forming kPivot / 4 of its b entries. W holds kPivot rows of wStride
// doubles per mesh.
__global__ __launch_bounds__(256) void CoarsePivotRows(double const* __restrict__ Ak, double* __restrict__ W,
                                                       int k, int wStride, CoarseArgs c)
{
    int const m = static_cast<int>(blockIdx.y);
    int const n = kCoarseDofs * (c.meshAggBegin[m + 1] - c.meshAggBegin[m]);
    int const j0 = static_cast<int>(blockIdx.x) * 64;
    if (k >= n || j0 >= n)
        return;
    int const b = min(kPivot, n - k);
    double const* A = Ak + c.meshOffset[m];
    __shared__ double P[kPivot][kPivot + 1];
    int const t = static_cast<int>(threadIdx.x);
    for (int e = t; e < kPivot * kPivot; e += 256)
    {
        int const r = e / kPivot, q = e % kPivot;
        P[r][q] = (r < b && q < b) ? A[static_cast<size_t>(k + r) * n + k + q] : 0.;
    }
    __syncthreads();
    for (int q = 0; q < b; ++q)
    {
        double v[kPivot * kPivot / 256];
#pragma unroll
        for (int u = 0; u < kPivot * kPivot / 256; ++u)
        {
            int const e = t + 256 * u, r = e / kPivot, cc = e % kPivot;
            double const inv = 1. / P[q][q];
            if (r >= b || cc >= b)
                v[u] = 0.;
            else if (r == q && cc == q)
                v[u] = inv;
            else if (r == q)
                v[u] = P[q][cc] * inv;
            else if (cc == q)
                v[u] = -P[r][q] * inv;
            else
                v[u] = P[r][cc] - P[r][q] * P[q][cc] * inv;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPivot * kPivot / 256; ++u)
        {
            int const e = t + 256 * u;
            P[e / kPivot][e % kPivot] = v[u];
        }
        __syncthreads();
    }
    int const j = j0 + (t & 63), l0 = (t >> 6) * (kPivot / 4);
    if (j >= n)
        return;
    double* Wm = W + static_cast<size_t>(m) * kPivot * wStride;
    if (j >= k && j < k + b)
    {
        for (int l = l0; l < l0 + kPivot / 4 && l < b; ++l)
            Wm[static_cast<size_t>(l) * wStride + j] = P[l][j - k];
        return;
    }
    double w[kPivot / 4];
#pragma unroll
    for (int l = 0; l < kPivot / 4; ++l)
        w[l] = 0.;
    for (int q = 0; q < b; ++q)
    {
        double const aq = A[static_cast<size_t>(k + q) * n + j];
#pragma unroll
        for (int l = 0; l < kPivot / 4; ++l)
            w[l] += P[l0 + l][q] * aq;
    }
#pragma unroll
    for (int l = 0; l < kPivot / 4; ++l)
        if (l0 + l < b)
            Wm[static_cast<size_t>(l0 + l) * wStride + j] = w[l];
}

// The other rows of step k, one 64 x 64 tile of (i, j) per workgroup
// (blockIdx.x over the tiles of the largest block, blockIdx.y = mesh): A[i, K]
// and W[:, j] staged in LDS, a 4 x 4 group of entries per thread. Rows in K
// copy W.
__global__ __launch_bounds__(256) void CoarsePivotUpdate(double const* __restrict__ Ak, double* __restrict__ Bk,
                                                         double const* __restrict__ W, int k, int wStride,
                                                         int tilesPerRow, CoarseArgs c)
{
    int const m = static_cast<int>(blockIdx.y);
    int const n = kCoarseDofs * (c.meshAggBegin[m + 1] - c.meshAggBegin[m]);
    int const i0 = (static_cast<int>(blockIdx.x) / tilesPerRow) * 64, j0 = (static_cast<int>(blockIdx.x) % tilesPerRow) * 64;
    if (k >= n || i0 >= n || j0 >= n)
        return;
    int const b = min(kPivot, n - k);
    double const* A = Ak + c.meshOffset[m];
    double* B = Bk + c.meshOffset[m];
    double const* Wm = W + static_cast<size_t>(m) * kPivot * wStride;
    __shared__ double Ai[64][kPivot + 1];
    __shared__ double Wj[kPivot][64 + 1];
    int const t = static_cast<int>(threadIdx.x);
    for (int e = t; e < 64 * kPivot; e += 256)
    {
        int const r = e / kPivot, q = e % kPivot, i = i0 + r;
        Ai[r][q] = (i < n && q < b) ? A[static_cast<size_t>(i) * n + k + q] : 0.;
    }
    for (int e = t; e < kPivot * 64; e += 256)
    {
        int const q = e / 64, cc = e % 64, j = j0 + cc;
        Wj[q][cc] = (j < n && q < b) ? Wm[static_cast<size_t>(q) * wStride + j] : 0.;
    }
    __syncthreads();
    // the 4 x 4 entries (r0 + u, c0 + v) of this thread, A[i, K] W[K, j] summed
    // over K from registers
    int const r0 = (t >> 4) * 4, c0 = (t & 15) * 4;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            acc[u][v] = 0.;
#pragma unroll 8
    for (int q = 0; q < kPivot; ++q)
    {
        double ai[4], wj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            ai[u] = Ai[r0 + u][q];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            wj[v] = Wj[q][c0 + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                acc[u][v] += ai[u] * wj[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        int const i = i0 + r0 + u;
        if (i >= n)
            break;
        bool const iK = i >= k && i < k + b;
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            int const j = j0 + c0 + v;
            if (j >= n)
                break;
            double out;
            if (iK)
                out = Wj[i - k][c0 + v];
            else
                out = ((j >= k && j < k + b) ? 0. : A[static_cast<size_t>(i) * n + j]) - acc[u][v];
            B[static_cast<size_t>(i) * n + j] = out;
        }
    }
}

// Guard after inversion: a mesh block whose inverse has a non-positive or
// non-finite diagonal entry (an indefinite coarse operator) is zeroed, which
// leaves that mesh with the block-Jacobi preconditioner. One workgroup per mesh.
__global__ __launch_bounds__(256) void CoarseGuard(double* __restrict__ buf0, double* __restrict__ buf1,
                                                   int const* __restrict__ finalBuf, CoarseArgs c)
{
    __shared__ int bad;
    int const m = static_cast<int>(blockIdx.x);
    int const n = kCoarseDofs * (c.meshAggBegin[m + 1] - c.meshAggBegin[m]);
    double* A = (finalBuf[m] ? buf1 : buf0) + c.meshOffset[m];
    if (threadIdx.x == 0)
        bad = 0;
    __syncthreads();
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x))
    {
        double const d = A[static_cast<size_t>(i) * n + i];
        if (!(d > 0.) || !isfinite(d))
            bad = 1;
    }
    __syncthreads();
    if (!bad)
        return;
    for (int64_t k = threadIdx.x; k < static_cast<int64_t>(n) * n; k += blockDim.x)
        A[k] = 0.;
}

// The guarded inverse, symmetrised and rounded to FP32 for the CG's coarse
// products (grid.y = mesh): it halves the bytes every iteration reads; the
// preconditioner stays symmetric.
__global__ __launch_bounds__(256) void CoarseToFloat(double const* __restrict__ buf0, double const* __restrict__ buf1,
                                                     int const* __restrict__ finalBuf, CoarseArgs c,
                                                     float* __restrict__ out)
{
    int const m = static_cast<int>(blockIdx.y);
    int const n = kCoarseDofs * (c.meshAggBegin[m + 1] - c.meshAggBegin[m]);
    int64_t const idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(n) * n)
        return;
    int64_t const i = idx / n, j = idx % n;
    double const* A = (finalBuf[m] ? buf1 : buf0) + c.meshOffset[m];
    out[c.meshOffset[m] + idx] = static_cast<float>(0.5 * (A[i * n + j] + A[j * n + i]));
}

__global__ __launch_bounds__(256) void ZeroDoublesLong(double* __restrict__ p, long long n)
{
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        p[i] = 0.;
}

// Clear n doubles and m ints
__global__ __launch_bounds__(256) void ZeroScalars(double* __restrict__ S, int n, int* __restrict__ F, int m)
{
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n + m;
         i += static_cast<int>(gridDim.x * blockDim.x))
    {
        if (i < n)
            S[i] = 0.;
        else
            F[i - n] = 0;
    }
}

__global__ __launch_bounds__(256) void Copy4(float4* __restrict__ dst, float4 const* __restrict__ src, int n)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        dst[i] = src[i];
}

__global__ __launch_bounds__(256) void Fill4(float4* __restrict__ dst, int n)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        dst[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Rescale per-tet Lame coefficients to Young's moduli exp(logY) (nu fixed):
// w*mu and w*lambda of the tet constants and of its incidence records: the
// adjoint's records (w*mu in the id words, w*lambda in g2.y) and the sweep's,
// rebuilt from their rest-shape constants shape = (gamma_a, gamma_b, gamma_c,
// kappa): A = w lambda kappa^2 in the id words, sgeo = (w mu gamma_k,
// kappa (w lambda + w mu)), in FP64 as the host builds them.
struct RecordSet
{
    int4 const* __restrict__ slots; // record slot of each (tet, local vertex), -1: none
    int4* __restrict__ ids;
    float2* __restrict__ g2;
    int4* __restrict__ sids;
    float4* __restrict__ sgeo;
    float4 const* __restrict__ shape;
};

__global__ __launch_bounds__(256) void SetLogYoung(float const* __restrict__ logY, float2 const* __restrict__ base,
                                                   float4* __restrict__ tgp, RecordSet rows, RecordSet ell, int nT)
{
    int const e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= nT)
        return;
    float const Y = expf(logY[e]);
    float2 const b = base[e]; // (w mu / Y, w lambda / Y)
    float const wmu = Y * b.x, wlam = Y * b.y;
    float4 t = tgp[3 * e + 2];
    t.y = wmu, t.z = wlam;
    tgp[3 * e + 2] = t;
    double const dmu = wmu, dlam = wlam;
    auto update = [&](RecordSet const& R) {
        int4 const sl = R.slots[e];
        int const s4[4] = {sl.x, sl.y, sl.z, sl.w};
        for (int k = 0; k < 4; ++k)
        {
            int const q = s4[k];
            if (q < 0)
                continue;
            int4 r = R.ids[q];
            r.w = __float_as_int(wmu);
            R.ids[q] = r;
            R.g2[q].y = wlam;
            float4 const sh = R.shape[q];
            double const kap = sh.w;
            int4 si = R.sids[q];
            si.w = __float_as_int(static_cast<float>(dlam * kap * kap));
            R.sids[q] = si;
            R.sgeo[q] = make_float4(static_cast<float>(dmu * sh.x), static_cast<float>(dmu * sh.y),
                                    static_cast<float>(dmu * sh.z), static_cast<float>(kap * (dlam + dmu)));
        }
    };
    update(rows);
    if (ell.slots)
        update(ell);
}

} // namespace pbat_vbd::device
