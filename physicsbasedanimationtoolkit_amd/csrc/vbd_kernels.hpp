// Device kernels of the MI355X VBD integrator (gfx950, wave64).
//
// Data layout in HBM (internal vertex numbering = colour-major, see
// vbd_integrator.hip):
//   xm[v]    float4 (x, y, z, m)      position and lumped mass
//   xtil[v]  float4 (x~, 0)           inertial target
//   xt[v]    float4 (x_t, 0)          positions at the start of the substep
//   vel[v], vprev[v], aext[v]  float4 velocities and external acceleration
//   vtp[v]   int32 CSR prefix over vertex-tet incidences
//   incidence records (32 B, structure of arrays, in sweep order): the other
//            three vertex ids of the tet and four constants of its GP rows and
//            material (see SweepArgs)
//   tgp[3e..3e+2]  float4 x3 (48 B) per tet: rows 1..3 of the shape function
//            gradients GP (row 0 = -(row1+row2+row3)), w*mu, w*lambda, alpha
//   tetv[e]  int4 vertex ids of tet e
//   (tgp and tetv serve the energy reduction and traces, not the sweep)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vbd_contact_kernels.hpp"

namespace pbat_vbd::device {

// ---- cross-lane sums over groups of L consecutive lanes (L | 64) ----------
template <int kCtrl>
__device__ __forceinline__ float DppRead(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xF, 0xF, false));
}

// Butterfly over L lanes: every lane of the group ends with the group sum.
//   quad_perm[1,0,3,2] (0xB1), quad_perm[2,3,0,1] (0x4E) -> quad sums,
//   row_half_mirror (0x141) pairs the two quads of 8 lanes,
//   row_mirror (0x140) pairs the two halves of a 16-lane row,
//   then ds_swizzle / bpermute for 32 and 64.
template <int L>
__device__ __forceinline__ float GroupSum(float x)
{
    if constexpr (L >= 2)
        x += DppRead<0xB1>(x);
    if constexpr (L >= 4)
        x += DppRead<0x4E>(x);
    if constexpr (L >= 8)
        x += DppRead<0x141>(x);
    if constexpr (L >= 16)
        x += DppRead<0x140>(x);
    if constexpr (L >= 32)
        x += __shfl_xor(x, 16);
    if constexpr (L >= 64)
        x += __shfl_xor(x, 32);
    return x;
}

// 3x3 symmetric block and gradient of one vertex.
struct Block
{
    float h00, h01, h02, h11, h12, h22;
    float g0, g1, g2;
};

template <int L>
__device__ __forceinline__ void GroupSum(Block& B)
{
    B.h00 = GroupSum<L>(B.h00);
    B.h01 = GroupSum<L>(B.h01);
    B.h02 = GroupSum<L>(B.h02);
    B.h11 = GroupSum<L>(B.h11);
    B.h12 = GroupSum<L>(B.h12);
    B.h22 = GroupSum<L>(B.h22);
    B.g0 = GroupSum<L>(B.g0);
    B.g1 = GroupSum<L>(B.g1);
    B.g2 = GroupSum<L>(B.g2);
}

// Elastic contribution of one incident tet to vertex i's block (Stable
// Neo-Hookean, physics/StableNeoHookeanEnergy.h), with c = cof(F) g:
//   grad_i += w mu F g + w lambda (J - alpha) c
//   Hess_i += w mu |g|^2 I + w lambda c c^T
// (the d2J/dF2 term of the SNH Hessian contracts to zero for one vertex:
// det(F0 + x_i g^T) is affine in x_i). alpha = 1 + mu/lambda, so
// w lambda (J - alpha) = w lambda (J - 1) - w mu.
// With a, b, c the tet's other vertices, d_k = x_k - x, G the matrix of their
// shape-function gradient rows g_a, g_b, g_c, g = -(g_a + g_b + g_c) the own
// row and kappa = det G, F = D G with D = [d_a d_b d_c], so
//   F g = sum_k gamma_k d_k (gamma_k = g_k . g),   J = kappa (d_a . n),
//   c = cof(D) cof(G) g = -kappa n,   |g|^2 = -sum_k gamma_k,
// where n = (x_b - x_a) x (x_c - x_a) (the column sum of cof D) and G^-T g =
// -(1, 1, 1). With the record's A = w lambda kappa^2, B = kappa (w lambda + w mu)
// and gamma'_k = w mu gamma_k (rest-shape and material constants, FP64 on the
// host), the terms are
//   grad_i += sum_k gamma'_k d_k - (A (d_a . n) - B) n
//   Hess_i += -(sum_k gamma'_k) I + A n n^T.
__device__ __forceinline__ void AccumulateElastic(float4 const& geo, float A, float4 const& own, float4 const& pa,
                                                  float4 const& pb, float4 const& pc, Block& B)
{
    float const dax = pa.x - own.x, day = pa.y - own.y, daz = pa.z - own.z;
    float const dbx = pb.x - own.x, dby = pb.y - own.y, dbz = pb.z - own.z;
    float const dcx = pc.x - own.x, dcy = pc.y - own.y, dcz = pc.z - own.z;
    float const ex = pb.x - pa.x, ey = pb.y - pa.y, ez = pb.z - pa.z;
    float const fx = pc.x - pa.x, fy = pc.y - pa.y, fz = pc.z - pa.z;
    float const nx = ey * fz - ez * fy, ny = ez * fx - ex * fz, nz = ex * fy - ey * fx;
    float const coef = A * (dax * nx + day * ny + daz * nz) - geo.w;
    B.g0 += geo.x * dax + geo.y * dbx + geo.z * dcx - coef * nx;
    B.g1 += geo.x * day + geo.y * dby + geo.z * dcy - coef * ny;
    B.g2 += geo.x * daz + geo.y * dbz + geo.z * dcz - coef * nz;
    float const dg = -(geo.x + geo.y + geo.z);
    float const anx = A * nx, any = A * ny, anz = A * nz;
    B.h00 += dg + anx * nx;
    B.h11 += dg + any * ny;
    B.h22 += dg + anz * nz;
    B.h01 += anx * ny;
    B.h02 += anx * nz;
    B.h12 += any * nz;
}

struct StepConstants
{
    float dt;       // substep h
    float dt2;      // h^2
    float kD;       // Rayleigh damping coefficient
    float detHZero; // skip threshold on |det H|
    unsigned int* fp64Solves; // counts 3x3 solves taken by the FP64 path (nullable)
};

// Rayleigh damping of the elastic block (sim/vbd/Kernels.h:173-191).
__device__ __forceinline__ void Damp(Block& B, float4 const& xv, float4 const& xtv, StepConstants const& c)
{
    if (c.kD != 0.f)
    {
        float const D = c.kD / c.dt;
        float const ux = xv.x - xtv.x, uy = xv.y - xtv.y, uz = xv.z - xtv.z;
        B.g0 += D * (B.h00 * ux + B.h01 * uy + B.h02 * uz);
        B.g1 += D * (B.h01 * ux + B.h11 * uy + B.h12 * uz);
        B.g2 += D * (B.h02 * ux + B.h12 * uy + B.h22 * uz);
        float const s = 1.f + D;
        B.h00 *= s, B.h01 *= s, B.h02 *= s, B.h11 *= s, B.h12 *= s, B.h22 *= s;
    }
}

__device__ __forceinline__ float3 InertiaNewtonStep(Block B, float4 const& xv, float4 const& xl,
                                                    StepConstants const& c);

// Damping, inertia and the 3x3 Newton step of one vertex
// (sim/vbd/Kernels.h:173-191, 304-339). Returns the displacement to subtract.
__device__ __forceinline__ float3 NewtonStep(Block B, float4 const& xv, float4 const& xl, float4 const& xtv,
                                             StepConstants const& c)
{
    Damp(B, xv, xtv, c);
    return InertiaNewtonStep(B, xv, xl, c);
}

// Inertia and the 3x3 Newton step (sim/vbd/Kernels.h:304-339).
__device__ __forceinline__ float3 InertiaNewtonStep(Block B, float4 const& xv, float4 const& xl,
                                                    StepConstants const& c)
{
    float const K = xv.w / c.dt2;
    B.h00 += K, B.h11 += K, B.h22 += K;
    B.g0 += K * (xv.x - xl.x);
    B.g1 += K * (xv.y - xl.y);
    B.g2 += K * (xv.z - xl.z);
    // adjugate of the symmetric block
    float const a00 = B.h11 * B.h22 - B.h12 * B.h12;
    float const a01 = B.h02 * B.h12 - B.h01 * B.h22;
    float const a02 = B.h01 * B.h12 - B.h02 * B.h11;
    float const det = B.h00 * a00 + B.h01 * a01 + B.h02 * a02;
    if (!(fabsf(det) > c.detHZero))
        return make_float3(0.f, 0.f, 0.f);
    float const dscale = fabsf(B.h00 * B.h11 * B.h22);
    if (fabsf(det) < 1e-4f * dscale)
    {
        // ill-conditioned block: solve in FP64 from the FP32 entries
        if (c.fp64Solves)
            atomicAdd(c.fp64Solves, 1u);
        double const h00 = B.h00, h01 = B.h01, h02 = B.h02, h11 = B.h11, h12 = B.h12, h22 = B.h22;
        double const b00 = h11 * h22 - h12 * h12, b01 = h02 * h12 - h01 * h22, b02 = h01 * h12 - h02 * h11;
        double const b11 = h00 * h22 - h02 * h02, b12 = h01 * h02 - h00 * h12, b22 = h00 * h11 - h01 * h01;
        double const dd = h00 * b00 + h01 * b01 + h02 * b02;
        double const gx = B.g0, gy = B.g1, gz = B.g2;
        return make_float3(static_cast<float>((b00 * gx + b01 * gy + b02 * gz) / dd),
                           static_cast<float>((b01 * gx + b11 * gy + b12 * gz) / dd),
                           static_cast<float>((b02 * gx + b12 * gy + b22 * gz) / dd));
    }
    float const a11 = B.h00 * B.h22 - B.h02 * B.h02;
    float const a12 = B.h01 * B.h02 - B.h00 * B.h12;
    float const a22 = B.h00 * B.h11 - B.h01 * B.h01;
    float const inv = 1.f / det;
    return make_float3((a00 * B.g0 + a01 * B.g1 + a02 * B.g2) * inv, (a01 * B.g0 + a11 * B.g1 + a12 * B.g2) * inv,
                       (a02 * B.g0 + a12 * B.g1 + a22 * B.g2) * inv);
}

// Incidence records, structure of arrays, row k of vertex v's incidences
// (CSR rows through vtp, or fixed-stride rows of ellStride entries), 32 B:
//   ids[k] int4   (a, b, c, A as bits): the tet's other three vertices in
//                 local order; a = -1 marks an unused fixed-stride slot
//   geo[k] float4 (gamma'_a, gamma'_b, gamma'_c, B) (AccumulateElastic)
// Records are stored in sweep order, so a colour pass streams them.
struct SweepArgs
{
    int const* __restrict__ vtp;
    int4 const* __restrict__ ids;
    float4 const* __restrict__ geo;
    float4* __restrict__ xm;
    float4 const* __restrict__ xtil;
    float4 const* __restrict__ xt;
    int vbeg;
    int vend;
    int xcdChunk;  // > 0: blocks per XCD for the XCD-aware block order
    int ntStream;  // != 0: read the incidence records with non-temporal loads
    int ellStride; // > 0: fixed-stride rows of this many records
    StepConstants c;
    ContactSweep contact; // contact sweeps only
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));

// Loads of the incidence stream, optionally with the non-temporal hint.
__device__ __forceinline__ int4 LoadIds(int4 const* p, bool nt)
{
    if (nt)
    {
        v4i const r = __builtin_nontemporal_load(reinterpret_cast<v4i const*>(p));
        return make_int4(r.x, r.y, r.z, r.w);
    }
    return *p;
}

__device__ __forceinline__ float4 LoadConst(float4 const* p, bool nt)
{
    if (nt)
    {
        v4f const r = __builtin_nontemporal_load(reinterpret_cast<v4f const*>(p));
        return make_float4(r.x, r.y, r.z, r.w);
    }
    return *p;
}

__device__ __forceinline__ float2 LoadConst(float2 const* p, bool nt)
{
    if (nt)
    {
        v2f const r = __builtin_nontemporal_load(reinterpret_cast<v2f const*>(p));
        return make_float2(r.x, r.y);
    }
    return *p;
}

// One colour of one VBD sweep (sim/vbd/Integrator.cpp:73-136,
// gpu/impl/vbd/Kernels.cuh:148-233). L lanes cooperate on each vertex: lane l
// of the group takes incidences l, l+L, ... of the vertex's row, first
// issuing the loads of its first K incidences together, then accumulating.
// The group reduces its 9 partial sums with DPP and its first lane applies the
// Newton update in place: vertices of one colour share no tet, so none reads
// what another vertex of the colour writes.
// With C (scenes in contact), the vertex's contact terms join its block after
// damping and before inertia (gpu/impl/vbd/Kernels.cuh:200-230), and the new
// positions go to contact.xb, so contact terms of the colour read positions from
// before its pass (the reference's write buffer, gpu/impl/vbd/Integrator.cu:303-326).
// The previous colour's positions are read from xb and copied back to xm by
// this launch; the last colour of an iteration is copied back by the host.
template <int L, int K, bool C = false>
__global__ __launch_bounds__(256) void VbdSweep(SweepArgs a)
{
    int block = static_cast<int>(blockIdx.x);
    if (a.xcdChunk > 0)
    {
        // Blocks are dealt to the 8 XCDs round-robin; give each XCD a
        // contiguous range of vertices so neighbouring vertices share its L2.
        // The host pads the grid to 8 * xcdChunk blocks, which makes this a
        // permutation of [0, gridDim.x).
        block = (block & 7) * a.xcdChunk + (block >> 3);
    }
    int const tid = block * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    int const lane = tid & (L - 1);
    int const v = a.vbeg + tid / L;
    bool const active = v < a.vend;
    int kb = 0, ke = 0;
    float4 own = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 xl = own;
    if (active)
    {
        if (a.ellStride > 0)
        {
            // fixed stride: the incidence address needs no prefix load
            kb = v * a.ellStride;
            ke = kb + a.ellStride;
        }
        else
        {
            kb = a.vtp[v];
            ke = a.vtp[v + 1];
        }
        own = a.xm[v];
        if (lane == 0 && K < 3)
            xl = a.xtil[v]; // issued now, used after the reduction
    }
    Block B{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool const nt = a.ntStream != 0;
    // contact sweeps: the start position and the lane's contact records (slots
    // lane, lane + L, ...), issued now; then copy the previous colour's new
    // positions back
    constexpr int kSlots = C ? (kMaxContacts + L - 1) / L : 1;
    int4 crec[kSlots];
    float4 xto = own;
    if constexpr (C)
    {
#pragma unroll
        for (int j = 0; j < kSlots; ++j)
        {
            int const k = lane + j * L;
            crec[j] = (active && k < kMaxContacts) ? a.contact.rec[static_cast<size_t>(v) * kMaxContacts + k]
                                                   : make_int4(-1, -1, -1, 0);
        }
        if (active)
            xto = a.xt[v];
        int const pend = a.contact.pendBeg + tid;
        if (pend < a.contact.pendEnd)
            a.xm[pend] = a.contact.xb[pend];
    }
    int4 I[K];
    float4 q0[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
    {
        int const k = kb + lane + j * L;
        if (k < ke)
        {
            I[j] = LoadIds(a.ids + k, nt);
            q0[j] = LoadConst(a.geo + k, nt);
        }
        else
            I[j] = make_int4(-1, 0, 0, 0);
    }
    float4 pa[K], pb[K], pc[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
    {
        if (I[j].x >= 0)
        {
            if constexpr (C)
            {
                pa[j] = ContactPosition(a.contact, a.xm, I[j].x);
                pb[j] = ContactPosition(a.contact, a.xm, I[j].y);
                pc[j] = ContactPosition(a.contact, a.xm, I[j].z);
            }
            else
            {
                pa[j] = a.xm[I[j].x];
                pb[j] = a.xm[I[j].y];
                pc[j] = a.xm[I[j].z];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
        if (I[j].x >= 0)
            AccumulateElastic(q0[j], __int_as_float(I[j].w), own, pa[j], pb[j], pc[j], B);
    // incidences beyond K per lane (vertices of unusually high degree)
    for (int k = kb + lane + K * L; k < ke; k += L)
    {
        int4 const J = LoadIds(a.ids + k, nt);
        if (J.x < 0)
            continue;
        if constexpr (C)
            AccumulateElastic(LoadConst(a.geo + k, nt), __int_as_float(J.w), own,
                              ContactPosition(a.contact, a.xm, J.x), ContactPosition(a.contact, a.xm, J.y),
                              ContactPosition(a.contact, a.xm, J.z), B);
        else
            AccumulateElastic(LoadConst(a.geo + k, nt), __int_as_float(J.w), own, a.xm[J.x], a.xm[J.y],
                              a.xm[J.z], B);
    }
    if constexpr (C)
    {
        // contact terms (after damping, before inertia): the group's lanes take
        // the vertex's contact slots k = lane, lane + L, ... and sum them apart
        // from the elastic block, which is damped alone
        ContactSweep const& cs = a.contact;
        Block Cb{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float3 const xo = make_float3(own.x, own.y, own.z), xt3 = make_float3(xto.x, xto.y, xto.z);
#pragma unroll
        for (int j = 0; j < kSlots; ++j)
            if (crec[j].x >= 0)
                AccumulateContact(xt3, xo, cs, a.xm, a.xt, crec[j], a.c.dt, __int_as_float(crec[j].w), cs.muF,
                                  cs.epsv, Cb);
        GroupSum<L>(B);
        GroupSum<L>(Cb);
        if (!active || lane != 0)
            return;
        if constexpr (K >= 3)
            xl = a.xtil[v];
        Damp(B, own, xto, a.c);
        B.h00 += Cb.h00, B.h01 += Cb.h01, B.h02 += Cb.h02, B.h11 += Cb.h11, B.h12 += Cb.h12, B.h22 += Cb.h22;
        B.g0 += Cb.g0, B.g1 += Cb.g1, B.g2 += Cb.g2;
        float3 const dx = InertiaNewtonStep(B, own, xl, a.c);
        cs.xb[v] = make_float4(own.x - dx.x, own.y - dx.y, own.z - dx.z, own.w);
        return;
    }
    GroupSum<L>(B);
    if (!active || lane != 0)
        return;
    if constexpr (K >= 3)
        xl = a.xtil[v];
    // the start position only matters with damping: loaded here rather than
    // held in registers through the incidence loop
    float3 const dx = NewtonStep(B, own, xl, (a.c.kD != 0.f) ? a.xt[v] : own, a.c);
    a.xm[v] = make_float4(own.x - dx.x, own.y - dx.y, own.z - dx.z, own.w);
}

// Substep begin (gpu/impl/vbd/Integrator.cu:190-248; sim/vbd/Kernels.h:24-94):
// inertial target x~ = xt + h v + h^2 a and the initial iterate.
[[maybe_unused]] static __global__ __launch_bounds__(256) void SubstepBegin(float4* __restrict__ xm, float4* __restrict__ xtil,
                                                    float4 const* __restrict__ xt, float4 const* __restrict__ vel,
                                                    float4 const* __restrict__ vprev, float4 const* __restrict__ aext,
                                                    int n, float dt, float dt2, int strategy)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const p = xt[i];
    float4 const u = vel[i];
    float4 const g = aext[i];
    xtil[i] = make_float4(p.x + dt * u.x + dt2 * g.x, p.y + dt * u.y + dt2 * g.y, p.z + dt * u.z + dt2 * g.z, 0.f);
    float3 x0;
    if (strategy == 0) // Position
        x0 = make_float3(p.x, p.y, p.z);
    else if (strategy == 1) // Inertia
        x0 = make_float3(p.x + dt * u.x, p.y + dt * u.y, p.z + dt * u.z);
    else if (strategy == 2) // KineticEnergyMinimum
        x0 = make_float3(p.x + dt * u.x + dt2 * g.x, p.y + dt * u.y + dt2 * g.y, p.z + dt * u.z + dt2 * g.z);
    else
    {
        float const an2 = g.x * g.x + g.y * g.y + g.z * g.z;
        float at = 0.f;
        if (an2 != 0.f)
        {
            if (strategy == 3) // AdaptiveVbd
            {
                float4 const w = vprev[i];
                float const dot = ((u.x - w.x) / dt) * g.x + ((u.y - w.y) / dt) * g.y + ((u.z - w.z) / dt) * g.z;
                at = fminf(fmaxf(dot / an2, 0.f), 1.f);
            }
            else // AdaptivePbat
            {
                float const vn = sqrtf(u.x * u.x + u.y * u.y + u.z * u.z) + 1.17549435e-38f;
                float const dot = (u.x / vn) * g.x + (u.y / vn) * g.y + (u.z / vn) * g.z;
                at = fminf(fabsf(dot / an2), 1.f);
            }
        }
        float const s = dt2 * at;
        x0 = make_float3(p.x + dt * u.x + s * g.x, p.y + dt * u.y + s * g.y, p.z + dt * u.z + s * g.z);
    }
    float const mass = xm[i].w;
    xm[i] = make_float4(x0.x, x0.y, x0.z, mass);
}

// Substep end (gpu/impl/vbd/Integrator.cu:329-347): vprev <- v, v = (x - xt)/h, xt <- x.
// Every vertex passes here once per substep, so this is where a non-finite
// position or velocity (a diverged solve, a NaN set by the caller) is flagged:
// *nonfinite = 1 (nullable; written only then).
[[maybe_unused]] static __global__ __launch_bounds__(256) void SubstepEnd(float4 const* __restrict__ xm, float4* __restrict__ xt,
                                                  float4* __restrict__ vel, float4* __restrict__ vprev, int n,
                                                  float dt, int* __restrict__ nonfinite)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const x = xm[i];
    float4 const p = xt[i];
    vprev[i] = vel[i];
    float4 const u = make_float4((x.x - p.x) / dt, (x.y - p.y) / dt, (x.z - p.z) / dt, 0.f);
    vel[i] = u;
    xt[i] = make_float4(x.x, x.y, x.z, 0.f);
    if (nonfinite && !(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(u.x) && isfinite(u.y) &&
                       isfinite(u.z)))
        *nonfinite = 1;
}

// Chebyshev semi-iterative update after iteration k (sim/vbd/Kernels.h:104-118,
// gpu/impl/vbd/ChebyshevIntegrator.cu:38-59): x = omega (x - x_{k-2}) + x_{k-2}
// for k > 1, then x_{k-2} = x_{k-1}, x_{k-1} = x.
[[maybe_unused]] static __global__ __launch_bounds__(256) void ChebyshevUpdate(float4* __restrict__ xm,
                                                                               float4* __restrict__ xkm1,
                                                                               float4* __restrict__ xkm2, int n,
                                                                               int k, float omega)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 x = xm[i];
    if (k > 1)
    {
        float4 const a = xkm2[i];
        x.x = omega * (x.x - a.x) + a.x;
        x.y = omega * (x.y - a.y) + a.y;
        x.z = omega * (x.z - a.z) + a.z;
        xm[i] = x;
    }
    xkm2[i] = xkm1[i];
    xkm1[i] = x;
}

// ---- Anderson acceleration (gpu/impl/vbd/AndersonIntegrator.cu:36-105) ------
// After VBD iteration k >= 1 of a substep, with G = x, F = G - x_{k-1} and the
// window of the last min(k, m) differences DG, DF (column (k-1) mod m is the
// newest): alpha = argmin |DF alpha - F| and x = G - DG alpha. The least
// squares is solved from the FP64 Gram matrix DF^T DF and DF^T F; only the
// newest column's row of the Gram matrix is recomputed each iteration.

constexpr int kMaxAndersonWindow = 16;
constexpr int kAndersonSpread = 16; // atomic targets per accumulated value
constexpr int kAndersonStride = 16; // doubles between targets (128 B)

// Accumulated values of one iteration: 2 kMaxAndersonWindow values
// (DF_j . DF_new, then DF_j . F), each over kAndersonSpread targets.
__host__ __device__ constexpr int AndersonAccDoubles()
{
    return 2 * kMaxAndersonWindow * kAndersonSpread * kAndersonStride;
}

// Before iteration 0 of a substep: x_{k-1} = x; clear both accumulator sets.
[[maybe_unused]] static __global__ __launch_bounds__(256) void AndersonStart(float4 const* __restrict__ xm,
                                                                             float4* __restrict__ xkm1, int n,
                                                                             double* __restrict__ acc)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        xkm1[i] = xm[i];
    if (blockIdx.x == 0)
        for (int j = static_cast<int>(threadIdx.x); j < 2 * AndersonAccDoubles(); j += static_cast<int>(blockDim.x))
            acc[j] = 0.;
}

// After iteration 0: G_{k-1} = x, F_{k-1} = x - x_{k-1}, then x_{k-1} = x.
[[maybe_unused]] static __global__ __launch_bounds__(256) void AndersonFirst(float4 const* __restrict__ xm,
                                                                             float4* __restrict__ xkm1,
                                                                             float4* __restrict__ gkm1,
                                                                             float4* __restrict__ fkm1, int n)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const g = xm[i], p = xkm1[i];
    gkm1[i] = make_float4(g.x, g.y, g.z, 0.f);
    fkm1[i] = make_float4(g.x - p.x, g.y - p.y, g.z - p.z, 0.f);
    xkm1[i] = make_float4(g.x, g.y, g.z, 0.f);
}

__device__ __forceinline__ double Dot3d(float4 const& u, float4 const& w)
{
    return double(u.x) * w.x + double(u.y) * w.y + double(u.z) * w.z;
}

// After iteration k >= 1: update the window (column col) and accumulate
// DF_j . DF_col and DF_j . F for j < mk into acc. blockDim = 256.
template <int M>
__global__ __launch_bounds__(256) void AndersonWindow(float4 const* __restrict__ xm, float4 const* __restrict__ xkm1,
                                                      float4* __restrict__ gkm1, float4* __restrict__ fkm1,
                                                      float4* __restrict__ dgk, float4* __restrict__ dfk, int n,
                                                      int col, int mk, double* __restrict__ acc)
{
    __shared__ double red[4][2 * M];
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    double vals[2 * M];
#pragma unroll
    for (int j = 0; j < 2 * M; ++j)
        vals[j] = 0.;
    if (i < n)
    {
        float4 const g = xm[i], p = xkm1[i], g1 = gkm1[i], f1 = fkm1[i];
        float4 const f = make_float4(g.x - p.x, g.y - p.y, g.z - p.z, 0.f);
        float4 const dg = make_float4(g.x - g1.x, g.y - g1.y, g.z - g1.z, 0.f);
        float4 const df = make_float4(f.x - f1.x, f.y - f1.y, f.z - f1.z, 0.f);
        dgk[static_cast<size_t>(col) * n + i] = dg;
        dfk[static_cast<size_t>(col) * n + i] = df;
        gkm1[i] = make_float4(g.x, g.y, g.z, 0.f);
        fkm1[i] = f;
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < mk)
            {
                float4 const dj = (j == col) ? df : dfk[static_cast<size_t>(j) * n + i];
                vals[j] = Dot3d(dj, df);
                vals[M + j] = Dot3d(dj, f);
            }
    }
    int const lane = static_cast<int>(threadIdx.x & 63), wave = static_cast<int>(threadIdx.x >> 6);
#pragma unroll
    for (int j = 0; j < 2 * M; ++j)
    {
        if ((j < M ? j : j - M) >= mk)
            continue; // a window column not filled yet (uniform over the launch)
        double v = vals[j];
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off);
        if (lane == 0)
            red[wave][j] = v;
    }
    __syncthreads();
    int const nw = static_cast<int>((blockDim.x + 63) / 64);
    int const t = static_cast<int>(threadIdx.x);
    if (t < 2 * M && ((t < M) ? t : t - M) < mk)
    {
        double sum = 0.;
        for (int w = 0; w < nw; ++w)
            sum += red[w][t];
        int const slot = static_cast<int>(blockIdx.x) & (kAndersonSpread - 1);
        int const idx = (t < M) ? t : kMaxAndersonWindow + (t - M);
        atomicAdd(acc + (static_cast<size_t>(idx) * kAndersonSpread + slot) * kAndersonStride, sum);
    }
}

// After AndersonWindow: every workgroup forms the mk x mk Gram matrix (the
// previous one, gramOld, with row and column col replaced) and solves for
// alpha by an FP64 Cholesky factorisation that drops columns whose pivot is
// below 1e-13 of the largest diagonal (a zero or dependent difference); the
// first workgroup stores the Gram matrix (gramNew) and clears accNext. Then
// x = G - DG alpha and x_{k-1} = x.
[[maybe_unused]] static __global__ __launch_bounds__(256) void AndersonApply(float4* __restrict__ xm, float4* __restrict__ xkm1,
                                                     float4 const* __restrict__ gkm1,
                                                     float4 const* __restrict__ dgk, int n, int m, int col, int mk,
                                                     double const* __restrict__ acc, double* __restrict__ accNext,
                                                     double const* __restrict__ gramOld, double* __restrict__ gramNew)
{
    constexpr int W = kMaxAndersonWindow;
    __shared__ double sG[W * W], sL[W * W];
    __shared__ double sb[W], sy[W], sa[W];
    __shared__ int sKeep[W];
    __shared__ float sAlpha[W];
    int const t = static_cast<int>(threadIdx.x);
    auto sum = [&](int idx) {
        double s = 0.;
        for (int k = 0; k < kAndersonSpread; ++k)
            s += acc[(static_cast<size_t>(idx) * kAndersonSpread + k) * kAndersonStride];
        return s;
    };
    if (t < mk * mk)
    {
        int const r = t / mk, c = t % mk;
        double g;
        if (r == col)
            g = sum(c);
        else if (c == col)
            g = sum(r);
        else
            g = gramOld[r * m + c];
        sG[r * W + c] = g;
        if (blockIdx.x == 0)
            gramNew[r * m + c] = g;
    }
    if (t < mk)
        sb[t] = sum(kMaxAndersonWindow + t);
    __syncthreads();
    if (t == 0)
    {
        // Cholesky factor L (column j of L in sL[. * W + j]) with dropped columns
        double dmax = 0.;
        for (int j = 0; j < mk; ++j)
            dmax = fmax(dmax, sG[j * W + j]);
        for (int j = 0; j < mk; ++j)
        {
            double d = sG[j * W + j];
            for (int p = 0; p < j; ++p)
                d -= sL[j * W + p] * sL[j * W + p];
            int const keep = (d > 1e-13 * dmax && d > 0.) ? 1 : 0;
            sKeep[j] = keep;
            double const ljj = keep ? sqrt(d) : 0.;
            sL[j * W + j] = ljj;
            for (int r = j + 1; r < mk; ++r)
            {
                double v = sG[r * W + j];
                for (int p = 0; p < j; ++p)
                    v -= sL[r * W + p] * sL[j * W + p];
                sL[r * W + j] = keep ? v / ljj : 0.;
            }
        }
        for (int j = 0; j < mk; ++j)
        {
            double v = sb[j];
            for (int p = 0; p < j; ++p)
                v -= sL[j * W + p] * sy[p];
            sy[j] = sKeep[j] ? v / sL[j * W + j] : 0.;
        }
        for (int j = mk - 1; j >= 0; --j)
        {
            double v = sy[j];
            for (int p = j + 1; p < mk; ++p)
                v -= sL[p * W + j] * sa[p];
            sa[j] = sKeep[j] ? v / sL[j * W + j] : 0.;
        }
        for (int j = 0; j < mk; ++j)
            sAlpha[j] = static_cast<float>(sa[j]);
    }
    if (blockIdx.x == 0)
        for (int j = t; j < AndersonAccDoubles(); j += static_cast<int>(blockDim.x))
            accNext[j] = 0.;
    __syncthreads();
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const g = gkm1[i];
    float x0 = g.x, x1 = g.y, x2 = g.z;
    for (int j = 0; j < mk; ++j)
    {
        float4 const d = dgk[static_cast<size_t>(j) * n + i];
        float const aj = sAlpha[j];
        x0 -= aj * d.x, x1 -= aj * d.y, x2 -= aj * d.z;
    }
    float const mass = xm[i].w;
    xm[i] = make_float4(x0, x1, x2, mass);
    xkm1[i] = make_float4(x0, x1, x2, 0.f);
}

// w Psi(F) of tet e in FP64: Stable Neo-Hookean (physics/StableNeoHookeanEnergy.h:737-755)
// with the quadrature weight folded into the tgp record's (w mu, w lambda).
__device__ inline double TetElasticEnergy(int4 const* __restrict__ tetv, float4 const* __restrict__ tgp,
                                          float4 const* __restrict__ xm, int e)
{
    int4 const t = tetv[e];
    float4 const q0 = tgp[3 * e], q1 = tgp[3 * e + 1], q2 = tgp[3 * e + 2];
    float4 const p0 = xm[t.x], p1 = xm[t.y], p2 = xm[t.z], p3 = xm[t.w];
    double const r[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
    double const d[3][3] = {{double(p1.x) - p0.x, double(p1.y) - p0.y, double(p1.z) - p0.z},
                            {double(p2.x) - p0.x, double(p2.y) - p0.y, double(p2.z) - p0.z},
                            {double(p3.x) - p0.x, double(p3.y) - p0.y, double(p3.z) - p0.z}};
    double F[3][3]; // F[row][col] = sum_k d[k][row] r[k][col]
    double I2 = 0.;
    for (int rr = 0; rr < 3; ++rr)
        for (int c = 0; c < 3; ++c)
        {
            F[rr][c] = d[0][rr] * r[0][c] + d[1][rr] * r[1][c] + d[2][rr] * r[2][c];
            I2 += F[rr][c] * F[rr][c];
        }
    double const J = F[0][0] * (F[1][1] * F[2][2] - F[2][1] * F[1][2]) -
                     F[0][1] * (F[1][0] * F[2][2] - F[2][0] * F[1][2]) +
                     F[0][2] * (F[1][0] * F[2][1] - F[2][0] * F[1][1]);
    double const q = J - double(q2.w);
    return 0.5 * double(q2.y) * (I2 - 3.) + 0.5 * double(q2.z) * q * q;
}

// Sum three doubles over a 256-thread workgroup; the result is valid in thread 0.
__device__ inline void BlockSum3(double& a, double& b, double& c)
{
    __shared__ double red[3][4];
    for (int off = 32; off > 0; off >>= 1)
    {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    int const w = static_cast<int>(threadIdx.x) / 64;
    if ((threadIdx.x & 63) == 0)
    {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        b = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        c = red[2][0] + red[2][1] + red[2][2] + red[2][3];
    }
}

// FP64 energy reduction: out[0] += 1/2 m |v|^2, out[2] += -m a.x over vertices,
// out[1] += w Psi(F) over tets. One atomic per block and term. blockDim = 256.
[[maybe_unused]] static __global__ __launch_bounds__(256) void EnergyTerms(float4 const* __restrict__ xm, float4 const* __restrict__ vel,
                                                   float4 const* __restrict__ aext, int4 const* __restrict__ tetv,
                                                   float4 const* __restrict__ tgp, int nV, int nT,
                                                   double* __restrict__ out)
{
    double K = 0., U = 0., P = 0.;
    int const stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < nV; i += stride)
    {
        float4 const x = xm[i], u = vel[i], g = aext[i];
        double const m = x.w;
        K += 0.5 * m * (double(u.x) * u.x + double(u.y) * u.y + double(u.z) * u.z);
        P -= m * (double(g.x) * x.x + double(g.y) * x.y + double(g.z) * x.z);
    }
    for (int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); e < nT; e += stride)
        U += TetElasticEnergy(tetv, tgp, xm, e);
    BlockSum3(K, U, P);
    if (threadIdx.x == 0)
    {
        atomicAdd(out + 0, K);
        atomicAdd(out + 1, U);
        atomicAdd(out + 2, P);
    }
}

// Trust-region reductions (gpu/impl/vbd/TrustRegionIntegrator.cu:265-364 objective
// terms without contact, :471-489 squared step): per workgroup partial sums of
// m |x - x~|^2, w Psi(F) and, when xprev is given, |x - xprev|^2, written to
// part[3 * blockIdx.x + j]. No atomics, so SumPartials3 gives the same bits on
// every run. blockDim = 256.
[[maybe_unused]] static __global__ __launch_bounds__(256) void ObjectiveTerms(
    float4 const* __restrict__ xm, float4 const* __restrict__ xtil, float4 const* __restrict__ xprev,
    int4 const* __restrict__ tetv, float4 const* __restrict__ tgp, int nV, int nT, double* __restrict__ part,
    double const* __restrict__ gate)
{
    if (gate && *gate == 0.)
        return; // not needed this time: the partials keep their last values
    double K = 0., U = 0., D = 0.;
    int const stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < nV; i += stride)
    {
        float4 const x = xm[i], y = xtil[i];
        double const dx = double(x.x) - y.x, dy = double(x.y) - y.y, dz = double(x.z) - y.z;
        K += double(x.w) * (dx * dx + dy * dy + dz * dz);
        if (xprev)
        {
            float4 const p = xprev[i];
            double const ex = double(x.x) - p.x, ey = double(x.y) - p.y, ez = double(x.z) - p.z;
            D += ex * ex + ey * ey + ez * ez;
        }
    }
    for (int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); e < nT; e += stride)
        U += TetElasticEnergy(tetv, tgp, xm, e);
    BlockSum3(K, U, D);
    if (threadIdx.x == 0)
    {
        part[3 * blockIdx.x + 0] = K;
        part[3 * blockIdx.x + 1] = U;
        part[3 * blockIdx.x + 2] = D;
    }
}

// Gradient of the incremental potential f = 1/2 |x - x~|_M^2 + h^2 sum_e w_e Psi_e
// (sim::vbd::Integrator::ObjectiveFunctionGradient, sim/vbd/Integrator.cpp:168-200) in
// FP64 from the FP32 state, one thread per vertex v of [0, nV):
//   g_v = m_v (x_v - x~_v) + h^2 sum_{(e, l) incident to v} (w_e P(F_e)) gp_{e,l}
// with P(F) = mu F + lambda (J - alpha) cof F the Stable Neo-Hookean stress
// (physics/StableNeoHookeanEnergy.h:766-795) and gp_{e,l} row l of the tet's shape
// function gradients. inc lists 4 e + l per vertex (CSR vtp), ascending e, so each
// vertex's sum has a fixed order (no atomics). grad: 3 doubles per vertex (internal
// order). part[3 b] receives workgroup b's sum of |g_v|^2 over v < nUpd (the
// non-Dirichlet vertices), part[3 b + 1..2] = 0. blockDim = 256.
[[maybe_unused]] static __global__ __launch_bounds__(256) void ObjectiveGradient(
    float4 const* __restrict__ xm, float4 const* __restrict__ xtil, int const* __restrict__ vtp,
    int const* __restrict__ inc, int4 const* __restrict__ tetv, float4 const* __restrict__ tgp, int nV, int nUpd,
    double dt2, double* __restrict__ grad, double* __restrict__ part)
{
    int const v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    double r2 = 0.;
    if (v < nV)
    {
        double gx = 0., gy = 0., gz = 0.;
        for (int k = vtp[v]; k < vtp[v + 1]; ++k)
        {
            int const e = inc[k] >> 2, l = inc[k] & 3;
            int4 const t = tetv[e];
            float4 const q0 = tgp[3 * e], q1 = tgp[3 * e + 1], q2 = tgp[3 * e + 2];
            float4 const p0 = xm[t.x], p1 = xm[t.y], p2 = xm[t.z], p3 = xm[t.w];
            double const r[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
            double const d[3][3] = {{double(p1.x) - p0.x, double(p1.y) - p0.y, double(p1.z) - p0.z},
                                    {double(p2.x) - p0.x, double(p2.y) - p0.y, double(p2.z) - p0.z},
                                    {double(p3.x) - p0.x, double(p3.y) - p0.y, double(p3.z) - p0.z}};
            double F[3][3]; // F[row][col] = sum_k d[k][row] r[k][col]
            for (int rr = 0; rr < 3; ++rr)
                for (int c = 0; c < 3; ++c)
                    F[rr][c] = d[0][rr] * r[0][c] + d[1][rr] * r[1][c] + d[2][rr] * r[2][c];
            double cof[3][3]; // column c = f_{c+1} x f_{c+2}
            for (int c = 0; c < 3; ++c)
            {
                int const c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                cof[0][c] = F[1][c1] * F[2][c2] - F[2][c1] * F[1][c2];
                cof[1][c] = F[2][c1] * F[0][c2] - F[0][c1] * F[2][c2];
                cof[2][c] = F[0][c1] * F[1][c2] - F[1][c1] * F[0][c2];
            }
            double const J = F[0][0] * cof[0][0] + F[1][0] * cof[1][0] + F[2][0] * cof[2][0];
            double const wmu = q2.y, s = double(q2.z) * (J - double(q2.w));
            double gp[3];
            for (int c = 0; c < 3; ++c)
                gp[c] = (l == 0) ? -(r[0][c] + r[1][c] + r[2][c]) : r[l - 1][c];
            double out[3];
            for (int rr = 0; rr < 3; ++rr)
                out[rr] = (wmu * F[rr][0] + s * cof[rr][0]) * gp[0] + (wmu * F[rr][1] + s * cof[rr][1]) * gp[1] +
                          (wmu * F[rr][2] + s * cof[rr][2]) * gp[2];
            gx += out[0], gy += out[1], gz += out[2];
        }
        float4 const x = xm[v], y = xtil[v];
        double const m = x.w;
        gx = m * (double(x.x) - y.x) + dt2 * gx;
        gy = m * (double(x.y) - y.y) + dt2 * gy;
        gz = m * (double(x.z) - y.z) + dt2 * gz;
        grad[3 * v] = gx, grad[3 * v + 1] = gy, grad[3 * v + 2] = gz;
        if (v < nUpd)
            r2 = gx * gx + gy * gy + gz * gz;
    }
    double z0 = 0., z1 = 0.;
    BlockSum3(r2, z0, z1);
    if (threadIdx.x == 0)
    {
        part[3 * blockIdx.x] = r2;
        part[3 * blockIdx.x + 1] = 0.;
        part[3 * blockIdx.x + 2] = 0.;
    }
}

// ---- linear trust-region path (gpu/impl/vbd/TrustRegionIntegrator.cu:36-263) ----
// The model's FP64 scalars live in device memory; TrModel and the epilogue of
// the objective's last reduction (SumPartials3) update them between the sweeps,
// so a step reads nothing back to the host and can be replayed from a graph.
enum TrSlot
{
    kTrF = 0,   // f_k
    kTrT,       // t_k
    kTrFm1,     // f_{k-1}
    kTrFm2,     // f_{k-2}
    kTrTm1,     // t_{k-1}
    kTrTm2,     // t_{k-2}
    kTrR2,      // squared trust radius
    kTrQa,      // model m(t) = qa t^2 + qb t + qc
    kTrQb,
    kTrQc,
    kTrFVbd,    // f at the VBD iterate
    kTrStep,    // step length t tried this iteration
    kTrUpper,   // upper bound of t
    kTrAttempt, // 1 when t is tried (t != 1)
    kTrRevert,  // 1 when the tried step is rolled back
    kTrSlots
};

struct TrConstants
{
    double dt2, tau2, eta, small, fltMin, fltMax;
};

// f = 1/2 sum m |x - x~|^2 + h^2 sum w Psi from the reduced sums (K, U, .)
__device__ inline double TrObjective(double const* sums, double dt2) { return 0.5 * sums[0] + dt2 * sums[1]; }

// Start of a substep: f_0 at the initial iterate, empty history, R = 0.
__device__ inline void TrBegin(double* __restrict__ s, double const* __restrict__ sums, TrConstants const& c)
{
    s[kTrF] = TrObjective(sums, c.dt2);
    s[kTrT] = -1.;
    s[kTrFm1] = s[kTrFm2] = s[kTrTm1] = s[kTrTm2] = s[kTrR2] = 0.;
}

// After a plain iteration k < 2: the history shifts, then f_k, t_k = k.
__device__ inline void TrPlain(double* __restrict__ s, double const* __restrict__ sums, TrConstants const& c, int k)
{
    s[kTrFm2] = s[kTrFm1], s[kTrFm1] = s[kTrF];
    s[kTrTm2] = s[kTrTm1], s[kTrTm1] = s[kTrT];
    s[kTrF] = TrObjective(sums, c.dt2);
    s[kTrT] = static_cast<double>(k);
}

// Before the sweep of iteration k >= 2 (ConstructModel and UpdateIterates): times
// shifted so that t_k = 0, the quadratic through the last three (t, f) by
// Cramer's rule, then the history shifts.
[[maybe_unused]] static __global__ void TrModel(double* __restrict__ s)
{
    if (threadIdx.x != 0)
        return;
    double const tk = s[kTrT];
    double const tkm2 = s[kTrTm2] - tk, tkm1 = s[kTrTm1] - tk;
    double const fk = s[kTrF], fkm1 = s[kTrFm1], fkm2 = s[kTrFm2];
    double const det = tkm2 * tkm2 * tkm1 - tkm1 * tkm1 * tkm2;
    double const r2 = fkm2 - fk, r1 = fkm1 - fk;
    s[kTrQa] = (r2 * tkm1 - r1 * tkm2) / det;
    s[kTrQb] = (tkm2 * tkm2 * r1 - tkm1 * tkm1 * r2) / det;
    s[kTrQc] = fk;
    s[kTrFm2] = fkm1, s[kTrFm1] = fk;
    s[kTrTm2] = tkm1, s[kTrTm1] = 0.;
    s[kTrT] = 0.;
}

// After the sweep (sums at the VBD iterate, sums[2] = |dx|^2): the radius, the
// model's optimal step clamped to [1, R/|dx|] (ModelOptimalStep), and whether
// it is tried.
__device__ inline void TrDecide(double* __restrict__ s, double const* __restrict__ sums, TrConstants const& c)
{
    double const fVbd = TrObjective(sums, c.dt2), dx2 = sums[2];
    double R2 = s[kTrR2];
    if (R2 < dx2 + c.small)
        R2 = c.tau2 * dx2;
    double const lower = 1., upper = sqrt(R2 / dx2);
    double const qa = s[kTrQa], qb = s[kTrQb];
    double tstar = 0.;
    if (fabs(qa) > c.fltMin)
        tstar = qa > c.fltMin ? -qb / (2. * qa) : c.fltMax;
    else if (fabs(qb) > c.fltMin)
        tstar = qb > 0. ? -c.fltMax : c.fltMax;
    double const t = tstar < lower ? lower : (upper < tstar ? upper : tstar);
    s[kTrR2] = R2, s[kTrFVbd] = fVbd, s[kTrUpper] = upper, s[kTrStep] = t;
    s[kTrAttempt] = (fabs(t - lower) < c.small) ? 0. : 1.;
}

// After the tried step (sums at x(t), read only when a step was tried): keep it
// when the actual over the predicted decrease exceeds eta, growing R at the
// bound; otherwise shrink R and roll back (or keep the VBD iterate when nothing
// was tried). Logs (t, accepted) at slot.
__device__ inline void TrAccept(double* __restrict__ s, double const* __restrict__ sums, TrConstants const& c,
                                double* __restrict__ logT, int* __restrict__ logAccepted, int slot)
{
    bool const attempt = s[kTrAttempt] != 0.;
    double const t = s[kTrStep];
    bool accepted = false;
    if (attempt)
    {
        double const fk = TrObjective(sums, c.dt2);
        s[kTrF] = fk;
        double const mNext = s[kTrQa] * t * t + s[kTrQb] * t + s[kTrQc];
        double const fkm1 = s[kTrFm1];
        accepted = (fkm1 - fk) / (fkm1 - mNext) > c.eta;
    }
    if (accepted)
    {
        if ((s[kTrUpper] - t) < c.small)
            s[kTrR2] *= c.tau2;
        s[kTrT] = t;
        s[kTrRevert] = 0.;
    }
    else
    {
        s[kTrR2] /= c.tau2;
        s[kTrRevert] = attempt ? 1. : 0.;
        if (!attempt)
            s[kTrF] = s[kTrFVbd];
        s[kTrT] = s[kTrTm1] + 1.;
    }
    logT[slot] = t;
    logAccepted[slot] = accepted ? 1 : 0;
}

// After a rollback: f_k at the restored VBD iterate.
__device__ inline void TrRevert(double* __restrict__ s, double const* __restrict__ sums, TrConstants const& c)
{
    if (s[kTrRevert] != 0.)
        s[kTrF] = TrObjective(sums, c.dt2);
}

// Trust-region update that runs after a reduction, on its sums (SumPartials3).
enum TrOp : int
{
    kTrOpNone = 0,
    kTrOpBegin,
    kTrOpPlain,
    kTrOpDecide,
    kTrOpAccept,
    kTrOpRevert
};

struct TrEpilogue
{
    int op;          // TrOp
    double* s;       // trust-region state
    TrConstants c;
    double* logT;    // kTrOpAccept: decision log
    int* logAccepted;
    int slot;        // kTrOpAccept: log slot
    int k;           // kTrOpPlain: iteration
};

// out[j] = sum over b < nb of part[3 b + j], in a fixed order. One workgroup of 256.
// The sums are skipped when gate points at 0; the epilogue's trust-region update
// runs either way (it reads the sums only when they were formed).
[[maybe_unused]] static __global__ __launch_bounds__(256) void SumPartials3(double const* __restrict__ part, int nb,
                                                                           double* __restrict__ out,
                                                                           double const* __restrict__ gate,
                                                                           TrEpilogue e)
{
    if (!(gate && *gate == 0.))
    {
        double a = 0., b = 0., c = 0.;
        for (int k = static_cast<int>(threadIdx.x); k < nb; k += static_cast<int>(blockDim.x))
        {
            a += part[3 * k];
            b += part[3 * k + 1];
            c += part[3 * k + 2];
        }
        BlockSum3(a, b, c);
        if (threadIdx.x == 0)
        {
            out[0] = a;
            out[1] = b;
            out[2] = c;
        }
    }
    if (threadIdx.x != 0)
        return;
    switch (e.op)
    {
    case kTrOpBegin:
        TrBegin(e.s, out, e.c);
        break;
    case kTrOpPlain:
        TrPlain(e.s, out, e.c, e.k);
        break;
    case kTrOpDecide:
        TrDecide(e.s, out, e.c);
        break;
    case kTrOpAccept:
        TrAccept(e.s, out, e.c, e.logT, e.logAccepted, e.slot);
        break;
    case kTrOpRevert:
        TrRevert(e.s, out, e.c);
        break;
    default:
        break;
    }
}

// Trust-region step along the line through x_k (xkm1) and the VBD iterate
// (gpu/impl/vbd/TrustRegionIntegrator.cu:491-522): x = x_k + t (x - x_k), or, with
// rollback, x = x_k + (x - x_k) / t, t from the device state; nothing happens
// unless the state says the step is tried (rolled back). Masses in xm.w are kept.
[[maybe_unused]] static __global__ void LinearPathStep(float4* __restrict__ xm, float4 const* __restrict__ xkm1, int n,
                                                       double const* __restrict__ s, int rollback)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || s[rollback ? kTrRevert : kTrAttempt] == 0.)
        return;
    float const t = static_cast<float>(s[kTrStep]);
    float4 const x = xm[i], p = xkm1[i];
    float4 r;
    if (rollback)
        r = make_float4(p.x + (x.x - p.x) / t, p.y + (x.y - p.y) / t, p.z + (x.z - p.z) / t, x.w);
    else
        r = make_float4(p.x + t * (x.x - p.x), p.y + t * (x.y - p.y), p.z + t * (x.z - p.z), x.w);
    xm[i] = r;
}

[[maybe_unused]] static __global__ void ZeroDoubles(double* p, int n)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        p[i] = 0.;
}

} // namespace pbat_vbd::device
