// Device kernels of the MI355X VBD integrator (gfx950, wave64).
//
// Data layout in HBM (internal vertex numbering = colour-major, see
// vbd_integrator.hip):
//   xm[v]    float4 (x, y, z, m)      position and lumped mass
//   xtil[v]  float4 (x~, 0)           inertial target
//   xt[v]    float4 (x_t, 0)          positions at the start of the substep
//   vel[v], vprev[v], aext[v]  float4 velocities and external acceleration
//   vtp[v]   int32 CSR prefix into vtref; vtref[k] = (tet << 2) | ilocal
//   tets[e]  TetRecord, 64 B = one aligned half cache line:
//            4 vertex ids, GP rows 1..3 (row 0 = -(row1+row2+row3)),
//            w*mu, w*lambda, alpha = 1 + mu/lambda.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pbat_vbd::device {

struct alignas(64) TetRecord
{
    int4 v;    // vertex indices
    float4 q0; // GP(1,0) GP(1,1) GP(1,2) GP(2,0)
    float4 q1; // GP(2,1) GP(2,2) GP(3,0) GP(3,1)
    float4 q2; // GP(3,2) w*mu    w*lam   alpha
};
static_assert(sizeof(TetRecord) == 64, "TetRecord must be 64 bytes");

struct SweepArgs
{
    int const* __restrict__ vtp;
    int const* __restrict__ vtref;
    TetRecord const* __restrict__ tets;
    float4* __restrict__ xm;
    float4 const* __restrict__ xtil;
    float4 const* __restrict__ xt;
    int vbeg;
    int vend;
    float dt;       // substep h
    float dt2;      // h^2
    float kD;       // Rayleigh damping coefficient
    float detHZero; // skip threshold on |det H|
};

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

// Elastic contribution of one incident tet to vertex i's 3x3 block.
// F = Ds Dm^{-1} with Ds = [x1-x0, x2-x0, x3-x0]; the own row g of GP gives
//   grad_i  += w mu F g + w lambda (J - alpha) c,   c = cof(F) g
//   Hess_i  += w mu |g|^2 I + w lambda c c^T
// (the d2J/dF2 term contracts to zero because det(F0 + x_i g^T) is affine in x_i).
struct Block
{
    float h00, h01, h02, h11, h12, h22;
    float g0, g1, g2;
};

__device__ __forceinline__ void AccumulateTet(TetRecord const& T, int il, float4 const& x0, float4 const& x1,
                                              float4 const& x2, float4 const& x3, Block& B)
{
    // rows of GP: r1, r2, r3 stored; r0 = -(r1 + r2 + r3)
    float const r1x = T.q0.x, r1y = T.q0.y, r1z = T.q0.z;
    float const r2x = T.q0.w, r2y = T.q1.x, r2z = T.q1.y;
    float const r3x = T.q1.z, r3y = T.q1.w, r3z = T.q2.x;
    float const wmu = T.q2.y, wlam = T.q2.z, alpha = T.q2.w;
    // edge vectors
    float const d1x = x1.x - x0.x, d1y = x1.y - x0.y, d1z = x1.z - x0.z;
    float const d2x = x2.x - x0.x, d2y = x2.y - x0.y, d2z = x2.z - x0.z;
    float const d3x = x3.x - x0.x, d3y = x3.y - x0.y, d3z = x3.z - x0.z;
    // columns of F
    float const f0x = d1x * r1x + d2x * r2x + d3x * r3x;
    float const f0y = d1y * r1x + d2y * r2x + d3y * r3x;
    float const f0z = d1z * r1x + d2z * r2x + d3z * r3x;
    float const f1x = d1x * r1y + d2x * r2y + d3x * r3y;
    float const f1y = d1y * r1y + d2y * r2y + d3y * r3y;
    float const f1z = d1z * r1y + d2z * r2y + d3z * r3y;
    float const f2x = d1x * r1z + d2x * r2z + d3x * r3z;
    float const f2y = d1y * r1z + d2y * r2z + d3y * r3z;
    float const f2z = d1z * r1z + d2z * r2z + d3z * r3z;
    // cofactor columns
    float const ax = f1y * f2z - f1z * f2y, ay = f1z * f2x - f1x * f2z, az = f1x * f2y - f1y * f2x; // f1 x f2
    float const bx = f2y * f0z - f2z * f0y, by = f2z * f0x - f2x * f0z, bz = f2x * f0y - f2y * f0x; // f2 x f0
    float const cx = f0y * f1z - f0z * f1y, cy = f0z * f1x - f0x * f1z, cz = f0x * f1y - f0y * f1x; // f0 x f1
    float const J = f0x * ax + f0y * ay + f0z * az;
    // own GP row
    float gx, gy, gz;
    if (il == 0)
    {
        gx = -(r1x + r2x + r3x);
        gy = -(r1y + r2y + r3y);
        gz = -(r1z + r2z + r3z);
    }
    else if (il == 1)
    {
        gx = r1x, gy = r1y, gz = r1z;
    }
    else if (il == 2)
    {
        gx = r2x, gy = r2y, gz = r2z;
    }
    else
    {
        gx = r3x, gy = r3y, gz = r3z;
    }
    float const kx = gx * ax + gy * bx + gz * cx;
    float const ky = gx * ay + gy * by + gz * cy;
    float const kz = gx * az + gy * bz + gz * cz;
    float const Fgx = f0x * gx + f1x * gy + f2x * gz;
    float const Fgy = f0y * gx + f1y * gy + f2y * gz;
    float const Fgz = f0z * gx + f1z * gy + f2z * gz;
    float const s = wlam * (J - alpha);
    B.g0 += wmu * Fgx + s * kx;
    B.g1 += wmu * Fgy + s * ky;
    B.g2 += wmu * Fgz + s * kz;
    float const dg = wmu * (gx * gx + gy * gy + gz * gz);
    float const lkx = wlam * kx, lky = wlam * ky, lkz = wlam * kz;
    B.h00 += dg + lkx * kx;
    B.h11 += dg + lky * ky;
    B.h22 += dg + lkz * kz;
    B.h01 += lkx * ky;
    B.h02 += lkx * kz;
    B.h12 += lky * kz;
}

// Damping, inertia and the 3x3 Newton step of one vertex
// (sim/vbd/Kernels.h:173-191, 304-339). Returns the displacement to subtract.
__device__ __forceinline__ float3 NewtonStep(Block B, float4 const& xv, float4 const& xl, float4 const& xtv,
                                             SweepArgs const& a)
{
    if (a.kD != 0.f)
    {
        float const D = a.kD / a.dt;
        float const ux = xv.x - xtv.x, uy = xv.y - xtv.y, uz = xv.z - xtv.z;
        B.g0 += D * (B.h00 * ux + B.h01 * uy + B.h02 * uz);
        B.g1 += D * (B.h01 * ux + B.h11 * uy + B.h12 * uz);
        B.g2 += D * (B.h02 * ux + B.h12 * uy + B.h22 * uz);
        float const s = 1.f + D;
        B.h00 *= s, B.h01 *= s, B.h02 *= s, B.h11 *= s, B.h12 *= s, B.h22 *= s;
    }
    float const K = xv.w / a.dt2;
    B.h00 += K, B.h11 += K, B.h22 += K;
    B.g0 += K * (xv.x - xl.x);
    B.g1 += K * (xv.y - xl.y);
    B.g2 += K * (xv.z - xl.z);
    // adjugate of the symmetric block
    float const a00 = B.h11 * B.h22 - B.h12 * B.h12;
    float const a01 = B.h02 * B.h12 - B.h01 * B.h22;
    float const a02 = B.h01 * B.h12 - B.h02 * B.h11;
    float const det = B.h00 * a00 + B.h01 * a01 + B.h02 * a02;
    if (!(fabsf(det) > a.detHZero))
        return make_float3(0.f, 0.f, 0.f);
    float const dscale = fabsf(B.h00 * B.h11 * B.h22);
    if (fabsf(det) < 1e-4f * dscale)
    {
        // ill-conditioned block: solve in FP64 from the FP32 entries
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

// One colour of one VBD sweep. L lanes cooperate on each vertex: lane l of
// the group handles incidences l, l+L, ... of the vertex's CSR row, the group
// reduces its 9 partial sums with DPP, and the group's first lane applies the
// Newton update in place (same-colour vertices share no tet).
template <int L>
__global__ __launch_bounds__(256) void VbdSweep(SweepArgs a)
{
    int const tid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    int const lane = tid & (L - 1);
    int const v = a.vbeg + tid / L;
    bool const active = v < a.vend;
    int kb = 0, ke = 0;
    if (active)
    {
        kb = a.vtp[v];
        ke = a.vtp[v + 1];
    }
    Block B{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = kb + lane; k < ke; k += L)
    {
        int const ref = a.vtref[k];
        TetRecord const T = a.tets[ref >> 2];
        float4 const x0 = a.xm[T.v.x];
        float4 const x1 = a.xm[T.v.y];
        float4 const x2 = a.xm[T.v.z];
        float4 const x3 = a.xm[T.v.w];
        AccumulateTet(T, ref & 3, x0, x1, x2, x3, B);
    }
    B.h00 = GroupSum<L>(B.h00);
    B.h01 = GroupSum<L>(B.h01);
    B.h02 = GroupSum<L>(B.h02);
    B.h11 = GroupSum<L>(B.h11);
    B.h12 = GroupSum<L>(B.h12);
    B.h22 = GroupSum<L>(B.h22);
    B.g0 = GroupSum<L>(B.g0);
    B.g1 = GroupSum<L>(B.g1);
    B.g2 = GroupSum<L>(B.g2);
    if (!active || lane != 0)
        return;
    float4 const xv = a.xm[v];
    float4 const xl = a.xtil[v];
    float4 const xtv = (a.kD != 0.f) ? a.xt[v] : xv;
    float3 const dx = NewtonStep(B, xv, xl, xtv, a);
    a.xm[v] = make_float4(xv.x - dx.x, xv.y - dx.y, xv.z - dx.z, xv.w);
}

// Substep begin (gpu/impl/vbd/Integrator.cu:190-248; sim/vbd/Kernels.h:24-94):
// inertial target x~ = xt + h v + h^2 a and the initial iterate.
__global__ __launch_bounds__(256) void SubstepBegin(float4* __restrict__ xm, float4* __restrict__ xtil,
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
__global__ __launch_bounds__(256) void SubstepEnd(float4 const* __restrict__ xm, float4* __restrict__ xt,
                                                  float4* __restrict__ vel, float4* __restrict__ vprev, int n,
                                                  float dt)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const x = xm[i];
    float4 const p = xt[i];
    vprev[i] = vel[i];
    vel[i] = make_float4((x.x - p.x) / dt, (x.y - p.y) / dt, (x.z - p.z) / dt, 0.f);
    xt[i] = make_float4(x.x, x.y, x.z, 0.f);
}

// FP64 energy reduction: out[0] += 1/2 m |v|^2, out[2] += -m a.x over vertices,
// out[1] += w Psi(F) over tets. One atomic per block and term.
__global__ __launch_bounds__(256) void EnergyTerms(float4 const* __restrict__ xm, float4 const* __restrict__ vel,
                                                   float4 const* __restrict__ aext,
                                                   TetRecord const* __restrict__ tets, int nV, int nT,
                                                   double* __restrict__ out)
{
    __shared__ double red[3][256 / 64];
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
    {
        TetRecord const T = tets[e];
        float4 const p0 = xm[T.v.x], p1 = xm[T.v.y], p2 = xm[T.v.z], p3 = xm[T.v.w];
        double const r[3][3] = {{T.q0.x, T.q0.y, T.q0.z}, {T.q0.w, T.q1.x, T.q1.y}, {T.q1.z, T.q1.w, T.q2.x}};
        double const d[3][3] = {{double(p1.x) - p0.x, double(p1.y) - p0.y, double(p1.z) - p0.z},
                                {double(p2.x) - p0.x, double(p2.y) - p0.y, double(p2.z) - p0.z},
                                {double(p3.x) - p0.x, double(p3.y) - p0.y, double(p3.z) - p0.z}};
        double F[3][3]; // F[row][col]
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
        double const q = J - double(T.q2.w);
        U += 0.5 * double(T.q2.y) * (I2 - 3.) + 0.5 * double(T.q2.z) * q * q;
    }
    // wave reduction then block reduction
    for (int off = 32; off > 0; off >>= 1)
    {
        K += __shfl_xor(K, off);
        U += __shfl_xor(U, off);
        P += __shfl_xor(P, off);
    }
    int const w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
    {
        red[0][w] = K;
        red[1][w] = U;
        red[2][w] = P;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s0 = 0., s1 = 0., s2 = 0.;
        for (int k = 0; k < static_cast<int>(blockDim.x / 64); ++k)
        {
            s0 += red[0][k];
            s1 += red[1][k];
            s2 += red[2][k];
        }
        atomicAdd(out + 0, s0);
        atomicAdd(out + 1, s1);
        atomicAdd(out + 2, s2);
    }
}

__global__ void ZeroDoubles(double* p, int n)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        p[i] = 0.;
}

} // namespace pbat_vbd::device
