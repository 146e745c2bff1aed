// Vertex-triangle contact terms of the VBD sweep (gpu/impl/vbd/Kernels.cuh:80-114,
// 204-223; sim/vbd/Kernels.h:215-302) and the contact-sweep write buffer. The
// active-set detection lives in vbd_contact.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace pbat_vbd::device
{

constexpr int kMaxContacts = 8;     // contacting triangles per vertex (Kernels.cuh:64-65)
constexpr float kFltEpsilon = 1.1920928955078125e-07f; // FLT_EPSILON

__device__ inline float3 Sub(float4 a, float4 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float Dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 Cross(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Contact terms of vertex v in the sweep (ContactPenalty, Kernels.cuh:80-114, and
// AccumulateVertexTriangleContact, sim/vbd/Kernels.h:215-302): the penalty muC is
// scaled by the vertex area over the summed areas of its contacting triangles,
// times each triangle's area; the normal penalty 1/2 muC' d^2 on the penetration
// d = min(0, (x - x_b).n) below a triangle the vertex projects into, and lagged
// IPC friction on the substep's relative tangential displacement. Triangle
// positions are read from xm, which holds the positions from before the colour
// pass (the contact sweep writes to a separate buffer).
struct ContactSweep
{
    // kMaxContacts records per vertex, one per contact slot: the triangle's
    // vertices and, as bits in .w, its scaled penalty; x = -1 for an empty slot
    int4 const* __restrict__ rec;
    float4* __restrict__ xb; // new positions of the colour
    float muF, epsv;
    // The previous colour's range, whose current positions are still in xb: the
    // sweep reads them there and copies them back to xm (nobody reads xm there).
    int pendBeg, pendEnd;
};

// Position of vertex j during a contact sweep: in xb when j belongs to the
// previous colour's pending range, else in xm.
__device__ __forceinline__ float4 ContactPosition(ContactSweep const& cs, float4 const* xm, int j)
{
    return (j >= cs.pendBeg && j < cs.pendEnd) ? cs.xb[j] : xm[j];
}

template <class TBlock>
__device__ inline void AccumulateContact(float3 xtv, float3 xv, ContactSweep const& cs, float4 const* xm,
                                         float4 const* xt, int4 t, float dt, float muC, float muF, float epsv,
                                         TBlock& B)
{
    float4 const A4 = ContactPosition(cs, xm, t.x), B4 = ContactPosition(cs, xm, t.y),
                 C4 = ContactPosition(cs, xm, t.z);
    float3 const A = make_float3(A4.x, A4.y, A4.z);
    float3 const e0 = make_float3(B4.x - A4.x, B4.y - A4.y, B4.z - A4.z);
    float3 const e1 = make_float3(C4.x - A4.x, C4.y - A4.y, C4.z - A4.z);
    float3 n = Cross(e0, e1);
    float const dbl = sqrtf(Dot(n, n));
    if (dbl <= 1e-8f)
        return;
    n = make_float3(n.x / dbl, n.y / dbl, n.z / dbl);
    float3 ap = make_float3(xv.x - A.x, xv.y - A.y, xv.z - A.z);
    float const tt = Dot(n, ap);
    ap = make_float3(ap.x - tt * n.x, ap.y - tt * n.y, ap.z - tt * n.z);
    float const d00 = Dot(e0, e0), d01 = Dot(e0, e1), d11 = Dot(e1, e1);
    float const d20 = Dot(ap, e0), d21 = Dot(ap, e1);
    float const den = d00 * d11 - d01 * d01;
    float const bv = (d11 * d20 - d01 * d21) / den, bw = (d00 * d21 - d01 * d20) / den;
    float const bu = 1.f - bv - bw;
    if (!(bu >= 0.f && bu <= 1.f && bv >= 0.f && bv <= 1.f && bw >= 0.f && bw <= 1.f))
        return;
    float3 const xb = make_float3(bu * A4.x + bv * B4.x + bw * C4.x, bu * A4.y + bv * B4.y + bw * C4.y,
                                  bu * A4.z + bv * B4.z + bw * C4.z);
    float const dpen = fminf(0.f, (xv.x - xb.x) * n.x + (xv.y - xb.y) * n.y + (xv.z - xb.z) * n.z);
    float const lambda = muC * dpen;
    B.g0 += lambda * n.x, B.g1 += lambda * n.y, B.g2 += lambda * n.z;
    B.h00 += muC * n.x * n.x, B.h01 += muC * n.x * n.y, B.h02 += muC * n.x * n.z;
    B.h11 += muC * n.y * n.y, B.h12 += muC * n.y * n.z, B.h22 += muC * n.z * n.z;
    // friction in the tangent frame (e0, n x e0)
    float4 const tA = xt[t.x], tB = xt[t.y], tC = xt[t.z];
    float3 const xtb = make_float3(bu * tA.x + bv * tB.x + bw * tC.x, bu * tA.y + bv * tB.y + bw * tC.y,
                                   bu * tA.z + bv * tB.z + bw * tC.z);
    float3 const b1 = Cross(n, e0);
    float3 const du = make_float3((xv.x - xtv.x) - (xb.x - xtb.x), (xv.y - xtv.y) - (xb.y - xtb.y),
                                  (xv.z - xtv.z) - (xb.z - xtb.z));
    float const u0 = Dot(e0, du), u1 = Dot(b1, du);
    float const unorm = sqrtf(u0 * u0 + u1 * u1) + kFltEpsilon;
    float const epsvh = epsv * dt;
    float const y = unorm / epsvh;
    float const f1 = (y < 1.f) ? 2.f * y - y * y : 1.f;
    float const c1 = muF * fabsf(lambda) * f1 / unorm;
    B.g0 += c1 * (e0.x * u0 + b1.x * u1);
    B.g1 += c1 * (e0.y * u0 + b1.y * u1);
    B.g2 += c1 * (e0.z * u0 + b1.z * u1);
    B.h00 += c1 * (e0.x * e0.x + b1.x * b1.x), B.h01 += c1 * (e0.x * e0.y + b1.x * b1.y);
    B.h02 += c1 * (e0.x * e0.z + b1.x * b1.z), B.h11 += c1 * (e0.y * e0.y + b1.y * b1.y);
    B.h12 += c1 * (e0.y * e0.z + b1.y * b1.z), B.h22 += c1 * (e0.z * e0.z + b1.z * b1.z);
}

// xm[v] = xb[v] over a colour's range, after the last contact sweep of an iteration.
[[maybe_unused]] static __global__ __launch_bounds__(256) void CopyColour(float4* __restrict__ xm,
                                                                          float4 const* __restrict__ xb, int vbeg,
                                                                          int vend)
{
    int const v = vbeg + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (v < vend)
        xm[v] = xb[v];
}

} // namespace pbat_vbd::device
