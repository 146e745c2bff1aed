// Vertex-triangle contact of the GPU VBD integrator: active-set detection
// (gpu/impl/contact/VertexTriangleMixedCcdDcd.cu) and the per-vertex contact
// terms of the sweep (gpu/impl/vbd/Kernels.cuh:80-114, 204-223;
// sim/vbd/Kernels.h:215-302).
//
// Detection is an all-pairs search over chunks of 64 consecutive triangles,
// each with a bounding box: a wave takes kPerWave collision vertices, skips the
// chunks whose box lies beyond all of their query bounds, and for the others
// gives one triangle to each lane. It is exhaustive up to those exact bounds,
// which makes its answers independent of any traversal order: a vertex's search
// radius is the largest over its overlapping triangles, and its nearest
// triangles come in triangle order.
#pragma once

#include <hip/hip_runtime.h>

namespace pbat_vbd::device
{

constexpr int kMaxContacts = 8;     // contacting triangles per vertex (Kernels.cuh:64-65)
constexpr int kContactBlock = 256;  // workgroup size of the detection kernels (4 waves)
constexpr int kChunk = 64;          // triangles per chunk box = lanes per wave
constexpr int kPerWave = 4;         // collision vertices per wave in the detection kernels
constexpr int kGroup = 4 * kPerWave; // collision vertices per workgroup
constexpr float kFltEpsilon = 1.1920928955078125e-07f; // FLT_EPSILON

struct ContactArgs
{
    int const* __restrict__ cvIdx;     // internal vertex of each collision vertex
    int const* __restrict__ cvBody;    // body of each collision vertex
    int4 const* __restrict__ tri;      // internal vertices of each triangle, body of its first vertex in .w
    float const* __restrict__ triArea; // triangle areas
    float const* __restrict__ xva;     // vertex areas (internal order)
    int* __restrict__ fc;              // kMaxContacts contacting triangles per internal vertex, -1 padded
    int* __restrict__ active;          // per collision vertex, kept across steps
    int* __restrict__ inAv;            // per collision vertex: in this step's active list
    float* __restrict__ dupper;        // per collision vertex: squared search radius
    float4* __restrict__ triLo;        // swept triangle boxes
    float4* __restrict__ triHi;
    float4* __restrict__ chunkLo;      // box of each chunk of kChunk consecutive triangles
    float4* __restrict__ chunkHi;
    int nCV, nF;
};

__device__ inline float3 Sub(float4 a, float4 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float Dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 Cross(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// x1 = x0 + h v + h^2 a over all vertices (gpu/impl/vbd/Integrator.cu:163-188).
[[maybe_unused]] static __global__ __launch_bounds__(256) void ContactFullStep(
    float4 const* __restrict__ xt, float4 const* __restrict__ vel, float4 const* __restrict__ aext, int n, float dt,
    float dt2, float4* __restrict__ x1)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n)
        return;
    float4 const p = xt[i], u = vel[i], g = aext[i];
    x1[i] = make_float4(p.x + dt * u.x + dt2 * g.x, p.y + dt * u.y + dt2 * g.y, p.z + dt * u.z + dt2 * g.z, 0.f);
}

// Boxes of the triangles swept from x0 to x1 (VertexTriangleMixedCcdDcd.cu:87-101).
[[maybe_unused]] static __global__ __launch_bounds__(256) void ContactSweptBoxes(ContactArgs c,
                                                                                 float4 const* __restrict__ x0,
                                                                                 float4 const* __restrict__ x1)
{
    int const f = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (f >= c.nF)
        return;
    int4 const t = c.tri[f];
    float4 const p[6] = {x0[t.x], x0[t.y], x0[t.z], x1[t.x], x1[t.y], x1[t.z]};
    float lx = p[0].x, ly = p[0].y, lz = p[0].z, hx = lx, hy = ly, hz = lz;
    for (int k = 1; k < 6; ++k)
    {
        lx = fminf(lx, p[k].x), ly = fminf(ly, p[k].y), lz = fminf(lz, p[k].z);
        hx = fmaxf(hx, p[k].x), hy = fmaxf(hy, p[k].y), hz = fmaxf(hz, p[k].z);
    }
    c.triLo[f] = make_float4(lx, ly, lz, __int_as_float(t.w));
    c.triHi[f] = make_float4(hx, hy, hz, 0.f);
}

// Box of each chunk of kChunk consecutive triangles: of their swept boxes
// (swept != 0) or of their vertices at x. One wave per chunk.
[[maybe_unused]] static __global__ __launch_bounds__(kContactBlock) void ContactChunkBoxes(
    ContactArgs c, float4 const* __restrict__ x, int swept)
{
    int const chunk = static_cast<int>(blockIdx.x) * (kContactBlock / kChunk) + static_cast<int>(threadIdx.x >> 6);
    int const f = chunk * kChunk + static_cast<int>(threadIdx.x & 63);
    float lx = 3.402823466e+38f, ly = lx, lz = lx, hx = -lx, hy = -lx, hz = -lx;
    if (f < c.nF)
    {
        if (swept)
        {
            float4 const L = c.triLo[f], U = c.triHi[f];
            lx = L.x, ly = L.y, lz = L.z, hx = U.x, hy = U.y, hz = U.z;
        }
        else
        {
            int4 const t = c.tri[f];
            float4 const a = x[t.x], b = x[t.y], cc = x[t.z];
            lx = fminf(a.x, fminf(b.x, cc.x)), ly = fminf(a.y, fminf(b.y, cc.y)), lz = fminf(a.z, fminf(b.z, cc.z));
            hx = fmaxf(a.x, fmaxf(b.x, cc.x)), hy = fmaxf(a.y, fmaxf(b.y, cc.y)), hz = fmaxf(a.z, fmaxf(b.z, cc.z));
        }
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        lx = fminf(lx, __shfl_xor(lx, off)), ly = fminf(ly, __shfl_xor(ly, off));
        lz = fminf(lz, __shfl_xor(lz, off)), hx = fmaxf(hx, __shfl_xor(hx, off));
        hy = fmaxf(hy, __shfl_xor(hy, off)), hz = fmaxf(hz, __shfl_xor(hz, off));
    }
    if ((threadIdx.x & 63) == 0 && chunk * kChunk < c.nF)
    {
        c.chunkLo[chunk] = make_float4(lx, ly, lz, 0.f);
        c.chunkHi[chunk] = make_float4(hx, hy, hz, 0.f);
    }
}

// Squared distance from p to the box [L, U] (0 inside).
__device__ inline float PointBoxSqDist(float4 p, float4 L, float4 U)
{
    float const dx = p.x - fminf(fmaxf(p.x, L.x), U.x);
    float const dy = p.y - fminf(fmaxf(p.y, L.y), U.y);
    float const dz = p.z - fminf(fmaxf(p.z, L.z), U.z);
    return dx * dx + dy * dy + dz * dz;
}

// A tile can be skipped for a query when its box lies beyond the query's bound;
// the margin keeps the test conservative against the rounding of the exact
// per-triangle distances.
__device__ inline bool TileWithin(float dbox, float bound)
{
    return dbox <= bound + 1e-5f * (bound + dbox) + 1e-30f;
}

__device__ inline float WaveMax(float x)
{
    for (int off = 32; off > 0; off >>= 1)
        x = fmaxf(x, __shfl_xor(x, off));
    return x;
}

__device__ inline float WaveMin(float x)
{
    for (int off = 32; off > 0; off >>= 1)
        x = fminf(x, __shfl_xor(x, off));
    return x;
}

__device__ inline int WaveMaxInt(int x)
{
    for (int off = 32; off > 0; off >>= 1)
        x = max(x, __shfl_xor(x, off));
    return x;
}

// Broad phase of InitializeActiveSet (VertexTriangleMixedCcdDcd.cu:53-142): a
// collision vertex whose segment x0 -> x1 has a box that overlaps (closed boxes)
// the swept box of a triangle of another body becomes active, with search radius
// the squared diagonal of the union of the two boxes, largest over such
// triangles. Vertices active from earlier steps stay active. inAv receives the
// step's active list. kPerWave collision vertices per wave.
[[maybe_unused]] static __global__ __launch_bounds__(kContactBlock) void ContactBroadphase(
    ContactArgs c, float4 const* __restrict__ x0, float4 const* __restrict__ x1)
{
    int const wave = static_cast<int>(threadIdx.x) >> 6, lane = static_cast<int>(threadIdx.x) & 63;
    int const v0 = static_cast<int>(blockIdx.x) * kGroup + wave * kPerWave;
    if (v0 >= c.nCV)
        return;
    float3 lo[kPerWave], hi[kPerWave];
    int body[kPerWave];
    bool hit[kPerWave];
    float d2[kPerWave];
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
    {
        int const v = v0 + j;
        lo[j] = hi[j] = make_float3(0.f, 0.f, 0.f);
        body[j] = 0;
        hit[j] = false;
        d2[j] = 0.f;
        if (v < c.nCV)
        {
            int const i = c.cvIdx[v];
            float4 const a = x0[i], b = x1[i];
            lo[j] = make_float3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z));
            hi[j] = make_float3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
            body[j] = c.cvBody[v];
        }
    }
    int const nChunks = (c.nF + kChunk - 1) / kChunk;
    // 64 chunk boxes at a time, one per lane: the lanes test them against the
    // wave's segments, and the wave visits the chunks that any segment overlaps
    for (int cb = 0; cb < nChunks; cb += 64)
    {
        int bits = 0;
        if (cb + lane < nChunks)
        {
            float4 const TL = c.chunkLo[cb + lane], TU = c.chunkHi[cb + lane];
#pragma unroll
            for (int j = 0; j < kPerWave; ++j)
                if (v0 + j < c.nCV && lo[j].x <= TU.x && lo[j].y <= TU.y && lo[j].z <= TU.z && TL.x <= hi[j].x &&
                    TL.y <= hi[j].y && TL.z <= hi[j].z)
                    bits |= 1 << j;
        }
        unsigned long long todo = __ballot(bits != 0);
        while (todo)
        {
            int const k = __ffsll(static_cast<long long>(todo)) - 1;
            todo &= todo - 1ull;
            int const need = __shfl(bits, k);
            int const f = (cb + k) * kChunk + lane;
            if (f >= c.nF)
                continue;
            float4 const L = c.triLo[f], U = c.triHi[f];
#pragma unroll
            for (int j = 0; j < kPerWave; ++j)
            {
                if (!((need >> j) & 1) || __float_as_int(L.w) == body[j])
                    continue;
                bool const overlap = lo[j].x <= U.x && lo[j].y <= U.y && lo[j].z <= U.z && L.x <= hi[j].x &&
                                     L.y <= hi[j].y && L.z <= hi[j].z;
                if (!overlap)
                    continue;
                float const dx = fmaxf(hi[j].x, U.x) - fminf(lo[j].x, L.x);
                float const dy = fmaxf(hi[j].y, U.y) - fminf(lo[j].y, L.y);
                float const dz = fmaxf(hi[j].z, U.z) - fminf(lo[j].z, L.z);
                float const q = dx * dx + dy * dy + dz * dz;
                d2[j] = hit[j] ? fmaxf(d2[j], q) : q;
                hit[j] = true;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
    {
        int const v = v0 + j;
        if (v >= c.nCV)
            continue;
        bool const any = __ballot(hit[j]) != 0ull;
        float const dmax = WaveMax(hit[j] ? d2[j] : -1.f);
        if (lane != 0)
            continue;
        int act = c.active[v];
        if (any)
        {
            act = 1;
            c.active[v] = 1;
            c.dupper[v] = dmax;
        }
        c.inAv[v] = act;
    }
}

// Squared distance from p to triangle (a, b, c): closest point by Voronoi
// regions (Ericson, Real-Time Collision Detection, 5.1.5), as
// geometry/DistanceQueries.h:217-222 evaluates it.
__device__ inline float PointTriangleSqDist(float4 p, float4 a, float4 b, float4 c)
{
    float3 const ab = Sub(b, a), ac = Sub(c, a), ap = Sub(p, a);
    float const d1 = Dot(ab, ap), d2 = Dot(ac, ap);
    float3 const bp = Sub(p, b);
    float const d3 = Dot(ab, bp), d4 = Dot(ac, bp);
    float3 const cp = Sub(p, c);
    float const d5 = Dot(ab, cp), d6 = Dot(ac, cp);
    float const vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    float u, v, w;
    if (d1 <= 0.f && d2 <= 0.f)
        u = 1.f, v = 0.f, w = 0.f;
    else if (d3 >= 0.f && d4 <= d3)
        u = 0.f, v = 1.f, w = 0.f;
    else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f)
        v = d1 / (d1 - d3), u = 1.f - v, w = 0.f;
    else if (d6 >= 0.f && d5 <= d6)
        u = 0.f, v = 0.f, w = 1.f;
    else if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f)
        w = d2 / (d2 - d6), u = 1.f - w, v = 0.f;
    else if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f)
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6)), u = 0.f, v = 1.f - w;
    else
    {
        float const den = 1.f / (va + vb + vc);
        v = vb * den, w = vc * den, u = 1.f - v - w;
    }
    float const qx = p.x - (u * a.x + v * b.x + w * c.x);
    float const qy = p.y - (u * a.y + v * b.y + w * c.y);
    float const qz = p.z - (u * a.z + v * b.z + w * c.z);
    return qx * qx + qy * qy + qz * qz;
}

// Nearest triangles of the step's active vertices (Bvh::NearestNeighbours,
// gpu/impl/geometry/Bvh.cuh:355-466, as VertexTriangleMixedCcdDcd.cu:117-215
// uses it): squared distances to the triangles of other bodies, d* the
// smallest, r the vertex's search radius; with d* < r - eps the triangles
// within d* + eps, with d* within r +- eps those within r + eps, none
// otherwise; in triangle order, at most kMaxContacts. finalize = 0
// (UpdateActiveSet, gpu/impl/vbd/Integrator.cu:250-271): they become the
// vertex's contacts. finalize = 1 (FinalizeActiveSet): the vertex stays active
// when it lies below the plane of the last of them. kPerWave collision vertices
// per wave; two passes over the triangle chunks, the second collecting in
// triangle order with wave ballots.
[[maybe_unused]] static __global__ __launch_bounds__(kContactBlock) void ContactNearest(
    ContactArgs c, float4 const* __restrict__ x, int finalize)
{
    int const wave = static_cast<int>(threadIdx.x) >> 6, lane = static_cast<int>(threadIdx.x) & 63;
    int const v0 = static_cast<int>(blockIdx.x) * kGroup + wave * kPerWave;
    float4 p[kPerWave];
    int body[kPerWave];
    bool mine[kPerWave];
    float r[kPerWave], dmin[kPerWave], thr[kPerWave];
    int count[kPerWave], last[kPerWave];
    bool anyWork = false;
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
    {
        int const v = v0 + j;
        mine[j] = v < c.nCV && c.inAv[v] != 0;
        p[j] = make_float4(0.f, 0.f, 0.f, 0.f);
        body[j] = 0;
        r[j] = 0.f;
        if (mine[j])
        {
            p[j] = x[c.cvIdx[v]];
            body[j] = c.cvBody[v];
            r[j] = c.dupper[v];
        }
        dmin[j] = 3.402823466e+38f;
        thr[j] = -1.f;
        count[j] = 0;
        last[j] = -1;
        anyWork = anyWork || mine[j];
    }
    if (!anyWork)
        return;
    int const nChunks = (c.nF + kChunk - 1) / kChunk;
    for (int pass = 0; pass < 2; ++pass)
    {
        if (pass == 1)
        {
#pragma unroll
            for (int j = 0; j < kPerWave; ++j)
            {
                if (!mine[j])
                    continue;
                float const ds = WaveMin(dmin[j]);
                if (!(ds < 3.402823466e+38f))
                    continue; // no triangle of another body within reach
                if (ds < r[j] - kFltEpsilon)
                    thr[j] = ds + kFltEpsilon;
                else if (ds <= r[j] + kFltEpsilon)
                    thr[j] = r[j] + kFltEpsilon;
            }
        }
        // 64 chunk boxes at a time, one per lane; the wave visits, in order, the
        // chunks whose box lies within some query's bound: r + eps in pass 0 (no
        // triangle farther can matter), the collection threshold in pass 1
        for (int cb = 0; cb < nChunks; cb += 64)
        {
            int bits = 0;
            if (cb + lane < nChunks)
            {
                float4 const TL = c.chunkLo[cb + lane], TU = c.chunkHi[cb + lane];
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                {
                    float const bound = pass == 0 ? r[j] + kFltEpsilon : thr[j];
                    if (mine[j] && (pass == 0 || thr[j] >= 0.f) && TileWithin(PointBoxSqDist(p[j], TL, TU), bound))
                        bits |= 1 << j;
                }
            }
            unsigned long long todo = __ballot(bits != 0);
            while (todo)
            {
                int const k = __ffsll(static_cast<long long>(todo)) - 1;
                todo &= todo - 1ull;
                int const needBits = __shfl(bits, k);
                bool need[kPerWave];
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                    need[j] = ((needBits >> j) & 1) && (pass == 0 || count[j] < kMaxContacts);
                int const f = (cb + k) * kChunk + lane;
                bool const valid = f < c.nF;
                float4 A = make_float4(0.f, 0.f, 0.f, 0.f), Bv = A, Cv = A;
                int tb = 0;
                if (valid)
                {
                    int4 const t = c.tri[f];
                    A = x[t.x], Bv = x[t.y], Cv = x[t.z];
                    tb = t.w;
                }
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                {
                    if (!need[j])
                        continue;
                    bool const other = valid && tb != body[j];
                    if (pass == 0)
                    {
                        if (other)
                            dmin[j] = fminf(dmin[j], PointTriangleSqDist(p[j], A, Bv, Cv));
                        continue;
                    }
                    bool const take = other && PointTriangleSqDist(p[j], A, Bv, Cv) <= thr[j];
                    unsigned long long const mask = __ballot(take);
                    int const slot = count[j] + __popcll(mask & ((1ull << lane) - 1ull));
                    bool const kept = take && slot < kMaxContacts;
                    if (kept && !finalize)
                        c.fc[static_cast<size_t>(c.cvIdx[v0 + j]) * kMaxContacts + slot] = f;
                    last[j] = max(last[j], WaveMaxInt(kept ? f : -1));
                    count[j] += __popcll(mask);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
    {
        if (!mine[j])
            continue;
        int const v = v0 + j;
        int const n = min(count[j], kMaxContacts);
        if (!finalize)
        {
            if (lane >= n && lane < kMaxContacts)
                c.fc[static_cast<size_t>(c.cvIdx[v]) * kMaxContacts + lane] = -1;
            continue;
        }
        if (lane == 0 && last[j] >= 0)
        {
            int4 const t = c.tri[last[j]];
            float4 const a = x[t.x], b = x[t.y], cc = x[t.z];
            float3 const nrm = Cross(Sub(b, a), Sub(cc, a));
            c.active[v] = Dot(Sub(p[j], a), nrm) < 0.f ? 1 : 0;
        }
    }
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
    int const* __restrict__ fc;
    int4 const* __restrict__ tri;
    float const* __restrict__ triArea;
    float const* __restrict__ xva;
    float4* __restrict__ xb; // new positions of the colour
    float muC, muF, epsv;
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
