// Vertex-triangle contact of the GPU VBD integrator: set-up from the collision
// mesh, active-set detection (gpu/impl/contact/VertexTriangleMixedCcdDcd.cu;
// gpu/impl/vbd/Integrator.cu:163-188, 250-271) and the contact-sweep write
// buffer's last copy. The contact terms of the sweep are in
// vbd_contact_kernels.hpp.
//
// Detection is an all-pairs search over chunks of 64 consecutive triangles,
// each with a bounding box: a wave takes kPerWave collision vertices, skips the
// chunks whose box lies beyond all of their query bounds, and for the others
// gives one triangle to each lane. The triangles are stored grouped by body and
// along a Morton curve of their rest centroids, so a chunk's box is tight. The
// search is exhaustive up to those exact bounds, which makes its answers
// independent of the storage and traversal order: a vertex's search radius is
// the largest over its overlapping triangles, and its nearest triangles are kept
// as the kMaxContacts smallest triangle indices, ascending (triangle order).

#include "pbat_vbd.h"
#include "status.hpp"
#include "vbd_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "integrator.hpp"

namespace pbat_vbd::device
{

constexpr int kContactBlock = 256;   // workgroup size of the detection kernels (4 waves)
constexpr int kChunk = 64;           // triangles per chunk box = lanes per wave
// Collision vertices per wave in the detection kernels. One: a wave's chunk
// visits then serve one query, so the few queries with far nearest triangles do
// not hold up three others (8 to 200 cubes: 2-15% faster steps than with 4).
constexpr int kPerWave = 1;
constexpr int kGroup = 4 * kPerWave; // collision vertices per workgroup

struct ContactArgs
{
    int const* __restrict__ cvIdx;     // internal vertex of each collision vertex
    int const* __restrict__ cvBody;    // body of each collision vertex
    int4 const* __restrict__ tri;      // internal vertices of each triangle, body of its first vertex in .w
    int4 const* __restrict__ triS;     // the same in storage order
    int const* __restrict__ triId;     // triangle index of each stored triangle
    float const* __restrict__ triArea; // triangle areas
    float const* __restrict__ xva;     // vertex areas (internal order)
    int* __restrict__ fc;              // kMaxContacts contacting triangles per internal vertex, -1 padded
    int* __restrict__ active;          // per collision vertex, kept across steps
    int* __restrict__ inAv;            // per collision vertex: in this step's active list
    float* __restrict__ dupper;        // per collision vertex: squared search radius
    float4* __restrict__ triLo;        // swept triangle boxes
    float4* __restrict__ triHi;
    float4* __restrict__ chunkLo;      // box of each chunk of kChunk consecutive triangles; when all
    float4* __restrict__ chunkHi;      // belong to one body, chunkHi.w != 0 and chunkLo.w = that body
    float4* __restrict__ superLo;      // the same for each super-chunk of 64 consecutive chunks
    float4* __restrict__ superHi;
    // triangle vertex positions at the nearest pass's x (body of the triangle in triA.w)
    float4* __restrict__ triA;
    float4* __restrict__ triB;
    float4* __restrict__ triC;
    int nCV, nF;
};

constexpr int kChunksPerRound = 4; // chunks whose triangles a nearest-pass round loads together

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
    int4 const t = c.triS[f];
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

// Box of each chunk of kChunk consecutive stored triangles: of their swept boxes
// (swept != 0) or of their vertices at x, whose positions per triangle it
// also stores for the nearest pass; and the body of the chunk's triangles when
// they all belong to one. One wave per chunk.
[[maybe_unused]] static __global__ __launch_bounds__(kContactBlock) void ContactChunkBoxes(
    ContactArgs c, float4 const* __restrict__ x, int swept)
{
    int const chunk = static_cast<int>(blockIdx.x) * (kContactBlock / kChunk) + static_cast<int>(threadIdx.x >> 6);
    int const f = chunk * kChunk + static_cast<int>(threadIdx.x & 63);
    float lx = 3.402823466e+38f, ly = lx, lz = lx, hx = -lx, hy = -lx, hz = -lx;
    int blo = 0x7fffffff, bhi = static_cast<int>(0x80000000u);
    if (f < c.nF)
    {
        blo = bhi = c.triS[f].w;
        if (swept)
        {
            float4 const L = c.triLo[f], U = c.triHi[f];
            lx = L.x, ly = L.y, lz = L.z, hx = U.x, hy = U.y, hz = U.z;
        }
        else
        {
            int4 const t = c.triS[f];
            float4 const a = x[t.x], b = x[t.y], cc = x[t.z];
            lx = fminf(a.x, fminf(b.x, cc.x)), ly = fminf(a.y, fminf(b.y, cc.y)), lz = fminf(a.z, fminf(b.z, cc.z));
            hx = fmaxf(a.x, fmaxf(b.x, cc.x)), hy = fmaxf(a.y, fmaxf(b.y, cc.y)), hz = fmaxf(a.z, fmaxf(b.z, cc.z));
            c.triA[f] = make_float4(a.x, a.y, a.z, __int_as_float(t.w));
            c.triB[f] = b;
            c.triC[f] = cc;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        lx = fminf(lx, __shfl_xor(lx, off)), ly = fminf(ly, __shfl_xor(ly, off));
        lz = fminf(lz, __shfl_xor(lz, off)), hx = fmaxf(hx, __shfl_xor(hx, off));
        hy = fmaxf(hy, __shfl_xor(hy, off)), hz = fmaxf(hz, __shfl_xor(hz, off));
        blo = min(blo, __shfl_xor(blo, off)), bhi = max(bhi, __shfl_xor(bhi, off));
    }
    if ((threadIdx.x & 63) == 0 && chunk * kChunk < c.nF)
    {
        c.chunkLo[chunk] = make_float4(lx, ly, lz, __int_as_float(blo));
        c.chunkHi[chunk] = make_float4(hx, hy, hz, blo == bhi ? 1.f : 0.f);
    }
}

// Box of each super-chunk of 64 consecutive chunk boxes, and its body when all
// its triangles belong to one. One wave per super-chunk.
[[maybe_unused]] static __global__ __launch_bounds__(kContactBlock) void ContactSuperBoxes(ContactArgs c)
{
    int const nChunks = (c.nF + kChunk - 1) / kChunk, nSuper = (nChunks + 63) / 64;
    int const sc = static_cast<int>(blockIdx.x) * (kContactBlock / 64) + static_cast<int>(threadIdx.x >> 6);
    int const k = sc * 64 + static_cast<int>(threadIdx.x & 63);
    float lx = 3.402823466e+38f, ly = lx, lz = lx, hx = -lx, hy = -lx, hz = -lx;
    int blo = 0x7fffffff, bhi = static_cast<int>(0x80000000u);
    bool mixed = false;
    if (k < nChunks)
    {
        float4 const L = c.chunkLo[k], U = c.chunkHi[k];
        lx = L.x, ly = L.y, lz = L.z, hx = U.x, hy = U.y, hz = U.z;
        blo = bhi = __float_as_int(L.w);
        mixed = U.w == 0.f;
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        lx = fminf(lx, __shfl_xor(lx, off)), ly = fminf(ly, __shfl_xor(ly, off));
        lz = fminf(lz, __shfl_xor(lz, off)), hx = fmaxf(hx, __shfl_xor(hx, off));
        hy = fmaxf(hy, __shfl_xor(hy, off)), hz = fmaxf(hz, __shfl_xor(hz, off));
        blo = min(blo, __shfl_xor(blo, off)), bhi = max(bhi, __shfl_xor(bhi, off));
    }
    bool const anyMixed = __ballot(mixed) != 0ull;
    if ((threadIdx.x & 63) == 0 && sc < nSuper)
    {
        c.superLo[sc] = make_float4(lx, ly, lz, __int_as_float(blo));
        c.superHi[sc] = make_float4(hx, hy, hz, (!anyMixed && blo == bhi) ? 1.f : 0.f);
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
    int const nChunks = (c.nF + kChunk - 1) / kChunk, nSuper = (nChunks + 63) / 64;
    // 64 chunk boxes at a time (one super-chunk), one per lane: the lanes test
    // them against the wave's segments, and the wave visits the chunks that any
    // segment overlaps. Super-chunk boxes, tested 64 at a time, skip the groups
    // that no segment meets (a chunk box lies inside its super-chunk's box).
    unsigned long long sTodo = 0ull;
    int sbits = 0;
    for (int cb = 0; cb < nChunks; cb += 64)
    {
        int const sIdx = cb >> 6;
        if ((sIdx & 63) == 0)
        {
            sbits = 0;
            if (sIdx + lane < nSuper)
            {
                float4 const SL = c.superLo[sIdx + lane], SU = c.superHi[sIdx + lane];
                bool const oneBody = SU.w != 0.f;
                int const superBody = __float_as_int(SL.w);
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                    if (v0 + j < c.nCV && !(oneBody && superBody == body[j]) && lo[j].x <= SU.x &&
                        lo[j].y <= SU.y && lo[j].z <= SU.z && SL.x <= hi[j].x && SL.y <= hi[j].y && SL.z <= hi[j].z)
                        sbits |= 1 << j;
            }
            sTodo = __ballot(sbits != 0);
        }
        if (!((sTodo >> (sIdx & 63)) & 1ull))
            continue;
        int const qmask = __shfl(sbits, sIdx & 63);
        int bits = 0;
        if (cb + lane < nChunks)
        {
            float4 const TL = c.chunkLo[cb + lane], TU = c.chunkHi[cb + lane];
            // a chunk of the vertex's own body holds no pair
            bool const oneBody = TU.w != 0.f;
            int const chunkBody = __float_as_int(TL.w);
#pragma unroll
            for (int j = 0; j < kPerWave; ++j)
                if (((qmask >> j) & 1) && !(oneBody && chunkBody == body[j]) && lo[j].x <= TU.x && lo[j].y <= TU.y && lo[j].z <= TU.z && TL.x <= hi[j].x &&
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
// otherwise; the kMaxContacts with the smallest triangle indices, ascending
// (triangle order), whatever the storage order scanned. finalize = 0
// (UpdateActiveSet, gpu/impl/vbd/Integrator.cu:250-271): they become the
// vertex's contacts. finalize = 1 (FinalizeActiveSet): the vertex stays active
// when it lies below the plane of the last of them. kPerWave collision vertices
// per wave; two passes over the triangle chunks, the second collecting with
// wave ballots into a sorted list held by the first kMaxContacts lanes. A round
// loads the triangles of up to kChunksPerRound chunks at once (positions stored
// by ContactChunkBoxes).
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
    int kept[kPerWave]; // lane s < kMaxContacts: the s-th smallest triangle index collected
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
        kept[j] = 0x7fffffff;
        anyWork = anyWork || mine[j];
    }
    if (!anyWork)
        return;
    // seed the nearest distance with the vertex's first contact of its last list
    // when that triangle lies well within reach (by 0.1%, far beyond the rounding
    // of the distances): it belongs to the searched set, so the minimum is
    // unchanged, and the search starts from a tight bound
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
    {
        if (!mine[j])
            continue;
        int const f0 = c.fc[static_cast<size_t>(c.cvIdx[v0 + j]) * kMaxContacts];
        if (f0 >= 0)
        {
            int4 const t0 = c.tri[f0];
            float const d0 = PointTriangleSqDist(p[j], x[t0.x], x[t0.y], x[t0.z]);
            if (d0 <= (r[j] + kFltEpsilon) * 0.999f)
                dmin[j] = d0;
        }
    }
    int const nChunks = (c.nF + kChunk - 1) / kChunk, nSuper = (nChunks + 63) / 64;
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
        // chunks of other bodies whose box lies within some query's bound: r + eps
        // in pass 0 (no triangle farther can matter), the collection threshold in
        // pass 1. In pass 0 the bound shrinks to 4 d + 1e-6 with d the nearest
        // squared distance found so far (before each group of boxes and after each
        // round): a box beyond it holds no triangle nearer than d, by a margin far
        // wider than the rounding of either distance, so the minimum is the one of
        // the full search. Super-chunk boxes, tested 64 at a time against the bound
        // (widened by 1e-4 against rounding), skip the groups of 64 chunks that no
        // query can reach.
        unsigned long long sTodo = 0ull;
        int sbits = 0;
        for (int cb = 0; cb < nChunks; cb += 64)
        {
            int const sIdx = cb >> 6;
            if ((sIdx & 63) == 0)
            {
                float sBound[kPerWave];
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                    sBound[j] = (pass == 0 ? fminf(r[j] + kFltEpsilon, 4.f * WaveMin(dmin[j]) + 1e-6f) : thr[j]) *
                                (1.f + 1e-4f);
                sbits = 0;
                if (sIdx + lane < nSuper)
                {
                    float4 const SL = c.superLo[sIdx + lane], SU = c.superHi[sIdx + lane];
                    bool const oneBody = SU.w != 0.f;
                    int const superBody = __float_as_int(SL.w);
#pragma unroll
                    for (int j = 0; j < kPerWave; ++j)
                        if (mine[j] && !(oneBody && superBody == body[j]) && (pass == 0 || thr[j] >= 0.f) &&
                            TileWithin(PointBoxSqDist(p[j], SL, SU), sBound[j]))
                            sbits |= 1 << j;
                }
                sTodo = __ballot(sbits != 0);
            }
            if (!((sTodo >> (sIdx & 63)) & 1ull))
                continue;
            int const qmask = __shfl(sbits, sIdx & 63);
            float bound[kPerWave];
#pragma unroll
            for (int j = 0; j < kPerWave; ++j)
                bound[j] = pass == 0 ? fminf(r[j] + kFltEpsilon, 4.f * WaveMin(dmin[j]) + 1e-6f) : thr[j];
            int bits = 0;
            float4 TL = make_float4(0.f, 0.f, 0.f, 0.f), TU = TL;
            if (cb + lane < nChunks)
            {
                TL = c.chunkLo[cb + lane], TU = c.chunkHi[cb + lane];
                bool const oneBody = TU.w != 0.f;
                int const chunkBody = __float_as_int(TL.w);
#pragma unroll
                for (int j = 0; j < kPerWave; ++j)
                {
                    if (((qmask >> j) & 1) && !(oneBody && chunkBody == body[j]) && (pass == 0 || thr[j] >= 0.f) &&
                        TileWithin(PointBoxSqDist(p[j], TL, TU), bound[j]))
                        bits |= 1 << j;
                }
            }
            unsigned long long todo = __ballot(bits != 0);
            while (todo)
            {
                int ks[kChunksPerRound];
                float4 A[kChunksPerRound], Bv[kChunksPerRound], Cv[kChunksPerRound];
#pragma unroll
                for (int i = 0; i < kChunksPerRound; ++i)
                {
                    ks[i] = -1;
                    A[i] = Bv[i] = Cv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (todo)
                    {
                        ks[i] = __ffsll(static_cast<long long>(todo)) - 1;
                        todo &= todo - 1ull;
                        int const f = (cb + ks[i]) * kChunk + lane;
                        if (f < c.nF)
                            A[i] = c.triA[f], Bv[i] = c.triB[f], Cv[i] = c.triC[f];
                    }
                }
#pragma unroll
                for (int i = 0; i < kChunksPerRound; ++i)
                {
                    if (ks[i] < 0)
                        break;
                    int const needBits = __shfl(bits, ks[i]);
                    bool need[kPerWave];
#pragma unroll
                    for (int j = 0; j < kPerWave; ++j)
                        need[j] = (needBits >> j) & 1;
                    int const f = (cb + ks[i]) * kChunk + lane;
                    bool const valid = f < c.nF;
                    int const tb = __float_as_int(A[i].w);
#pragma unroll
                    for (int j = 0; j < kPerWave; ++j)
                    {
                        if (!need[j])
                            continue;
                        bool const other = valid && tb != body[j];
                        if (pass == 0)
                        {
                            if (other)
                                dmin[j] = fminf(dmin[j], PointTriangleSqDist(p[j], A[i], Bv[i], Cv[i]));
                            continue;
                        }
                        bool const take = other && PointTriangleSqDist(p[j], A[i], Bv[i], Cv[i]) <= thr[j];
                        unsigned long long mask = __ballot(take);
                        count[j] += __popcll(mask);
                        int const id = take ? c.triId[f] : 0;
                        // insert each taken index into the ascending list of the
                        // kMaxContacts smallest (lanes 0 .. kMaxContacts - 1)
                        while (mask)
                        {
                            int const l = __ffsll(static_cast<long long>(mask)) - 1;
                            mask &= mask - 1ull;
                            int const t = __shfl(id, l);
                            int const pos = __popcll(__ballot(lane < kMaxContacts && kept[j] < t));
                            int const up = __shfl_up(kept[j], 1);
                            if (lane < kMaxContacts && pos < kMaxContacts)
                                kept[j] = lane < pos ? kept[j] : (lane == pos ? t : up);
                        }
                    }
                }
                if (pass == 0 && todo)
                {
#pragma unroll
                    for (int j = 0; j < kPerWave; ++j)
                        bound[j] = fminf(r[j] + kFltEpsilon, 4.f * WaveMin(dmin[j]) + 1e-6f);
                    int left = 0;
                    if ((todo >> lane) & 1ull)
                    {
#pragma unroll
                        for (int j = 0; j < kPerWave; ++j)
                            if (((bits >> j) & 1) && TileWithin(PointBoxSqDist(p[j], TL, TU), bound[j]))
                                left |= 1 << j;
                    }
                    bits = left;
                    todo = __ballot(left != 0);
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
        last[j] = n > 0 ? __shfl(kept[j], n - 1) : -1;
        if (!finalize)
        {
            if (lane < kMaxContacts)
                c.fc[static_cast<size_t>(c.cvIdx[v]) * kMaxContacts + lane] = lane < n ? kept[j] : -1;
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

// The sweep's per-slot contact records from the contact lists (kept, stale or
// not, for every collision vertex): the triangle's vertices and its penalty
// muC XVA_v / sum_k FA_k * FA_f (ContactPenalty, gpu/impl/vbd/Kernels.cuh:80-114;
// the sum over the vertex's contacting triangles in slot order), or x = -1 for
// an empty slot. One thread per collision vertex.
[[maybe_unused]] static __global__ __launch_bounds__(256) void ContactPack(ContactArgs c, float muC,
                                                                           int4* __restrict__ rec)
{
    int const k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= c.nCV)
        return;
    size_t const v = static_cast<size_t>(c.cvIdx[k]);
    int4 const fa = *reinterpret_cast<int4 const*>(c.fc + v * kMaxContacts);
    int4 const fb = *reinterpret_cast<int4 const*>(c.fc + v * kMaxContacts + 4);
    int const f[kMaxContacts] = {fa.x, fa.y, fa.z, fa.w, fb.x, fb.y, fb.z, fb.w};
    float sumfa = 0.f;
#pragma unroll
    for (int j = 0; j < kMaxContacts; ++j)
        if (f[j] >= 0)
            sumfa += c.triArea[f[j]];
    float const kC = c.xva[v] * muC / sumfa;
#pragma unroll
    for (int j = 0; j < kMaxContacts; ++j)
    {
        int4 r = make_int4(-1, -1, -1, 0);
        if (f[j] >= 0)
        {
            int4 const t = c.tri[f[j]];
            r = make_int4(t.x, t.y, t.z, __float_as_int(kC * c.triArea[f[j]]));
        }
        rec[v * kMaxContacts + j] = r;
    }
}

} // namespace pbat_vbd::device

namespace pbat_vbd {

// Copy the last contact sweep's colour back from the write buffer.
void FlushContactSweep(pbat_vbd_integrator& h)
{
    int const n = h.pendEnd - h.pendBeg;
    if (n > 0)
        hipLaunchKernelGGL(device::CopyColour, dim3((n + 255) / 256), dim3(256), 0, h.stream, h.xm.ptr,
                           static_cast<float4 const*>(h.xb.ptr), h.pendBeg, h.pendEnd);
    h.pendBeg = h.pendEnd = 0;
}

// Whether the collision mesh of d holds vertices or triangles of two bodies.
bool CollisionMeshSpansBodies(pbat_vbd_data const& d)
{
    if (d.nCollisionVertices <= 0 || d.nCollisionTriangles <= 0 || !d.B || !d.V || !d.F)
        return false;
    int64_t const b0 = d.B[d.V[0]];
    for (int64_t k = 0; k < d.nCollisionVertices; ++k)
        if (d.B[d.V[k]] != b0)
            return true;
    for (int64_t k = 0; k < d.nCollisionTriangles; ++k)
        if (d.B[d.F[3 * k]] != b0)
            return true;
    return false;
}

// Collision mesh of the data (sim/vbd/Data.cpp:34-60): contact is on when the
// collision mesh holds vertices or triangles of two bodies; otherwise the
// reference's detector can never find a pair and nothing is set up.
void SetUpContact(pbat_vbd_integrator& h, pbat_vbd_data const* d)
{
    int64_t const nV = h.nV, nCV = d->nCollisionVertices, nF = d->nCollisionTriangles;
    if (nCV < 0 || nF < 0 || (nCV > 0 && !d->V) || (nF > 0 && !d->F))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "invalid collision mesh");
    if (nCV == 0 || nF == 0 || !d->B)
        return;
    for (int64_t k = 0; k < nCV; ++k)
        if (d->V[k] < 0 || d->V[k] >= nV)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "collision vertex index out of range");
    for (int64_t k = 0; k < 3 * nF; ++k)
        if (d->F[k] < 0 || d->F[k] >= nV)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "collision triangle vertex index out of range");
    if (!CollisionMeshSpansBodies(*d))
        return;
    if (!d->XVA || !d->FA)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "contact needs vertex areas XVA and triangle areas FA");
    if (nCV > (int64_t{1} << 30) || nF > (int64_t{1} << 30))
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "collision mesh too large");
    if (d->activeSetUpdateFrequency < 1)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "active set update frequency must be >= 1");
    if (h.acceleration == PBAT_VBD_ACCELERATION_TRUST_REGION && !h.curved)
        Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT,
             "the linear trust-region path is not available with contact (its objective has no contact term)");
    h.contact = true;
    h.nCV = static_cast<int>(nCV), h.nF = static_cast<int>(nF);
    h.muC = static_cast<float>(d->muC), h.muF = static_cast<float>(d->muF), h.epsv = static_cast<float>(d->epsv);
    h.activeSetFrequency = d->activeSetUpdateFrequency;
    std::vector<int> cvIdx(static_cast<size_t>(nCV)), cvBody(cvIdx.size());
    h.cvOrig.assign(d->V, d->V + nCV);
    for (int64_t k = 0; k < nCV; ++k)
    {
        cvIdx[k] = h.iperm[d->V[k]];
        cvBody[k] = static_cast<int>(d->B[d->V[k]]);
    }
    std::vector<int4> tri(static_cast<size_t>(nF));
    std::vector<float> area(tri.size()), xva(static_cast<size_t>(nV));
    for (int64_t f = 0; f < nF; ++f)
    {
        tri[f] = make_int4(h.iperm[d->F[3 * f]], h.iperm[d->F[3 * f + 1]], h.iperm[d->F[3 * f + 2]],
                           static_cast<int>(d->B[d->F[3 * f]]));
        area[f] = static_cast<float>(d->FA[f]);
    }
    for (int64_t i = 0; i < nV; ++i)
        xva[h.iperm[i]] = static_cast<float>(d->XVA[i]);
    // storage order of the detection kernels: by body, then along a Morton curve
    // of the rest centroids over the collision mesh's box (10 bits per axis)
    std::vector<int> order(static_cast<size_t>(nF));
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t k = 0; k < 3 * nF; ++k)
            for (int a = 0; a < 3; ++a)
            {
                double const x = d->x[3 * d->F[k] + a];
                lo[a] = std::min(lo[a], x), hi[a] = std::max(hi[a], x);
            }
        auto spread = [](uint32_t v) { // 10 bits -> every third of 30 bits
            v &= 0x3ffu;
            v = (v | (v << 16)) & 0x030000ffu;
            v = (v | (v << 8)) & 0x0300f00fu;
            v = (v | (v << 4)) & 0x030c30c3u;
            v = (v | (v << 2)) & 0x09249249u;
            return v;
        };
        std::vector<uint32_t> code(static_cast<size_t>(nF));
        for (int64_t f = 0; f < nF; ++f)
        {
            uint32_t key = 0;
            for (int a = 0; a < 3; ++a)
            {
                double const cen = (d->x[3 * d->F[3 * f] + a] + d->x[3 * d->F[3 * f + 1] + a] +
                                    d->x[3 * d->F[3 * f + 2] + a]) / 3.;
                double const ext = hi[a] - lo[a];
                double const t = ext > 0. ? (cen - lo[a]) / ext : 0.;
                key |= spread(static_cast<uint32_t>(std::min(1023., std::max(0., std::floor(t * 1024.))))) << a;
            }
            code[f] = key;
            order[f] = static_cast<int>(f);
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return tri[a].w != tri[b].w ? tri[a].w < tri[b].w : code[a] < code[b];
        });
    }
    std::vector<int4> triS(static_cast<size_t>(nF));
    for (int64_t s = 0; s < nF; ++s)
        triS[s] = tri[order[s]];
    h.cvIdx.upload(cvIdx, h.stream);
    h.cvBody.upload(cvBody, h.stream);
    h.tri.upload(tri, h.stream);
    h.triS.upload(triS, h.stream);
    h.triId.upload(order, h.stream);
    h.triArea.upload(area, h.stream);
    h.xva.upload(xva, h.stream);
    // no contacts, nothing active, search radii at FLT_MAX
    // (gpu/impl/contact/VertexTriangleMixedCcdDcd.cu:22-48, gpu/impl/vbd/Integrator.cu:60)
    h.fc.upload(std::vector<int>(static_cast<size_t>(nV) * device::kMaxContacts, -1), h.stream);
    h.crec.upload(std::vector<int4>(static_cast<size_t>(nV) * device::kMaxContacts, make_int4(-1, -1, -1, 0)),
                  h.stream);
    h.cvActive.upload(std::vector<int>(static_cast<size_t>(nCV), 0), h.stream);
    h.cvInAv.upload(std::vector<int>(static_cast<size_t>(nCV), 0), h.stream);
    h.dupper.upload(std::vector<float>(static_cast<size_t>(nCV), std::numeric_limits<float>::max()), h.stream);
    h.triLo.allocate(static_cast<size_t>(nF));
    h.triHi.allocate(static_cast<size_t>(nF));
    size_t const nChunks = static_cast<size_t>((nF + device::kChunk - 1) / device::kChunk);
    h.chunkLo.allocate(nChunks);
    h.chunkHi.allocate(nChunks);
    h.superLo.allocate((nChunks + 63) / 64);
    h.superHi.allocate((nChunks + 63) / 64);
    h.triA.allocate(static_cast<size_t>(nF));
    h.triB.allocate(static_cast<size_t>(nF));
    h.triC.allocate(static_cast<size_t>(nF));
    h.xfull.allocate(static_cast<size_t>(nV));
    h.xb.allocate(static_cast<size_t>(nV));
}

// Workgroups of ContactSuperBoxes: one wave per super-chunk of 64 chunks.
int SuperBlocks(pbat_vbd_integrator const& h)
{
    int const nChunks = (h.nF + device::kChunk - 1) / device::kChunk, nSuper = (nChunks + 63) / 64;
    return (nSuper + device::kContactBlock / 64 - 1) / (device::kContactBlock / 64);
}

device::ContactArgs MakeContactArgs(pbat_vbd_integrator& h)
{
    device::ContactArgs c{};
    c.cvIdx = h.cvIdx.ptr, c.cvBody = h.cvBody.ptr, c.tri = h.tri.ptr, c.triArea = h.triArea.ptr;
    c.triS = h.triS.ptr, c.triId = h.triId.ptr;
    c.xva = h.xva.ptr, c.fc = h.fc.ptr, c.active = h.cvActive.ptr, c.inAv = h.cvInAv.ptr;
    c.dupper = h.dupper.ptr, c.triLo = h.triLo.ptr, c.triHi = h.triHi.ptr;
    c.chunkLo = h.chunkLo.ptr, c.chunkHi = h.chunkHi.ptr, c.superLo = h.superLo.ptr, c.superHi = h.superHi.ptr;
    c.triA = h.triA.ptr, c.triB = h.triB.ptr, c.triC = h.triC.ptr;
    c.nCV = h.nCV, c.nF = h.nF;
    return c;
}

// InitializeActiveSet (gpu/impl/vbd/Integrator.cu:163-188): the active set of the
// step from the sweep of every vertex over the full step, x_t -> x_t + dt v + dt^2 a.
void ContactBeginStep(pbat_vbd_integrator& h, float dt)
{
    int const n = static_cast<int>(h.nV);
    hipLaunchKernelGGL(device::ContactFullStep, dim3((n + 255) / 256), dim3(256), 0, h.stream,
                       static_cast<float4 const*>(h.xt.ptr), static_cast<float4 const*>(h.vel.ptr),
                       static_cast<float4 const*>(h.aext.ptr), n, dt, dt * dt, h.xfull.ptr);
    device::ContactArgs const c = MakeContactArgs(h);
    int const boxBlocks = (h.nF + device::kContactBlock - 1) / device::kContactBlock;
    hipLaunchKernelGGL(device::ContactSweptBoxes, dim3((h.nF + 255) / 256), dim3(256), 0, h.stream, c,
                       static_cast<float4 const*>(h.xt.ptr), static_cast<float4 const*>(h.xfull.ptr));
    hipLaunchKernelGGL(device::ContactChunkBoxes, dim3(boxBlocks), dim3(device::kContactBlock), 0, h.stream, c,
                       static_cast<float4 const*>(nullptr), 1);
    hipLaunchKernelGGL(device::ContactSuperBoxes, dim3(SuperBlocks(h)), dim3(device::kContactBlock), 0, h.stream, c);
    hipLaunchKernelGGL(device::ContactBroadphase, dim3((h.nCV + device::kGroup - 1) / device::kGroup),
                       dim3(device::kContactBlock), 0, h.stream, c, static_cast<float4 const*>(h.xt.ptr),
                       static_cast<float4 const*>(h.xfull.ptr));
}

// UpdateActiveSet (finalize = 0, gpu/impl/vbd/Integrator.cu:250-271; then the
// sweep's contact records) or FinalizeActiveSet (finalize = 1,
// gpu/impl/contact/VertexTriangleMixedCcdDcd.cu:185-215) at the current positions.
void ContactNearestPass(pbat_vbd_integrator& h, int finalize)
{
    device::ContactArgs const c = MakeContactArgs(h);
    int const boxBlocks = (h.nF + device::kContactBlock - 1) / device::kContactBlock;
    hipLaunchKernelGGL(device::ContactChunkBoxes, dim3(boxBlocks), dim3(device::kContactBlock), 0, h.stream, c,
                       static_cast<float4 const*>(h.xm.ptr), 0);
    hipLaunchKernelGGL(device::ContactSuperBoxes, dim3(SuperBlocks(h)), dim3(device::kContactBlock), 0, h.stream, c);
    hipLaunchKernelGGL(device::ContactNearest, dim3((h.nCV + device::kGroup - 1) / device::kGroup),
                       dim3(device::kContactBlock), 0, h.stream, c, static_cast<float4 const*>(h.xm.ptr), finalize);
    if (!finalize)
        hipLaunchKernelGGL(device::ContactPack, dim3((h.nCV + 255) / 256), dim3(256), 0, h.stream, c, h.muC,
                           h.crec.ptr);
}

} // namespace pbat_vbd

using namespace pbat_vbd;

extern "C" int pbat_vbd_integrator_contacts(pbat_vbd_integrator* h, int32_t* active, int32_t* fc)
{
    return Guard([&] {
        auto& H = Handle(h);
        if (!H.contact)
            Fail(PBAT_VBD_ERROR_INVALID_ARGUMENT, "contact is off for this integrator");
        PBAT_HIP(hipSetDevice(H.device));
        PBAT_HIP(hipStreamSynchronize(H.stream));
        if (active)
        {
            std::vector<int> a(static_cast<size_t>(H.nCV));
            PBAT_HIP(hipMemcpy(a.data(), H.cvActive.ptr, a.size() * sizeof(int), hipMemcpyDeviceToHost));
            for (int k = 0; k < H.nCV; ++k)
                active[k] = a[k];
        }
        if (fc)
        {
            size_t const n = static_cast<size_t>(H.nV) * device::kMaxContacts;
            std::vector<int> f(n);
            PBAT_HIP(hipMemcpy(f.data(), H.fc.ptr, n * sizeof(int), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < H.nV; ++i)
                for (int k = 0; k < device::kMaxContacts; ++k)
                    fc[device::kMaxContacts * i + k] = f[device::kMaxContacts * static_cast<size_t>(H.iperm[i]) + k];
        }
    });
}

