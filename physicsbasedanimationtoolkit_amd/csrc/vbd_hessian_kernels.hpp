// Device kernels of the adjoint's assembled step Hessian and of its fused
// conjugate-gradient iteration.
//
// The step Hessian H = M / h^2 + sum_e w_e d2Psi_e/dx2 is fixed during one
// linear solve (it is evaluated at the step's minimiser), so it is assembled
// once per solve into 3x3 blocks over the updated vertices (block-sparse rows,
// columns ascending) and every product of the solve is a block-sparse
// matrix-vector product: 40 bytes per block (36 B of values + the column)
// instead of re-evaluating the Stable Neo-Hookean Hessian of every incident tet
// per product. RowAssemble forms each tet's blocks once per row vertex and
// sums every block's terms in a fixed order. For tet e with shape-function gradient rows
// g_a, g_b of two of its vertices, F its deformation gradient and J = det F, the
// block is
//   w mu (g_a . g_b) I + w lambda (cof F g_a)(cof F g_b)^T
//     - w lambda (J - alpha) [F (g_a x g_b)]_x
// (d2Psi/dF2 = mu I + lambda cof F (x) cof F + lambda (J - alpha) d2J/dF2
// contracted with g_a and g_b; d2J/dF2 contracts to the cross-product matrix).
// The terms are formed in FP64 and kept in FP32; each block sums them in FP64
// in a fixed order (ascending tet) and is stored in FP32.
#pragma once

#include "vbd_adjoint_kernels.hpp"

namespace pbat_vbd::device {

// Values of the block-sparse step Hessian, structure of arrays per block:
// a = (h00, h01, h02, h10), b = (h11, h12, h20, h21), c = h22.
struct BsrValues
{
    float4* __restrict__ a;
    float4* __restrict__ b;
    float* __restrict__ c;
};

// Sparsity pattern and assembly sources: block k sits in row blockRow[k] at
// column col[k]; its sources src[srcPtr[k] .. srcPtr[k+1]) are 16 e + 4 la + lb
// (tet e, local vertex la of the row vertex, lb of the column vertex),
// ascending in e.
struct BsrPattern
{
    int const* __restrict__ rowPtr; // nRows + 1
    int const* __restrict__ col;
    int const* __restrict__ blockRow;
    int const* __restrict__ srcPtr;
    int const* __restrict__ src;
    int nRows;
    int nBlocks;
};

// General assembly (rows longer than RowAssemble takes): one thread per block,
// each source's tet recomputed. Diagonal blocks also receive m_i / h^2 I and
// write the FP64-inverted block to dinv[2 i], dinv[2 i + 1] as
// (00, 01, 02, 11), (12, 22).
__global__ __launch_bounds__(256) void BsrAssemble(float4 const* __restrict__ xs, int4 const* __restrict__ tetv,
                                                   float4 const* __restrict__ tgp, BsrPattern p, double invDt2,
                                                   BsrValues v, float4* __restrict__ dinv,
                                                   int const* __restrict__ rowVertex)
{
    int const k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= p.nBlocks)
        return;
    int const i = p.blockRow[k], j = p.col[k];
    double H[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    for (int s = p.srcPtr[k]; s < p.srcPtr[k + 1]; ++s)
    {
        int const code = p.src[s];
        int const e = code >> 4, la = (code >> 2) & 3, lb = code & 3;
        int4 const t = tetv[e];
        float4 const q0 = tgp[3 * e], q1 = tgp[3 * e + 1], q2 = tgp[3 * e + 2];
        float4 const x0 = xs[t.x], x1 = xs[t.y], x2 = xs[t.z], x3 = xs[t.w];
        double g[4][3] = {{0., 0., 0.}, {q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
        for (int c = 0; c < 3; ++c)
            g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
        double const d[3][3] = {{double(x1.x) - x0.x, double(x1.y) - x0.y, double(x1.z) - x0.z},
                                {double(x2.x) - x0.x, double(x2.y) - x0.y, double(x2.z) - x0.z},
                                {double(x3.x) - x0.x, double(x3.y) - x0.y, double(x3.z) - x0.z}};
        double F[3][3]; // F[row][col] = sum_k d[k][row] g[k + 1][col]
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                F[r][c] = d[0][r] * g[1][c] + d[1][r] * g[2][c] + d[2][r] * g[3][c];
        double cof[3][3]; // column c = f_{c+1} x f_{c+2}
        for (int c = 0; c < 3; ++c)
        {
            int const c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            cof[0][c] = F[1][c1] * F[2][c2] - F[2][c1] * F[1][c2];
            cof[1][c] = F[2][c1] * F[0][c2] - F[0][c1] * F[2][c2];
            cof[2][c] = F[0][c1] * F[1][c2] - F[1][c1] * F[0][c2];
        }
        double const J = F[0][0] * cof[0][0] + F[1][0] * cof[1][0] + F[2][0] * cof[2][0];
        double const* ga = g[la];
        double const* gb = g[lb];
        double ca[3], cb[3], w[3];
        double const gx[3] = {ga[1] * gb[2] - ga[2] * gb[1], ga[2] * gb[0] - ga[0] * gb[2],
                              ga[0] * gb[1] - ga[1] * gb[0]};
        for (int r = 0; r < 3; ++r)
        {
            ca[r] = cof[r][0] * ga[0] + cof[r][1] * ga[1] + cof[r][2] * ga[2];
            cb[r] = cof[r][0] * gb[0] + cof[r][1] * gb[1] + cof[r][2] * gb[2];
            w[r] = F[r][0] * gx[0] + F[r][1] * gx[1] + F[r][2] * gx[2];
        }
        double const wmu = q2.y, wlam = q2.z, sk = double(q2.z) * (J - double(q2.w));
        double const dg = wmu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
        // - sk [w]_x with [w]_x = [[0, -w2, w1], [w2, 0, -w0], [-w1, w0, 0]]
        double const S[3][3] = {{0., sk * w[2], -sk * w[1]}, {-sk * w[2], 0., sk * w[0]}, {sk * w[1], -sk * w[0], 0.}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                H[r][c] += (r == c ? dg : 0.) + wlam * ca[r] * cb[c] + S[r][c];
    }
    if (i == j)
    {
        double const m = double(xs[rowVertex[i]].w) * invDt2;
        H[0][0] += m, H[1][1] += m, H[2][2] += m;
        Inverse3 const C(H[0][0], 0.5 * (H[0][1] + H[1][0]), 0.5 * (H[0][2] + H[2][0]), H[1][1],
                         0.5 * (H[1][2] + H[2][1]), H[2][2]);
        dinv[2 * i] = make_float4(static_cast<float>(C.c00 * C.inv), static_cast<float>(C.c01 * C.inv),
                                  static_cast<float>(C.c02 * C.inv), static_cast<float>(C.c11 * C.inv));
        dinv[2 * i + 1] = make_float4(static_cast<float>(C.c12 * C.inv), static_cast<float>(C.c22 * C.inv), 0.f, 0.f);
    }
    v.a[k] = make_float4(static_cast<float>(H[0][0]), static_cast<float>(H[0][1]), static_cast<float>(H[0][2]),
                         static_cast<float>(H[1][0]));
    v.b[k] = make_float4(static_cast<float>(H[1][1]), static_cast<float>(H[1][2]), static_cast<float>(H[2][0]),
                         static_cast<float>(H[2][1]));
    v.c[k] = static_cast<float>(H[2][2]);
}

// The FP32 state's tet (F, cof F, J in FP64) and the Stable Neo-Hookean
// constants, for the blocks
//   H(a, b) = w mu (g_a . g_b) I + w lambda (cof F g_a)(cof F g_b)^T
//             - w lambda (J - alpha) [F (g_a x g_b)]_x.
struct TetHessian
{
    double g[4][3];  // shape-function gradient rows
    double F[3][3];  // F[row][col]
    double cg[4][3]; // cof F g_l
    double wmu, wlam, sk;
    __device__ __forceinline__ TetHessian(float4 const* __restrict__ xs, int4 const& t, float4 const& q0,
                                          float4 const& q1, float4 const& q2)
    {
        float4 const x0 = xs[t.x], x1 = xs[t.y], x2 = xs[t.z], x3 = xs[t.w];
        double const G[3][3] = {{q0.x, q0.y, q0.z}, {q0.w, q1.x, q1.y}, {q1.z, q1.w, q2.x}};
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            g[1][c] = G[0][c], g[2][c] = G[1][c], g[3][c] = G[2][c];
            g[0][c] = -(G[0][c] + G[1][c] + G[2][c]);
        }
        double const d[3][3] = {{double(x1.x) - x0.x, double(x1.y) - x0.y, double(x1.z) - x0.z},
                                {double(x2.x) - x0.x, double(x2.y) - x0.y, double(x2.z) - x0.z},
                                {double(x3.x) - x0.x, double(x3.y) - x0.y, double(x3.z) - x0.z}};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                F[r][c] = d[0][r] * G[0][c] + d[1][r] * G[1][c] + d[2][r] * G[2][c];
        double cof[3][3]; // column c = f_{c+1} x f_{c+2}
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            int const c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            cof[0][c] = F[1][c1] * F[2][c2] - F[2][c1] * F[1][c2];
            cof[1][c] = F[2][c1] * F[0][c2] - F[0][c1] * F[2][c2];
            cof[2][c] = F[0][c1] * F[1][c2] - F[1][c1] * F[0][c2];
        }
        double const J = F[0][0] * cof[0][0] + F[1][0] * cof[1][0] + F[2][0] * cof[2][0];
        wmu = q2.y, wlam = q2.z, sk = double(q2.z) * (J - double(q2.w));
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int r = 0; r < 3; ++r)
                cg[l][r] = cof[r][0] * g[l][0] + cof[r][1] * g[l][1] + cof[r][2] * g[l][2];
    }
    // H(a, b) for the row's gradient row ga and cof F ga, column lb
    template <int lb>
    __device__ __forceinline__ void Block(double const (&ga)[3], double const (&ca)[3], double (&H)[9]) const
    {
        double const* gb = g[lb];
        double const gx[3] = {ga[1] * gb[2] - ga[2] * gb[1], ga[2] * gb[0] - ga[0] * gb[2],
                              ga[0] * gb[1] - ga[1] * gb[0]};
        double w[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            w[r] = sk * (F[r][0] * gx[0] + F[r][1] * gx[1] + F[r][2] * gx[2]);
        double const dg = wmu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
        // - sk [F (g_a x g_b)]_x with [w]_x = [[0, -w2, w1], [w2, 0, -w0], [-w1, w0, 0]]
        double const S[9] = {0., w[2], -w[1], -w[2], 0., w[0], w[1], -w[0], 0.};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                H[3 * r + c] = (r == c ? dg : 0.) + wlam * ca[r] * cg[lb][c] + S[3 * r + c];
    }
};

// Row-major assembly: incidences of row i (tet e, local vertex la of the row
// vertex) in incCode[incPtr[i] .. incPtr[i + 1]) as 4 e + la, ascending e;
// incSlot holds, per incidence, the position in the row of the tet's local
// vertices 0..3 (-1: not updated).
struct RowAssembleArgs
{
    float4 const* __restrict__ xs;
    int4 const* __restrict__ tetv;
    float4 const* __restrict__ tgp;
    int const* __restrict__ incPtr;
    int const* __restrict__ incCode;
    int4 const* __restrict__ incSlot;
    int const* __restrict__ rowPtr;
    int const* __restrict__ col;
    int nRows;
    double invDt2;
    BsrValues v;
    float4* __restrict__ dinv;
    int const* __restrict__ rowVertex; // vertex of each row
};

// 32 lanes per row, two rows per wave, one wave per workgroup. The row's
// incidences are taken 32 at a time, one per lane: the lane forms its tet's
// F, cof F and J in FP64 and the four blocks (la, lb), lb = 0..3. The diagonal
// block (la, la) of every incidence is summed over the row's lanes by a
// fixed-order butterfly in FP64; the three others are staged in LDS as FP32,
// and lane l then adds, for each off-diagonal block it owns (l, l + 32, ..., NB
// of them, below the row length), the staged blocks that fall on it in
// incidence order (ascending tet), in FP64, visiting only the incidences its
// block appears in (a bit mask from the staged row positions). Diagonal blocks
// also receive m_i / h^2 I and write the FP64-inverted block to dinv[2 i],
// dinv[2 i + 1] as (00, 01, 02, 11), (12, 22).
template <int NB>
__global__ __launch_bounds__(64) void RowAssemble(RowAssembleArgs a)
{
    constexpr int L = 32;
    __shared__ float stage[2][L][4][9];
    __shared__ int4 slotS[2][L];
    int const lane = static_cast<int>(threadIdx.x) & (L - 1), half = static_cast<int>(threadIdx.x) >> 5;
    int const i = static_cast<int>(blockIdx.x) * 2 + half;
    int kb = 0, ke = 0, rb = 0, nb = 0;
    if (i < a.nRows)
        kb = a.incPtr[i], ke = a.incPtr[i + 1], rb = a.rowPtr[i], nb = a.rowPtr[i + 1] - rb;
    int const mine = (ke - kb + L - 1) / L;
    int const batches = max(mine, __shfl_xor(mine, 32)); // uniform over the wave
    double acc[NB][9];
#pragma unroll
    for (int o = 0; o < NB; ++o)
#pragma unroll
        for (int c = 0; c < 9; ++c)
            acc[o][c] = 0.;
    double diag[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    int diagSlot = -1; // the row vertex's position in its row (from its first incidence)
    for (int bt = 0; bt < batches; ++bt)
    {
        int const k = kb + bt * L + lane;
        double Hd[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
        if (k < ke)
        {
            int const code = a.incCode[k];
            int const e = code >> 2, la = code & 3;
            int4 const sl = a.incSlot[k];
            TetHessian const T(a.xs, a.tetv[e], a.tgp[3 * e], a.tgp[3 * e + 1], a.tgp[3 * e + 2]);
            double ga[3], ca[3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
            {
                ga[c] = la == 0 ? T.g[0][c] : la == 1 ? T.g[1][c] : la == 2 ? T.g[2][c] : T.g[3][c];
                ca[c] = la == 0 ? T.cg[0][c] : la == 1 ? T.cg[1][c] : la == 2 ? T.cg[2][c] : T.cg[3][c];
            }
            double H[9];
            auto keep = [&](int lb) {
                if (lb == la)
                {
#pragma unroll
                    for (int c = 0; c < 9; ++c)
                        Hd[c] = H[c];
                }
                else
                {
#pragma unroll
                    for (int c = 0; c < 9; ++c)
                        stage[half][lane][lb][c] = static_cast<float>(H[c]);
                }
            };
            T.Block<0>(ga, ca, H);
            keep(0);
            T.Block<1>(ga, ca, H);
            keep(1);
            T.Block<2>(ga, ca, H);
            keep(2);
            T.Block<3>(ga, ca, H);
            keep(3);
            slotS[half][lane] = sl;
        }
        // diagonal terms: butterfly over the row's 32 lanes (idle lanes add 0)
#pragma unroll
        for (int c = 0; c < 9; ++c)
        {
            double v = Hd[c];
#pragma unroll
            for (int off = 1; off < L; off <<= 1)
                v += __shfl_xor(v, off, L);
            diag[c] += v;
        }
        __syncthreads();
        if (bt == 0)
        {
            // every row has an incidence; lane 0 of the row holds its first
            int const code0 = kb < ke ? a.incCode[kb] : 0;
            int4 const s0 = slotS[half][0];
            int const la0 = code0 & 3;
            diagSlot = kb < ke ? (la0 == 0 ? s0.x : la0 == 1 ? s0.y : la0 == 2 ? s0.z : s0.w) : -1;
        }
        int const cnt = min(L, max(0, ke - kb - bt * L));
#pragma unroll
        for (int o = 0; o < NB; ++o)
        {
            int const b = lane + o * L;
            if (b >= nb || b == diagSlot)
                continue;
            unsigned mask = 0u;
            for (int j = 0; j < cnt; ++j)
            {
                int4 const sl = slotS[half][j];
                if (sl.x == b || sl.y == b || sl.z == b || sl.w == b)
                    mask |= 1u << j;
            }
            while (mask)
            {
                int const j = __builtin_ctz(mask);
                mask &= mask - 1u;
                int4 const sl = slotS[half][j];
                int const s = sl.x == b ? 0 : sl.y == b ? 1 : sl.z == b ? 2 : 3;
#pragma unroll
                for (int c = 0; c < 9; ++c)
                    acc[o][c] += stage[half][j][s][c];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int o = 0; o < NB; ++o)
    {
        int const b = lane + o * L;
        if (b >= nb)
            continue;
        int const kBlock = rb + b;
        double* H = acc[o];
        if (b == diagSlot)
        {
            double const m = double(a.xs[a.rowVertex[i]].w) * a.invDt2;
#pragma unroll
            for (int c = 0; c < 9; ++c)
                H[c] = diag[c];
            H[0] += m, H[4] += m, H[8] += m;
            Inverse3 const C(H[0], 0.5 * (H[1] + H[3]), 0.5 * (H[2] + H[6]), H[4], 0.5 * (H[5] + H[7]), H[8]);
            a.dinv[2 * i] = make_float4(static_cast<float>(C.c00 * C.inv), static_cast<float>(C.c01 * C.inv),
                                        static_cast<float>(C.c02 * C.inv), static_cast<float>(C.c11 * C.inv));
            a.dinv[2 * i + 1] =
                make_float4(static_cast<float>(C.c12 * C.inv), static_cast<float>(C.c22 * C.inv), 0.f, 0.f);
        }
        a.v.a[kBlock] = make_float4(static_cast<float>(H[0]), static_cast<float>(H[1]), static_cast<float>(H[2]),
                                    static_cast<float>(H[3]));
        a.v.b[kBlock] = make_float4(static_cast<float>(H[4]), static_cast<float>(H[5]), static_cast<float>(H[6]),
                                    static_cast<float>(H[7]));
        a.v.c[kBlock] = static_cast<float>(H[8]);
    }
}

enum class BsrMode : int
{
    kProduct = 0,        // out = H y
    kProductDot = 1,     // out = H y, acc[mesh] += y . H y
    kGaussSeidel = 2,    // rows of one colour: y_i += D_i^-1 (b_i - (H y)_i)
    kGaussSeidelDot = 3, // same, then acc[mesh] += b_i . y_i
    kDirection = 4       // p = z + beta p_old formed on the fly (written to y);
                         // out = H p, p . H p per row chunk (BsrDirection)
};

struct BsrArgs
{
    int const* __restrict__ rowPtr;
    int const* __restrict__ col;
    float4 const* __restrict__ va;
    float4 const* __restrict__ vb;
    float const* __restrict__ vc;
    float4* __restrict__ y;          // vector multiplied (Gauss-Seidel: updated in place)
    float4 const* __restrict__ b;    // right-hand side (Gauss-Seidel)
    float4* __restrict__ out;        // product
    float4 const* __restrict__ dinv; // inverse diagonal blocks (Gauss-Seidel)
    int const* __restrict__ meshOf;
    // kDirection: p_j = z_j + beta p_old_j of mesh m (p_old_j when m is done); z
    // holds the whole preconditioned residual, its coarse part included
    float4 const* __restrict__ z;
    float4 const* __restrict__ pOld;
    CgState cg;
    double* __restrict__ acc; // accumulators (dot modes), array accArray
    int accArray;
    int vbeg, vend;
    int xcdChunk;
};

// p_j = z_j + beta p_old_j (p_old_j when the mesh has converged)
__device__ __forceinline__ float4 DirectionVector(BsrArgs const& a, int j, float beta, bool frozen)
{
    float4 const po = a.pOld[j];
    if (frozen)
        return po;
    float4 const zj = a.z[j];
    return make_float4(zj.x + beta * po.x, zj.y + beta * po.y, zj.z + beta * po.z, 0.f);
}

// (H v)_row over blocks [kb, ke) of one row: lane l of the row's L lanes takes
// blocks l, l + L, ..., the first K of them loaded together before
// accumulating; the group sum leaves the product in every lane of the group.
template <int L, int K, BsrMode M>
__device__ __forceinline__ float3 RowProduct(BsrArgs const& a, int kb, int ke, int lane, float beta, bool frozen)
{
    constexpr bool kDir = M == BsrMode::kDirection;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    int cj[K];
    float4 A[K], B[K];
    float C[K];
#pragma unroll
    for (int u = 0; u < K; ++u)
    {
        int const k = kb + lane + u * L;
        cj[u] = -1;
        if (k < ke)
        {
            cj[u] = a.col[k];
            A[u] = a.va[k];
            B[u] = a.vb[k];
            C[u] = a.vc[k];
        }
    }
    float4 P[K];
#pragma unroll
    for (int u = 0; u < K; ++u)
        if (cj[u] >= 0)
        {
            if constexpr (kDir)
                P[u] = DirectionVector(a, cj[u], beta, frozen);
            else
                P[u] = a.y[cj[u]];
        }
#pragma unroll
    for (int u = 0; u < K; ++u)
        if (cj[u] >= 0)
        {
            hx += A[u].x * P[u].x + A[u].y * P[u].y + A[u].z * P[u].z;
            hy += A[u].w * P[u].x + B[u].x * P[u].y + B[u].y * P[u].z;
            hz += B[u].z * P[u].x + B[u].w * P[u].y + C[u] * P[u].z;
        }
    for (int k = kb + lane + K * L; k < ke; k += L)
    {
        float4 const Ak = a.va[k], Bk = a.vb[k];
        float const Ck = a.vc[k];
        float4 Pk;
        if constexpr (kDir)
            Pk = DirectionVector(a, a.col[k], beta, frozen);
        else
            Pk = a.y[a.col[k]];
        hx += Ak.x * Pk.x + Ak.y * Pk.y + Ak.z * Pk.z;
        hy += Ak.w * Pk.x + Bk.x * Pk.y + Bk.y * Pk.z;
        hz += Bk.z * Pk.x + Bk.w * Pk.y + Ck * Pk.z;
    }
    return make_float3(GroupSum<L>(hx), GroupSum<L>(hy), GroupSum<L>(hz));
}

// L lanes per row of [vbeg, vend): products and Gauss-Seidel passes.
template <int L, int K, BsrMode M>
__global__ __launch_bounds__(256) void BsrKernel(BsrArgs a)
{
    static_assert(M != BsrMode::kDirection, "the direction product runs over row chunks (BsrDirection)");
    constexpr bool kGS = M == BsrMode::kGaussSeidel || M == BsrMode::kGaussSeidelDot;
    constexpr bool kDot = M == BsrMode::kGaussSeidelDot || M == BsrMode::kProductDot;
    int block = static_cast<int>(blockIdx.x);
    if (a.xcdChunk > 0)
        block = (block & 7) * a.xcdChunk + (block >> 3);
    int const tid = block * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    int const lane = tid & (L - 1);
    int const row = a.vbeg + tid / L;
    bool const active = row < a.vend;
    int kb = 0, ke = 0;
    if (active)
    {
        kb = a.rowPtr[row];
        ke = a.rowPtr[row + 1];
    }
    float3 const h = RowProduct<L, K, M>(a, kb, ke, lane, 0.f, false);
    bool const leader = active && lane == 0;
    double dot = 0.;
    if (leader)
    {
        if constexpr (kGS)
        {
            float4 const bo = a.b[row], yo = a.y[row];
            float const rx = bo.x - h.x, ry = bo.y - h.y, rz = bo.z - h.z;
            float4 const D0 = a.dinv[2 * row], D1 = a.dinv[2 * row + 1];
            float4 const r = make_float4(yo.x + D0.x * rx + D0.y * ry + D0.z * rz,
                                         yo.y + D0.y * rx + D0.w * ry + D1.x * rz,
                                         yo.z + D0.z * rx + D1.x * ry + D1.y * rz, 0.f);
            a.y[row] = r;
            dot = double(bo.x) * r.x + double(bo.y) * r.y + double(bo.z) * r.z;
        }
        else
        {
            a.out[row] = make_float4(h.x, h.y, h.z, 0.f);
            if constexpr (M == BsrMode::kProductDot)
            {
                float4 const yr = a.y[row];
                dot = double(yr.x) * h.x + double(yr.y) * h.y + double(yr.z) * h.z;
            }
        }
    }
    if constexpr (kDot)
        BlockAccumulate(dot, leader ? a.meshOf[row] : -1, a.acc, a.accArray, a.cg.nM);
}

// Row chunks of the fused CG iteration: each (mesh, colour) run of rows is cut
// into chunks of at most 256 / L rows, numbered mesh-major, so the chunks of
// a mesh are consecutive (aggChunks holds each aggregate's mesh's range) and every chunk holds
// rows of one mesh. A chunk's sums depend only on its own rows, so a mesh gets
// the same partial sums whether it is stepped alone or in a batch.
//
// One workgroup per chunk (XCD-aware order when xcdChunk > 0): the direction
// p = z + beta p_old of each of its rows is formed on the fly for every block
// (z includes the coarse correction), written to y, q = H p to out, and the chunk's p . q, summed over its
// rows in order, to pqPart[chunk]. Chunks of a converged mesh do nothing.
template <int L, int K>
__global__ __launch_bounds__(256) void BsrDirection(BsrArgs a, int4 const* __restrict__ chunks, int nChunks,
                                                    double* __restrict__ pqPart)
{
    constexpr int R = 256 / L;
    int const nM = a.cg.nM;
    int w = static_cast<int>(blockIdx.x);
    if (a.xcdChunk > 0)
        w = (w & 7) * a.xcdChunk + (w >> 3);
    if (w >= nChunks)
        return; // padding of the XCD order
    // the chunk (rows and mesh), then its mesh's done flag and beta together
    // with the row's extent
    int4 const ch = chunks[w];
    int const m = ch.z;
    constexpr bool frozen = false;
    int const r = static_cast<int>(threadIdx.x) / L, lane = static_cast<int>(threadIdx.x) & (L - 1);
    int const row = ch.x + r;
    bool const active = row < ch.y;
    int const done = a.cg.F[kDone * nM + m];
    float const beta = static_cast<float>(a.cg.P[kBeta * nM + m]);
    int kb = 0, ke = 0;
    if (active)
    {
        kb = a.rowPtr[row];
        ke = a.rowPtr[row + 1];
    }
    if (done != 0)
        return; // the chunk's mesh has converged: its vectors no longer change
    float4 pr = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active && lane == 0) // the row's own direction, in flight during the product
        pr = DirectionVector(a, row, beta, frozen);
    float3 const h = RowProduct<L, K, BsrMode::kDirection>(a, kb, ke, lane, beta, frozen);
    __shared__ double sdot[R];
    if (lane == 0)
    {
        double dot = 0.;
        if (active)
        {
            a.y[row] = pr; // the direction of this iteration (the p_old of the next)
            a.out[row] = make_float4(h.x, h.y, h.z, 0.f);
            dot = double(pr.x) * h.x + double(pr.y) * h.y + double(pr.z) * h.z;
        }
        sdot[r] = dot;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s = 0.;
        for (int k = 0; k < R; ++k)
            s += sdot[k];
        pqPart[w] = s;
    }
}

// Thread t of a group of N threads sums v[b + t], v[b + t + N], ... in that
// order, issuing the loads U at a time (the order, and so the bits, do not
// depend on U: out-of-range entries add an exact 0).
template <int N, int U = 16>
__device__ __forceinline__ double StridedSum(double const* __restrict__ v, int b, int e, int t)
{
    double s = 0.;
    for (int k0 = b + t; k0 < e; k0 += U * N)
    {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            int const k = k0 + u * N;
            x[u] = k < e ? v[k] : 0.;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            s += x[u];
    }
    return s;
}

// Sums of V values over the 128 threads of a workgroup in a fixed order (each
// wave's butterfly, then wave 0 + wave 1); every thread calls it and gets the
// sums in v.
template <int V>
__device__ __forceinline__ void Sum128(double (&v)[V], double (*lds)[V])
{
#pragma unroll
    for (int i = 0; i < V; ++i)
        for (int off = 32; off > 0; off >>= 1)
            v[i] += __shfl_xor(v[i], off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < V; ++i)
            lds[threadIdx.x >> 6][i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < V; ++i)
        v[i] = lds[0][i] + lds[1][i];
}

// One workgroup of 128 threads per aggregate g of mesh m, alpha = rz / pq with
// pq the sum of m's chunk partials in order (0 once m has converged); over the
// aggregate's vertices:
//   init:   r = b - q               (q = H y0)
//   update: y += alpha p, r -= alpha q
// then z = D^-1 r, and the aggregate's r . z, r . r (and b . b at init) and,
// when rc is given, its restriction rc = Z^T r = (sum r_i, sum d_i x r_i),
// summed in a fixed order and written to the per-aggregate partials. Every updated vertex belongs to
// exactly one aggregate and every aggregate to one mesh.
struct AggUpdateArgs
{
    int const* __restrict__ aggPtr;
    int const* __restrict__ aggVerts;
    int const* __restrict__ aggMesh;
    float4* __restrict__ y;
    float4* __restrict__ r;
    float4* __restrict__ z;
    float4 const* __restrict__ p;
    float4 const* __restrict__ q;
    float4 const* __restrict__ b; // init only
    float4 const* __restrict__ dinv;
    double* __restrict__ rc;      // nullable
    float4 const* __restrict__ coarseD; // per row: d_i of the coarse rotations (with rc)
    double const* __restrict__ pqPart;
    int2 const* __restrict__ aggChunks; // chunks of the aggregate's mesh
    int const* __restrict__ aggFirst;   // kAggFirst vertices per aggregate (-1: none)
    int const* __restrict__ aggDone;    // done flag of the aggregate's mesh
    double* __restrict__ rzPart;
    double* __restrict__ rrPart;
    double* __restrict__ bbPart; // init only
    CgState cg;
    int init;
};

constexpr int kAggFirst = 128; // = AggUpdate's workgroup size

__global__ __launch_bounds__(128) void AggUpdate(AggUpdateArgs u)
{
    int const nM = u.cg.nM;
    __shared__ double lds1[2][1];
    __shared__ double lds9[2][9];
    int const g = static_cast<int>(blockIdx.x), t = static_cast<int>(threadIdx.x);
    // independent loads first: the mesh, its chunk range, the aggregate's vertex
    // range, its mesh's done flag and the thread's first vertex. A converged
    // mesh's workgroups leave after those; the others load the mesh's chunk
    // partials, r . z and the first vertex's data together.
    int const m = u.aggMesh[g];
    int2 const cr = u.aggChunks[g];
    int const kBeg = u.aggPtr[g] + t, kEnd = u.aggPtr[g + 1];
    int const done = u.init ? 0 : u.aggDone[g];
    int const i0 = u.aggFirst[static_cast<size_t>(g) * kAggFirst + t];
    if (done != 0)
        return; // the aggregate's mesh has converged (uniform over the workgroup)
    double pq[1] = {u.init ? 0. : StridedSum<128, 32>(u.pqPart, cr.x, cr.y, t)};
    double const rzOld = u.init ? 0. : u.cg.P[kRz * nM + m];
    // the thread's first vertex: loads issued before alpha is known
    float4 const zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 q0 = zero4, a0 = zero4, y0 = zero4, r0 = zero4, d0 = zero4, d1 = zero4, o0 = zero4;
    if (i0 >= 0)
    {
        if (u.rc)
            o0 = u.coarseD[i0];
        q0 = u.q[i0];
        a0 = u.init ? u.b[i0] : u.p[i0];
        if (!u.init)
        {
            y0 = u.y[i0];
            r0 = u.r[i0];
        }
        d0 = u.dinv[2 * i0];
        d1 = u.dinv[2 * i0 + 1];
    }
    float alpha = 0.f;
    if (!u.init)
    {
        Sum128<1>(pq, lds1);
        alpha = pq[0] != 0. ? static_cast<float>(rzOld / pq[0]) : 0.f;
    }
    double sx = 0., sy = 0., sz = 0., ox = 0., oy = 0., oz = 0., rz = 0., rr = 0., bb = 0.;
    // init: a = b, r = b - q; update: a = p, y += alpha p, r -= alpha q
    auto update = [&](int i, float4 const& qi, float4 const& ai, float4 yi, float4 ri, float4 const& D0,
                      float4 const& D1, float4 const& d) {
        if (u.init)
        {
            ri = make_float4(ai.x - qi.x, ai.y - qi.y, ai.z - qi.z, 0.f);
            bb += Dot3(ai, ai);
        }
        else
        {
            yi.x += alpha * ai.x, yi.y += alpha * ai.y, yi.z += alpha * ai.z;
            ri.x -= alpha * qi.x, ri.y -= alpha * qi.y, ri.z -= alpha * qi.z;
            u.y[i] = yi;
        }
        u.r[i] = ri;
        float4 const zi = make_float4(D0.x * ri.x + D0.y * ri.y + D0.z * ri.z,
                                      D0.y * ri.x + D0.w * ri.y + D1.x * ri.z,
                                      D0.z * ri.x + D1.x * ri.y + D1.y * ri.z, 0.f);
        u.z[i] = zi;
        sx += ri.x, sy += ri.y, sz += ri.z;
        ox += d.y * ri.z - d.z * ri.y, oy += d.z * ri.x - d.x * ri.z, oz += d.x * ri.y - d.y * ri.x;
        rz += Dot3(ri, zi);
        rr += Dot3(ri, ri);
    };
    if (i0 >= 0)
        update(i0, q0, a0, y0, r0, d0, d1, o0);
    for (int k = kBeg + 128; k < kEnd; k += 128)
    {
        int const i = u.aggVerts[k];
        float4 const qi = u.q[i], ai = u.init ? u.b[i] : u.p[i];
        float4 yi = qi, ri = qi;
        if (!u.init)
        {
            yi = u.y[i];
            ri = u.r[i];
        }
        update(i, qi, ai, yi, ri, u.dinv[2 * i], u.dinv[2 * i + 1], u.rc ? u.coarseD[i] : zero4);
    }
    double sums[9] = {rz, rr, bb, sx, sy, sz, ox, oy, oz};
    Sum128<9>(sums, lds9);
    if (t == 0)
    {
        if (u.rc)
            for (int c = 0; c < kCoarseDofs; ++c)
                u.rc[kCoarseDofs * g + c] = sums[3 + c];
        u.rzPart[g] = sums[0];
        u.rrPart[g] = sums[1];
        if (u.init)
            u.bbPart[g] = sums[2];
    }
}

// zc = (Z^T H Z)^-1 rc per mesh and each coarse row's share of rc . zc, the
// coarse part of the mesh's r . z, into rczcPart; then the coarse correction
// Z_i zc_g = t_g + w_g x d_i (zc_g = (t_g, w_g)) added to z_i for the
// aggregate's rows, so that z = D^-1 r + Z zc. One workgroup of kCoarseWaves
// waves per aggregate g computes its kCoarseDofs rows: wave w takes the
// 64-column chunks w, w + kCoarseWaves, ... of its mesh's n columns, loads rc
// and the rows' entries of kU of them at once, then accumulates (kU chosen for
// the largest mesh; the order of the terms does not depend on it); the waves'
// sums are added in order. The order depends on n only, so a mesh gets the
// same zc alone or in a batch. The aggregate's descriptor (its mesh's first
// coarse row b0 and size n, its rows' range) and the offset of its first
// inverse row are precomputed, and the rows' z and d are loaded before the
// product, so the kernel waits on memory about twice. The fused solvers number
// the rows aggregate-major: aggregate g's rows are the range info.z .. info.w
// itself.
constexpr int kCoarseWaves = 4;

struct CoarseApplyArgs
{
    float const* __restrict__ coarseF;
    double const* __restrict__ rc;
    double* __restrict__ zc;
    double* __restrict__ rczcPart;
    int4 const* __restrict__ info;            // per aggregate: (b0, n, first row, end row)
    long long const* __restrict__ rowOffset;  // per aggregate: its first row in coarseF
    int const* __restrict__ aggDone;          // per aggregate: its mesh's done flag
    float4 const* __restrict__ coarseD;
    float4* __restrict__ z;
};

template <int kU>
__global__ __launch_bounds__(64 * kCoarseWaves) void CoarseApplyRows(CoarseApplyArgs a, int first)
{
    constexpr int kR = kCoarseDofs, kW = kCoarseWaves;
    __shared__ double part[kW][kR];
    __shared__ float zcS[kR];
    int const g = static_cast<int>(blockIdx.x), t = static_cast<int>(threadIdx.x);
    int const wave = t >> 6, lane = t & 63;
    int4 const inf = a.info[g];
    long long const off = a.rowOffset[g];
    int const done = first ? 0 : a.aggDone[g];
    if (done != 0)
        return; // the mesh has converged
    int const b0 = inf.x, n = inf.y;
    // the thread's first row of the aggregate, its d and z, in flight during the product
    int const k0 = inf.z + t;
    int i0 = -1;
    float4 d0 = make_float4(0.f, 0.f, 0.f, 0.f), z0 = d0;
    if (k0 < inf.w)
    {
        i0 = k0;
        d0 = a.coarseD[i0];
        z0 = a.z[i0];
    }
    float const* A = a.coarseF + off;
    double s[kR];
#pragma unroll
    for (int i = 0; i < kR; ++i)
        s[i] = 0.;
    for (int c0 = wave; c0 * 64 < n; c0 += kW * kU)
    {
        double x[kU];
        float v[kR][kU];
        // branch-free: out-of-range lanes load column n - 1 and drop it, so
        // every load of the pass is in flight together
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
            int const j = min((c0 + kW * u) * 64 + lane, n - 1);
            x[u] = a.rc[b0 + j];
#pragma unroll
            for (int i = 0; i < kR; ++i)
                v[i][u] = A[static_cast<size_t>(i) * n + j];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if ((c0 + kW * u) * 64 + lane >= n)
                x[u] = 0.;
#pragma unroll
        for (int i = 0; i < kR; ++i)
#pragma unroll
            for (int u = 0; u < kU; ++u)
                s[i] += double(v[i][u]) * x[u];
    }
#pragma unroll
    for (int i = 0; i < kR; ++i)
        for (int o = 32; o > 0; o >>= 1)
            s[i] += __shfl_xor(s[i], o);
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < kR; ++i)
            part[wave][i] = s[i];
    __syncthreads();
    if (t < kR)
    {
        int const row = kR * g + t;
        double tot = part[0][t];
        for (int w = 1; w < kW; ++w)
            tot += part[w][t];
        a.zc[row] = tot;
        a.rczcPart[row] = tot * a.rc[row];
        zcS[t] = static_cast<float>(tot);
    }
    __syncthreads();
    float const tx = zcS[0], ty = zcS[1], tz = zcS[2], wx = zcS[3], wy = zcS[4], wz = zcS[5];
    auto correct = [&](int i, float4 const& d, float4 zi) {
        zi.x += tx + (wy * d.z - wz * d.y);
        zi.y += ty + (wz * d.x - wx * d.z);
        zi.z += tz + (wx * d.y - wy * d.x);
        a.z[i] = zi;
    };
    if (i0 >= 0)
        correct(i0, d0, z0);
    for (int i = k0 + 64 * kW; i < inf.w; i += 64 * kW)
        correct(i, a.coarseD[i], a.z[i]);
}

// A solve's statistics without a host round trip: out = (largest per-mesh
// iteration count, largest per-mesh |r| / |b| at its last iteration).
__global__ __launch_bounds__(64) void SolveStats(CgState cg, double* __restrict__ out)
{
    int const nM = cg.nM, lane = static_cast<int>(threadIdx.x);
    int it = 0;
    double rel = 0.;
    for (int m = lane; m < nM; m += 64)
    {
        it = max(it, cg.F[kIters * nM + m]);
        double const bb = cg.P[kBb * nM + m];
        if (bb > 0.)
            rel = fmax(rel, sqrt(cg.P[kRrLast * nM + m] / bb));
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        it = max(it, __shfl_xor(it, off));
        rel = fmax(rel, __shfl_xor(rel, off));
    }
    if (lane == 0)
    {
        out[0] = it;
        out[1] = rel;
    }
}

// Vectors between the vertices' order and the solver's row order.
__global__ __launch_bounds__(256) void GatherRows(float4* __restrict__ out, float4 const* __restrict__ in,
                                                  int const* __restrict__ rowVertex, int n)
{
    int const s = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s < n)
        out[s] = in[rowVertex[s]];
}

__global__ __launch_bounds__(256) void ScatterRows(float4* __restrict__ out, float4 const* __restrict__ in,
                                                   int const* __restrict__ rowVertex, int n)
{
    int const s = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (s < n)
        out[rowVertex[s]] = in[s];
}

// ---- initial guess of a backward solve from the previous solutions ----
// With s_1, ..., s_k the previous solves' solutions (newest first), the basis
// b_j = the j-th backward difference at s_1 (b_0 = s_1, b_1 = s_1 - s_2,
// b_2 = s_1 - 2 s_2 + s_3, ...: the same span, far better conditioned than the
// nearly parallel solutions themselves) and y0 = sum_j c_j b_j with, per mesh,
// (B^T H B) c = B^T b: the element of the span closest to the solution in the
// H-norm, which is the norm CG minimises. Polynomial extrapolation of the
// solutions is one element of that span.
constexpr int kProjMax = 8;
constexpr int kProjPairs = kProjMax * (kProjMax + 1) / 2;
constexpr int kProjTerms = kProjPairs + kProjMax; // b_p . H b_q (p <= q), b_p . b

struct ProjArgs
{
    float4 const* hist[kProjMax]; // previous solutions, newest first
    float4* basis[kProjMax];
    float4* hb[kProjMax];         // H b_j
    float4 const* __restrict__ rhs;
    float4* __restrict__ y;
    int const* __restrict__ aggPtr;
    int const* __restrict__ aggVerts;
    int const* __restrict__ meshAggBegin;
    int const* __restrict__ meshOf;
    double* __restrict__ part; // kProjTerms x nAgg, term-major
    double* __restrict__ coef; // kProjMax x nM
    int k, n, nM, nAgg;
};

// H b_j for the k basis vectors at once, 16 lanes per row: every block of the
// row is read once and applied to the k vectors (one pass over the Hessian
// instead of k); lane sums in block order, then the 16-lane butterfly.
__global__ __launch_bounds__(256) void ProjProducts(ProjArgs a, int const* __restrict__ rowPtr,
                                                    int const* __restrict__ col, float4 const* __restrict__ va,
                                                    float4 const* __restrict__ vb, float const* __restrict__ vc)
{
    constexpr int L = 16;
    int const tid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    int const row = tid / L, lane = tid & (L - 1);
    int kb = 0, ke = 0;
    if (row < a.n)
        kb = rowPtr[row], ke = rowPtr[row + 1];
    float hx[kProjMax], hy[kProjMax], hz[kProjMax];
#pragma unroll
    for (int v = 0; v < kProjMax; ++v)
        hx[v] = hy[v] = hz[v] = 0.f;
    for (int k = kb + lane; k < ke; k += L)
    {
        float4 const A = va[k], B = vb[k];
        float const C = vc[k];
        int const j = col[k];
#pragma unroll
        for (int v = 0; v < kProjMax; ++v)
        {
            if (v >= a.k)
                break;
            float4 const P = a.basis[v][j];
            hx[v] += A.x * P.x + A.y * P.y + A.z * P.z;
            hy[v] += A.w * P.x + B.x * P.y + B.y * P.z;
            hz[v] += B.z * P.x + B.w * P.y + C * P.z;
        }
    }
#pragma unroll
    for (int v = 0; v < kProjMax; ++v)
    {
        if (v >= a.k)
            break;
        float const x = GroupSum<L>(hx[v]), y = GroupSum<L>(hy[v]), z = GroupSum<L>(hz[v]);
        if (lane == 0 && row < a.n)
            a.hb[v][row] = make_float4(x, y, z, 0.f);
    }
}

__global__ __launch_bounds__(256) void ProjBasis(ProjArgs a)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.n)
        return;
    float4 s[kProjMax];
#pragma unroll
    for (int j = 0; j < kProjMax; ++j)
        s[j] = j < a.k ? a.hist[j][i] : make_float4(0.f, 0.f, 0.f, 0.f);
    a.basis[0][i] = s[0];
    // d holds the differences of the current order, b their first entry
#pragma unroll
    for (int o = 1; o < kProjMax; ++o)
    {
        if (o >= a.k)
            break;
#pragma unroll
        for (int j = 0; j + o < kProjMax; ++j)
            s[j] = make_float4(s[j].x - s[j + 1].x, s[j].y - s[j + 1].y, s[j].z - s[j + 1].z, 0.f);
        a.basis[o][i] = s[0];
    }
}

// One 128-thread workgroup per aggregate: its vertices' b_p . H b_q and
// b_p . b in FP64, summed in a fixed order.
__global__ __launch_bounds__(128) void ProjGram(ProjArgs a)
{
    __shared__ double lds[2][kProjTerms];
    int const g = static_cast<int>(blockIdx.x), t = static_cast<int>(threadIdx.x);
    double acc[kProjTerms];
#pragma unroll
    for (int u = 0; u < kProjTerms; ++u)
        acc[u] = 0.;
    for (int kk = a.aggPtr[g] + t; kk < a.aggPtr[g + 1]; kk += 128)
    {
        int const i = a.aggVerts[kk];
        float4 b[kProjMax], hb[kProjMax];
#pragma unroll
        for (int j = 0; j < kProjMax; ++j)
        {
            b[j] = j < a.k ? a.basis[j][i] : make_float4(0.f, 0.f, 0.f, 0.f);
            hb[j] = j < a.k ? a.hb[j][i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        float4 const r = a.rhs[i];
        int u = 0;
#pragma unroll
        for (int p = 0; p < kProjMax; ++p)
#pragma unroll
            for (int q = p; q < kProjMax; ++q, ++u)
                if (q < a.k)
                    acc[u] += Dot3(b[p], hb[q]);
#pragma unroll
        for (int p = 0; p < kProjMax; ++p)
            if (p < a.k)
                acc[kProjPairs + p] += Dot3(b[p], r);
    }
    Sum128<kProjTerms>(acc, lds);
    if (t < kProjTerms)
    {
        double v = 0.;
#pragma unroll
        for (int u = 0; u < kProjTerms; ++u)
            v = u == t ? acc[u] : v;
        a.part[static_cast<size_t>(t) * a.nAgg + g] = v;
    }
}

// Sum of v[b .. e) over one wave in a fixed order; every lane gets it.
__device__ __forceinline__ double MeshSum(double const* __restrict__ v, int b, int e, int lane)
{
    double s = StridedSum<64>(v, b, e, lane);
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    return s;
}

// One workgroup per mesh: the mesh's sums over its aggregates in a fixed order
// (each by one wave), then (B^T H B) c = B^T b scaled to a unit diagonal and
// solved by LDL^T, dropping a basis vector whose pivot falls below 1e-6
// (numerically in the span of the others) or whose diagonal is not positive.
__global__ __launch_bounds__(256) void ProjSolve(ProjArgs a)
{
    __shared__ double sums[kProjTerms];
    int const m = static_cast<int>(blockIdx.x);
    int const wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
    int const a0 = a.meshAggBegin[m], a1 = a.meshAggBegin[m + 1];
    // term u (in use: both basis indices below k) summed by wave u mod 4
    {
        int u = 0;
#pragma unroll
        for (int p = 0; p < kProjMax; ++p)
#pragma unroll
            for (int q = p; q < kProjMax; ++q, ++u)
                if ((u & 3) == wave)
                {
                    double const v = q < a.k ? MeshSum(a.part + static_cast<size_t>(u) * a.nAgg, a0, a1, lane) : 0.;
                    if (lane == 0)
                        sums[u] = v;
                }
#pragma unroll
        for (int p = 0; p < kProjMax; ++p, ++u)
            if ((u & 3) == wave)
            {
                double const v = p < a.k ? MeshSum(a.part + static_cast<size_t>(u) * a.nAgg, a0, a1, lane) : 0.;
                if (lane == 0)
                    sums[u] = v;
            }
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    double sum[kProjTerms];
#pragma unroll
    for (int u = 0; u < kProjTerms; ++u)
        sum[u] = sums[u];

    int const k = a.k;
    double G[kProjMax][kProjMax], f[kProjMax], sc[kProjMax], L[kProjMax][kProjMax], d[kProjMax], c[kProjMax];
    bool use[kProjMax];
    {
        int u = 0;
        for (int p = 0; p < kProjMax; ++p)
            for (int q = p; q < kProjMax; ++q, ++u)
                G[p][q] = G[q][p] = sum[u];
        for (int p = 0; p < kProjMax; ++p)
            f[p] = sum[kProjPairs + p];
    }
    for (int p = 0; p < kProjMax; ++p)
    {
        use[p] = p < k && G[p][p] > 0.;
        sc[p] = use[p] ? 1. / sqrt(G[p][p]) : 0.;
    }
    for (int p = 0; p < kProjMax; ++p)
    {
        d[p] = 0.;
        for (int q = 0; q < kProjMax; ++q)
            L[p][q] = 0.;
    }
    for (int p = 0; p < kProjMax; ++p)
    {
        if (!use[p])
            continue;
        double dp = G[p][p] * sc[p] * sc[p];
        for (int q = 0; q < p; ++q)
            if (use[q])
                dp -= L[p][q] * L[p][q] * d[q];
        if (!(dp > 1e-6))
        {
            use[p] = false;
            continue;
        }
        d[p] = dp;
        for (int r = p + 1; r < kProjMax; ++r)
        {
            if (!use[r])
                continue;
            double v = G[r][p] * sc[r] * sc[p];
            for (int q = 0; q < p; ++q)
                if (use[q])
                    v -= L[r][q] * L[p][q] * d[q];
            L[r][p] = v / dp;
        }
    }
    // L D L^T c~ = f~ over the kept vectors
    for (int p = 0; p < kProjMax; ++p)
    {
        double v = use[p] ? f[p] * sc[p] : 0.;
        for (int q = 0; q < p; ++q)
            v -= L[p][q] * c[q];
        c[p] = use[p] ? v : 0.;
    }
    for (int p = 0; p < kProjMax; ++p)
        c[p] = use[p] ? c[p] / d[p] : 0.;
    for (int p = kProjMax - 1; p >= 0; --p)
    {
        double v = c[p];
        for (int q = p + 1; q < kProjMax; ++q)
            v -= L[q][p] * c[q];
        c[p] = use[p] ? v : 0.;
    }
    for (int p = 0; p < kProjMax; ++p)
        a.coef[static_cast<size_t>(m) * kProjMax + p] = c[p] * sc[p];
}

__global__ __launch_bounds__(256) void ProjCombine(ProjArgs a)
{
    int const i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.n)
        return;
    double const* c = a.coef + static_cast<size_t>(a.meshOf[i]) * kProjMax;
    double x = 0., y = 0., z = 0.;
#pragma unroll
    for (int j = 0; j < kProjMax; ++j)
    {
        if (j >= a.k)
            break;
        float4 const b = a.basis[j][i];
        x += c[j] * b.x, y += c[j] * b.y, z += c[j] * b.z;
    }
    a.y[i] = make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z), 0.f);
}

// End of an iteration of the fused CG, one workgroup per mesh: the mesh's
// partials summed in a fixed order (one sum per wave), r . z = the aggregates'
// r . D^-1 r + the coarse rows' rc . zc, then the stopping test and beta
// (FinishMesh). Converged meshes are skipped; the host counts live meshes from
// their done flags.
struct CgFinishArgs
{
    CgState cg;
    int const* __restrict__ meshAggBegin;
    double const* __restrict__ rzPart;
    double const* __restrict__ rrPart;
    double const* __restrict__ bbPart;
    double const* __restrict__ rczcPart; // nullable (block Jacobi)
    int* __restrict__ doneOut;           // nullable: host-mapped copy of the done flags
    int* __restrict__ aggDone;           // per aggregate: its mesh's done flag
};

__global__ __launch_bounds__(256) void CgFinish(CgFinishArgs f, int first)
{
    __shared__ double part[4];
    int const nM = f.cg.nM;
    int const m = static_cast<int>(blockIdx.x);
    int const wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
    if (!first && f.cg.F[kDone * nM + m] != 0)
        return;
    int const a0 = f.meshAggBegin[m], a1 = f.meshAggBegin[m + 1];
    // one sum per wave, all in flight together: r . D^-1 r, r . r, rc . zc, b . b
    double v = 0.;
    if (wave == 0)
        v = StridedSum<64>(f.rzPart, a0, a1, lane);
    else if (wave == 1)
        v = StridedSum<64>(f.rrPart, a0, a1, lane);
    else if (wave == 2 && f.rczcPart)
        v = StridedSum<64, 32>(f.rczcPart, kCoarseDofs * a0, kCoarseDofs * a1, lane);
    else if (wave == 3 && first)
        v = StridedSum<64>(f.bbPart, a0, a1, lane);
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    if (lane == 0)
        part[wave] = v;
    __syncthreads();
    __shared__ int doneS;
    if (threadIdx.x == 0)
    {
        bool const live = FinishMesh(f.cg, m, part[1], part[0] + part[2], part[3], first);
        if (f.doneOut)
            f.doneOut[m] = live ? 0 : 1;
        doneS = live ? 0 : 1;
    }
    __syncthreads();
    for (int g = a0 + static_cast<int>(threadIdx.x); g < a1; g += static_cast<int>(blockDim.x))
        f.aggDone[g] = doneS;
}

// Rotation offsets of the coarse space, one workgroup of 128 per aggregate g
// (rows aggRows[aggPtr[g] .. aggPtr[g + 1]), positions xs of their vertices):
// c the centroid, s the RMS radius, coarseD[row] = ((x - c) / s, 1), or 0 when
// the aggregate has no extent or its vertices are nearly collinear (the sum of
// the 2x2 principal minors of sum (x - c)(x - c)^T below 1e-3 of its trace
// squared). All sums in a fixed order.
__global__ __launch_bounds__(128) void CoarseBasis(int const* __restrict__ aggPtr, int const* __restrict__ aggRows,
                                                   int const* __restrict__ rowVertex,
                                                   float4 const* __restrict__ xs, float4* __restrict__ coarseD)
{
    __shared__ double lds3[2][3];
    __shared__ double lds6[2][6];
    int const g = static_cast<int>(blockIdx.x), t = static_cast<int>(threadIdx.x);
    int const kb = aggPtr[g], ke = aggPtr[g + 1];
    double c[3] = {0., 0., 0.};
    for (int k = kb + t; k < ke; k += 128)
    {
        float4 const x = xs[rowVertex[aggRows[k]]];
        c[0] += x.x, c[1] += x.y, c[2] += x.z;
    }
    Sum128<3>(c, lds3);
    double const n = static_cast<double>(ke - kb);
    for (double& v : c)
        v /= n;
    double S[6] = {0., 0., 0., 0., 0., 0.}; // xx, yy, zz, xy, xz, yz
    for (int k = kb + t; k < ke; k += 128)
    {
        float4 const x = xs[rowVertex[aggRows[k]]];
        double const dx = x.x - c[0], dy = x.y - c[1], dz = x.z - c[2];
        S[0] += dx * dx, S[1] += dy * dy, S[2] += dz * dz;
        S[3] += dx * dy, S[4] += dx * dz, S[5] += dy * dz;
    }
    Sum128<6>(S, lds6);
    double const tr = S[0] + S[1] + S[2];
    double const minors = S[0] * S[1] - S[3] * S[3] + S[0] * S[2] - S[4] * S[4] + S[1] * S[2] - S[5] * S[5];
    bool const use = tr > 0. && minors >= 1e-3 * tr * tr;
    double const inv = use ? 1. / sqrt(tr / n) : 0.;
    for (int k = kb + t; k < ke; k += 128)
    {
        int const row = aggRows[k];
        float4 const x = xs[rowVertex[row]];
        coarseD[row] = use ? make_float4(static_cast<float>((x.x - c[0]) * inv), static_cast<float>((x.y - c[1]) * inv),
                                         static_cast<float>((x.z - c[2]) * inv), 1.f)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Galerkin coarse operator Z^T H Z from the assembled blocks: coarse block k
// (aggregates cbA[k], cbB[k] of one mesh) is the sum over the fine blocks H_ij
// with i in cbA[k] and j in cbB[k] of Z_i^T H_ij Z_j, with D = [d]_x:
//   [ H        -H D_j      ]
//   [ D_i H    -D_i H D_j  ]
// One wave per coarse block: lane l takes the group's fine blocks l, l + 64, ...
// in ascending order, then a fixed butterfly; the wave writes the 6x6 entries
// into the mesh's dense row-major matrix (zeroed beforehand), so the operator
// is the same on every run. An aggregate without rotations (d = 0) gets an
// identity rotation diagonal, decoupled from everything else.
struct CoarseFromBsrArgs
{
    int const* __restrict__ cbA;
    int const* __restrict__ cbB;
    int const* __restrict__ cbSrcPtr;
    int const* __restrict__ cbSrc;
    int nCoarseBlocks;
    float4 const* __restrict__ va;
    float4 const* __restrict__ vb;
    float const* __restrict__ vc;
    int const* __restrict__ blockRow;
    int const* __restrict__ col;
    float4 const* __restrict__ coarseD;
    CoarseArgs c;
    double* __restrict__ Ac;
};

__device__ __forceinline__ void Skew(float4 const& d, double (&D)[3][3])
{
    D[0][0] = 0., D[0][1] = -d.z, D[0][2] = d.y;
    D[1][0] = d.z, D[1][1] = 0., D[1][2] = -d.x;
    D[2][0] = -d.y, D[2][1] = d.x, D[2][2] = 0.;
}

__global__ __launch_bounds__(256) void CoarseFromBsr(CoarseFromBsrArgs a)
{
    constexpr int kN = kCoarseDofs;
    int const wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
    int const k = static_cast<int>(blockIdx.x) * 4 + wave;
    if (k >= a.nCoarseBlocks)
        return;
    double s[kN * kN];
#pragma unroll
    for (int e = 0; e < kN * kN; ++e)
        s[e] = 0.;
    int const q0 = a.cbSrcPtr[k], q1 = a.cbSrcPtr[k + 1];
    for (int q = q0 + lane; q < q1; q += 64)
    {
        int const f = a.cbSrc[q];
        float4 const A = a.va[f], B = a.vb[f];
        float const C = a.vc[f];
        float4 const di = a.coarseD[a.blockRow[f]], dj = a.coarseD[a.col[f]];
        double const H[3][3] = {{A.x, A.y, A.z}, {A.w, B.x, B.y}, {B.z, B.w, C}};
        double Di[3][3], Dj[3][3], HD[3][3], DH[3][3];
        Skew(di, Di);
        Skew(dj, Dj);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
            {
                HD[r][cc] = H[r][0] * Dj[0][cc] + H[r][1] * Dj[1][cc] + H[r][2] * Dj[2][cc];
                DH[r][cc] = Di[r][0] * H[0][cc] + Di[r][1] * H[1][cc] + Di[r][2] * H[2][cc];
            }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
            {
                s[r * kN + cc] += H[r][cc];
                s[r * kN + 3 + cc] -= HD[r][cc];
                s[(3 + r) * kN + cc] += DH[r][cc];
                s[(3 + r) * kN + 3 + cc] -= Di[r][0] * HD[0][cc] + Di[r][1] * HD[1][cc] + Di[r][2] * HD[2][cc];
            }
    }
#pragma unroll
    for (int e = 0; e < kN * kN; ++e)
        for (int off = 32; off > 0; off >>= 1)
            s[e] += __shfl_xor(s[e], off);
    if (lane != 0)
        return;
    int const ga = a.cbA[k], gb = a.cbB[k];
    if (ga == gb && a.coarseD[a.blockRow[a.cbSrc[q0]]].w == 0.f)
#pragma unroll
        for (int r = 3; r < kN; ++r)
#pragma unroll
            for (int cc = 3; cc < kN; ++cc)
                s[r * kN + cc] = r == cc ? 1. : 0.;
    int const m = a.c.aggMesh[ga];
    int const a0 = a.c.meshAggBegin[m];
    int const n = kN * (a.c.meshAggBegin[m + 1] - a0);
    double* const Am = a.Ac + a.c.meshOffset[m];
    int const ra = kN * (ga - a0), cb = kN * (gb - a0);
    for (int r = 0; r < kN; ++r)
        for (int cc = 0; cc < kN; ++cc)
            Am[static_cast<size_t>(ra + r) * n + cb + cc] = s[r * kN + cc];
}

} // namespace pbat_vbd::device
