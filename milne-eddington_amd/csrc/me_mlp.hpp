// me_mlp.hpp — learned initialiser: MLP 4*nl -> 256 -> 256 -> 256 -> 10
// (GELU), BF16 operands on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation, FP32 master weights and Adam.  Trained on profiles the
// device map generator synthesises from the truth ranges; its prediction
// seeds one LM run per pixel (init_mode = ME_INIT_LEARNED).
//
// Normalisation: inputs are I/Imax - 1 and 10*Q/Imax, 10*U/Imax, 10*V/Imax
// (4*nl values, padded to a multiple of 32); outputs are the truths mapped to
// [0,1] by their ranges, except phi (sin 2phi, cos 2phi) and S (S/Imax).
#pragma once

#include <hip/hip_bf16.h>

#include <cstring>
#include <vector>

#include "../../include/meinv.h"
#include "me_device.hpp"

namespace me {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int MLP_H = 256;      // hidden width
constexpr int MLP_OUT = 10;     // 9 parameters with phi as (sin 2phi, cos 2phi)
constexpr int MLP_OUTP = 16;    // output width padded for the MFMA tile
constexpr int MLP_NL = 4;       // layers
constexpr int MLP_KSPLIT = 32;  // batch (K) chunks of a weight-gradient GEMM

__device__ __forceinline__ __bf16 to_bf16(float x) { return __float2bfloat16(x); }
__device__ __forceinline__ float from_bf16(__bf16 x) { return __bfloat162float(x); }

// GELU in its tanh form, z * sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)): one
// exp2 and one reciprocal, within 5e-4 of z Phi(z) everywhere (activations
// are then rounded to bfloat16).  For large |z| the exponential saturates to
// 0 or +inf and the result to z or -0.
constexpr float kGeluA = 2.3022082f;    // 2 sqrt(2/pi) log2(e)
constexpr float kGeluB = 0.10294324f;   // kGeluA * 0.044715
__device__ __forceinline__ float gelu(float z)
{
    const float u = z * fmaf(-kGeluB, z * z, -kGeluA);   // -log2(e) * argument of the sigmoid
    return z * __builtin_amdgcn_rcpf(1.0f + exp2f(u));
}
// its derivative: s + z s (1 - s) k (1 + 3 * 0.044715 z^2), s the sigmoid,
// k = 2 sqrt(2/pi)
__device__ __forceinline__ float gelu_grad(float z)
{
    const float z2 = z * z;
    const float s = __builtin_amdgcn_rcpf(1.0f + exp2f(z * fmaf(-kGeluB, z2, -kGeluA)));
    return s + z * s * (1.0f - s) * (1.5957691f * fmaf(0.134145f, z2, 1.0f));
}

enum { EPI_FWD_GELU = 0, EPI_FWD_LINEAR = 1, EPI_BWD_GELU = 2, EPI_F32 = 3 };

// C[M][N] = A[M][K] * B[K][N] on bf16 operands, FP32 accumulation.
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]: transposes are strides.
// 256 threads compute a 64x64 tile; each wave owns 16 rows and 4 MFMA
// 16x16 tiles; K advances 32 at a time through LDS (M, N and K are
// bounds-checked; out-of-range operands load as zero).  Epilogues:
//   EPI_FWD_GELU    Z = C + bias (fp32 out), Y = gelu(Z) (bf16 out)
//   EPI_FWD_LINEAR  Z = C + bias (fp32 out)
//   EPI_BWD_GELU    Y = C * gelu'(aux) (bf16 out; aux = pre-activation fp32)
//   EPI_F32         Z[z] = C * scale, the partial product of K chunk z =
//                   blockIdx.z (K split over gridDim.z); Z[z] starts at
//                   z * zs, and sum_partials adds them in chunk order
// Operand layouts of the staging loads: LAY_ANY reads element by element
// through the strides; LAY_K: the 8 consecutive k of a row (A) or column (B)
// are contiguous in memory; LAY_MN: 8 consecutive m (A) or n (B) of one k are
// contiguous.  The contiguous forms move 16 bytes per load (the caller
// guarantees 16-byte alignment) and fall back to single elements on a tile's
// ragged edge.
enum { LAY_ANY = 0, LAY_K = 1, LAY_MN = 2 };

// A thread's share of one 64x32 operand tile: 8 values, loaded from global
// memory (tile_load) and written to the LDS tile (tile_store).  LAY_K and
// LAY_ANY: row r = tid / 4 (m or n), k = 8 (tid % 4) .. +7; LAY_MN: k = tid / 8,
// rows 8 (tid % 8) .. +7.
template <int LAY>
__device__ __forceinline__ bf16x8 tile_load(const __bf16 *__restrict__ P, int64_t s_mn, int64_t s_k, int mn0, int MN,
                                            int k0, int kend, int tid)
{
    bf16x8 v;
    if (LAY == LAY_K) {
        const int mn = mn0 + (tid >> 2), k = k0 + (tid & 3) * 8;
        if (mn < MN && k + 7 < kend) {
            v = *(const bf16x8 *)(P + (int64_t)mn * s_mn + k);
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++)
                v[q] = (mn < MN && k + q < kend) ? P[(int64_t)mn * s_mn + (int64_t)(k + q) * s_k] : to_bf16(0.f);
        }
    } else if (LAY == LAY_MN) {
        const int mn = mn0 + (tid & 7) * 8, k = k0 + (tid >> 3);
        if (mn + 7 < MN && k < kend) {
            v = *(const bf16x8 *)(P + (int64_t)k * s_k + mn);
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++)
                v[q] = (mn + q < MN && k < kend) ? P[(int64_t)(mn + q) * s_mn + (int64_t)k * s_k] : to_bf16(0.f);
        }
    } else {
        const int mn = mn0 + (tid >> 2), k = k0 + (tid & 3) * 8;
#pragma unroll
        for (int q = 0; q < 8; q++)
            v[q] = (mn < MN && k + q < kend) ? P[(int64_t)mn * s_mn + (int64_t)(k + q) * s_k] : to_bf16(0.f);
    }
    return v;
}

template <int LAY>
__device__ __forceinline__ void tile_store(__bf16 (*sT)[40], const bf16x8 &v, int tid)
{
    if (LAY == LAY_MN) {
        const int kk = tid >> 3, r = (tid & 7) * 8;
#pragma unroll
        for (int q = 0; q < 8; q++) sT[r + q][kk] = v[q];
    } else {
        *(bf16x8 *)&sT[tid >> 2][(tid & 3) * 8] = v;
    }
}

// k slabs of 32 staged per barrier pair: every global load of the KS slabs
// is in flight at once, and a GEMM with K <= 32 KS takes one round trip
constexpr int GEMM_KS = 4;

template <int EPI, int ALAY = LAY_ANY, int BLAY = LAY_ANY>
__global__ __launch_bounds__(256) void gemm_bf16(const __bf16 *__restrict__ A, int64_t sam, int64_t sak,
                                                 const __bf16 *__restrict__ B, int64_t sbk, int64_t sbn, int M, int N,
                                                 int K, float *__restrict__ Zf, __bf16 *__restrict__ Yb, int ldc,
                                                 const float *__restrict__ bias, const float *__restrict__ aux,
                                                 float scale, int64_t zs)
{
    __shared__ __attribute__((aligned(16))) __bf16 sA[GEMM_KS][64][32 + 8];
    __shared__ __attribute__((aligned(16))) __bf16 sB[GEMM_KS][64][32 + 8];  // stored transposed: sB[n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    // split K over gridDim.z (weight gradients: K is the batch)
    const int kchunk = ((K + 32 * gridDim.z - 1) / (32 * gridDim.z)) * 32;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = K < kbeg + kchunk ? K : kbeg + kchunk;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += 32 * GEMM_KS) {
        // GEMM_KS 64x32 tiles of A and B^T (a slab past kend is skipped, as
        // the whole block sees); the MFMAs then run in k order, so every
        // output is the same chain of products whatever GEMM_KS is
        bf16x8 va[GEMM_KS], vb[GEMM_KS];
#pragma unroll
        for (int q = 0; q < GEMM_KS; q++) {
            if (k0 + 32 * q < kend) {
                va[q] = tile_load<ALAY>(A, sam, sak, m0, M, k0 + 32 * q, kend, tid);
                vb[q] = tile_load<BLAY>(B, sbn, sbk, n0, N, k0 + 32 * q, kend, tid);
            }
        }
#pragma unroll
        for (int q = 0; q < GEMM_KS; q++) {
            if (k0 + 32 * q < kend) {
                tile_store<ALAY>(sA[q], va[q], tid);
                tile_store<BLAY>(sB[q], vb[q], tid);
            }
        }
        __syncthreads();
        // lane l: row/col (l % 16), k group 8*(l / 16) .. +7
        const int rr = lane & 15, kg = (lane >> 4) * 8;
#pragma unroll
        for (int q = 0; q < GEMM_KS; q++) {
            if (k0 + 32 * q < kend) {
                const bf16x8 a = *(const bf16x8 *)&sA[q][wave * 16 + rr][kg];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const bf16x8 b = *(const bf16x8 *)&sB[q][t * 16 + rr][kg];
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D layout: lane l holds column (l % 16), rows 4*(l/16) .. +3
    const int col = lane & 15, rb = (lane >> 4) * 4;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = n0 + t * 16 + col;
        if (n >= N) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int m = m0 + wave * 16 + rb + v;
            if (m >= M) continue;
            const float c = acc[t][v];
            const int64_t o = (int64_t)m * ldc + n;
            if (EPI == EPI_FWD_GELU) {
                const float z = c + bias[n];
                Zf[o] = z;
                Yb[o] = to_bf16(gelu(z));
            } else if (EPI == EPI_FWD_LINEAR) {
                Zf[o] = c + bias[n];
            } else if (EPI == EPI_BWD_GELU) {
                Yb[o] = to_bf16(c * gelu_grad(aux[o]));
            } else {
                Zf[(int64_t)blockIdx.z * zs + o] = c * scale;  // this K chunk's partial
            }
        }
    }
}

// profiles -> normalised bf16 inputs [n][nin_pad]: one wave per pixel (the
// I maximum by a wave reduction, each lane writes every 64th input), so the
// loads and the bf16 stores are coalesced whatever nl is
constexpr int MLP_IN_PX = 4;   // pixels per 256-thread block
template <typename TIn>
__global__ __launch_bounds__(256) void mlp_inputs(const TIn *__restrict__ prof, int64_t n, int nl, int nin_pad,
                                                  __bf16 *__restrict__ X)
{
    const int64_t p = (int64_t)blockIdx.x * MLP_IN_PX + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= n) return;  // wave-uniform
    const TIn *pr = prof + p * 4 * nl;
    float imax = 0.f;
    for (int i = lane; i < nl; i += 64) imax = fmaxf(imax, (float)pr[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) imax = fmaxf(imax, __shfl_xor(imax, off));
    const float inv = imax > 0.f ? 1.0f / imax : 0.f;
    for (int i = lane; i < nin_pad; i += 64) {
        float x = 0.f;
        if (i < nl) x = (float)pr[i] * inv - 1.0f;
        else if (i < 4 * nl) x = 10.0f * (float)pr[i] * inv;
        if (!isfinite(x)) x = 0.f;
        X[p * nin_pad + i] = to_bf16(x);
    }
}

struct MlpNorm {
    double lo[9], hi[9];
};

// truths -> normalised targets [n][16] (10 used)
__global__ void mlp_targets(const double *__restrict__ truth, const float *__restrict__ prof, int64_t n, int nl,
                            MlpNorm nm, float *__restrict__ Y)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double *t = truth + p * 9;
    float imax = 0.f;
    for (int i = 0; i < nl; i++) imax = fmaxf(imax, prof[p * 4 * nl + i]);
    float *y = Y + p * MLP_OUTP;
    int o = 0;
    for (int k = 0; k < 9; k++) {
        if (k == 2) {
            y[o++] = (float)sin(2.0 * t[2]);
            y[o++] = (float)cos(2.0 * t[2]);
        } else if (k == 7) {
            y[o++] = imax > 0.f ? (float)(t[7] / imax) : 0.f;
        } else {
            y[o++] = (float)((t[k] - nm.lo[k]) / (nm.hi[k] - nm.lo[k]));
        }
    }
    for (; o < MLP_OUTP; o++) y[o] = 0.f;
}

// network outputs -> starting points [n][9] (double)
template <typename TIn>
__global__ void mlp_decode(const float *__restrict__ Z, const TIn *__restrict__ prof, int64_t n, int nl, MlpNorm nm,
                           double *__restrict__ init)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float *z = Z + p * MLP_OUTP;
    float imax = 0.f;
    for (int i = 0; i < nl; i++) imax = fmaxf(imax, (float)prof[p * 4 * nl + i]);
    double *o = init + p * 9;
    int j = 0;
    for (int k = 0; k < 9; k++) {
        if (k == 2) {
            double s = z[j++], c = z[j++];
            double ph = 0.5 * atan2(s, c);
            if (ph < 0) ph += kPi;
            o[2] = ph;
        } else if (k == 7) {
            o[7] = (double)z[j++] * imax;
        } else {
            double u = z[j++];
            u = u < 0 ? 0 : (u > 1 ? 1 : u);
            o[k] = nm.lo[k] + u * (nm.hi[k] - nm.lo[k]);
        }
    }
}

// loss = mean over rows and the 10 outputs of (Z - Y)^2; dZ = 2 (Z - Y) / (n*10).
// Each wave writes its share of the loss to loss_part[wave] (4 waves per
// block of 256 rows); mlp_loss_total adds them in wave order
__global__ void mlp_loss(const float *__restrict__ Z, const float *__restrict__ Y, int n, __bf16 *__restrict__ dZ,
                         float *__restrict__ loss_part)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < n) {
        for (int o = 0; o < MLP_OUTP; o++) {
            float d = 0.f;
            if (o < MLP_OUT) {
                d = Z[p * MLP_OUTP + o] - Y[p * MLP_OUTP + o];
                l += d * d;
            }
            dZ[p * MLP_OUTP + o] = to_bf16(2.0f * d / (float)(n * MLP_OUT));
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) loss_part[blockIdx.x * 4 + (threadIdx.x >> 6)] = l / (float)(n * MLP_OUT);
}

// column sums of a bf16 [n][w] matrix (bias gradients): COL_SUM_G waves =
// 64 columns x COL_SUM_G row groups per block; blockIdx.y selects a chunk of
// rows and its partial column sums go to part[chunk * zs + column] (a chunk
// past the last row writes zeros).  Wave g takes rows g, g + COL_SUM_G, ...
// of the chunk and the group partials are added in wave order; sum_partials
// then adds the chunks in order, so the bias gradients come out the same in
// every run
constexpr int COL_SUM_G = 16;
__global__ __launch_bounds__(64 * COL_SUM_G) void col_sum(const __bf16 *__restrict__ D, int n, int w,
                                                          int rows_per_chunk, float *__restrict__ part, int64_t zs)
{
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int g = threadIdx.x >> 6;
    const int r0 = blockIdx.y * rows_per_chunk;
    const int r1 = n < r0 + rows_per_chunk ? n : r0 + rows_per_chunk;
    __shared__ float ps[COL_SUM_G][64];
    float s = 0.f;
    if (c < w)
        for (int r = r0 + g; r < r1; r += COL_SUM_G) s += from_bf16(D[(int64_t)r * w + c]);
    ps[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0 && c < w) {
        float t = ps[0][threadIdx.x];
        for (int k = 1; k < COL_SUM_G; k++) t += ps[k][threadIdx.x];
        part[(int64_t)blockIdx.y * zs + c] = t;
    }
}

// out[i] = sum over z of part[z][i], z = 0 .. nz-1 in order
__global__ __launch_bounds__(256) void sum_partials(const float *__restrict__ part, int nz, int64_t n,
                                                    float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = part[i];
    for (int z = 1; z < nz; z++) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// Adam on FP32 masters, then the bf16 copy used by the GEMMs.  The gradient
// of entry i is the sum of its nz chunk partials part[z][i], added in chunk
// order exactly as sum_partials does (one launch fewer per training step)
__global__ void adam_step(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                          const float *__restrict__ part, int nz, __bf16 *__restrict__ pb, int64_t n, float lr, float b1,
                          float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = part[i];
    for (int z = 1; z < nz; z++) gi += part[(int64_t)z * n + i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = to_bf16(pi);
}

__global__ void to_bf16_kernel(const float *__restrict__ x, __bf16 *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// ------------------------------------------------ fused inference kernel
// All four layers in one kernel: each wave keeps MI_C tiles of 16 pixels as
// MFMA B operands (k x pixels) in registers, a layer's weights are staged in
// LDS as A-operand fragments (out channels x k), and the D fragments of one
// layer become the next layer's B fragments in place.  D gives a lane 4
// consecutive output channels per 16-channel tile; the next layer reads its
// k index in that order, so the packed weights of layers 1..3 permute their
// input rows to match (mlp_in_channel) and no data moves between lanes.
// Two tiles per wave and eight waves per block (two per SIMD: 256 VGPRs
// each, no spills) run faster than four tiles and four waves (one per SIMD),
// whose MFMA, LDS and GELU phases do not overlap.
constexpr int MI_C = 2;                  // pixel tiles (16 each) per wave
constexpr int MI_W = 8;                  // waves per block
constexpr int MI_T = 64 * MI_W;          // threads per block
constexpr int MI_PX = MI_W * 16 * MI_C;  // pixels per block pass

// input channel that k element q of k-block kb, lane group g reads
__device__ __forceinline__ int mlp_in_channel(int layer, int kb, int g, int q)
{
    return layer == 0 ? 32 * kb + 8 * g + q : 16 * (2 * kb + (q >> 2)) + 4 * g + (q & 3);
}

struct MlpPack {
    int64_t offW[MLP_NL], offB[MLP_NL];   // masters (Pb / P)
    int din[MLP_NL], dout[MLP_NL];
    int T[MLP_NL], KB[MLP_NL];            // 16-row tiles and 32-wide k blocks
    int64_t frag[MLP_NL];                 // first fragment of each layer in wpk
};

// one thread per (layer fragment, lane): 8 bf16 of the A fragment
__global__ void mlp_pack_kernel(const __bf16 *__restrict__ Pb, const float *__restrict__ P, MlpPack d, int64_t total,
                                bf16x8 *__restrict__ wpk, float *__restrict__ bpk)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < MLP_NL * 256) {
        const int l = (int)(f / 256), o = (int)(f % 256);
        bpk[f] = o < d.dout[l] ? P[d.offB[l] + o] : 0.f;
    }
    if (f >= total * 64) return;
    const int64_t fr = f / 64;
    const int lane = (int)(f % 64);
    int l = MLP_NL - 1;
    while (l > 0 && fr < d.frag[l]) l--;
    const int64_t r = fr - d.frag[l];
    const int t = (int)(r / d.KB[l]), kb = (int)(r % d.KB[l]);
    const int o = 16 * t + (lane & 15), g = lane >> 4;
    bf16x8 v;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int i = mlp_in_channel(l, kb, g, q);
        v[q] = (i < d.din[l] && o < d.dout[l]) ? Pb[d.offW[l] + (int64_t)i * d.dout[l] + o] : to_bf16(0.f);
    }
    wpk[f] = v;
}

// one hidden layer (16 output tiles of 16 channels, GELU) for the wave's
// MI_C pixel tiles; KB (k blocks of 32) is a compile-time constant so the
// MFMA chain of a tile has no branches between its steps
template <int KB>
__device__ __forceinline__ void mlp_hidden_layer(const bf16x8 *__restrict__ sW, const float *__restrict__ sb, int lane,
                                                 int g, bf16x8 (&act)[MI_C][8])
{
    bf16x8 nxt[MI_C][8];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        // one output tile at a time: keeps the scheduler from hoisting every
        // tile's LDS reads into registers
        __builtin_amdgcn_sched_barrier(0);
        // the accumulators start at the bias (a lane's four rows are output
        // channels 16 t + 4 g .. + 3 for every pixel column)
        const f32x4 b = f32x4{sb[16 * t + 4 * g], sb[16 * t + 4 * g + 1], sb[16 * t + 4 * g + 2], sb[16 * t + 4 * g + 3]};
        f32x4 acc[MI_C];
#pragma unroll
        for (int c = 0; c < MI_C; c++) acc[c] = b;
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
            const bf16x8 a = sW[(t * KB + kb) * 64 + lane];
#pragma unroll
            for (int c = 0; c < MI_C; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, act[c][kb], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < MI_C; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) nxt[c][t >> 1][4 * (t & 1) + v] = to_bf16(gelu(acc[c][v]));
    }
#pragma unroll
    for (int c = 0; c < MI_C; c++)
#pragma unroll
        for (int kb = 0; kb < 8; kb++) act[c][kb] = nxt[c][kb];
}

// k blocks of the first layer: 4 nl inputs in blocks of 32, rounded up to a
// power of two (the packed weights are zero beyond the inputs)
inline int mlp_kb0_padded(int ninp)
{
    const int kb = ninp / 32;
    return kb <= 1 ? 1 : (kb <= 2 ? 2 : (kb <= 4 ? 4 : 8));
}

template <typename TIn>
__global__ __launch_bounds__(MI_T) void mlp_infer_kernel(const TIn *__restrict__ prof, int64_t n, int nl, int kb0,
                                                        const bf16x8 *__restrict__ wpk, MlpPack d,
                                                        const float *__restrict__ bpk, MlpNorm nm,
                                                        double *__restrict__ init)
{
    __shared__ bf16x8 sW[16 * 8 * 64];   // 128 KiB: the fragments of one layer
    __shared__ float sb[256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, col = lane & 15, g = lane >> 4;
    for (int64_t base = (int64_t)blockIdx.x * MI_PX; base < n; base += (int64_t)gridDim.x * MI_PX) {
        bf16x8 act[MI_C][8];
        float imax[MI_C];
        // layer-0 operands: normalised profile values, k = input index
#pragma unroll
        for (int c = 0; c < MI_C; c++) {
            const int64_t p = base + (wave * MI_C + c) * 16 + col;
            const bool ok = p < n;
            const TIn *pr = prof + (ok ? p : 0) * 4 * nl;
            float im = 0.f;
            if (ok)
                for (int i = 0; i < nl; i++) im = fmaxf(im, (float)pr[i]);
            imax[c] = im;
            const float inv = im > 0.f ? 1.0f / im : 0.f;
#pragma unroll
            for (int kb = 0; kb < 8; kb++) {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int i = 32 * kb + 8 * g + q;
                    float x = 0.f;
                    if (ok && kb < kb0) {
                        if (i < nl) x = (float)pr[i] * inv - 1.0f;
                        else if (i < 4 * nl) x = 10.0f * (float)pr[i] * inv;
                        if (!isfinite(x)) x = 0.f;
                    }
                    act[c][kb][q] = to_bf16(x);
                }
            }
        }
#pragma unroll 1
        for (int l = 0; l < MLP_NL; l++) {
            const int KB = d.KB[l], T = d.T[l];
            __syncthreads();
            const bf16x8 *src = wpk + d.frag[l] * 64;
            // the layer's fragments go straight from global memory to LDS
            // (global_load_lds: no VGPR round trip, every load of the wave in
            // flight at once); a wave copies 64 consecutive fragments per
            // instruction, and fragment counts are multiples of 64
            const int nfr = T * KB * 64;
            for (int i0 = wave * 64; i0 < nfr; i0 += MI_T)
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + i0 + lane),
                                                 (__attribute__((address_space(3))) void *)(sW + i0), 16, 0, 0);
            for (int i = tid; i < 256; i += MI_T) sb[i] = bpk[l * 256 + i];
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (l < MLP_NL - 1) {
                if (KB == 8) mlp_hidden_layer<8>(sW, sb, lane, g, act);
                else if (KB == 4) mlp_hidden_layer<4>(sW, sb, lane, g, act);
                else if (KB == 2) mlp_hidden_layer<2>(sW, sb, lane, g, act);
                else mlp_hidden_layer<1>(sW, sb, lane, g, act);
            } else {
                // output layer: 16 channels (10 used); lane (g, col) holds
                // channels 4g..4g+3 of pixel col; lanes of group 0 gather
                // and decode
#pragma unroll
                for (int c = 0; c < MI_C; c++) {
                    f32x4 acc = f32x4{sb[4 * g], sb[4 * g + 1], sb[4 * g + 2], sb[4 * g + 3]};   // starts at the bias
#pragma unroll
                    for (int kb = 0; kb < 8; kb++)
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sW[kb * 64 + lane], act[c][kb], acc, 0, 0, 0);
                    float z[16];
#pragma unroll
                    for (int j = 0; j < 12; j++) z[j] = __shfl(acc[j & 3], (j >> 2) * 16 + col);
                    const int64_t p = base + (wave * MI_C + c) * 16 + col;
                    if (g == 0 && p < n) {
                        double *o = init + p * 9;
                        int j = 0;
                        for (int k = 0; k < 9; k++) {
                            if (k == 2) {
                                const double sn = z[j++], cs = z[j++];
                                double ph = 0.5 * atan2(sn, cs);
                                if (ph < 0) ph += kPi;
                                o[2] = ph;
                            } else if (k == 7) {
                                o[7] = (double)z[j++] * imax[c];
                            } else {
                                double u = z[j++];
                                u = u < 0 ? 0 : (u > 1 ? 1 : u);
                                o[k] = nm.lo[k] + u * (nm.hi[k] - nm.lo[k]);
                            }
                        }
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- host side
struct MlpState {
    int nin = 0, ninp = 0, trained = 0;
    int dims[MLP_NL + 1] = {0, MLP_H, MLP_H, MLP_H, MLP_OUTP};
    // parameters: for each layer W [din][dout] then b [dout], FP32 masters
    float *P = nullptr, *Mo = nullptr, *Vo = nullptr, *G = nullptr;
    float *Gp = nullptr;   // [MLP_KSPLIT][np] chunk partials of the gradient
    __bf16 *Pb = nullptr;
    int64_t np = 0;
    int64_t offW[MLP_NL] = {0}, offB[MLP_NL] = {0};
    MlpNorm norm;
    // activations for up to `cap` rows
    int64_t cap = 0;
    __bf16 *X = nullptr, *A[MLP_NL] = {nullptr}, *D[MLP_NL] = {nullptr};
    float *Z[MLP_NL] = {nullptr}, *Y = nullptr, *loss = nullptr;
    float *prof_tmp = nullptr;
    double *truth_tmp = nullptr;
    // the training set's normalised inputs and targets, built once per training
    __bf16 *Xall = nullptr;
    float *Yall = nullptr;
    int64_t tmp_cap = 0;
    // fused-inference weights (fragment layout), rebuilt after parameter changes
    bf16x8 *wpk = nullptr;
    float *bpk = nullptr;
    MlpPack pack;
    int packed = 0;
};

// output decoding ranges of the 9 parameters: [lo, hi] per parameter, as the
// training sampler's truth ranges (an empty range is widened to width 1)
inline void mlp_set_ranges(MlpState &m, const double *ranges)
{
    for (int k = 0; k < 9; k++) {
        m.norm.lo[k] = ranges[2 * k];
        m.norm.hi[k] = ranges[2 * k + 1] > ranges[2 * k] ? ranges[2 * k + 1] : ranges[2 * k] + 1.0;
    }
}

// free a device buffer and forget it (a failing free at this point leaves
// nothing to recover)
template <typename T> inline void dev_release(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline int64_t mlp_num_params(const MlpState &m)
{
    // exported layout: only the real input rows (nin) and output columns (10)
    if (m.nin == 0) return 0;
    return (int64_t)m.nin * MLP_H + MLP_H + 2 * ((int64_t)MLP_H * MLP_H + MLP_H) + (int64_t)MLP_H * MLP_OUT + MLP_OUT;
}

inline void mlp_free(MlpState &m)
{
    dev_release(m.P); dev_release(m.Mo); dev_release(m.Vo); dev_release(m.G); dev_release(m.Gp); dev_release(m.Pb);
    dev_release(m.X); dev_release(m.Y); dev_release(m.loss); dev_release(m.prof_tmp); dev_release(m.truth_tmp);
    dev_release(m.Xall); dev_release(m.Yall);
    dev_release(m.wpk); dev_release(m.bpk);
    for (int l = 0; l < MLP_NL; l++) { dev_release(m.A[l]); dev_release(m.D[l]); dev_release(m.Z[l]); }
    m = MlpState();
}

inline void mlp_init(MlpState &m, int nin)
{
    m.nin = nin;
    m.ninp = (nin + 31) / 32 * 32;
    m.dims[0] = m.ninp;
}

inline me_status mlp_alloc_params(MlpState &m)
{
    if (m.P) return ME_OK;
    int64_t off = 0;
    for (int l = 0; l < MLP_NL; l++) {
        m.offW[l] = off;
        off += (int64_t)m.dims[l] * m.dims[l + 1];
        m.offB[l] = off;
        off += m.dims[l + 1];
    }
    m.np = off;
    hipError_t e = hipMalloc(&m.P, sizeof(float) * off);
    e = e ? e : hipMalloc(&m.Mo, sizeof(float) * off);
    e = e ? e : hipMalloc(&m.Vo, sizeof(float) * off);
    e = e ? e : hipMalloc(&m.G, sizeof(float) * off);
    e = e ? e : hipMalloc(&m.Gp, sizeof(float) * MLP_KSPLIT * off);
    e = e ? e : hipMalloc(&m.Pb, sizeof(__bf16) * off);
    return e == hipSuccess ? ME_OK : ME_ENOMEM;
}

// waves of one mlp_loss launch over a batch of rows; m.loss holds their
// partials followed by the total
inline int mlp_loss_waves(int64_t rows) { return (int)((rows + 255) / 256) * 4; }

inline me_status mlp_alloc_acts(MlpState &m, int64_t rows)
{
    if (rows <= m.cap) return ME_OK;
    m.cap = 0;
    dev_release(m.X); dev_release(m.Y); dev_release(m.loss);
    for (int l = 0; l < MLP_NL; l++) { dev_release(m.A[l]); dev_release(m.D[l]); dev_release(m.Z[l]); }
    hipError_t e = hipMalloc(&m.X, sizeof(__bf16) * rows * m.ninp);
    e = e ? e : hipMalloc(&m.loss, sizeof(float) * (mlp_loss_waves(rows) + 1));
    e = e ? e : hipMalloc(&m.Y, sizeof(float) * rows * MLP_OUTP);
    for (int l = 0; l < MLP_NL; l++) {
        const int w = m.dims[l + 1];
        e = e ? e : hipMalloc(&m.A[l], sizeof(__bf16) * rows * w);
        e = e ? e : hipMalloc(&m.D[l], sizeof(__bf16) * rows * w);
        e = e ? e : hipMalloc(&m.Z[l], sizeof(float) * rows * w);
    }
    if (e != hipSuccess) return ME_ENOMEM;
    m.cap = rows;
    return ME_OK;
}

// He-style uniform initialisation from a host LCG (deterministic)
inline me_status mlp_init_params(MlpState &m, hipStream_t s, uint64_t seed)
{
    std::vector<float> h(m.np, 0.f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    };
    for (int l = 0; l < MLP_NL; l++) {
        const int din = m.dims[l], dout = m.dims[l + 1];
        const int rin = l == 0 ? m.nin : din, rout = l == MLP_NL - 1 ? MLP_OUT : dout;
        const double a = sqrt(6.0 / rin);
        for (int i = 0; i < rin; i++)
            for (int j = 0; j < rout; j++) h[m.offW[l] + (int64_t)i * dout + j] = (float)(a * rnd() * (l == MLP_NL - 1 ? 0.5 : 1.0));
    }
    hipError_t e = hipMemcpyAsync(m.P, h.data(), sizeof(float) * m.np, hipMemcpyHostToDevice, s);
    e = e ? e : hipMemsetAsync(m.Mo, 0, sizeof(float) * m.np, s);
    e = e ? e : hipMemsetAsync(m.Vo, 0, sizeof(float) * m.np, s);
    if (e != hipSuccess) return ME_EDEVICE;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((m.np + 255) / 256), dim3(256), 0, s, m.P, m.Pb, m.np);
    return hipGetLastError() == hipSuccess ? ME_OK : ME_EDEVICE;
}

inline dim3 gemm_grid(int M, int N) { return dim3((M + 63) / 64, (N + 63) / 64); }

// forward pass over `rows` rows of X (m.X unless given)
inline void mlp_forward(MlpState &m, hipStream_t s, int rows, const __bf16 *X = nullptr)
{
    const __bf16 *in = X ? X : m.X;
    for (int l = 0; l < MLP_NL; l++) {
        const int din = m.dims[l], dout = m.dims[l + 1];
        const __bf16 *W = m.Pb + m.offW[l];
        const float *b = m.P + m.offB[l];
        if (l < MLP_NL - 1)
            hipLaunchKernelGGL((gemm_bf16<EPI_FWD_GELU, LAY_K, LAY_MN>), gemm_grid(rows, dout), dim3(256), 0, s, in, (int64_t)din,
                               (int64_t)1, W, (int64_t)dout, (int64_t)1, rows, dout, din, m.Z[l], m.A[l], dout, b,
                               (const float *)nullptr, 1.0f, (int64_t)0);
        else
            hipLaunchKernelGGL((gemm_bf16<EPI_FWD_LINEAR, LAY_K, LAY_MN>), gemm_grid(rows, dout), dim3(256), 0, s, in, (int64_t)din,
                               (int64_t)1, W, (int64_t)dout, (int64_t)1, rows, dout, din, m.Z[l], (__bf16 *)nullptr,
                               dout, b, (const float *)nullptr, 1.0f, (int64_t)0);
        in = m.A[l];
    }
}

// K chunks of a weight-gradient GEMM (and row chunks of the bias sums) for a
// batch of `rows`
inline int mlp_ksplit(int rows) { return rows >= 4096 ? MLP_KSPLIT : (rows >= 1024 ? 8 : 1); }

// backward pass (loss gradient already in m.D[last]); gradients into m.G, or
// with sum = false only their chunk partials into m.Gp (adam_step adds them)
inline void mlp_backward(MlpState &m, hipStream_t s, int rows, const __bf16 *X = nullptr, bool sum = true)
{
    if (!X) X = m.X;
    // every gradient entry is the sum of ksplit chunk partials (weights: the
    // K chunks of the GEMM; biases: row chunks), which sit in Gp at stride np
    // at the entry's own offset; one sum_partials adds them all in chunk
    // order at the end, so the gradient is the same in every run
    const int ksplit = mlp_ksplit(rows);
    for (int l = MLP_NL - 1; l >= 0; l--) {
        const int din = m.dims[l], dout = m.dims[l + 1];
        const __bf16 *act = l == 0 ? X : m.A[l - 1];
        // dW = act^T dZ  (M = din, N = dout, K = rows split over gridDim.z)
        dim3 g = gemm_grid(din, dout);
        g.z = ksplit;
        hipLaunchKernelGGL((gemm_bf16<EPI_F32, LAY_MN, LAY_MN>), g, dim3(256), 0, s, act, (int64_t)1, (int64_t)din,
                           (const __bf16 *)m.D[l], (int64_t)dout, (int64_t)1, din, dout, rows, m.Gp + m.offW[l],
                           (__bf16 *)nullptr, dout, (const float *)nullptr, (const float *)nullptr, 1.0f, m.np);
        // db = column sums of dZ over ksplit row chunks
        const int rchunk = (rows + ksplit - 1) / ksplit;
        hipLaunchKernelGGL(col_sum, dim3((dout + 63) / 64, ksplit), dim3(64 * COL_SUM_G), 0, s, (const __bf16 *)m.D[l], rows,
                           dout, rchunk, m.Gp + m.offB[l], m.np);
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W^T) * gelu'(Z_{l-1})
            hipLaunchKernelGGL((gemm_bf16<EPI_BWD_GELU, LAY_K, LAY_K>), gemm_grid(rows, din), dim3(256), 0, s,
                               (const __bf16 *)m.D[l], (int64_t)dout, (int64_t)1, (const __bf16 *)(m.Pb + m.offW[l]),
                               (int64_t)1, (int64_t)dout, rows, din, dout, (float *)nullptr, m.D[l - 1], din,
                               (const float *)nullptr, (const float *)m.Z[l - 1], 1.0f, (int64_t)0);
        }
    }
    if (sum)
        hipLaunchKernelGGL(sum_partials, dim3((unsigned)((m.np + 255) / 256)), dim3(256), 0, s, (const float *)m.Gp,
                           ksplit, m.np, m.G);
}

// the batch loss of the last mlp_loss launch (its wave partials added in
// order, so the same batch gives the same value in every run)
inline hipError_t mlp_loss_total(MlpState &m, hipStream_t s, int rows, float *host)
{
    const int nw = mlp_loss_waves(rows);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(64), 0, s, (const float *)m.loss, nw, (int64_t)1, m.loss + nw);
    hipError_t e = hipMemcpyAsync(host, m.loss + nw, sizeof(float), hipMemcpyDeviceToHost, s);
    return e ? e : hipStreamSynchronize(s);
}

inline me_status mlp_tmp(MlpState &m, int64_t rows, int nl)
{
    if (rows <= m.tmp_cap) return ME_OK;
    m.tmp_cap = 0;
    dev_release(m.prof_tmp);
    dev_release(m.truth_tmp);
    dev_release(m.Xall);
    dev_release(m.Yall);
    hipError_t e = hipMalloc(&m.prof_tmp, sizeof(float) * rows * 4 * nl);
    e = e ? e : hipMalloc(&m.truth_tmp, sizeof(double) * rows * 9);
    e = e ? e : hipMalloc(&m.Xall, sizeof(__bf16) * rows * m.ninp);
    e = e ? e : hipMalloc(&m.Yall, sizeof(float) * rows * MLP_OUTP);
    if (e != hipSuccess) return ME_ENOMEM;
    m.tmp_cap = rows;
    return ME_OK;
}

// Training: the dataset (n_train profiles + truths) is synthesised once by
// gen (the device map generator); epochs sweep it in batches.
template <typename GenFn>
me_status mlp_train_impl(MlpState &m, const DevCfg &C, hipStream_t s, uint64_t seed, const double *ranges,
                         int64_t n_train, int batch, int epochs, double lr, double *final_loss, GenFn &&gen)
{
    batch = (batch + 63) / 64 * 64;
    if (batch > n_train) batch = (int)((n_train / 64) * 64);
    if (batch < 64) return ME_EINVAL;
    const int nl = C.nl;
    me_status st = mlp_alloc_params(m);
    if (st != ME_OK) return st;
    if ((st = mlp_alloc_acts(m, batch)) != ME_OK) return st;
    if ((st = mlp_tmp(m, n_train, nl)) != ME_OK) return st;
    mlp_set_ranges(m, ranges);
    if ((st = mlp_init_params(m, s, seed)) != ME_OK) return st;
    if ((st = gen(m.prof_tmp, m.truth_tmp, n_train)) != ME_OK) return st;
    // normalised inputs and targets of the whole training set, once
    hipLaunchKernelGGL(mlp_inputs<float>, dim3((unsigned)((n_train + MLP_IN_PX - 1) / MLP_IN_PX)), dim3(256), 0, s,
                       (const float *)m.prof_tmp, n_train, nl, m.ninp, m.Xall);
    hipLaunchKernelGGL(mlp_targets, dim3((unsigned)((n_train + 255) / 256)), dim3(256), 0, s,
                       (const double *)m.truth_tmp, (const float *)m.prof_tmp, n_train, nl, m.norm, m.Yall);
    const int64_t nb = n_train / batch;
    const int64_t total = nb * epochs;
    float b1 = 0.9f, b2 = 0.999f;
    double lossv = 0;
    for (int64_t step = 0; step < total; step++) {
        const int64_t r0 = (step % nb) * batch;
        const __bf16 *Xb = m.Xall + r0 * m.ninp;
        mlp_forward(m, s, batch, Xb);
        hipLaunchKernelGGL(mlp_loss, dim3((batch + 255) / 256), dim3(256), 0, s, (const float *)m.Z[MLP_NL - 1],
                           (const float *)(m.Yall + r0 * MLP_OUTP), batch, m.D[MLP_NL - 1], m.loss);
        mlp_backward(m, s, batch, Xb, false);
        // cosine decay to 5% of lr
        const double frac = (double)step / (double)total;
        const float lrs = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(kPi * frac))));
        const float c1 = 1.f - powf(b1, (float)(step + 1)), c2 = 1.f - powf(b2, (float)(step + 1));
        hipLaunchKernelGGL(adam_step, dim3((m.np + 255) / 256), dim3(256), 0, s, m.P, m.Mo, m.Vo, (const float *)m.Gp,
                           mlp_ksplit(batch), m.Pb, m.np, lrs, b1, b2, 1e-8f, c1, c2);
        if (step == total - 1) {
            float lh = 0;
            if (mlp_loss_total(m, s, batch, &lh) != hipSuccess) return ME_EDEVICE;
            lossv = lh;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) return ME_EDEVICE;
    if (final_loss) *final_loss = lossv;
    m.trained = 1;
    m.packed = 0;
    return ME_OK;
}

// predictions for n device profiles, processed in chunks of the activation capacity
inline me_status mlp_predict(MlpState &m, const DevCfg &C, hipStream_t s, const void *prof, int prof_is_f64, int64_t n,
                             double *init)
{
    if (!m.trained) return ME_ENOTREADY;
    const int nl = C.nl;
    if (m.ninp <= 256) {
        if (!m.packed) {
            MlpPack &d = m.pack;
            int64_t fr = 0;
            for (int l = 0; l < MLP_NL; l++) {
                d.offW[l] = m.offW[l];
                d.offB[l] = m.offB[l];
                d.din[l] = m.dims[l];
                d.dout[l] = m.dims[l + 1];
                d.T[l] = m.dims[l + 1] / 16;
                d.KB[l] = l == 0 ? mlp_kb0_padded(m.dims[0]) : m.dims[l] / 32;
                d.frag[l] = fr;
                fr += (int64_t)d.T[l] * d.KB[l];
            }
            if (!m.wpk) {
                hipError_t e = hipMalloc(&m.wpk, sizeof(bf16x8) * 64 * fr);
                e = e ? e : hipMalloc(&m.bpk, sizeof(float) * MLP_NL * 256);
                if (e != hipSuccess) return ME_ENOMEM;
            }
            const int64_t work = fr * 64 > MLP_NL * 256 ? fr * 64 : MLP_NL * 256;
            hipLaunchKernelGGL(mlp_pack_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s,
                               (const __bf16 *)m.Pb, (const float *)m.P, d, fr, m.wpk, m.bpk);
            m.packed = 1;
        }
        int dev = 0, ncu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return ME_EDEVICE;
        const int64_t passes = (n + MI_PX - 1) / MI_PX;
        const unsigned grid = (unsigned)(passes < ncu ? passes : ncu);
        if (grid == 0) return ME_OK;
        if (prof_is_f64)
            hipLaunchKernelGGL(mlp_infer_kernel<double>, dim3(grid), dim3(MI_T), 0, s, (const double *)prof, n, nl,
                               m.ninp / 32, (const bf16x8 *)m.wpk, m.pack, (const float *)m.bpk, m.norm, init);
        else
            hipLaunchKernelGGL(mlp_infer_kernel<float>, dim3(grid), dim3(MI_T), 0, s, (const float *)prof, n, nl,
                               m.ninp / 32, (const bf16x8 *)m.wpk, m.pack, (const float *)m.bpk, m.norm, init);
        return hipGetLastError() == hipSuccess ? ME_OK : ME_EDEVICE;
    }
    const int64_t chunk = 1 << 20;
    me_status st = mlp_alloc_acts(m, n < chunk ? ((n + 63) / 64) * 64 : chunk);
    if (st != ME_OK) return st;
    for (int64_t r0 = 0; r0 < n; r0 += m.cap) {
        const int rows = (int)((n - r0) < m.cap ? (n - r0) : m.cap);
        if (prof_is_f64)
            hipLaunchKernelGGL(mlp_inputs<double>, dim3((rows + MLP_IN_PX - 1) / MLP_IN_PX), dim3(256), 0, s,
                               (const double *)prof + r0 * 4 * nl, (int64_t)rows, nl, m.ninp, m.X);
        else
            hipLaunchKernelGGL(mlp_inputs<float>, dim3((rows + MLP_IN_PX - 1) / MLP_IN_PX), dim3(256), 0, s,
                               (const float *)prof + r0 * 4 * nl, (int64_t)rows, nl, m.ninp, m.X);
        mlp_forward(m, s, rows);
        if (prof_is_f64)
            hipLaunchKernelGGL(mlp_decode<double>, dim3((rows + 255) / 256), dim3(256), 0, s,
                               (const float *)m.Z[MLP_NL - 1], (const double *)prof + r0 * 4 * nl, (int64_t)rows, nl,
                               m.norm, init + r0 * 9);
        else
            hipLaunchKernelGGL(mlp_decode<float>, dim3((rows + 255) / 256), dim3(256), 0, s,
                               (const float *)m.Z[MLP_NL - 1], (const float *)prof + r0 * 4 * nl, (int64_t)rows, nl,
                               m.norm, init + r0 * 9);
    }
    return hipGetLastError() == hipSuccess ? ME_OK : ME_EDEVICE;
}

// exported parameter order: per layer W (real rows x real cols, row-major) then b (real cols)
template <typename F> inline void mlp_for_each_exported(const MlpState &m, F &&f)
{
    int64_t e = 0;
    for (int l = 0; l < MLP_NL; l++) {
        const int din = m.dims[l], dout = m.dims[l + 1];
        const int rin = l == 0 ? m.nin : din, rout = l == MLP_NL - 1 ? MLP_OUT : dout;
        for (int i = 0; i < rin; i++)
            for (int j = 0; j < rout; j++) f(e++, m.offW[l] + (int64_t)i * dout + j);
        for (int j = 0; j < rout; j++) f(e++, m.offB[l] + j);
    }
}

inline me_status mlp_get_params(MlpState &m, hipStream_t s, float *out)
{
    if (!m.P) return ME_ENOTREADY;
    std::vector<float> h(m.np);
    if (hipMemcpyAsync(h.data(), m.P, sizeof(float) * m.np, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ME_EDEVICE;
    mlp_for_each_exported(m, [&](int64_t e, int64_t i) { out[e] = h[i]; });
    return ME_OK;
}

inline me_status mlp_set_params(MlpState &m, hipStream_t s, const float *in)
{
    me_status st = mlp_alloc_params(m);
    if (st != ME_OK) return st;
    std::vector<float> h(m.np, 0.f);
    mlp_for_each_exported(m, [&](int64_t e, int64_t i) { h[i] = in[e]; });
    if (hipMemcpyAsync(m.P, h.data(), sizeof(float) * m.np, hipMemcpyHostToDevice, s) != hipSuccess) return ME_EDEVICE;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((m.np + 255) / 256), dim3(256), 0, s, m.P, m.Pb, m.np);
    if (hipStreamSynchronize(s) != hipSuccess) return ME_EDEVICE;
    m.packed = 0;
    return ME_OK;
}

// a network trained elsewhere: its exported weights and the truth ranges it
// was trained on; the context's learned mode can use it at once
inline me_status mlp_import(MlpState &m, hipStream_t s, const float *params, const double *ranges)
{
    me_status st = mlp_set_params(m, s, params);
    if (st != ME_OK) return st;
    mlp_set_ranges(m, ranges);
    m.trained = 1;
    return ME_OK;
}

// one forward + backward pass on a host batch (inputs already normalised,
// [batch][4*nl]; targets [batch][10]); gradient in the exported order
inline me_status mlp_grad_host(MlpState &m, hipStream_t s, const float *x, const float *y, int batch, float *grad,
                               double *loss)
{
    if (!m.P) return ME_ENOTREADY;
    if (batch % 32) return ME_EINVAL;  // batch is the K of the weight-gradient GEMM
    me_status st = mlp_alloc_acts(m, batch);
    if (st != ME_OK) return st;
    std::vector<__bf16> hx((size_t)batch * m.ninp);
    std::vector<float> hy((size_t)batch * MLP_OUTP, 0.f);
    for (int r = 0; r < batch; r++) {
        for (int i = 0; i < m.ninp; i++) hx[(size_t)r * m.ninp + i] = __float2bfloat16(i < m.nin ? x[(size_t)r * m.nin + i] : 0.f);
        for (int o = 0; o < MLP_OUT; o++) hy[(size_t)r * MLP_OUTP + o] = y[(size_t)r * MLP_OUT + o];
    }
    hipError_t e = hipMemcpyAsync(m.X, hx.data(), sizeof(__bf16) * hx.size(), hipMemcpyHostToDevice, s);
    e = e ? e : hipMemcpyAsync(m.Y, hy.data(), sizeof(float) * hy.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return ME_EDEVICE;
    mlp_forward(m, s, batch);
    hipLaunchKernelGGL(mlp_loss, dim3((batch + 255) / 256), dim3(256), 0, s, (const float *)m.Z[MLP_NL - 1],
                       (const float *)m.Y, batch, m.D[MLP_NL - 1], m.loss);
    mlp_backward(m, s, batch);
    std::vector<float> hg(m.np);
    float lh = 0;
    e = hipMemcpyAsync(hg.data(), m.G, sizeof(float) * m.np, hipMemcpyDeviceToHost, s);
    e = e ? e : mlp_loss_total(m, s, batch, &lh);
    if (e != hipSuccess) return ME_EDEVICE;
    mlp_for_each_exported(m, [&](int64_t ei, int64_t i) { grad[ei] = hg[i]; });
    if (loss) *loss = lh;
    return ME_OK;
}

}  // namespace me
