// HIP kernels of the learned surrogate for MI355X (gfx950): a message-passing network over
// the particle graph (edges between particles closer than the kernel support), BF16 weights
// and activations, FP32 accumulation on the matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Network (hidden width 128, L message-passing layers):
//   encoder   h0 = MLP_n(x_i)                e0 = MLP_e(u_ij)
//   layer l   e' = e + MLP([e, h_recv, h_send])      (edge update, residual)
//             a_i = mean over edges into i of e'      (aggregation, CSR by receiver)
//             h' = h + MLP([h, a])                    (node update, residual)
//   decoder   y = MLP_d(h_L)                          (per-particle state change)
// every MLP is Linear -> ReLU -> Linear.  The graph is symmetric, so the edges leaving node j
// are the reverses of the edges entering it, and gradients reach senders by a gather over
// reverse-edge indices instead of atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gnn_dev.h"

namespace psgnn {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// float -> bf16, round to nearest even (NaN stays NaN): one v_cvt_pk_bf16_f32 on gfx950
// (FP32 denormals are kept in these kernels, so the conversion rounds them too)
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------------------------------
// tiled GEMM on the matrix cores: C[m][n] = sum_k A(m, k) B(k, n), 128 x 128 output tile per
// workgroup of 4 waves (2 x 2 waves of 64 x 64, each 2 x 2 MFMA 32 x 32 tiles), K in steps of
// 32 staged through LDS.  A and B come from loader objects that return 8 consecutive k of a
// row (A) or of a column (B), so gathers and transposed operands share one kernel; the
// epilogue object consumes each FP32 result.  K is a multiple of 32 and rows of the loaded
// operands are padded to it.
// ------------------------------------------------------------------------------------------
constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;

template <class AL, class BL, class EP>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, AL al, BL bl, EP ep)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[BM*LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN*LDK];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x*BM, n0 = blockIdx.y*BN;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            int r = (tid >> 2) + 64*p, kc = (tid & 3)*8;
            uint4 va = (m0 + r < M) ? al.load8(m0 + r, k0 + kc) : make_uint4(0, 0, 0, 0);
            *(uint4 *)&As[r*LDK + kc] = va;
            uint4 vb = (n0 + r < N) ? bl.load8(n0 + r, k0 + kc) : make_uint4(0, 0, 0, 0);
            *(uint4 *)&Bs[r*LDK + kc] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                int ra = wm*64 + t*32 + (lane & 31);
                a[t] = *(const bf16x8 *)&As[ra*LDK + ks + (lane >> 5)*8];
                int rb = wn*64 + t*32 + (lane & 31);
                b[t] = *(const bf16x8 *)&Bs[rb*LDK + ks + (lane >> 5)*8];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // result (i, j, q) of lane l: row 8(q/4) + 4(l/32) + q%4, column l%32 of tile (i, j)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                int r = m0 + wm*64 + i*32 + 8*(q >> 2) + 4*(lane >> 5) + (q & 3);
                int c = n0 + wn*64 + j*32 + (lane & 31);
                if (r < M && c < N) ep.store(r, c, acc[i][j][q]);
            }
}

// Weight gradient dW[k][n] += sum over rows m in this workgroup's chunk of X(m, k) dY[m][n]:
// the reduction runs over rows, so both operands are staged transposed into LDS.  Partial
// sums of the row chunks meet in dW by FP32 atomic adds.
template <class XL>
__global__ void __launch_bounds__(256) k_gemm_wgrad(int M, int Kin, int chunk, XL xl, const uint16_t *dY,
                                                    float *dW, float *db)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[BM*LDK];   // [k][m]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN*LDK];   // [n][m]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int k0 = blockIdx.x*BM;
    const int mb = blockIdx.y*chunk, me = min(M, mb + chunk);
    f32x16 acc[2][2];
    float csum = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;
    for (int m0 = mb; m0 < me; m0 += BK) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            int mm = tid & 31, c8 = ((tid >> 5) + 8*p)*8;
            int m = m0 + mm;
            uint4 vx = make_uint4(0, 0, 0, 0), vy = make_uint4(0, 0, 0, 0);
            if (m < me) {
                if (k0 + c8 < Kin) vx = xl.load8(m, k0 + c8);
                vy = *(const uint4 *)(dY + (size_t)m*128 + c8);
            }
            const uint16_t *px = (const uint16_t *)&vx, *py = (const uint16_t *)&vy;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                As[(c8 + q)*LDK + mm] = px[q];
                Bs[(c8 + q)*LDK + mm] = py[q];
            }
        }
        __syncthreads();
        if (blockIdx.x == 0) {
            // the bias gradient: column sums of the staged dY rows (rows past the end are zeros)
            const uint4 *col = reinterpret_cast<const uint4 *>(&Bs[(tid & 127)*LDK + (tid >> 7)*16]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint4 v = col[h];
                const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
                for (int q = 0; q < 8; ++q) csum += bf2f(e[q]);
            }
        }
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                int ra = wm*64 + t*32 + (lane & 31);
                a[t] = *(const bf16x8 *)&As[ra*LDK + ks + (lane >> 5)*8];
                int rb = wn*64 + t*32 + (lane & 31);
                b[t] = *(const bf16x8 *)&Bs[rb*LDK + ks + (lane >> 5)*8];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                int r = k0 + wm*64 + i*32 + 8*(q >> 2) + 4*(lane >> 5) + (q & 3);
                int c = wn*64 + j*32 + (lane & 31);
                if (r < Kin) atomicAdd(&dW[(size_t)r*128 + c], acc[i][j][q]);
            }
    if (blockIdx.x == 0) atomicAdd(&db[tid & 127], csum);
}

// ---- operand loaders ----------------------------------------------------------------------
// rows of a dense row-major bf16 matrix with leading dimension ld
struct LdRows {
    const uint16_t *p; int ld;
    __device__ uint4 load8(int r, int k) const { return *(const uint4 *)(p + (size_t)r*ld + k); }
};
// edge-layer input [e | h_recv | h_send] (3 x 128 columns)
struct LdEdgeCat {
    const uint16_t *E, *H; const int32_t *recv, *send;
    __device__ uint4 load8(int r, int k) const
    {
        if (k < 128) return *(const uint4 *)(E + (size_t)r*128 + k);
        if (k < 256) return *(const uint4 *)(H + (size_t)recv[r]*128 + (k - 128));
        return *(const uint4 *)(H + (size_t)send[r]*128 + (k - 256));
    }
};
// node-layer input [h | a] (2 x 128 columns)
struct LdNodeCat {
    const uint16_t *H, *A;
    __device__ uint4 load8(int r, int k) const
    {
        if (k < 128) return *(const uint4 *)(H + (size_t)r*128 + k);
        return *(const uint4 *)(A + (size_t)r*128 + (k - 128));
    }
};

// ---- epilogues ----------------------------------------------------------------------------
struct EpBias {           // Y = bf16(act(acc + b))
    uint16_t *Y; int ld; const float *b; int relu;
    __device__ void store(int r, int c, float v) const
    {
        v += b[c];
        if (relu) v = v > 0.0f ? v : 0.0f;
        Y[(size_t)r*ld + c] = f2bf(v);
    }
};
struct EpResid {          // Y = bf16(R + acc + b)
    uint16_t *Y; const uint16_t *R; const float *b;
    __device__ void store(int r, int c, float v) const
    {
        size_t o = (size_t)r*128 + c;
        Y[o] = f2bf(bf2f(R[o]) + v + b[c]);
    }
};
struct EpOut {            // O = acc + b for the first `dout` columns (FP32)
    float *O; int dout; const float *b;
    __device__ void store(int r, int c, float v) const
    {
        if (c < dout) O[(size_t)r*dout + c] = v + b[c];
    }
};
struct EpMask {           // dZ = bf16(acc) where the ReLU output Z was positive, else 0
    uint16_t *dZ; const uint16_t *Z;
    __device__ void store(int r, int c, float v) const
    {
        size_t o = (size_t)r*128 + c;
        dZ[o] = (bf2f(Z[o]) > 0.0f) ? f2bf(v) : (uint16_t)0;
    }
};
struct EpF32 {            // D = acc (FP32), leading dimension ld
    float *D; int ld;
    __device__ void store(int r, int c, float v) const { D[(size_t)r*ld + c] = v; }
};

// ---- fused edge update (inference) -------------------------------------------------------
// e' = e + W2^T relu(W1^T [e | h_recv | h_send] + b1) + b2.  One workgroup per CU keeps both
// layers' weights in LDS for its whole life (W1^T 96 KB, W2^T 32 KB), and each of its four
// waves takes 32-edge tiles in turn, with no workgroup barrier after the weights are in:
//   - the tile's A operand (the 32 x 384 concatenated input) goes from memory straight into the
//     MFMA operand registers, one 16-B piece of one row per lane and k-step, and the next
//     tile's pieces are in flight while this tile is multiplied;
//   - layer 1 (K = 384) accumulates four 32 x 32 output tiles, its ReLU'd BF16 activations go
//     to the wave's own 8 KB of LDS, which is layer 2's A operand (K = 128);
//   - the residual input (the first 128 columns of the A operand, still in registers) is put
//     in the same LDS, added to the result there, and the tile leaves as row-contiguous 16-B
//     stores.
// Every weight fragment is read from LDS once per 32 edges and used by one MFMA; the 16-B
// chunks of each LDS row are XOR-swizzled by the row, so the 32 rows a fragment read touches
// spread over all banks.  Each output element sees the MFMAs of k_gemm<LdEdgeCat, ., EpBias>
// and k_gemm<LdRows, ., EpResid> in the same k order with the same operands, and the same
// bias, ReLU, rounding and residual additions, so the result is the same to the bit.
constexpr int kPwW1 = 128*384, kPwW2 = 128*128, kPwZ = 32*128;     // bf16 elements
constexpr int kPwLds = kPwW1 + kPwW2 + 4*kPwZ;                      // 160 KB

// element offset of 16-B chunk c of row r (rows of `len` elements)
__device__ __forceinline__ int pw_at(int r, int c, int len) { return r*len + ((c ^ (r & 15)) << 3); }

__global__ void __launch_bounds__(256) k_edge_mlp(int M, LdEdgeCat al, const uint16_t *W1t, const float *b1,
                                                  const uint16_t *W2t, const float *b2, uint16_t *Eout)
{
    __shared__ __attribute__((aligned(16))) uint16_t lds[kPwLds];
    uint16_t *W1s = lds, *W2s = lds + kPwW1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint16_t *Zs = lds + kPwW1 + kPwW2 + w*kPwZ;
    for (int e = tid; e < 128*48; e += 256) {
        const int n = e / 48, c = e - n*48;
        *(uint4 *)&W1s[pw_at(n, c, 384)] = *(const uint4 *)(W1t + (size_t)n*384 + c*8);
    }
    for (int e = tid; e < 128*16; e += 256) {
        const int n = e >> 4, c = e & 15;
        *(uint4 *)&W2s[pw_at(n, c, 128)] = *(const uint4 *)(W2t + (size_t)n*128 + c*8);
    }
    __syncthreads();
    const int l31 = lane & 31, h = lane >> 5;
    float bias1[4], bias2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { bias1[j] = b1[32*j + l31]; bias2[j] = b2[32*j + l31]; }
    const int ntiles = (M + 31) >> 5, stride = gridDim.x*4;
    // this lane's 16-B pieces of row l31 of a tile: k-step s covers k = 16 s + 8 h .. + 7.
    // Rows past the edges read the last edge (the loads stay unconditional, so that the waits
    // for them are exact); each output row depends on its own input row only, and such rows
    // are not stored.
    // The receiver and sender of the row are loaded one fetch ahead (ri, si: those of `tile`
    // on entry, of the tile after it on return), so no fetch waits on its own index loads.
    auto row_of = [&](int tile) { const int r = tile*32 + l31; return r < M ? r : M - 1; };
    auto fetch = [&](int tile, uint4 *A, int &ri, int &si) {
        const int r = row_of(tile);
        const uint16_t *pe = al.E + (size_t)r*128 + 8*h;
        const uint16_t *pr = al.H + (size_t)ri*128 + 8*h;
        const uint16_t *ps = al.H + (size_t)si*128 + 8*h;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            A[s] = *(const uint4 *)(pe + 16*s);
            A[8 + s] = *(const uint4 *)(pr + 16*s);
            A[16 + s] = *(const uint4 *)(ps + 16*s);
        }
        const int rn = row_of(tile + stride);
        ri = al.recv[rn];
        si = al.send[rn];
    };
    auto tile_pass = [&](int tile, const uint4 *A) {
        f32x16 acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[j][q] = 0.0f;
        // layer 1: K = 384; the weight fragments of step s + 1 are read while step s multiplies
        bf16x8 bw[2][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) bw[0][j] = *(const bf16x8 *)&W1s[pw_at(32*j + l31, h, 384)];
#pragma unroll
        for (int s = 0; s < 24; ++s) {
            if (s + 1 < 24) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bw[(s + 1) & 1][j] = *(const bf16x8 *)&W1s[pw_at(32*j + l31, 2*(s + 1) + h, 384)];
            }
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&A[s]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[s & 1][j], acc[j], 0, 0, 0);
        }
        // Z = bf16(relu(acc + b1)), row r = 8 (q / 4) + 4 h + q % 4, column c = 32 j + l31
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int r = 8*(q >> 2) + 4*h + (q & 3), c = 32*j + l31;
                float v = acc[j][q] + bias1[j];
                v = v > 0.0f ? v : 0.0f;
                Zs[pw_at(r, c >> 3, 128) + (c & 7)] = f2bf(v);
                acc[j][q] = 0.0f;
            }
        __builtin_amdgcn_wave_barrier();
        // layer 2: K = 128 from the wave's LDS, fragments one step ahead
        bf16x8 az[2];
        az[0] = *(const bf16x8 *)&Zs[pw_at(l31, h, 128)];
#pragma unroll
        for (int j = 0; j < 4; ++j) bw[0][j] = *(const bf16x8 *)&W2s[pw_at(32*j + l31, h, 128)];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s + 1 < 8) {
                az[(s + 1) & 1] = *(const bf16x8 *)&Zs[pw_at(l31, 2*(s + 1) + h, 128)];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    bw[(s + 1) & 1][j] = *(const bf16x8 *)&W2s[pw_at(32*j + l31, 2*(s + 1) + h, 128)];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az[s & 1], bw[s & 1][j], acc[j], 0, 0, 0);
        }
        __builtin_amdgcn_wave_barrier();
        // the residual input (columns 0..127 of the A operand) into the same LDS
#pragma unroll
        for (int s = 0; s < 8; ++s) *(uint4 *)&Zs[pw_at(l31, 2*s + h, 128)] = A[s];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int r = 8*(q >> 2) + 4*h + (q & 3), c = 32*j + l31;
                const int o = pw_at(r, c >> 3, 128) + (c & 7);
                Zs[o] = f2bf(bf2f(Zs[o]) + acc[j][q] + bias2[j]);
            }
        __builtin_amdgcn_wave_barrier();
    };
    // the tile's result rows from the wave's LDS to memory
    auto store = [&](int tile) {
        const int r = tile*32 + l31;
        if (r < M) {
#pragma unroll
            for (int s = 0; s < 8; ++s)
                *(uint4 *)(Eout + (size_t)r*128 + 16*s + 8*h) = *(const uint4 *)&Zs[pw_at(l31, 2*s + h, 128)];
        }
        __builtin_amdgcn_wave_barrier();
    };
    uint4 A0[24], A1[24];
    int tile = blockIdx.x*4 + w;
    int ri = al.recv[row_of(tile)], si = al.send[row_of(tile)];
    fetch(tile, A0, ri, si);
    // two tiles per round, so that each buffer keeps fixed registers; the scheduling barriers
    // keep the next tile's loads ahead of this tile's multiplications
    while (tile < ntiles) {
        fetch(tile + stride, A1, ri, si);
        __builtin_amdgcn_sched_barrier(0);
        tile_pass(tile, A0);
        store(tile);
        tile += stride;
        if (tile >= ntiles) break;
        fetch(tile + stride, A0, ri, si);
        __builtin_amdgcn_sched_barrier(0);
        tile_pass(tile, A1);
        store(tile);
        tile += stride;
    }
}

// ---- elementwise / graph kernels ----------------------------------------------------------

// a[i] = mean of E over the CSR segment of i (bf16 out, FP32 sums in edge order); one wave per
// node, eight edge rows in flight per wave (the loads of a group are issued before its adds)
__global__ void __launch_bounds__(256) k_segsum(int n, const int32_t *rp, const uint16_t *E, uint16_t *a)
{
    int i = blockIdx.x*4 + (threadIdx.x >> 6);
    if (i >= n) return;
    int lane = threadIdx.x & 63;
    float s0 = 0.0f, s1 = 0.0f;
    const uint16_t *col = E + 2*lane;
    int e = rp[i];
    const int e1 = rp[i + 1];
    for (; e + 8 <= e1; e += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = *(const uint32_t *)(col + (size_t)(e + k)*128);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            s0 += bf2f((uint16_t)(v[k] & 0xffffu));
            s1 += bf2f((uint16_t)(v[k] >> 16));
        }
    }
    for (; e < e1; ++e) {
        uint32_t v = *(const uint32_t *)(col + (size_t)e*128);
        s0 += bf2f((uint16_t)(v & 0xffffu));
        s1 += bf2f((uint16_t)(v >> 16));
    }
    int deg = rp[i + 1] - rp[i];
    float inv = 1.0f/(float)(deg > 0 ? deg : 1);
    s0 *= inv;
    s1 *= inv;
    uint32_t o = (uint32_t)f2bf(s0) | ((uint32_t)f2bf(s1) << 16);
    *(uint32_t *)(a + (size_t)i*128 + 2*lane) = o;
}

// gradient into an edge latent: g[e] = gnext[e] + da[recv[e]] / deg(recv[e]) (the mean's
// share; FP32, da with leading dimension lda), also as bf16
__global__ void __launch_bounds__(256) k_edge_grad(int ne, const int32_t *rp, const int32_t *recv,
                                                   const float *gnext, const float *da, int lda, float *g,
                                                   uint16_t *gb)
{
    size_t t = (size_t)blockIdx.x*256 + threadIdx.x;
    if (t >= (size_t)ne*128) return;
    size_t e = t >> 7;
    int c = (int)(t & 127);
    int r = recv[e];
    int deg = rp[r + 1] - rp[r];
    float v = (gnext ? gnext[t] : 0.0f) + da[(size_t)r*lda + c]/(float)(deg > 0 ? deg : 1);
    g[t] = v;
    gb[t] = f2bf(v);
}

// node gradient through an edge layer's input: dh[i] = dh_next[i] + dNode[i][0:128]
//   + sum over edges e into i of dEdge[e][128:256] + sum over edges e into i of
//   dEdge[rev[e]][256:384]  (the reverse edge of (i <- j) is (j <- i), whose sender is i)
__global__ void __launch_bounds__(256) k_node_grad(int n, const int32_t *rp, const int32_t *rev,
                                                   const float *dh_next, const float *dNode,
                                                   const float *dEdge, float *dh, uint16_t *dhb)
{
    int i = blockIdx.x*4 + (threadIdx.x >> 6);
    if (i >= n) return;
    int lane = threadIdx.x & 63;
    for (int h = 0; h < 2; ++h) {
        int c = lane + 64*h;
        float s = dh_next[(size_t)i*128 + c] + dNode[(size_t)i*256 + c];
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            s += dEdge[(size_t)e*384 + 128 + c];
            int r = rev[e];
            if (r >= 0) s += dEdge[(size_t)r*384 + 256 + c];
        }
        dh[(size_t)i*128 + c] = s;
        dhb[(size_t)i*128 + c] = f2bf(s);
    }
}

// split of an FP32 [M][ld] gradient: columns [c0, c0+128) added to g (FP32) and mirrored as bf16
__global__ void __launch_bounds__(256) k_add_cols(int M, int ld, int c0, const float *src, const float *base,
                                                  float *g, uint16_t *gb)
{
    size_t t = (size_t)blockIdx.x*256 + threadIdx.x;
    if (t >= (size_t)M*128) return;
    size_t m = t >> 7;
    int c = (int)(t & 127);
    float v = src[m*ld + c0 + c] + (base ? base[t] : 0.0f);
    g[t] = v;
    gb[t] = f2bf(v);
}

// mean-square loss over the first dout columns; dy = 2 (y - t) / count (bf16, zero-padded to 128)
__global__ void __launch_bounds__(256) k_mse(int n, int dout, const float *y, const float *tgt, const float *wgt,
                                             uint16_t *dy, float *loss, float inv_count)
{
    int i = blockIdx.x*2 + (threadIdx.x >> 7);
    int c = threadIdx.x & 127;
    float l = 0.0f;
    if (i < n) {
        float g = 0.0f;
        if (c < dout) {
            float w = wgt ? wgt[i] : 1.0f;
            float d = y[(size_t)i*dout + c] - tgt[(size_t)i*dout + c];
            l = w*d*d*inv_count;
            g = 2.0f*w*d*inv_count;
        }
        dy[(size_t)i*128 + c] = f2bf(g);
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, sh[0]);
}

// Adam step on FP32 master parameters (bias-corrected, step t >= 1)
__global__ void __launch_bounds__(256) k_adam(size_t n, float *p, const float *g, float *m, float *v,
                                              float lr, float b1, float b2, float eps, float c1, float c2)
{
    size_t i = (size_t)blockIdx.x*256 + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1*m[i] + (1.0f - b1)*gi;
    float vi = b2*v[i] + (1.0f - b2)*gi*gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr*(mi*c1)/(sqrtf(vi*c2) + eps);
}

// bf16 copies of one layer's weights: W [kin][128] and Wt [128][kpad] (zero-padded k >= kin)
__global__ void __launch_bounds__(256) k_weights_bf16(int kin, int kpad, const float *W, uint16_t *Wb, uint16_t *Wtb)
{
    int t = blockIdx.x*256 + threadIdx.x;
    if (t >= kpad*128) return;
    int k = t >> 7, n = t & 127;
    uint16_t b = (k < kin) ? f2bf(W[(size_t)k*128 + n]) : (uint16_t)0;
    Wb[(size_t)k*128 + n] = b;
    Wtb[(size_t)n*kpad + k] = b;
}

// ---- graph construction: edges between particles closer than R, CSR by receiver -----------
__device__ inline int gcell(const GraphGrid &g, double x, double y, double z, int *c3)
{
    double xs[3] = {x, y, z};
    int c[3] = {0, 0, 0};
    for (int d = 0; d < g.dim; ++d) {
        int v = (int)floor((xs[d] - g.lo[d])*g.inv_cell);
        v = v < 0 ? 0 : (v >= g.nc[d] ? g.nc[d] - 1 : v);
        c[d] = v;
    }
    if (c3) { c3[0] = c[0]; c3[1] = c[1]; c3[2] = c[2]; }
    return c[0] + g.nc[0]*(c[1] + g.nc[1]*c[2]);
}

__global__ void __launch_bounds__(256) k_g_bin(int n, GraphGrid g, const double *x, int32_t *key, int32_t *cnt)
{
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n) return;
    int k = gcell(g, x[3*i], x[3*i+1], x[3*i+2], nullptr);
    key[i] = k;
    atomicAdd(&cnt[k], 1);
}

// scatter into cell order (order inside a cell is fixed afterwards by k_g_cellorder)
__global__ void __launch_bounds__(256) k_g_scatter(int n, const int32_t *key, const int32_t *cstart,
                                                   int32_t *fill, int32_t *sorted)
{
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n) return;
    int k = key[i];
    sorted[cstart[k] + atomicAdd(&fill[k], 1)] = i;
}

// sort the particle indices of each cell ascending (cells hold few particles)
__global__ void __launch_bounds__(256) k_g_cellorder(int ncell, const int32_t *cstart, int32_t *sorted)
{
    int c = blockIdx.x*256 + threadIdx.x;
    if (c >= ncell) return;
    int b = cstart[c], e = cstart[c + 1];
    for (int i = b + 1; i < e; ++i) {
        int v = sorted[i], j = i - 1;
        while (j >= b && sorted[j] > v) { sorted[j + 1] = sorted[j]; --j; }
        sorted[j + 1] = v;
    }
}

// pass 0: count neighbours of each node; pass 1: write them (ascending index) at rp[i]
template <int PASS>
__global__ void __launch_bounds__(256) k_g_edges(int n, GraphGrid g, const double *x, const int32_t *cstart,
                                                 const int32_t *sorted, int32_t *deg, const int32_t *rp,
                                                 int32_t *cols)
{
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n) return;
    double xi = x[3*i], yi = x[3*i+1], zi = x[3*i+2];
    int c3[3];
    gcell(g, xi, yi, zi, c3);
    const double r2max = g.radius*g.radius;
    int cnt = 0;
    int base = PASS ? rp[i] : 0;
    int zlo = g.dim == 3 ? -1 : 0, zhi = g.dim == 3 ? 1 : 0;
    for (int dz = zlo; dz <= zhi; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int cx = c3[0] + dx, cy = c3[1] + dy, cz = c3[2] + dz;
                if (cx < 0 || cx >= g.nc[0] || cy < 0 || cy >= g.nc[1] || cz < 0 || cz >= g.nc[2]) continue;
                int cell = cx + g.nc[0]*(cy + g.nc[1]*cz);
                for (int s = cstart[cell]; s < cstart[cell + 1]; ++s) {
                    int j = sorted[s];
                    if (j == i) continue;
                    double a = xi - x[3*j], b = yi - x[3*j+1], cz2 = zi - x[3*j+2];
                    if (a*a + b*b + cz2*cz2 <= r2max) {
                        if (PASS) cols[base + cnt] = j;
                        ++cnt;
                    }
                }
            }
    if (!PASS) deg[i] = cnt;
    else {
        // ascending neighbour order inside the row (insertion sort, rows hold tens of entries)
        for (int a = base + 1; a < base + cnt; ++a) {
            int v = cols[a], b = a - 1;
            while (b >= base && cols[b] > v) { cols[b + 1] = cols[b]; --b; }
            cols[b + 1] = v;
        }
    }
}

// receiver/sender arrays and the reverse-edge index (binary search in the sender's row); one
// wave per node, its lanes taking the node's edges, so the searches run side by side
__global__ void __launch_bounds__(256) k_g_finish(int n, const int32_t *rp, const int32_t *cols,
                                                  int32_t *recv, int32_t *send, int32_t *rev)
{
    int i = blockIdx.x*4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int lane = threadIdx.x & 63;
    for (int e = rp[i] + lane; e < rp[i + 1]; e += 64) {
        int j = cols[e];
        recv[e] = i;
        send[e] = j;
        int lo = rp[j], hi = rp[j + 1] - 1, at = -1;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            int v = cols[mid];
            if (v == i) { at = mid; break; }
            if (v < i) lo = mid + 1; else hi = mid - 1;
        }
        rev[e] = at;
    }
}

// exclusive scan of int32 counts (single workgroup of 1024 threads, chunked)
__global__ void __launch_bounds__(1024) k_scan1(int n, const int32_t *in, int32_t *out)
{
    __shared__ int sh[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < n; b0 += 1024*8) {
        int v[8], s = 0;
        int base = b0 + threadIdx.x*8;
#pragma unroll
        for (int q = 0; q < 8; ++q) { v[q] = (base + q < n) ? in[base + q] : 0; s += v[q]; }
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int t = (threadIdx.x >= (unsigned)o) ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        int ex = sh[threadIdx.x] - s + carry;
#pragma unroll
        for (int q = 0; q < 8; ++q) { if (base + q < n) out[base + q] = ex; ex += v[q]; }
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// ---- features and state update ------------------------------------------------------------
// node inputs (padded to 32 columns): velocity / vs (dim), (rho/rho0 - 1) / ds, 1 for free
// particles, neighbour count / expected count, then zeros
__global__ void __launch_bounds__(256) k_node_in(int n, int dim, const double *v, const double *rho,
                                                 const int32_t *isfree, const int32_t *rp, FeatScale fs,
                                                 uint16_t *X)
{
    int t = blockIdx.x*256 + threadIdx.x;
    if (t >= n*32) return;
    int i = t >> 5, c = t & 31;
    float f = 0.0f;
    if (c < dim) f = (float)(v[3*i + c]*fs.inv_v);
    else if (c == dim) f = (float)((rho[i]/fs.rho0 - 1.0)*fs.inv_rho);
    else if (c == dim + 1) f = isfree[i] ? 1.0f : 0.0f;
    else if (c == dim + 2) f = (float)((rp[i + 1] - rp[i])*fs.inv_deg);
    X[t] = f2bf(f);
}

// edge inputs (padded to 32 columns): (x_recv - x_send) / R (dim), |x_recv - x_send| / R; one
// thread per edge, its 64-B row written as four 16-B stores
__global__ void __launch_bounds__(256) k_edge_in(int ne, int dim, const int32_t *recv, const int32_t *send,
                                                 const double *x, double inv_r, uint16_t *U)
{
    size_t e = (size_t)blockIdx.x*256 + threadIdx.x;
    if (e >= (size_t)ne) return;
    const int i = recv[e], j = send[e];
    double a[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d) a[d] = x[3*i + d] - x[3*j + d];
    double s = 0.0;
    for (int d = 0; d < dim; ++d) s += a[d]*a[d];
    uint32_t w[4] = {0u, 0u, 0u, 0u};          // columns 0..7 as bf16 pairs
    for (int d = 0; d < dim; ++d)
        w[d >> 1] |= (uint32_t)f2bf((float)(a[d]*inv_r)) << (16*(d & 1));
    w[dim >> 1] |= (uint32_t)f2bf((float)(sqrt(s)*inv_r)) << (16*(dim & 1));
    uint4 *row = reinterpret_cast<uint4 *>(U + e*32);
    row[0] = make_uint4(w[0], w[1], w[2], w[3]);
    row[1] = make_uint4(0u, 0u, 0u, 0u);
    row[2] = make_uint4(0u, 0u, 0u, 0u);
    row[3] = make_uint4(0u, 0u, 0u, 0u);
}

// state after k solver steps: x += dx, v += dv, rho += drho for free particles, from the
// network output y (normalised: dx = drift_t v + y[0:dim] sx, dv = y[dim:2dim] sv,
// drho = y[2dim] sr)
__global__ void __launch_bounds__(256) k_apply(int n, int dim, const float *y, const int32_t *isfree,
                                               FeatScale fs, double *x, double *v, double *rho)
{
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n || !isfree[i]) return;
    const float *yi = y + (size_t)i*(2*dim + 1);
    for (int d = 0; d < dim; ++d) {
        double v0 = v[3*i + d];
        x[3*i + d] += fs.drift_t*v0 + (double)yi[d]*fs.out_dx;
        v[3*i + d] = v0 + (double)yi[dim + d]*fs.out_dv;
    }
    rho[i] += (double)yi[2*dim]*fs.out_drho;
}

// targets for training: the normalised change between two states, position net of the drift
// drift_t v0 (free particles weighted 1)
__global__ void __launch_bounds__(256) k_targets(int n, int dim, const double *x0, const double *v0,
                                                 const double *r0, const double *x1, const double *v1,
                                                 const double *r1, const int32_t *isfree, FeatScale fs,
                                                 float *tgt, float *wgt)
{
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n) return;
    float *t = tgt + (size_t)i*(2*dim + 1);
    for (int d = 0; d < dim; ++d) {
        t[d] = (float)((x1[3*i + d] - x0[3*i + d] - fs.drift_t*v0[3*i + d])/fs.out_dx);
        t[dim + d] = (float)((v1[3*i + d] - v0[3*i + d])/fs.out_dv);
    }
    t[2*dim] = (float)((r1[i] - r0[i])/fs.out_drho);
    wgt[i] = isfree[i] ? 1.0f : 0.0f;
}

// ---- launchers ----------------------------------------------------------------------------
static inline unsigned nb(size_t n, int b) { return (unsigned)((n + b - 1)/b); }

void gemm_fwd_rows(hipStream_t s, int M, int K, const uint16_t *A, int lda, const uint16_t *Wt, const float *b,
                   int relu, uint16_t *Y)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdRows, LdRows, EpBias>), g, dim3(256), 0, s, M, 128, K, LdRows{A, lda},
                       LdRows{Wt, K}, EpBias{Y, 128, b, relu});
}
void gemm_fwd_edgecat(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                      const int32_t *send, const uint16_t *Wt, const float *b, uint16_t *Y)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdEdgeCat, LdRows, EpBias>), g, dim3(256), 0, s, M, 128, 384,
                       LdEdgeCat{E, H, recv, send}, LdRows{Wt, 384}, EpBias{Y, 128, b, 1});
}
void gemm_fwd_nodecat(hipStream_t s, int M, const uint16_t *H, const uint16_t *A, const uint16_t *Wt,
                      const float *b, uint16_t *Y)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdNodeCat, LdRows, EpBias>), g, dim3(256), 0, s, M, 128, 256,
                       LdNodeCat{H, A}, LdRows{Wt, 256}, EpBias{Y, 128, b, 1});
}
void edge_mlp_fused(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                    const int32_t *send, const uint16_t *W1t, const float *b1, const uint16_t *W2t,
                    const float *b2, uint16_t *Eout)
{
    // one workgroup per CU (each holds 160 KB of LDS), none idle on small graphs
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
    }
    const int tiles = (M + 31)/32;
    const int grid = std::max(1, std::min(ncu, (tiles + 3)/4));
    hipLaunchKernelGGL(k_edge_mlp, dim3(grid), dim3(256), 0, s, M, LdEdgeCat{E, H, recv, send}, W1t, b1,
                       W2t, b2, Eout);
}
void gemm_fwd_resid(hipStream_t s, int M, const uint16_t *Z, const uint16_t *Wt, const float *b,
                    const uint16_t *R, uint16_t *Y)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdRows, LdRows, EpResid>), g, dim3(256), 0, s, M, 128, 128, LdRows{Z, 128},
                       LdRows{Wt, 128}, EpResid{Y, R, b});
}
void gemm_fwd_out(hipStream_t s, int M, const uint16_t *Z, const uint16_t *Wt, const float *b, int dout, float *O)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdRows, LdRows, EpOut>), g, dim3(256), 0, s, M, 128, 128, LdRows{Z, 128},
                       LdRows{Wt, 128}, EpOut{O, dout, b});
}
// dZ = (dY W^T) masked by Z > 0, W [128][128]
void gemm_bwd_mask(hipStream_t s, int M, const uint16_t *dY, const uint16_t *W, const uint16_t *Z, uint16_t *dZ)
{
    dim3 g(nb(M, BM), 1);
    hipLaunchKernelGGL((k_gemm<LdRows, LdRows, EpMask>), g, dim3(256), 0, s, M, 128, 128, LdRows{dY, 128},
                       LdRows{W, 128}, EpMask{dZ, Z});
}
// dX = dY W^T (FP32 [M][kin]), W [kin][128]
void gemm_bwd_input(hipStream_t s, int M, int kin, const uint16_t *dY, const uint16_t *W, float *dX)
{
    dim3 g(nb(M, BM), nb(kin, BN));
    hipLaunchKernelGGL((k_gemm<LdRows, LdRows, EpF32>), g, dim3(256), 0, s, M, kin, 128, LdRows{dY, 128},
                       LdRows{W, 128}, EpF32{dX, kin});
}
template <class XL>
static void wgrad(hipStream_t s, int M, int kin, XL xl, const uint16_t *dY, float *dW, float *db)
{
    const int chunk = 2048;
    dim3 g(nb(kin, BM), nb(M, chunk));
    hipLaunchKernelGGL((k_gemm_wgrad<XL>), g, dim3(256), 0, s, M, kin, chunk, xl, dY, dW, db);
}
void gemm_wgrad_rows(hipStream_t s, int M, int kin, const uint16_t *X, int ldx, const uint16_t *dY, float *dW, float *db)
{ wgrad(s, M, kin, LdRows{X, ldx}, dY, dW, db); }
void gemm_wgrad_edgecat(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                        const int32_t *send, const uint16_t *dY, float *dW, float *db)
{ wgrad(s, M, 384, LdEdgeCat{E, H, recv, send}, dY, dW, db); }
void gemm_wgrad_nodecat(hipStream_t s, int M, const uint16_t *H, const uint16_t *A, const uint16_t *dY, float *dW, float *db)
{ wgrad(s, M, 256, LdNodeCat{H, A}, dY, dW, db); }

void segsum(hipStream_t s, int n, const int32_t *rp, const uint16_t *E, uint16_t *a)
{ hipLaunchKernelGGL(k_segsum, dim3(nb(n, 4)), dim3(256), 0, s, n, rp, E, a); }
void edge_grad(hipStream_t s, int ne, const int32_t *rp, const int32_t *recv, const float *gnext, const float *da,
               int lda, float *g, uint16_t *gb)
{ hipLaunchKernelGGL(k_edge_grad, dim3(nb((size_t)ne*128, 256)), dim3(256), 0, s, ne, rp, recv, gnext, da, lda, g, gb); }
void node_grad(hipStream_t s, int n, const int32_t *rp, const int32_t *rev, const float *dh_next, const float *dNode,
               const float *dEdge, float *dh, uint16_t *dhb)
{ hipLaunchKernelGGL(k_node_grad, dim3(nb(n, 4)), dim3(256), 0, s, n, rp, rev, dh_next, dNode, dEdge, dh, dhb); }
void add_cols(hipStream_t s, int M, int ld, int c0, const float *src, const float *base, float *g, uint16_t *gb)
{ hipLaunchKernelGGL(k_add_cols, dim3(nb((size_t)M*128, 256)), dim3(256), 0, s, M, ld, c0, src, base, g, gb); }
void mse(hipStream_t s, int n, int dout, const float *y, const float *tgt, const float *wgt, uint16_t *dy,
         float *loss, float inv_count)
{ hipLaunchKernelGGL(k_mse, dim3(nb(n, 2)), dim3(256), 0, s, n, dout, y, tgt, wgt, dy, loss, inv_count); }
void adam(hipStream_t s, size_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
          float eps, float c1, float c2)
{ hipLaunchKernelGGL(k_adam, dim3(nb(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1, c2); }
void weights_bf16(hipStream_t s, int kin, int kpad, const float *W, uint16_t *Wb, uint16_t *Wtb)
{ hipLaunchKernelGGL(k_weights_bf16, dim3(nb((size_t)kpad*128, 256)), dim3(256), 0, s, kin, kpad, W, Wb, Wtb); }

void graph_bin(hipStream_t s, int n, const GraphGrid &g, const double *x, int32_t *key, int32_t *cnt)
{ hipLaunchKernelGGL(k_g_bin, dim3(nb(n, 256)), dim3(256), 0, s, n, g, x, key, cnt); }
void graph_scatter(hipStream_t s, int n, const int32_t *key, const int32_t *cstart, int32_t *fill, int32_t *sorted)
{ hipLaunchKernelGGL(k_g_scatter, dim3(nb(n, 256)), dim3(256), 0, s, n, key, cstart, fill, sorted); }
void graph_cellorder(hipStream_t s, int ncell, const int32_t *cstart, int32_t *sorted)
{ hipLaunchKernelGGL(k_g_cellorder, dim3(nb(ncell, 256)), dim3(256), 0, s, ncell, cstart, sorted); }
void graph_edges(hipStream_t s, int pass, int n, const GraphGrid &g, const double *x, const int32_t *cstart,
                 const int32_t *sorted, int32_t *deg, const int32_t *rp, int32_t *cols)
{
    if (pass == 0)
        hipLaunchKernelGGL(k_g_edges<0>, dim3(nb(n, 256)), dim3(256), 0, s, n, g, x, cstart, sorted, deg, rp, cols);
    else
        hipLaunchKernelGGL(k_g_edges<1>, dim3(nb(n, 256)), dim3(256), 0, s, n, g, x, cstart, sorted, deg, rp, cols);
}
void graph_finish(hipStream_t s, int n, const int32_t *rp, const int32_t *cols, int32_t *recv, int32_t *send, int32_t *rev)
{ hipLaunchKernelGGL(k_g_finish, dim3(nb(n, 4)), dim3(256), 0, s, n, rp, cols, recv, send, rev); }
void scan1(hipStream_t s, int n, const int32_t *in, int32_t *out)
{ hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, n, in, out); }

void node_in(hipStream_t s, int n, int dim, const double *v, const double *rho, const int32_t *isfree,
             const int32_t *rp, const FeatScale &fs, uint16_t *X)
{ hipLaunchKernelGGL(k_node_in, dim3(nb((size_t)n*32, 256)), dim3(256), 0, s, n, dim, v, rho, isfree, rp, fs, X); }
void edge_in(hipStream_t s, int ne, int dim, const int32_t *recv, const int32_t *send, const double *x, double inv_r,
             uint16_t *U)
{ hipLaunchKernelGGL(k_edge_in, dim3(nb((size_t)ne, 256)), dim3(256), 0, s, ne, dim, recv, send, x, inv_r, U); }
void apply_out(hipStream_t s, int n, int dim, const float *y, const int32_t *isfree, const FeatScale &fs,
               double *x, double *v, double *rho)
{ hipLaunchKernelGGL(k_apply, dim3(nb(n, 256)), dim3(256), 0, s, n, dim, y, isfree, fs, x, v, rho); }
void targets(hipStream_t s, int n, int dim, const double *x0, const double *v0, const double *r0, const double *x1,
             const double *v1, const double *r1, const int32_t *isfree, const FeatScale &fs, float *tgt, float *wgt)
{ hipLaunchKernelGGL(k_targets, dim3(nb(n, 256)), dim3(256), 0, s, n, dim, x0, v0, r0, x1, v1, r1, isfree, fs, tgt, wgt); }

}  // namespace psgnn
