// rgbnet on the matrix cores (gfx950): Linear(+ReLU) layers of width 192 for the colour
// MLP of lib/voxurf_fine.py:132-149, forward and backward, FP32 (v_mfma_f32_32x32x2f32)
// and BF16 inputs with FP32 accumulation (v_mfma_f32_32x32x16_bf16).
//
// Layout trick.  A 32x32 MFMA leaves output element (row i, col j) in lane
// j + 32*((i % 8) / 4), register (i % 4) + 4*(i / 8).  The layers here compute
// H_out^T = W * H_in^T, i.e. the sample index is the MFMA column (one sample per lane
// pair) and the feature index the row, so after a layer each lane already holds 96 of
// its sample's 192 output features.  The next layer consumes them in place by
// permuting its reduction order: reduction step s of lane half h uses input feature
// dfeat(s / 16, s % 16, h), and the weights are packed once per call in exactly that
// order (and in lane order, so every LDS read is conflict-free).  Activations therefore
// never pass through LDS; only the weights do, loaded once per workgroup.
//
// Per 32-sample tile and layer: 6 output tiles x 96 reduction steps = 576 MFMAs (FP32)
// or 6 x 12 = 72 (BF16).  Weight gradients dW = dY^T X are a separate split-M GEMM
// through LDS (its reduction runs over samples).
#include "common.h"

namespace vx {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kWidth = 192;

// feature held by accumulator register v of output tile t in lane half h
__host__ __device__ constexpr int dfeat(int t, int v, int h) {
  return 32 * t + 8 * (v >> 2) + 4 * h + (v & 3);
}

// ---- weight packing --------------------------------------------------------------
// Packed A operand, FP32: element ((to * S/4 + s/4) * 64 + lane) * 4 + s%4 holds
//   fwd: A[i = 32 to + lane%32][k = dfeat(s/16, s%16, lane/32)] = Wt[i][k]
//   bwd: A[i = 32 to + lane%32][k = ...] = Wt[k][i]        (W transposed)
// with zero outside [0, n_out) x [0, k_in).  S = 16 * KT reduction steps.
__global__ void k_pack_f32(const float* __restrict__ Wt, int n_out, int k_in, int KT, int NTO,
                           int transpose, float* __restrict__ out) {
  const int S = 16 * KT;
  const int total = NTO * S * 64;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int lane = idx & 63, s = (idx >> 6) % S, to = (idx >> 6) / S;
    const int i = 32 * to + (lane & 31);
    const int k = dfeat(s >> 4, s & 15, lane >> 5);
    float v = 0.f;
    if (!transpose) {
      if (i < n_out && k < k_in) v = Wt[i * k_in + k];
    } else {
      // A[i][k] = Wt[k][i]: i indexes the layer's inputs (< k_in), k its outputs (< n_out)
      if (k < n_out && i < k_in) v = Wt[k * k_in + i];
    }
    out[(((to * (S / 4)) + (s >> 2)) * 64 + lane) * 4 + (s & 3)] = v;
  }
}

// Packed A operand, BF16 (K = 16 per MFMA, 8 values per lane): element
// ((to * S + s) * 64 + lane) * 8 + e holds A[i = 32 to + lane%32][k] with
// k = dfeat(s / 2, 8 (s % 2) + e, lane / 32), S = 2 * KT steps.
__global__ void k_pack_bf16(const float* __restrict__ Wt, int n_out, int k_in, int KT, int NTO,
                            int transpose, __bf16* __restrict__ out) {
  const int S = 2 * KT;
  const int total = NTO * S * 64 * 8;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int e = idx & 7, lane = (idx >> 3) & 63, s = (idx >> 9) % S, to = (idx >> 9) / S;
    const int i = 32 * to + (lane & 31);
    const int k = dfeat(s >> 1, 8 * (s & 1) + e, lane >> 5);
    float v = 0.f;
    if (!transpose) {
      if (i < n_out && k < k_in) v = Wt[i * k_in + k];
    } else {
      if (k < n_out && i < k_in) v = Wt[k * k_in + i];
    }
    out[idx] = (__bf16)v;
  }
}

// ---- one layer over all rows -------------------------------------------------------
// MODE 0: out = relu(W x + b)            (hidden layer, forward)
// MODE 1: as 0, plus logits = W3 out + b3 (last hidden layer, forward)
// MODE 2: out = (W^T dy) * [mask > 0]    (backward data through a ReLU layer)
// MODE 3: out = W^T dy                   (backward data to the network input)
// Input rows: `in` [M, in_stride] (float32 or bfloat16 per TIN), read at the permuted
// feature order.  Output rows [M, out_stride] in the same element type as TIN for
// activations (MODE 0/1) and float32 for gradients (MODE 2/3).
template <int KT, int NTO, int MODE, bool BF>
__global__ void __launch_bounds__(512) k_mlp_layer(
    const void* __restrict__ in_, int in_stride, int64_t M, const void* __restrict__ wpk,
    const float* __restrict__ bias, const void* __restrict__ mask_, int mask_stride,
    void* __restrict__ out_, int out_stride, const float* __restrict__ w3,
    const float* __restrict__ b3, float* __restrict__ logits) {
  constexpr int S = BF ? 2 * KT : 16 * KT;     // reduction steps
  constexpr int WELEMS = NTO * S * 64 * (BF ? 8 : 1);
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[WELEMS * (BF ? 2 : 4)];
  __shared__ float lb[NTO * 32];
  __shared__ float lw3[3 * kWidth];
  const int tid = threadIdx.x;
  {
    const float4* src = reinterpret_cast<const float4*>(wpk);
    float4* dst = reinterpret_cast<float4*>(lds_raw);
    constexpr int n4 = WELEMS * (BF ? 2 : 4) / 16;
    for (int i = tid; i < n4; i += blockDim.x) dst[i] = src[i];
    if (MODE <= 1)
      for (int i = tid; i < NTO * 32; i += blockDim.x) lb[i] = bias[i];
    if (MODE == 1)
      for (int i = tid; i < 3 * kWidth; i += blockDim.x) lw3[i] = w3[i];
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6, h = lane >> 5, rl = lane & 31;
  const int64_t ntiles = (M + 31) / 32;
  const int nwaves = blockDim.x >> 6;
  for (int64_t tile = (int64_t)blockIdx.x * nwaves + wave; tile < ntiles;
       tile += (int64_t)gridDim.x * nwaves) {
    const int64_t row = tile * 32 + rl;
    const bool valid = row < M;
    f32x16 acc[NTO];
#pragma unroll
    for (int t = 0; t < NTO; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
    if constexpr (!BF) {
      const float* in = reinterpret_cast<const float*>(in_);
      float b[S];
#pragma unroll
      for (int t = 0; t < KT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
          if (valid) f = *reinterpret_cast<const float4*>(in + row * in_stride + 32 * t + 8 * u + 4 * h);
          b[t * 16 + 4 * u + 0] = f.x;
          b[t * 16 + 4 * u + 1] = f.y;
          b[t * 16 + 4 * u + 2] = f.z;
          b[t * 16 + 4 * u + 3] = f.w;
        }
      const float4* lw = reinterpret_cast<const float4*>(lds_raw);
#pragma unroll
      for (int sq = 0; sq < S / 4; ++sq) {
#pragma unroll
        for (int to = 0; to < NTO; ++to) {
          const float4 a = lw[(to * (S / 4) + sq) * 64 + lane];
          acc[to] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[4 * sq + 0], acc[to], 0, 0, 0);
          acc[to] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[4 * sq + 1], acc[to], 0, 0, 0);
          acc[to] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[4 * sq + 2], acc[to], 0, 0, 0);
          acc[to] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b[4 * sq + 3], acc[to], 0, 0, 0);
        }
      }
    } else {
      // BF16: step s takes the 8 features dfeat(s/2, 8 (s%2) + e, h), e = 0..7, which
      // are 8 consecutive registers of the producing layer's output.
      bf16x8 b[S];
      if (MODE <= 1) {
        const __bf16* in = reinterpret_cast<const __bf16*>(in_);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          // features 32 t + 8 u + 4 h + w for u in {2 (s%2), 2 (s%2) + 1}, w = 0..3
          const int t = s >> 1, u0 = 2 * (s & 1);
          uint2 lo = make_uint2(0, 0), hi = make_uint2(0, 0);
          if (valid) {
            lo = *reinterpret_cast<const uint2*>(in + row * in_stride + 32 * t + 8 * u0 + 4 * h);
            hi = *reinterpret_cast<const uint2*>(in + row * in_stride + 32 * t + 8 * (u0 + 1) + 4 * h);
          }
          uint4 q = make_uint4(lo.x, lo.y, hi.x, hi.y);
          b[s] = *reinterpret_cast<bf16x8*>(&q);
        }
      } else {
        // backward: the incoming gradient rows are float32; round to bf16 for the MFMA
        const float* in = reinterpret_cast<const float*>(in_);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int t = s >> 1, u0 = 2 * (s & 1);
          float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
          if (valid) {
            lo = *reinterpret_cast<const float4*>(in + row * in_stride + 32 * t + 8 * u0 + 4 * h);
            hi = *reinterpret_cast<const float4*>(in + row * in_stride + 32 * t + 8 * (u0 + 1) + 4 * h);
          }
          b[s][0] = (__bf16)lo.x;
          b[s][1] = (__bf16)lo.y;
          b[s][2] = (__bf16)lo.z;
          b[s][3] = (__bf16)lo.w;
          b[s][4] = (__bf16)hi.x;
          b[s][5] = (__bf16)hi.y;
          b[s][6] = (__bf16)hi.z;
          b[s][7] = (__bf16)hi.w;
        }
      }
      const bf16x8* lw = reinterpret_cast<const bf16x8*>(lds_raw);
#pragma unroll
      for (int s = 0; s < S; ++s)
#pragma unroll
        for (int to = 0; to < NTO; ++to)
          acc[to] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lw[(to * S + s) * 64 + lane], b[s],
                                                             acc[to], 0, 0, 0);
    }
    // ---- epilogue ----
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int to = 0; to < NTO; ++to) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float v4[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const int v = 4 * u + w;
          float x = acc[to][v];
          if (MODE <= 1) {
            x += lb[dfeat(to, v, h)];
            x = x > 0.f ? x : 0.f;
          }
          v4[w] = x;
        }
        const int64_t col = 32 * to + 8 * u + 4 * h;
        if (MODE == 2 && valid) {
          if constexpr (BF) {
            const __bf16* mk = reinterpret_cast<const __bf16*>(mask_) + row * mask_stride + col;
            for (int w = 0; w < 4; ++w) v4[w] = (float)mk[w] > 0.f ? v4[w] : 0.f;
          } else {
            const float4 mk = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(mask_) +
                                                              row * mask_stride + col);
            v4[0] = mk.x > 0.f ? v4[0] : 0.f;
            v4[1] = mk.y > 0.f ? v4[1] : 0.f;
            v4[2] = mk.z > 0.f ? v4[2] : 0.f;
            v4[3] = mk.w > 0.f ? v4[3] : 0.f;
          }
        }
        if (MODE == 1) {
          // logits from the activations as stored (bf16-rounded in the BF16 path)
#pragma unroll
          for (int w = 0; w < 4; ++w) {
            const float a = BF ? (float)(__bf16)v4[w] : v4[w];
            const int f = (int)col + w;
#pragma unroll
            for (int c = 0; c < 3; ++c) p[c] += a * lw3[c * kWidth + f];
          }
        }
        if (valid) {
          if (BF && MODE <= 1) {
            __bf16* o = reinterpret_cast<__bf16*>(out_) + row * out_stride + col;
            __bf16 q[4] = {(__bf16)v4[0], (__bf16)v4[1], (__bf16)v4[2], (__bf16)v4[3]};
            *reinterpret_cast<uint2*>(o) = *reinterpret_cast<uint2*>(q);
          } else {
            float* o = reinterpret_cast<float*>(out_) + row * out_stride + col;
            *reinterpret_cast<float4*>(o) = make_float4(v4[0], v4[1], v4[2], v4[3]);
          }
        }
      }
    }
    if (MODE == 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) p[c] += __shfl_xor(p[c], 32, 64);
      if (h == 0 && valid)
        for (int c = 0; c < 3; ++c) logits[row * 3 + c] = p[c] + b3[c];
    }
  }
}

// dH3 = (dlogit W3) * [H3 > 0] (the K = 3 layer, on the vector ALUs).
template <bool BF>
__global__ void k_mlp_dh_last(const float* __restrict__ dlogit, const void* __restrict__ H_,
                              int h_stride, int64_t M, const float* __restrict__ w3,
                              float* __restrict__ dH, int dh_stride) {
  const int64_t total = M * kWidth;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / kWidth;
    const int f = (int)(i % kWidth);
    const float hv = BF ? (float)reinterpret_cast<const __bf16*>(H_)[r * h_stride + f]
                        : reinterpret_cast<const float*>(H_)[r * h_stride + f];
    const float g = dlogit[3 * r] * w3[f] + dlogit[3 * r + 1] * w3[kWidth + f] +
                    dlogit[3 * r + 2] * w3[2 * kWidth + f];
    dH[r * dh_stride + f] = hv > 0.f ? g : 0.f;
  }
}

// ---- weight gradients: part[blk] = dY[rows]^T X[rows], partb[blk] = sum dY[rows] --------
// dY [M, 192] float32, X [M, 32*KT] (float32 or bf16).  Each block takes a contiguous
// range of rows; output tiles (tn, tk) of 32x32 are spread over the 4 waves.
template <int KT, bool BF>
__global__ void __launch_bounds__(256) k_mlp_wgrad(const float* __restrict__ dY, int dys,
                                                   const void* __restrict__ X_, int xs, int64_t M,
                                                   int64_t rows_per_block, float* __restrict__ part,
                                                   float* __restrict__ partb) {
  constexpr int K = 32 * KT;
  constexpr int NTILE = 6 * KT;
  constexpr int NTW = (NTILE + 3) / 4;
  __shared__ float sY[32][kWidth];
  __shared__ float sX[32][K];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = lane >> 5, rl = lane & 31;
  f32x16 acc[NTW];
#pragma unroll
  for (int q = 0; q < NTW; ++q)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[q][v] = 0.f;
  float bacc = 0.f;
  const int64_t r_begin = blockIdx.x * rows_per_block;
  const int64_t r_end = (r_begin + rows_per_block < M) ? r_begin + rows_per_block : M;
  for (int64_t r0 = r_begin; r0 < r_end; r0 += 32) {
    for (int i = tid; i < 32 * kWidth; i += blockDim.x) {
      const int rr = i / kWidth, c = i % kWidth;
      sY[rr][c] = (r0 + rr < r_end) ? dY[(r0 + rr) * dys + c] : 0.f;
    }
    for (int i = tid; i < 32 * K; i += blockDim.x) {
      const int rr = i / K, c = i % K;
      float v = 0.f;
      if (r0 + rr < r_end)
        v = BF ? (float)reinterpret_cast<const __bf16*>(X_)[(r0 + rr) * xs + c]
               : reinterpret_cast<const float*>(X_)[(r0 + rr) * xs + c];
      sX[rr][c] = v;
    }
    __syncthreads();
    if (tid < kWidth)
      for (int rr = 0; rr < 32; ++rr) bacc += sY[rr][tid];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
#pragma unroll
      for (int q = 0; q < NTW; ++q) {
        const int tile = wave + 4 * q;
        if (tile < NTILE) {
          const int tn = tile / KT, tk = tile % KT;
          acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(sY[2 * j + h][32 * tn + rl],
                                                        sX[2 * j + h][32 * tk + rl], acc[q], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  float* pb = part + (int64_t)blockIdx.x * kWidth * K;
#pragma unroll
  for (int q = 0; q < NTW; ++q) {
    const int tile = wave + 4 * q;
    if (tile < NTILE) {
      const int tn = tile / KT, tk = tile % KT;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int n = 32 * tn + 8 * (v >> 2) + 4 * h + (v & 3);
        pb[n * K + 32 * tk + rl] = acc[q][v];
      }
    }
  }
  if (tid < kWidth) partb[(int64_t)blockIdx.x * kWidth + tid] = bacc;
}

// dW3 [3, 192] and db3 [3] partials over row blocks: part3[blk][c*192 + f], partb3[blk][c].
template <bool BF>
__global__ void __launch_bounds__(192) k_mlp_wgrad_last(const float* __restrict__ dlogit,
                                                        const void* __restrict__ H_, int hs,
                                                        int64_t M, int64_t rows_per_block,
                                                        float* __restrict__ part3,
                                                        float* __restrict__ partb3) {
  const int f = threadIdx.x;
  const int64_t r_begin = blockIdx.x * rows_per_block;
  const int64_t r_end = (r_begin + rows_per_block < M) ? r_begin + rows_per_block : M;
  float a0 = 0.f, a1 = 0.f, a2 = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
  for (int64_t r = r_begin; r < r_end; ++r) {
    const float hv = BF ? (float)reinterpret_cast<const __bf16*>(H_)[r * hs + f]
                        : reinterpret_cast<const float*>(H_)[r * hs + f];
    const float g0 = dlogit[3 * r], g1 = dlogit[3 * r + 1], g2 = dlogit[3 * r + 2];
    a0 += g0 * hv;
    a1 += g1 * hv;
    a2 += g2 * hv;
    c0 += g0;
    c1 += g1;
    c2 += g2;
  }
  float* pb = part3 + (int64_t)blockIdx.x * 3 * kWidth;
  pb[f] = a0;
  pb[kWidth + f] = a1;
  pb[2 * kWidth + f] = a2;
  if (f == 0) {
    partb3[3 * blockIdx.x] = c0;
    partb3[3 * blockIdx.x + 1] = c1;
    partb3[3 * blockIdx.x + 2] = c2;
  }
}

// out[i] = sum_b part[b * n + i]   (i < n_valid_cols per row: the first k_real columns of
// each row of width k_pad are kept)
__global__ void k_reduce_parts(const float* __restrict__ part, int nblk, int rows, int k_pad,
                               int k_real, float* __restrict__ out) {
  const int n = rows * k_real;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = i / k_real, c = i % k_real;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * rows * k_pad + r * k_pad + c];
    out[i] = s;
  }
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, int64_t n, __bf16* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (__bf16)x[i];
}

static int num_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

template <int KT, int NTO, int MODE, bool BF>
static void launch_layer(const void* in, int in_stride, int64_t M, const void* wpk,
                         const float* bias, const void* mask, int mask_stride, void* out,
                         int out_stride, const float* w3, const float* b3, float* logits,
                         hipStream_t st) {
  const int64_t tiles = (M + 31) / 32;
  int64_t blocks = (tiles + 7) / 8;
  const int64_t cap = num_cus();
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  k_mlp_layer<KT, NTO, MODE, BF><<<(unsigned)blocks, 512, 0, st>>>(
      in, in_stride, M, wpk, bias, mask, mask_stride, out, out_stride, w3, b3, logits);
}

}  // namespace vx

using namespace vx;

extern "C" {

int vx_mlp_pack(int bf16, const float* W, int n_out, int k_in, int kt, int nto, int transpose,
                void* out, void* stream) {
  VX_CHECK_ARG(kt >= 1 && nto >= 1 && n_out >= 1 && k_in >= 1, "mlp_pack: bad sizes");
  hipStream_t st = (hipStream_t)stream;
  if (bf16)
    k_pack_bf16<<<grid_for((int64_t)nto * kt * 2 * 64 * 8, 256), 256, 0, st>>>(
        W, n_out, k_in, kt, nto, transpose, (__bf16*)out);
  else
    k_pack_f32<<<grid_for((int64_t)nto * kt * 16 * 64, 256), 256, 0, st>>>(W, n_out, k_in, kt,
                                                                             nto, transpose,
                                                                             (float*)out);
  return launch_status("mlp_pack");
}

int64_t vx_mlp_pack_elems(int kt, int nto) { return (int64_t)nto * kt * 16 * 64; }

int vx_mlp_layer_fwd(int bf16, int kt, int last, const void* in, int in_stride, int64_t M,
                     const void* wpk, const float* bias, void* out, int out_stride,
                     const float* w3, const float* b3, float* logits, void* stream) {
  VX_CHECK_ARG(M >= 0, "mlp_layer_fwd: M < 0");
  VX_CHECK_ARG(kt == 2 || kt == 3 || kt == 6, "mlp_layer_fwd: input width must be 64, 96 or 192");
  if (M == 0) return VX_OK;
  hipStream_t st = (hipStream_t)stream;
#define L(KT, MODE, BF) launch_layer<KT, 6, MODE, BF>(in, in_stride, M, wpk, bias, nullptr, 0, out, out_stride, w3, b3, logits, st)
  if (!bf16) {
    if (kt == 2) { if (last) L(2, 1, false); else L(2, 0, false); }
    else if (kt == 3) { if (last) L(3, 1, false); else L(3, 0, false); }
    else { if (last) L(6, 1, false); else L(6, 0, false); }
  } else {
    if (kt == 2) { if (last) L(2, 1, true); else L(2, 0, true); }
    else if (kt == 3) { if (last) L(3, 1, true); else L(3, 0, true); }
    else { if (last) L(6, 1, true); else L(6, 0, true); }
  }
#undef L
  return launch_status("mlp_layer_fwd");
}

int vx_mlp_layer_bwd(int bf16, int nto, const float* dy, int dy_stride, int64_t M,
                     const void* wpk_t, const void* mask, int mask_stride, float* dx,
                     int dx_stride, void* stream) {
  VX_CHECK_ARG(M >= 0, "mlp_layer_bwd: M < 0");
  VX_CHECK_ARG(nto == 2 || nto == 3 || nto == 6, "mlp_layer_bwd: output width must be 64, 96 or 192");
  if (M == 0) return VX_OK;
  hipStream_t st = (hipStream_t)stream;
#define L(NTO, MODE, BF) launch_layer<6, NTO, MODE, BF>(dy, dy_stride, M, wpk_t, nullptr, mask, mask_stride, dx, dx_stride, nullptr, nullptr, nullptr, st)
  if (!bf16) {
    if (nto == 6) { if (mask) L(6, 2, false); else L(6, 3, false); }
    else if (nto == 3) L(3, 3, false);
    else L(2, 3, false);
  } else {
    if (nto == 6) { if (mask) L(6, 2, true); else L(6, 3, true); }
    else if (nto == 3) L(3, 3, true);
    else L(2, 3, true);
  }
#undef L
  return launch_status("mlp_layer_bwd");
}

int vx_mlp_dh_last(int bf16, const float* dlogit, const void* H, int h_stride, int64_t M,
                   const float* w3, float* dH, int dh_stride, void* stream) {
  if (M == 0) return VX_OK;
  hipStream_t st = (hipStream_t)stream;
  if (bf16)
    k_mlp_dh_last<true><<<grid_for(M * kWidth, 256), 256, 0, st>>>(dlogit, H, h_stride, M, w3, dH, dh_stride);
  else
    k_mlp_dh_last<false><<<grid_for(M * kWidth, 256), 256, 0, st>>>(dlogit, H, h_stride, M, w3, dH, dh_stride);
  return launch_status("mlp_dh_last");
}

// Number of row blocks the weight-gradient kernels use for M rows (the partial buffers
// hold that many [192 x K] slices).
int vx_mlp_wgrad_blocks(int64_t M) {
  int64_t b = (M + 255) / 256;
  const int64_t cap = 4 * (int64_t)num_cus();
  if (b > cap) b = cap;
  return (int)(b < 1 ? 1 : b);
}

int vx_mlp_wgrad(int bf16, int kt, int k_real, const float* dY, int dys, const void* X, int xs,
                 int64_t M, float* part, float* partb, float* dW, float* db, void* stream) {
  VX_CHECK_ARG(kt == 2 || kt == 3 || kt == 6, "mlp_wgrad: input width must be 64, 96 or 192");
  hipStream_t st = (hipStream_t)stream;
  const int nblk = vx_mlp_wgrad_blocks(M);
  const int64_t rpb = ((M + nblk - 1) / nblk + 31) / 32 * 32;
  const int K = 32 * kt;
  if (M > 0) {
    if (kt == 2) {
      if (bf16) k_mlp_wgrad<2, true><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
      else k_mlp_wgrad<2, false><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
    } else if (kt == 3) {
      if (bf16) k_mlp_wgrad<3, true><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
      else k_mlp_wgrad<3, false><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
    } else {
      if (bf16) k_mlp_wgrad<6, true><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
      else k_mlp_wgrad<6, false><<<nblk, 256, 0, st>>>(dY, dys, X, xs, M, rpb, part, partb);
    }
    k_reduce_parts<<<grid_for((int64_t)kWidth * k_real, 256), 256, 0, st>>>(part, nblk, kWidth, K, k_real, dW);
    k_reduce_parts<<<1, 256, 0, st>>>(partb, nblk, 1, kWidth, kWidth, db);
  } else {
    k_reduce_parts<<<grid_for((int64_t)kWidth * k_real, 256), 256, 0, st>>>(part, 0, kWidth, K, k_real, dW);
    k_reduce_parts<<<1, 256, 0, st>>>(partb, 0, 1, kWidth, kWidth, db);
  }
  return launch_status("mlp_wgrad");
}

int vx_mlp_wgrad_last(int bf16, const float* dlogit, const void* H, int hs, int64_t M,
                      float* part, float* partb, float* dW3, float* db3, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int nblk = vx_mlp_wgrad_blocks(M);
  const int64_t rpb = (M + nblk - 1) / nblk;
  if (M > 0) {
    if (bf16) k_mlp_wgrad_last<true><<<nblk, 192, 0, st>>>(dlogit, H, hs, M, rpb, part, partb);
    else k_mlp_wgrad_last<false><<<nblk, 192, 0, st>>>(dlogit, H, hs, M, rpb, part, partb);
  }
  k_reduce_parts<<<grid_for(3 * kWidth, 256), 256, 0, st>>>(part, M > 0 ? nblk : 0, 3, kWidth, kWidth, dW3);
  k_reduce_parts<<<1, 64, 0, st>>>(partb, M > 0 ? nblk : 0, 1, 3, 3, db3);
  return launch_status("mlp_wgrad_last");
}

int vx_f32_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  if (n == 0) return VX_OK;
  k_f32_to_bf16<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(x, n, (__bf16*)y);
  return launch_status("f32_to_bf16");
}

}  // extern "C"
