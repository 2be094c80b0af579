// Dense per-step grid operators of the fine stage (gfx950): the smoothed-gradient TV
// term, separable Gaussian smoothing with replicate padding, and trilinear resampling.
//
// smooth_grad_tv (lib/voxurf_fine.py:412-421 with neus_sdf_gradient :440-450 and the
// tv_smooth_conv of :204-240): the reference materialises the [1,3,X,Y,Z] central
// difference grid, pads and convolves it with a 3x3x3 kernel, subtracts, gathers the
// masked entries and reduces; autograd then runs all of it backwards.  The 3x3x3
// kernel is an outer product of three 3-tap filters (kernel0 = [1,2,1]^3 times
// exp(-sigma (i^2+j^2+k^2-1))), so here one kernel reads an SDF tile with a 2-voxel
// halo into LDS, forms the gradient on a 1-voxel halo, smooths it there and writes the
// loss gradient with respect to the gradient field (dG) plus a per-block loss partial;
// the adjoint of the central difference (vx_dense_sdf_gradient_bwd) then adds dG's
// contribution to the SDF gradient.  HBM traffic: SDF and mask once, dG written and
// read once, the SDF gradient read-modify-written once.
#include "common.h"

namespace vx {

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Output tile (x, y, z) of one block; z is the contiguous axis.
constexpr int kTX = 4, kTY = 8, kTZ = 64;
constexpr int kSX = kTX + 4, kSY = kTY + 4, kSZ = kTZ + 4;  // SDF tile, 2-voxel halo
constexpr int kGX = kTX + 2, kGY = kTY + 2, kGZ = kTZ + 2;  // gradient tile, 1-voxel halo
constexpr int kSgtvThreads = 1024;

template <typename T>
struct SgtvArgs {
  const T* sdf;
  int X, Y, Z;
  T vs;
  T w[3];               // 1-D smoothing taps (offsets -1, 0, +1)
  const uint8_t* mask;  // [X, Y, Z] or nullptr (every voxel counts)
  T scale;              // dL/dG = scale * E on masked voxels (scale = -2 weight / count)
  T* dg;                // [3, X, Y, Z]
  T* partial;           // [blocks] sum of masked E^2 per block
};

// Gradient channel ch (0: x, 1: y, 2: z) of the SDF tile at grid point q, read from the
// LDS tile S whose element (0, 0, 0) is grid point (ox - 2, oy - 2, oz - 2).
template <typename T>
__device__ __forceinline__ T tile_grad(const T* S, int ch, int qx, int qy, int qz, int ox, int oy,
                                       int oz, int X, int Y, int Z, T vs) {
#pragma clang fp contract(off)
  const int c = ((qx - ox + 2) * kSY + (qy - oy + 2)) * kSZ + (qz - oz + 2);
  if (ch == 0) return (qx > 0 && qx < X - 1) ? ((S[c + kSY * kSZ] - S[c - kSY * kSZ]) / (T)2) / vs : (T)0;
  if (ch == 1) return (qy > 0 && qy < Y - 1) ? ((S[c + kSZ] - S[c - kSZ]) / (T)2) / vs : (T)0;
  return (qz > 0 && qz < Z - 1) ? ((S[c + 1] - S[c - 1]) / (T)2) / vs : (T)0;
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_smooth_grad_tv(SgtvArgs<T> a) {
#pragma clang fp contract(off)
  constexpr int kOut = kTX * kTY * kTZ;
  constexpr int kPer = kOut / NT;                  // outputs per thread
  static_assert(kOut % NT == 0, "tile must split evenly over the block");
  __shared__ T S[kSX * kSY * kSZ];                 // SDF, 2-voxel halo
  __shared__ T G[kGX * kGY * kGZ];                 // one gradient channel, 1-voxel halo
  __shared__ T Gz[kGX * kGY * kTZ];                // ... smoothed along z
  __shared__ T Gyz[kGX * kTY * kTZ];               // ... and along y
  __shared__ T red[NT / 64];
  const int X = a.X, Y = a.Y, Z = a.Z;
  const int nbz = (Z + kTZ - 1) / kTZ, nby = (Y + kTY - 1) / kTY;
  const int b = blockIdx.x;
  const int ox = (b / (nbz * nby)) * kTX, oy = (b / nbz % nby) * kTY, oz = (b % nbz) * kTZ;
  const int64_t sx = (int64_t)Y * Z, sy = Z;
  const int64_t N = (int64_t)X * Y * Z;
  // 1. SDF at raw coordinates origin-2 .. origin+T+1, clamped into the grid; every
  // load of the thread is issued before the first LDS store
  constexpr int kSN = kSX * kSY * kSZ, kSIt = (kSN + NT - 1) / NT;
  T sv[kSIt];
#pragma unroll
  for (int k = 0; k < kSIt; ++k) {
    const int i = threadIdx.x + NT * k;
    if (i < kSN) {
      const int lz = i % kSZ, ly = i / kSZ % kSY, lx = i / (kSZ * kSY);
      const int x = clampi(ox - 2 + lx, 0, X - 1), y = clampi(oy - 2 + ly, 0, Y - 1),
                z = clampi(oz - 2 + lz, 0, Z - 1);
      sv[k] = a.sdf[x * sx + y * sy + z];
    }
  }
#pragma unroll
  for (int k = 0; k < kSIt; ++k) {
    const int i = threadIdx.x + NT * k;
    if (i < kSN) S[i] = sv[k];
  }
  // the thread's voxels: in-grid and mask bits
  uint32_t inside = 0, mbits = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = threadIdx.x + NT * k;
    const int lz = i % kTZ, ly = i / kTZ % kTY, lx = i / (kTZ * kTY);
    const int x = ox + lx, y = oy + ly, z = oz + lz;
    if (x < X && y < Y && z < Z) {
      inside |= 1u << k;
      if (!a.mask || a.mask[x * sx + y * sy + z]) mbits |= 1u << k;
    }
  }
  T acc = 0;
  for (int ch = 0; ch < 3; ++ch) {
    __syncthreads();  // S is loaded; the previous channel's reads of G and Gyz are done
    // 2. gradient channel ch at q = clamp(p) for raw p = origin-1 .. origin+T (the
    // replicate padding of the gradient grid), each computed once (fixed trip counts,
    // unrolled, so a thread's LDS reads are in flight together)
#pragma unroll
    for (int k = 0; k < (kGX * kGY * kGZ + NT - 1) / NT; ++k) {
      const int i = threadIdx.x + NT * k;
      if (i >= kGX * kGY * kGZ) break;
      const int lz = i % kGZ, ly = i / kGZ % kGY, lx = i / (kGZ * kGY);
      G[i] = tile_grad(S, ch, clampi(ox - 1 + lx, 0, X - 1), clampi(oy - 1 + ly, 0, Y - 1),
                       clampi(oz - 1 + lz, 0, Z - 1), ox, oy, oz, X, Y, Z, a.vs);
    }
    __syncthreads();
    // 3. separable smoothing: along z, then y (x last, with the output)
#pragma unroll
    for (int k = 0; k < (kGX * kGY * kTZ + NT - 1) / NT; ++k) {
      const int i = threadIdx.x + NT * k;
      if (i >= kGX * kGY * kTZ) break;
      const int lz = i % kTZ, r = i / kTZ;  // r = lx * kGY + ly
      const int g0 = r * kGZ + lz + 1;
      Gz[i] = (a.w[0] * G[g0 - 1] + a.w[1] * G[g0]) + a.w[2] * G[g0 + 1];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (kGX * kTY * kTZ + NT - 1) / NT; ++k) {
      const int i = threadIdx.x + NT * k;
      if (i >= kGX * kTY * kTZ) break;
      const int lz = i % kTZ, ly = i / kTZ % kTY, lx = i / (kTZ * kTY);
      const int g0 = (lx * kGY + ly) * kTZ + lz;
      Gyz[i] = (a.w[0] * Gz[g0] + a.w[1] * Gz[g0 + kTZ]) + a.w[2] * Gz[g0 + 2 * kTZ];
    }
    __syncthreads();
    // 4. along x, residual E = S(G) - G, loss partial and dG for the tile's voxels
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      if (!(inside >> k & 1u)) continue;
      const int i = threadIdx.x + NT * k;
      const int lz = i % kTZ, ly = i / kTZ % kTY, lx = i / (kTZ * kTY);
      const int g0 = (lx * kTY + ly) * kTZ + lz;
      const T sm = (a.w[0] * Gyz[g0] + a.w[1] * Gyz[g0 + kTY * kTZ]) + a.w[2] * Gyz[g0 + 2 * kTY * kTZ];
      const int x = ox + lx, y = oy + ly, z = oz + lz;
      const T e = sm - G[((lx + 1) * kGY + (ly + 1)) * kGZ + (lz + 1)];
      const bool m = mbits >> k & 1u;
      if (m) acc += e * e;
      a.dg[ch * N + (x * sx + y * sy + z)] = m ? a.scale * e : (T)0;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    a.partial[b] = t;
  }
}

// loss = factor * sum(partial), summed in a fixed order by one block.
template <typename T>
__global__ void __launch_bounds__(1024) k_sum_partials(const T* __restrict__ p, int64_t n, T factor,
                                                       T* __restrict__ out) {
  __shared__ T red[16];
  T s = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[0] = t * factor;
  }
}

// ---- separable smoothing with replicate padding ---------------------------------------
// One axis of F.conv3d(pad(x, replicate), w_x (x) w_y (x) w_z): out[u] = sum_t w_t
// in[clamp(u + t - r)] along `axis` (r = k / 2), or its adjoint (transpose = 1):
// out[u] = sum_t w_t sum_{v : clamp(v + t - r) = u} in[v].

constexpr int kMaxTaps = 9;

template <typename T>
struct Taps {
  T w[kMaxTaps];
};

// Launched as a 3-D grid: (z chunks of 256 threads, y, x + X * channel).
template <typename T>
__global__ void __launch_bounds__(256) k_smooth_axis(const T* __restrict__ in, T* __restrict__ out,
                                                     int C, int X, int Y, int Z, int axis,
                                                     Taps<T> taps, int k, int transpose) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const int x = blockIdx.z % X, c = blockIdx.z / X;
  if (z >= Z) return;
  const int n = axis == 0 ? X : (axis == 1 ? Y : Z);
  const int u = axis == 0 ? x : (axis == 1 ? y : z);
  const int64_t stride = axis == 0 ? (int64_t)Y * Z : (axis == 1 ? (int64_t)Z : 1);
  const int64_t i = (((int64_t)c * X + x) * Y + y) * Z + z;
  const int r = k / 2;
  const T* line = in + (i - (int64_t)u * stride);
  T acc = 0;
  if (!transpose) {
    for (int t = 0; t < k; ++t) acc += taps.w[t] * line[(int64_t)clampi(u + t - r, 0, n - 1) * stride];
  } else {
    for (int t = 0; t < k; ++t) {
      // the v whose tap t lands on u
      int lo, hi;
      if (u == 0 && u == n - 1) {
        lo = 0;
        hi = 0;
      } else if (u == 0) {
        lo = 0;
        hi = min(n - 1, r - t);
      } else if (u == n - 1) {
        lo = max(0, n - 1 - t + r);
        hi = n - 1;
      } else {
        lo = hi = u - t + r;
        if (lo < 0 || lo > n - 1) continue;
      }
      T s = 0;
      for (int v = lo; v <= hi; ++v) s += line[(int64_t)v * stride];
      acc += taps.w[t] * s;
    }
  }
  out[i] = acc;
}

// ---- trilinear resampling, align_corners = True (F.interpolate, lib/grid.py:60-65) ----

template <typename T>
__device__ __forceinline__ void src_index(int d, int in_n, int out_n, int& i0, int& i1, T& l0, T& l1) {
  const T scale = out_n > 1 ? (T)(in_n - 1) / (T)(out_n - 1) : (T)0;
  const T real = scale * (T)d;
  i0 = (int)real;
  if (i0 > in_n - 1) i0 = in_n - 1;
  T lam = real - (T)i0;
  lam = lam < (T)0 ? (T)0 : (lam > (T)1 ? (T)1 : lam);
  i1 = i0 < in_n - 1 ? i0 + 1 : i0;
  l1 = lam;
  l0 = (T)1 - lam;
}

// Launched as a 3-D grid: (z chunks of 256 threads, y, x + X2 * channel).
template <typename T>
__global__ void __launch_bounds__(256) k_resample(const T* __restrict__ in, int C, int X, int Y, int Z,
                                                  T* __restrict__ out, int X2, int Y2, int Z2) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  const int x = blockIdx.z % X2, c = blockIdx.z / X2;
  if (z >= Z2) return;
  int x0, x1, y0, y1, z0, z1;
  T ax0, ax1, ay0, ay1, az0, az1;
  src_index<T>(x, X, X2, x0, x1, ax0, ax1);
  src_index<T>(y, Y, Y2, y0, y1, ay0, ay1);
  src_index<T>(z, Z, Z2, z0, z1, az0, az1);
  const T* g = in + (int64_t)c * X * Y * Z;
  auto at = [&](int a, int b, int d) { return g[((int64_t)a * Y + b) * Z + d]; };
  out[(((int64_t)c * X2 + x) * Y2 + y) * Z2 + z] =
      ax0 * (ay0 * (az0 * at(x0, y0, z0) + az1 * at(x0, y0, z1)) +
             ay1 * (az0 * at(x0, y1, z0) + az1 * at(x0, y1, z1))) +
      ax1 * (ay0 * (az0 * at(x1, y0, z0) + az1 * at(x1, y0, z1)) +
             ay1 * (az0 * at(x1, y1, z0) + az1 * at(x1, y1, z1)));
}

}  // namespace vx

using namespace vx;

#define SMOOTH_DISPATCH(dtype, ...)                                   \
  do {                                                                \
    if ((dtype) == VX_F32) {                                          \
      using scalar_t = float;                                         \
      __VA_ARGS__;                                                    \
    } else if ((dtype) == VX_F64) {                                   \
      using scalar_t = double;                                        \
      __VA_ARGS__;                                                    \
    } else {                                                          \
      ::vx::set_error("unsupported dtype code %d", (int)(dtype));     \
      return VX_ERR_DTYPE;                                            \
    }                                                                 \
  } while (0)

static int64_t sgtv_blocks(int64_t X, int64_t Y, int64_t Z) {
  return cdiv(X, kTX) * cdiv(Y, kTY) * cdiv(Z, kTZ);
}

extern "C" {

int64_t vx_smooth_grad_tv_partials(int64_t X, int64_t Y, int64_t Z) {
  if (X < 1 || Y < 1 || Z < 1) return 0;
  return sgtv_blocks(X, Y, Z);
}

int vx_smooth_grad_tv(int dtype, const void* sdf, int X, int Y, int Z, double voxel_size,
                      const double* taps, const uint8_t* mask, int64_t mask_count, double weight,
                      void* dgrad, void* partials, void* loss, void* sdf_grad, void* stream) {
  VX_CHECK_ARG(X >= 1 && Y >= 1 && Z >= 1, "smooth_grad_tv: bad shape");
  VX_CHECK_ARG(taps && sdf && dgrad && partials, "smooth_grad_tv: missing argument");
  VX_CHECK_ARG(mask_count >= 1, "smooth_grad_tv: the mask selects no voxel");
  const int64_t nb = sgtv_blocks(X, Y, Z);
  VX_CHECK_ARG(nb <= 0x7fffffff, "smooth_grad_tv: grid too large");
  hipStream_t st = (hipStream_t)stream;
  // mean over the 3 * mask_count selected entries of E^2
  const double count = 3.0 * (double)mask_count;
  SMOOTH_DISPATCH(dtype, {
    SgtvArgs<scalar_t> a;
    a.sdf = (const scalar_t*)sdf;
    a.X = X;
    a.Y = Y;
    a.Z = Z;
    a.vs = (scalar_t)voxel_size;
    for (int i = 0; i < 3; ++i) a.w[i] = (scalar_t)taps[i];
    a.mask = mask;
    a.scale = (scalar_t)(-2.0 * weight / count);
    a.dg = (scalar_t*)dgrad;
    a.partial = (scalar_t*)partials;
    k_smooth_grad_tv<scalar_t, kSgtvThreads><<<(unsigned)nb, kSgtvThreads, 0, st>>>(a);
    if (loss)
      k_sum_partials<scalar_t><<<1, 1024, 0, st>>>((const scalar_t*)partials, nb,
                                                   (scalar_t)(weight / count), (scalar_t*)loss);
  });
  if (int e = launch_status("smooth_grad_tv")) return e;
  // the adjoint of the central difference adds dG's contribution to the SDF gradient
  if (sdf_grad) return vx_dense_sdf_gradient_bwd(dtype, dgrad, X, Y, Z, voxel_size, sdf_grad, stream);
  return VX_OK;
}

int vx_smooth3d(int dtype, const void* in, void* out, void* tmp, int C, int X, int Y, int Z,
                const double* taps, int k, int transpose, void* stream) {
  VX_CHECK_ARG(C >= 1 && X >= 1 && Y >= 1 && Z >= 1, "smooth3d: bad shape");
  VX_CHECK_ARG(k >= 1 && k <= kMaxTaps && (k & 1), "smooth3d: taps must be odd and <= 9");
  VX_CHECK_ARG(in && out && tmp && taps && in != out && out != tmp && in != tmp,
               "smooth3d: in, out and tmp must be distinct buffers");
  VX_CHECK_ARG(Y <= 65535 && (int64_t)X * C <= 65535, "smooth3d: grid too large");
  hipStream_t st = (hipStream_t)stream;
  SMOOTH_DISPATCH(dtype, {
    Taps<scalar_t> w{};
    for (int i = 0; i < k; ++i) w.w[i] = (scalar_t)taps[i];
    const scalar_t* src = (const scalar_t*)in;
    const dim3 grid((unsigned)cdiv(Z, 256), (unsigned)Y, (unsigned)((int64_t)X * C));
    // z into out, y into tmp, x into out
    k_smooth_axis<scalar_t><<<grid, 256, 0, st>>>(src, (scalar_t*)out, C, X, Y, Z, 2, w, k,
                                                  transpose);
    k_smooth_axis<scalar_t><<<grid, 256, 0, st>>>((const scalar_t*)out, (scalar_t*)tmp, C, X, Y, Z,
                                                  1, w, k, transpose);
    k_smooth_axis<scalar_t><<<grid, 256, 0, st>>>((const scalar_t*)tmp, (scalar_t*)out, C, X, Y, Z,
                                                  0, w, k, transpose);
  });
  return launch_status("smooth3d");
}

int vx_resample_trilinear(int dtype, const void* in, int C, int X, int Y, int Z, void* out, int X2,
                          int Y2, int Z2, void* stream) {
  VX_CHECK_ARG(C >= 0 && X >= 1 && Y >= 1 && Z >= 1 && X2 >= 1 && Y2 >= 1 && Z2 >= 1,
               "resample_trilinear: bad shape");
  if (C == 0) return VX_OK;
  VX_CHECK_ARG(Y2 <= 65535 && (int64_t)X2 * C <= 65535, "resample_trilinear: grid too large");
  const dim3 grid((unsigned)cdiv(Z2, 256), (unsigned)Y2, (unsigned)((int64_t)X2 * C));
  SMOOTH_DISPATCH(dtype, k_resample<scalar_t><<<grid, 256, 0, (hipStream_t)stream>>>(
                             (const scalar_t*)in, C, X, Y, Z, (scalar_t*)out, X2, Y2, Z2));
  return launch_status("resample_trilinear");
}

}  // extern "C"
