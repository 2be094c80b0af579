// Per-sample elementwise kernels of the Voxurf forward (gfx950): NeuS alpha from the
// SDF and its gradient (lib/voxurf_fine.py:463-500) with its adjoint, and the sorted-
// index segment sum used for ray compositing (torch_scatter.segment_coo, reduce='sum')
// with its adjoint (a gather).
#include "voxel.h"

namespace vx {

template <typename T>
__global__ void k_neus_fwd(const T* __restrict__ viewdirs, const int64_t* __restrict__ ray_id,
                           const T* __restrict__ sdf, const T* __restrict__ grad, int64_t n,
                           T dist, T inv_s, T* __restrict__ alpha) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const NeusFwd<T> f = neus_fwd(viewdirs + 3 * ray_id[i], grad + 3 * i, sdf[i], dist, inv_s);
    alpha[i] = f.a < (T)0 ? (T)0 : (f.a > (T)1 ? (T)1 : f.a);
  }
}

// Adjoint: clip passes the gradient on [0, 1] inclusive; relu(x) only for x > 0;
// sigmoid' = y (1 - y).
template <typename T>
__global__ void k_neus_bwd(const T* __restrict__ viewdirs, const int64_t* __restrict__ ray_id,
                           const T* __restrict__ sdf, const T* __restrict__ grad, int64_t n,
                           T dist, T inv_s, const T* __restrict__ galpha, T* __restrict__ gsdf,
                           T* __restrict__ ggrad) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T* dir = viewdirs + 3 * ray_id[i];
    const NeusFwd<T> f = neus_fwd(dir, grad + 3 * i, sdf[i], dist, inv_s);
    T ga = galpha[i];
    if (!(f.a >= (T)0 && f.a <= (T)1)) ga = 0;
    const T den = f.c + (T)1e-5;
    const T gp = ga / den;                                   // d/dp
    const T gc = -ga * ((f.c - f.nx + (T)1e-5) / (den * den)) + gp;  // d/dc (through c and p)
    const T gn = -gp;                                        // d/dnext_cdf
    const T gprev = gc * (f.c * ((T)1 - f.c)) * inv_s;
    const T gnext = gn * (f.nx * ((T)1 - f.nx)) * inv_s;
    gsdf[i] = gprev + gnext;
    // d ic = (gnext - gprev) * dist * 0.5 ; d cos = d ic * [cos < 0]
    const T gic = (gnext - gprev) * dist * (T)0.5;
    const T gcos = (-f.cos > (T)0) ? gic : (T)0;
    ggrad[3 * i + 0] = gcos * dir[0];
    ggrad[3 * i + 1] = gcos * dir[1];
    ggrad[3 * i + 2] = gcos * dir[2];
  }
}

// segment_coo(src [n, C], index [n] sorted, out [n_out, C], 'sum'): out += segment sums.
// Each wavefront takes 64 consecutive rows; lanes holding the same index are reduced
// with a segmented shuffle scan and the last lane of each run adds it atomically, so a
// run of samples of one ray costs one atomic per channel per wavefront.
template <typename T>
__global__ void __launch_bounds__(256) k_segment_sum(const T* __restrict__ src,
                                                     const int64_t* __restrict__ index, int64_t n,
                                                     int C, int64_t n_out, T* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t wid = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t base = wid * 64; base < n; base += nw * 64) {
    const int64_t i = base + lane;
    const bool valid = i < n;
    const int64_t key = valid ? index[i] : -1;
    // run boundaries: lane l starts a run if l == 0 or key differs from lane l-1
    const int64_t prev_key = __shfl_up(key, 1, 64);
    const int64_t next_key = __shfl_down(key, 1, 64);
    const bool last_of_run = valid && (lane == 63 || next_key != key || i + 1 >= n);
    const bool first_of_run = (lane == 0) || prev_key != key;
    const uint64_t starts = __ballot(first_of_run);
    // run start of this lane = highest set bit of `starts` at or below lane
    const uint64_t below = starts & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int rs = 63 - __clzll(below);
    for (int c = 0; c < C; ++c) {
      T v = valid ? src[i * C + c] : (T)0;
      // segmented inclusive scan
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const T u = __shfl_up(v, off, 64);
        if (lane - off >= rs) v += u;
      }
      if (last_of_run && key >= 0 && key < n_out) unsafeAtomicAdd(&out[key * C + c], v);
    }
  }
}

template <typename T>
__global__ void k_segment_gather(const T* __restrict__ gout, const int64_t* __restrict__ index,
                                 int64_t n, int C, int64_t n_out, T* __restrict__ gsrc) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * C;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / C;
    const int64_t k = index[r];
    gsrc[i] = (k >= 0 && k < n_out) ? gout[k * C + (i % C)] : (T)0;
  }
}

}  // namespace vx

using namespace vx;

#define ST(p) ((hipStream_t)(p))
#define NDISPATCH(dtype, ...)                                         \
  do {                                                                \
    if ((dtype) == VX_F32) {                                          \
      using scalar_t = float;                                         \
      __VA_ARGS__;                                                    \
    } else if ((dtype) == VX_F64) {                                   \
      using scalar_t = double;                                        \
      __VA_ARGS__;                                                    \
    } else {                                                          \
      vx::set_error("unsupported dtype code %d", (int)(dtype));       \
      return VX_ERR_DTYPE;                                            \
    }                                                                 \
  } while (0)
#define CP(p) ((const scalar_t*)(p))
#define MP(p) ((scalar_t*)(p))

extern "C" {

int vx_neus_alpha_fwd(int dtype, const void* viewdirs, const int64_t* ray_id, const void* sdf,
                      const void* gradient, int64_t n, double dist, double inv_s, void* alpha,
                      void* stream) {
  VX_CHECK_ARG(n >= 0, "neus_alpha: n < 0");
  if (n == 0) return VX_OK;
  NDISPATCH(dtype, k_neus_fwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                       CP(viewdirs), ray_id, CP(sdf), CP(gradient), n, (scalar_t)dist,
                       (scalar_t)inv_s, MP(alpha)));
  return launch_status("neus_alpha_fwd");
}

int vx_neus_alpha_bwd(int dtype, const void* viewdirs, const int64_t* ray_id, const void* sdf,
                      const void* gradient, int64_t n, double dist, double inv_s,
                      const void* galpha, void* gsdf, void* ggradient, void* stream) {
  VX_CHECK_ARG(n >= 0, "neus_alpha: n < 0");
  if (n == 0) return VX_OK;
  NDISPATCH(dtype, k_neus_bwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                       CP(viewdirs), ray_id, CP(sdf), CP(gradient), n, (scalar_t)dist,
                       (scalar_t)inv_s, CP(galpha), MP(gsdf), MP(ggradient)));
  return launch_status("neus_alpha_bwd");
}

int vx_segment_sum(int dtype, const void* src, const int64_t* index, int64_t n, int C,
                   int64_t n_out, void* out, void* stream) {
  VX_CHECK_ARG(n >= 0 && C >= 1 && n_out >= 0, "segment_sum: bad sizes");
  if (n == 0) return VX_OK;
  NDISPATCH(dtype, k_segment_sum<scalar_t><<<grid_for(cdiv(n, 64), 4, 1 << 16), 256, 0,
                                             ST(stream)>>>(CP(src), index, n, C, n_out, MP(out)));
  return launch_status("segment_sum");
}

int vx_segment_gather(int dtype, const void* gout, const int64_t* index, int64_t n, int C,
                      int64_t n_out, void* gsrc, void* stream) {
  VX_CHECK_ARG(n >= 0 && C >= 1 && n_out >= 0, "segment_gather: bad sizes");
  if (n == 0) return VX_OK;
  NDISPATCH(dtype, k_segment_gather<scalar_t><<<grid_for(n * C, 256), 256, 0, ST(stream)>>>(
                       CP(gout), index, n, C, n_out, MP(gsrc)));
  return launch_status("segment_gather");
}

}  // extern "C"
