// Dense voxel-grid kernels (gfx950): trilinear gather and its scatter-add adjoint, the
// 6-neighbour SDF stencil (centre value + clamped central differences), the dense
// central-difference gradient grid, and the total-variation gradient.
//
// A grid is [C, X, Y, Z] with arbitrary element strides (sc, sx, sy, sz): the
// reference's contiguous [1, C, X, Y, Z] layout (sc = XYZ, sz = 1) and the
// channels-last layout used for multi-channel feature grids (sc = 1, sz = C) are both
// served by the same kernels.  The interpolation reproduces F.grid_sample(...,
// mode='bilinear', align_corners=True, padding_mode='zeros') on a 5-D input: the same
// normalise/unnormalise chain, corner order and weight products (contraction off), so
// results agree with PyTorch's CPU kernel to rounding of the final sum.
#include "voxel.h"

namespace vx {

// ---- DenseGrid.forward / backward (lib/grid.py:47-58) ----------------------------

template <typename T>
__global__ void k_tri_fwd(const T* __restrict__ grid, GridDesc g, const T* __restrict__ xyz,
                          const T* __restrict__ mn, const T* __restrict__ mx, int64_t n,
                          T* __restrict__ out) {
  const T ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    Tri<T> t;
    t.init(world_to_index(xyz[3 * p], mn[0], ex, g.X), world_to_index(xyz[3 * p + 1], mn[1], ey, g.Y),
           world_to_index(xyz[3 * p + 2], mn[2], ez, g.Z));
    for (int c = 0; c < g.C; ++c) out[p * g.C + c] = tri_gather(grid, g, t, c);
  }
}

template <typename T>
__global__ void k_tri_bwd(T* __restrict__ grad, GridDesc g, const T* __restrict__ xyz,
                          const T* __restrict__ mn, const T* __restrict__ mx, int64_t n,
                          const T* __restrict__ gout) {
  const T ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    Tri<T> t;
    t.init(world_to_index(xyz[3 * p], mn[0], ex, g.X), world_to_index(xyz[3 * p + 1], mn[1], ey, g.Y),
           world_to_index(xyz[3 * p + 2], mn[2], ez, g.Z));
    for (int c = 0; c < g.C; ++c) tri_scatter(grad, g, t, c, gout[p * g.C + c]);
  }
}

// ---- sample_sdfs (lib/voxurf_fine.py:537-577) -------------------------------------
// For each point and displacement d in `disp[0..nd)`, the SDF is sampled at the 6
// points ind +- d along each index axis (zyx order), each component clamped to
// [0, size-1] in index space, then mapped back through the normalised coordinate
// (all_ind / (size-1)) * 2 - 1 before grid_sample unnormalises it again.
//   feat[p, k*nd + j]   k = 0..5 : (-z, +z, -y, +y, -x, +x)
//   grad[p, a*nd + j]   a = 0..2 : (z, y, x),  (f[+] - f[-]) / (ind[+] - ind[-]) / voxel_size
// optionally normalised per displacement: grad / (|grad| + 1e-5).

template <typename T>
__global__ void k_sdfs_fwd(const T* __restrict__ grid, GridDesc g, const T* __restrict__ xyz,
                           const T* __restrict__ mn, const T* __restrict__ mx,
                           const T* __restrict__ disp, int nd, T voxel_size, int use_norm,
                           int64_t n, T* __restrict__ feat, T* __restrict__ grad) {
#pragma clang fp contract(off)
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    T ind[3];
    zyx_index(xyz + 3 * p, mn, mx, g, ind);
    for (int j = 0; j < nd; ++j) {
      const T d = disp[j];
      T gr[3];
      for (int a = 0; a < 3; ++a) {
        T cm, cp;
        const T fm = sample_displaced(grid, g, ind, a, -d, cm);
        const T fp = sample_displaced(grid, g, ind, a, d, cp);
        feat[p * 6 * nd + (2 * a) * nd + j] = fm;
        feat[p * 6 * nd + (2 * a + 1) * nd + j] = fp;
        const T df = clamped_diff(ind, a, d, g);
        gr[a] = ((fp - fm) / df) / voxel_size;
      }
      if (use_norm) {
        const T nrm = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]) + (T)1e-5;
        for (int a = 0; a < 3; ++a) gr[a] = gr[a] / nrm;
      }
      for (int a = 0; a < 3; ++a) grad[p * 3 * nd + a * nd + j] = gr[a];
    }
  }
}

template <typename T>
__global__ void k_sdfs_bwd(T* __restrict__ ggrid, const T* __restrict__ grid, GridDesc g,
                           const T* __restrict__ xyz, const T* __restrict__ mn,
                           const T* __restrict__ mx, const T* __restrict__ disp, int nd,
                           T voxel_size, int use_norm, int64_t n, const T* __restrict__ gfeat,
                           const T* __restrict__ ggrad) {
#pragma clang fp contract(off)
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    T ind[3];
    zyx_index(xyz + 3 * p, mn, mx, g, ind);
    for (int j = 0; j < nd; ++j) {
      const T d = disp[j];
      T df[3], gg[3];
      for (int a = 0; a < 3; ++a) {
        df[a] = clamped_diff(ind, a, d, g);
        gg[a] = ggrad ? ggrad[p * 3 * nd + a * nd + j] : (T)0;
      }
      if (use_norm && ggrad) {
        // y = r / (|r| + eps):  dr = (gy - r (r . gy) / (|r| (|r| + eps))) / (|r| + eps)
        T r[3];
        for (int a = 0; a < 3; ++a) {
          T cm, cp;
          const T fm = sample_displaced(grid, g, ind, a, -d, cm);
          const T fp = sample_displaced(grid, g, ind, a, d, cp);
          r[a] = ((fp - fm) / df[a]) / voxel_size;
        }
        const T nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const T den = nr + (T)1e-5;
        const T dot = r[0] * gg[0] + r[1] * gg[1] + r[2] * gg[2];
        const T k = nr > (T)0 ? dot / (nr * den) : (T)0;
        for (int a = 0; a < 3; ++a) gg[a] = (gg[a] - r[a] * k) / den;
      }
      for (int a = 0; a < 3; ++a) {
        const T gd = gg[a] / df[a] / voxel_size;  // d grad / d f[+] ; -gd for f[-]
        const T gm = (gfeat ? gfeat[p * 6 * nd + (2 * a) * nd + j] : (T)0) - gd;
        const T gp = (gfeat ? gfeat[p * 6 * nd + (2 * a + 1) * nd + j] : (T)0) + gd;
        scatter_displaced(ggrid, g, ind, a, -d, gm);
        scatter_displaced(ggrid, g, ind, a, d, gp);
      }
    }
  }
}

// ---- Voxurf.grid_sampler(sample_ret, sample_grad) (lib/voxurf_fine.py:502-534) ------
// Centre SDF value plus the displacement-1 stencil in one pass:
//   sdf [n], grad [n,3] (xyz order, unnormalised), feat [n,6] (-x,+x,-y,+y,-z,+z).

template <typename T>
__global__ void k_sdf_grad_fwd(const T* __restrict__ grid, GridDesc g, const T* __restrict__ xyz,
                               const T* __restrict__ mn, const T* __restrict__ mx, T voxel_size,
                               int64_t n, T* __restrict__ sdf, T* __restrict__ grad,
                               T* __restrict__ feat) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T q[3] = {xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]};
    T s, gr[3], f[6];
    sdf7(grid, g, q, mn, mx, voxel_size, s, gr, f);
    sdf[p] = s;
    for (int c = 0; c < 3; ++c) grad[3 * p + c] = gr[c];
    if (feat)
      for (int c = 0; c < 6; ++c) feat[6 * p + c] = f[c];
  }
}

template <typename T>
__global__ void k_sdf_grad_bwd(T* __restrict__ ggrid, GridDesc g, const T* __restrict__ xyz,
                               const T* __restrict__ mn, const T* __restrict__ mx, T voxel_size,
                               int64_t n, const T* __restrict__ gsdf, const T* __restrict__ ggrad,
                               const T* __restrict__ gfeat) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T q[3] = {xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]};
    const T gg[3] = {ggrad ? ggrad[3 * p] : (T)0, ggrad ? ggrad[3 * p + 1] : (T)0,
                     ggrad ? ggrad[3 * p + 2] : (T)0};
    sdf7_scatter(ggrid, g, q, mn, mx, voxel_size, gsdf ? gsdf[p] : (T)0, gg,
                 gfeat ? gfeat + 6 * p : (const T*)nullptr);
  }
}

// ---- neus_sdf_gradient 'interpolate' (lib/voxurf_fine.py:440-450) ------------------
// out[a] = (s[+1] - s[-1]) / 2 / voxel_size along axis a, 0 on that axis' boundary faces.
// Launched as a 3-D grid (z chunks of 256 threads, y, x): no index division.

template <typename T>
__global__ void __launch_bounds__(256) k_dense_grad_fwd(const T* __restrict__ s, int X, int Y, int Z,
                                                        T vs, T* __restrict__ out) {
#pragma clang fp contract(off)
  const int z = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, x = blockIdx.z;
  if (z >= Z) return;
  const int64_t N = (int64_t)X * Y * Z;
  const int64_t sx = (int64_t)Y * Z, sy = Z;
  const int64_t i = x * sx + y * sy + z;
  out[i] = (x > 0 && x < X - 1) ? ((s[i + sx] - s[i - sx]) / (T)2) / vs : (T)0;
  out[N + i] = (y > 0 && y < Y - 1) ? ((s[i + sy] - s[i - sy]) / (T)2) / vs : (T)0;
  out[2 * N + i] = (z > 0 && z < Z - 1) ? ((s[i + 1] - s[i - 1]) / (T)2) / vs : (T)0;
}

// Adjoint: gs[i] += sum_a (g_a[i - e_a] - g_a[i + e_a]) / 2 / vs over interior entries.
template <typename T>
__global__ void __launch_bounds__(256) k_dense_grad_bwd(const T* __restrict__ go, int X, int Y, int Z,
                                                        T vs, T* __restrict__ gs) {
#pragma clang fp contract(off)
  const int z = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, x = blockIdx.z;
  if (z >= Z) return;
  const int64_t N = (int64_t)X * Y * Z;
  const int64_t sx = (int64_t)Y * Z, sy = Z;
  const int64_t i = x * sx + y * sy + z;
  T acc = 0;
  if (x - 1 > 0) acc += (go[i - sx] / (T)2) / vs;
  if (x + 1 < X - 1) acc -= (go[i + sx] / (T)2) / vs;
  if (y - 1 > 0) acc += (go[N + i - sy] / (T)2) / vs;
  if (y + 1 < Y - 1) acc -= (go[N + i + sy] / (T)2) / vs;
  if (z - 1 > 0) acc += (go[2 * N + i - 1] / (T)2) / vs;
  if (z + 1 < Z - 1) acc -= (go[2 * N + i + 1] / (T)2) / vs;
  gs[i] += acc;
}

// ---- total_variation_add_grad (lib/cuda/total_variation_kernel.cu:13-66) ----------
// Weights arrive already divided by 6.  Axis weights follow the reference exactly: the
// plain kernel uses wz on both the contiguous (k) and outermost (i) axes and wy on j;
// the masked kernel uses wx, wy, wz on k, j, i.

template <typename T, bool kMasked>
__global__ void k_tv(const T* __restrict__ param, T* __restrict__ grad, const T* __restrict__ mask,
                     GridDesc g, float wk, float wj, float wi, int dense) {
#pragma clang fp contract(off)
  const int64_t N = (int64_t)g.C * g.X * g.Y * g.Z;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N;
       idx += (int64_t)gridDim.x * blockDim.x) {
    // logical (c, i, j, k) -> address via strides
    const int k = (int)(idx % g.Z);
    const int j = (int)(idx / g.Z % g.Y);
    const int i = (int)(idx / g.Z / g.Y % g.X);
    const int c = (int)(idx / g.Z / g.Y / g.X);
    const int64_t a = c * g.sc + i * g.sx + j * g.sy + k * g.sz;
    if (!dense && grad[a] == (T)0) continue;
    const T p = param[a];
    float add = 0;
    // grad_to_add (float) += w * clamp(p - p_nb, -1, 1) [* mask[a] * mask[nb]], the
    // product formed in the tensor type.
    auto term = [&](bool has, int64_t nb, float w) {
      if (!has) return;
      T v = p - param[nb];
      v = v < (T)-1 ? (T)-1 : (v > (T)1 ? (T)1 : v);
      T t = (T)w * v;
      if (kMasked) t = (t * mask[a]) * mask[nb];
      add = (float)((T)add + t);
    };
    term(k > 0, a - g.sz, wk);
    term(k < g.Z - 1, a + g.sz, wk);
    term(j > 0, a - g.sy, wj);
    term(j < g.Y - 1, a + g.sy, wj);
    term(i > 0, a - g.sx, wi);
    term(i < g.X - 1, a + g.sx, wi);
    grad[a] += (T)add;
  }
}

// ---- MaskCache.forward (lib/voxurf_fine.py:930-942, lib/dvgo_ori.py:440-473) --------
// density at the point (trilinear, or nearest lattice point with round-half-to-even as
// grid_sample's 'nearest'), alpha = 1 - exp(-softplus(density + act_shift) * ratio)
// with torch's softplus (threshold 20), out = alpha >= thres.
template <typename T>
__device__ __forceinline__ T softplus20(T x) {
  return x > (T)20 ? x : log1p(exp(x));
}

template <typename T>
__global__ void k_maskcache_alpha(const T* __restrict__ dens, GridDesc g, const T* __restrict__ xyz,
                                  const T* __restrict__ mn, const T* __restrict__ mx, T act_shift,
                                  T ratio, T thres, int nearest, int64_t n,
                                  uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
  const T ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T fx = world_to_index(xyz[3 * p], mn[0], ex, g.X);
    const T fy = world_to_index(xyz[3 * p + 1], mn[1], ey, g.Y);
    const T fz = world_to_index(xyz[3 * p + 2], mn[2], ez, g.Z);
    T d;
    if (nearest) {
      const int x = sat_f2i(rint(fx)), y = sat_f2i(rint(fy)), z = sat_f2i(rint(fz));
      d = in_grid<T>(x, y, z, g) ? dens[x * g.sx + y * g.sy + z * g.sz] : (T)0;
    } else {
      Tri<T> t;
      t.init(fx, fy, fz);
      d = tri_gather(dens, g, t, 0);
    }
    const T alpha = (T)1 - exp(-softplus20(d + act_shift) * ratio);
    out[p] = alpha >= thres;
  }
}

template <typename T>
__global__ void k_zero(T* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

}  // namespace vx

using namespace vx;

#define ST(p) ((hipStream_t)(p))

static int make_desc(const int64_t* shape_strides, GridDesc& g) {
  // shape_strides = {C, X, Y, Z, sc, sx, sy, sz}
  g.C = (int)shape_strides[0];
  g.X = (int)shape_strides[1];
  g.Y = (int)shape_strides[2];
  g.Z = (int)shape_strides[3];
  g.sc = shape_strides[4];
  g.sx = shape_strides[5];
  g.sy = shape_strides[6];
  g.sz = shape_strides[7];
  VX_CHECK_ARG(g.C >= 1 && g.X >= 1 && g.Y >= 1 && g.Z >= 1, "grid: non-positive dimension");
  return VX_OK;
}

#define GRID_DISPATCH(dtype, ...)                                     \
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

int vx_grid_sample_fwd(int dtype, const void* grid, const int64_t* geom, const void* xyz,
                       const void* xyz_min, const void* xyz_max, int64_t n, void* out,
                       void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(n >= 0, "grid_sample: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_tri_fwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           CP(grid), g, CP(xyz), CP(xyz_min), CP(xyz_max), n, MP(out)));
  return launch_status("grid_sample_fwd");
}

int vx_grid_sample_bwd(int dtype, void* grad, const int64_t* geom, const void* xyz,
                       const void* xyz_min, const void* xyz_max, int64_t n, const void* gout,
                       void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(n >= 0, "grid_sample: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_tri_bwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           MP(grad), g, CP(xyz), CP(xyz_min), CP(xyz_max), n, CP(gout)));
  return launch_status("grid_sample_bwd");
}

int vx_sample_sdfs_fwd(int dtype, const void* grid, const int64_t* geom, const void* xyz,
                       const void* xyz_min, const void* xyz_max, const void* disp, int nd,
                       double voxel_size, int use_grad_norm, int64_t n, void* feat, void* grad,
                       void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(g.C == 1, "sample_sdfs: the SDF grid must have one channel");
  VX_CHECK_ARG(n >= 0 && nd >= 1, "sample_sdfs: bad sizes");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_sdfs_fwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           CP(grid), g, CP(xyz), CP(xyz_min), CP(xyz_max), CP(disp), nd,
                           (scalar_t)voxel_size, use_grad_norm, n, MP(feat), MP(grad)));
  return launch_status("sample_sdfs_fwd");
}

int vx_sample_sdfs_bwd(int dtype, void* ggrid, const void* grid, const int64_t* geom,
                       const void* xyz, const void* xyz_min, const void* xyz_max,
                       const void* disp, int nd, double voxel_size, int use_grad_norm, int64_t n,
                       const void* gfeat, const void* ggrad, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(g.C == 1, "sample_sdfs: the SDF grid must have one channel");
  VX_CHECK_ARG(n >= 0 && nd >= 1, "sample_sdfs: bad sizes");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_sdfs_bwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           MP(ggrid), CP(grid), g, CP(xyz), CP(xyz_min), CP(xyz_max), CP(disp),
                           nd, (scalar_t)voxel_size, use_grad_norm, n, CP(gfeat), CP(ggrad)));
  return launch_status("sample_sdfs_bwd");
}

int vx_sdf_grad_fwd(int dtype, const void* grid, const int64_t* geom, const void* xyz,
                    const void* xyz_min, const void* xyz_max, double voxel_size, int64_t n,
                    void* sdf, void* grad, void* feat, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(g.C == 1, "sdf_grad: the SDF grid must have one channel");
  VX_CHECK_ARG(n >= 0, "sdf_grad: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_sdf_grad_fwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           CP(grid), g, CP(xyz), CP(xyz_min), CP(xyz_max), (scalar_t)voxel_size,
                           n, MP(sdf), MP(grad), MP(feat)));
  return launch_status("sdf_grad_fwd");
}

int vx_sdf_grad_bwd(int dtype, void* ggrid, const int64_t* geom, const void* xyz,
                    const void* xyz_min, const void* xyz_max, double voxel_size, int64_t n,
                    const void* gsdf, const void* ggrad, const void* gfeat, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(g.C == 1, "sdf_grad: the SDF grid must have one channel");
  VX_CHECK_ARG(n >= 0, "sdf_grad: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_sdf_grad_bwd<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           MP(ggrid), g, CP(xyz), CP(xyz_min), CP(xyz_max), (scalar_t)voxel_size,
                           n, CP(gsdf), CP(ggrad), CP(gfeat)));
  return launch_status("sdf_grad_bwd");
}

int vx_dense_sdf_gradient_fwd(int dtype, const void* sdf, int X, int Y, int Z, double voxel_size,
                              void* out, void* stream) {
  VX_CHECK_ARG(X >= 1 && Y >= 1 && Z >= 1, "dense gradient: bad shape");
  VX_CHECK_ARG(X <= 65535 && Y <= 65535, "dense gradient: X and Y must be <= 65535");
  const dim3 grid((unsigned)cdiv(Z, 256), (unsigned)Y, (unsigned)X);
  GRID_DISPATCH(dtype, k_dense_grad_fwd<scalar_t><<<grid, 256, 0, ST(stream)>>>(
                           CP(sdf), X, Y, Z, (scalar_t)voxel_size, MP(out)));
  return launch_status("dense_sdf_gradient_fwd");
}

int vx_dense_sdf_gradient_bwd(int dtype, const void* gout, int X, int Y, int Z,
                              double voxel_size, void* gsdf, void* stream) {
  VX_CHECK_ARG(X >= 1 && Y >= 1 && Z >= 1, "dense gradient: bad shape");
  VX_CHECK_ARG(X <= 65535 && Y <= 65535, "dense gradient: X and Y must be <= 65535");
  const dim3 grid((unsigned)cdiv(Z, 256), (unsigned)Y, (unsigned)X);
  GRID_DISPATCH(dtype, k_dense_grad_bwd<scalar_t><<<grid, 256, 0, ST(stream)>>>(
                           CP(gout), X, Y, Z, (scalar_t)voxel_size, MP(gsdf)));
  return launch_status("dense_sdf_gradient_bwd");
}

int vx_total_variation_add_grad(int dtype, const void* param, void* grad, const int64_t* geom,
                                float wx, float wy, float wz, int dense_mode, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  const int64_t N = (int64_t)g.C * g.X * g.Y * g.Z;
  // wx /= 6; wy /= 6; wz /= 6 (total_variation_kernel.cu:76-78); the k and i axes both
  // take wz (total_variation_kernel.cu:27-32).
  wx /= 6;
  wy /= 6;
  wz /= 6;
  (void)wx;
  GRID_DISPATCH(dtype, k_tv<scalar_t, false><<<grid_for(N, 256), 256, 0, ST(stream)>>>(
                           CP(param), MP(grad), nullptr, g, wz, wy, wz, dense_mode));
  return launch_status("total_variation_add_grad");
}

int vx_total_variation_add_grad_new(int dtype, const void* param, void* grad, const void* mask,
                                    const int64_t* geom, float wx, float wy, float wz,
                                    int dense_mode, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  const int64_t N = (int64_t)g.C * g.X * g.Y * g.Z;
  wx /= 6;
  wy /= 6;
  wz /= 6;
  GRID_DISPATCH(dtype, k_tv<scalar_t, true><<<grid_for(N, 256), 256, 0, ST(stream)>>>(
                           CP(param), MP(grad), CP(mask), g, wx, wy, wz, dense_mode));
  return launch_status("total_variation_add_grad_new");
}

int vx_maskcache_alpha(int dtype, const void* density, const int64_t* geom, const void* xyz,
                       const void* xyz_min, const void* xyz_max, double act_shift,
                       double voxel_size_ratio, double thres, int nearest, int64_t n,
                       uint8_t* out, void* stream) {
  GridDesc g;
  if (int st = make_desc(geom, g)) return st;
  VX_CHECK_ARG(g.C == 1, "maskcache: the density grid must have one channel");
  VX_CHECK_ARG(n >= 0, "maskcache: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_maskcache_alpha<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(
                           CP(density), g, CP(xyz), CP(xyz_min), CP(xyz_max),
                           (scalar_t)act_shift, (scalar_t)voxel_size_ratio, (scalar_t)thres,
                           nearest, n, out));
  return launch_status("maskcache_alpha");
}

int vx_zero(int dtype, void* p, int64_t n, void* stream) {
  VX_CHECK_ARG(n >= 0, "zero: n < 0");
  if (n == 0) return VX_OK;
  GRID_DISPATCH(dtype, k_zero<scalar_t><<<grid_for(n, 256), 256, 0, ST(stream)>>>(MP(p), n));
  return launch_status("zero");
}

}  // extern "C"
