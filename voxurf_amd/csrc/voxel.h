// Device helpers shared by the drop-in ops and the fused per-ray kernels, so both
// paths evaluate every sample with the same operations in the same order.
#pragma once
#include "common.h"

namespace vx {

// ---- ray geometry (lib/cuda/render_utils_kernel.cu:12-79, 184-192) ----------------

template <typename T>
__device__ __forceinline__ void ray_t_minmax(const T* o, const T* d, const T* xyz_min,
                                             const T* xyz_max, float near_, float far_,
                                             float& tmin, float& tmax) {
#pragma clang fp contract(off)
  // `(d==0) ? 1e-6 : d` has type double in the reference; the result is stored in a float.
  const float vx = (d[0] == 0) ? (float)1e-6 : (float)d[0];
  const float vy = (d[1] == 0) ? (float)1e-6 : (float)d[1];
  const float vz = (d[2] == 0) ? (float)1e-6 : (float)d[2];
  const float ax = (float)((xyz_max[0] - o[0]) / (T)vx);
  const float ay = (float)((xyz_max[1] - o[1]) / (T)vy);
  const float az = (float)((xyz_max[2] - o[2]) / (T)vz);
  const float bx = (float)((xyz_min[0] - o[0]) / (T)vx);
  const float by = (float)((xyz_min[1] - o[1]) / (T)vy);
  const float bz = (float)((xyz_min[2] - o[2]) / (T)vz);
  tmin = fmaxf(fminf(fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz)), far_), near_);
  tmax = fmaxf(fminf(fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)), far_), near_);
}

template <typename T>
__device__ __forceinline__ float ray_norm(const T* d) {
#pragma clang fp contract(off)
  const T s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return (float)sqrt(s);
}

template <typename T>
__device__ __forceinline__ int64_t ray_n_samples(const T* d, T tmin, T tmax, float stepdist) {
#pragma clang fp contract(off)
  const float rnorm = ray_norm(d);
  const T q = (tmax - tmin) * (T)rnorm / (T)stepdist;
  const double c = ceil((double)q);
  return (int64_t)(c > 1.0 ? c : 1.0);
}

template <typename T>
__device__ __forceinline__ void ray_start_dir(const T* o, const T* d, T tmin, T* start, T* dir) {
#pragma clang fp contract(off)
  const float rnorm = ray_norm(d);
  start[0] = o[0] + d[0] * tmin;
  start[1] = o[1] + d[1] * tmin;
  start[2] = o[2] + d[2] * tmin;
  dir[0] = d[0] / (T)rnorm;
  dir[1] = d[1] / (T)rnorm;
  dir[2] = d[2] / (T)rnorm;
}

// Point of step `s` on a ray (render_utils_kernel.cu:184-192): the coordinates are
// rounded to float whatever T is, and the bbox test compares in T.
template <typename T>
__device__ __forceinline__ bool sample_point(const T* start, const T* dir, const T* xyz_min,
                                             const T* xyz_max, float stepdist, int64_t s,
                                             float& px, float& py, float& pz) {
#pragma clang fp contract(off)
  const float dist = stepdist * (float)(int)s;
  px = (float)(start[0] + dir[0] * (T)dist);
  py = (float)(start[1] + dir[1] * (T)dist);
  pz = (float)(start[2] + dir[2] * (T)dist);
  return (xyz_min[0] > (T)px) | (xyz_min[1] > (T)py) | (xyz_min[2] > (T)pz) |
         (xyz_max[0] < (T)px) | (xyz_max[1] < (T)py) | (xyz_max[2] < (T)pz);
}


// ---- trilinear lookups on dense grids -----------------------------------------------

struct GridDesc {
  int64_t sc, sx, sy, sz;  // element strides
  int C, X, Y, Z;
};

// World coordinate -> continuous index along an axis of `size` lattice points spanning
// [mn, mx]:  n = ((x - mn) / (mx - mn)) * 2 - 1;  i = ((n + 1) / 2) * (size - 1)
// (lib/grid.py:53 followed by grid_sample's align_corners unnormalisation).
template <typename T>
__device__ __forceinline__ T world_to_index(T x, T mn, T ext, int size) {
#pragma clang fp contract(off)
  const T n = ((x - mn) / ext) * (T)2 - (T)1;
  return ((n + (T)1) / (T)2) * (T)(size - 1);
}

// The 8 corners of a trilinear lookup at continuous index (fx, fy, fz).
template <typename T>
struct Tri {
  int x0, y0, z0;
  T wx[2], wy[2], wz[2];
  __device__ __forceinline__ void init(T fx, T fy, T fz) {
#pragma clang fp contract(off)
    const T ffx = floor(fx), ffy = floor(fy), ffz = floor(fz);
    x0 = sat_f2i(ffx);
    y0 = sat_f2i(ffy);
    z0 = sat_f2i(ffz);
    wx[0] = (ffx + (T)1) - fx;
    wx[1] = fx - ffx;
    wy[0] = (ffy + (T)1) - fy;
    wy[1] = fy - ffy;
    wz[0] = (ffz + (T)1) - fz;
    wz[1] = fz - ffz;
  }
  // weight of corner (bx, by, bz); PyTorch multiplies W-, H-, then D-axis terms, i.e.
  // z, y, x in world axes.
  __device__ __forceinline__ T w(int bx, int by, int bz) const {
#pragma clang fp contract(off)
    return wz[bz] * wy[by] * wx[bx];
  }
};

template <typename T>
__device__ __forceinline__ bool in_grid(int x, int y, int z, const GridDesc& g) {
  return (unsigned)x < (unsigned)g.X && (unsigned)y < (unsigned)g.Y && (unsigned)z < (unsigned)g.Z;
}

__device__ __forceinline__ int clamp_index(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Gather one channel: sum over corners in grid_sample's order (x outer, z inner),
// corners outside the grid skipped (zero padding).  The eight loads are unconditional
// (addresses clamped into the grid) so they are all in flight together; the sum then
// selects the in-grid terms, with the same operations in the same order.
template <typename T>
__device__ __forceinline__ void tri_load(const T* __restrict__ grid, const GridDesc& g,
                                         const Tri<T>& t, int c, T v[8], unsigned& in) {
  in = 0u;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = t.x0 + (k >> 2), y = t.y0 + (k >> 1 & 1), z = t.z0 + (k & 1);
    if (in_grid<T>(x, y, z, g)) in |= 1u << k;
    v[k] = grid[c * g.sc + clamp_index(x, g.X) * g.sx + clamp_index(y, g.Y) * g.sy +
                clamp_index(z, g.Z) * g.sz];
  }
}

template <typename T>
__device__ __forceinline__ T tri_sum(const Tri<T>& t, const T v[8], unsigned in) {
#pragma clang fp contract(off)
  T acc = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (in >> k & 1u) acc += v[k] * t.w(k >> 2, k >> 1 & 1, k & 1);
  return acc;
}

template <typename T>
__device__ __forceinline__ T tri_gather(const T* __restrict__ grid, const GridDesc& g,
                                        const Tri<T>& t, int c) {
  T v[8];
  unsigned in;
  tri_load(grid, g, t, c, v, in);
  return tri_sum(t, v, in);
}

// Four consecutive channels c0 .. c0+3 of a channels-innermost grid (channel stride 1,
// 16-byte aligned rows): one 16-byte load per corner, each channel summed exactly as
// tri_gather sums it.
__device__ __forceinline__ float4 tri_gather4(const float* __restrict__ grid, const GridDesc& g,
                                             const Tri<float>& t, int c0) {
#pragma clang fp contract(off)
  float4 v[8];
  bool in[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = t.x0 + (k >> 2), y = t.y0 + (k >> 1 & 1), z = t.z0 + (k & 1);
    in[k] = in_grid<float>(x, y, z, g);
    v[k] = *reinterpret_cast<const float4*>(grid + c0 + clamp_index(x, g.X) * g.sx +
                                            clamp_index(y, g.Y) * g.sy + clamp_index(z, g.Z) * g.sz);
  }
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (!in[k]) continue;
    const float w = t.w(k >> 2, k >> 1 & 1, k & 1);
    acc.x += v[k].x * w;
    acc.y += v[k].y * w;
    acc.z += v[k].z * w;
    acc.w += v[k].w * w;
  }
  return acc;
}

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

template <typename T>
__device__ __forceinline__ void tri_scatter(T* __restrict__ grad, const GridDesc& g,
                                            const Tri<T>& t, int c, T go) {
#pragma clang fp contract(off)
  if (go == (T)0) return;
  for (int bx = 0; bx < 2; ++bx)
    for (int by = 0; by < 2; ++by)
      for (int bz = 0; bz < 2; ++bz) {
        const int x = t.x0 + bx, y = t.y0 + by, z = t.z0 + bz;
        if (in_grid<T>(x, y, z, g))
          atomic_add(&grad[c * g.sc + x * g.sx + y * g.sy + z * g.sz], t.w(bx, by, bz) * go);
      }
}

template <typename T>
__device__ __forceinline__ T roundtrip_index(T a, int size) {
#pragma clang fp contract(off)
  const T n = (a / (T)(size - 1)) * (T)2 - (T)1;
  return ((n + (T)1) / (T)2) * (T)(size - 1);
}

template <typename T>
__device__ __forceinline__ T clampT(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Continuous (zyx) index of a world point, as sample_sdfs computes it.
template <typename T>
__device__ __forceinline__ void zyx_index(const T* p, const T* mn, const T* mx, const GridDesc& g,
                                          T ind[3]) {
#pragma clang fp contract(off)
  // ind_norm (zyx) then ((ind_norm + 1) / 2) * (size_zyx - 1)
  const int sz[3] = {g.Z, g.Y, g.X};
  const int ax[3] = {2, 1, 0};
  for (int k = 0; k < 3; ++k) {
    const int a = ax[k];
    const T n = ((p[a] - mn[a]) / (mx[a] - mn[a])) * (T)2 - (T)1;
    ind[k] = ((n + (T)1) / (T)2) * (T)(sz[k] - 1);
  }
}

// Sample of the single-channel grid at the zyx index with per-component displacement;
// returns the value and the clamped displaced index along `axis`.
template <typename T>
__device__ __forceinline__ T sample_displaced(const T* __restrict__ grid, const GridDesc& g,
                                              const T ind[3], int axis, T off, T& clamped) {
#pragma clang fp contract(off)
  const int sz[3] = {g.Z, g.Y, g.X};
  T pos[3];
  for (int k = 0; k < 3; ++k) {
    T a = ind[k] + (k == axis ? off : (T)0);
    a = clampT(a, (T)0, (T)(sz[k] - 1));
    if (k == axis) clamped = a;
    pos[k] = roundtrip_index(a, sz[k]);
  }
  Tri<T> t;
  t.init(pos[2], pos[1], pos[0]);
  return tri_gather(grid, g, t, 0);
}

template <typename T>
__device__ __forceinline__ void scatter_displaced(T* __restrict__ grad, const GridDesc& g,
                                                  const T ind[3], int axis, T off, T go) {
#pragma clang fp contract(off)
  const int sz[3] = {g.Z, g.Y, g.X};
  T pos[3];
  for (int k = 0; k < 3; ++k) {
    T a = ind[k] + (k == axis ? off : (T)0);
    a = clampT(a, (T)0, (T)(sz[k] - 1));
    pos[k] = roundtrip_index(a, sz[k]);
  }
  Tri<T> t;
  t.init(pos[2], pos[1], pos[0]);
  tri_scatter(grad, g, t, 0, go);
}

// Clamped displaced index pair along `axis` (no sampling): (ind[+], ind[-]).
template <typename T>
__device__ __forceinline__ T clamped_diff(const T ind[3], int axis, T d, const GridDesc& g) {
#pragma clang fp contract(off)
  const int sz[3] = {g.Z, g.Y, g.X};
  const T hi = (T)(sz[axis] - 1);
  const T ap = clampT(ind[axis] + d, (T)0, hi);
  const T am = clampT(ind[axis] + (-d), (T)0, hi);
  // diff.max(dim=-1) over the 3 components: the other two components are 0
  const T df = ap - am;
  return df > (T)0 ? df : (T)0;
}


// ---- NeuS alpha (lib/voxurf_fine.py:463-500) ------------------------------------------

template <typename T>
__device__ __forceinline__ T sigmoidT(T x) {
  return (T)1 / ((T)1 + exp(-x));
}

// NeuS alpha, use_mid=True, cos_anneal_ratio=1:
//   cos  = <viewdir[ray], grad>
//   ic   = -relu(-cos)
//   prev = sdf - ic*dist*0.5,  next = sdf + ic*dist*0.5
//   c    = sigmoid(prev*inv_s), p = c - sigmoid(next*inv_s)
//   alpha= clip((p + 1e-5) / (c + 1e-5), 0, 1)
template <typename T>
struct NeusFwd {
  T cos, ic, c, nx, a;
};

template <typename T>
__device__ __forceinline__ NeusFwd<T> neus_fwd(const T* dir, const T* gr, T sdf, T dist, T inv_s) {
#pragma clang fp contract(off)
  NeusFwd<T> f;
  f.cos = dir[0] * gr[0] + dir[1] * gr[1] + dir[2] * gr[2];
  const T r = -f.cos > (T)0 ? -f.cos : (T)0;
  f.ic = -r;
  const T h = (f.ic * dist) * (T)0.5;
  const T nxt = sdf + h;
  const T prv = sdf - h;
  f.c = sigmoidT(prv * inv_s);
  f.nx = sigmoidT(nxt * inv_s);
  const T p = f.c - f.nx;
  f.a = (p + (T)1e-5) / (f.c + (T)1e-5);
  return f;
}

template <typename T>
__device__ __forceinline__ T clip01(T a) {
  return a < (T)0 ? (T)0 : (a > (T)1 ? (T)1 : a);
}

// Adjoint of clip(neus alpha) w.r.t. sdf and the gradient vector: clip passes the
// gradient on [0, 1] inclusive, relu(x) only for x > 0, sigmoid' = y (1 - y).
template <typename T>
__device__ __forceinline__ void neus_bwd(const T* dir, const NeusFwd<T>& f, T dist, T inv_s, T ga,
                                         T& gsdf, T gg[3]) {
#pragma clang fp contract(off)
  if (!(f.a >= (T)0 && f.a <= (T)1)) ga = 0;
  const T den = f.c + (T)1e-5;
  const T gp = ga / den;
  const T gc = -ga * ((f.c - f.nx + (T)1e-5) / (den * den)) + gp;
  const T gn = -gp;
  const T gprev = gc * (f.c * ((T)1 - f.c)) * inv_s;
  const T gnext = gn * (f.nx * ((T)1 - f.nx)) * inv_s;
  gsdf = gprev + gnext;
  const T gic = (gnext - gprev) * dist * (T)0.5;
  const T gcos = (-f.cos > (T)0) ? gic : (T)0;
  gg[0] = gcos * dir[0];
  gg[1] = gcos * dir[1];
  gg[2] = gcos * dir[2];
}

// ---- centre SDF + 6-neighbour gradient (lib/voxurf_fine.py:502-534, 537-577) ----------
// sdf at the point, grad[3] in xyz order (unnormalised), optional feat[6]
// (-x,+x,-y,+y,-z,+z).

// Trilinear cell of the single-channel grid at the zyx index displaced along `axis`
// (clamped), as sample_displaced / scatter_displaced compute it.
template <typename T>
__device__ __forceinline__ void displaced_cell(const GridDesc& g, const T ind[3], int axis, T off,
                                               Tri<T>& t) {
#pragma clang fp contract(off)
  const int sz[3] = {g.Z, g.Y, g.X};
  T pos[3];
  for (int k = 0; k < 3; ++k) {
    T a = ind[k] + (k == axis ? off : (T)0);
    a = clampT(a, (T)0, (T)(sz[k] - 1));
    pos[k] = roundtrip_index(a, sz[k]);
  }
  t.init(pos[2], pos[1], pos[0]);
}

template <typename T>
__device__ __forceinline__ void sdf7(const T* __restrict__ grid, const GridDesc& g, const T q[3],
                                     const T* mn, const T* mx, T voxel_size, T& sdf, T grad[3],
                                     T* feat, bool* shared_out = nullptr) {
#pragma clang fp contract(off)
  // the seven cells first, then all corner loads, then the sums: one memory latency per
  // sample instead of seven
  Tri<T> t[7];
  t[0].init(world_to_index(q[0], mn[0], mx[0] - mn[0], g.X),
            world_to_index(q[1], mn[1], mx[1] - mn[1], g.Y),
            world_to_index(q[2], mn[2], mx[2] - mn[2], g.Z));
  T ind[3];
  zyx_index(q, mn, mx, g, ind);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    displaced_cell(g, ind, a, (T)-1, t[1 + 2 * a]);
    displaced_cell(g, ind, a, (T)1, t[2 + 2 * a]);
  }
  // In the interior a displaced cell is the centre cell shifted by one voxel along its
  // axis (or, where the index was clamped at a face, the centre cell itself), so the
  // seven cells cover 32 distinct voxels: the centre's 8 and, per axis and side, the 4
  // beyond it.  When that holds for every lane, only those 32 are loaded and each cell's
  // 8 corner values are selected from them (the same values, weights and sums as 56
  // separate loads); otherwise every cell loads its own corners.
  int dd[6];
  bool shared = true;
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    const int c = 2 - (i - 1) / 2, s = (i - 1) % 2 ? 1 : -1;  // world axis, side
    const int d[3] = {t[i].x0 - t[0].x0, t[i].y0 - t[0].y0, t[i].z0 - t[0].z0};
    dd[i - 1] = d[c];
    shared = shared && d[(c + 1) % 3] == 0 && d[(c + 2) % 3] == 0 && (d[c] == 0 || d[c] == s);
  }
  if (shared_out) *shared_out = shared;
  T f[7];  // the seven lookups: centre, then -/+ along z, y, x of the zyx index
  // element offsets in 32 bits (the grid is below 2^31 elements; uniform)
  const bool idx32 = (int64_t)(g.X - 1) * g.sx + (int64_t)(g.Y - 1) * g.sy +
                         (int64_t)(g.Z - 1) * g.sz < ((int64_t)1 << 31);
  if (idx32 && __all(shared)) {
    // per axis, the clamped offsets and in-grid flags of lattice lines base-1 .. base+2
    int ox[4], oy[4], oz[4];
    unsigned vx = 0u, vy = 0u, vz = 0u;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int x = t[0].x0 + d - 1, y = t[0].y0 + d - 1, z = t[0].z0 + d - 1;
      vx |= ((unsigned)x < (unsigned)g.X ? 1u : 0u) << d;
      vy |= ((unsigned)y < (unsigned)g.Y ? 1u : 0u) << d;
      vz |= ((unsigned)z < (unsigned)g.Z ? 1u : 0u) << d;
      ox[d] = clamp_index(x, g.X) * (int)g.sx;
      oy[d] = clamp_index(y, g.Y) * (int)g.sy;
      oz[d] = clamp_index(z, g.Z) * (int)g.sz;
    }
    auto ld = [&](int dx, int dy, int dz) { return grid[ox[dx] + oy[dy] + oz[dz]]; };
    auto ok = [&](int dx, int dy, int dz) {
      return ((vx >> dx) & (vy >> dy) & (vz >> dz) & 1u) != 0u;
    };
    T v0[8], vo[3][2][4];
    unsigned in0 = 0u, ino[3][2];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int dx = 1 + (k >> 2), dy = 1 + ((k >> 1) & 1), dz = 1 + (k & 1);
      v0[k] = ld(dx, dy, dz);
      if (ok(dx, dy, dz)) in0 |= 1u << k;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int lo = (c + 1) % 3 < (c + 2) % 3 ? (c + 1) % 3 : (c + 2) % 3;
        const int hi = (c + 1) % 3 < (c + 2) % 3 ? (c + 2) % 3 : (c + 1) % 3;
        ino[c][side] = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          // the voxel beyond the centre cell along axis c; j holds the corner bits of
          // the other two axes (the lower-numbered axis first)
          int o[3];
          o[c] = side ? 3 : 0;
          o[lo] = 1 + (j >> 1);
          o[hi] = 1 + (j & 1);
          vo[c][side][j] = ld(o[0], o[1], o[2]);
          if (ok(o[0], o[1], o[2])) ino[c][side] |= 1u << j;
        }
      }
    f[0] = tri_sum(t[0], v0, in0);
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      const int c = 2 - (i - 1) / 2, side = (i - 1) % 2, s = side ? 1 : -1;
      const int lo = (c + 1) % 3 < (c + 2) % 3 ? (c + 1) % 3 : (c + 2) % 3;
      const int hi = (c + 1) % 3 < (c + 2) % 3 ? (c + 2) % 3 : (c + 1) % 3;
      const bool same = dd[i - 1] == 0;
      T vi[8];
      unsigned m = 0u;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int bb[3] = {b >> 2, (b >> 1) & 1, b & 1};
        const int cb = bb[c] + s;  // the corner's index along c in the centre cell when shifted
        T val;
        bool inb;
        if (cb == 0 || cb == 1) {
          const int k = b ^ ((bb[c] ^ cb) << (2 - c));
          val = v0[k];
          inb = (in0 >> k) & 1u;
        } else {
          const int j = bb[lo] * 2 + bb[hi];
          val = vo[c][side][j];
          inb = (ino[c][side] >> j) & 1u;
        }
        // not shifted (clamped at a face): the centre cell's own corner b
        vi[b] = same ? v0[b] : val;
        if (same ? ((in0 >> b) & 1u) : inb) m |= 1u << b;
      }
      f[i] = tri_sum(t[i], vi, m);
    }
  } else {
    T v[7][8];
    unsigned in[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) tri_load(grid, g, t[i], 0, v[i], in[i]);
#pragma unroll
    for (int i = 0; i < 7; ++i) f[i] = tri_sum(t[i], v[i], in[i]);
  }
  sdf = f[0];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T fm = f[1 + 2 * a];
    const T fp = f[2 + 2 * a];
    const T df = clamped_diff(ind, a, (T)1, g);
    grad[2 - a] = ((fp - fm) / df) / voxel_size;
    if (feat) {
      feat[2 * (2 - a)] = fm;
      feat[2 * (2 - a) + 1] = fp;
    }
  }
}

// Set the page flag (pages[element >> kPageShift]) of every voxel the 7-point stencil of
// q reaches, i.e. every voxel sdf7_scatter / sdf7_terms can add to: the 32 voxels around
// the centre cell when every displaced cell is the shifted or clamped centre cell
// (`shared`, as sdf7 reports it), else each cell's own 8 corners besides.
template <typename T>
__device__ __forceinline__ void sdf7_mark_pages(uint32_t* __restrict__ pages, const GridDesc& g,
                                                const T q[3], const T* mn, const T* mx,
                                                bool shared) {
#pragma clang fp contract(off)
  Tri<T> t0;
  t0.init(world_to_index(q[0], mn[0], mx[0] - mn[0], g.X),
          world_to_index(q[1], mn[1], mx[1] - mn[1], g.Y),
          world_to_index(q[2], mn[2], mx[2] - mn[2], g.Z));
  auto mark = [&](int x, int y, int z) {
    if (in_grid<T>(x, y, z, g)) pages[(x * g.sx + y * g.sy + z * g.sz) >> kPageShift] = 1u;
  };
#pragma unroll
  for (int k = 0; k < 8; ++k) mark(t0.x0 + (k >> 2), t0.y0 + ((k >> 1) & 1), t0.z0 + (k & 1));
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int lo = (c + 1) % 3 < (c + 2) % 3 ? (c + 1) % 3 : (c + 2) % 3;
    const int hi = (c + 1) % 3 < (c + 2) % 3 ? (c + 2) % 3 : (c + 1) % 3;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int v[3] = {t0.x0, t0.y0, t0.z0};
        v[c] += sd ? 2 : -1;
        v[lo] += j >> 1;
        v[hi] += j & 1;
        mark(v[0], v[1], v[2]);
      }
  }
  if (!shared) {
    T ind[3];
    zyx_index(q, mn, mx, g, ind);
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        Tri<T> td;
        displaced_cell(g, ind, a, side ? (T)1 : (T)-1, td);
        for (int k = 0; k < 8; ++k) mark(td.x0 + (k >> 2), td.y0 + ((k >> 1) & 1), td.z0 + (k & 1));
      }
  }
}

// Adjoint of sdf7: ggrid += d(sdf, grad, feat)^T (gsdf, ggrad, gfeat); gfeat may be null.
// The seven trilinear scatters (centre, then -/+ one voxel along z, y, x of the zyx
// index) add exactly the reference's per-corner terms; where a displaced cell is the
// centre cell shifted by one voxel (the interior case), the four corners it shares with
// the centre are summed in registers first, so a sample issues 32 atomic adds instead of
// 56.
template <typename T>
__device__ __forceinline__ void sdf7_scatter(T* __restrict__ ggrid, const GridDesc& g,
                                             const T q[3], const T* mn, const T* mx,
                                             T voxel_size, T gsdf, const T ggrad[3],
                                             const T* gfeat) {
#pragma clang fp contract(off)
  Tri<T> t0;
  t0.init(world_to_index(q[0], mn[0], mx[0] - mn[0], g.X),
          world_to_index(q[1], mn[1], mx[1] - mn[1], g.Y),
          world_to_index(q[2], mn[2], mx[2] - mn[2], g.Z));
  T acc[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) acc[b] = gsdf != (T)0 ? t0.w(b >> 2, (b >> 1) & 1, b & 1) * gsdf : (T)0;
  T ind[3];
  zyx_index(q, mn, mx, g, ind);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = 2 - a;  // world axis of zyx axis a (0 = x)
    const T df = clamped_diff(ind, a, (T)1, g);
    const T gd = (ggrad[c] / df) / voxel_size;
    const T gm = (gfeat ? gfeat[2 * c] : (T)0) - gd;
    const T gp = (gfeat ? gfeat[2 * c + 1] : (T)0) + gd;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int s = side ? 1 : -1;
      const T go = side ? gp : gm;
      if (go == (T)0) continue;
      Tri<T> td;
      displaced_cell(g, ind, a, (T)s, td);
      const int d[3] = {td.x0 - t0.x0, td.y0 - t0.y0, td.z0 - t0.z0};
      const bool shifted = d[c] == s && d[(c + 1) % 3] == 0 && d[(c + 2) % 3] == 0;
      if (!shifted) {
        tri_scatter(ggrid, g, td, 0, go);
        continue;
      }
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int bb[3] = {b >> 2, (b >> 1) & 1, b & 1};
        const T w = td.w(bb[0], bb[1], bb[2]) * go;
        const int cb = bb[c] + s;  // this corner's index along axis c in the centre cell
        if (cb == 0 || cb == 1) {
          int k[3] = {bb[0], bb[1], bb[2]};
          k[c] = cb;
          acc[k[0] * 4 + k[1] * 2 + k[2]] += w;
        } else {
          const int x = td.x0 + bb[0], y = td.y0 + bb[1], z = td.z0 + bb[2];
          if (in_grid<T>(x, y, z, g)) atomic_add(&ggrid[x * g.sx + y * g.sy + z * g.sz], w);
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int x = t0.x0 + (b >> 2), y = t0.y0 + ((b >> 1) & 1), z = t0.z0 + (b & 1);
    if (acc[b] != (T)0 && in_grid<T>(x, y, z, g)) atomic_add(&ggrid[x * g.sx + y * g.sy + z * g.sz], acc[b]);
  }
}


// sdf7_scatter in two halves, so that consecutive samples of a ray in the same cell can
// add their terms before the atomics.  sdf7_terms: the 56 corner terms of one sample
// (centre, then -/+ one voxel along z, y, x of the zyx index) summed per voxel of the
// 4x4x4 block around the centre cell that the seven cells can reach: acc[8] for the
// centre cell's corners, out[c][side][j] for the four voxels beyond it along world axis
// c (side 0: below, 1: above; j the corner bits of the other two axes, lower-numbered
// axis first).  A displaced cell that is neither the centre cell shifted by one voxel
// along its axis nor the centre cell itself (clamped at a face) is scattered directly.
// sdf7_emit adds a block's terms to the gradient grid.
template <typename T>
struct Sdf7Terms {
  int x0, y0, z0;
  T acc[8];
  T out[3][2][4];
};

template <typename T>
__device__ __forceinline__ void sdf7_terms(T* __restrict__ ggrid, const GridDesc& g, const T q[3],
                                           const T* mn, const T* mx, T voxel_size, T gsdf,
                                           const T ggrad[3], Sdf7Terms<T>& o) {
#pragma clang fp contract(off)
  Tri<T> t0;
  t0.init(world_to_index(q[0], mn[0], mx[0] - mn[0], g.X),
          world_to_index(q[1], mn[1], mx[1] - mn[1], g.Y),
          world_to_index(q[2], mn[2], mx[2] - mn[2], g.Z));
  o.x0 = t0.x0;
  o.y0 = t0.y0;
  o.z0 = t0.z0;
#pragma unroll
  for (int b = 0; b < 8; ++b) o.acc[b] = gsdf != (T)0 ? t0.w(b >> 2, (b >> 1) & 1, b & 1) * gsdf : (T)0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
      for (int j = 0; j < 4; ++j) o.out[c][sd][j] = (T)0;
  T ind[3];
  zyx_index(q, mn, mx, g, ind);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = 2 - a;  // world axis of zyx axis a (0 = x)
    const int lo = (c + 1) % 3 < (c + 2) % 3 ? (c + 1) % 3 : (c + 2) % 3;
    const int hi = (c + 1) % 3 < (c + 2) % 3 ? (c + 2) % 3 : (c + 1) % 3;
    const T df = clamped_diff(ind, a, (T)1, g);
    const T gd = (ggrad[c] / df) / voxel_size;
    const T gm = -gd, gp = gd;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int s = side ? 1 : -1;
      const T go = side ? gp : gm;
      if (go == (T)0) continue;
      Tri<T> td;
      displaced_cell(g, ind, a, (T)s, td);
      const int d[3] = {td.x0 - t0.x0, td.y0 - t0.y0, td.z0 - t0.z0};
      const bool others = d[lo] == 0 && d[hi] == 0;
      if (others && d[c] == 0) {
#pragma unroll
        for (int b = 0; b < 8; ++b) o.acc[b] += td.w(b >> 2, (b >> 1) & 1, b & 1) * go;
      } else if (others && d[c] == s) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const int bb[3] = {b >> 2, (b >> 1) & 1, b & 1};
          const T w = td.w(bb[0], bb[1], bb[2]) * go;
          const int cb = bb[c] + s;  // this corner's index along c in the centre cell
          if (cb == 0 || cb == 1)
            o.acc[b ^ ((bb[c] ^ cb) << (2 - c))] += w;
          else
            o.out[c][side][bb[lo] * 2 + bb[hi]] += w;
        }
      } else {
        tri_scatter(ggrid, g, td, 0, go);
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void sdf7_emit(T* __restrict__ ggrid, const GridDesc& g,
                                          const Sdf7Terms<T>& o) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int x = o.x0 + (b >> 2), y = o.y0 + ((b >> 1) & 1), z = o.z0 + (b & 1);
    if (o.acc[b] != (T)0 && in_grid<T>(x, y, z, g)) atomic_add(&ggrid[x * g.sx + y * g.sy + z * g.sz], o.acc[b]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int lo = (c + 1) % 3 < (c + 2) % 3 ? (c + 1) % 3 : (c + 2) % 3;
    const int hi = (c + 1) % 3 < (c + 2) % 3 ? (c + 2) % 3 : (c + 1) % 3;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        int v[3] = {o.x0, o.y0, o.z0};
        v[c] += sd ? 2 : -1;
        v[lo] += j >> 1;
        v[hi] += j & 1;
        if (o.out[c][sd][j] != (T)0 && in_grid<T>(v[0], v[1], v[2], g))
          atomic_add(&ggrid[v[0] * g.sx + v[1] * g.sy + v[2] * g.sz], o.out[c][sd][j]);
      }
  }
}

}  // namespace vx
