// Camera rays and the MaskCache ray pre-filter of the fine stage (gfx950).
//
// get_rays_of_a_view (lib/voxurf_fine.py:1001-1068) builds the rays of one pinhole view
// with ~15 full-image torch passes; here one thread makes one pixel's ray in registers.
// hit_coarse_geo (lib/voxurf_fine.py:579-591) samples every step of every ray, then looks
// up the MaskCache at every in-box point and scatters the hits; here one wavefront walks
// one ray, 64 steps per iteration, and stops at the first step that passes the mask.
// get_training_rays_in_maskcache_sampling (lib/voxurf_fine.py:1127-1164) becomes, per
// view: hit flags straight from the camera (no ray arrays), an inclusive scan, and a
// compaction that rebuilds the kept rays with the same arithmetic and writes them in
// pixel order, as the reference's boolean indexing does.
#include "fine.h"

namespace vx {

struct Camera {
  int H, W;
  float fx, fy, cx, cy;  // K[0][0], K[1][1], K[0][2], K[1][2]
  float R[9];            // c2w[:3, :3], row-major
  float t[3];            // c2w[:3, 3]
  float off;             // 0.5 (pixel centres) or 0 (left-top corners)
  int inverse_y, flip_x, flip_y;
  int ndc;
  float ndc_c0, ndc_c1, ndc_near;  // -1/(W/(2 focal)), -1/(H/(2 focal)), near plane
};

// NDC warp of one ray (lib/voxurf_fine.py:1042-1059), in the reference's operation
// order: shift the origin to z = -near, then project.
__device__ __forceinline__ void ndc_ray(float c0, float c1, float near_, float o[3], float d[3]) {
#pragma clang fp contract(off)
  const float t = -(near_ + o[2]) / d[2];
  const float ox = o[0] + t * d[0], oy = o[1] + t * d[1], oz = o[2] + t * d[2];
  const float o0 = c0 * ox / oz;
  const float o1 = c1 * oy / oz;
  const float o2 = 1.f + 2.f * near_ / oz;
  const float d0 = c0 * (d[0] / d[2] - ox / oz);
  const float d1 = c1 * (d[1] / d[2] - oy / oz);
  const float d2 = -2.f * near_ / oz;
  o[0] = o0;
  o[1] = o1;
  o[2] = o2;
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
}

// Ray of pixel (row, col) (lib/voxurf_fine.py:1001-1029, 1062-1068): camera-frame
// direction from the pixel coordinate, rotated by c2w (sum of the three products in
// index order), viewdir = d / |d| before the optional NDC warp.
__device__ __forceinline__ void camera_ray(const Camera& c, int row, int col, float o[3],
                                           float d[3], float v[3]) {
#pragma clang fp contract(off)
  const float i = (float)(c.flip_x ? c.W - 1 - col : col) + c.off;
  const float j = (float)(c.flip_y ? c.H - 1 - row : row) + c.off;
  float dir[3];
  dir[0] = (i - c.cx) / c.fx;
  if (c.inverse_y) {
    dir[1] = (j - c.cy) / c.fy;
    dir[2] = 1.f;
  } else {
    dir[1] = -((j - c.cy) / c.fy);
    dir[2] = -1.f;
  }
  for (int r = 0; r < 3; ++r) {
    d[r] = (dir[0] * c.R[3 * r] + dir[1] * c.R[3 * r + 1]) + dir[2] * c.R[3 * r + 2];
    o[r] = c.t[r];
  }
  const float nrm = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  v[0] = d[0] / nrm;
  v[1] = d[1] / nrm;
  v[2] = d[2] / nrm;
  if (c.ndc) ndc_ray(c.ndc_c0, c.ndc_c1, c.ndc_near, o, d);
}

__global__ void __launch_bounds__(256) k_camera_rays(Camera c, float* __restrict__ rays_o,
                                                     float* __restrict__ rays_d,
                                                     float* __restrict__ viewdirs) {
  const int64_t n = (int64_t)c.H * c.W;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    float o[3], d[3], v[3];
    camera_ray(c, (int)(p / c.W), (int)(p % c.W), o, d, v);
    for (int k = 0; k < 3; ++k) {
      rays_o[3 * p + k] = o[k];
      rays_d[3 * p + k] = d[k];
      viewdirs[3 * p + k] = v[k];
    }
  }
}

__global__ void __launch_bounds__(256) k_ndc_rays(int64_t n, float c0, float c1, float near_,
                                                  const float* __restrict__ o_in,
                                                  const float* __restrict__ d_in,
                                                  float* __restrict__ o_out,
                                                  float* __restrict__ d_out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    float o[3] = {o_in[3 * p], o_in[3 * p + 1], o_in[3 * p + 2]};
    float d[3] = {d_in[3 * p], d_in[3 * p + 1], d_in[3 * p + 2]};
    ndc_ray(c0, c1, near_, o, d);
    for (int k = 0; k < 3; ++k) {
      o_out[3 * p + k] = o[k];
      d_out[3 * p + k] = d[k];
    }
  }
}

// Does any in-box sample of the ray pass the MaskCache?  Samples exactly as
// sample_pts_on_rays; the wave leaves at the first chunk holding a passing sample.
__device__ __forceinline__ bool ray_hits(const FusedGeom& G, const float* o, const float* d,
                                         int lane) {
  float start[3], dir[3];
  int64_t n;
  ray_setup(G, o, d, start, dir, n);
  for (int64_t base = 0; base < n; base += 64) {
    const int64_t s = base + lane;
    bool pass = false;
    if (s < n) {
      float q[3];
      if (!sample_point<float>(start, dir, G.mn, G.mx, G.stepdist, s, q[0], q[1], q[2]))
        pass = maskcache_pass(G, q);
    }
    if (__ballot(pass)) return true;
  }
  return false;
}

// One wavefront per ray: hit (bool bytes) and/or flag (int64 0/1) per ray.
__global__ void __launch_bounds__(256) k_hit_rays(FusedGeom G, const float* __restrict__ rays_o,
                                                  const float* __restrict__ rays_d, int64_t n_rays,
                                                  uint8_t* __restrict__ hit,
                                                  int64_t* __restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); r < n_rays;
       r += waves) {
    const bool h = ray_hits(G, rays_o + 3 * r, rays_d + 3 * r, lane);
    if (lane == 0) {
      if (hit) hit[r] = h;
      if (flag) flag[r] = h;
    }
  }
}

// The same, with each pixel's ray made by the wave itself (every lane computes it: a
// few dozen flops, cheaper than a broadcast).
__global__ void __launch_bounds__(256) k_hit_camera(FusedGeom G, Camera c,
                                                    int64_t* __restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int64_t n = (int64_t)c.H * c.W;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t p = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6); p < n;
       p += waves) {
    float o[3], d[3], v[3];
    camera_ray(c, (int)(p / c.W), (int)(p % c.W), o, d, v);
    const bool h = ray_hits(G, o, d, lane);
    if (lane == 0) flag[p] = h;
  }
}

// Kept pixels of one view, in pixel order: row incl[p] - 1 of the outputs receives the
// pixel's colour and rebuilt ray when its flag (incl[p] - incl[p - 1]) is set.
__global__ void __launch_bounds__(256) k_compact_view(Camera c, const float* __restrict__ rgb,
                                                      const int64_t* __restrict__ incl,
                                                      float* __restrict__ rgb_out,
                                                      float* __restrict__ o_out,
                                                      float* __restrict__ d_out,
                                                      float* __restrict__ v_out) {
  const int64_t n = (int64_t)c.H * c.W;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t prev = p ? incl[p - 1] : 0;
    if (incl[p] == prev) continue;
    float o[3], d[3], v[3];
    camera_ray(c, (int)(p / c.W), (int)(p % c.W), o, d, v);
    const int64_t k = prev;
    for (int j = 0; j < 3; ++j) {
      rgb_out[3 * k + j] = rgb[3 * p + j];
      o_out[3 * k + j] = o[j];
      d_out[3 * k + j] = d[j];
      v_out[3 * k + j] = v[j];
    }
  }
}

}  // namespace vx

using namespace vx;

// cam = {H, W, fx, fy, cx, cy, R[9], t[3], off, inverse_y, flip_x, flip_y, ndc, ndc_c0,
// ndc_c1, ndc_near} as 26 floats (host memory).
static int camera_from(Camera& c, const float* cam) {
  VX_CHECK_ARG(cam, "camera: parameters missing");
  c.H = (int)cam[0];
  c.W = (int)cam[1];
  VX_CHECK_ARG(c.H >= 0 && c.W >= 0 && (float)c.H == cam[0] && (float)c.W == cam[1],
               "camera: H and W must be non-negative integers");
  c.fx = cam[2];
  c.fy = cam[3];
  c.cx = cam[4];
  c.cy = cam[5];
  for (int i = 0; i < 9; ++i) c.R[i] = cam[6 + i];
  for (int i = 0; i < 3; ++i) c.t[i] = cam[15 + i];
  c.off = cam[18];
  c.inverse_y = cam[19] != 0.f;
  c.flip_x = cam[20] != 0.f;
  c.flip_y = cam[21] != 0.f;
  c.ndc = cam[22] != 0.f;
  c.ndc_c0 = cam[23];
  c.ndc_c1 = cam[24];
  c.ndc_near = cam[25];
  return VX_OK;
}

// Sampling geometry of the pre-filter: box {mn, mx}, params {stepdist, near}.
static int prefilter_geom(FusedGeom& G, const float* box, const float* params, const float* mc,
                          const int64_t* mc_geom, const float* mc_params) {
  VX_CHECK_ARG(box && params, "hit_coarse_geo: box/parameters missing");
  VX_CHECK_ARG(mc, "hit_coarse_geo: the mask cache is required");
  G = FusedGeom{};
  for (int i = 0; i < 3; ++i) {
    G.mn[i] = box[i];
    G.mx[i] = box[3 + i];
  }
  G.stepdist = params[0];
  G.near_ = params[1];
  G.far_ = 1e9f;
  return fused_maskcache(G, mc, mc_geom, mc_params);
}

extern "C" {

int vx_camera_rays(const float* cam, float* rays_o, float* rays_d, float* viewdirs, void* stream) {
  Camera c;
  if (int st = camera_from(c, cam)) return st;
  const int64_t n = (int64_t)c.H * c.W;
  if (n == 0) return VX_OK;
  k_camera_rays<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(c, rays_o, rays_d, viewdirs);
  return launch_status("camera_rays");
}

int vx_ndc_rays(int64_t n, float c0, float c1, float near_, const float* rays_o,
                const float* rays_d, float* o_out, float* d_out, void* stream) {
  VX_CHECK_ARG(n >= 0, "ndc_rays: n < 0");
  if (n == 0) return VX_OK;
  k_ndc_rays<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, c0, c1, near_, rays_o, rays_d,
                                                                o_out, d_out);
  return launch_status("ndc_rays");
}

int vx_hit_coarse_geo(const float* rays_o, const float* rays_d, int64_t n_rays, const float* box,
                      const float* params, const float* mc, const int64_t* mc_geom,
                      const float* mc_params, uint8_t* hit, int64_t* flag, void* stream) {
  FusedGeom G;
  if (int st = prefilter_geom(G, box, params, mc, mc_geom, mc_params)) return st;
  VX_CHECK_ARG(n_rays >= 0, "hit_coarse_geo: n_rays < 0");
  if (n_rays == 0) return VX_OK;
  VX_CHECK_ARG(hit || flag, "hit_coarse_geo: no output");
  k_hit_rays<<<grid_for(n_rays, 4, 1 << 16), 256, 0, (hipStream_t)stream>>>(G, rays_o, rays_d,
                                                                           n_rays, hit, flag);
  return launch_status("hit_coarse_geo");
}

int vx_view_hit_coarse_geo(const float* cam, const float* box, const float* params,
                           const float* mc, const int64_t* mc_geom, const float* mc_params,
                           int64_t* flag, void* stream) {
  Camera c;
  if (int st = camera_from(c, cam)) return st;
  FusedGeom G;
  if (int st = prefilter_geom(G, box, params, mc, mc_geom, mc_params)) return st;
  const int64_t n = (int64_t)c.H * c.W;
  if (n == 0) return VX_OK;
  k_hit_camera<<<grid_for(n, 4, 1 << 16), 256, 0, (hipStream_t)stream>>>(G, c, flag);
  return launch_status("view_hit_coarse_geo");
}

int vx_view_compact_rays(const float* cam, const float* rgb, const int64_t* incl, float* rgb_out,
                         float* o_out, float* d_out, float* v_out, void* stream) {
  Camera c;
  if (int st = camera_from(c, cam)) return st;
  const int64_t n = (int64_t)c.H * c.W;
  if (n == 0) return VX_OK;
  k_compact_view<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(c, rgb, incl, rgb_out, o_out,
                                                                    d_out, v_out);
  return launch_status("view_compact_rays");
}

}  // extern "C"
