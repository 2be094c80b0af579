/*
 * voxhip.h -- C ABI of the voxurf_amd HIP library (libvoxhip.so, gfx950).
 *
 * Every entry point takes plain device pointers, element counts, scalars and the
 * hipStream_t to launch on (passed as void*, NULL = the null stream), and returns a
 * status code (VX_OK = 0).  On failure vx_last_error() returns a thread-local message.
 * Nothing here allocates or frees caller-visible memory: outputs and workspaces are
 * provided by the caller.  Index outputs are int64, matching the reference's tensors.
 *
 * `dtype` selects the floating type of the float arrays: VX_F32 or VX_F64 (the
 * reference dispatches both through AT_DISPATCH_FLOATING_TYPES).
 *
 * Each declaration cites the reference interface it replaces (paths relative to the
 * Voxurf repository).
 */
#ifndef VOXHIP_H
#define VOXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VX_OK = 0, VX_ERR_ARG = 1, VX_ERR_LAUNCH = 2, VX_ERR_SIZE = 3, VX_ERR_DTYPE = 4 };
enum { VX_F32 = 0, VX_F64 = 1 };

const char* vx_last_error(void);
int vx_version(void);

/* ---- render_utils: ray sampling ------------------------------------------------ */

/* lib/cuda/render_utils.cpp:50-58 infer_t_minmax (kernel render_utils_kernel.cu:12-35). */
int vx_infer_t_minmax(int dtype, const void* rays_o, const void* rays_d, const void* xyz_min,
                      const void* xyz_max, float near_, float far_, int64_t n_rays,
                      void* t_min, void* t_max, void* stream);

/* lib/cuda/render_utils.cpp:60-65 infer_n_samples (render_utils_kernel.cu:38-55). */
int vx_infer_n_samples(int dtype, const void* rays_d, const void* t_min, const void* t_max,
                       float stepdist, int64_t n_rays, int64_t* n_samples, void* stream);

/* lib/cuda/render_utils.cpp:67-72 infer_ray_start_dir (render_utils_kernel.cu:58-79). */
int vx_infer_ray_start_dir(int dtype, const void* rays_o, const void* rays_d, const void* t_min,
                           int64_t n_rays, void* rays_start, void* rays_dir, void* stream);

/* lib/cuda/render_utils.cpp:74-85 sample_pts_on_rays, phase 1 of 2
 * (render_utils_kernel.cu:196-212: t_min/t_max, N_steps, N_steps.cumsum, total).
 * Writes t_min/t_max [n_rays], n_steps [n_rays], n_steps_cumsum [n_rays] (inclusive)
 * and total [1] (device).  `scan_ws` is a device workspace of
 * vx_scan_workspace_elems(n_rays) int64 elements. */
int vx_sample_pts_count(int dtype, const void* rays_o, const void* rays_d, const void* xyz_min,
                        const void* xyz_max, float near_, float far_, float stepdist,
                        int64_t n_rays, void* t_min, void* t_max, int64_t* n_steps,
                        int64_t* n_steps_cumsum, int64_t* total, int64_t* scan_ws,
                        void* stream);

/* Phase 2 (render_utils_kernel.cu:213-241): fills rays_pts [M,3], mask_outbbox [M]
 * (bool, 1 byte), ray_id [M], step_id [M] for M = n_steps_cumsum[n_rays-1]. */
int vx_sample_pts_fill(int dtype, const void* rays_o, const void* rays_d, const void* xyz_min,
                       const void* xyz_max, const void* t_min, float stepdist, int64_t n_rays,
                       const int64_t* n_steps_cumsum, int64_t total, void* rays_pts,
                       uint8_t* mask_outbbox, int64_t* ray_id, int64_t* step_id, void* stream);

/* Scratch size (int64 elements) the inclusive scan of vx_sample_pts_count needs. */
int64_t vx_scan_workspace_elems(int64_t n);

/* Inclusive prefix sum of n int64 values (used by sample_pts_on_rays and segment
 * offsets); `ws` as for vx_sample_pts_count. */
int vx_inclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, int64_t* ws, void* stream);

/* lib/cuda/render_utils.cpp:87-98 sample_ndc_pts_on_rays (render_utils_kernel.cu:244-293).
 * rays_pts [n_rays,N_samples,3], mask_outbbox [n_rays,N_samples]. */
int vx_sample_ndc_pts_on_rays(int dtype, const void* rays_o, const void* rays_d,
                              const void* xyz_min, const void* xyz_max, int n_samples,
                              int64_t n_rays, void* rays_pts, uint8_t* mask_outbbox,
                              void* stream);

/* lib/cuda/render_utils.cpp:100-107 sample_bg_pts_on_rays (render_utils_kernel.cu:300-360). */
int vx_sample_bg_pts_on_rays(int dtype, const void* rays_o, const void* rays_d,
                             const void* t_max, float bg_preserve, int n_samples,
                             int64_t n_rays, void* rays_pts, void* stream);

/* lib/cuda/render_utils.cpp:109-118 maskcache_lookup (render_utils_kernel.cu:367-424).
 * world [sz_i,sz_j,sz_k] bool bytes; out [n_pts] bool bytes (fully written). */
int vx_maskcache_lookup(int dtype, const uint8_t* world, const void* xyz, const void* scale,
                        const void* shift, int sz_i, int sz_j, int sz_k, int64_t n_pts,
                        uint8_t* out, void* stream);

/* ---- render_utils: alpha and transmittance ------------------------------------- */

/* lib/cuda/render_utils.cpp:120-124 raw2alpha (render_utils_kernel.cu:431-481). */
int vx_raw2alpha(int dtype, const void* density, float shift, float interval, int64_t n_pts,
                 void* exp_d, void* alpha, void* stream);
/* lib/cuda/render_utils.cpp:125-129 raw2alpha_nonuni (render_utils_kernel.cu:446-504). */
int vx_raw2alpha_nonuni(int dtype, const void* density, float shift, const void* interval,
                        int64_t n_pts, void* exp_d, void* alpha, void* stream);
/* lib/cuda/render_utils.cpp:131-135 raw2alpha_backward (render_utils_kernel.cu:507-552). */
int vx_raw2alpha_backward(int dtype, const void* exp_d, const void* grad_back, float interval,
                          int64_t n_pts, void* grad, void* stream);
/* lib/cuda/render_utils.cpp:136-140 raw2alpha_nonuni_backward (render_utils_kernel.cu:520-574). */
int vx_raw2alpha_nonuni_backward(int dtype, const void* exp_d, const void* grad_back,
                                 const void* interval, int64_t n_pts, void* grad, void* stream);

/* lib/cuda/render_utils.cpp:142-149 alpha2weight (render_utils_kernel.cu:577-651).
 * Outputs weight [n_pts], T [n_pts], alphainv_last [n_rays], i_start/i_end [n_rays];
 * the function initialises all of them itself (weight=0, T=1, alphainv_last=1,
 * i_start=i_end=0 before the scan) and needs no host synchronisation. */
int vx_alpha2weight(int dtype, const void* alpha, const int64_t* ray_id, int64_t n_pts,
                    int64_t n_rays, void* weight, void* T, void* alphainv_last,
                    int64_t* i_start, int64_t* i_end, void* stream);

/* lib/cuda/render_utils.cpp:151-167 alpha2weight_backward (render_utils_kernel.cu:654-707).
 * grad [n_pts] is fully written (zero outside every ray's [i_start, i_end)). */
int vx_alpha2weight_backward(int dtype, const void* alpha, const void* weight, const void* T,
                             const void* alphainv_last, const int64_t* i_start,
                             const int64_t* i_end, int64_t n_pts, int64_t n_rays,
                             const void* grad_weights, const void* grad_last, void* grad,
                             void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VOXHIP_H */
