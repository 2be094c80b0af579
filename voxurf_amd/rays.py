"""Camera rays and the MaskCache ray pre-filter of the fine stage on the HIP kernels
(csrc/rays.hip).

Mirrors lib/voxurf_fine.py: get_rays (:1001-1029), ndc_rays (:1042-1059),
get_rays_of_a_view (:1062-1068), Voxurf.hit_coarse_geo (:579-591) and
get_training_rays_in_maskcache_sampling (:1127-1164), with the same arguments and
returns.  Rays are made on the GPU (c2w's device when that is a GPU, else the current
one); K and c2w may be numpy arrays or tensors and are read as float32, as the
reference's float32 arithmetic reads them.
"""
import numpy as np
import torch

from ._lib import call, lib, ptr, stream_of
from .fused import maskcache_args

_MODES = {"center": 0.5, "lefttop": 0.0}


def _device(c2w):
    if torch.is_tensor(c2w) and c2w.is_cuda:
        return c2w.device
    if not torch.cuda.is_available():
        raise RuntimeError("voxurf_amd.rays needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _ndc_coeffs(H, W, focal):
    focal = float(focal)
    c0 = np.float32(-1.0 / (W / (2.0 * focal)))
    c1 = np.float32(-1.0 / (H / (2.0 * focal)))
    return float(c0), float(c1)


def camera_params(H, W, K, c2w, inverse_y, flip_x, flip_y, mode="center", ndc=False):
    """The 26 host floats a camera is passed as (see vx_camera_rays)."""
    if mode not in _MODES:
        raise NotImplementedError(f"ray mode {mode!r} (center and lefttop are supported)")
    K = np.asarray(K.detach().cpu() if torch.is_tensor(K) else K, dtype=np.float64)
    M = np.asarray(c2w.detach().cpu() if torch.is_tensor(c2w) else c2w, dtype=np.float32)
    c0, c1 = _ndc_coeffs(H, W, K[0][0]) if ndc else (0.0, 0.0)
    vals = [H, W, K[0][0], K[1][1], K[0][2], K[1][2], *M[:3, :3].reshape(-1).tolist(),
            *M[:3, 3].tolist(), _MODES[mode], float(bool(inverse_y)), float(bool(flip_x)),
            float(bool(flip_y)), float(bool(ndc)), c0, c1, 1.0]
    return torch.tensor(np.asarray(vals, dtype=np.float64).astype(np.float32))


def _camera_rays(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y, mode):
    dev = _device(c2w)
    cam = camera_params(H, W, K, c2w, inverse_y, flip_x, flip_y, mode, ndc)
    o = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    d = torch.empty_like(o)
    v = torch.empty_like(o)
    call("vx_camera_rays", cam.data_ptr(), ptr(o), ptr(d), ptr(v), stream_of(o))
    return o, d, v


def get_rays(H, W, K, c2w, inverse_y, flip_x, flip_y, mode="center"):
    """rays_o, rays_d [H, W, 3] of a pinhole view (lib/voxurf_fine.py:1001-1029)."""
    o, d, _ = _camera_rays(H, W, K, c2w, False, inverse_y, flip_x, flip_y, mode)
    return o, d


def ndc_rays(H, W, focal, near, rays_o, rays_d):
    """lib/voxurf_fine.py:1042-1059 on CUDA tensors [..., 3]."""
    if not (rays_o.is_cuda and rays_d.is_cuda):
        raise RuntimeError("ndc_rays: CUDA (HIP) tensors expected")
    if rays_o.dtype != torch.float32 or rays_d.dtype != torch.float32 or \
            rays_o.shape != rays_d.shape:
        raise RuntimeError("ndc_rays: rays_o and rays_d must be float32 of one shape")
    o_in, d_in = rays_o.contiguous(), rays_d.contiguous()
    c0, c1 = _ndc_coeffs(H, W, focal)
    o, d = torch.empty_like(o_in), torch.empty_like(d_in)
    call("vx_ndc_rays", o_in.numel() // 3, c0, c1, float(near), ptr(o_in), ptr(d_in), ptr(o),
         ptr(d), stream_of(o))
    return o, d


def get_rays_of_a_view(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y, mode="center"):
    """rays_o, rays_d, viewdirs [H, W, 3] (lib/voxurf_fine.py:1062-1068; NDC with near 1)."""
    return _camera_rays(H, W, K, c2w, ndc, inverse_y, flip_x, flip_y, mode)


def _prefilter_args(mask_cache, xyz_min, xyz_max, near, stepdist):
    if mask_cache is None:
        raise RuntimeError("hit_coarse_geo needs a MaskCache")
    if mask_cache.density.dtype != torch.float32 or not mask_cache.density.is_cuda:
        raise RuntimeError("hit_coarse_geo: the MaskCache density must be float32 on the GPU")
    box = torch.tensor([*torch.as_tensor(xyz_min).tolist(), *torch.as_tensor(xyz_max).tolist()],
                       dtype=torch.float32)
    params = torch.tensor([stepdist, near], dtype=torch.float32)
    return box, params, maskcache_args(mask_cache)


def hit_coarse_geo(mask_cache, xyz_min, xyz_max, rays_o, rays_d, near, stepdist):
    """Whether each ray hits the coarse geometry (lib/voxurf_fine.py:579-591): some
    in-box sample (sample_pts_on_rays with far = 1e9) passes the MaskCache.  bool [...]."""
    shape = rays_o.shape[:-1]
    o = rays_o.reshape(-1, 3).contiguous()
    d = rays_d.reshape(-1, 3).contiguous()
    if not (o.is_cuda and d.is_cuda) or o.dtype != torch.float32 or d.dtype != torch.float32:
        raise RuntimeError("hit_coarse_geo: float32 CUDA (HIP) rays expected")
    if o.shape != d.shape:
        raise RuntimeError("hit_coarse_geo: rays_o and rays_d differ in shape")
    box, params, mc = _prefilter_args(mask_cache, xyz_min, xyz_max, near, stepdist)
    hit = torch.empty(o.shape[0], dtype=torch.bool, device=o.device)
    call("vx_hit_coarse_geo", ptr(o), ptr(d), o.shape[0], box.data_ptr(), params.data_ptr(),
         *mc, ptr(hit), None, stream_of(o))
    return hit.reshape(shape)


@torch.no_grad()
def get_training_rays_in_maskcache_sampling(rgb_tr_ori, train_poses, HW, Ks, ndc, inverse_y,
                                            flip_x, flip_y, model, render_kwargs,
                                            rgbnet_sup_reduce=1):
    """The training rays of every view that hit the coarse geometry, in pixel order
    (lib/voxurf_fine.py:1127-1164).  `model` provides mask_cache, xyz_min, xyz_max and
    voxel_size; render_kwargs provides near and stepsize.  Returns rgb_tr, rays_o_tr,
    rays_d_tr, viewdirs_tr [N_kept, 3] and imsz (kept rays per view)."""
    assert len(rgb_tr_ori) == len(train_poses) and len(rgb_tr_ori) == len(Ks) and \
        len(rgb_tr_ori) == len(HW)
    dev = rgb_tr_ori[0].device
    if not dev.type == "cuda":
        raise RuntimeError("get_training_rays_in_maskcache_sampling: images must be on the GPU")
    stepdist = render_kwargs["stepsize"] * model.voxel_size
    box, params, mc = _prefilter_args(model.mask_cache, model.xyz_min, model.xyz_max,
                                      render_kwargs["near"], stepdist)
    N = sum(im.shape[0] * im.shape[1] for im in rgb_tr_ori)
    rgb_tr = torch.zeros([N, 3], device=dev)
    rays_o_tr = torch.zeros_like(rgb_tr)
    rays_d_tr = torch.zeros_like(rgb_tr)
    viewdirs_tr = torch.zeros_like(rgb_tr)
    imsz = []
    top = 0
    st = stream_of(rgb_tr)
    for c2w, img, (H, W), K in zip(train_poses, rgb_tr_ori, HW, Ks):
        assert img.shape[:2] == (H, W)
        cam = camera_params(H, W, K, c2w, inverse_y, flip_x, flip_y, "center", ndc)
        n = H * W
        flag = torch.empty(n, dtype=torch.int64, device=dev)
        call("vx_view_hit_coarse_geo", cam.data_ptr(), box.data_ptr(), params.data_ptr(), *mc,
             ptr(flag), st)
        incl = torch.empty_like(flag)
        ws = torch.empty(max(1, lib().vx_scan_workspace_elems(n)), dtype=torch.int64, device=dev)
        call("vx_inclusive_scan_i64", ptr(flag), ptr(incl), n, ptr(ws), st)
        kept = int(incl[-1]) if n else 0
        rgb = img.reshape(-1, 3).to(device=dev, dtype=torch.float32).contiguous()
        call("vx_view_compact_rays", cam.data_ptr(), ptr(rgb), ptr(incl), ptr(rgb_tr[top:]),
             ptr(rays_o_tr[top:]), ptr(rays_d_tr[top:]), ptr(viewdirs_tr[top:]), st)
        imsz.append(kept)
        top += kept
    return rgb_tr[:top], rays_o_tr[:top], rays_d_tr[:top], viewdirs_tr[:top], imsz
