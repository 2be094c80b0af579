"""Drop-in for the reference's `render_utils_cuda` extension (lib/cuda/render_utils.cpp).

Same function names, positional signatures, return values and input checks as the
pybind module at lib/cuda/render_utils.cpp:170-184; every call launches hand-written
HIP kernels of libvoxhip.so on torch's current stream.  Outputs are allocated here on
the inputs' device (the reference allocates inside its C++ wrappers).

Differences, all stricter: dtype and shape mismatches raise instead of reading out of
bounds; ray ids outside [0, n_rays) are ignored by alpha2weight instead of writing out
of bounds; sizes are 64-bit.
"""
import torch

from ._lib import call, dtype_code, ptr, stream_of, lib

__all__ = [
    "infer_t_minmax", "infer_n_samples", "infer_ray_start_dir", "sample_pts_on_rays",
    "sample_ndc_pts_on_rays", "sample_bg_pts_on_rays", "maskcache_lookup", "raw2alpha",
    "raw2alpha_backward", "raw2alpha_nonuni", "raw2alpha_nonuni_backward", "alpha2weight",
    "alpha2weight_backward",
]


def _check_input(x, name):
    # CHECK_INPUT (render_utils.cpp:46-48)
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not x.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _same_float(ref, *others):
    code = dtype_code(ref)
    for name, t in others:
        if t.dtype != ref.dtype:
            raise RuntimeError(f"{name} has dtype {t.dtype}, expected {ref.dtype}")
    return code


def _check_index(t, name):
    if t.dtype != torch.int64:
        raise RuntimeError(f"{name} must be int64")


def infer_t_minmax(rays_o, rays_d, xyz_min, xyz_max, near, far):
    """render_utils.cpp:50-58 -> [t_min, t_max]"""
    for t, n in ((rays_o, "rays_o"), (rays_d, "rays_d"), (xyz_min, "xyz_min"), (xyz_max, "xyz_max")):
        _check_input(t, n)
    code = _same_float(rays_o, ("rays_d", rays_d), ("xyz_min", xyz_min), ("xyz_max", xyz_max))
    n = rays_o.size(0)
    t_min = torch.empty(n, dtype=rays_o.dtype, device=rays_o.device)
    t_max = torch.empty_like(t_min)
    call("vx_infer_t_minmax", code, ptr(rays_o), ptr(rays_d), ptr(xyz_min), ptr(xyz_max),
         float(near), float(far), n, ptr(t_min), ptr(t_max), stream_of(rays_o))
    return [t_min, t_max]


def infer_n_samples(rays_d, t_min, t_max, stepdist):
    """render_utils.cpp:60-65 -> N_steps (int64)"""
    for t, n in ((rays_d, "rays_d"), (t_min, "t_min"), (t_max, "t_max")):
        _check_input(t, n)
    code = _same_float(t_min, ("rays_d", rays_d), ("t_max", t_max))
    n = t_min.size(0)
    out = torch.empty(n, dtype=torch.int64, device=t_min.device)
    call("vx_infer_n_samples", code, ptr(rays_d), ptr(t_min), ptr(t_max), float(stepdist), n,
         ptr(out), stream_of(t_min))
    return out


def infer_ray_start_dir(rays_o, rays_d, t_min):
    """render_utils.cpp:67-72 -> [rays_start, rays_dir]"""
    for t, n in ((rays_o, "rays_o"), (rays_d, "rays_d"), (t_min, "t_min")):
        _check_input(t, n)
    code = _same_float(rays_o, ("rays_d", rays_d), ("t_min", t_min))
    n = rays_o.size(0)
    start = torch.empty_like(rays_o)
    dirs = torch.empty_like(rays_o)
    call("vx_infer_ray_start_dir", code, ptr(rays_o), ptr(rays_d), ptr(t_min), n, ptr(start),
         ptr(dirs), stream_of(rays_o))
    return [start, dirs]


def _sample_count(rays_o, rays_d, xyz_min, xyz_max, near, far, stepdist):
    code = dtype_code(rays_o)
    n = rays_o.size(0)
    dev = rays_o.device
    t_min = torch.empty(n, dtype=rays_o.dtype, device=dev)
    t_max = torch.empty_like(t_min)
    n_steps = torch.empty(n, dtype=torch.int64, device=dev)
    csum = torch.empty(n, dtype=torch.int64, device=dev)
    total = torch.empty(1, dtype=torch.int64, device=dev)
    ws = torch.empty(max(1, lib().vx_scan_workspace_elems(n)), dtype=torch.int64, device=dev)
    call("vx_sample_pts_count", code, ptr(rays_o), ptr(rays_d), ptr(xyz_min), ptr(xyz_max),
         float(near), float(far), float(stepdist), n, ptr(t_min), ptr(t_max), ptr(n_steps),
         ptr(csum), ptr(total), ptr(ws), stream_of(rays_o))
    return t_min, t_max, n_steps, csum, total


def sample_pts_on_rays(rays_o, rays_d, xyz_min, xyz_max, near, far, stepdist):
    """render_utils.cpp:74-85 ->
    [rays_pts, mask_outbbox, ray_id, step_id, N_steps, t_min, t_max]

    One host read of the sample total, as the reference's `.item()`
    (render_utils_kernel.cu:212); everything else stays on the device.
    """
    for t, n in ((rays_o, "rays_o"), (rays_d, "rays_d"), (xyz_min, "xyz_min"), (xyz_max, "xyz_max")):
        _check_input(t, n)
    if rays_o.dim() != 2 or rays_o.size(1) != 3:
        raise RuntimeError("rays_o must have shape [N, 3]")
    if rays_d.shape != rays_o.shape:
        raise RuntimeError("rays_d must have the shape of rays_o")
    code = _same_float(rays_o, ("rays_d", rays_d), ("xyz_min", xyz_min), ("xyz_max", xyz_max))
    t_min, t_max, n_steps, csum, total = _sample_count(rays_o, rays_d, xyz_min, xyz_max, near,
                                                       far, stepdist)
    m = int(total.item())
    dev = rays_o.device
    pts = torch.empty((m, 3), dtype=rays_o.dtype, device=dev)
    mask = torch.empty(m, dtype=torch.bool, device=dev)
    ray_id = torch.empty(m, dtype=torch.int64, device=dev)
    step_id = torch.empty(m, dtype=torch.int64, device=dev)
    call("vx_sample_pts_fill", code, ptr(rays_o), ptr(rays_d), ptr(xyz_min), ptr(xyz_max),
         ptr(t_min), float(stepdist), rays_o.size(0), ptr(csum), m, ptr(pts), ptr(mask),
         ptr(ray_id), ptr(step_id), stream_of(rays_o))
    return [pts, mask, ray_id, step_id, n_steps, t_min, t_max]


def sample_ndc_pts_on_rays(rays_o, rays_d, xyz_min, xyz_max, N_samples):
    """render_utils.cpp:87-98 -> [rays_pts [N,S,3], mask_outbbox [N,S]]"""
    for t, n in ((rays_o, "rays_o"), (rays_d, "rays_d"), (xyz_min, "xyz_min"), (xyz_max, "xyz_max")):
        _check_input(t, n)
    code = _same_float(rays_o, ("rays_d", rays_d), ("xyz_min", xyz_min), ("xyz_max", xyz_max))
    n = rays_o.size(0)
    S = int(N_samples)
    pts = torch.empty((n, S, 3), dtype=rays_o.dtype, device=rays_o.device)
    mask = torch.empty((n, S), dtype=torch.bool, device=rays_o.device)
    call("vx_sample_ndc_pts_on_rays", code, ptr(rays_o), ptr(rays_d), ptr(xyz_min), ptr(xyz_max),
         S, n, ptr(pts), ptr(mask), stream_of(rays_o))
    return [pts, mask]


def sample_bg_pts_on_rays(rays_o, rays_d, t_max, bg_preserve, N_samples):
    """render_utils.cpp:100-107 -> rays_pts [N,S,3]"""
    for t, n in ((rays_o, "rays_o"), (rays_d, "rays_d"), (t_max, "t_max")):
        _check_input(t, n)
    code = _same_float(rays_o, ("rays_d", rays_d), ("t_max", t_max))
    n = rays_o.size(0)
    S = int(N_samples)
    pts = torch.empty((n, S, 3), dtype=rays_o.dtype, device=rays_o.device)
    call("vx_sample_bg_pts_on_rays", code, ptr(rays_o), ptr(rays_d), ptr(t_max),
         float(bg_preserve), S, n, ptr(pts), stream_of(rays_o))
    return pts


def maskcache_lookup(world, xyz, xyz2ijk_scale, xyz2ijk_shift):
    """render_utils.cpp:109-118 -> bool [n_pts]"""
    for t, n in ((world, "world"), (xyz, "xyz"), (xyz2ijk_scale, "xyz2ijk_scale"),
                 (xyz2ijk_shift, "xyz2ijk_shift")):
        _check_input(t, n)
    if world.dim() != 3 or xyz.dim() != 2 or xyz.size(1) != 3:
        raise RuntimeError("maskcache_lookup: world must be 3-D and xyz [N, 3]")
    if world.dtype != torch.bool:
        raise RuntimeError("world must be a bool tensor")
    code = _same_float(xyz, ("xyz2ijk_scale", xyz2ijk_scale), ("xyz2ijk_shift", xyz2ijk_shift))
    n = xyz.size(0)
    out = torch.empty(n, dtype=torch.bool, device=xyz.device)
    call("vx_maskcache_lookup", code, ptr(world), ptr(xyz), ptr(xyz2ijk_scale),
         ptr(xyz2ijk_shift), world.size(0), world.size(1), world.size(2), n, ptr(out),
         stream_of(xyz))
    return out


def raw2alpha(density, shift, interval):
    """render_utils.cpp:120-124 -> [exp_d, alpha]"""
    _check_input(density, "density")
    if density.dim() != 1:
        raise RuntimeError("density must be 1-D")
    code = dtype_code(density)
    exp_d = torch.empty_like(density)
    alpha = torch.empty_like(density)
    call("vx_raw2alpha", code, ptr(density), float(shift), float(interval), density.numel(),
         ptr(exp_d), ptr(alpha), stream_of(density))
    return [exp_d, alpha]


def raw2alpha_nonuni(density, shift, interval):
    """render_utils.cpp:125-129 -> [exp_d, alpha] with a per-point interval."""
    _check_input(density, "density")
    _check_input(interval, "interval")
    if density.dim() != 1 or interval.numel() != density.numel():
        raise RuntimeError("density must be 1-D and interval of the same length")
    code = _same_float(density, ("interval", interval))
    exp_d = torch.empty_like(density)
    alpha = torch.empty_like(density)
    call("vx_raw2alpha_nonuni", code, ptr(density), float(shift), ptr(interval),
         density.numel(), ptr(exp_d), ptr(alpha), stream_of(density))
    return [exp_d, alpha]


def raw2alpha_backward(exp, grad_back, interval):
    """render_utils.cpp:131-135 -> grad"""
    _check_input(exp, "exp")
    _check_input(grad_back, "grad_back")
    if grad_back.numel() != exp.numel():
        raise RuntimeError("grad_back must have the length of exp")
    code = _same_float(exp, ("grad_back", grad_back))
    grad = torch.empty_like(exp)
    call("vx_raw2alpha_backward", code, ptr(exp), ptr(grad_back), float(interval), exp.numel(),
         ptr(grad), stream_of(exp))
    return grad


def raw2alpha_nonuni_backward(exp, grad_back, interval):
    """render_utils.cpp:136-140 -> grad"""
    _check_input(exp, "exp")
    _check_input(grad_back, "grad_back")
    _check_input(interval, "interval")
    if grad_back.numel() != exp.numel() or interval.numel() != exp.numel():
        raise RuntimeError("grad_back and interval must have the length of exp")
    code = _same_float(exp, ("grad_back", grad_back), ("interval", interval))
    grad = torch.empty_like(exp)
    call("vx_raw2alpha_nonuni_backward", code, ptr(exp), ptr(grad_back), ptr(interval),
         exp.numel(), ptr(grad), stream_of(exp))
    return grad


def alpha2weight(alpha, ray_id, n_rays):
    """render_utils.cpp:142-149 -> [weight, T, alphainv_last, i_start, i_end]

    No host synchronisation (the reference's `i_end[ray_id[n_pts-1]] = n_pts` is done
    on the device).
    """
    _check_input(alpha, "alpha")
    _check_input(ray_id, "ray_id")
    if alpha.dim() != 1 or ray_id.dim() != 1 or alpha.shape != ray_id.shape:
        raise RuntimeError("alpha and ray_id must be 1-D of equal length")
    _check_index(ray_id, "ray_id")
    code = dtype_code(alpha)
    n_rays = int(n_rays)
    n_pts = alpha.numel()
    dev = alpha.device
    weight = torch.empty_like(alpha)
    T = torch.empty_like(alpha)
    last = torch.empty(n_rays, dtype=alpha.dtype, device=dev)
    i_start = torch.empty(n_rays, dtype=torch.int64, device=dev)
    i_end = torch.empty(n_rays, dtype=torch.int64, device=dev)
    call("vx_alpha2weight", code, ptr(alpha), ptr(ray_id), n_pts, n_rays, ptr(weight), ptr(T),
         ptr(last), ptr(i_start), ptr(i_end), stream_of(alpha))
    return [weight, T, last, i_start, i_end]


def alpha2weight_backward(alpha, weight, T, alphainv_last, i_start, i_end, n_rays,
                          grad_weights, grad_last):
    """render_utils.cpp:151-167 -> grad"""
    for t, n in ((alpha, "alpha"), (weight, "weight"), (T, "T"), (alphainv_last, "alphainv_last"),
                 (i_start, "i_start"), (i_end, "i_end"), (grad_weights, "grad_weights"),
                 (grad_last, "grad_last")):
        _check_input(t, n)
    code = _same_float(alpha, ("weight", weight), ("T", T), ("alphainv_last", alphainv_last),
                       ("grad_weights", grad_weights), ("grad_last", grad_last))
    _check_index(i_start, "i_start")
    _check_index(i_end, "i_end")
    n_rays = int(n_rays)
    n_pts = alpha.numel()
    if not (weight.numel() == T.numel() == grad_weights.numel() == n_pts):
        raise RuntimeError("weight, T and grad_weights must have the length of alpha")
    if not (alphainv_last.numel() >= n_rays and grad_last.numel() >= n_rays
            and i_start.numel() >= n_rays and i_end.numel() >= n_rays):
        raise RuntimeError("per-ray inputs shorter than n_rays")
    grad = torch.empty_like(alpha)
    call("vx_alpha2weight_backward", code, ptr(alpha), ptr(weight), ptr(T), ptr(alphainv_last),
         ptr(i_start), ptr(i_end), n_pts, n_rays, ptr(grad_weights), ptr(grad_last), ptr(grad),
         stream_of(alpha))
    return grad
