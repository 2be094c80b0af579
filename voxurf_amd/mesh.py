"""Mesh extraction on the HIP kernels (csrc/mcubes.hip): the lattice SDF query and
marching cubes of Voxurf.extract_geometry (lib/voxurf_fine.py:894-910,
lib/dvgo_ori.py:679-703).

extract_fields(...)   -> u [n, n, n] on the device (float32), the field the reference
                         builds chunk by chunk on the host
marching_cubes(u, t)  -> (vertices [V, 3] float64, triangles [F, 3] int64) on the device,
                         vertices in lattice index space as mcubes.marching_cubes returns
extract_geometry(...) -> numpy (vertices in world space, triangles), as the reference

Triangles face the side where the field is below the threshold; for the field -sdf at
threshold 0 (what the reference meshes) that is outward.
"""
import numpy as np
import torch

from ._lib import call, lib, ptr, stream_of
from .total_variation_cuda import grid_geom

_TABLES = {}


def mc_tables_host():
    """(count [256] int8, edges [256, 3 * max_tris] int8) of the generated table."""
    m = lib().vx_mc_max_tris()
    count = np.zeros(256, np.int8)
    edges = np.zeros((256, 3 * m), np.int8)
    call("vx_mc_tables", count.ctypes.data, edges.ctypes.data)
    return count, edges


def _tables(dev):
    key = str(dev)
    if key not in _TABLES:
        count, edges = mc_tables_host()
        _TABLES[key] = (torch.from_numpy(count).to(dev), torch.from_numpy(edges).to(dev))
    return _TABLES[key]


def extract_fields(sdf_grid, xyz_min, xyz_max, bound_min, bound_max, resolution, scale=-1.0):
    """u[i, j, k] = scale * trilinear(sdf_grid) at (linspace(bmin, bmax, res)[i], ...)
    (lib/dvgo_ori.py:679-693 with query_func = grid_sampler(pts, -sdf_grid))."""
    s = sdf_grid.detach()
    if not s.is_cuda or s.dtype != torch.float32 or s.dim() != 5 or s.shape[1] != 1:
        raise RuntimeError("extract_fields: float32 [1, 1, X, Y, Z] grid on the GPU expected")
    res = [int(resolution)] * 3 if np.isscalar(resolution) else [int(r) for r in resolution]
    geom = grid_geom(s)
    box = torch.tensor([*torch.as_tensor(xyz_min).tolist(), *torch.as_tensor(xyz_max).tolist()],
                       dtype=torch.float32)
    bounds = torch.tensor([*torch.as_tensor(bound_min).tolist(),
                           *torch.as_tensor(bound_max).tolist()], dtype=torch.float32)
    u = torch.empty(res, dtype=torch.float32, device=s.device)
    call("vx_extract_fields", ptr(s), geom.data_ptr(), box.data_ptr(), bounds.data_ptr(), *res,
         float(scale), ptr(u), stream_of(s))
    return u


def marching_cubes(u, threshold=0.0, bound_min=None, bound_max=None):
    """Marching cubes of the field u [nx, ny, nz] (float32, on the GPU) at `threshold`
    (mcubes.marching_cubes(u, threshold), lib/dvgo_ori.py:698).  With bound_min and
    bound_max, vertices are mapped to world space as lib/dvgo_ori.py:699-702 does."""
    if not u.is_cuda or u.dtype != torch.float32 or u.dim() != 3:
        raise RuntimeError("marching_cubes: float32 [nx, ny, nz] field on the GPU expected")
    u = u.contiguous()
    nx, ny, nz = u.shape
    dev = u.device
    st = stream_of(u)
    count, edges = _tables(dev)
    n = nx * ny * nz
    vcnt = torch.empty(n, dtype=torch.int64, device=dev)
    tcnt = torch.empty(n, dtype=torch.int64, device=dev)
    call("vx_mc_count", ptr(u), nx, ny, nz, float(threshold), ptr(count), ptr(edges), ptr(vcnt),
         ptr(tcnt), st)
    ws = torch.empty(max(1, lib().vx_scan_workspace_elems(n)), dtype=torch.int64, device=dev)
    vincl = torch.empty_like(vcnt)
    tincl = torch.empty_like(tcnt)
    call("vx_inclusive_scan_i64", ptr(vcnt), ptr(vincl), n, ptr(ws), st)
    call("vx_inclusive_scan_i64", ptr(tcnt), ptr(tincl), n, ptr(ws), st)
    totals = torch.stack([vincl[-1], tincl[-1]]).tolist()
    nv, nf = int(totals[0]), int(totals[1])
    verts = torch.empty((max(nv, 1), 3), dtype=torch.float64, device=dev)
    faces = torch.empty((max(nf, 1), 3), dtype=torch.int64, device=dev)
    bounds = None
    if bound_min is not None:
        bounds = np.asarray([*torch.as_tensor(bound_min).tolist(),
                             *torch.as_tensor(bound_max).tolist()], dtype=np.float64)
    call("vx_mc_emit", ptr(u), nx, ny, nz, float(threshold), ptr(count), ptr(edges), ptr(vincl),
         ptr(tincl), None if bounds is None else bounds.ctypes.data, ptr(verts), ptr(faces), st)
    return verts[:nv], faces[:nf]


@torch.no_grad()
def extract_geometry(sdf_grid, xyz_min, xyz_max, bound_min, bound_max, resolution=128,
                     threshold=0.0, smooth=True, sigma=0.5, smooth_conv=None):
    """Voxurf.extract_geometry (lib/voxurf_fine.py:894-910): optional smoothing (the
    model's smooth_conv, else a 3^3 Gaussian of `sigma` when smooth), the -sdf field on
    a resolution^3 lattice over [bound_min, bound_max], marching cubes, world-space
    vertices.  Returns numpy (vertices [V, 3] float64, triangles [F, 3] int64)."""
    from .smooth import GaussianConv3d
    g = sdf_grid.detach()
    if smooth_conv is not None:
        g = smooth_conv(g)
    elif smooth:
        g = GaussianConv3d(3, sigma)(g)
    u = extract_fields(g.contiguous(), xyz_min, xyz_max, bound_min, bound_max, resolution)
    verts, faces = marching_cubes(u, threshold, bound_min, bound_max)
    return verts.cpu().numpy(), faces.cpu().numpy()
