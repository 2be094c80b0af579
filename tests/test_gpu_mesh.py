"""Mesh extraction on the GPU (csrc/mcubes.hip) against the oracle's restatement of
lib/dvgo_ori.py:679-703 and lib/voxurf_fine.py:894-910.

The lattice query equals the drop-in grid_sample at torch.linspace points bit for bit
and the CPU restatement within 1e-5.  Marching cubes is integer work plus one float64
division per vertex: the vertex set equals the brute-force crossing-edge points and the
triangles equal the CPU assembly with the same table, both exactly; world-space vertices
follow the reference's scaling exactly.  PyMCubes (what the reference calls) is not
installed, so its own triangle choice is not compared (parity unpinned there); the
table is pinned by mesh invariants (tests/test_mc_tables.py and the closed, genus-0,
outward-facing sphere below)."""
import numpy as np
import pytest
import torch

from oracle import torch_ref as tr
from voxurf_amd import functional, mesh, scene

pytestmark = pytest.mark.gpu


def _canon(tris):
    """Triangles [F, 3, 3] rotated to start at their smallest vertex (orientation kept),
    then sorted."""
    out = []
    for t in tris:
        keys = [tuple(v) for v in t]
        r = keys.index(min(keys))
        out.append(tuple(keys[r:] + keys[:r]))
    return sorted(out)


def _closed(faces):
    directed = {}
    for f in faces.tolist():
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            directed[(a, b)] = directed.get((a, b), 0) + 1
    return all(c == 1 for c in directed.values()) and all((b, a) in directed for a, b in directed), \
        len(directed) // 2


def test_extract_fields(dev):
    R = 40
    sdf = scene.sdf_grid(R)
    mn, mx = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
    bmin, bmax = [-0.9, -1.0, -0.8], [0.95, 0.7, 1.0]
    u = mesh.extract_fields(sdf.to(dev), mn, mx, bmin, bmax, 37)
    ref = tr.extract_fields(sdf, mn, mx, bmin, bmax, 37)
    assert float((u.cpu() - ref).abs().max()) <= 1e-5 * float(ref.abs().max())
    # the same lookup as the drop-in grid_sample at torch.linspace's points, bit for bit
    axes = [torch.linspace(a, b, 37) for a, b in zip(bmin, bmax)]
    pts = torch.stack(torch.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    g = functional.grid_sample(-sdf.to(dev), pts.to(dev), torch.tensor(mn, device=dev),
                               torch.tensor(mx, device=dev))
    assert torch.equal(u.reshape(-1), g.reshape(-1))


def test_marching_cubes_matches_cpu_assembly(dev):
    g = torch.Generator().manual_seed(0)
    u = torch.rand((14, 11, 17), generator=g) * 2 - 1
    u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 1, 1, 1, 1, 1, 1
    verts, faces = mesh.marching_cubes(u.to(dev), 0.0)
    v, f = verts.cpu().numpy(), faces.cpu().numpy()
    assert np.array_equal(v[np.lexsort((v[:, 2], v[:, 1], v[:, 0]))], tr.mc_vertices(u.numpy(), 0.0))
    count, edges = mesh.mc_tables_host()
    ref = tr.mc_assemble(u.numpy(), 0.0, count, edges)
    assert _canon(v[f]) == _canon(ref)
    ok, _ = _closed(f)
    assert ok
    # threshold away from zero and world-space vertices
    verts2, faces2 = mesh.marching_cubes(u.to(dev), 0.3, [-1.0, -2.0, 0.5], [1.0, 2.0, 1.5])
    vi, _ = mesh.marching_cubes(u.to(dev), 0.3)
    n = np.array(u.shape, dtype=np.float64)
    ext = (np.float32([1.0, 2.0, 1.5]) - np.float32([-1.0, -2.0, 0.5])).astype(np.float64)
    want = vi.cpu().numpy() / (n - 1.0) * ext + np.float32([-1.0, -2.0, 0.5]).astype(np.float64)
    assert np.array_equal(verts2.cpu().numpy(), want)
    # no crossing: empty mesh
    ve, fe = mesh.marching_cubes(torch.ones((5, 6, 7), device=dev), 0.0)
    assert ve.shape == (0, 3) and fe.shape == (0, 3)


def test_extract_geometry_sphere(dev):
    R = 64
    sdf = scene.sdf_grid(R)
    verts, faces = mesh.extract_geometry(sdf.to(dev), [-1, -1, -1], [1, 1, 1],
                                         torch.tensor([-1.0, -1.0, -1.0]),
                                         torch.tensor([1.0, 1.0, 1.0]), resolution=96)
    ok, E = _closed(faces)
    assert ok and len(verts) - E + len(faces) == 2
    p = verts[faces]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert ((n * p.mean(1)).sum(1) > 0).mean() > 0.999      # outward
    area = 0.5 * np.linalg.norm(n, axis=1).sum()
    assert abs(area - 4 * np.pi * 0.25) < 0.05 * 4 * np.pi * 0.25   # rippled sphere, r 0.5
