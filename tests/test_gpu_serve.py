"""Serving loop (GPU): continuous batching over resident scenes.

  * every request's result in a concurrent mix (several scenes, renders of uneven ray
    counts split across and packed into batches, SDF queries, 3 worker streams) equals
    the same request rendered alone, bit for bit (rendering has no cross-ray
    arithmetic);
  * admission under an HBM budget smaller than the requests' outputs still completes
    every request (one admitted at a time);
  * the fused render equals the op-by-op model's forward within 1e-5 relative, and an
    SDF query equals the trilinear SDF / 6-neighbour normal of the drop-in ops.
"""
import dataclasses

import pytest
import torch

from voxurf_amd import scene
from voxurf_amd.model import VoxurfFine
from voxurf_amd.serve import Request, ResidentScene, Server, Workspace, camera_rays, sequential

pytestmark = pytest.mark.gpu

CFG = dataclasses.replace(scene.FINE, R=96)


@pytest.fixture(scope="module")
def scenes(dev):
    out = []
    for s in range(2):
        out.append(ResidentScene(CFG, scene.sdf_grid(CFG.R), scene.feature_grid(CFG.R, 12, seed=s),
                                 scene.mlp_params(CFG.mlp_in, 192, 4, seed=s), device=dev))
    return out


def make_requests(dev, sizes, n_sdf=2):
    reqs = []
    for i, n in enumerate(sizes):
        reqs.append(Request(rid=i, kind="render", scene=i % 2, rays=camera_rays(n, seed=i, device=dev)))
    for j in range(n_sdf):
        reqs.append(Request(rid=len(sizes) + j, kind="sdf", scene=j % 2, tile=(512, (64 * j, 128, 192))))
    return reqs


def alone(scenes, req, dev):
    ws = Workspace(dev, scenes[0].cap)
    sc = scenes[req.scene]
    if req.kind == "sdf":
        return sc.sdf_query(*req.tile)
    return sc.render(*req.rays, ws)


def test_mix_equals_alone(dev, scenes):
    sizes = [1000, 70000, 5, 33333, 200000, 4096, 1]
    reqs = make_requests(dev, sizes)
    srv = Server(scenes, n_workers=3, batch_rays=65536, hbm_budget=1 << 30)
    for r in reqs:
        srv.submit(r)
    srv.drain(len(reqs), timeout=60)
    srv.close()
    for r in reqs:
        ref = alone(scenes, r, dev)
        if r.kind == "sdf":
            assert torch.equal(r.out[0], ref[0]) and torch.equal(r.out[1], ref[1])
        else:
            assert torch.equal(r.out, ref), f"request {r.rid} differs in the mix"


def test_budget_admission(dev, scenes):
    reqs = make_requests(dev, [50000, 60000, 70000], n_sdf=1)
    srv = Server(scenes, n_workers=2, batch_rays=32768, hbm_budget=1)  # admits one at a time
    for r in reqs:
        srv.submit(r)
    srv.drain(len(reqs), timeout=60)
    srv.close()
    for r in reqs:
        assert r.done_t >= r.admitted >= r.arrival
        ref = alone(scenes, r, dev)
        if r.kind == "render":
            assert torch.equal(r.out, ref)
    seq = sequential(scenes, make_requests(dev, [50000, 60000, 70000], n_sdf=1), chunk=8192)
    for a, b in zip(seq, reqs):
        if a.kind == "render":
            assert torch.equal(a.out, b.out)


def test_render_matches_op_path(dev, scenes):
    sc = scenes[0]
    o, d, v = camera_rays(5000, seed=3, device=dev)
    ws = Workspace(dev, sc.cap)
    rgb = sc.render(o, d, v, ws)
    model = VoxurfFine(CFG, scene.sdf_grid(CFG.R), scene.feature_grid(CFG.R, 12, seed=0),
                       scene.mlp_params(CFG.mlp_in, 192, 4, seed=0)).to(dev)
    with torch.no_grad():
        ref = model(o, d, v)["rgb_marched"]
    torch.testing.assert_close(rgb, ref, rtol=1e-5, atol=1e-6)
    # SDF query: the lattice points' SDF and normal from the drop-in ops
    from voxurf_amd.functional import SdfGradSample
    sdf, nrm = sc.sdf_query(512, (0, 64, 128))
    ax = [(-1.0 + 2.0 * (torch.arange(64, device=dev, dtype=torch.float32) + o_) / 511) for o_ in (0, 64, 128)]
    X, Y, Z = torch.meshgrid(*ax, indexing="ij")
    pts = torch.stack([X, Y, Z], -1).reshape(-1, 3)
    s2, g2, _ = SdfGradSample.apply(sc.sdf, pts, sc.xyz_min, sc.xyz_max, CFG.voxel_size)
    assert torch.equal(sdf, s2)
    torch.testing.assert_close(nrm, g2 / (g2.norm(dim=-1, keepdim=True) + 1e-5))
