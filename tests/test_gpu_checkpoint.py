"""Checkpoints on the GPU: a trained fused trainer saved in the reference's format
(run.py:786-794) and loaded back (lib/utils.py:234-300) holds bit-identical grids,
rgbnet, MaskCache and Adam state, so its next forward is bit-identical too; the same
file loads into the server, whose renders equal those of a scene built from the
trainer's tensors; a narrower rgbnet.0.weight is zero-padded as lib/utils.py:237-241
does."""
import dataclasses

import pytest
import torch

from voxurf_amd import checkpoint as ck
from voxurf_amd import scene
from voxurf_amd.fused import FusedFineTrainer
from voxurf_amd.maskcache import MaskCache
from voxurf_amd.serve import ResidentScene, Workspace

pytestmark = pytest.mark.gpu

CFG = dataclasses.replace(scene.FINE, R=48, n_rays=512, seed=2)


def _trained(dev):
    t = FusedFineTrainer(CFG, scene.sdf_grid(CFG.R), scene.feature_grid(CFG.R, CFG.feat_dim),
                         scene.mlp_params(CFG.mlp_in, CFG.mlp_width, CFG.mlp_depth), device=dev)
    x = torch.linspace(-1, 1, 24)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    dens = torch.where((X * X + Y * Y + Z * Z).sqrt() < 0.7, 5.0, -20.0)[None, None]
    t.mask_cache = MaskCache(density=dens, xyz_min=[-1, -1, -1], xyz_max=[1, 1, 1],
                             act_shift=-4.6, voxel_size_ratio=1.0).to(dev)
    for s in range(3):
        o, d, v = (a.to(dev) for a in scene.rays(CFG.n_rays, seed=10 + s))
        t.step(o, d, v, scene.targets(CFG.n_rays, seed=10 + s).to(dev))
    torch.cuda.synchronize()
    return t


def test_save_load_trainer_and_scene(dev, tmp_path):
    t = _trained(dev)
    path = tmp_path / "fine_last.tar"
    ck.save_checkpoint(path, t, global_step=3)
    u, step = ck.load_trainer(path, device=dev)
    assert step == 3
    assert torch.equal(u.sdf, t.sdf) and torch.equal(u.k0, t.k0)
    for a, b in zip(u.rgbnet.parameters(), t.rgbnet.parameters()):
        assert torch.equal(a, b)
    for name in ("sdf", "k0", "mlp"):
        assert torch.equal(u.adam[name][0], t.adam[name][0])
        assert torch.equal(u.adam[name][1], t.adam[name][1])
        assert u.adam[name][2] == t.adam[name][2]
    assert torch.equal(u.mask_cache.density, t.mask_cache.density)
    # the next forward of both is bit-identical (no atomics in the forward)
    o, d, v = (a.to(dev) for a in scene.rays(CFG.n_rays, seed=99))
    mt, mu = t.forward(o, d, v), u.forward(o, d, v)
    assert mt == mu
    assert torch.equal(t.P4[:mt], u.P4[:mu]) and torch.equal(t.X[:mt], u.X[:mu])
    # load_checkpoint into an existing trainer gives the same state
    w = FusedFineTrainer(CFG, torch.zeros(1, 1, CFG.R, CFG.R, CFG.R),
                         torch.zeros(1, CFG.feat_dim, CFG.R, CFG.R, CFG.R), device=dev)
    w, step = ck.load_checkpoint(w, path)
    assert step == 3 and torch.equal(w.k0, t.k0) and torch.equal(w.adam["k0"][1], t.adam["k0"][1])
    for a, b in zip(w.rgbnet.parameters(), t.rgbnet.parameters()):
        assert torch.equal(a, b)
    # the server loads the same file
    sc = ck.load_scene(path, device=dev)
    ref = ResidentScene(CFG, t.sdf, t.k0, [(l.weight.detach(), l.bias.detach())
                                           for l in t.rgbnet.linears()], device=dev)
    ref.mask_cache = t.mask_cache
    ws = Workspace(dev, sc.cap)
    a = sc.render(o, d, v, ws).clone()
    b = ref.render(o, d, v, ws).clone()
    assert torch.equal(a, b)


def test_narrower_first_layer_is_zero_padded(dev, tmp_path):
    t = _trained(dev)
    path = tmp_path / "narrow.tar"
    ckpt = ck.save_checkpoint(path, t, global_step=1, with_optimizer=False)
    W0 = ckpt["model_state_dict"]["rgbnet.0.weight"]
    ckpt["model_state_dict"]["rgbnet.0.weight"] = W0[:, :40].clone()
    torch.save(ckpt, path)
    u = FusedFineTrainer(CFG, torch.zeros(1, 1, CFG.R, CFG.R, CFG.R),
                         torch.zeros(1, CFG.feat_dim, CFG.R, CFG.R, CFG.R), device=dev)
    u, _ = ck.load_checkpoint(u, path, no_reload_optimizer=True)
    got = u.rgbnet.linears()[0].weight.detach().cpu()
    assert torch.equal(got[:, :40], W0[:, :40]) and not bool(got[:, 40:].any())
