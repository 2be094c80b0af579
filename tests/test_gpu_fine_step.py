"""End-to-end training-step parity on the GPU: the fused per-ray path and the op-by-op
path against the CPU restatement of the reference step (oracle/torch_ref.train_step).

Scene: the analytic SDF scene at reduced size (64^3, 12-channel features, the fine
rgbnet 42-192x3-3, 1024 rays, s = 256) so the CPU restatement runs in seconds.

Bars:
  * survivor sets (samples with weight > 1e-4) equal: fused == op-by-op exactly, and
    each equal to the CPU restatement's per-ray counts except for rays where a weight
    sits within float rounding of the threshold (at most 0.5% of rays);
  * rgb_marched and the loss within 1e-5 relative (floor 1e-6 of the batch maximum);
  * gradients of sdf, k0 and rgbnet within 1e-4 relative on the same footing, on the
    rays whose survivor sets agree;
  * parameters after the Adam step within 1e-5 relative.
"""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import torch_ref as tr
from voxurf_amd import scene
from voxurf_amd.fused import FusedFineTrainer
from voxurf_amd.model import Trainer, VoxurfFine

pytestmark = pytest.mark.gpu

CFG = dataclasses.replace(scene.FINE, R=64, n_rays=1024, seed=3)


def close(a, b, rtol, floor, what):
    a = a.detach().float().cpu()
    b = b.detach().float().cpu()
    scale = float(b.abs().max()) if b.numel() else 1.0
    err = (a - b).abs()
    bound = rtol * b.abs() + floor * scale
    bad = err > bound
    assert not bool(bad.any()), f"{what}: {int(bad.sum())} of {b.numel()} out of tolerance, " \
                                f"max err {float(err.max()):.3g} (scale {scale:.3g})"


@pytest.fixture(scope="module")
def setup():
    torch.manual_seed(0)
    st = tr.make_state(CFG)
    o, d, v = scene.rays(CFG.n_rays, seed=CFG.seed)
    tg = scene.targets(CFG.n_rays, seed=CFG.seed)
    ref_state = {k: (val.clone() if torch.is_tensor(val) else [(w.clone(), b.clone()) for w, b in val])
                 for k, val in st.items()}
    res = tr.train_step(ref_state, {}, o, d, v, tg, CFG, 1)
    return st, (o, d, v, tg), res, ref_state


def _per_ray_counts(ray_id, n):
    return torch.bincount(ray_id.long().cpu(), minlength=n)


@pytest.mark.parametrize("path", ["fused", "ops"])
def test_fine_step_matches_restatement(dev, setup, path):
    st, (o, d, v, tg), ref, ref_after = setup
    O, D, V, TG = (t.to(dev) for t in (o, d, v, tg))
    if path == "fused":
        t = FusedFineTrainer(CFG, st["sdf"], st["k0"], st["mlp"], device=dev)
        out = t.step(O, D, V, TG, profile=True)
        rid = t.rid[: t.m]
        grads = t.grads_snapshot
        sdf_after, k0_after = t.sdf, t.k0
        mlp_after = list(t.rgbnet.parameters())
    else:
        model = VoxurfFine(CFG, st["sdf"], st["k0"], st["mlp"]).to(dev)
        t = Trainer(model, CFG)
        out = t.step(O, D, V, TG, profile=True)
        rid = out["ray_id"]
        grads = t.grads_snapshot
        sdf_after, k0_after = model.sdf.grid, model.k0.grid
        mlp_after = list(model.rgbnet.parameters())
    n = CFG.n_rays
    c_gpu = _per_ray_counts(rid, n)
    c_ref = _per_ray_counts(ref["ray_id"], n)
    agree = c_gpu == c_ref
    print(path, "survivors", int(c_gpu.sum()), "ref", int(c_ref.sum()), "rays differing",
          int((~agree).sum()))
    assert int((~agree).sum()) <= max(1, n // 200)
    rgb = out["rgb_marched"].cpu()
    close(rgb[agree], ref["rgb"][agree], 1e-5, 1e-6, "rgb_marched")
    if bool(agree.all()):
        assert abs(float(out["loss"]) - ref["loss"]) <= 1e-5 * abs(ref["loss"]) + 1e-9
        close(grads["sdf"], ref["grads"]["sdf"], 1e-4, 1e-4, "sdf grad")
        close(grads["k0"], ref["grads"]["k0"], 1e-4, 1e-4, "k0 grad")
        names = [f"mlp{i // 2}.{'w' if i % 2 == 0 else 'b'}" for i in range(len(grads["mlp"]))]
        for g, nm in zip(grads["mlp"], names):
            close(g, ref["grads"][nm], 1e-4, 1e-4, nm + " grad")
        # Adam at step 1 moves each element by about -lr * g / (|g| + eps), so elements
        # whose gradient is within rounding noise of zero may move differently from the
        # restatement.  The optimiser is therefore checked against utils.Adam applied
        # to this path's own gradients (1e-5), and against the restatement's parameters
        # only where the gradient is well above its noise floor: the step's sensitivity
        # to g is lr * eps / (|g| + eps)^2, so a 1e-4 relative gradient difference moves
        # the parameter by up to lr * 2.5e-5 when |g| ~ eps; from |g| > 1e3 * eps on it
        # is below lr * 1e-7.
        pairs = [("sdf", st["sdf"], grads["sdf"], sdf_after, CFG.lr_sdf, ref_after["sdf"], ref["grads"]["sdf"]),
                 ("k0", st["k0"], grads["k0"], k0_after, CFG.lr_k0, ref_after["k0"], ref["grads"]["k0"])]
        flat0 = [x for wb in st["mlp"] for x in wb]
        flat1 = [x for wb in ref_after["mlp"] for x in wb]
        for i, (p0, g, p1, q1) in enumerate(zip(flat0, grads["mlp"], mlp_after, flat1)):
            pairs.append((f"rgbnet[{i}]", p0, g, p1, CFG.lr_mlp, q1, ref["grads"][names[i]]))
        for name, p0, g, p1, lr, q1, gref in pairs:
            exp = p0.clone().double()
            tr.adam_step(exp, g.detach().cpu().double(), torch.zeros_like(exp), torch.zeros_like(exp),
                         1, lr, 0.9, 0.99, 1e-8)
            close(p1, exp, 1e-5, 1e-7, name + " after Adam (own gradient)")
            big = (gref.abs() > 1e-3 * float(gref.abs().max())) & (gref.abs() > 1e-5)
            close(p1.detach().cpu()[big.cpu()], q1[big], 1e-5, 1e-6, name + " after Adam vs restatement")


def test_fused_equals_ops_survivors_exactly(dev, setup):
    st, (o, d, v, tg), _, _ = setup
    O, D, V, TG = (t.to(dev) for t in (o, d, v, tg))
    f = FusedFineTrainer(CFG, st["sdf"], st["k0"], st["mlp"], device=dev)
    f.forward(O, D, V)
    model = VoxurfFine(CFG, st["sdf"], st["k0"], st["mlp"]).to(dev)
    with torch.no_grad():
        out = model(O, D, V)
    assert f.m == out["n_samples"]
    assert torch.equal(f.rid[: f.m].long(), out["ray_id"])
    assert torch.equal(f.P4[: f.m, 3], out["weights"])
    assert torch.equal(f.last, out["alphainv_cum"])


def test_split_k0_update_over_consecutive_steps(dev, setup):
    """Three consecutive fused steps with the k0 Adam pass split by page flags and run
    on the second stream (alternating flag buffers, page lists built by phase 1): after
    every step, k0 and its Adam moments equal utils.Adam applied, in order and to the
    whole grid, to the state before the step and that step's k0 gradient.  (Comparing
    two multi-step runs directly is ill-posed: the gradient atomics perturb the weights
    by rounding, a sample whose weight sits at the 1e-4 threshold can flip, and Adam
    turns a flipped 1e-10 gradient into a full-size step on the 8 corners involved.)"""
    st, (o, d, v, tg), _, _ = setup
    O, D, V, TG = (x.to(dev) for x in (o, d, v, tg))
    t = FusedFineTrainer(CFG, st["sdf"], st["k0"], st["mlp"], device=dev)
    assert t.overlap_adam
    p = t.k0.detach().cpu().double()
    m = torch.zeros_like(p)
    vv = torch.zeros_like(p)
    for i in range(3):
        k = (i * 97) % CFG.n_rays  # a different batch each step: rotate the rays
        t.step(torch.roll(O, k, 0), torch.roll(D, k, 0), torch.roll(V, k, 0),
               torch.roll(TG, k, 0), profile=True)
        g = t.grads_snapshot["k0"].cpu().double()
        tr.adam_step(p, g, m, vv, i + 1, CFG.lr_k0, 0.9, 0.99, 1e-8)
        close(t.k0, p, 1e-5, 1e-7, f"k0 after step {i + 1}")
        close(t.adam["k0"][0], m, 1e-5, 1e-7, f"k0 exp_avg after step {i + 1}")
        close(t.adam["k0"][1], vv, 1e-5, 1e-7, f"k0 exp_avg_sq after step {i + 1}")
        # the next step starts from the trainer's own state (removes float32 drift)
        p = t.k0.detach().cpu().double()
        m = t.adam["k0"][0].detach().cpu().double()
        vv = t.adam["k0"][1].detach().cpu().double()


def test_touched_page_skip_over_changing_batches(dev, setup):
    """Three steps on different ray batches: pages touched by an earlier step but not
    flagged now are swept with their decaying moments, pages never touched are skipped
    (their update is the identity); after every step k0 and the SDF grid, with their
    moments, equal utils.Adam applied in order to the whole grid."""
    st, _, _, _ = setup
    t = FusedFineTrainer(CFG, st["sdf"], st["k0"], st["mlp"], device=dev)
    p = t.k0.detach().cpu().double()
    m = torch.zeros_like(p)
    vv = torch.zeros_like(p)
    ps = t.sdf.detach().cpu().double()
    ms = torch.zeros_like(ps)
    vs = torch.zeros_like(ps)
    for i in range(3):
        o, d, v = (x.to(dev) for x in scene.rays(CFG.n_rays, seed=50 + i))
        t.step(o, d, v, scene.targets(CFG.n_rays, seed=50 + i).to(dev), profile=True)
        g = t.grads_snapshot["k0"].cpu().double()
        tr.adam_step(p, g, m, vv, i + 1, CFG.lr_k0, 0.9, 0.99, 1e-8)
        close(t.k0, p, 1e-5, 1e-7, f"k0 after step {i + 1}")
        close(t.adam["k0"][0], m, 1e-5, 1e-7, f"k0 exp_avg after step {i + 1}")
        close(t.adam["k0"][1], vv, 1e-5, 1e-7, f"k0 exp_avg_sq after step {i + 1}")
        # the SDF grid's two-phase pass (page flags from the forward) likewise
        gs = t.grads_snapshot["sdf"].cpu().double()
        tr.adam_step(ps, gs, ms, vs, i + 1, CFG.lr_sdf, 0.9, 0.99, 1e-8)
        close(t.sdf, ps, 1e-5, 1e-7, f"sdf after step {i + 1}")
        close(t.adam["sdf"][0], ms, 1e-5, 1e-7, f"sdf exp_avg after step {i + 1}")
        close(t.adam["sdf"][1], vs, 1e-5, 1e-7, f"sdf exp_avg_sq after step {i + 1}")
        if i > 0:
            assert t.last_swept_pages > 0   # touched earlier, not flagged now
        p = t.k0.detach().cpu().double()
        m = t.adam["k0"][0].detach().cpu().double()
        vv = t.adam["k0"][1].detach().cpu().double()
        ps = t.sdf.detach().cpu().double()
        ms = t.adam["sdf"][0].detach().cpu().double()
        vs = t.adam["sdf"][1].detach().cpu().double()
    assert 0 < t.touched_pages() < t.n_pages
    # untouched pages: parameters exactly as initialised, moments exactly zero
    untouched = (t._touched == 0).cpu()
    flat = lambda x: x.detach().permute(0, 2, 3, 4, 1).reshape(-1).cpu()  # noqa: E731
    k0_init = st["k0"].permute(0, 2, 3, 4, 1).reshape(-1)
    n = k0_init.numel()
    sel = untouched.repeat_interleave(t.page)[:n]
    assert torch.equal(flat(t.k0)[sel], k0_init[sel])
    assert not bool(flat(t.adam["k0"][0])[sel].any()) and not bool(flat(t.adam["k0"][1])[sel].any())
