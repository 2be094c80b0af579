"""Writes the golden fixtures of tests/golden/*.npz from the oracle (seeded inputs and
the oracle's outputs).  Run: python tests/golden/gen_golden.py

The reference ships no test vectors; these fixtures freeze the pinned oracle's
behaviour on small hand-built and seeded cases so that (a) a change to the oracle is
caught by tests/test_golden.py on the CPU and (b) the HIP library is checked against
fixed vectors on the GPU.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests.raysets import mixed_rays  # noqa: E402


def sampling():
    out = {}
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        o, d = mixed_rays(48, seed=21, dtype=dt)
        mn = np.array([-1.0, -0.8, -1.2], dt)
        mx = np.array([1.0, 0.9, 1.1], dt)
        step = np.float32(0.05)
        res = oracle.sample_pts_on_rays(o, d, mn, mx, 0.1, 1e9, step)
        out.update({f"{tag}_rays_o": o, f"{tag}_rays_d": d, f"{tag}_xyz_min": mn, f"{tag}_xyz_max": mx})
        for name, v in zip(["rays_pts", "mask_outbbox", "ray_id", "step_id", "N_steps", "t_min", "t_max"], res):
            out[f"{tag}_{name}"] = v
    out["near"] = np.float32(0.1)
    out["stepdist"] = np.float32(0.05)
    return out


def alpha2weight():
    rng = np.random.default_rng(22)
    lens = rng.integers(0, 60, 40)
    rid = np.repeat(np.arange(40), lens).astype(np.int64)
    alpha = (rng.random(len(rid)) ** 6).astype(np.float32)
    alpha[rng.random(len(rid)) < 0.05] = 0.95
    w, T, last, i_s, i_e = oracle.alpha2weight(alpha, rid, 40)
    gw = rng.normal(size=len(rid)).astype(np.float32)
    gl = rng.normal(size=40).astype(np.float32)
    g = oracle.alpha2weight_backward(alpha, w, T, last, i_s, i_e, 40, gw, gl)
    return dict(alpha=alpha, ray_id=rid, weight=w, T=T, alphainv_last=last, i_start=i_s, i_end=i_e,
                grad_weights=gw, grad_last=gl, grad=g)


def maskcache():
    rng = np.random.default_rng(23)
    world = rng.random((7, 6, 5)) > 0.5
    xyz = rng.uniform(-1.1, 1.1, (300, 3)).astype(np.float32)
    xyz[:3] = [[0.5, 0.5, 0.5], [-0.5, 0.0, 0.25], [1.0, 1.0, 1.0]]
    scale = np.array([3.0, 2.5, 2.0], np.float32)
    shift = np.array([3.0, 2.5, 2.0], np.float32)
    return dict(world=world, xyz=xyz, scale=scale, shift=shift,
                out=oracle.maskcache_lookup(world, xyz, scale, shift))


def raw2alpha():
    rng = np.random.default_rng(24)
    dens = rng.normal(0, 3, 200).astype(np.float32)
    e, a = oracle.raw2alpha(dens, -1.5, 0.5)
    gb = rng.normal(size=200).astype(np.float32)
    g = oracle.raw2alpha_backward(e, gb, 0.5)
    return dict(density=dens, exp_d=e, alpha=a, grad_back=gb, grad=g)


def main():
    for name, fn in (("sampling", sampling), ("alpha2weight", alpha2weight),
                     ("maskcache", maskcache), ("raw2alpha", raw2alpha)):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **fn())
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
