"""The committed golden fixtures (tests/golden/*.npz, written by gen_golden.py):
the oracle must reproduce them exactly (CPU), and the HIP ops must reproduce them
exactly for integer/index outputs and the float outputs that share the oracle's
operation order, within 1e-5 relative for exp/pow results and 1e-4 for adjoints (GPU).
"""
import os

import numpy as np
import pytest
import torch

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False))


NAMES = ["rays_pts", "mask_outbbox", "ray_id", "step_id", "N_steps", "t_min", "t_max"]


@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_oracle_sampling_fixture(tag):
    g = load("sampling")
    res = oracle.sample_pts_on_rays(g[f"{tag}_rays_o"], g[f"{tag}_rays_d"], g[f"{tag}_xyz_min"],
                                    g[f"{tag}_xyz_max"], float(g["near"]), 1e9, g["stepdist"])
    for n, v in zip(NAMES, res):
        np.testing.assert_array_equal(v, g[f"{tag}_{n}"], err_msg=n)


def test_oracle_other_fixtures():
    a = load("alpha2weight")
    res = oracle.alpha2weight(a["alpha"], a["ray_id"], 40)
    for n, v in zip(["weight", "T", "alphainv_last", "i_start", "i_end"], res):
        np.testing.assert_array_equal(v, a[n], err_msg=n)
    g = oracle.alpha2weight_backward(a["alpha"], *res, 40, a["grad_weights"], a["grad_last"])
    np.testing.assert_array_equal(g, a["grad"])
    m = load("maskcache")
    np.testing.assert_array_equal(oracle.maskcache_lookup(m["world"], m["xyz"], m["scale"], m["shift"]), m["out"])
    r = load("raw2alpha")
    e, al = oracle.raw2alpha(r["density"], -1.5, 0.5)
    np.testing.assert_array_equal(al, r["alpha"])
    np.testing.assert_array_equal(oracle.raw2alpha_backward(e, r["grad_back"], 0.5), r["grad"])


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["f32", "f64"])
def test_hip_sampling_fixture(dev, tag):
    from voxurf_amd import render_utils_cuda as ru
    g = load("sampling")
    t = lambda k: torch.from_numpy(g[f"{tag}_{k}"]).to(dev)  # noqa: E731
    res = ru.sample_pts_on_rays(t("rays_o"), t("rays_d"), t("xyz_min"), t("xyz_max"),
                                float(g["near"]), 1e9, float(g["stepdist"]))
    for n, v in zip(NAMES, res):
        np.testing.assert_array_equal(v.cpu().numpy(), g[f"{tag}_{n}"], err_msg=n)


@pytest.mark.gpu
def test_hip_other_fixtures(dev):
    from voxurf_amd import render_utils_cuda as ru
    a = load("alpha2weight")
    T = lambda x: torch.from_numpy(x).to(dev)  # noqa: E731
    res = ru.alpha2weight(T(a["alpha"]), T(a["ray_id"]), 40)
    for n, v in zip(["weight", "T", "alphainv_last", "i_start", "i_end"], res):
        np.testing.assert_array_equal(v.cpu().numpy(), a[n], err_msg=n)
    g = ru.alpha2weight_backward(T(a["alpha"]), *res, 40, T(a["grad_weights"]), T(a["grad_last"]))
    np.testing.assert_allclose(g.cpu().numpy(), a["grad"], rtol=1e-4, atol=1e-6 * np.abs(a["grad"]).max())
    m = load("maskcache")
    out = ru.maskcache_lookup(T(m["world"]), T(m["xyz"]), T(m["scale"]), T(m["shift"]))
    np.testing.assert_array_equal(out.cpu().numpy(), m["out"])
    r = load("raw2alpha")
    e, al = ru.raw2alpha(T(r["density"]), -1.5, 0.5)
    np.testing.assert_allclose(al.cpu().numpy(), r["alpha"], rtol=1e-5, atol=1e-7)
    gr = ru.raw2alpha_backward(e, T(r["grad_back"]), 0.5)
    np.testing.assert_allclose(gr.cpu().numpy(), r["grad"], rtol=1e-4, atol=1e-7)
