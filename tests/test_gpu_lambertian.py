"""The GPU's Lambertian BRDF (muon/materials.cc:17-47) against the oracle's.

The reference's parser never builds it (`brdf lambertian` falls through to
Phong, muon/parser.cc:165-178), so the GPU gets it through the scene
descriptor's material tag MUONX_BRDF_LAMBERTIAN and the oracle through its
test hook load_text_lambertian (tests/golden/make_lambertian_fixture.py
wrote the CPU side). The Cornell box's walls and boxes carry a specular
colour of 0.5, which the Lambertian BRDF ignores and a Phong BRDF would
turn into a lobe about three times as bright: the GPU mean over 64 render
keys must equal the oracle's over 64 seeds within 5 standard errors,
overall and in 2 x 2 blocks."""
import importlib.util
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene  # noqa: E402
from muon_amd.serving import render_batch  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_lambertian_fixture.py"))
mk = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mk)

MUONX_BRDF_LAMBERTIAN, MUONX_BRDF_PHONG = 0, 1


def test_lambertian_descriptor_unbiased():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    ctx = Context(0)
    fx = np.load(os.path.join(HERE, "golden", "lambertian_cornell32.npz"))
    cpu, n, skip = fx["block_means"], fx["block_means"].shape[0], int(fx["meta"][5])
    hs = HostScene.from_text(mk.scene_text(0))
    mats = hs.arrays()["materials"].copy()
    assert (mats[:, 14] == MUONX_BRDF_PHONG).all()  # what the parser builds for `brdf lambertian`
    assert (mats[:, 6:9] == 0.5).any()
    mats[:, 14] = MUONX_BRDF_LAMBERTIAN
    ds = DeviceScene.from_arrays(ctx, hs, materials=mats.reshape(-1))
    accs, _, _ = render_batch(ctx, [(ds, hs.params, k) for k in range(n)])
    gpu = np.array([mk.block_means(a.cpu().numpy() / mk.SPP, skip) for a in accs])
    se = np.sqrt(cpu.var(axis=0) / n + gpu.var(axis=0) / n)
    z = (gpu.mean(axis=0) - cpu.mean(axis=0)) / se
    assert np.abs(z).max() < 5.0, (z, gpu.mean(axis=0) / cpu.mean(axis=0))
    # the same descriptor with the parser's Phong tag renders the specular lobe
    ds_phong = DeviceScene(ctx, hs)
    (acc,), _, _ = render_batch(ctx, [(ds_phong, hs.params, 0)])
    assert acc.cpu().numpy().mean() > 2 * gpu.mean() * mk.SPP
