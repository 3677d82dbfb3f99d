"""Every path-tracer estimator option of the GPU engine against the oracle
(the CPU restatement of muon/integration.cc): next-event estimation off / on
/ MIS x hemisphere / cosine / BRDF importance sampling, plus Russian
roulette, several (stratified) light samples and a minimum depth, on the
Cornell box with a GGX and a Phong sphere added
(tests/golden/make_estimator_fixture.py wrote the CPU side). For each option
the mean radiance over 64 render seeds, overall and in 2 x 2 image blocks,
must agree between the GPU and the oracle within 5 standard errors of the
difference. Both are unbiased estimators of the same image (the reference's
options do not agree with each other: its MIS adds more light than its NEE,
and that is what is matched), so a wrong term, weight or pdf in any option
shows up as a bias."""
import importlib.util
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene  # noqa: E402
from muon_amd.serving import render_batch  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_estimator_fixture",
                                               os.path.join(HERE, "golden", "make_estimator_fixture.py"))
fx = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(fx)


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


@pytest.fixture(scope="module")
def cpu_side():
    return np.load(os.path.join(HERE, "golden", "estimators_cornell32.npz"))


@pytest.mark.parametrize("k", range(len(fx.OPTIONS)),
                         ids=[f"{n}-{i}{'-' + e.split()[0] if e else ''}" for n, i, e in fx.OPTIONS])
def test_estimator_option_unbiased(ctx, cpu_side, k):
    cpu = cpu_side["block_means"][k]
    n, skip = cpu.shape[0], int(cpu_side["meta"][5])
    texts = [fx.scene_text(s, *fx.OPTIONS[k]) for s in range(n)]
    ds = DeviceScene(ctx, HostScene.from_text(texts[0]))
    accs, _, _ = render_batch(ctx, [(ds, HostScene.from_text(t).params, 0) for t in texts])
    gpu = np.array([fx.block_means(a.cpu().numpy() / fx.SPP, skip) for a in accs])
    assert np.isfinite(gpu).all() and (gpu >= 0).all()
    se = np.sqrt(cpu.var(axis=0) / n + gpu.var(axis=0) / n)
    z = (gpu.mean(axis=0) - cpu.mean(axis=0)) / se
    assert np.abs(z).max() < 5.0, (z, gpu.mean(axis=0) / cpu.mean(axis=0))
    g, c = gpu.mean(axis=(1, 2)), cpu.mean(axis=(1, 2))
    zt = (g.mean() - c.mean()) / np.sqrt(g.var() / n + c.var() / n)
    assert abs(zt) < 5.0, (zt, g.mean() / c.mean())
