"""GPU trace counters against hand counts.

A 1x1 film's only camera sample sits at the pixel centre, where the camera
ray is exactly the viewing direction (muon/camera.cc:37-62 with a = b = 0),
so a one-pixel `depth` render traces exactly the ray from the camera eye
towards look_at. The counters of that render (TraceStats, muon/stats.h:11-45)
must equal the hand counts of the reference's traversal over the six-sphere
SAH tree derived in tests/test_oracle_pins.py, and the oracle's counters for
the same render."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene  # noqa: E402
from muon_amd.renderer import render_device  # noqa: E402
from test_oracle_pins import TRACE_RAYS, six_spheres_text  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


def _camera_line(o, d):
    up = (0, 1, 0) if abs(d[1]) < 0.5 else (0, 0, 1)
    look = [o[i] + d[i] for i in range(3)]
    return "camera %r %r %r %r %r %r %d %d %d 45" % (tuple(o) + tuple(look) + up)


@pytest.mark.parametrize("k", range(len(TRACE_RAYS)))
def test_depth_render_counters_hand_counted(ctx, oracle, k):
    o, d, pid, dist, (bt, bh, ot, oh) = TRACE_RAYS[k]
    text = six_spheres_text("integrator depth\npixel_samples 1\n" + _camera_line(o, d))
    hs = HostScene.from_text(text)
    ds = DeviceScene(ctx, hs)
    acc, st = render_device(ds, hs.params)
    torch.cuda.synchronize()
    assert st.primary_rays == 1 and st.secondary_rays == 0
    got = (st.bounds_tests, st.bounds_hits, st.object_tests, st.object_hits)
    assert got == (bt, bh, ot, oh)
    _, _, ost, _ = oracle.load_text(text).render(1)
    assert (int(ost[4]), int(ost[5]), int(ost[2]), int(ost[3])) == got
    # the depth film holds 1 / (1 + distance) (muon/integration.cc:67-72), 0 for a miss
    a = acc.cpu().numpy().reshape(-1)
    assert (a == 0).all() if pid < 0 else (a == np.float32(1.0) / (np.float32(1.0) + np.float32(dist))).all()
