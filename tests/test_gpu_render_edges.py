"""Edge cases of the GPU render path against the oracle: film sizes of one
pixel and ragged sizes (not multiples of any block or wavefront size),
depth limits of the path tracer (muon/integration.cc:29-33 returns before
tracing past the cutoff) and a scene without lights."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene, scenes  # noqa: E402
from muon_amd.renderer import render_device  # noqa: E402
from muon_amd.serving import render_batch  # noqa: E402

SIZES = [(1, 1), (1, 7), (7, 1), (13, 5), (33, 17), (65, 3)]


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


def _cornell(w, h, spp, depth, nee="mis", seed=1):
    return scenes.cornell_text(w, h, spp, depth, seed=seed).replace("next_event_estimation mis",
                                                                    f"next_event_estimation {nee}")


def _render(ctx, text):
    hs = HostScene.from_text(text)
    acc, st = render_device(DeviceScene(ctx, hs), hs.params)
    torch.cuda.synchronize()
    return acc.cpu().numpy(), st


@pytest.mark.parametrize("integrator", ["normals", "depth"])
@pytest.mark.parametrize("size", SIZES, ids=[f"{w}x{h}" for w, h in SIZES])
def test_ragged_films_bit_exact(ctx, oracle, size, integrator):
    w, h = size
    text = _cornell(w, h, 1, 5).replace("integrator pathtracer", f"integrator {integrator}")
    acc, st = _render(ctx, text)
    ref, _, ref_st, _ = oracle.load_text(text).render(1)
    assert acc.shape == (h, w, 3)
    assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))
    assert st.primary_rays == ref_st[0] == w * h


@pytest.mark.parametrize("size", SIZES, ids=[f"{w}x{h}" for w, h in SIZES])
def test_ragged_films_path_tracer(ctx, size):
    """Finite, non-negative films; a batch of the same job under two keys
    and alone agree where the key matches."""
    w, h = size
    hs = HostScene.from_text(_cornell(w, h, 3, 5))
    ds = DeviceScene(ctx, hs)
    alone, st = render_device(ds, hs.params)
    torch.cuda.synchronize()
    accs, _, bst = render_batch(ctx, [(ds, hs.params, 0), (ds, hs.params, 5)])
    a = alone.cpu().numpy()
    assert np.isfinite(a).all() and (a >= 0).all() and st.primary_rays == w * h * 3
    assert np.array_equal(accs[0].cpu().numpy().view(np.uint32), a.view(np.uint32))
    assert bst.primary_rays == 2 * w * h * 3


def test_depth_limits(ctx, oracle):
    w, h, spp = 16, 16, 2
    # max_depth 0 with NEE: the cutoff is -1, nothing is traced
    text = _cornell(w, h, spp, 0, "mis")
    acc, st = _render(ctx, text)
    _, _, ref_st, _ = oracle.load_text(text).render(1)
    assert not acc.any() and st.primary_rays == ref_st[0] == 0 and st.secondary_rays == 0 and st.shadow_rays == 0
    # max_depth 0 without NEE: camera rays only, emission seen directly
    text = _cornell(w, h, spp, 0, "off")
    acc, st = _render(ctx, text)
    ref, _, ref_st, _ = oracle.load_text(text).render(1)
    assert st.primary_rays == ref_st[0] == w * h * spp and st.secondary_rays == ref_st[1] == 0
    assert np.isfinite(acc).all() and (acc >= 0).all()
    # max_depth 1 with MIS: every camera ray hits the box and draws one MIS ray
    text = _cornell(w, h, spp, 1, "mis")
    acc, st = _render(ctx, text)
    _, _, ref_st, _ = oracle.load_text(text).render(1)
    assert st.primary_rays == ref_st[0] == w * h * spp
    assert st.secondary_rays == ref_st[1] == w * h * spp


def test_scene_without_lights(ctx, oracle):
    text = _cornell(24, 24, 4, 5)
    text = "\n".join(line for line in text.splitlines() if not line.startswith("quad_light")) + "\n"
    acc, st = _render(ctx, text)
    ref, _, _, _ = oracle.load_text(text).render(1)
    assert not ref.any() and not acc.any()
    assert st.primary_rays == 24 * 24 * 4 and st.shadow_rays == 0
