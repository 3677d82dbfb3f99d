"""The muonx command line (muon's flags, muon/muon.cc:12-49) as a drop-in
for the reference binary: run as a child process on reference scenes, its
PNG equals the reference's golden image, a scene that loads an OBJ mesh
renders as the same triangles written inline, and flag errors exit with
status 1 before any device work."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from PIL import Image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "muon_amd", "lib", "muonx")


def _run(args, timeout=120):
    if not os.path.exists(CLI):
        pytest.fail("muonx CLI not built: run __graft_entry__.build()")
    return subprocess.run([CLI] + args, capture_output=True, text=True, timeout=timeout)


def _png(path):
    return np.asarray(Image.open(path).convert("RGB"))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cornell_normals", "cornell_depth", "spheres_raytrace"])
def test_cli_matches_reference_golden(tmp_path, ref_dir, name):
    shutil.copy(os.path.join(ref_dir, name + ".muon"), tmp_path / "scene.muon")
    out = tmp_path / "out.png"
    r = _run(["--scene=%s" % (tmp_path / "scene.muon"), "--output=%s" % out, "--seed=3", "--stats=true"])
    assert r.returncode == 0, r.stderr
    for label in ("Primary rays", "Ray-object tests", "Ray-object hits", "Bounds tests", "Bounds hits"):
        assert label in r.stderr
    got, gold = _png(out), _png(os.path.join(ref_dir, name + ".png"))
    assert got.shape == gold.shape
    assert int((got != gold).any(axis=2).sum()) == 0


@pytest.mark.gpu
def test_cli_load_renders_like_inline_mesh(tmp_path):
    obj = "v -1 -1 0\nv 1 -1 0\nv 1 1 0\nv -1 1 0\nv 0 0 1\nf 1 2 3 4\nf 1 2 5\nf 2 3 5\n"
    (tmp_path / "mesh.obj").write_text(obj)
    head = ("film_size 96 64\npixel_samples 1\nintegrator normals\ncamera 0 -3 2 0 0 0.3 0 0 1 45\n"
            "rotate 0 0 1 20\n")
    (tmp_path / "load.muon").write_text(head + "load mesh.obj\n")
    inline = ("vertex -1 -1 0\nvertex 1 -1 0\nvertex 1 1 0\nvertex -1 1 0\nvertex 0 0 1\n"
              "tri 0 1 2\ntri 0 2 3\ntri 0 1 4\ntri 1 2 4\n")
    (tmp_path / "inline.muon").write_text(head + inline)
    for n in ("load", "inline"):
        r = _run(["--scene=%s" % (tmp_path / (n + ".muon")), "--output=%s" % (tmp_path / (n + ".png"))])
        assert r.returncode == 0, r.stderr
    a, b = _png(tmp_path / "load.png"), _png(tmp_path / "inline.png")
    assert np.array_equal(a, b)
    assert (a.sum(axis=2) > 0).mean() > 0.2


@pytest.mark.gpu
@pytest.mark.parametrize("feexcept", [False, True], ids=["default", "enable_feexcept"])
def test_cli_writes_the_scenes_output(tmp_path, ref_dir, feexcept):
    """Without --output the render goes to the scene's `output` file
    (muon/renderer.cc:26-27, relative to the working directory); with
    --enable_feexcept (muon/debug.cc:7-17) the host-side work raises no
    floating-point trap and the image is the same."""
    text = open(os.path.join(ref_dir, "cornell_normals.muon")).read() + "\noutput from_scene.png\n"
    (tmp_path / "scene.muon").write_text(text)
    args = [CLI, "--scene=scene.muon"] + (["--enable_feexcept"] if feexcept else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert not (tmp_path / "raytrace.png").exists()
    got, gold = _png(tmp_path / "from_scene.png"), _png(os.path.join(ref_dir, "cornell_normals.png"))
    assert int((got != gold).any(axis=2).sum()) == 0


@pytest.mark.gpu
def test_python_renderer_writes_the_scenes_output(tmp_path, ref_dir, monkeypatch):
    from muon_amd import Options, Renderer
    text = open(os.path.join(ref_dir, "cornell_depth.muon")).read() + "\noutput depth_out.png\n"
    (tmp_path / "scene.muon").write_text(text)
    monkeypatch.chdir(tmp_path)
    r = Renderer(str(tmp_path / "scene.muon"), Options(show_stats=False))
    r.Render()
    assert r.output == "depth_out.png"
    got, gold = _png(tmp_path / "depth_out.png"), _png(os.path.join(ref_dir, "cornell_depth.png"))
    assert int((got != gold).any(axis=2).sum()) == 0
    Renderer(str(tmp_path / "scene.muon"), Options(output=str(tmp_path / "o.png"), show_stats=False)).Render()
    assert np.array_equal(_png(tmp_path / "o.png"), got)


def test_cli_flag_errors():
    r = _run([])
    assert r.returncode == 1 and "scene file is required" in r.stderr
    r = _run(["--scene=x.muon", "--bogus=1"])
    assert r.returncode == 1 and "unknown flag" in r.stderr
    r = _run(["--scene=x.muon", "--acceleration=octree"])
    assert r.returncode == 1


@pytest.mark.gpu
def test_cli_stats_count_path_tracer_work(tmp_path, ref_dir):
    """--stats on a path-traced scene prints the traversal's test counters
    (the device tree's work), not zeros."""
    import re
    shutil.copy(os.path.join(ref_dir, "cornell_brdf.muon"), tmp_path / "scene.muon")
    r = _run(["--scene=%s" % (tmp_path / "scene.muon"), "--output=%s" % (tmp_path / "o.png"), "--stats=true"])
    assert r.returncode == 0, r.stderr
    got = {k: int(re.search(k + r"\s*:\s*(\d+)", r.stderr).group(1))
           for k in ("Primary rays", "Ray-object tests", "Ray-object hits", "Bounds tests", "Bounds hits")}
    assert got["Primary rays"] > 0 and got["Ray-object tests"] > got["Ray-object hits"] > 0
    assert got["Bounds tests"] > got["Bounds hits"] > 0
