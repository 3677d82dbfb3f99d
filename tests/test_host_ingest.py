"""The product's host ingest (parser + SAH BVH build behind the C ABI)
matches the oracle bit for bit on the reference's scenes and on the
synthetic workloads: primitive order and geometry, transform bits, material
values, lights, camera, path seed, and the exact SAH tree and leaf order."""
import os

import numpy as np
import pytest

from muon_amd import HostScene, scenes

REF_SCENES = ["intersection", "transforms", "spheres_raytrace", "cornell_raytrace", "analytic", "cornell_normals",
              "cornell_depth", "cornell_mesh", "partial_emission", "random_seed", "sphere", "cornell_phong",
              "cornell_brdf", "ggx", "mis", "partial_direct", "partial_indirect"]


def _compare(host, orc):
    a = host.arrays()
    assert host.desc.num_prims == orc.num_prims
    op = orc.prims()
    n = orc.num_prims
    assert np.array_equal(a["prim_type"], op["type"].astype(np.uint8))
    # geometry in parse order
    g = a["prim_geom"]
    assert np.array_equal(g.view(np.uint32), op["geom"].view(np.uint32))
    # transform records by bit pattern
    xf = a["transforms"][a["prim_transform"]]
    assert np.array_equal(xf.view(np.uint32), op["xform"].view(np.uint32))
    # materials: ambient, diffuse, specular, emission, shininess, roughness, brdf
    m = a["materials"][a["prim_material"]]
    assert np.array_equal(m[:, :14].view(np.uint32), op["mat"][:, :14].view(np.uint32))
    assert np.array_equal(m[:, 14], op["mat"][:, 14])
    assert np.array_equal(a["prim_light"], op["light"])
    # BVH: pre-order nodes, bounds bit-exact, leaf ranges, split axes
    b = orc.bvh()
    assert np.array_equal(a["node_bounds"].view(np.uint32), b["bounds"].view(np.uint32))
    assert np.array_equal(a["node_count"], b["count"].astype(np.int32))
    leaf = b["count"] > 0
    assert np.array_equal(a["node_start"][leaf], b["start"][leaf].astype(np.int32))
    assert np.array_equal(a["node_axis"][~leaf], b["axis"][~leaf])
    assert np.array_equal(a["node_left"], b["left"].astype(np.int32))
    assert np.array_equal(a["node_right"], b["right"].astype(np.int32))
    assert np.array_equal(a["leaf_prims"], b["prim_order"])
    # render settings and camera
    p = host.params
    assert (p.width, p.height, p.spp) == (orc.width, orc.height, orc.spp)
    assert (p.min_depth, p.max_depth, p.nee, p.importance_sampling) == (orc.min_depth, orc.max_depth, orc.nee,
                                                                         orc.importance_sampling)
    assert (p.russian_roulette, p.light_samples, p.light_stratify) == (orc.russian_roulette, orc.light_samples,
                                                                        orc.light_stratify)
    assert np.float32(p.gamma) == np.float32(orc.gamma)
    cam = orc.camera()
    if cam is not None:
        mine = np.array(list(p.eye) + list(p.cam_u) + list(p.cam_v) + list(p.cam_w) + [p.tan_fov_x, p.tan_fov_y],
                        np.float32)
        assert np.array_equal(mine.view(np.uint32), cam.view(np.uint32))
    assert host.warnings == orc.warnings


@pytest.mark.parametrize("name", REF_SCENES)
def test_reference_scene_ingest_matches_oracle(oracle, ref_dir, name):
    path = os.path.join(ref_dir, name + ".muon")
    host = HostScene.from_file(path, default_seed=7)
    orc = oracle.load(path, seed=7)
    _compare(host, orc)
    if host.params.integrator == 5:
        assert host.params.seed == orc.pathtracer_seed()


def test_dragon_ingest_matches_oracle(oracle, ref_dir):
    """The reference's 100k-triangle dragon scene (stored compressed)."""
    import gzip
    with gzip.open(os.path.join(ref_dir, "dragon.muon.gz"), "rt") as f:
        text = f.read()
    host = HostScene.from_text(text, default_seed=7)
    assert host.desc.num_prims == 100008
    _compare(host, oracle.load_text(text, seed=7))
    assert host.params.seed == oracle.load_text(text, seed=7).pathtracer_seed()


@pytest.mark.parametrize("part", [0, 1, 2])
def test_partition_strategies_match_oracle(oracle, ref_dir, part):
    path = os.path.join(ref_dir, "spheres_raytrace.muon")
    _compare(HostScene.from_file(path, partition_strategy=part, default_seed=1), oracle.load(path, partition=part, seed=1))


def test_linear_structure_is_one_leaf_in_parse_order(ref_dir):
    h = HostScene.from_file(os.path.join(ref_dir, "cornell_normals.muon"), acceleration=0)
    a = h.arrays()
    assert h.desc.num_nodes == 1 and a["node_count"][0] == h.desc.num_prims
    assert np.array_equal(a["leaf_prims"], np.arange(h.desc.num_prims))


def test_cornell_workload_ingest(oracle):
    text = scenes.cornell_text()
    _compare(HostScene.from_text(text), oracle.load_text(text))


def test_icosphere_workload_ingest_small(oracle):
    text = scenes.icosphere_text(level=3)
    h = HostScene.from_text(text)
    _compare(h, oracle.load_text(text))
    assert h.desc.num_prims == 20 * 4 ** 3 + 10 + 2


def test_icosphere_workload_ingest_threaded_build(oracle):
    """At 327,692 primitives the host build splits large subtrees over
    threads (and parses with the fast number reader): the tree, the leaf
    order and every array stay the sequential oracle's."""
    text = scenes.icosphere_text(level=7)
    _compare(HostScene.from_text(text), oracle.load_text(text))


def test_icosphere_faces_point_outward():
    v, f = scenes.icosphere(3)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    assert (np.einsum("ij,ij->i", n, v[f[:, 0]]) > 0).all()
    assert len(f) == 20 * 4 ** 3


def test_sphere_field_ingest_small(oracle):
    text = scenes.sphere_field_text(n=300)
    _compare(HostScene.from_text(text), oracle.load_text(text))


def test_quirks(oracle):
    # random_seed falls through into film_size; lambertian builds Phong;
    # camera pushes the transform stack; unknown commands are counted.
    text = ("random_seed 5 320 200\nintegrator pathtracer\nbrdf lambertian\ndiffuse 1 0 0\n"
            "camera 0 0 5 0 0 0 0 1 0 45\nbogus 1 2 3\nsphere 0 0 0 1\npop_transform\nsphere 1 1 1 1\n")
    h = HostScene.from_text(text)
    o = oracle.load_text(text)
    assert (h.params.width, h.params.height) == (320, 200)
    assert h.arrays()["materials"][0, 14] == 1  # Phong
    _compare(h, o)


def test_scene_output_name():
    """The `output` command sets the scene's output file (muon/parser.cc:309-318);
    without it the default is raytrace.png (muon/defaults.h); a bare `output`
    line is a malformed line (logged, output unchanged)."""
    from muon_amd import HostScene
    assert HostScene.from_text("film_size 4 4\n").output == "raytrace.png"
    assert HostScene.from_text("output first.png\noutput second.png\n").output == "second.png"
    hs = HostScene.from_text("output keep.png\noutput\n")
    assert hs.output == "keep.png" and hs.warnings == 1


def test_device_bvh_parse_skips_the_host_tree(ref_dir):
    """device_bvh=True parses the same scene (primitives, transforms,
    materials, lights, render settings) with no host tree; the linear
    structure is still built on the host (one leaf)."""
    import os
    path = os.path.join(ref_dir, "cornell_phong.muon")
    a = HostScene.from_file(path, default_seed=1)
    b = HostScene.from_file(path, default_seed=1, device_bvh=True)
    assert b.desc.num_nodes == 0 and b.desc.num_prims == a.desc.num_prims > 0
    aa, bb = a.arrays(), b.arrays()
    for k in ("prim_type", "prim_geom", "prim_transform", "prim_material", "prim_light", "transforms", "materials",
              "lights"):
        assert np.array_equal(aa[k], bb[k]), k
    assert bytes(a.params) == bytes(b.params)
    c = HostScene.from_file(path, acceleration=0, device_bvh=True)
    assert c.desc.num_nodes == 1
