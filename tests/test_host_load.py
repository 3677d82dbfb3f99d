"""The .muon `load` command (muon/parser.cc:465-561): OBJ and PLY meshes are
added with the current transform and material, exactly as the same
triangles written as vertex / tri commands; the three PLY encodings read
identically; loaded vertex normals are the unnormalised sums of the face
normals of the triangles that share the vertex (loaded vertices live outside
the meshes that muon/parser.cc:862-871 normalises); a file that cannot be
read is logged (a warning) and skipped. assimp itself is not available, so
its face reordering is not reproduced; these tests pin everything else
against the expanded scene and the oracle's parse of it."""
import struct

import numpy as np
import pytest

from muon_amd import HostScene

HEAD = ("film_size 32 32\ncamera 0 0 6 0 0 0 0 1 0 45\ndiffuse 0.3 0.4 0.5\n"
        "translate 0.5 -0.25 0\nrotate 0 1 0 30\nscale 1.5 1.5 1.5\n")
TAIL = "diffuse 0.9 0.1 0.1\nsphere 0 0 -3 0.5\n"

# a cube (6 quads) written with every OBJ corner form, a pentagon and a
# degenerate triangle (two equal corners: dropped)
OBJ = """# cube
v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
f 1 4 3 2
f 5/1 6/2 7/3 8/1
f 1//1 2//1 6//1 5//1
f -5/1/1 -6/2/1 -2/3/1 -1/1/1
f 2 3 7 6
f 1 5 8 4
f 1 2 2
v 0 0 2
v 1 0 2.5
v 1.5 1 2.5
v 0.5 1.5 2
v -0.5 1 2
f 9 10 11 12 13
"""


def _obj_tris():
    """The triangles the loader must produce from OBJ, in file order."""
    v = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1),
         (0, 0, 2), (1, 0, 2.5), (1.5, 1, 2.5), (0.5, 1.5, 2), (-0.5, 1, 2)]
    faces = [[1, 4, 3, 2], [5, 6, 7, 8], [1, 2, 6, 5], [4, 3, 7, 8], [2, 3, 7, 6], [1, 5, 8, 4], [9, 10, 11, 12, 13]]
    tris = []
    for f in faces:
        for k in range(1, len(f) - 1):
            tris.append((f[0] - 1, f[k] - 1, f[k + 1] - 1))
    return [tuple(map(float, p)) for p in v], tris


def _expanded(v, tris):
    lines = ["vertex %r %r %r" % p for p in v] + ["tri %d %d %d" % t for t in tris]
    return "\n".join(lines) + "\n"


def _same_arrays(a, b):
    ka, kb = a.arrays(), b.arrays()
    for k in ka:
        assert np.array_equal(np.asarray(ka[k]).view(np.uint8), np.asarray(kb[k]).view(np.uint8)), k


def test_obj_equals_expanded(tmp_path, oracle):
    (tmp_path / "mesh.obj").write_text(OBJ)
    (tmp_path / "scene.muon").write_text(HEAD + "load mesh.obj\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "scene.muon"))
    v, tris = _obj_tris()
    text = HEAD + _expanded(v, tris) + TAIL
    ref = HostScene.from_text(text)
    assert hs.desc.num_prims == len(tris) + 1 == 16
    assert hs.warnings == 0
    _same_arrays(hs, ref)
    # and the expanded scene is what the oracle parses
    orc = oracle.load_text(text)
    a, op = hs.arrays(), orc.prims()
    assert np.array_equal(a["prim_geom"].view(np.uint32), op["geom"].view(np.uint32))
    assert np.array_equal(a["transforms"][a["prim_transform"]].view(np.uint32), op["xform"].view(np.uint32))
    assert np.array_equal(a["leaf_prims"], orc.bvh()["prim_order"])


def _ply(fmt):
    """A mesh with extra vertex properties, an extra element and a quad, a
    triangle, a degenerate face and a pentagon."""
    verts = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1), (2, 0, 0), (3, 0, 0.5), (3.5, 1, 0.5),
             (2.5, 1.5, 0), (1.5, 1, 0)]
    faces = [[0, 1, 2, 3], [0, 1, 4], [1, 1, 2], [5, 6, 7, 8, 9]]
    head = ["ply", "format %s 1.0" % fmt, "comment made by a test", "element vertex %d" % len(verts),
            "property float x", "property float y", "property float z", "property float nx",
            "property uchar red", "element face %d" % len(faces), "property list uchar int vertex_indices",
            "property uchar flags", "element edge 2", "property int vertex1", "property int vertex2", "end_header"]
    data = ("\n".join(head) + "\n").encode()
    if fmt == "ascii":
        body = ["%r %r %r 0.5 %d" % (p + (i,)) for i, p in enumerate(verts)]
        body += ["%d %s 7" % (len(f), " ".join(map(str, f))) for f in faces]
        body += ["0 1", "1 2"]
        return data + ("\n".join(body) + "\n").encode()
    e = "<" if fmt == "binary_little_endian" else ">"
    for i, p in enumerate(verts):
        data += struct.pack(e + "ffffB", *p, 0.5, i)
    for f in faces:
        data += struct.pack(e + "B%diB" % len(f), len(f), *f, 7)
    data += struct.pack(e + "iiii", 0, 1, 1, 2)
    return data, verts, faces


def _ply_tris():
    data, verts, faces = _ply("binary_little_endian")
    tris = []
    for f in faces:
        for k in range(1, len(f) - 1):
            t = (f[0], f[k], f[k + 1])
            if len({verts[j] for j in t}) == 3:
                tris.append(t)
    return [tuple(map(float, p)) for p in verts], tris


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_equals_expanded(tmp_path, fmt):
    blob = _ply(fmt)
    (tmp_path / "m.PLY").write_bytes(blob if fmt == "ascii" else blob[0])
    (tmp_path / "scene.muon").write_text(HEAD + "load m.PLY\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "scene.muon"))
    v, tris = _ply_tris()
    ref = HostScene.from_text(HEAD + _expanded(v, tris) + TAIL)
    assert hs.warnings == 0
    assert hs.desc.num_prims == len(tris) + 1 == 7
    _same_arrays(hs, ref)


def test_loaded_vertex_normals_stay_unnormalised(tmp_path):
    """compute_vertex_normals on: a loaded vertex carries the float32 sum of
    cross(v1 - v0, v2 - v0) over the triangles that share it, in triangle
    order, and is not normalised; corners of the same position with
    different texture coordinates are different vertices."""
    (tmp_path / "mesh.obj").write_text(OBJ)
    (tmp_path / "scene.muon").write_text("compute_vertex_normals on\n" + HEAD + "load mesh.obj\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "scene.muon"))
    a = hs.arrays()
    n = hs.desc.num_prims - 1
    assert (a["prim_use_vnormal"][:n] == 1).all()
    # rebuild the joined vertices: key = (position, texcoord or none, normal or none)
    vt = [(0, 0), (1, 0), (1, 1)]
    corners = [[(1, None, None), (4, None, None), (3, None, None), (2, None, None)],
               [(5, 1, None), (6, 2, None), (7, 3, None), (8, 1, None)],
               [(1, None, 1), (2, None, 1), (6, None, 1), (5, None, 1)],
               [(4, 1, 1), (3, 2, 1), (7, 3, 1), (8, 1, 1)],
               [(2, None, None), (3, None, None), (7, None, None), (6, None, None)],
               [(1, None, None), (5, None, None), (8, None, None), (4, None, None)],
               [(9, None, None), (10, None, None), (11, None, None), (12, None, None), (13, None, None)]]
    pos, _ = _obj_tris()
    keys, tri_keys = {}, []
    for f in corners:
        ids = []
        for c in f:
            k = (pos[c[0] - 1], None if c[1] is None else vt[c[1] - 1], c[2])
            ids.append(keys.setdefault(k, len(keys)))
        for j in range(1, len(f) - 1):
            tri_keys.append((ids[0], ids[j], ids[j + 1]))
    P = np.zeros((len(keys), 3), np.float32)
    for k, i in keys.items():
        P[i] = k[0]
    N = np.zeros_like(P)
    for t in tri_keys:
        e1 = P[t[1]] - P[t[0]]
        e2 = P[t[2]] - P[t[0]]
        c = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]],
                     np.float32)
        for j in t:
            N[j] = N[j] + c
    want = np.stack([np.concatenate([N[t[0]], N[t[1]], N[t[2]]]) for t in tri_keys])
    assert np.array_equal(a["prim_vnormal"][:n].view(np.uint32), want.view(np.uint32))
    # not unit length (the expanded scene's mesh vertices would be)
    lengths = np.linalg.norm(want.reshape(-1, 3), axis=1)
    assert np.abs(lengths - 1).max() > 0.5


def test_load_failures_warn_and_skip(tmp_path, monkeypatch):
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")  # index out of range
    (tmp_path / "mesh.xyz").write_text("0 0 0\n")
    text = HEAD + "load missing.obj\nload bad.obj\nload mesh.xyz\nload\n" + TAIL
    monkeypatch.chdir(tmp_path)
    hs = HostScene.from_text(text)
    assert hs.warnings == 4
    assert hs.desc.num_prims == 1  # only the sphere


def test_text_scenes_load_relative_to_working_directory(tmp_path, monkeypatch):
    (tmp_path / "mesh.obj").write_text(OBJ)
    monkeypatch.chdir(tmp_path)
    hs = HostScene.from_text(HEAD + "load mesh.obj\n" + TAIL)
    assert hs.warnings == 0 and hs.desc.num_prims == 16


def test_obj_groups_are_separate_meshes(tmp_path):
    """A position shared by two groups is two vertices: each accumulates the
    normals of its own group's triangles only."""
    obj = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\ng a\nf 1 2 3\ng b\nf 1 3 4\n"
    (tmp_path / "g.obj").write_text(obj)
    (tmp_path / "s.muon").write_text("compute_vertex_normals on\nload g.obj\n")
    a = HostScene.from_file(str(tmp_path / "s.muon")).arrays()
    vn = a["prim_vnormal"].reshape(2, 3, 3)
    assert np.array_equal(vn[0, 0], np.float32([0, 0, 1]))   # group a: only its face
    assert np.array_equal(vn[1, 0], np.float32([1, 0, 0]))   # group b: only its face


def test_obj_face_forms_read_alike(tmp_path):
    """Face lines taken by the fast reader and by the stream reader (an empty
    texture index after a slash, a 10-digit index) give the same triangles."""
    obj = ("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.5\n"
           "f 1 2 3\nf 1/ 2/ 3/\nf 0000000001 2 0000000003\nf -3/1 -2/1 -1/1\n")
    (tmp_path / "m.obj").write_text(obj)
    (tmp_path / "s.muon").write_text("load m.obj\n")
    hs = HostScene.from_file(str(tmp_path / "s.muon"))
    assert hs.warnings == 0 and hs.desc.num_prims == 4
    g = hs.arrays()["prim_geom"][:, :9]
    assert all(np.array_equal(g[0], g[k]) for k in range(1, 4))


def test_ply_double_coordinates_and_wide_list_counts(tmp_path):
    """Binary PLY with double coordinates, a ushort list count, uint indices
    and a trailing element: read as the same triangles as inline."""
    verts = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.5), (0.0, 1.0, 0.25)]
    head = ["ply", "format binary_big_endian 1.0", "element vertex 4", "property double x", "property double y",
            "property double z", "element face 1", "property list ushort uint vertex_index", "element extra 1",
            "property list uchar float values", "end_header"]
    data = ("\n".join(head) + "\n").encode()
    for p in verts:
        data += struct.pack(">ddd", *p)
    data += struct.pack(">HIIII", 4, 0, 1, 2, 3)
    data += struct.pack(">Bff", 2, 1.0, 2.0)
    (tmp_path / "q.ply").write_bytes(data)
    (tmp_path / "s.muon").write_text(HEAD + "load q.ply\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "s.muon"))
    ref = HostScene.from_text(HEAD + _expanded(verts, [(0, 1, 2), (0, 2, 3)]) + TAIL)
    assert hs.warnings == 0
    _same_arrays(hs, ref)


def test_ply_truncated_body_is_a_warning(tmp_path):
    data = b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n" \
           b"property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n" + \
           struct.pack("<fff", 0, 0, 0)
    (tmp_path / "t.ply").write_bytes(data)
    (tmp_path / "s.muon").write_text(HEAD + "load t.ply\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "s.muon"))
    assert hs.warnings == 1 and hs.desc.num_prims == 1


PLY_VERTS = "element vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
PLY_V_ASCII = "0 0 0\n1 0 0\n0 1 0\n"


@pytest.mark.parametrize("header,body", [
    # float list counts are refused at the header
    ("element face 1\nproperty list float int vertex_indices\n", "3 0 1 2\n"),
    # non-integral, non-finite, negative and huge face indices
    ("element face 1\nproperty list uchar float vertex_indices\n", "3 0 1 1.5\n"),
    ("element face 1\nproperty list uchar double vertex_indices\n", "3 0 1 nan\n"),
    ("element face 1\nproperty list uchar double vertex_indices\n", "3 0 1 -1\n"),
    ("element face 1\nproperty list uchar double vertex_indices\n", "3 0 1 1e30\n"),
    # a list count beyond the body
    ("element face 1\nproperty list uint int vertex_indices\n", "4000000000 0 1 2\n"),
    # an element declaring more rows than the body can hold
    ("element face 99999999999\nproperty list uchar int vertex_indices\n", "3 0 1 2\n"),
], ids=["float-count", "fractional-index", "nan-index", "negative-index", "huge-index", "huge-count", "huge-rows"])
def test_ply_bad_counts_and_indices_are_warnings(tmp_path, header, body):
    """Malformed counts and indices reject the file (a warning; the scene's
    other primitives stay) instead of being cast unchecked."""
    data = "ply\nformat ascii 1.0\n" + PLY_VERTS + header + "end_header\n" + PLY_V_ASCII + body
    (tmp_path / "t.ply").write_text(data)
    (tmp_path / "s.muon").write_text(HEAD + "load t.ply\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "s.muon"))
    assert hs.warnings == 1 and hs.desc.num_prims == 1


def test_ply_property_less_element_with_huge_count_is_skipped(tmp_path):
    """An element without properties reads nothing whatever its declared
    count (it used to loop over the count)."""
    data = ("ply\nformat ascii 1.0\nelement junk 18446744073709551615\n" + PLY_VERTS +
            "element face 1\nproperty list uchar int vertex_indices\nend_header\n" + PLY_V_ASCII + "3 0 1 2\n")
    (tmp_path / "t.ply").write_text(data)
    (tmp_path / "s.muon").write_text(HEAD + "load t.ply\n" + TAIL)
    hs = HostScene.from_file(str(tmp_path / "s.muon"))
    assert hs.warnings == 0 and hs.desc.num_prims == 2


def test_icosphere_from_binary_ply_equals_inline(tmp_path):
    """The benchmark's icospheres loaded from binary PLY (scenes.write_ply)
    parse to the same primitives, transforms and materials as the inline
    mesh text."""
    from muon_amd import scenes
    inline = HostScene.from_text(scenes.icosphere_text(64, 64, 4, 8, level=4, seed=9))
    path = str(tmp_path / "ico.ply")
    loaded = HostScene.from_text(scenes.icosphere_text(64, 64, 4, 8, level=4, seed=9, mesh_file=path),
                                 device_bvh=True)
    # (one warning each: `random_seed` falls through into `film_size`, muon/parser.cc:267-276)
    assert loaded.warnings == inline.warnings and loaded.desc.num_nodes == 0
    a, b = inline.arrays(), loaded.arrays()
    for k in ("prim_type", "prim_geom", "prim_transform", "prim_material", "prim_light", "transforms", "materials",
              "lights"):
        assert np.array_equal(a[k], b[k]), k


def test_sphere_field_arrays_equal_the_text_scene():
    """The benchmark's sphere fields handed over as arrays hold the spheres,
    transforms and material records the parser makes of the same field as
    text (the arrays put the 8 emitters first)."""
    from muon_amd import scenes, workloads
    for n, ne in ((700, 8), (25000, 20)):  # below and above the density scaling
        text = HostScene.from_text(scenes.sphere_field_text(64, 64, 4, 8, n=n, seed=4, scaled=True)).arrays()
        hs, f = workloads.sphere_field_arrays(n, 4, 64, 64, 4, 8)
        assert f["num_prims"] == n + ne and hs.desc.num_prims == ne
        geom = f["prim_geom"].reshape(-1, 12)
        mats = f["materials"].reshape(-1, 16)
        assert np.array_equal(geom[ne:], text["prim_geom"][:n])
        assert np.array_equal(geom[:ne], text["prim_geom"][n:])
        tx = text["transforms"]
        assert np.array_equal(hs.arrays()["transforms"][f["prim_transform"][ne:]], tx[text["prim_transform"][:n]])
        assert np.array_equal(mats[f["prim_material"][ne:]], text["materials"][text["prim_material"][:n]])
    assert (f["prim_light"] == -1).all() and (f["prim_type"] == 1).all()
