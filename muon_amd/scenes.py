"""Seeded synthetic workloads, emitted as .muon scene text so that the GPU
path and the CPU restatement parse identical inputs.

  cornell_text()          Cornell box: 5 wall quads, two boxes, one 0.25 quad light
                          (256x256, 16 spp, depth 5, BRDF IS + NEE + MIS, seed 0)
  icosphere_text()        displaced level-7 icosphere (327,680 tris), GGX roughness
                          0.3, inside the Cornell enclosure (512x512, 64 spp, depth 8)
  sphere_field_text()     10,000 spheres (Lambert / Phong / GGX) + 8 spherical
                          emitters (1024x1024, 32 spp, depth 8, NEE off)
  microbench_rays()       2^24 rays, origins uniform in a box, directions uniform
                          on the sphere

Floats are written with 9 significant digits, which round-trips float32.
"""
import functools
import math

import numpy as np

F = "%.9g"


def _fmt(*xs):
    return " ".join(F % float(np.float32(x)) for x in xs)


# ---------------------------------------------------------------- Cornell box
WALL_GREY = (0.73, 0.73, 0.73)
WALL_RED = (0.65, 0.05, 0.05)
WALL_GREEN = (0.12, 0.45, 0.15)


def _box_vertices(cx, cz, sx, sz, h, angle_deg):
    """8 corners of a y-aligned box rotated about y, bottom at y=0."""
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)
    out = []
    for y in (0.0, h):
        for dx, dz in ((-sx, -sz), (sx, -sz), (sx, sz), (-sx, sz)):
            x = cx + ca * dx + sa * dz
            z = cz - sa * dx + ca * dz
            out.append((x, y, z))
    return out


# outward-facing triangles of a box with vertices 0-3 bottom, 4-7 top
_BOX_TRIS = [(0, 1, 2), (0, 2, 3), (4, 6, 5), (4, 7, 6), (0, 4, 5), (0, 5, 1),
             (1, 5, 6), (1, 6, 2), (2, 6, 7), (2, 7, 3), (3, 7, 4), (3, 4, 0)]


def cornell_walls(lines):
    """The 5 walls of the enclosure x in [-1,1], y in [0,2], z in [-1,1]."""
    walls = [
        (WALL_GREY, [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)], [(0, 2, 1), (0, 3, 2)]),    # floor
        (WALL_GREY, [(-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)], [(0, 1, 2), (0, 2, 3)]),    # ceiling
        (WALL_GREY, [(-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)], [(0, 1, 2), (0, 2, 3)]),  # back
        (WALL_RED, [(-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)], [(0, 1, 2), (0, 2, 3)]),   # left
        (WALL_GREEN, [(1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)], [(0, 1, 2), (0, 2, 3)]),     # right
    ]
    for color, verts, tris in walls:
        lines.append("diffuse " + _fmt(*color))
        lines.append("start_mesh")
        for v in verts:
            lines.append("vertex " + _fmt(*v))
        for t in tris:
            lines.append("tri %d %d %d" % t)
        lines.append("end_mesh")


def cornell_header(lines, width, height, spp, max_depth, seed=0, nee="mis"):
    lines += [
        f"random_seed {seed}",
        f"film_size {width} {height}",
        "integrator pathtracer",
        "importance_sampling brdf",
        f"next_event_estimation {nee}",
        "russian_roulette off",
        f"max_depth {max_depth}",
        f"pixel_samples {spp}",
        "camera 0 1 6.8  0 1 0  0 1 0  19.5",
        "quad_light -0.125 1.98 0.125  0.25 0 0  0 0 -0.25  15 15 15",
        "brdf lambertian",
        "ambient 0 0 0",
        "specular 0 0 0",
    ]


def cornell_text(width=256, height=256, spp=16, max_depth=5, seed=0):
    lines = ["# Cornell box: 5 walls, two boxes, one quad light"]
    cornell_header(lines, width, height, spp, max_depth, seed)
    cornell_walls(lines)
    lines.append("diffuse " + _fmt(*WALL_GREY))
    for box in (_box_vertices(0.33, 0.37, 0.3, 0.3, 0.6, -17.0), _box_vertices(-0.33, -0.3, 0.3, 0.3, 1.2, 17.0)):
        lines.append("start_mesh")
        for v in box:
            lines.append("vertex " + _fmt(*v))
        for t in _BOX_TRIS:
            lines.append("tri %d %d %d" % t)
        lines.append("end_mesh")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- icosphere
@functools.lru_cache(maxsize=8)
def icosphere(level):
    """Unit icosphere: (vertices [V,3] float64, faces [F,3] int64); F = 20*4^level."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // (len(v) + 1), uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)  # midpoints of edges (01, 12, 20)
        v = np.concatenate([v, mid])
        f = np.concatenate([
            np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), np.stack([m[:, 0], m[:, 1], m[:, 2]], 1)])
    return v, f


def value_noise(p, seed, octaves=3, base_freq=4.0):
    """Seeded 3-octave value noise in [-1,1]: trilinear interpolation
    (smoothstep-weighted) of per-lattice-point random values; octave
    frequencies base*(1,2,4), amplitudes (1,1/2,1/4), normalised."""
    total = np.zeros(len(p))
    norm = 0.0
    for o in range(octaves):
        freq, amp = base_freq * 2 ** o, 0.5 ** o
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ix, iy, iz = i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz
                    h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761 + o * 97)
                    h = (h ^ (h >> 13)) * 1274126177
                    h = (h ^ (h >> 16)) & 0xFFFFFF
                    val = h / float(0xFFFFFF) * 2.0 - 1.0
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * val
        total += amp * acc
        norm += amp
    return total / norm


def displaced_icosphere(level=7, radius=50.0, amplitude=0.05, seed=0):
    v, f = icosphere(level)
    n = value_noise(v, seed)
    verts = v * (radius * (1.0 + amplitude * n))[:, None]
    return verts.astype(np.float32), f


def _mesh_lines(lines, verts, faces):
    lines.append("start_mesh")
    lines.extend("vertex " + s for s in _fmt_rows(verts))
    lines.extend("tri %d %d %d" % (a, b, c) for a, b, c in faces.tolist())
    lines.append("end_mesh")


def _fmt_rows(arr):
    arr = np.asarray(arr, np.float32)
    return [" ".join(F % x for x in row) for row in arr.astype(np.float64).tolist()]


def icosphere_text(width=512, height=512, spp=64, max_depth=8, level=7, seed=0, radius=50.0, render_seed=None,
                   mesh_file=None):
    """Displaced icosphere emitted at object radius 50 and placed with
    `translate 0 0.6 0; scale 0.008`: at world scale every triangle's
    2*area would be below the reference's 1e-4 parallel threshold
    (muon/objects.cc:104-108) and the mesh would be invisible. `seed` seeds
    the displacement and, unless `render_seed` is given, the render. With
    `mesh_file` the mesh is written there as binary PLY and the scene loads
    it (`load`) instead of listing its vertices and triangles."""
    verts, faces = displaced_icosphere(level, radius, 0.05, seed)
    lines = [f"# displaced icosphere, level {level}, {len(faces)} tris"]
    cornell_header(lines, width, height, spp, max_depth, seed if render_seed is None else render_seed)
    cornell_walls(lines)
    lines += ["push_transform", "translate 0 0.6 0", "scale 0.008 0.008 0.008", "brdf ggx", "roughness 0.3",
              "diffuse 0.1 0.1 0.1", "specular 0.6 0.6 0.6"]
    if mesh_file is None:
        _mesh_lines(lines, verts, faces)
    else:
        write_ply(mesh_file, verts, faces)
        lines.append(f"load {mesh_file}")
    lines.append("pop_transform")
    return "\n".join(lines) + "\n"


def write_ply(path, verts, faces):
    """Binary little-endian PLY: float x y z per vertex, a uchar-counted int
    list per face."""
    verts = np.ascontiguousarray(verts, "<f4")
    faces = np.asarray(faces, np.int64)
    head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(verts)}\nproperty float x\n"
            f"property float y\nproperty float z\nelement face {len(faces)}\n"
            "property list uchar int vertex_indices\nend_header\n").encode()
    rec = np.zeros(len(faces), dtype=[("n", "u1"), ("i", "<i4", (3,))])
    rec["n"] = 3
    rec["i"] = faces
    with open(path, "wb") as f:
        f.write(head)
        f.write(verts.tobytes())
        f.write(rec.tobytes())


def icosphere_world_bounds(level=7, seed=0, radius=50.0):
    verts, _ = displaced_icosphere(level, radius, 0.05, seed)
    w = verts.astype(np.float64) * 0.008 + np.array([0.0, 0.6, 0.0])
    return w.min(axis=0).astype(np.float32), w.max(axis=0).astype(np.float32)


# ---------------------------------------------------------------- sphere field
def sphere_field_extent(n, scaled):
    """Half-size of the sphere field's box and its number of emitters: the
    configuration's 10 and 8; a density-scaled field of more than 10,000
    spheres grows both with the count, so that spheres and emitters keep the
    10,000-sphere field's density (its emitters are not buried)."""
    if not scaled or n <= 10000:
        return 10.0, 8
    return 10.0 * (n / 10000.0) ** (1.0 / 3.0), int(round(8 * n / 10000.0))


def sphere_field_draw(n, seed, half=10.0):
    """The sphere field's seeded draws: centres, radii, material kinds
    (0 Lambert, 1 Phong, 2 GGX), colours, Phong exponents, GGX roughness."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-half, half, (n, 3))
    r = np.exp(rng.uniform(math.log(0.05), math.log(0.5), n))
    kind = rng.integers(0, 3, n)
    col = rng.uniform(0.2, 0.8, (n, 3))
    expo = np.exp(rng.uniform(math.log(10), math.log(1000), n))
    rough = rng.uniform(0.05, 0.8, n)
    return rng, c, r, kind, col, expo, rough


def sphere_field_prologue(width, height, spp, max_depth, n, seed, half=10.0, emitters=8):
    cam = "camera 0 0 32  0 0 0  0 1 0  45" if half == 10.0 else f"camera 0 0 {_fmt(3.2 * half)}  0 0 0  0 1 0  45"
    return [
        f"# sphere field: {n} spheres + {emitters} emitters",
        f"random_seed {seed}", f"film_size {width} {height}", "integrator pathtracer",
        "importance_sampling brdf", "next_event_estimation off", "russian_roulette off",
        f"max_depth {max_depth}", f"pixel_samples {spp}", cam,
        "ambient 0 0 0", "emission 0 0 0",
    ]


def sphere_field_emitters(rng, half=10.0, count=8):
    """The emitting spheres (radius 0.5), drawn after the field's spheres."""
    lines = ["brdf lambertian", "diffuse 0 0 0", "specular 0 0 0"]
    ec = rng.uniform(-0.9 * half, 0.9 * half, (count, 3))
    ecol = rng.uniform(0.5, 1.0, (count, 3))
    for i in range(count):
        lines.append("emission " + _fmt(*(20 * ecol[i])))
        lines.append("sphere " + _fmt(ec[i, 0], ec[i, 1], ec[i, 2], 0.5))
    return lines


def sphere_field_text(width=1024, height=1024, spp=32, max_depth=8, n=10000, seed=0, scaled=False):
    """The sphere field: n spheres uniform in a box of half-size 10, radii
    log-uniform in [0.05, 0.5], a third each Lambert / Phong / GGX, lit by 8
    emitting spheres (NEE off: the reference cannot NEE-sample sphere
    emitters). `scaled`: box and emitter count grow with n beyond 10,000
    (sphere_field_extent)."""
    half, ne = sphere_field_extent(n, scaled)
    rng, c, r, kind, col, expo, rough = sphere_field_draw(n, seed, half)
    lines = sphere_field_prologue(width, height, spp, max_depth, n, seed, half, ne)
    for i in range(n):
        if kind[i] == 0:
            lines += ["brdf lambertian", "specular 0 0 0", "diffuse " + _fmt(*col[i])]
        elif kind[i] == 1:
            lines += ["brdf phong", "diffuse " + _fmt(*(0.4 * col[i])), "specular 0.4 0.4 0.4",
                      "shininess " + _fmt(expo[i])]
        else:
            lines += ["brdf ggx", "diffuse " + _fmt(*(0.1 * col[i])), "specular 0.6 0.6 0.6",
                      "roughness " + _fmt(rough[i])]
        lines.append("sphere " + _fmt(c[i, 0], c[i, 1], c[i, 2], r[i]))
    lines += sphere_field_emitters(rng, half, ne)
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- microbench rays
def microbench_rays(n, lo, hi, seed=0):
    """n rays: origins uniform in [lo, hi], directions uniform on the unit
    sphere (z = 1 - 2u, phi = 2 pi v), float32 [n,3] arrays."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    o = (lo + rng.random((n, 3), dtype=np.float32) * (hi - lo)).astype(np.float32)
    u = rng.random(n, dtype=np.float32)
    v = rng.random(n, dtype=np.float32)
    z = 1.0 - 2.0 * u
    rr = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    d = np.stack([rr * np.cos(phi), rr * np.sin(phi), z], axis=1).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    return o, d.astype(np.float32)


# ---------------------------------------------------------------- cameras and request mixes
def set_camera(params, eye, look, up=(0.0, 1.0, 0.0), fov_deg=None, width=None, height=None):
    """Sets the pinhole camera of a muonx_render_params (muon/camera.cc:15-35:
    w = norm(eye - look), u = norm(up x w), v = w x u, tan(fov/2) on y and
    scaled by W/H on x). fov_deg None keeps the params' vertical field of view."""
    if width is not None:
        params.width = int(width)
    if height is not None:
        params.height = int(height)
    f32 = np.float32
    e, l, u = (np.asarray(x, f32) for x in (eye, look, up))
    w = e - l
    w = w / f32(np.sqrt(np.dot(w, w)))
    cu = np.cross(u, w).astype(f32)
    cu = cu / f32(np.sqrt(np.dot(cu, cu)))
    cv = np.cross(w, cu).astype(f32)
    tfy = f32(params.tan_fov_y) if fov_deg is None else f32(math.tan(math.radians(fov_deg) / 2.0))
    params.tan_fov_y = float(tfy)
    params.tan_fov_x = float(tfy * (f32(params.width) / f32(params.height)))
    for i in range(3):
        params.eye[i] = float(e[i])
        params.cam_u[i] = float(cu[i])
        params.cam_v[i] = float(cv[i])
        params.cam_w[i] = float(w[i])
    params.has_camera = 1
    return params


# Resident scenes for the batched workload: displaced icospheres of levels
# 3..8 (1,280 .. 1,310,720 tris) in the Cornell enclosure and sphere fields,
# spanning ~1k to ~1.3M primitives (sizes spread log-uniformly).
POOL = [("ico", 3), ("ico", 4), ("ico", 5), ("ico", 6), ("ico", 7), ("ico", 8), ("spheres", 3000),
        ("spheres", 30000)]

# Resident scenes for the served workload: 8 scenes with sizes in the same
# range, about 5.7M primitives in all (four level-8 icospheres with different
# displacement seeds); the third entry is the seed.
SERVE_POOL = [("ico", 5, 0), ("ico", 6, 0), ("ico", 7, 0), ("ico", 8, 0), ("ico", 8, 1), ("ico", 8, 2),
              ("ico", 8, 3), ("spheres", 30000, 0)]


def pool_scene_text(kind, size, seed=0):
    if kind == "ico":
        return icosphere_text(256, 256, 16, 8, level=size, seed=seed)
    return sphere_field_text(256, 256, 16, 8, n=size, seed=seed)


def random_pose(params, kind, rng, width, height, field_half=10.0):
    """A seeded camera pose: the eye on a sphere around the scene's centre,
    looking at it (inside the enclosure, radius 0.9 around the icosphere;
    radius 3.2 x the box half-size, 32 for the configuration's field, around
    the sphere field), elevation within +-35 degrees."""
    az = rng.uniform(0.0, 2.0 * math.pi)
    el = math.radians(rng.uniform(-35.0, 35.0))
    dirv = (math.sin(az) * math.cos(el), math.sin(el), math.cos(az) * math.cos(el))
    if kind == "ico":
        c, rad, fov = (0.0, 0.6, 0.0), 0.9, 45.0
    else:
        c, rad, fov = (0.0, 0.0, 0.0), 3.2 * field_half, 45.0
    eye = tuple(c[i] + rad * dirv[i] for i in range(3))
    return set_camera(params, eye, c, fov_deg=fov, width=width, height=height)
