"""Seeded procedural batch workloads for the renderer.

Every generator returns a :class:`BatchSpec`: one FileData holding all mesh
templates, materials and textures (added once, like the reference's
``addFile()``), plus per environment a list of instances
(template name, bake transform, semantic id) and a camera.  ``populate()``
replays a spec through any renderer with the reference's method names.

Seeds: config ``i`` uses ``0x5EED0000 + i``; environment ``e`` draws from a
PCG64 stream seeded with SplitMix64(seed ^ e).  Everything is float32 and
reproducible across machines.

* ``small_rooms`` (16 envs): an 8 x 6 x 3 m box room (12 inward triangles),
  20 cuboids of uniform random size 0.2-1.5 m on the floor, one
  subdivision-3 icosphere (1280 triangles); per-vertex RGB uniform random
  uint8; semantic ids 1..22; one 128x128 pinhole camera per env (HFOV 90,
  height 1.5 m, position uniform in the room, yaw uniform, pitch +-15 deg).
* ``apartments`` (512 envs): 3 x 3 grid of 4 x 4 m rooms with 1 m doorways,
  300 instances of 20 shared template meshes (noisy-displaced UV spheres and
  cuboids, 500-5000 triangles, template picks weighted to ~200k triangles per
  scene), each with a 256x256 RGBA8 uniform-noise texture layer with a full
  mip chain (trilinear), FP32 positions, FP16 texture coordinates.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import math3d as m3
from .assets import (FileData, MaterialData, MeshData, MeshEntry, NodeData, TextureData,
                     GL_LINEAR, GL_LINEAR_MIPMAP_LINEAR, GL_REPEAT, generate_mipmaps)

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def env_rng(config_index: int, env: int) -> np.random.Generator:
    seed = 0x5EED0000 + config_index
    return np.random.Generator(np.random.PCG64(splitmix64(seed ^ env)))


@dataclass
class Instance:
    template: str
    bake: np.ndarray
    semantic: int


@dataclass
class EnvSpec:
    instances: List[Instance]
    camera_pose: np.ndarray      # camera-to-world
    projection: np.ndarray
    viewport: Optional[Tuple[int, int]] = None  # (w, h) inside the tile; None = the tile
    planes: int = 0              # sensor set (PLANE_* bits; 0 = every plane)
    # the instances as arrays (template names, bakes (n, 4, 4), semantic ids),
    # for adding them in one call where the renderer offers it
    arrays: Optional[Tuple[list, np.ndarray, np.ndarray]] = None


@dataclass
class BatchSpec:
    name: str
    tile: Tuple[int, int]
    file: FileData
    envs: List[EnvSpec]
    template_tris: dict = field(default_factory=dict)

    @property
    def grid(self) -> Tuple[int, int]:
        n = len(self.envs)
        x = int(math.ceil(math.sqrt(np.float32(n))))
        return x, (n + x - 1) // x

    def triangles(self, env: int) -> int:
        return sum(self.template_tris[i.template] for i in self.envs[env].instances)


def populate(renderer, spec: BatchSpec, first_scene: int = 0, add_file: bool = True) -> None:
    """Adds the spec's file and instances to a renderer exposing the
    reference's method names (product or oracle)."""
    if add_file:
        assert renderer.add_file_data(spec.file), "add_file_data failed"
    batched = hasattr(renderer, "add_node_hierarchies")
    for e, env in enumerate(spec.envs):
        s = first_scene + e
        if batched and env.arrays is not None:
            renderer.add_node_hierarchies(s, *env.arrays)
        else:
            for inst in env.instances:
                renderer.add_node_hierarchy(s, inst.template, inst.bake, inst.semantic)
        renderer.update_camera(s, env.projection, m3.inverted_rigid(env.camera_pose))
        if env.viewport is not None:
            renderer.set_scene_viewport(s, env.viewport)
        if env.planes and hasattr(renderer, "set_scene_planes"):
            renderer.set_scene_planes(s, env.planes)


# ---------------------------------------------------------------------------
# Mesh builders (CCW front faces seen from outside, or inside for rooms)
# ---------------------------------------------------------------------------

def _check_winding(pos: np.ndarray, idx: np.ndarray, outward_from: np.ndarray, inward=False):
    t = idx.reshape(-1, 3)
    a, b, c = pos[t[:, 0]], pos[t[:, 1]], pos[t[:, 2]]
    n = np.cross(b - a, c - a)
    d = ((a + b + c) / 3.0) - outward_from
    s = np.einsum("ij,ij->i", n, d)
    assert (s < 0).all() if inward else (s > 0).all(), "winding"


def box_mesh(lo, hi, inward: bool) -> Tuple[np.ndarray, np.ndarray]:
    """8-vertex box, 12 triangles, faces pointing out (or in)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    p = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1],
                   lo[2] if i & 4 == 0 else hi[2]] for i in range(8)])
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    idx = []
    for a, b, c, d in quads:
        idx += [a, b, c, a, c, d]
    idx = np.array(idx, np.uint32)
    centre = (lo + hi) / 2
    t = idx.reshape(-1, 3)
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    d = p[t].mean(1) - centre
    outward = np.einsum("ij,ij->i", n, d) > 0
    flip = outward if inward else ~outward
    t[flip] = t[flip][:, [0, 2, 1]]
    idx = t.reshape(-1).astype(np.uint32)
    _check_winding(p, idx, centre, inward)
    return p.astype(np.float32), idx


def icosphere(subdiv: int) -> Tuple[np.ndarray, np.ndarray]:
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts)
    idx = np.array(faces, np.uint32).reshape(-1)
    _check_winding(p, idx, np.zeros(3))
    return p.astype(np.float32), idx


def _noise_field(rng: np.random.Generator, n_waves: int = 6):
    dirs = rng.normal(size=(n_waves, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    freq = rng.uniform(2.0, 9.0, n_waves)
    phase = rng.uniform(0, 2 * math.pi, n_waves)

    def f(p):
        return np.sin(p @ (dirs * freq[:, None]).T + phase).sum(-1) / n_waves

    return f


def uv_sphere(rings: int, segs: int, rng: np.random.Generator, amp: float = 0.06):
    """Noisy-displaced UV sphere of radius ~0.5: 2*segs*(rings-1) triangles."""
    pos, uv = [], []
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(segs + 1):
            ph = 2 * math.pi * s / segs
            pos.append((math.sin(th) * math.cos(ph), math.cos(th), -math.sin(th) * math.sin(ph)))
            uv.append((s / segs, 1.0 - r / rings))
    pos = np.array(pos)
    f = _noise_field(rng)
    pos = pos * (0.5 * (1.0 + amp * f(pos)))[:, None]
    idx = []
    W = segs + 1
    for r in range(rings):
        for s in range(segs):
            a, b, c, d = r * W + s, r * W + s + 1, (r + 1) * W + s, (r + 1) * W + s + 1
            if r != 0:
                idx += [a, c, b]
            if r != rings - 1:
                idx += [b, c, d]
    idx = np.array(idx, np.uint32)
    _check_winding(pos, idx, np.zeros(3))
    return pos.astype(np.float32), np.array(uv, np.float32), idx


def grid_cuboid(n: int, rng: np.random.Generator, amp: float = 0.04):
    """Unit cuboid ([-0.5,0.5]^3) with an n x n grid per face, displaced along
    the face normal by a smooth noise field: 12 n^2 triangles."""
    f = _noise_field(rng)
    pos, uv, idx = [], [], []
    axes = [(0, 1, 2), (0, 2, 1), (1, 2, 0)]
    base = 0
    for (u_ax, v_ax, n_ax) in axes:
        for sign in (-1.0, 1.0):
            g = np.linspace(-0.5, 0.5, n + 1)
            U, V = np.meshgrid(g, g, indexing="xy")
            P = np.zeros((n + 1, n + 1, 3))
            P[..., u_ax], P[..., v_ax], P[..., n_ax] = U, V, 0.5 * sign
            P = P.reshape(-1, 3)
            nrm = np.zeros(3)
            nrm[n_ax] = sign
            P = P + amp * f(P)[:, None] * nrm
            pos.append(P)
            uv.append(np.stack([(U.reshape(-1) + 0.5), (V.reshape(-1) + 0.5)], 1))
            for j in range(n):
                for i in range(n):
                    a = base + j * (n + 1) + i
                    b, c, d = a + 1, a + n + 1, a + n + 2
                    idx += [a, b, d, a, d, c]
            base += (n + 1) ** 2
    pos = np.concatenate(pos)
    idx = np.array(idx, np.uint32).reshape(-1, 3)
    # orient every face outwards
    a, b, c = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    s = np.einsum("ij,ij->i", np.cross(b - a, c - a), (a + b + c) / 3)
    idx[s < 0] = idx[s < 0][:, [0, 2, 1]]
    idx = idx.reshape(-1).astype(np.uint32)
    return pos.astype(np.float32), np.concatenate(uv).astype(np.float32), idx


def random_pose_in_box(rng, lo, hi, height=1.5, pitch_deg=15.0) -> np.ndarray:
    x = rng.uniform(lo[0], hi[0])
    z = rng.uniform(lo[1], hi[1])
    yaw = rng.uniform(0.0, 2 * math.pi)
    pitch = math.radians(rng.uniform(-pitch_deg, pitch_deg))
    return m3.camera_pose((x, height, z), yaw, pitch)


# ---------------------------------------------------------------------------
# small_rooms: 16 envs, vertex-coloured, untextured
# ---------------------------------------------------------------------------

def small_rooms(n_envs: int = 16, tile: int = 128, config_index: int = 0) -> BatchSpec:
    fd = FileData()
    fd.materials.append(MaterialData(flat=True))
    P = m3.pinhole_projection((tile, tile), math.radians(90.0), 0.01, 1000.0)
    envs, tris = [], {}
    sphere_p, sphere_i = icosphere(3)

    def add_template(name, pos, idx, rng):
        col = rng.integers(0, 256, size=(len(pos), 4), dtype=np.uint8)
        col[:, 3] = 255
        mi = len(fd.meshes)
        fd.meshes.append(MeshData(positions=np.ascontiguousarray(pos, np.float32),
                                  indices=idx, colors=col))
        root = len(fd.nodes)
        fd.nodes.append(NodeData(parent=-1, name=name))
        fd.nodes.append(NodeData(parent=root))
        fd.mesh_entries.append(MeshEntry(node=root + 1, mesh=mi, material=0))
        tris[name] = len(idx) // 3

    for e in range(n_envs):
        rng = env_rng(config_index, e)
        insts = []
        room_p, room_i = box_mesh((-4, 0, -3), (4, 3, 3), inward=True)
        add_template(f"e{e}/room", room_p, room_i, rng)
        insts.append(Instance(f"e{e}/room", m3.identity(), 1))
        cube_p, cube_i = box_mesh((-0.5, 0, -0.5), (0.5, 1, 0.5), inward=False)
        for k in range(20):
            size = rng.uniform(0.2, 1.5, 3)
            yaw = rng.uniform(0, 2 * math.pi)
            x = rng.uniform(-4 + 0.75, 4 - 0.75)
            z = rng.uniform(-3 + 0.75, 3 - 0.75)
            name = f"e{e}/cuboid{k}"
            add_template(name, cube_p, cube_i, rng)
            bake = m3.mul(m3.mul(m3.translation((x, 0.0, z)), m3.rotation_y(yaw)), m3.scaling(size))
            insts.append(Instance(name, bake, 2 + k))
        rad = rng.uniform(0.3, 0.8)
        x = rng.uniform(-4 + 1, 4 - 1)
        z = rng.uniform(-3 + 1, 3 - 1)
        add_template(f"e{e}/sphere", sphere_p, sphere_i, rng)
        insts.append(Instance(f"e{e}/sphere", m3.mul(m3.translation((x, rad, z)), m3.scaling(rad)), 22))
        pose = random_pose_in_box(rng, (-3.5, -2.5), (3.5, 2.5))
        envs.append(EnvSpec(insts, pose, P))
    return BatchSpec("small_rooms", (tile, tile), fd, envs, tris)


# ---------------------------------------------------------------------------
# apartments: 512 envs, 20 shared textured templates + apartment shell
# ---------------------------------------------------------------------------

ROOM = 4.0
GRID = 3
HEIGHT = 3.0
DOOR = 1.0


def apartment_mesh() -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Floor and ceiling (12x12 grids), outer walls facing in, inner walls as
    double-sided panels with a 1 m doorway centred in every 4 m segment."""
    size = ROOM * GRID
    pos, uv, idx = [], [], []

    def quad(a, b, c, d, uvs):
        base = sum(len(p) for p in pos)
        pos.append(np.array([a, b, c, d], np.float64))
        uv.append(np.array(uvs, np.float64))
        idx.extend([base, base + 1, base + 2, base, base + 2, base + 3])

    n = 12
    for j in range(n):
        for i in range(n):
            x0, x1 = size * i / n, size * (i + 1) / n
            z0, z1 = size * j / n, size * (j + 1) / n
            # floor faces up: CCW seen from above
            quad((x0, 0, z0), (x0, 0, z1), (x1, 0, z1), (x1, 0, z0),
                 [(x0 / 2, z0 / 2), (x0 / 2, z1 / 2), (x1 / 2, z1 / 2), (x1 / 2, z0 / 2)])
            # ceiling faces down
            quad((x0, HEIGHT, z0), (x1, HEIGHT, z0), (x1, HEIGHT, z1), (x0, HEIGHT, z1),
                 [(x0 / 2, z0 / 2), (x1 / 2, z0 / 2), (x1 / 2, z1 / 2), (x0 / 2, z1 / 2)])

    def wall(p0, p1, both):
        (x0, z0), (x1, z1) = p0, p1
        L = math.hypot(x1 - x0, z1 - z0)
        uvs = [(0, 0), (L / 2, 0), (L / 2, HEIGHT / 2), (0, HEIGHT / 2)]
        quad((x0, 0, z0), (x1, 0, z1), (x1, HEIGHT, z1), (x0, HEIGHT, z0), uvs)
        if both:
            quad((x1, 0, z1), (x0, 0, z0), (x0, HEIGHT, z0), (x1, HEIGHT, z1), uvs)

    # outer walls, facing inwards (normal (b-a)x(c-a) points into the box)
    wall((0, 0), (size, 0), False)
    wall((size, 0), (size, size), False)
    wall((size, size), (0, size), False)
    wall((0, size), (0, 0), False)
    # inner walls with doorways
    for k in (1, 2):
        c = ROOM * k
        for seg in range(GRID):
            a, b = ROOM * seg, ROOM * (seg + 1)
            m0, m1 = (a + b) / 2 - DOOR / 2, (a + b) / 2 + DOOR / 2
            for (s0, s1) in ((a, m0), (m1, b)):
                wall((s0, c), (s1, c), True)
                wall((c, s0), (c, s1), True)
    return (np.concatenate(pos).astype(np.float32), np.concatenate(uv).astype(np.float32),
            np.array(idx, np.uint32))


def noise_texture(rng: np.random.Generator, size: int, layers: int) -> TextureData:
    lv0 = rng.integers(0, 256, size=(layers, size, size, 4), dtype=np.uint8)
    lv0[..., 3] = 255
    return TextureData(width=size, height=size, layers=layers, levels=generate_mipmaps(lv0),
                       min_filter=GL_LINEAR_MIPMAP_LINEAR, mag_filter=GL_LINEAR,
                       wrap_s=GL_REPEAT, wrap_t=GL_REPEAT)


def apartment_templates(config_index: int = 1, n_templates: int = 20, tex_size: int = 256):
    """Shared FileData: 'apartment' + 'obj0'..'obj{n-1}' templates; returns
    (file, template triangle counts, template pick weights)."""
    rng = env_rng(config_index, 0xFFFFFFFF)
    fd = FileData()
    fd.textures.append(noise_texture(rng, tex_size, n_templates + 1))
    tris = {}

    def add(name, pos, uv, idx, layer):
        mi = len(fd.meshes)
        fd.meshes.append(MeshData(positions=pos, indices=idx, texcoords=uv, texcoord_half=True))
        fd.materials.append(MaterialData(flat=True, texture=0, layer=layer))
        root = len(fd.nodes)
        fd.nodes.append(NodeData(parent=-1, name=name))
        fd.nodes.append(NodeData(parent=root))
        fd.mesh_entries.append(MeshEntry(node=root + 1, mesh=mi, material=mi))
        tris[name] = len(idx) // 3

    ap, auv, ai = apartment_mesh()
    add("apartment", ap, auv, ai, n_templates)
    targets = [500.0 * 10 ** (k / (n_templates - 1)) for k in range(n_templates)]
    for k, t in enumerate(targets):
        if k % 2 == 0:
            rings = max(3, int(round(math.sqrt(t / 4.0))))
            p, uv, idx = uv_sphere(rings, 2 * rings, rng)
        else:
            n = max(2, int(round(math.sqrt(t / 12.0))))
            p, uv, idx = grid_cuboid(n, rng)
        add(f"obj{k}", p, uv, idx, k)
    counts = np.array([tris[f"obj{k}"] for k in range(n_templates)], np.float64)
    weights = counts ** -3.0
    weights /= weights.sum()
    return fd, tris, weights


_OBJ_NAMES = [f"obj{i}" for i in range(64)]


def _scatter(rng: np.random.Generator, picks: np.ndarray):
    """The instances of an apartment: the shell, then per pick a template
    scaled by s ~ U[0.3, 1.2], turned by yaw ~ U[0, 2 pi) and standing on
    the floor at (x, z) ~ U[0.5, 11.5]^2, drawn from the env's stream in the
    order s, x, z, yaw per instance.  Bakes are the float32 product
    translation * rotation_y * scaling written out (each entry one product,
    as the 4x4 product gives it).  Returns (instances, arrays)."""
    n = len(picks)
    u = rng.random((n, 4))
    span = ROOM * GRID - 0.5
    s = 0.3 + (1.2 - 0.3) * u[:, 0]
    x = 0.5 + (span - 0.5) * u[:, 1]
    z = 0.5 + (span - 0.5) * u[:, 2]
    yaw = 0.0 + (2 * math.pi - 0.0) * u[:, 3]
    c = np.array([math.cos(v) for v in yaw], np.float32)
    sn = np.array([math.sin(v) for v in yaw], np.float32)
    s32 = s.astype(np.float32)
    bakes = np.zeros((n + 1, 4, 4), np.float32)
    bakes[0] = m3.identity()
    b = bakes[1:]
    b[:, 0, 0] = c * s32
    b[:, 0, 2] = -sn * s32
    b[:, 1, 1] = s32
    b[:, 2, 0] = sn * s32
    b[:, 2, 2] = c * s32
    b[:, 3, 0] = x.astype(np.float32)
    b[:, 3, 1] = (0.5 * s).astype(np.float32)
    b[:, 3, 2] = z.astype(np.float32)
    b[:, 3, 3] = 1.0
    names = ["apartment"] + [_OBJ_NAMES[int(k)] for k in picks]
    sems = np.arange(1, n + 2, dtype=np.uint32)
    insts = [Instance(names[i], bakes[i], int(sems[i])) for i in range(n + 1)]
    return insts, (names, bakes, sems)


def apartments(n_envs: int = 512, tile: int = 128, instances: int = 300,
               config_index: int = 1, template_file=None) -> BatchSpec:
    fd, tris, weights = template_file or apartment_templates(config_index)
    P = m3.pinhole_projection((tile, tile), math.radians(90.0), 0.01, 1000.0)
    n_t = len(weights)
    envs = []
    for e in range(n_envs):
        rng = env_rng(config_index, e)
        picks = rng.choice(n_t, size=instances, p=weights)
        insts, arrays = _scatter(rng, picks)
        room = rng.integers(0, GRID, 2)
        lo = (room[0] * ROOM + 0.3, room[1] * ROOM + 0.3)
        hi = ((room[0] + 1) * ROOM - 0.3, (room[1] + 1) * ROOM - 0.3)
        envs.append(EnvSpec(insts, random_pose_in_box(rng, lo, hi), P, arrays=arrays))
    return BatchSpec("apartments", (tile, tile), fd, envs, tris)


# sensor planes (the library's PLANE_* bits)
PLANE_COLOR, PLANE_DEPTH, PLANE_LINEAR_DEPTH, PLANE_SEMANTIC = 0x1, 0x2, 0x4, 0x8


def heterogeneous(n_envs: int = 1024, tri_range=(1e4, 1e6), resolutions=(64, 128, 256),
                  res_p=(0.3, 0.5, 0.2), config_index: int = 2, template_file=None) -> BatchSpec:
    """Uneven batch: per env a triangle budget drawn log-uniformly from
    tri_range (apartment shell plus as many templates as fit), a square
    resolution from `resolutions` with probabilities res_p (the env's
    viewport in a tile of the largest resolution), and a sensor set of 1-3
    of colour, depth and semantic (linear depth comes with depth).  Cameras
    as in apartments()."""
    fd, tris, weights = template_file or apartment_templates(1)
    tile = max(resolutions)
    shell = tris["apartment"]
    n_t = len(weights)
    mean_t = float(sum(w * tris[f"obj{i}"] for i, w in enumerate(weights)))
    sensor_sets = [PLANE_COLOR, PLANE_DEPTH | PLANE_LINEAR_DEPTH, PLANE_SEMANTIC]
    envs = []
    for e in range(n_envs):
        rng = env_rng(config_index, e)
        budget = math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))
        count = max(0, int(round((budget - shell) / mean_t)))
        res = int(rng.choice(resolutions, p=res_p))
        k = int(rng.integers(1, 4))
        chosen = rng.choice(3, size=k, replace=False)
        planes = 0
        for c in chosen:
            planes |= sensor_sets[int(c)]
        P = m3.pinhole_projection((res, res), math.radians(90.0), 0.01, 1000.0)
        picks = rng.choice(n_t, size=count, p=weights)
        insts, arrays = _scatter(rng, picks)
        room = rng.integers(0, GRID, 2)
        lo = (room[0] * ROOM + 0.3, room[1] * ROOM + 0.3)
        hi = ((room[0] + 1) * ROOM - 0.3, (room[1] + 1) * ROOM - 0.3)
        vp = None if res == tile else (res, res)
        envs.append(EnvSpec(insts, random_pose_in_box(rng, lo, hi), P, vp, planes, arrays))
    return BatchSpec("heterogeneous", (tile, tile), fd, envs, tris)
