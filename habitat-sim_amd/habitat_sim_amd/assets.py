"""Batch-friendly glTF composite import for the batch renderer.

Parses the file formats the reference's ``Renderer::addFile()`` consumes
through Magnum's importers (src/esp/gfx_batch/Renderer.cpp:354-867) into a
plain :class:`FileData` description: meshes (triangle lists with SoA vertex
streams), textures (RGBA8 levels, bottom-up rows), materials, and scene nodes
carrying the ``meshViewIndexOffset`` / ``meshViewIndexCount`` /
``meshViewMaterial`` custom fields (Renderer.h:326-346).  The renderer's
``add_file_data()`` then applies ``addFile()``'s own logic (offsets, flat
materials, node-hierarchy templates, RendererFileFlag::Whole) natively.

Conventions reproduced from Magnum's glTF/KTX/PNG importers, pinned by the
reference fixtures ``data/test_assets/scenes/batch*.gltf`` and the PNG goldens
of ``src/tests/GfxBatchRendererTest.cpp``:

* images are stored with the first row at the bottom (Y up); PNG rows are
  flipped, KTX2 ``KTXorientation`` "u" data is kept, "d" data is flipped;
* glTF texture coordinates are flipped (v -> 1 - v) and KHR_texture_transform
  matrices are conjugated by the same flip (GfxBatchRendererTest.cpp:417-419
  vs. batch.gltf's "flat cyan" material);
* KHR_texture_ktx textures referencing the same (image, sampler) pair are one
  2D-array texture; the layer goes to the material (Renderer.cpp:595-597);
* non-indexed meshes become indexed, triangle fans / strips become lists in
  GL's primitive order (Renderer.cpp:545-547).
"""
from __future__ import annotations

import io
import json
import math
import os
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

# GL sampler enums (glTF uses the same values)
GL_NEAREST = 9728
GL_LINEAR = 9729
GL_NEAREST_MIPMAP_NEAREST = 9984
GL_LINEAR_MIPMAP_NEAREST = 9985
GL_NEAREST_MIPMAP_LINEAR = 9986
GL_LINEAR_MIPMAP_LINEAR = 9987
GL_REPEAT = 10497
GL_CLAMP_TO_EDGE = 33071
GL_MIRRORED_REPEAT = 33648

# file flags, Renderer.h:120-154
FILE_WHOLE = 1 << 0
FILE_GENERATE_MIPMAP = 1 << 1


class ImportError_(RuntimeError):
    """Raised when a file cannot be imported (addFile() returns false)."""


@dataclass
class MeshData:
    positions: np.ndarray                 # (V, 3) float32
    indices: np.ndarray                   # (3T,) uint32, triangle list
    texcoords: Optional[np.ndarray] = None  # (V, 2) float32
    colors: Optional[np.ndarray] = None   # (V, 4) float32 or uint8
    texcoord_half: bool = False           # store texture coordinates as fp16
    index_size: int = 4                   # byte size of the source index type
    normals: Optional[np.ndarray] = None  # (V, 3) float32 (shaded materials)


@dataclass
class TextureData:
    width: int
    height: int
    layers: int
    levels: List[np.ndarray]              # each (layers, h, w, 4) uint8, bottom-up rows
    min_filter: int = GL_LINEAR_MIPMAP_LINEAR
    mag_filter: int = GL_LINEAR
    wrap_s: int = GL_REPEAT
    wrap_t: int = GL_REPEAT
    compressed: bool = False              # source was block-compressed


@dataclass
class MaterialData:
    flat: bool
    base_color: Tuple[float, float, float, float] = (1.0, 1.0, 1.0, 1.0)
    texture: int = -1
    layer: int = 0
    texture_matrix: np.ndarray = field(
        default_factory=lambda: np.eye(3, dtype=np.float32))


@dataclass
class NodeData:
    parent: int = -1
    transform: np.ndarray = field(
        default_factory=lambda: np.eye(4, dtype=np.float32))  # [col][row]
    name: Optional[str] = None


@dataclass
class MeshEntry:
    """One entry of the scene's Mesh field (an object may have several)."""
    node: int
    mesh: int
    material: int = -1                           # material of the mesh primitive
    view: Optional[Tuple[int, int, int]] = None  # (offset bytes, count, material)


@dataclass
class FileData:
    meshes: List[MeshData] = field(default_factory=list)
    textures: List[TextureData] = field(default_factory=list)
    materials: List[MaterialData] = field(default_factory=list)
    nodes: List[NodeData] = field(default_factory=list)
    mesh_entries: List[MeshEntry] = field(default_factory=list)
    has_scene: bool = True


# ---------------------------------------------------------------------------
# Image decoding
# ---------------------------------------------------------------------------

def _expand_rgba(a: np.ndarray) -> np.ndarray:
    """(..., C) uint8 -> (..., 4) uint8 with alpha 255 where missing."""
    c = a.shape[-1]
    if c == 4:
        return a
    out = np.empty(a.shape[:-1] + (4,), np.uint8)
    if c == 3:
        out[..., :3] = a
    elif c == 2:  # luminance-alpha
        out[..., :3] = a[..., :1]
        out[..., 3] = a[..., 1]
        return out
    else:
        out[..., :3] = a
    out[..., 3] = 255
    return out


def decode_png(data: bytes) -> np.ndarray:
    """PNG -> (h, w, 4) uint8 with the first row at the bottom."""
    from PIL import Image  # decoding only; no other use of PIL

    img = Image.open(io.BytesIO(data))
    if img.mode not in ("RGB", "RGBA", "L", "LA"):
        img = img.convert("RGBA")
    a = np.asarray(img, dtype=np.uint8)
    if a.ndim == 2:
        a = a[..., None]
    return np.ascontiguousarray(_expand_rgba(a)[::-1])


def _rgb565(c: int) -> Tuple[int, int, int]:
    r = (c >> 11) & 31
    g = (c >> 5) & 63
    b = c & 31
    return ((r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2))


def decode_bc1(data: bytes, w: int, h: int, alpha: bool) -> np.ndarray:
    """BC1 (DXT1) blocks -> (h, w, 4) uint8; block row 0 is image row 0."""
    bw, bh = (w + 3) // 4, (h + 3) // 4
    out = np.zeros((bh * 4, bw * 4, 4), np.uint8)
    off = 0
    for by in range(bh):
        for bx in range(bw):
            c0, c1, bits = struct.unpack_from("<HHI", data, off)
            off += 8
            p0, p1 = _rgb565(c0), _rgb565(c1)
            if c0 > c1:
                pal = [p0 + (255,), p1 + (255,),
                       tuple((2 * a + b + 1) // 3 for a, b in zip(p0, p1)) + (255,),
                       tuple((a + 2 * b + 1) // 3 for a, b in zip(p0, p1)) + (255,)]
            else:
                pal = [p0 + (255,), p1 + (255,),
                       tuple((a + b) // 2 for a, b in zip(p0, p1)) + (255,),
                       (0, 0, 0, 0 if alpha else 255)]
            for py in range(4):
                for px in range(4):
                    idx = (bits >> (2 * (py * 4 + px))) & 3
                    out[by * 4 + py, bx * 4 + px] = pal[idx]
    return out[:h, :w]


_KTX2_MAGIC = b"\xabKTX 20\xbb\r\n\x1a\n"
_VK_FORMATS = {
    9: ("u8", 1), 16: ("u8", 2), 23: ("u8", 3), 37: ("u8", 4),
    131: ("bc1", 0), 132: ("bc1", 0), 133: ("bc1a", 0), 134: ("bc1a", 0),
}


def decode_ktx2(data: bytes) -> Tuple[List[np.ndarray], int, bool]:
    """KTX2 -> (levels [(layers, h, w, 4) uint8 bottom-up], layers, compressed).

    Returns the array-layer count (0 in the file means a plain 2D image)."""
    if data[:12] != _KTX2_MAGIC:
        raise ImportError_("not a KTX2 file")
    (vk, _ts, w, h, d, layers, faces, levels, sc) = struct.unpack_from("<9I", data, 12)
    if sc != 0:
        raise ImportError_("supercompressed KTX2 is not supported")
    if d > 1 or faces != 1:
        raise ImportError_("only 2D / 2D-array KTX2 images are supported")
    if vk not in _VK_FORMATS:
        raise ImportError_(f"unsupported KTX2 vkFormat {vk}")
    kind, comps = _VK_FORMATS[vk]
    kvd_off, kvd_len = struct.unpack_from("<2I", data, 56)
    orientation = "ru"
    kv = data[kvd_off:kvd_off + kvd_len]
    p = 0
    while p + 4 <= len(kv):
        (n,) = struct.unpack_from("<I", kv, p)
        entry = kv[p + 4:p + 4 + n]
        key, _, val = entry.partition(b"\0")
        if key == b"KTXorientation":
            orientation = val.rstrip(b"\0").decode()
        p += 4 + ((n + 3) & ~3)
    nlayers = max(layers, 1)
    out = []
    for lv in range(max(levels, 1)):
        off, ln, _uln = struct.unpack_from("<3Q", data, 80 + 24 * lv)
        lw, lh = max(w >> lv, 1), max(h >> lv, 1)
        blob = data[off:off + ln]
        imgs = []
        if kind == "u8":
            a = np.frombuffer(blob, np.uint8).reshape(nlayers, lh, lw, comps)
            for i in range(nlayers):
                img = _expand_rgba(a[i])
                if len(orientation) > 1 and orientation[1] == "d":
                    img = img[::-1]
                imgs.append(img)
        else:
            bsz = ((lw + 3) // 4) * ((lh + 3) // 4) * 8
            for i in range(nlayers):
                imgs.append(decode_bc1(blob[i * bsz:(i + 1) * bsz], lw, lh, kind == "bc1a"))
        out.append(np.ascontiguousarray(np.stack(imgs)))
    return out, layers, kind != "u8"


def generate_mipmaps(level0: np.ndarray) -> List[np.ndarray]:
    """Full mip chain by 2x2 box filter with rounding, (layers, h, w, 4) input.

    Stands in for glGenerateMipmap() (Renderer.cpp:476-477,526-527), whose
    filter is driver-defined."""
    levels = [level0]
    n = int(math.log2(min(level0.shape[1], level0.shape[2]))) + 1
    cur = level0.astype(np.uint32)
    for _ in range(1, n):
        h, w = cur.shape[1], cur.shape[2]
        h2, w2 = max(h // 2, 1), max(w // 2, 1)
        c = cur[:, :h2 * 2, :w2 * 2]
        s = c[:, 0::2, 0::2] + c[:, 1::2, 0::2] + c[:, 0::2, 1::2] + c[:, 1::2, 1::2]
        cur = (s + 2) >> 2
        levels.append(np.ascontiguousarray(cur.astype(np.uint8)))
    return levels


# ---------------------------------------------------------------------------
# glTF
# ---------------------------------------------------------------------------

_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16,
         5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}
_NORM = {5120: 127.0, 5121: 255.0, 5122: 32767.0, 5123: 65535.0}


def _flip_y3() -> np.ndarray:
    f = np.eye(3, dtype=np.float32)
    f[1, 1] = -1.0
    f[2, 1] = 1.0  # [col][row]: translation (0, 1)
    return f


def _mat3_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """3x3 product on [col][row] arrays, Magnum order, float32."""
    out = np.zeros((3, 3), np.float32)
    for c in range(3):
        for r in range(3):
            s = np.float32(0.0)
            for k in range(3):
                s = np.float32(s + np.float32(a[k, r] * b[c, k]))
            out[c, r] = s
    return out


def _mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    for c in range(4):
        for r in range(4):
            s = np.float32(0.0)
            for k in range(4):
                s = np.float32(s + np.float32(a[k, r] * b[c, k]))
            out[c, r] = s
    return out


class _Gltf:
    def __init__(self, path: str):
        self.path = path
        self.dir = os.path.dirname(os.path.abspath(path))
        with open(path, "rb") as f:
            raw = f.read()
        self.bin0: Optional[bytes] = None
        if raw[:4] == b"glTF":
            _, _, length = struct.unpack_from("<3I", raw, 0)
            p = 12
            self.j = None
            while p < length:
                clen, ctype = struct.unpack_from("<2I", raw, p)
                chunk = raw[p + 8:p + 8 + clen]
                if ctype == 0x4E4F534A:
                    self.j = json.loads(chunk.decode())
                elif ctype == 0x004E4942:
                    self.bin0 = chunk
                p += 8 + clen
            if self.j is None:
                raise ImportError_("glb without a JSON chunk")
        else:
            self.j = json.loads(raw.decode())
        self._buffers: Dict[int, bytes] = {}

    def buffer(self, i: int) -> bytes:
        if i not in self._buffers:
            b = self.j["buffers"][i]
            uri = b.get("uri")
            if uri is None:
                if self.bin0 is None:
                    raise ImportError_("buffer without uri and no GLB chunk")
                self._buffers[i] = self.bin0
            elif uri.startswith("data:"):
                import base64
                self._buffers[i] = base64.b64decode(uri.split(",", 1)[1])
            else:
                with open(os.path.join(self.dir, uri), "rb") as f:
                    self._buffers[i] = f.read()
        return self._buffers[i]

    def view_bytes(self, vi: int) -> Tuple[bytes, int, int]:
        v = self.j["bufferViews"][vi]
        return self.buffer(v["buffer"]), v.get("byteOffset", 0), v.get("byteStride", 0)

    def accessor(self, ai: int, normalized_float: bool = False) -> np.ndarray:
        a = self.j["accessors"][ai]
        dt = np.dtype(_COMP[a["componentType"]])
        n = _NCOMP[a["type"]]
        count = a["count"]
        buf, off, stride = self.view_bytes(a["bufferView"])
        off += a.get("byteOffset", 0)
        esz = dt.itemsize * n
        if stride == 0 or stride == esz:
            arr = np.frombuffer(buf, dt, count * n, off).reshape(count, n)
        else:
            arr = np.stack([np.frombuffer(buf, dt, n, off + i * stride) for i in range(count)])
        if normalized_float and a["componentType"] != 5126:
            arr = (arr.astype(np.float32) / np.float32(_NORM[a["componentType"]])).astype(np.float32)
        return np.array(arr)

    def image_bytes(self, ii: int) -> Tuple[bytes, str]:
        im = self.j["images"][ii]
        if "bufferView" in im:
            buf, off, _ = self.view_bytes(im["bufferView"])
            ln = self.j["bufferViews"][im["bufferView"]]["byteLength"]
            return buf[off:off + ln], im.get("mimeType", "")
        uri = im["uri"]
        with open(os.path.join(self.dir, uri), "rb") as f:
            data = f.read()
        mime = "image/ktx2" if uri.endswith(".ktx2") else "image/png"
        return data, im.get("mimeType", mime)


def _triangulate(idx: np.ndarray, mode: int) -> np.ndarray:
    if mode == 4:
        return idx
    n = len(idx)
    if mode == 6:  # fan: (v0, v_i+1, v_i+2)
        tris = [(idx[0], idx[i + 1], idx[i + 2]) for i in range(n - 2)]
    elif mode == 5:  # strip, odd triangles swap the first two vertices
        tris = [((idx[i], idx[i + 1], idx[i + 2]) if i % 2 == 0 else
                 (idx[i + 1], idx[i], idx[i + 2])) for i in range(n - 2)]
    else:
        raise ImportError_(f"unsupported primitive mode {mode}")
    return np.asarray(tris, np.uint32).reshape(-1)


def _node_transform(n: dict) -> np.ndarray:
    if "matrix" in n:
        return np.asarray(n["matrix"], np.float32).reshape(4, 4)  # column-major
    t = np.asarray(n.get("translation", [0, 0, 0]), np.float32)
    q = np.asarray(n.get("rotation", [0, 0, 0, 1]), np.float32)
    s = np.asarray(n.get("scale", [1, 1, 1]), np.float32)
    x, y, z, w = [np.float32(v) for v in q]
    rot = np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)],
        [2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)],
        [2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]],
        np.float32)  # [col][row]
    m = np.eye(4, dtype=np.float32)
    for c in range(3):
        m[c, :3] = rot[c] * s[c]
    m[3, :3] = t
    return m


def load_gltf(path: str) -> FileData:
    """Parse a glTF / glb file into a :class:`FileData` (Magnum conventions)."""
    g = _Gltf(path)
    j = g.j
    fd = FileData()

    # textures: KHR_texture_ktx layers of one image share a texture
    tex_map: Dict[int, int] = {}
    tex_layer: Dict[int, int] = {}
    dedup: Dict[Tuple[int, int], int] = {}
    image_cache: Dict[int, Tuple[List[np.ndarray], int, bool]] = {}
    samplers = j.get("samplers", [])
    for ti, t in enumerate(j.get("textures", [])):
        ext = t.get("extensions", {}).get("KHR_texture_ktx")
        src = ext["source"] if ext else t.get("source")
        layer = ext.get("layer", 0) if ext else 0
        if src is None:
            raise ImportError_(f"texture {ti} has no image")
        smp = samplers[t["sampler"]] if "sampler" in t else {}
        key = (src, t.get("sampler", -1))
        if ext and key in dedup:
            tex_map[ti] = dedup[key]
            tex_layer[ti] = layer
            continue
        if src not in image_cache:
            data, mime = g.image_bytes(src)
            if "ktx2" in mime or data[:12] == _KTX2_MAGIC:
                image_cache[src] = decode_ktx2(data)
            else:
                lv = decode_png(data)
                image_cache[src] = ([lv[None]], 0, False)
        levels, nlayers, compressed = image_cache[src]
        td = TextureData(
            width=levels[0].shape[2], height=levels[0].shape[1],
            layers=levels[0].shape[0], levels=list(levels),
            min_filter=smp.get("minFilter", GL_LINEAR_MIPMAP_LINEAR),
            mag_filter=smp.get("magFilter", GL_LINEAR),
            wrap_s=smp.get("wrapS", GL_REPEAT), wrap_t=smp.get("wrapT", GL_REPEAT),
            compressed=compressed)
        tex_map[ti] = len(fd.textures)
        tex_layer[ti] = layer
        if ext:
            dedup[key] = tex_map[ti]
        fd.textures.append(td)

    # materials
    flip = _flip_y3()
    for m in j.get("materials", []):
        pbr = m.get("pbrMetallicRoughness", {})
        flat = "KHR_materials_unlit" in m.get("extensions", {})
        col = tuple(float(np.float32(c)) for c in pbr.get("baseColorFactor", [1, 1, 1, 1]))
        md = MaterialData(flat=flat, base_color=col)
        bct = pbr.get("baseColorTexture")
        if bct is not None:
            md.texture = tex_map[bct["index"]]
            md.layer = tex_layer[bct["index"]]
            tt = bct.get("extensions", {}).get("KHR_texture_transform")
            if tt is not None:
                off = tt.get("offset", [0.0, 0.0])
                sc = tt.get("scale", [1.0, 1.0])
                r = float(tt.get("rotation", 0.0))
                cs, sn = np.float32(math.cos(r)), np.float32(math.sin(r))
                gm = np.zeros((3, 3), np.float32)  # glTF T * R * S, [col][row]
                gm[0, 0] = cs * np.float32(sc[0])
                gm[0, 1] = -sn * np.float32(sc[0])
                gm[1, 0] = sn * np.float32(sc[1])
                gm[1, 1] = cs * np.float32(sc[1])
                gm[2, 0] = np.float32(off[0])
                gm[2, 1] = np.float32(off[1])
                gm[2, 2] = 1.0
                md.texture_matrix = _mat3_mul(_mat3_mul(flip, gm), flip)
        fd.materials.append(md)

    # meshes: one entry per primitive
    prim_ids: List[List[Tuple[int, int]]] = []
    for mesh in j.get("meshes", []):
        ids = []
        for prim in mesh["primitives"]:
            at = prim["attributes"]
            pos = g.accessor(at["POSITION"]).astype(np.float32)
            mode = prim.get("mode", 4)
            if "indices" in prim:
                acc = j["accessors"][prim["indices"]]
                isz = np.dtype(_COMP[acc["componentType"]]).itemsize
                idx = g.accessor(prim["indices"]).reshape(-1).astype(np.uint32)
            else:
                isz = 4
                idx = np.arange(len(pos), dtype=np.uint32)
            idx = _triangulate(idx, mode)
            uv = None
            if "TEXCOORD_0" in at:
                uv = g.accessor(at["TEXCOORD_0"], normalized_float=True).astype(np.float32)
                uv[:, 1] = np.float32(1.0) - uv[:, 1]
            nrm = None
            if "NORMAL" in at:
                nrm = np.ascontiguousarray(g.accessor(at["NORMAL"]).astype(np.float32))
            col = None
            if "COLOR_0" in at:
                c = g.accessor(at["COLOR_0"], normalized_float=True).astype(np.float32)
                if c.shape[1] == 3:
                    c = np.concatenate([c, np.ones((len(c), 1), np.float32)], 1)
                col = c
            ids.append((len(fd.meshes), prim.get("material", -1)))
            fd.meshes.append(MeshData(positions=np.ascontiguousarray(pos), indices=idx,
                                      texcoords=uv, colors=col, index_size=isz, normals=nrm))
        prim_ids.append(ids)

    # scene: nodes in index order, parents from children lists
    nodes = j.get("nodes", [])
    scenes = j.get("scenes", [])
    if not scenes:
        fd.has_scene = False
        return fd
    if len(scenes) != 1:
        raise ImportError_(f"expected exactly one scene, got {len(scenes)}")
    parent = [-1] * len(nodes)
    for i, n in enumerate(nodes):
        for c in n.get("children", []):
            parent[c] = i
    # Magnum objects are the glTF nodes; the Mesh field lists (object, mesh)
    # pairs in object order, one per primitive of the node's mesh.
    for i, n in enumerate(nodes):
        fd.nodes.append(NodeData(parent=parent[i], transform=_node_transform(n),
                                 name=n.get("name")))
        if "mesh" not in n:
            continue
        extras = n.get("extras", {})
        view = None
        if "meshViewIndexOffset" in extras:
            if "meshViewIndexCount" not in extras or "meshViewMaterial" not in extras:
                raise ImportError_("incomplete mesh view fields")
            view = (int(extras["meshViewIndexOffset"]), int(extras["meshViewIndexCount"]),
                    int(extras["meshViewMaterial"]))
        for mesh_id, mat in prim_ids[n["mesh"]]:
            fd.mesh_entries.append(MeshEntry(node=i, mesh=mesh_id, material=mat, view=view))
    return fd
