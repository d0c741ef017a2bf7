"""Python mirror of esp::gfx_batch::RendererConfiguration / Renderer /
RendererStandalone (src/esp/gfx_batch/Renderer.h, RendererStandalone.h) over
the C ABI.  Method names follow the reference in snake_case; argument meaning
and error behaviour are the reference's (out-of-range scene ids raise, a file
that cannot be imported makes add_file() return False).
"""
from __future__ import annotations

import ctypes
from collections import namedtuple
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .assets import FileData, ImportError_, load_gltf

SceneStats = namedtuple("SceneStats", "node_count draw_count draw_batch_count")

RendererFlags_NO_TEXTURES = N.FLAG_NO_TEXTURES
RendererFileFlags_WHOLE = N.FILE_WHOLE
RendererFileFlags_GENERATE_MIPMAP = N.FILE_GENERATE_MIPMAP


class RendererConfiguration:
    """RendererConfiguration, Renderer.h:51-113 (defaults Renderer.cpp:58-64)."""

    def __init__(self):
        self.flags = 0
        self.tile_size = (128, 128)
        self.tile_count = (1, 1)
        self.max_light_count = 0
        self.ambient_factor = 0.1

    def set_flags(self, flags: int) -> "RendererConfiguration":
        self.flags = int(flags)
        return self

    def set_tile_size_count(self, tile_size, tile_count) -> "RendererConfiguration":
        self.tile_size = (int(tile_size[0]), int(tile_size[1]))
        self.tile_count = (int(tile_count[0]), int(tile_count[1]))
        return self

    def set_max_light_count(self, count: int) -> "RendererConfiguration":
        self.max_light_count = int(count)
        return self

    def set_ambient_factor(self, factor: float) -> "RendererConfiguration":
        self.ambient_factor = float(factor)
        return self


class RendererStandaloneConfiguration:
    """RendererStandaloneConfiguration, RendererStandalone.h."""

    def __init__(self):
        self.device = -1
        self.flags = 0

    def set_cuda_device(self, device: int) -> "RendererStandaloneConfiguration":
        """Device ordinal (the reference's CUDA device; a HIP device here)."""
        self.device = int(device)
        return self

    def set_flags(self, flags: int) -> "RendererStandaloneConfiguration":
        self.flags = int(flags)
        return self


def _mat16(m) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    return a


class _TransformView:
    """Mutable view of a scene's node transformations (Renderer.h:682-683):
    ``view[i]`` is a (4, 4) float32 [col][row] array aliasing the renderer's
    host storage; assignment writes through."""

    def __init__(self, array: np.ndarray):
        self._a = array

    def __len__(self):
        return self._a.shape[0]

    def __getitem__(self, i):
        return self._a[i]

    def __setitem__(self, i, value):
        self._a[i] = np.asarray(value, np.float32).reshape(4, 4)

    def __array__(self, dtype=None):
        return self._a if dtype is None else self._a.astype(dtype)


class RendererStandalone:
    """RendererStandalone (RendererStandalone.h:120-241): renderer + its own
    framebuffer of tile_size * tile_count pixels on one MI355X."""

    def __init__(self, configuration: Optional[RendererConfiguration] = None,
                 standalone_configuration: Optional[RendererStandaloneConfiguration] = None,
                 planes: int = N.PLANE_ALL, top_down: bool = False):
        configuration = configuration or RendererConfiguration()
        standalone_configuration = standalone_configuration or RendererStandaloneConfiguration()
        lib = N.lib()
        cfg = N.Config()
        N.check(lib.gbr_config_init(ctypes.byref(cfg)))
        cfg.tile_size[0], cfg.tile_size[1] = configuration.tile_size
        cfg.tile_count[0], cfg.tile_count[1] = configuration.tile_count
        cfg.flags = configuration.flags
        cfg.max_light_count = configuration.max_light_count
        cfg.ambient_factor = configuration.ambient_factor
        cfg.device = standalone_configuration.device
        cfg.standalone_flags = standalone_configuration.flags
        cfg.planes = planes
        cfg.top_down = 1 if top_down else 0
        h = ctypes.c_void_p()
        N.check(lib.gbr_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._cfg = cfg
        self._keep = []

    # -- lifecycle -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().gbr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- getters (Renderer.h:458-491) ---------------------------------------
    @property
    def flags(self) -> int:
        return self._cfg.flags

    @property
    def standalone_flags(self) -> int:
        return self._cfg.standalone_flags

    @property
    def tile_size(self):
        return (self._cfg.tile_size[0], self._cfg.tile_size[1])

    @property
    def tile_count(self):
        return (self._cfg.tile_count[0], self._cfg.tile_count[1])

    @property
    def scene_count(self) -> int:
        n = ctypes.c_uint64()
        N.call("gbr_scene_count", self._h, ctypes.byref(n))
        return n.value

    @property
    def max_light_count(self) -> int:
        return self._cfg.max_light_count

    @property
    def framebuffer_size(self):
        return (self.tile_size[0] * self.tile_count[0], self.tile_size[1] * self.tile_count[1])

    # -- assets ----------------------------------------------------------------
    def add_file(self, filename: str, flags: int = 0, name: Optional[str] = None) -> bool:
        """Renderer::addFile() through the library's native glTF importer;
        returns False if the file can't be imported (as the reference)."""
        st = N.lib().gbr_add_file(self._h, filename.encode(), flags,
                                  name.encode() if name else None)
        if st in (N.ERR_IO, N.ERR_STATE, N.ERR_ARGUMENT):
            import sys
            print(N.lib().gbr_last_error().decode(), file=sys.stderr)
            return False
        N.check(st)
        return True

    def add_file_python(self, filename: str, flags: int = 0, name: Optional[str] = None) -> bool:
        """Renderer::addFile() through the Python importer (assets.load_gltf)
        and gbr_add_file_data(); same result as add_file()."""
        try:
            fd = load_gltf(filename)
        except (ImportError_, OSError, ValueError, KeyError) as e:
            import sys
            print(f"Renderer::addFile(): can't open the file {filename}: {e}", file=sys.stderr)
            return False
        return self.add_file_data(fd, flags, name, filename)

    def add_file_data(self, fd: FileData, flags: int = 0, name: Optional[str] = None,
                      filename: str = "") -> bool:
        keep = []

        def ptr(a):
            keep.append(a)
            return a.ctypes.data

        meshes = (N.MeshDesc * max(len(fd.meshes), 1))()
        for i, m in enumerate(fd.meshes):
            d = meshes[i]
            pos = np.ascontiguousarray(m.positions, np.float32)
            d.vertex_count = len(pos)
            d.positions = ptr(pos)
            if m.texcoords is not None:
                d.texcoords = ptr(np.ascontiguousarray(m.texcoords, np.float32))
            if m.colors is not None:
                c = m.colors
                if c.dtype == np.uint8:
                    d.colors = ptr(np.ascontiguousarray(c, np.uint8))
                    d.colors_u8 = 1
                else:
                    d.colors = ptr(np.ascontiguousarray(c, np.float32))
            d.texcoords_half = 1 if m.texcoord_half else 0
            idx = np.ascontiguousarray(m.indices, np.uint32)
            d.index_count = len(idx)
            d.indices = ptr(idx)
            d.source_index_size = m.index_size
            if getattr(m, "normals", None) is not None:
                d.normals = ptr(np.ascontiguousarray(m.normals, np.float32))
        textures = (N.TextureDesc * max(len(fd.textures), 1))()
        for i, t in enumerate(fd.textures):
            d = textures[i]
            d.width, d.height, d.layers, d.levels = t.width, t.height, t.layers, len(t.levels)
            blob = np.concatenate([np.ascontiguousarray(lv, np.uint8).reshape(-1) for lv in t.levels])
            d.texels = ptr(blob)
            d.min_filter, d.mag_filter = t.min_filter, t.mag_filter
            d.wrap_s, d.wrap_t = t.wrap_s, t.wrap_t
            d.compressed = 1 if t.compressed else 0
        materials = (N.MaterialDesc * max(len(fd.materials), 1))()
        for i, m in enumerate(fd.materials):
            d = materials[i]
            d.flat = 1 if m.flat else 0
            for k in range(4):
                d.base_color[k] = m.base_color[k]
            d.texture = m.texture
            d.layer = m.layer
            tm = np.asarray(m.texture_matrix, np.float32).reshape(9)
            for k in range(9):
                d.texture_matrix[k] = tm[k]
        nodes = (N.NodeDesc * max(len(fd.nodes), 1))()
        names = []
        for i, n in enumerate(fd.nodes):
            d = nodes[i]
            d.parent = n.parent
            tf = np.asarray(n.transform, np.float32).reshape(16)
            for k in range(16):
                d.transform[k] = tf[k]
            if n.name is not None:
                b = n.name.encode()
                names.append(b)
                d.name = b
        entries = (N.MeshEntryDesc * max(len(fd.mesh_entries), 1))()
        for i, e in enumerate(fd.mesh_entries):
            d = entries[i]
            d.node, d.mesh, d.material = e.node, e.mesh, e.material
            if e.view is not None:
                d.has_view = 1
                d.view_index_offset_bytes, d.view_index_count, d.view_material = e.view
        f = N.FileDesc()
        f.mesh_count, f.meshes = len(fd.meshes), meshes
        f.texture_count, f.textures = len(fd.textures), textures
        f.material_count, f.materials = len(fd.materials), materials
        f.node_count, f.nodes = len(fd.nodes), nodes
        f.mesh_entry_count, f.mesh_entries = len(fd.mesh_entries), entries
        f.has_scene = 1 if fd.has_scene else 0
        st = N.lib().gbr_add_file_data(self._h, ctypes.byref(f), flags,
                                       name.encode() if name else None,
                                       filename.encode() if filename else None)
        if st in (N.ERR_IO, N.ERR_STATE, N.ERR_ARGUMENT):
            import sys
            print(N.lib().gbr_last_error().decode(), file=sys.stderr)
            return False
        N.check(st)
        return True

    def has_node_hierarchy(self, name: str) -> bool:
        out = ctypes.c_int32()
        N.call("gbr_has_node_hierarchy", self._h, name.encode(), ctypes.byref(out))
        return bool(out.value)

    # -- scenes ------------------------------------------------------------------
    def add_node_hierarchy(self, scene: int, name: str, bake=None, semantic_id: int = 0) -> int:
        out = ctypes.c_uint64()
        b = None if bake is None else _mat16(bake)
        N.call("gbr_add_node_hierarchy_semantic", self._h, scene, name.encode(),
               None if b is None else b.ctypes.data, semantic_id, ctypes.byref(out))
        return out.value

    def add_node_hierarchies(self, scene: int, names, bakes=None, semantic_ids=None) -> np.ndarray:
        """add_node_hierarchy() for many instances of one scene in one call:
        names[i], bakes (n, 4, 4) or None, semantic_ids (n,) or None; returns
        the top-level node ids."""
        n = len(names)
        enc = (ctypes.c_char_p * max(n, 1))(*[x if isinstance(x, bytes) else x.encode()
                                             for x in names])
        b = None if bakes is None else np.ascontiguousarray(bakes, np.float32).reshape(n, 16)
        sem = None if semantic_ids is None else np.ascontiguousarray(semantic_ids, np.uint32)
        out = np.empty(max(n, 1), np.uint64)
        N.call("gbr_add_node_hierarchies", self._h, scene, n, ctypes.cast(enc, ctypes.c_void_p),
               None if b is None else b.ctypes.data, None if sem is None else sem.ctypes.data,
               out.ctypes.data)
        return out[:n]

    def set_node_hierarchy_semantic(self, scene: int, node: int, semantic_id: int) -> None:
        """Semantic id of every draw of the hierarchy under top-level `node`."""
        N.call("gbr_set_node_hierarchy_semantic", self._h, scene, node, semantic_id)

    def add_empty_node(self, scene: int) -> int:
        out = ctypes.c_uint64()
        N.call("gbr_add_empty_node", self._h, scene, ctypes.byref(out))
        return out.value

    def add_light(self, scene: int, node: int, light_type: int) -> int:
        out = ctypes.c_uint64()
        N.call("gbr_add_light", self._h, scene, node, light_type, ctypes.byref(out))
        return out.value

    def clear(self, scene: int) -> None:
        N.call("gbr_clear", self._h, scene)

    def set_scene_viewport(self, scene: int, size) -> None:
        """Render the scene into the bottom-left size = (width, height)
        pixels of its tile (new: per-scene resolutions in one batch)."""
        N.call("gbr_set_scene_viewport", self._h, scene, int(size[0]), int(size[1]))

    def scene_viewport(self, scene: int) -> tuple:
        out = (ctypes.c_int32 * 2)()
        N.call("gbr_scene_viewport", self._h, scene, out)
        return int(out[0]), int(out[1])

    def set_scene_planes(self, scene: int, planes: int) -> None:
        """Sensor set of the scene: the PLANE_* bits draw() writes for it
        (0 = every configured plane)."""
        N.call("gbr_set_scene_planes", self._h, scene, planes)

    def clear_lights(self, scene: int) -> None:
        N.call("gbr_clear_lights", self._h, scene)

    def _light_view(self, fn: str, scene: int, cols: int) -> np.ndarray:
        data = N.c_float_p()
        count, stride = ctypes.c_uint64(), ctypes.c_uint64()
        N.call(fn, self._h, scene, ctypes.byref(data), ctypes.byref(count), ctypes.byref(stride))
        n, st = count.value, stride.value
        shape = (n, cols) if cols > 1 else (n,)
        if n == 0:
            return np.zeros(shape, np.float32)
        buf = (ctypes.c_char * ((n - 1) * st + 4 * cols)).from_address(
            ctypes.cast(data, ctypes.c_void_p).value)
        strides = (st, 4) if cols > 1 else (st,)
        return np.ndarray(shape, np.float32, buffer=buf, strides=strides)

    def light_colors(self, scene: int) -> np.ndarray:
        """Renderer::lightColors(): writable (lights, 3) view of the scene's
        light colours, valid until the next add_light(), clear_lights() or
        clear() of the scene."""
        return self._light_view("gbr_light_colors", scene, 3)

    def light_ranges(self, scene: int) -> np.ndarray:
        """Renderer::lightRanges(): writable (lights,) view of the ranges."""
        return self._light_view("gbr_light_ranges", scene, 1)

    def camera(self, scene: int) -> np.ndarray:
        out = np.empty(16, np.float32)
        N.call("gbr_camera", self._h, scene, out.ctypes.data)
        return out.reshape(4, 4)

    def camera_depth_unprojection(self, scene: int) -> np.ndarray:
        out = np.empty(2, np.float32)
        N.call("gbr_camera_depth_unprojection", self._h, scene, out.ctypes.data)
        return out

    def update_camera(self, scene: int, projection, view) -> None:
        p, v = _mat16(projection), _mat16(view)
        N.call("gbr_update_camera", self._h, scene, p.ctypes.data, v.ctypes.data)

    def update_cameras(self, first: int, projections, views) -> None:
        p = np.ascontiguousarray(projections, np.float32).reshape(-1, 16)
        v = np.ascontiguousarray(views, np.float32).reshape(-1, 16)
        assert len(p) == len(v)
        N.call("gbr_update_cameras", self._h, first, len(p), p.ctypes.data, v.ctypes.data)

    def transformations(self, scene: int) -> _TransformView:
        data = N.c_float_p()
        count = ctypes.c_uint64()
        N.call("gbr_transformations", self._h, scene, ctypes.byref(data), ctypes.byref(count))
        if count.value == 0:
            return _TransformView(np.zeros((0, 4, 4), np.float32))
        arr = np.ctypeslib.as_array(data, shape=(count.value, 4, 4))
        return _TransformView(arr)

    def scene_stats(self, scene: int) -> SceneStats:
        s = N.SceneStatsInfo()
        N.call("gbr_scene_stats", self._h, scene, ctypes.byref(s))
        return SceneStats(s.node_count, s.draw_count, s.draw_batch_count)

    # -- rendering ---------------------------------------------------------------
    def draw(self) -> None:
        N.call("gbr_draw", self._h)

    def synchronize(self) -> None:
        N.call("gbr_synchronize", self._h)

    def stream(self) -> int:
        out = ctypes.c_void_p()
        N.call("gbr_stream", self._h, ctypes.byref(out))
        return out.value or 0

    def _rect(self, rect):
        if rect is None:
            w, h = self.framebuffer_size
            rect = (0, 0, w, h)
        return np.asarray(rect, np.int32)

    def color_image_into(self, rect, image: np.ndarray) -> None:
        r = self._rect(rect)
        fmt = N.FORMAT_RGB8 if image.shape[-1] == 3 else N.FORMAT_RGBA8
        self._check_size(r, image)
        N.call("gbr_color_image_into", self._h, r.ctypes.data, image.ctypes.data,
               image.strides[0], fmt, N.MEMORY_HOST)

    def depth_image_into(self, rect, image: np.ndarray) -> None:
        r = self._rect(rect)
        self._check_size(r, image)
        N.call("gbr_depth_image_into", self._h, r.ctypes.data, image.ctypes.data,
               image.strides[0], N.MEMORY_HOST)

    def linear_depth_image_into(self, rect, image: np.ndarray) -> None:
        r = self._rect(rect)
        self._check_size(r, image)
        N.call("gbr_linear_depth_image_into", self._h, r.ctypes.data, image.ctypes.data,
               image.strides[0], N.MEMORY_HOST)

    def semantic_image_into(self, rect, image: np.ndarray) -> None:
        r = self._rect(rect)
        self._check_size(r, image)
        N.call("gbr_semantic_image_into", self._h, r.ctypes.data, image.ctypes.data,
               image.strides[0], N.MEMORY_HOST)

    @staticmethod
    def _check_size(r, image):
        if image.shape[0] != r[3] - r[1] or image.shape[1] != r[2] - r[0]:
            raise N.GbrError(N.ERR_SIZE, f"expected image size of {(r[2]-r[0], r[3]-r[1])} pixels "
                                         f"but got {(image.shape[1], image.shape[0])}")
        if not image.flags.writeable or image.strides[-1] != image.itemsize and image.ndim == 2:
            raise N.GbrError(N.ERR_ARGUMENT, "image must be writeable with contiguous rows")

    def color_image(self) -> np.ndarray:
        w, h = self.framebuffer_size
        out = np.empty((h, w, 4), np.uint8)
        self.color_image_into(None, out)
        return out

    def depth_image(self) -> np.ndarray:
        w, h = self.framebuffer_size
        out = np.empty((h, w), np.float32)
        self.depth_image_into(None, out)
        return out

    def linear_depth_image(self) -> np.ndarray:
        w, h = self.framebuffer_size
        out = np.empty((h, w), np.float32)
        self.linear_depth_image_into(None, out)
        return out

    def semantic_image(self) -> np.ndarray:
        w, h = self.framebuffer_size
        out = np.empty((h, w), np.uint32)
        self.semantic_image_into(None, out)
        return out

    def color_device_pointer(self) -> int:
        out = ctypes.c_void_p()
        N.call("gbr_color_device_pointer", self._h, ctypes.byref(out))
        return out.value

    def depth_device_pointer(self) -> int:
        out = ctypes.c_void_p()
        N.call("gbr_depth_device_pointer", self._h, ctypes.byref(out))
        return out.value

    def plane_device_pointer(self, plane: int) -> int:
        out = ctypes.c_void_p()
        N.call("gbr_plane_device_pointer", self._h, plane, ctypes.byref(out))
        return out.value

    def frame_visibility(self) -> dict:
        """Work items of the last frame owning a visible pixel, how many of
        them phase 1 set up, and the ratio (the proposal's recall)."""
        v, v1 = ctypes.c_uint64(), ctypes.c_uint64()
        N.call("gbr_frame_visibility", self._h, ctypes.byref(v), ctypes.byref(v1))
        return {"visible_items": v.value, "visible_in_phase1": v1.value,
                "recall": v1.value / v.value if v.value else 1.0}

    @property
    def device_id(self) -> int:
        """The HIP device ordinal the renderer runs on."""
        cfg = N.Config()
        N.call("gbr_get_config", self._h, ctypes.byref(cfg))
        return cfg.device

    def plane_dlpack(self, plane: int):
        """Zero-copy DLPack view of a whole output plane, one image per
        scene: colour (scenes, H, W, 4) uint8, depth and linear depth
        (scenes, H, W) float32, semantic (scenes, H, W) int32; valid until
        the next draw() (``torch.from_dlpack(r.plane_dlpack(PLANE_COLOR))``)."""
        from .dlpack import DeviceArray
        w, h = self.tile_size
        n = self.scene_count
        shape, dtype = {N.PLANE_COLOR: ((n, h, w, 4), np.uint8),
                        N.PLANE_DEPTH: ((n, h, w), np.float32),
                        N.PLANE_LINEAR_DEPTH: ((n, h, w), np.float32),
                        N.PLANE_SEMANTIC: ((n, h, w), np.int32)}[plane]
        return DeviceArray(self.plane_device_pointer(plane), shape, dtype, self.device_id, self,
                           sync=self.synchronize)

    # -- tuning / counters ---------------------------------------------------------
    def tuning(self) -> N.Tuning:
        t = N.Tuning()
        N.call("gbr_get_tuning", self._h, ctypes.byref(t))
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        N.call("gbr_set_tuning", self._h, ctypes.byref(t))

    def autotune(self, warmup: int = 2, frames: int = 4, passes: int = 1) -> dict:
        """Times draw() of the current scenes and cameras under candidate
        launch parameters and keeps the fastest (same pixels for all);
        returns the candidates timed, ms per frame before and after, and the
        tuning left set."""
        ap = N.AutotuneParams()
        N.call("gbr_autotune_params_init", ctypes.byref(ap))
        ap.warmup, ap.frames, ap.passes = warmup, frames, passes
        rep = N.AutotuneReport()
        N.call("gbr_autotune", self._h, ctypes.byref(ap), ctypes.byref(rep))
        best = {k: getattr(rep.best, k) for k, _ in rep.best._fields_}
        return {"candidates": rep.candidates, "initial_ms": rep.initial_ms,
                "best_ms": rep.best_ms, "best": best}

    # ---- learned visibility proposal ------------------------------------
    def vis_init(self, seed: int = 1) -> None:
        """Random visibility-model weights (He-uniform) from seed."""
        N.call("gbr_vis_init", self._h, seed)

    def vis_collect(self) -> int:
        """draw() with the current cameras and append one training row per
        chunk that passed the frustum test; returns the rows added."""
        n = ctypes.c_uint64()
        N.call("gbr_vis_collect", self._h, ctypes.byref(n))
        return n.value

    def vis_clear_rows(self) -> None:
        N.call("gbr_vis_clear_rows", self._h)

    def vis_set_feature_set(self, feature_set: int) -> None:
        """VIS_FEATURES_POSE (default: camera pose and the chunk's box
        relative to the camera) or VIS_FEATURES_HISTORY (ablation: earlier
        frames' visibility and the projected rectangle); switching drops the
        training rows."""
        N.call("gbr_vis_set_feature_set", self._h, feature_set)

    def vis_feature_set(self) -> int:
        out = ctypes.c_int32()
        N.call("gbr_vis_feature_set", self._h, ctypes.byref(out))
        return out.value

    def vis_rows(self) -> tuple:
        """(training rows, rows labelled visible)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        N.call("gbr_vis_rows", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def vis_fit(self, **kw) -> dict:
        """Adam on the training rows; keyword arguments override
        gbr_vis_fit_params_init() defaults (steps, batch, learning_rate,
        beta1, beta2, eps, pos_weight, seed)."""
        p = N.VisFitParams()
        N.call("gbr_vis_fit_params_init", ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        rep = N.VisFitReport()
        N.call("gbr_vis_fit", self._h, ctypes.byref(p), ctypes.byref(rep))
        return {k: getattr(rep, k) for k, _ in rep._fields_}

    def vis_weights(self) -> np.ndarray:
        out = np.empty(N.VIS_PARAMS, np.float32)
        N.call("gbr_vis_get_weights", self._h, out.ctypes.data)
        return out

    def set_vis_weights(self, w) -> None:
        w = np.ascontiguousarray(w, np.float32)
        if w.size != N.VIS_PARAMS:
            raise ValueError(f"expected {N.VIS_PARAMS} weights, got {w.size}")
        N.call("gbr_vis_set_weights", self._h, w.ctypes.data)

    def vis_loss_grad(self, x, labels, pos_weight: float = 1.0):
        """(mean weighted loss, gradient) of the model on caller rows."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(labels, np.uint8)
        loss = ctypes.c_float()
        grad = np.empty(N.VIS_PARAMS, np.float32)
        N.call("gbr_vis_loss_grad", self._h, x.ctypes.data, y.ctypes.data, len(y), pos_weight,
               ctypes.byref(loss), grad.ctypes.data)
        return loss.value, grad

    def vis_predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(len(x), np.float32)
        N.call("gbr_vis_predict", self._h, x.ctypes.data, len(x), out.ctypes.data)
        return out

    def frame_timings(self) -> dict:
        """Milliseconds per stage, averaged over the draw() calls since the
        previous call (needs tuning timing=1)."""
        ms = (ctypes.c_float * len(N.TIMING_STAGES))()
        N.call("gbr_frame_timings", self._h, ms)
        return dict(zip(N.TIMING_STAGES, list(ms)))

    def frame_stats(self) -> dict:
        s = N.FrameStatsInfo()
        N.call("gbr_frame_stats", self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}


def calculate_depth_unprojection(projection) -> np.ndarray:
    """calculateDepthUnprojection(), DepthUnprojection.cpp:98-101."""
    p = _mat16(projection)
    out = np.empty(2, np.float32)
    N.call("gbr_calculate_depth_unprojection", p.ctypes.data, out.ctypes.data)
    return out


def unproject_depth(unprojection, depth: np.ndarray) -> np.ndarray:
    """unprojectDepth(), DepthUnprojection.cpp:107-130 (in place on host
    memory; returns the array)."""
    u = np.ascontiguousarray(unprojection, np.float32)
    if depth.dtype != np.float32 or depth.ndim != 2 or depth.strides[1] != 4:
        raise ValueError("depth must be a 2D float32 array with contiguous rows")
    N.call("gbr_unproject_depth", u.ctypes.data, depth.ctypes.data, depth.shape[0],
           depth.shape[1], depth.strides[0])
    return depth
