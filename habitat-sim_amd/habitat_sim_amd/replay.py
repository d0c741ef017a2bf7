"""Batch replay-renderer surface: what habitat-sim's Python ``ReplayRenderer``
exposes for the batch path (src/esp/bindings/SimBindings.cpp:467-546,
src/esp/sim/AbstractReplayRenderer.{h,cpp}, BatchReplayRenderer.cpp), on top
of the MI355X renderer.

One environment = one scene + one camera = one tile of the renderer
(BatchReplayRenderer.cpp:34-36: the tile is the sensor resolution, the grid
is environment_grid_size(n)). All sensors of an environment share its
camera pose and resolution, as in the reference batch path; each sensor is
one output plane: colour RGBA8, depth (linear metres, far -> 0, the
reference's unprojectDepth()), semantic uint32.

``render(color_images, depth_images)`` fills caller arrays with GL rows
(bottom-up), like the reference (AbstractReplayRenderer.cpp:94-110 checks
the list sizes; an empty list skips that kind). ``get_observations()``
returns habitat's observation dict, rows flipped to top-down and dtypes as
habitat's sensor wrapper produces them (sensor_wrapper.py:120-219):
colour (H, W, 4) uint8, depth (H, W) float32, semantic (H, W) uint32.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from . import math3d as m3
from .replay_player import BatchPlayer, Keyframe, transform_matrix
from .renderer import RendererConfiguration, RendererStandalone, RendererStandaloneConfiguration

COLOR, DEPTH, SEMANTIC = "color", "depth", "semantic"
_PLANES = {COLOR: N.PLANE_COLOR, DEPTH: N.PLANE_LINEAR_DEPTH, SEMANTIC: N.PLANE_SEMANTIC}


def environment_grid_size(num_environments: int) -> tuple:
    """AbstractReplayRenderer::environmentGridSize (AbstractReplayRenderer.cpp:26-31):
    x = ceil(sqrt(n)), y = ceil(n / x)."""
    if num_environments < 1:
        raise ValueError("num_environments must be positive")
    x = int(math.ceil(math.sqrt(num_environments)))
    y = (num_environments + x - 1) // x
    return x, y


@dataclass
class Ray:
    """esp::geo::Ray (src/esp/geo/Geo.h:300-310): origin and direction."""
    origin: np.ndarray
    direction: np.ndarray


def _capsule(address: int):
    """A PyCapsule with no name around a device address, as pybind11's
    py::capsule(const void*) makes for the reference's pointer getters."""
    import ctypes
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    return new(ctypes.c_void_p(address), None, None)


def capsule_address(capsule) -> int:
    """The address a device-pointer capsule holds."""
    import ctypes
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get(capsule, None) or 0


@dataclass
class CameraSensorSpec:
    """The fields of habitat's CameraSensorSpec the batch path uses
    (src/esp/sensor/CameraSensor.cpp:45-66, VisualSensor.h:52-89)."""
    uuid: str
    sensor_type: str = COLOR               # "color" | "depth" | "semantic"
    resolution: Sequence[int] = (128, 128)  # (height, width), habitat order
    hfov: float = 90.0                     # degrees
    near: float = 0.01
    far: float = 1000.0

    def projection(self) -> np.ndarray:
        h, w = int(self.resolution[0]), int(self.resolution[1])
        return m3.pinhole_projection((h, w), math.radians(self.hfov), self.near, self.far)


@dataclass
class ReplayRendererConfiguration:
    """ReplayRendererConfiguration (AbstractReplayRenderer.h): number of
    environments, sensor specifications, device; plus this renderer's cull
    mode."""
    num_environments: int = 1
    sensor_specifications: List[CameraSensorSpec] = field(default_factory=list)
    gpu_device_id: int = -1
    cull_mode: int = N.CULL_NONE


class BatchReplayRenderer:
    def __init__(self, cfg: ReplayRendererConfiguration,
                 batch_renderer_configuration: Optional[RendererConfiguration] = None):
        """BatchReplayRenderer(cfg, gfx_batch::RendererConfiguration&&)
        (BatchReplayRenderer.cpp:20-56): the batch configuration (flags, max
        light count, ambient factor) gets the tile size and count from the
        sensors and the environment count."""
        if not cfg.sensor_specifications:
            raise ValueError("at least one sensor specification is needed")
        res = {tuple(int(v) for v in s.resolution) for s in cfg.sensor_specifications}
        if len(res) != 1:
            raise ValueError("all sensors of the batch path share one resolution")
        cams = {(s.hfov, s.near, s.far) for s in cfg.sensor_specifications}
        if len(cams) != 1:
            raise ValueError("all sensors of the batch path share one camera (hfov, near, far)")
        kinds = [s.sensor_type for s in cfg.sensor_specifications]
        for k in kinds:
            if k not in _PLANES:
                raise ValueError(f"unsupported sensor type {k!r}")
        if len(set(kinds)) != len(kinds):
            raise ValueError("one sensor per type")
        self.cfg = cfg
        self.sensors = {s.uuid: s for s in cfg.sensor_specifications}
        h, w = next(iter(res))
        self.resolution = (h, w)
        self.grid = environment_grid_size(cfg.num_environments)
        planes = 0
        for k in kinds:
            planes |= _PLANES[k]
        rc = (batch_renderer_configuration or RendererConfiguration()).set_tile_size_count(
            (w, h), self.grid)
        sc = RendererStandaloneConfiguration()
        if cfg.gpu_device_id >= 0:
            sc.set_cuda_device(cfg.gpu_device_id)
        self.renderer = RendererStandalone(rc, sc, planes=planes, top_down=True)
        if cfg.cull_mode != N.CULL_NONE:
            self.renderer.set_tuning(cull_mode=cfg.cull_mode)
        self._projection = cfg.sensor_specifications[0].projection()
        self._pose = [m3.identity() for _ in range(cfg.num_environments)]
        self._players: Dict[int, BatchPlayer] = {}
        self._depth_fb = None
        for e in range(cfg.num_environments):
            self.renderer.update_camera(e, self._projection, m3.identity())

    # ---- environments -------------------------------------------------------
    @staticmethod
    def create_batch_replay_renderer(cfg: ReplayRendererConfiguration) -> "BatchReplayRenderer":
        """ReplayRenderer.create_batch_replay_renderer(cfg) (SimBindings.cpp:
        478-483): the factory habitat-lab calls."""
        return BatchReplayRenderer(cfg)

    @property
    def num_environments(self) -> int:
        return self.cfg.num_environments

    @property
    def environment_count(self) -> int:
        """ReplayRenderer.environment_count (SimBindings.cpp:493-495,
        AbstractReplayRenderer::environmentCount)."""
        return self.cfg.num_environments

    def sensor_size(self, env: int) -> tuple:
        """ReplayRenderer.sensor_size(env) (SimBindings.cpp:496-497): the
        tile size (width, height), the same for every environment
        (BatchReplayRenderer::doSensorSize, BatchReplayRenderer.cpp:82-86).
        The environment index is checked as AbstractReplayRenderer::sensorSize
        asserts it (AbstractReplayRenderer.cpp:49-52)."""
        self._check_env(env)
        h, w = self.resolution
        return (w, h)

    @staticmethod
    def environment_grid_size(num_environments: int) -> tuple:
        return environment_grid_size(num_environments)

    def preload_file(self, path: str) -> bool:
        """AbstractReplayRenderer::preloadFile -> Renderer::addFile."""
        return self.renderer.add_file(path)

    def add_file_data(self, fd, name: Optional[str] = None) -> bool:
        return self.renderer.add_file_data(fd, name=name)

    def add_instance(self, env: int, name: str, transform=None, semantic_id: int = 0) -> int:
        """An instance of a named template (BatchPlayerImplementation.cpp:51-91);
        returns its node id."""
        self._check_env(env)
        return self.renderer.add_node_hierarchy(env, name, transform, semantic_id)

    def set_instance_transform(self, env: int, node: int, transform) -> None:
        """BatchPlayerImplementation::setNodeTransform (:106-118)."""
        self._check_env(env)
        t = self.renderer.transformations(env)
        t[node] = np.asarray(transform, np.float32)

    def clear_environment(self, env: int) -> None:
        """AbstractReplayRenderer::clearEnvironment (AbstractReplayRenderer.cpp:
        54-58): closes the environment's player, then Renderer::clear."""
        self._check_env(env)
        player = self._players.get(env)
        if player is not None:
            player.close()
        self.renderer.clear(env)

    # ---- gfx-replay keyframes -----------------------------------------------
    def player_for(self, env: int) -> BatchPlayer:
        """The environment's keyframe player (BatchReplayRenderer::doPlayerFor),
        made on first use."""
        self._check_env(env)
        player = self._players.get(env)
        if player is None:
            player = self._players[env] = BatchPlayer(self.renderer, env)
        return player

    def set_environment_keyframe(self, env: int, ser_keyframe: str) -> None:
        """AbstractReplayRenderer::setEnvironmentKeyframe (:60-66): a JSON
        document whose "keyframe" member is the keyframe
        (Player::keyframeFromString, Player.cpp:105-111); replaces the
        player's keyframes with it and applies it on top of the current
        frame (Player::setSingleKeyframe, Player.cpp:361-366)."""
        doc = json.loads(ser_keyframe)
        if not isinstance(doc, dict) or "keyframe" not in doc:
            raise ValueError('set_environment_keyframe: no "keyframe" member')
        self.player_for(env).set_single_keyframe(Keyframe.from_json(doc["keyframe"]))

    def set_environment_keyframe_unwrapped(self, env: int, ser_keyframe: str) -> None:
        """AbstractReplayRenderer::setEnvironmentKeyframeUnwrapped (:68-74):
        the document is the keyframe itself (keyframeFromStringUnwrapped,
        Player.cpp:113-120)."""
        self.player_for(env).set_single_keyframe(Keyframe.from_json(ser_keyframe))

    def set_sensor_transforms_from_keyframe(self, env: int, prefix: str) -> None:
        """AbstractReplayRenderer::setSensorTransformsFromKeyframe (:83-92) and
        BatchReplayRenderer::doSetSensorTransformsFromKeyframe
        (BatchReplayRenderer.cpp:100-114): the camera pose is the keyframe's
        user transform named prefix + sensor uuid.  The sensors of one
        environment share its camera, so the first sensor (in specification
        order) whose name is in the keyframe gives it."""
        self._check_env(env)
        player = self._players.get(env)
        if player is None or player.num_keyframes != 1:
            raise RuntimeError(f"set_sensor_transforms_from_keyframe: for environment {env}, "
                               "you have not yet called set_environment_keyframe.")
        for spec in self.cfg.sensor_specifications:
            found = player.get_user_transform(prefix + spec.uuid)
            if found is not None:
                translation, rotation = found
                self.set_sensor_transform(env, spec.uuid,
                                          transform_matrix(translation, rotation))
                return
        first = prefix + self.cfg.sensor_specifications[0].uuid
        raise KeyError(f'set_sensor_transforms_from_keyframe: couldn\'t find user transform '
                       f'"{first}" for environment {env}.')

    def set_sensor_transform(self, env: int, sensor_uuid: str, transform) -> None:
        """Camera-to-world pose of the environment's sensors
        (BatchReplayRenderer::doSetSensorTransform, :92-99: the view is the
        inverted transform)."""
        self._check_env(env)
        if sensor_uuid not in self.sensors:
            raise KeyError(sensor_uuid)
        pose = np.asarray(transform, np.float32)
        self._pose[env] = pose
        self.renderer.update_camera(env, self._projection, m3.inverted(pose))

    def set_sensor_transforms(self, transforms) -> None:
        """All environments' camera-to-world poses at once, (N, 4, 4)."""
        poses = np.asarray(transforms, np.float32).reshape(-1, 4, 4)
        if len(poses) != self.num_environments:
            raise ValueError("one pose per environment")
        views = np.stack([m3.inverted(p) for p in poses])
        self._pose = list(poses)
        self.renderer.update_cameras(0, np.broadcast_to(self._projection, views.shape), views)

    # ---- rendering ----------------------------------------------------------
    def render(self, color_images: Sequence[np.ndarray] = (),
               depth_images: Sequence[np.ndarray] = ()) -> None:
        """AbstractReplayRenderer::render(colorImageViews, depthImageViews):
        draw all environments, then copy each environment's tile into the
        caller's arrays (GL rows, bottom-up); an empty list skips that kind,
        otherwise there must be one array per environment
        (AbstractReplayRenderer.cpp:94-110). Depth is linear (unprojected)."""
        n = self.num_environments
        for lst, what in ((color_images, "color"), (depth_images, "depth")):
            if len(lst) and len(lst) != n:
                raise ValueError(f"expected {n} {what} images, got {len(lst)}")
        if len(color_images) and COLOR not in self._kinds():
            raise ValueError("no color sensor configured")
        if len(depth_images) and DEPTH not in self._kinds():
            raise ValueError("no depth sensor configured")
        self.renderer.draw()
        h, w = self.resolution
        gx = self.grid[0]
        for e in range(n):
            x0, y0 = (e % gx) * w, (e // gx) * h
            rect = (x0, y0, x0 + w, y0 + h)
            if len(color_images):
                self.renderer.color_image_into(rect, color_images[e])
            if len(depth_images):
                self.renderer.linear_depth_image_into(rect, depth_images[e])

    def draw(self) -> None:
        """Draw without a copy; read with get_observations() or the device
        pointers."""
        self.renderer.draw()

    def get_observations(self, env: Optional[int] = None) -> Dict[str, np.ndarray]:
        """Habitat's observation dict for one environment, or batched over
        all environments with a leading env axis (env=None): rows top-down,
        colour (H, W, 4) uint8, depth (H, W) float32, semantic (H, W) uint32."""
        self.renderer.synchronize()
        h, w = self.resolution
        n = self.num_environments
        out = {}
        for uuid, s in self.sensors.items():
            plane = _PLANES[s.sensor_type]
            full = self._plane_host(plane, s.sensor_type)
            out[uuid] = full[:n] if env is None else full[env]
        return out

    def observations_dlpack(self) -> Dict[str, object]:
        """Zero-copy DLPack views of every sensor's plane over all
        environments (leading env axis, rows top-down; semantic as int32),
        the analogue of the reference's gpu2gpu_transfer observations
        (sensor_wrapper.py:97-124); valid until the next draw()."""
        n = self.num_environments
        out = {}
        for uuid, s in self.sensors.items():
            full = self.renderer.plane_dlpack(_PLANES[s.sensor_type])
            full.shape = (n,) + full.shape[1:]  # the grid's trailing empty tiles dropped
            out[uuid] = full
        return out

    def get_observations_torch(self, env: Optional[int] = None, copy: bool = False):
        """get_observations() as torch tensors on the GPU, aliasing the
        renderer's planes (copy=True clones them); torch's GPU must be
        initialised before libgbr loads."""
        import torch
        out = {}
        for uuid, arr in self.observations_dlpack().items():
            t = torch.from_dlpack(arr)
            t = t if env is None else t[env]
            out[uuid] = t.clone() if copy else t
        return out

    def _plane_host(self, plane: int, kind: str) -> np.ndarray:
        from . import hip
        h, w = self.resolution
        n = self.num_environments
        dtype, shape = {COLOR: (np.uint8, (n, h, w, 4)), DEPTH: (np.float32, (n, h, w)),
                        SEMANTIC: (np.uint32, (n, h, w))}[kind]
        out = np.empty(shape, dtype)
        ptr = self.renderer.plane_device_pointer(plane)
        hip.memcpy_d2h(out, ptr)
        return out

    def cuda_color_buffer_device_pointer(self):
        """ReplayRenderer.cuda_color_buffer_device_pointer() (SimBindings.cpp:
        531-536): a capsule holding the device address of the colour
        framebuffer (RGBA8, tile grid, rows bottom-up, the layout of
        RendererStandalone::colorCudaBufferDevicePointer), valid until the
        next render()/draw()."""
        return _capsule(self.renderer.color_device_pointer())

    def cuda_depth_buffer_device_pointer(self):
        """ReplayRenderer.cuda_depth_buffer_device_pointer() (SimBindings.cpp:
        537-542): a capsule holding the device address of the depth
        framebuffer (float32, tile grid, rows bottom-up), valid until the
        next render()/draw().  The reference binding returns the colour
        pointer here by mistake; this returns depth.  The values are the
        depth sensor's: linear metres with far -> 0 (what render() writes),
        the window depth when the renderer keeps no linear plane."""
        if DEPTH not in self._kinds():
            return _capsule(self.renderer.depth_device_pointer())
        w, h = self.renderer.framebuffer_size
        if self._depth_fb is None:
            from . import hip
            self._depth_fb = hip.DeviceBuffer(w * h * 4)
        rect = np.asarray((0, 0, w, h), np.int32)
        N.call("gbr_linear_depth_image_into", self.renderer._h, rect.ctypes.data,
               self._depth_fb.value, w * 4, N.MEMORY_DEVICE)
        return _capsule(self._depth_fb.value)

    def unproject(self, env: int, viewport_position) -> Ray:
        """ReplayRenderer.unproject(env, viewport_position) (SimBindings.cpp:
        544-545, AbstractReplayRenderer::unproject :137-141): the world-space
        ray from the environment's camera through a viewport pixel given in
        top-down window coordinates.  The batch reference returns a
        placeholder here (BatchReplayRenderer::doUnproject,
        BatchReplayRenderer.cpp:161-171); this follows the classic renderer's
        RenderCamera::unproject (RenderCamera.cpp:186-216): origin = camera
        position, direction = (inverse(P) of the pixel's far-plane NDC point,
        through the camera pose) - origin, normalised."""
        self._check_env(env)
        h, w = self.resolution
        x, y = int(viewport_position[0]), int(viewport_position[1])
        ndc = np.array([2.0 * x / w - 1.0, 2.0 * (h - y - 1) / h - 1.0, 1.0, 1.0])
        pose = np.asarray(self._pose[env], np.float64)
        inv_p = np.linalg.inv(np.asarray(self._projection, np.float64).T)
        q = inv_p @ ndc
        q = q[:3] / q[3]
        world = (pose.T @ np.append(q, 1.0))[:3]
        origin = pose[3, :3].astype(np.float64)
        d = world - origin
        n = np.linalg.norm(d)
        return Ray(origin.astype(np.float32), (d / n if n > 0 else d).astype(np.float32))

    def plane_device_pointer(self, sensor_uuid: str) -> int:
        """Device address of a sensor's env-major plane [env][H][W] (rows
        top-down), valid until the next draw() (the analogue of
        getCudaColorBufferDevicePointer, BatchReplayRenderer.cpp:162-185)."""
        return self.renderer.plane_device_pointer(_PLANES[self.sensors[sensor_uuid].sensor_type])

    def close(self) -> None:
        self._players.clear()
        if self._depth_fb is not None:
            self._depth_fb.close()
            self._depth_fb = None
        self.renderer.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- helpers ------------------------------------------------------------
    def _kinds(self):
        return {s.sensor_type for s in self.sensors.values()}

    def _check_env(self, env: int) -> None:
        if not 0 <= env < self.num_environments:
            raise IndexError(f"environment {env} out of range for {self.num_environments}")
