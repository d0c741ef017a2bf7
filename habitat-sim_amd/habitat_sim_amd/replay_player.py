"""gfx-replay keyframe player for the batch renderer.

Plays habitat-sim's gfx-replay keyframes (the JSON of
src/esp/io/JsonEspTypes.cpp:104-200: ``loads``, ``creations``,
``deletions``, ``metadata``, ``stateUpdates``, ``userTransforms``,
``lightsChanged``/``lights``) into one scene of a renderer, with the
semantics of the reference's ``gfx::replay::Player``
(src/esp/gfx/replay/Player.cpp:152-330) driving
``BatchPlayerImplementation`` (src/esp/sim/BatchPlayerImplementation.cpp):

* an instance of ``filepath`` is a node hierarchy of that name; a name not
  yet known is added with ``addFile(filepath, Whole | GenerateMipmap)``;
  primitive render assets (``cubeSolid...``) are unsupported and skipped;
* the creation bakes ``scaling(scale) * rotation(asset frame to world)``
  into the hierarchy; state updates write ``from(rotation, translation)``
  into the instance's top-level transformation;
* any deletion in a keyframe clears the scene and re-creates the surviving
  instances with their latest transformations (the batch renderer cannot
  remove single hierarchies), so node ids stay dense;
* going back to an earlier keyframe clears the scene and replays from the
  first one;
* ``lightsChanged`` replaces the scene's lights (``changeLightSetup``); a
  deletion's clear removes them until the next light change, as in the
  reference.

Difference from the reference: instance metadata ``semanticId`` is forwarded
to the renderer's semantic plane (the reference's batch player drops it).
"""
from __future__ import annotations

import json
import math
import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from . import math3d as m3

FILE_WHOLE = 1
FILE_GENERATE_MIPMAP = 2
LIGHT_DIRECTIONAL, LIGHT_POINT = 0, 1  # RendererLightType

_PRIMITIVE_PREFIXES = ("capsule3DSolid", "capsule3DWireframe", "coneSolid", "coneWireframe",
                       "cubeSolid", "cubeWireframe", "cylinderSolid", "cylinderWireframe",
                       "icosphereSolid", "icosphereWireframe", "uvSphereSolid",
                       "uvSphereWireframe")


def is_supported_render_asset(filepath: str) -> bool:
    """BatchPlayerImplementation.cpp:10-37: primitives are not batch assets."""
    return not filepath.startswith(_PRIMITIVE_PREFIXES)


def quaternion_matrix(q_wxyz) -> np.ndarray:
    """3x3 rotation of a unit quaternion given as [w, x, y, z] (the JSON
    order, JsonMagnumTypes.cpp:133-158), [col][row], float32."""
    w, x, y, z = (np.float32(v) for v in q_wxyz)
    two = np.float32(2.0)
    one = np.float32(1.0)
    return np.array([
        [one - two * y * y - two * z * z, two * x * y + two * z * w, two * x * z - two * y * w],
        [two * x * y - two * z * w, one - two * x * x - two * z * z, two * y * z + two * x * w],
        [two * x * z + two * y * w, two * y * z - two * x * w, one - two * x * x - two * y * y]],
        np.float32)


def transform_matrix(translation, rotation_wxyz) -> np.ndarray:
    """Matrix4::from(rotation.toMatrix(), translation), [col][row]."""
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = quaternion_matrix(rotation_wxyz)
    m[3, :3] = np.asarray(translation, np.float32)
    return m


def _quat_rotation(a, b) -> np.ndarray:
    """Shortest-arc unit quaternion [w, x, y, z] turning unit vector a into b."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = float(np.dot(a, b))
    if d < -1.0 + 1e-7:  # opposite: half turn about any axis orthogonal to a
        axis = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, [0.0, 1.0, 0.0])
        axis /= np.linalg.norm(axis)
        return np.array([0.0, *axis])
    c = np.cross(a, b)
    q = np.array([1.0 + d, *c])
    return q / np.linalg.norm(q)


def _quat_mul(p, q) -> np.ndarray:
    w1, x1, y1, z1 = p
    w2, x2, y2, z2 = q
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _rotate(q, v) -> np.ndarray:
    p = np.array([0.0, *v])
    qc = np.array([q[0], -q[1], -q[2], -q[3]])
    return _quat_mul(_quat_mul(q, p), qc)[1:]


def frame_to_world(frame: Optional[dict]) -> np.ndarray:
    """CoordinateFrame::rotationFrameToWorld() (geo/CoordinateFrame.cpp:
    27-43) of an asset frame {"up", "front"}; world up +Y, front -Z.  Exact
    for the default frame (identity), float64 otherwise."""
    up_w, front_w = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, -1.0])
    if not frame:
        return np.array([1.0, 0.0, 0.0, 0.0])
    up = np.asarray(frame.get("up", up_w), np.float64)
    front = np.asarray(frame.get("front", front_w), np.float64)
    if np.allclose(up, up_w) and np.allclose(front, front_w):
        return np.array([1.0, 0.0, 0.0, 0.0])
    r_up = _quat_rotation(up_w, up)
    w2f = _quat_mul(_quat_rotation(_rotate(r_up, front_w), front), r_up)
    w2f /= np.linalg.norm(w2f)
    return np.array([w2f[0], -w2f[1], -w2f[2], -w2f[3]])  # inverse of a unit quaternion


@dataclass
class Keyframe:
    loads: List[dict] = field(default_factory=list)
    creations: List[Tuple[int, dict]] = field(default_factory=list)
    deletions: List[int] = field(default_factory=list)
    metadata: List[Tuple[int, dict]] = field(default_factory=list)
    state_updates: List[Tuple[int, Tuple[list, list]]] = field(default_factory=list)
    user_transforms: Dict[str, Tuple[list, list]] = field(default_factory=dict)
    lights_changed: bool = False
    lights: List[dict] = field(default_factory=list)

    @staticmethod
    def from_json(obj: Union[str, dict]) -> "Keyframe":
        """JsonEspTypes.cpp:104-200."""
        if isinstance(obj, str):
            obj = json.loads(obj)
        k = Keyframe()
        k.loads = list(obj.get("loads", []))
        k.creations = [(int(c["instanceKey"]), dict(c["creation"])) for c in obj.get("creations", [])]
        k.deletions = [int(d) for d in obj.get("deletions", [])]
        k.metadata = [(int(m["instanceKey"]), dict(m["metadata"])) for m in obj.get("metadata", [])]
        for s in obj.get("stateUpdates", []):
            t = s["state"]["absTransform"]
            k.state_updates.append((int(s["instanceKey"]), (t["translation"], t["rotation"])))
        for u in obj.get("userTransforms", []):
            t = u["transform"]
            k.user_transforms[u["name"]] = (t["translation"], t["rotation"])
        k.lights_changed = bool(obj.get("lightsChanged", False))
        if k.lights_changed:
            k.lights = list(obj.get("lights", []))
        return k


def load_keyframes(path: str) -> List[Keyframe]:
    """A gfx-replay file: {"keyframes": [...]}."""
    with open(path) as f:
        obj = json.load(f)
    return [Keyframe.from_json(k) for k in obj.get("keyframes", [])]


@dataclass
class _Record:
    creation: dict
    metadata: dict


class BatchPlayer:
    """Player + BatchPlayerImplementation for scene `scene` of `renderer`."""

    def __init__(self, renderer, scene: int):
        self.renderer = renderer
        self.scene = scene
        self.keyframes: List[Keyframe] = []
        self.frame_index = -1
        self._asset_infos: Dict[str, dict] = {}
        self._created: Dict[int, int] = {}       # instance key -> top-level node id
        self._records: Dict[int, _Record] = {}
        self._failed = set()

    # ---- Player ---------------------------------------------------------------
    def set_keyframes(self, keyframes: List[Keyframe]) -> None:
        self.clear_frame()
        self.keyframes = list(keyframes)

    def append_keyframe(self, keyframe: Union[Keyframe, str, dict]) -> None:
        if not isinstance(keyframe, Keyframe):
            keyframe = Keyframe.from_json(keyframe)
        self.keyframes.append(keyframe)

    def set_single_keyframe(self, keyframe: Union[Keyframe, str, dict]) -> None:
        self.keyframes = []
        self.frame_index = -1
        self.append_keyframe(keyframe)
        self.set_keyframe_index(0)

    @property
    def num_keyframes(self) -> int:
        return len(self.keyframes)

    def set_keyframe_index(self, index: int) -> None:
        if not (index == -1 or 0 <= index < len(self.keyframes)):
            raise IndexError(f"keyframe {index} out of range for {len(self.keyframes)}")
        if index < self.frame_index:
            self.clear_frame()
        while self.frame_index < index:
            self.frame_index += 1
            self._apply(self.keyframes[self.frame_index])

    def get_user_transform(self, name: str):
        """(translation, rotation [w, x, y, z]) of the current keyframe, or None."""
        if self.frame_index < 0:
            raise RuntimeError("no keyframe applied")
        return self.keyframes[self.frame_index].user_transforms.get(name)

    def clear_frame(self) -> None:
        if self._created:
            self.renderer.clear(self.scene)
        self._created.clear()
        self._asset_infos.clear()
        self._records.clear()
        self.frame_index = -1

    def close(self) -> None:
        self.clear_frame()
        self.keyframes = []

    # ---- keyframe application -------------------------------------------------
    def _apply(self, k: Keyframe) -> None:
        for info in k.loads:
            if info.get("filepath") not in self._failed:
                self._asset_infos[info.get("filepath")] = info
        meta = dict(k.metadata)
        for key, creation in k.creations:
            path = creation.get("filepath", "").split("?", 1)[0]
            if path not in self._asset_infos:
                self._failed.add(path)
                continue
            creation = dict(creation, filepath=path)
            node = self._create(self._asset_infos[path], creation, meta.get(key, {}))
            if node is None:
                self._failed.add(path)
                continue
            if key in self._created:
                raise RuntimeError(f"instance key {key} created twice")
            self._created[key] = node
            self._records[key] = _Record(creation, meta.get(key, {}))
        for key, md in k.metadata:
            if key in self._created:
                self._records[key].metadata = md
                self._set_metadata(self._created[key], md)
        self._deletions(k)
        for key, (t, r) in k.state_updates:
            if key in self._created:
                self._set_transform(self._created[key], transform_matrix(t, r))
        if k.lights_changed:
            self._change_light_setup(k.lights)

    def _deletions(self, k: Keyframe) -> None:
        """Player::hackProcessDeletions, batch branch (Player.cpp:291-325)."""
        if not k.deletions:
            return
        latest = {key: np.array(self.renderer.transformations(self.scene)[node])
                  for key, node in self._created.items()}
        self.renderer.clear(self.scene)
        for d in k.deletions:
            self._created.pop(d, None)
            self._records.pop(d, None)
        for key in list(self._created):
            rec = self._records[key]
            node = self._create(self._asset_infos[rec.creation["filepath"]], rec.creation,
                                rec.metadata)
            self._created[key] = node
            self._set_transform(node, latest[key])

    # ---- BatchPlayerImplementation ------------------------------------------
    def _create(self, info: dict, creation: dict, metadata: dict) -> Optional[int]:
        path = creation["filepath"]
        if not is_supported_render_asset(path):
            return None
        r = self.renderer
        if not r.has_node_hierarchy(path):
            if not r.add_file(path, FILE_WHOLE | FILE_GENERATE_MIPMAP):
                raise RuntimeError(f"addFile failed for {path}")
        scale = creation.get("scale")
        s = m3.scaling(tuple(scale) if scale is not None else (1.0, 1.0, 1.0))
        q = frame_to_world(info.get("frame"))
        rot = np.eye(4, dtype=np.float32)
        rot[:3, :3] = quaternion_matrix(q)
        bake = m3.mul(s, rot)
        sem = metadata.get("semanticId", -1)
        return r.add_node_hierarchy(self.scene, path, bake, max(int(sem), 0))

    def _change_light_setup(self, lights: List[dict]) -> None:
        """BatchPlayerImplementation::changeLightSetup (:127-170): the
        scene's lights are replaced; each light gets a new empty node, a
        point light (vector w != 0) translated to vector.xyz, a directional
        one looking along vector.xyz (lookAt with +Y up; only its Z axis
        matters, NaNs in the others are harmless).  Range stays infinite."""
        r = self.renderer
        if not r.max_light_count:
            warnings.warn(f"Attempted to change {len(lights)} lights for scene {self.scene} "
                          "but the renderer is configured without lights")
            return
        r.clear_lights(self.scene)
        for light in lights:
            model = str(light.get("model", "global")).lower()
            if model != "global":
                raise RuntimeError(f"light position model {model!r} is not global")
            v = [float(x) for x in light.get("vector", (0.0, 0.0, 1.0, 0.0))]
            node = r.add_empty_node(self.scene)
            if v[3]:
                xf, kind = m3.translation(tuple(v[:3])), LIGHT_POINT
            else:
                with np.errstate(invalid="ignore", divide="ignore"):
                    xf = m3.look_at((0.0, 0.0, 0.0), tuple(v[:3]), (0.0, 1.0, 0.0))
                kind = LIGHT_DIRECTIONAL
            r.transformations(self.scene)[node] = xf
            lid = r.add_light(self.scene, node, kind)
            r.light_colors(self.scene)[lid] = np.asarray(light.get("color", (1.0, 1.0, 1.0)),
                                                         np.float32)[:3]

    def _set_metadata(self, node: int, md: dict) -> None:
        sem = int(md.get("semanticId", -1))
        self.renderer.set_node_hierarchy_semantic(self.scene, node, max(sem, 0))

    def _set_transform(self, node: int, m: np.ndarray) -> None:
        self.renderer.transformations(self.scene)[node] = np.asarray(m, np.float32)
