"""Tiny synthetic FileData builders for rule-level tests (fill rule, depth
function, snapping, clipping, texture filtering)."""
from __future__ import annotations

import numpy as np

from habitat_sim_amd.assets import (FileData, MaterialData, MeshData, MeshEntry, NodeData,
                                    TextureData, GL_NEAREST, GL_REPEAT)


def mesh_file(meshes, materials=None, textures=None, names=None):
    """meshes: list of (positions (V,3), indices (3T,), texcoords or None,
    colors or None, material index).  Each mesh becomes a named template
    ("m0", "m1", ...) with one child mesh view covering it."""
    fd = FileData()
    fd.textures = list(textures or [])
    fd.materials = list(materials or [])
    names = names or [f"m{i}" for i in range(len(meshes))]
    for i, (pos, idx, uv, col, mat) in enumerate(meshes):
        fd.meshes.append(MeshData(positions=np.asarray(pos, np.float32),
                                  indices=np.asarray(idx, np.uint32),
                                  texcoords=None if uv is None else np.asarray(uv, np.float32),
                                  colors=None if col is None else np.asarray(col, np.float32)))
        root = len(fd.nodes)
        fd.nodes.append(NodeData(parent=-1, name=names[i]))
        fd.nodes.append(NodeData(parent=root))
        fd.mesh_entries.append(MeshEntry(node=root + 1, mesh=i, material=mat))
    return fd


def flat(rgba, texture=-1, layer=0, matrix=None):
    m = MaterialData(flat=True, base_color=tuple(float(c) for c in rgba), texture=texture,
                     layer=layer)
    if matrix is not None:
        m.texture_matrix = np.asarray(matrix, np.float32)
    return m


def texture(levels, min_filter=GL_NEAREST, mag_filter=GL_NEAREST, wrap=GL_REPEAT):
    lv = [np.asarray(x, np.uint8) for x in levels]
    if lv[0].ndim == 3:
        lv = [x[None] for x in lv]
    return TextureData(width=lv[0].shape[2], height=lv[0].shape[1], layers=lv[0].shape[0],
                       levels=lv, min_filter=min_filter, mag_filter=mag_filter,
                       wrap_s=wrap, wrap_t=wrap)


def window_to_ndc(xw, yw, w, h):
    """Window coordinate -> NDC for an identity camera (exact for dyadic
    inputs and power-of-two tile sizes)."""
    return (np.float32(xw) - np.float32(w / 2)) / np.float32(w / 2), \
           (np.float32(yw) - np.float32(h / 2)) / np.float32(h / 2)
