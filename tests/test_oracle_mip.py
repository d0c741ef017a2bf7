"""Pins the oracle's mip-mapped texture filtering to GL 4.6 section 8.14
with closed forms (tests/mip_cases.py): level choice of *_MIPMAP_NEAREST at
lambda = k +- 0.4, the (1 - frac) / frac blend of *_MIPMAP_LINEAR at
lambda = k + {0.25, 0.5, 0.75}, rho as the larger of the two screen axes
(anisotropic mapping), and rho from both texture axes (rotated mapping).
The reference enables mip-mapped minification (Renderer.cpp:452-456) and
its own tests leave the filtering unchecked (GfxBatchRendererTest.cpp:
106-107,125-126).  tests/test_oracle_mutants.py checks that each of these
fails under a planted mistake in the oracle.

Also the oracle's stand-in for glGenerateMipmap(): its level count is the
reference's (log2 of the smaller edge + 1, Renderer.cpp:446-448,493-495),
its level sizes halve, constant and 2x2-block-constant images filter to
themselves and a linear ramp to its exact average.  How the GL driver rounds
a 2x2 mean that falls between two 8-bit values is not specified (GL 4.6
section 8.14.4 leaves the filter to the implementation), so that rounding
stays unpinned."""
import numpy as np
import pytest

import mip_cases as MC
from habitat_sim_amd import math3d as m3
from oracle.oracle import OracleRenderer, generate_mipmaps
import synth


def render_case(case):
    tile, meshes, mats, texs, expected, tol = MC.scene(case)
    r = OracleRenderer((tile, tile), (1, 1))
    assert r.add_file_data(synth.mesh_file(meshes, mats, texs))
    r.add_node_hierarchy(0, "m0", None, 1)
    r.update_camera(0, m3.identity(), m3.identity())
    out = r.render(want_count=True, want_ids=True)
    assert out["count"][0].min() == 1  # the quad covers every pixel once
    return out["color"][0], expected, tol


@pytest.mark.parametrize("case", MC.CASES, ids=MC.case_id)
def test_mip_filtering_closed_form(case):
    color, expected, tol = render_case(case)
    assert MC.check(color, expected, tol), (
        f"got red {np.unique(color[..., 0])}, green {np.unique(color[..., 1])}, "
        f"expected {expected}")


@pytest.mark.parametrize("w,h,layers", [(16, 16, 1), (32, 8, 2), (8, 32, 1), (5, 12, 1), (1, 1, 3)])
def test_generated_mip_chain_shape(w, h, layers):
    rng = np.random.default_rng(w * 100 + h)
    img = rng.integers(0, 256, (layers, h, w, 4), dtype=np.uint8)
    levels = generate_mipmaps(img)
    assert len(levels) == int(np.floor(np.log2(min(w, h)))) + 1
    for i, lv in enumerate(levels):
        assert lv.shape == (layers, max(h >> i, 1), max(w >> i, 1), 4), i
    np.testing.assert_array_equal(levels[0], img)


def test_generated_mips_of_block_constant_images():
    # 2^k x 2^k blocks of one colour stay that colour on levels 0..k
    rng = np.random.default_rng(7)
    base = rng.integers(0, 256, (2, 4, 4, 4), dtype=np.uint8)
    img = base.repeat(8, axis=1).repeat(8, axis=2)  # 32x32, 8x8 blocks
    levels = generate_mipmaps(img)
    for k in range(4):
        b = 8 >> k
        np.testing.assert_array_equal(levels[k], base.repeat(b, axis=1).repeat(b, axis=2))


def test_generated_mips_average_a_ramp_exactly():
    # texel x of level 0 = 4x: a 2x2 mean of (4x, 4x + 4) is 4x + 2, exact
    img = np.zeros((1, 8, 64, 4), np.uint8)
    img[0, :, :, 0] = (4 * np.arange(64)).astype(np.uint8)
    img[..., 3] = 255
    lv1 = generate_mipmaps(img)[1]
    np.testing.assert_array_equal(lv1[0, :, :, 0], np.broadcast_to(8 * np.arange(32) + 2, (4, 32)))
