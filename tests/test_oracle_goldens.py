"""The CPU oracle reproduces the reference's own PNG goldens and pixel /
depth known-answers (src/tests/GfxBatchRendererTest.cpp), pinning its GL
restatement before anything is compared against it."""
import pytest

import scenarios as S
from oracle.oracle import OracleRenderer, unproject_depth


def make(tile_size, tile_count, flags=0, max_lights=0):
    return OracleRenderer(tile_size, tile_count, flags, max_light_count=max_lights)


@pytest.mark.parametrize("case", list(S.FILE_DATA))
def test_defaults(case):
    files = S.FILE_DATA[case][0]
    checks, images = S.defaults(make, files)
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.FILE_DATA))
def test_single_mesh(case):
    files, _, _, thr = S.FILE_DATA[case]
    checks, images = S.single_mesh(make, files, thr)
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.MESH_HIERARCHY_DATA))
def test_mesh_hierarchy(case):
    checks, images = S.mesh_hierarchy(make, *S.MESH_HIERARCHY_DATA[case])
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.FILE_DATA))
def test_multiple_meshes(case):
    files, bc, _, thr = S.FILE_DATA[case]
    checks, images = S.multiple_meshes(make, files, bc, thr)
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.FILE_DATA))
def test_multiple_scenes(case):
    files, _, bcs, thr = S.FILE_DATA[case]
    checks, images = S.multiple_scenes(make, files, bcs, thr)
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.FILE_DATA))
def test_clear_scene(case):
    files, _, _, thr = S.FILE_DATA[case]
    checks, images = S.clear_scene(make, files, thr)
    S.assert_checks(checks)
    S.assert_images(images)


def test_lights_disabled():
    checks, images = S.lights_disabled(make)
    S.assert_checks(checks)
    S.assert_images(images)


@pytest.mark.parametrize("case", list(S.LIGHT_DATA))
def test_lights(case):
    checks, images = S.lights(make, *S.LIGHT_DATA[case], (0.75, 0.005))
    S.assert_checks(checks)
    S.assert_images(images)


def test_clear_lights():
    checks, images = S.clear_lights(make)
    S.assert_checks(checks)
    S.assert_images(images)


def test_depth_unprojection():
    checks, _ = S.depth_unprojection(make, unproject_depth)
    S.assert_checks(checks)
