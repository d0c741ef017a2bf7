"""The mip-mapped filtering cases of tests/mip_cases.py through the HIP
renderer: bit-exact with the oracle and within the GL closed form, so the
GPU's LOD selection and blend are pinned to GL 4.6 section 8.14, not only to
the oracle."""
import numpy as np
import pytest

import mip_cases as MC
import synth
from habitat_sim_amd import RendererConfiguration, RendererStandalone
from habitat_sim_amd import math3d as m3
from test_oracle_mip import render_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", MC.CASES, ids=MC.case_id)
def test_gpu_mip_filtering(case):
    tile, meshes, mats, texs, expected, tol = MC.scene(case)
    r = RendererStandalone(RendererConfiguration().set_tile_size_count((tile, tile), (1, 1)))
    assert r.add_file_data(synth.mesh_file(meshes, mats, texs))
    r.add_node_hierarchy(0, "m0", None, 1)
    r.update_camera(0, m3.identity(), m3.identity())
    r.draw()
    got = r.color_image()
    r.close()
    ref, _, _ = render_case(case)
    np.testing.assert_array_equal(got, ref)
    assert MC.check(got, expected, tol)
