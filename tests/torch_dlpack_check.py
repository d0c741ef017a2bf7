"""Run as its own process by test_gpu_dlpack.py (torch's GPU is initialised
before libgbr loads, as torch brings its own HIP runtime): the replay
renderer's observations exported through DLPack alias the renderer's planes
and equal the host observations; a later draw shows through the views."""
import os
import sys

import numpy as np
import torch

torch.cuda.init()
torch.zeros(1, device="cuda")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [ROOT, os.path.join(ROOT, "habitat-sim_amd")]

from habitat_sim_amd import _native as N  # noqa: E402
from habitat_sim_amd import math3d as m3  # noqa: E402
from habitat_sim_amd import replay, scenes  # noqa: E402

spec = scenes.small_rooms(n_envs=5, tile=64)
S = replay.CameraSensorSpec
cfg = replay.ReplayRendererConfiguration(
    5, [S("rgba", "color", (64, 64)), S("depth", "depth", (64, 64)),
        S("semantic", "semantic", (64, 64))])
rr = replay.BatchReplayRenderer(cfg)
assert rr.add_file_data(spec.file)
for e, env in enumerate(spec.envs):
    for inst in env.instances:
        rr.add_instance(e, inst.template, inst.bake, inst.semantic)
poses = np.stack([env.camera_pose for env in spec.envs])
rr.set_sensor_transforms(poses)
rr.draw()
host = rr.get_observations()
dev = rr.get_observations_torch()
planes = {"rgba": N.PLANE_COLOR, "depth": N.PLANE_LINEAR_DEPTH, "semantic": N.PLANE_SEMANTIC}
for uuid, t in dev.items():
    assert t.is_cuda and t.shape[0] == 5, (uuid, t.shape)
    assert t.data_ptr() == rr.renderer.plane_device_pointer(planes[uuid]), uuid  # zero-copy
    got = t.cpu().numpy()
    if uuid == "semantic":
        got = got.view(np.uint32)
    np.testing.assert_array_equal(got, host[uuid], err_msg=uuid)
one = rr.get_observations_torch(2, copy=True)
np.testing.assert_array_equal(one["depth"].cpu().numpy(), host["depth"][2])
# torch work on the views: mean depth per env through torch equals numpy's
np.testing.assert_allclose(dev["depth"].mean(dim=(1, 2)).cpu().numpy(),
                           host["depth"].mean(axis=(1, 2)), rtol=1e-5)
# views alias the planes: the next frame shows through, the clone keeps the old one
rr.set_sensor_transforms(np.stack([m3.mul(p, m3.rotation_y(0.7)) for p in poses]))
rr.draw()
rr.renderer.synchronize()
new = rr.get_observations()
np.testing.assert_array_equal(dev["rgba"].cpu().numpy(), new["rgba"])
assert not np.array_equal(new["depth"][2], one["depth"].cpu().numpy())
rr.close()
print("DLPACK OK", flush=True)
