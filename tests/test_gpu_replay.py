"""The replay-renderer surface on the GPU: render(color_images,
depth_images) fills caller arrays with GL rows like the reference's
AbstractReplayRenderer::render, get_observations() gives habitat's top-down
observation dict, and both equal the oracle's planes."""
import math

import numpy as np
import pytest

from habitat_sim_amd import replay, scenes
from habitat_sim_amd import math3d as m3
from oracle.oracle import OracleRenderer

pytestmark = pytest.mark.gpu


def test_render_and_observations_match_oracle():
    spec = scenes.small_rooms(n_envs=5)
    S = replay.CameraSensorSpec
    cfg = replay.ReplayRendererConfiguration(
        5, [S("rgba", "color", (128, 128)), S("depth", "depth", (128, 128)),
            S("semantic", "semantic", (128, 128))])
    rr = replay.BatchReplayRenderer(cfg)
    assert rr.grid == (3, 2)
    assert rr.add_file_data(spec.file)
    o = OracleRenderer((128, 128), rr.grid)
    o.add_file_data(spec.file)
    for e, env in enumerate(spec.envs):
        for inst in env.instances:
            rr.add_instance(e, inst.template, inst.bake, inst.semantic)
            o.add_node_hierarchy(e, inst.template, inst.bake, inst.semantic)
    rr.set_sensor_transforms(np.stack([env.camera_pose for env in spec.envs]))
    proj = S("x").projection()
    for e, env in enumerate(spec.envs):
        o.update_camera(e, proj, m3.inverted(env.camera_pose))
    colors = [np.zeros((128, 128, 4), np.uint8) for _ in range(5)]
    depths = [np.zeros((128, 128), np.float32) for _ in range(5)]
    rr.render(colors, depths)
    planes = o.render(scenes=list(range(5)))
    obs = rr.get_observations()
    assert obs["rgba"].dtype == np.uint8 and obs["rgba"].shape == (5, 128, 128, 4)
    assert obs["depth"].dtype == np.float32 and obs["semantic"].dtype == np.uint32
    for e in range(5):
        np.testing.assert_array_equal(colors[e], planes["color"][e])
        np.testing.assert_array_equal(depths[e], planes["lindepth"][e])
        # observations: the same, rows top-down
        np.testing.assert_array_equal(obs["rgba"][e], planes["color"][e][::-1])
        np.testing.assert_array_equal(obs["depth"][e], planes["lindepth"][e][::-1])
        np.testing.assert_array_equal(obs["semantic"][e], planes["semantic"][e][::-1])
    one = rr.get_observations(3)
    np.testing.assert_array_equal(one["semantic"], obs["semantic"][3])
    # empty lists skip; wrong counts are errors
    rr.render([], depths)
    with pytest.raises(ValueError):
        rr.render(colors[:2], [])
    rr.close()


def test_keyframe_player_on_the_gpu_matches_oracle():
    """The gfx-replay deletion sequence played into the HIP renderer and into
    the oracle: identical node layouts and identical pixels per keyframe."""
    from habitat_sim_amd import RendererConfiguration, RendererStandalone
    from habitat_sim_amd.replay_player import BatchPlayer
    from test_replay_player import deletion_keyframes
    r = RendererStandalone(RendererConfiguration().set_tile_size_count((64, 64), (1, 1)))
    o = OracleRenderer((64, 64), (1, 1))
    P = m3.orthographic_projection((6.0, 6.0), 0.1, 10.0)
    V = m3.inverted_rigid(m3.translation((0.5, 0.5, 5.0)))
    players = []
    for x in (r, o):
        x.update_camera(0, P, V)
        p = BatchPlayer(x, 0)
        for k in deletion_keyframes():
            p.append_keyframe(k)
        players.append(p)
    for idx in (0, 1, 2, 0):
        for p in players:
            p.set_keyframe_index(idx)
        r.draw()
        o.draw()
        np.testing.assert_array_equal(r.color_image(), o.color_image(), err_msg=f"keyframe {idx}")
        np.testing.assert_array_equal(r.depth_image(), o.depth_image(), err_msg=f"keyframe {idx}")
    r.close()


def test_keyframe_lights_on_the_gpu_match_oracle():
    """Keyframes that change the light setup (changeLightSetup) over two
    shaded spheres of batch.gltf: every keyframe's pixels identical to the
    oracle's, and each light change visible."""
    import json
    import os
    from habitat_sim_amd import RendererConfiguration, RendererStandalone
    from habitat_sim_amd.replay_player import BatchPlayer
    asset = os.path.join(os.path.dirname(__file__), "golden", "scenes", "batch.gltf")
    l1 = [{"vector": [-1.0, 1.5, 1.0, 1.0], "color": [0.6, 0.6, 3.0], "model": "global"},
          {"vector": [3.0, 1.0, 3.0, 0.0], "color": [0.8, 1.0, 1.0], "model": "global"}]
    l2 = [{"vector": [0.0, 0.3, 1.0, 0.0], "color": [1.0, 0.5, 0.25], "model": "global"}]
    frames = [json.dumps({"lightsChanged": True, "lights": ls}) for ls in (l1, l2, [])]
    r = RendererStandalone(RendererConfiguration().set_tile_size_count((96, 96), (1, 1))
                           .set_max_light_count(2))
    o = OracleRenderer((96, 96), (1, 1), max_light_count=2)
    P = m3.perspective_projection(math.radians(60.0), 1.0, 0.1, 20.0)
    V = m3.inverted(m3.translation((0.0, 0.0, 5.0)))
    players = []
    for x in (r, o):
        assert x.add_file(asset)
        a = x.add_node_hierarchy(0, "shaded yellow sphere")
        b = x.add_node_hierarchy(0, "shaded checkerboard sphere")
        x.transformations(0)[a] = m3.translation((-1.1, 0.0, 0.0))
        x.transformations(0)[b] = m3.translation((1.1, 0.0, 0.0))
        x.update_camera(0, P, V)
        p = BatchPlayer(x, 0)
        for k in frames:
            p.append_keyframe(k)
        players.append(p)
    images = []
    for idx in range(len(frames)):
        for p in players:
            p.set_keyframe_index(idx)
        r.draw()
        o.draw()
        np.testing.assert_array_equal(r.color_image(), o.color_image(), err_msg=f"keyframe {idx}")
        images.append(r.color_image())
    assert not np.array_equal(images[0], images[1]) and not np.array_equal(images[1], images[2])
    r.close()


def test_environment_keyframes_and_sensor_transforms_from_keyframe():
    """setEnvironmentKeyframe (wrapped in {"keyframe": ...}),
    setEnvironmentKeyframeUnwrapped, setSensorTransformsFromKeyframe and
    clearEnvironment of AbstractReplayRenderer: the pixels of each step equal
    the oracle's, driven by one player per environment with the camera taken
    from the same user transform."""
    import json
    from habitat_sim_amd.replay_player import BatchPlayer, transform_matrix
    from test_replay_player import deletion_keyframes
    S = replay.CameraSensorSpec
    cfg = replay.ReplayRendererConfiguration(
        2, [S("rgba", "color", (64, 64), hfov=70.0), S("depth", "depth", (64, 64), hfov=70.0)])
    rr = replay.BatchReplayRenderer(cfg)
    o = OracleRenderer((64, 64), rr.grid)
    proj = cfg.sensor_specifications[0].projection()
    half = math.radians(8.0)
    cams = {0: ("sensor_rgba", [0.5, 0.5, 6.0], [math.cos(half), 0.0, math.sin(half), 0.0]),
            # only the depth sensor's name: the first sensor that is found gives the pose
            1: ("agent_depth", [0.2, 0.8, 7.0], [math.cos(half), math.sin(half), 0.0, 0.0])}
    k0, k1, _ = (json.loads(k) for k in deletion_keyframes())
    players = [BatchPlayer(o, e) for e in range(2)]

    def frame(env, k):
        name, t, q = cams[env]
        return dict(k, userTransforms=[{"name": name, "transform": {"translation": t,
                                                                    "rotation": q}}])

    def check(step):
        for e in range(2):
            name, t, q = cams[e]
            o.update_camera(e, proj, m3.inverted(transform_matrix(t, q)))
        colors = [np.zeros((64, 64, 4), np.uint8) for _ in range(2)]
        depths = [np.zeros((64, 64), np.float32) for _ in range(2)]
        rr.render(colors, depths)
        planes = o.render(scenes=[0, 1])
        for e in range(2):
            np.testing.assert_array_equal(colors[e], planes["color"][e], err_msg=f"{step} env {e}")
            np.testing.assert_array_equal(depths[e], planes["lindepth"][e],
                                          err_msg=f"{step} env {e}")
        return colors

    with pytest.raises(RuntimeError, match="set_environment_keyframe"):
        rr.set_sensor_transforms_from_keyframe(0, "sensor_")
    with pytest.raises(ValueError, match="keyframe"):
        rr.set_environment_keyframe(0, json.dumps(frame(0, k0)))
    rr.set_environment_keyframe(0, json.dumps({"keyframe": frame(0, k0)}))
    rr.set_environment_keyframe_unwrapped(1, json.dumps(frame(1, k0)))
    players[0].set_single_keyframe(frame(0, k0))
    players[1].set_single_keyframe(frame(1, k0))
    rr.set_sensor_transforms_from_keyframe(0, "sensor_")
    rr.set_sensor_transforms_from_keyframe(1, "agent_")
    with pytest.raises(KeyError, match="other_rgba"):
        rr.set_sensor_transforms_from_keyframe(0, "other_")
    first = check("first keyframe")
    assert first[0].any() and first[1].any() and not np.array_equal(first[0], first[1])
    # a second single keyframe applies on top of the first (X deleted, Z made and deleted)
    rr.set_environment_keyframe(0, json.dumps({"keyframe": frame(0, k1)}))
    players[0].set_single_keyframe(frame(0, k1))
    rr.set_sensor_transforms_from_keyframe(0, "sensor_")
    second = check("second keyframe")
    assert not np.array_equal(first[0], second[0])
    np.testing.assert_array_equal(first[1], second[1])
    # clearEnvironment closes the player: environment 0 renders empty
    rr.clear_environment(0)
    players[0].close()
    cleared = check("cleared")
    assert (cleared[0] == cleared[0][0, 0]).all() and not np.array_equal(cleared[0], first[0])
    with pytest.raises(RuntimeError):
        rr.set_sensor_transforms_from_keyframe(0, "sensor_")
    rr.close()


def test_reference_binding_names():
    """The names habitat-lab calls on ReplayRenderer (SimBindings.cpp:
    478-545): create_batch_replay_renderer, environment_count, sensor_size,
    cuda_{color,depth}_buffer_device_pointer (capsules of framebuffer
    addresses; depth is depth, not the reference binding's colour pointer)
    and unproject."""
    from habitat_sim_amd import hip

    spec = scenes.small_rooms(n_envs=3)
    S = replay.CameraSensorSpec
    cfg = replay.ReplayRendererConfiguration(
        3, [S("rgba", "color", (96, 128)), S("depth", "depth", (96, 128))])
    rr = replay.BatchReplayRenderer.create_batch_replay_renderer(cfg)
    assert rr.environment_count == 3
    assert rr.sensor_size(2) == (128, 96)  # (width, height)
    with pytest.raises(IndexError):
        rr.sensor_size(3)
    assert rr.add_file_data(spec.file)
    for e, env in enumerate(spec.envs):
        for inst in env.instances:
            rr.add_instance(e, inst.template, inst.bake, inst.semantic)
    rr.set_sensor_transforms(np.stack([env.camera_pose for env in spec.envs]))
    colors = [np.zeros((96, 128, 4), np.uint8) for _ in range(3)]
    depths = [np.zeros((96, 128), np.float32) for _ in range(3)]
    rr.render(colors, depths)
    gx, gy = rr.grid
    fw, fh = 128 * gx, 96 * gy
    c_ptr = replay.capsule_address(rr.cuda_color_buffer_device_pointer())
    d_ptr = replay.capsule_address(rr.cuda_depth_buffer_device_pointer())
    assert c_ptr and d_ptr and c_ptr != d_ptr
    fb_c = np.empty((fh, fw, 4), np.uint8)
    fb_d = np.empty((fh, fw), np.float32)
    hip.memcpy_d2h(fb_c, c_ptr)
    hip.memcpy_d2h(fb_d, d_ptr)
    for e in range(3):
        x0, y0 = (e % gx) * 128, (e // gx) * 96
        np.testing.assert_array_equal(fb_c[y0:y0 + 96, x0:x0 + 128], colors[e])
        np.testing.assert_array_equal(fb_d[y0:y0 + 96, x0:x0 + 128], depths[e])
    # unproject: the ray through the far-plane point of a pixel's lower-left
    # corner (RenderCamera.cpp:191-197 maps the pixel index, not its
    # centre), from the camera position; projected back it lands on that
    # corner, and the middle pixel's ray is the camera's forward axis
    proj = S("x", resolution=(96, 128)).projection()
    for e, env in enumerate(spec.envs):
        pose = env.camera_pose
        view = m3.inverted(pose)
        for (px, py) in ((0, 0), (64, 48), (127, 95), (17, 80)):
            ray = rr.unproject(e, (px, py))
            np.testing.assert_allclose(ray.origin, pose[3, :3], atol=1e-6)
            assert abs(np.linalg.norm(ray.direction) - 1.0) < 1e-5
            p = np.append(ray.origin + 5.0 * ray.direction, 1.0)
            clip = proj.T.astype(np.float64) @ (view.T.astype(np.float64) @ p)
            ndc = clip[:2] / clip[3]
            assert abs((ndc[0] + 1) * 128 / 2 - px) < 1e-3
            assert abs((ndc[1] + 1) * 96 / 2 - (96 - py - 1)) < 1e-3
        mid = rr.unproject(e, (64, 47)).direction
        assert np.dot(mid, -pose[2, :3]) > 0.999
    with pytest.raises(IndexError):
        rr.unproject(3, (0, 0))
    rr.close()
