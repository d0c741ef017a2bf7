"""Host-side contract of the replay-renderer surface (no GPU): grid size
(AbstractReplayRenderer.cpp:26-31) and configuration checks."""
import math

import pytest

from habitat_sim_amd import replay
from oracle.oracle import environment_grid_size as oracle_grid


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 16, 17, 100, 512, 513, 4096])
def test_environment_grid_size(n):
    x, y = replay.environment_grid_size(n)
    assert x == math.ceil(math.sqrt(n)) and x * y >= n and x * (y - 1) < n
    assert (x, y) == tuple(oracle_grid(n))


def test_grid_size_of_benchmark():
    assert replay.environment_grid_size(512) == (23, 23)


def test_configuration_checks():
    S = replay.CameraSensorSpec
    with pytest.raises(ValueError):
        replay.BatchReplayRenderer(replay.ReplayRendererConfiguration(2, []))
    with pytest.raises(ValueError):  # resolutions differ
        replay.BatchReplayRenderer(replay.ReplayRendererConfiguration(
            2, [S("rgb", "color", (64, 64)), S("d", "depth", (32, 32))]))
    with pytest.raises(ValueError):  # two colour sensors
        replay.BatchReplayRenderer(replay.ReplayRendererConfiguration(
            2, [S("a", "color"), S("b", "color")]))
    with pytest.raises(ValueError):
        replay.BatchReplayRenderer(replay.ReplayRendererConfiguration(1, [S("x", "normal")]))
