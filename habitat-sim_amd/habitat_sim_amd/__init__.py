"""MI355X-native batch sensor renderer: drop-in for habitat-sim's
esp::gfx_batch::Renderer / RendererStandalone path."""
from .renderer import (RendererConfiguration, RendererStandalone,  # noqa: F401
                       RendererStandaloneConfiguration, SceneStats,
                       calculate_depth_unprojection, unproject_depth)
from .replay import (BatchReplayRenderer, CameraSensorSpec,  # noqa: F401
                     ReplayRendererConfiguration, environment_grid_size)
from .noise import RedwoodDepthNoiseModel  # noqa: F401
from .replay_player import BatchPlayer, Keyframe  # noqa: F401

__version__ = "0.1.0"
