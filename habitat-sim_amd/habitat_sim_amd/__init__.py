"""MI355X-native batch sensor renderer: drop-in for habitat-sim's
esp::gfx_batch::Renderer / RendererStandalone path."""
from .renderer import (RendererConfiguration, RendererStandalone,  # noqa: F401
                       RendererStandaloneConfiguration, SceneStats,
                       calculate_depth_unprojection, unproject_depth)
from .replay import (BatchReplayRenderer, CameraSensorSpec,  # noqa: F401
                     ReplayRendererConfiguration, environment_grid_size)

__version__ = "0.1.0"
