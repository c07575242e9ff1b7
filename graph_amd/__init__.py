"""graph_amd — MI355X-native force-simulation tick of cosmos.gl.

``graph_amd.simulation.Graph`` mirrors the reference's simulation API over the
C ABI in ``include/cosmosgl.h``; ``graph_amd._lib.Context`` is the thin binding.
"""
from ._lib import Context, CosmosGLError, default_params  # noqa: F401
from .simulation import Graph  # noqa: F401
