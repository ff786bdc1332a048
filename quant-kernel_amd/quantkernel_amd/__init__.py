"""MI355X (gfx950) Monte Carlo batch pricing behind the QuantKernel C ABI."""
from .engine import *  # noqa: F401,F403
from .engine import QuantKernel, load_library  # noqa: F401
