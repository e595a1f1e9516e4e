"""MI355X-native optical-photon transport (HIP kernels for gfx950 behind a C-ABI)."""
from .params import Fate, SimParams, Source, Surface, TallyConfig  # noqa: F401
