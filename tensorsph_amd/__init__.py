"""tensorsph_amd: MI355X-native particle pipeline behind the TensorSPH API."""
from .config import (Config, DamBreakConfig, DamBreak3DConfig, Kleefsman3DConfig,
                     SimulationUnstableError, VISCOUS_LAPLACIAN_FACTOR, validate_config)
