"""Host-side mirror of the reference's per-timestep interface over the
MI355X C-ABI (include/lsd.h).  See ``lsd.gpu`` for the binding."""
from .config import LsdConfig, make_config, config1, config3, config_small, ra_crit, ra_crit_estimate, grid_for, PLAIN_F64, MIXED_F32, MIXED_BF16X3, MIXED_F32_MAG64  # noqa: F401
from .initial import seeded_state, harmonics, FIELD_ID, FIELD_ORDER, SEEDS, AMPS  # noqa: F401
