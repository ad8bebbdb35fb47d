"""gtopen_amd: MI355X DCFR engine for heads-up postflop spots.

Host side: ``Spot`` (tree/range/river-table construction in the native
library) and ``GpuSolver`` (the HIP engine over the C ABI in
include/gtopen_amd.h). Everything routes through libgtopen_amd.so; there is
no CPU fallback.
"""
from ._lib import GtoError, lib  # noqa: F401
from .game_value import GameValue, game_value  # noqa: F401
from .solver import GpuSolver, Progress, vram_info  # noqa: F401
from .spot import (  # noqa: F401
    Spot,
    SpotConfig,
    StreetSizing,
    TreeConfig,
    canonical_flops,
    card_str,
    combo_from_index,
    combo_index,
    evaluate7,
    parse_cards,
    parse_sizes,
    range_parse,
)
