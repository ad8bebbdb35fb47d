"""ctypes binding of libgtopen_amd.so (declarations mirror include/gtopen_amd.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C gtopen_amd/csrc``). Loading fails loudly when it is missing: there
is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgtopen_amd.so")

c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)

GTO_OK = 0
KIND_ACTION, KIND_CHANCE, KIND_FOLD, KIND_SHOWDOWN = 0, 1, 2, 3
SENTINEL = 0xFFFFFFFF
STORAGE_F32, STORAGE_COMPRESSED, STORAGE_HALF, STORAGE_HALF_ROWS = 0, 1, 2, 3
STORAGE = {"f32": STORAGE_F32, "compressed": STORAGE_COMPRESSED, "half": STORAGE_HALF,
           "half_rows": STORAGE_HALF_ROWS}
ALGO_DCFR, ALGO_CFRPLUS, ALGO_PCFRPLUS = 0, 1, 2
SHOWDOWN = {"exact": 0, "bf16": 1, "fp8": 2, "sequential": 3}  # GTO_SHOWDOWN_*


def algo_code(name):
    """Algorithm::parse (cfr.rs:26-34): case-insensitive, '-', '_' and '+'
    ignored; dcfr, cfr / cfrplus, pcfr / pcfrplus."""
    key = name.lower().replace("-", "").replace("_", "").replace("+", "")
    codes = {"dcfr": ALGO_DCFR, "cfr": ALGO_CFRPLUS, "cfrplus": ALGO_CFRPLUS, "pcfr": ALGO_PCFRPLUS,
             "pcfrplus": ALGO_PCFRPLUS}
    if key not in codes:
        raise ValueError(f"unknown algorithm {name!r} (use dcfr, cfr+ or pcfr+)")
    return codes[key]


STRICT, LENIENT_LOAD = 0, 1
PROF_KINDS = 9


class SpotView(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_uint32),
        ("node_kind", c_u8p),
        ("node_player", c_u8p),
        ("node_street", c_u8p),
        ("node_num_children", c_u8p),
        ("node_children_start", c_u32p),
        ("node_data_offset", c_u64p),
        ("node_t_win", c_f64p),
        ("node_t_lose", c_f64p),
        ("node_t_tie", c_f64p),
        ("node_put", c_f64p),
        ("num_children_total", ctypes.c_uint32),
        ("children", c_u32p),
        ("data_size", ctypes.c_uint64 * 2),
        ("root_street", ctypes.c_uint32),
        ("starting_pot", ctypes.c_double),
        ("effective_stack", ctypes.c_double),
        ("rake_pct", ctypes.c_double),
        ("rake_cap", ctypes.c_double),
        ("board_len", ctypes.c_uint32),
        ("board", ctypes.c_uint8 * 5),
        ("board_mask", ctypes.c_uint64),
        ("num_hands", ctypes.c_uint32 * 2),
        ("hand_c1", c_u8p * 2),
        ("hand_c2", c_u8p * 2),
        ("weights", c_f32p * 2),
        ("same_combo", c_u32p * 2),
        ("num_river_keys", ctypes.c_uint32),
        ("river_off", c_i64p * 2),
        ("river_cnt", c_u32p * 2),
        ("river_str", c_u32p * 2),
        ("river_idx", c_u16p * 2),
        ("num_suit_perms", ctypes.c_uint32),
        ("suit_perms", c_u8p),
        ("hand_perm", c_u16p * 2),
    ]


class HipOptions(ctypes.Structure):
    _fields_ = [
        ("storage", ctypes.c_int),
        ("algo", ctypes.c_int),
        ("use_graphs", ctypes.c_int),
        ("device", ctypes.c_int),
        ("isomorphism", ctypes.c_int),
    ]


class HipProfile(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * PROF_KINDS), ("launches", ctypes.c_uint32 * PROF_KINDS)]


class Schedule(ctypes.Structure):
    _fields_ = [("show_nodes", ctypes.c_uint32), ("per_level", ctypes.c_uint32), ("river", ctypes.c_uint32),
                ("absorb", ctypes.c_uint32)]


SCHEDULE_CANDIDATES = 7


class RunOptions(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_uint32), ("target_exploit_pct", ctypes.c_double),
                ("check_every", ctypes.c_uint32)]


class Progress(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_uint32), ("exploit_chips", ctypes.c_double),
                ("exploit_pct_pot", ctypes.c_double), ("elapsed_secs", ctypes.c_double),
                ("gpu_secs", ctypes.c_double)]


PROGRESS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(Progress), ctypes.c_void_p)


# (name, restype, argtypes) for every symbol declared in include/gtopen_amd.h
SIGNATURES = [
    ("gto_last_error", ctypes.c_char_p, []),
    ("gto_spot_build", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("gto_spot_free", None, [ctypes.c_void_p]),
    ("gto_spot_view_of", ctypes.POINTER(SpotView), [ctypes.c_void_p]),
    ("gto_spot_num_action_nodes", ctypes.c_uint64, [ctypes.c_void_p]),
    ("gto_spot_arena_bytes", ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_int]),
    ("gto_spot_num_suit_perms", ctypes.c_uint32, [ctypes.c_void_p]),
    ("gto_combo_index", ctypes.c_uint32, [ctypes.c_uint8, ctypes.c_uint8]),
    ("gto_evaluate7", ctypes.c_uint32, [c_u8p]),
    ("gto_range_parse", ctypes.c_int, [ctypes.c_char_p, c_f32p]),
    ("gto_canonical_flops", ctypes.c_int, [ctypes.c_uint32, c_u8p, c_u32p, ctypes.c_uint32, c_u32p]),
    ("gto_hip_create", ctypes.c_int, [ctypes.POINTER(SpotView), ctypes.POINTER(HipOptions), ctypes.POINTER(ctypes.c_void_p)]),
    ("gto_hip_destroy", None, [ctypes.c_void_p]),
    ("gto_hip_iterate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gto_hip_get_iteration", ctypes.c_int, [ctypes.c_void_p, c_u32p]),
    ("gto_hip_set_iteration", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gto_hip_exploitability", ctypes.c_int, [ctypes.c_void_p, c_f64p]),
    ("gto_hip_root_cfv", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_f32p]),
    ("gto_hip_download_arenas", ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p, c_f32p, c_f32p]),
    ("gto_hip_upload_arenas", ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p, c_f32p, c_f32p]),
    ("gto_hip_download_compressed", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int16), c_u16p, c_f32p, c_f32p]),
    ("gto_hip_update_locks", ctypes.c_int, [ctypes.c_void_p, c_i64p, c_f32p, ctypes.c_uint64]),
    ("gto_hip_set_weights", ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p]),
    ("gto_hip_set_showdown", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("gto_hip_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Schedule)]),
    ("gto_hip_get_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Schedule)]),
    ("gto_hip_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Schedule), c_f32p]),
    ("gto_batch_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Schedule), c_f32p]),
    ("gto_hip_solved_strategy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_f32p]),
    ("gto_batch_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    ("gto_batch_iterate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("gto_batch_destroy", None, [ctypes.c_void_p]),
    ("gto_hip_node_values", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, c_u32p, c_f32p]),
    ("gto_save_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint32, c_f32p, c_f32p, c_f32p, c_f32p,
                                      ctypes.c_uint32, c_u32p, c_u64p, c_u64p, c_f32p, ctypes.c_uint32, c_u32p,
                                      ctypes.POINTER(ctypes.c_char_p)]),
    ("gto_save_open", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("gto_save_close", None, [ctypes.c_void_p]),
    ("gto_save_config_json", ctypes.c_char_p, [ctypes.c_void_p]),
    ("gto_save_iteration", ctypes.c_uint32, [ctypes.c_void_p]),
    ("gto_save_num_locks", ctypes.c_uint32, [ctypes.c_void_p]),
    ("gto_save_lock", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.POINTER(c_f32p), c_u64p]),
    ("gto_save_num_labels", ctypes.c_uint32, [ctypes.c_void_p]),
    ("gto_save_label", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, ctypes.POINTER(ctypes.c_char_p)]),
    ("gto_save_validate_locks", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("gto_save_read_arenas", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_f32p, c_f32p, c_f32p, c_f32p]),
    ("gto_validate_locks", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, c_u64p, c_u64p, c_f32p]),
    ("gto_hip_game_value", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_f64p, c_f32p, c_f32p, c_f32p, c_f32p]),
    ("gto_hip_game_value_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_f64p, c_f32p, c_f32p, c_f32p,
                                                  c_f32p]),
    # device addresses (e.g. torch.Tensor.data_ptr() of tensors on the engine's GPU)
    ("gto_hip_game_value_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_f64p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gto_hip_upload_strategy_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("gto_hip_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("gto_hip_profile_kind_name", ctypes.c_char_p, [ctypes.c_int]),
    ("gto_hip_profile_iterations", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(HipProfile)]),
    ("gto_hip_run", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RunOptions), ctypes.POINTER(ctypes.c_int),
                                   PROGRESS_FN, ctypes.c_void_p, ctypes.POINTER(Progress)]),
    ("gto_hip_reached_entries", ctypes.c_uint64, [ctypes.c_void_p]),
    ("gto_hip_river_entries", ctypes.c_uint64, [ctypes.c_void_p]),
    ("gto_hip_estimate_vram", ctypes.c_uint64, [ctypes.POINTER(SpotView), ctypes.c_int]),
    ("gto_hip_vram_info", ctypes.c_int, [ctypes.c_int, c_u64p, c_u64p]),
]

_lib = None


def lib():
    """Load libgtopen_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C gtopen_amd/csrc"
            )
        h = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, res, args in SIGNATURES:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


class GtoError(RuntimeError):
    pass


def check(rc, what):
    if rc != GTO_OK:
        msg = lib().gto_last_error()
        raise GtoError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
