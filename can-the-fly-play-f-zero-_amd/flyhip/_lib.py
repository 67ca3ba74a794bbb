"""ctypes binding of ``libflyhip.so`` (the C ABI in include/flyhip.h).

torch is imported first so that its HIP runtime (SONAME libamdhip64.so.7) is the one the
library binds to: device pointers from torch tensors and from the engine then live in one
runtime.  There is no CPU fallback: a missing library raises at import time of the engine.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library links against)

from .build import LIB_PATH

FZ_ABI_VERSION = 4
FZ_OK, FZ_E_ARG, FZ_E_STATE, FZ_E_HIP, FZ_E_OVERFLOW = 0, -1, -2, -3, -4

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int


class FzParams(ctypes.Structure):
    _fields_ = [("decay_syn", ctypes.c_float), ("decay_m", ctypes.c_float), ("u_th", ctypes.c_float),
                ("u_reset", ctypes.c_float), ("kick", ctypes.c_float),
                ("in_scale", ctypes.c_float), ("pw_qscale", ctypes.c_float),
                ("n_ref", c_int), ("n_delay", c_int)]


class FzState(ctypes.Structure):
    _fields_ = [("B", c_int), ("u", c_void_p), ("g", c_void_p), ("ref", c_void_p), ("ring", c_void_p),
                ("counts", c_void_p), ("clock", c_void_p), ("ring_slots", c_int), ("spk_idx", c_void_p),
                ("spk_cnt", c_void_p)]


class FzInputs(ctypes.Structure):
    _fields_ = [("n_ext", c_int), ("ext_idx", c_void_p), ("p_ext", c_void_p), ("bias_ext", c_void_p),
                ("pw", c_void_p), ("alive", c_void_p), ("seed", ctypes.c_uint), ("fly_base", ctypes.c_uint)]


class FzTune(ctypes.Structure):
    _fields_ = [("upd_threads", c_int), ("push_threads", c_int), ("push_chunk", c_int),
                ("push_grid", c_int), ("block_steps", c_int), ("tile_stage", c_int),
                ("tile_source", c_int), ("plastic_pull", c_int), ("block_parts", c_int),
                ("acc_bits", c_int)]


# name -> (restype, argtypes); every symbol declared in include/flyhip.h
SIGNATURES = {
    "fz_version": (c_int, []),
    "fz_build_info": (ctypes.c_char_p, []),
    "fz_last_error": (ctypes.c_char_p, []),
    "fz_create": (c_int, [c_int, ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_int, c_int,
                          ctypes.POINTER(FzParams), ctypes.POINTER(c_void_p)]),
    "fz_destroy": (c_int, [c_void_p]),
    "fz_set_plastic": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fz_bind": (c_int, [c_void_p, ctypes.POINTER(FzState)]),
    "fz_set_tiles": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "fz_run": (c_int, [c_void_p, ctypes.POINTER(FzInputs), c_int, c_void_p, c_void_p,
                       ctypes.POINTER(FzTune), c_void_p]),
    "fz_run_static": (c_int, [c_void_p, ctypes.POINTER(FzInputs), c_int, c_void_p, c_void_p,
                              ctypes.POINTER(FzTune), c_void_p]),
    "fz_run_timed": (c_int, [c_void_p, ctypes.POINTER(FzInputs), c_int, ctypes.POINTER(FzTune), c_void_p,
                             ctypes.POINTER(ctypes.c_float), ctypes.POINTER(c_int),
                             ctypes.POINTER(ctypes.c_float), ctypes.POINTER(c_int)]),
    "fz_graph_create": (c_int, [c_void_p, ctypes.POINTER(FzInputs), c_int, ctypes.POINTER(FzTune),
                                ctypes.POINTER(c_void_p)]),
    "fz_graph_launch": (c_int, [c_void_p, c_void_p]),
    "fz_graph_destroy": (c_int, [c_void_p]),
    "fz_propagate": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "fz_check": (c_int, [c_void_p, c_void_p]),
    "fz_reward_step": (c_int, [c_void_p, c_void_p]),
    "fz_sg_gemm": (c_int, [c_int, c_int, c_int, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int, c_void_p,
                           c_void_p, c_int, c_void_p, c_void_p]),
    "fz_sg_gemm_batched": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_int, ctypes.c_longlong, c_int, c_void_p,
                                   c_int, ctypes.c_longlong, c_int, c_int, c_void_p, c_void_p, c_int,
                                   ctypes.c_longlong, c_void_p, ctypes.c_longlong, c_void_p]),
    "fz_sg_ens_head": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                               ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_float, c_void_p,
                               c_void_p, c_void_p]),
    "fz_sg_gelu_bwd": (c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fz_sg_colsum": (c_int, [c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "fz_sg_gru_fwd": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fz_sg_gru_bwd": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p]),
    "fz_sg_mse": (c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_void_p, ctypes.c_float, c_void_p, c_void_p]),
    "fz_sg_adam": (c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float,
                           ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, c_void_p]),
    "fz_sg_cast": (c_int, [ctypes.c_longlong, c_void_p, c_void_p, c_void_p]),
    "fz_dn_rates": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                            c_void_p, c_void_p]),
    "fz_instruct_traces": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, ctypes.c_float,
                                   ctypes.c_float, c_int, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                                   c_void_p]),
    "fz_instruct_delta": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                  c_int, c_void_p, c_int, ctypes.c_float, c_void_p, c_void_p]),
    "fz_instruct_update": (c_int, [c_void_p, c_void_p]),
    "fz_eye_step": (c_int, [c_void_p, c_void_p]),
    "fz_instruct_bias": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                 c_void_p, c_int, ctypes.c_float, c_void_p, ctypes.c_float, c_void_p, c_void_p,
                                 c_int, c_void_p, c_void_p]),
}


class FzInstructUpdateArgs(ctypes.Structure):
    """fz_instruct_update_args of include/flyhip.h."""
    _fields_ = [(f, c_int) for f in ("mode", "n_flies", "n_sel", "n_pl", "n_upre", "n_upost", "n_groups", "n_l")] + \
               [(f, c_void_p) for f in ("sel", "pre_k", "group", "post_k", "mask", "post_l", "fly_ptr", "fly_slots",
                                       "learn", "coef", "trace", "rate", "tgt", "delta", "sign", "wmax", "pw")]

_lib = None


def load() -> ctypes.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(flyhip has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.fz_version() != FZ_ABI_VERSION:
            raise RuntimeError(f"libflyhip ABI {lib.fz_version()} != {FZ_ABI_VERSION}")
        _lib = lib
    return _lib


class FlyHipError(RuntimeError):
    pass


def library_info() -> dict:
    """The native library this process loaded: its path, the SHA-256 of the file, the source hash
    compiled into it and whether that equals the hash of the sources present now (the library
    is built in the development tree and travels with it; nothing rebuilds it on load)."""
    import hashlib

    from .build import source_hash

    lib = load()
    built = lib.fz_build_info().decode(errors="replace")
    src_now = source_hash()
    return {"path": LIB_PATH, "sha256": hashlib.sha256(open(LIB_PATH, "rb").read()).hexdigest()[:16],
            "build_info": built, "sources_now": src_now, "matches_sources": built.startswith(f"src {src_now};")}


def check(rc: int, what: str = ""):
    if rc != FZ_OK:
        msg = load().fz_last_error().decode(errors="replace")
        exc = ValueError if rc == FZ_E_ARG else FlyHipError
        raise exc(f"{what}: {msg} (code {rc})")


def ptr(t) -> int | None:
    """Device (or host) address of a tensor, None for None / empty."""
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class FzEyeArgs(ctypes.Structure):
    """fz_eye_args of include/flyhip.h."""
    _fields_ = [(f, c_int) for f in ("B", "h", "w", "cell", "gh", "gw", "shift", "n_cells")] + \
               [(f, c_void_p) for f in ("first", "frames", "lp", "d_on", "d_off", "key", "gy", "gx", "py")] + \
               [(f, ctypes.c_float) for f in ("y_cut", "a_hp", "a_d", "gain", "max_rate")] + \
               [("rates", c_void_p), ("p_ext", c_void_p), ("n_ext", c_int), ("slot", c_void_p),
                ("p_scale", ctypes.c_float), ("tr_gain", ctypes.c_float), ("su_gain", ctypes.c_float),
                ("ch_gain", ctypes.c_float), ("surround", c_int)]
