"""ctypes binding of libbats_amd.so (the C ABI declared in include/bats_amd.h).

The library is loaded lazily on first use. There is no fallback: if the library has not
been built, or no GPU is visible, every operator raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libbats_amd.so"

_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_f = C.c_float
_d = C.c_double
_sz = C.c_size_t

# name -> (restype, argtypes); must match include/bats_amd.h
SIGNATURES = {
    "bats_last_error": (C.c_char_p, []),
    "bats_version": (C.c_char_p, []),
    "bats_lif_forward_workspace_size": (_sz, [_i, _i, _i, _i]),
    "bats_lif_forward": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _f, _f, _f, _f, _f,
                              _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "bats_lif_forward_solver": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _f, _f, _f, _f, _f,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _sz]),
    "bats_event_lists_size": (_sz, [_i, _i]),
    "bats_lif_forward_chain": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _f, _f, _f, _f, _f,
                                    _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _vp, _vp,
                                    _vp, _sz, _vp, _sz]),
    "bats_lif_forward_events_workspace_size": (_sz, [_i, _i]),
    "bats_lif_forward_events": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _sz, _vp, _f, _f, _f, _f,
                                     _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _vp, _vp, _vp, _sz,
                                     _vp, _sz]),
    "bats_lif_forward_noalloc_max_bytes": (_sz, []),
    "bats_lif_forward_packed_workspace_size": (_sz, [_i, _i, _i, _i]),
    "bats_lif_forward_packed": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _f, _f, _f,
                                     _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "bats_lif_forward_packed_chain": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _f, _f,
                                           _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f,
                                           _vp, _vp, _vp, _sz, _vp, _sz]),
    "bats_lif_forward_noalloc": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _f, _f, _f, _f, _f,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bats_lif_backward_workspace_size": (_sz, [_i, _i, _i, _i, _i]),
    "bats_lif_backward": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _vp, _vp, _vp, _f, _f, _f, _f, _vp, _i, _vp, _vp, _sz]),
    "bats_conv_lif_forward_workspace_size": (_sz, [_i, _vp, _i, _vp]),
    "bats_conv_lif_forward": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _vp, _f, _f, _f, _f, _f,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz]),
    "bats_conv_lif_backward_workspace_size": (_sz, [_i, _vp, _i, _vp, _i]),
    "bats_conv_lif_backward": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _f, _f, _f, _f, _vp, _i, _vp, _vp,
                                    _sz]),
    "bats_pool2x2_forward": (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "bats_pool2x2_backward": (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    "bats_count_loss": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _f, _vp, _vp, _vp]),
    "bats_count_loss_targets": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "bats_ttfs_loss": (_i, [_vp, _i, _i, _i, _vp, _vp, _f, _vp, _vp, _vp]),
    "bats_adam_step": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _f,
                            _d, _d]),
    "bats_nadam_step": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _f,
                             _d, _d]),
    "bats_adam_step_dev": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _f,
                                _vp, _i]),
    "bats_sgd_step": (_i, [_vp, _i64, _vp, _vp, _vp, _f]),
    "bats_add_deltas": (_i, [_vp, _i64, _vp, _vp]),
    "bats_adam_step_multi": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _f,
                                  _d, _d, _vp, _i]),
    "bats_add_deltas_multi": (_i, [_vp, _i, _vp, _vp, _vp]),
    "bats_events_to_per_neuron": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "bats_events_csr_to_per_neuron": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bats_latency_encode": (_i, [_vp, _i64, _vp, _i, _d, _d, _vp, _vp]),
    "bats_pack_event_lists": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64]),
    "bats_timing_enable": (None, [_i]),
    "bats_timing_reset": (None, []),
    "bats_timing_query": (_i, [C.c_char_p, C.POINTER(_d), C.POINTER(C.c_longlong)]),
}

_lib = None


def load():
    """Load libbats_amd.so and declare every exported signature."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bats_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("bats_amd: " + load().bats_last_error().decode())
