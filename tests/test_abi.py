"""CPU checks of the C ABI: the library loads (no GPU needed) and exports exactly the entry
points include/bats_amd.h declares, and the ctypes binding declares each of them."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "bats_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bats_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("bats_lif_forward", "bats_lif_backward", "bats_count_loss", "bats_ttfs_loss",
                 "bats_adam_step", "bats_adam_step_dev", "bats_events_to_per_neuron",
                 "bats_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from bats_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libbats_amd.so is not built (python -m bats_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_ctypes_binding_covers_header():
    from bats_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()
    lib = _lib.load()
    assert lib.bats_version().startswith(b"bats_amd")


def test_errors_are_reported_without_a_gpu():
    # argument validation happens on the host before any device call
    from bats_amd import _lib
    lib = _lib.load()
    rc = lib.bats_lif_forward(None, 1, 1, 1, 1, 0, None, None, 0.1, 0.2, 1.0, 1.0, 1.0,
                              None, None, None, None, None, None, None, None, 0)
    assert rc != 0
    assert b"K_post" in lib.bats_last_error()
    rc = lib.bats_lif_forward_solver(None, 1, 1, 1, 1, 1, None, None, 0.1, 0.2, 1.0, 1.0, 1.0,
                                     None, None, None, None, None, None, None, 7, 0, None, 0)
    assert rc != 0 and b"solver_mode" in lib.bats_last_error()


def test_workspace_sizes_are_positive():
    from bats_amd import _lib
    lib = _lib.load()
    assert lib.bats_lif_forward_workspace_size(256, 784, 1, 800) > 256 * 784 * 48
    assert lib.bats_lif_backward_workspace_size(256, 784, 1, 800, 30) > 256 * 800 * 30 * 16


def test_plain_form_symbols():
    from bats_amd import _lib
    from bats_amd.plain import PLAIN_SIGNATURES
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "bats_amd_plain.h").read_text(), flags=re.S)
    declared = sorted(set(re.findall(r"\b(bats_plain_[a-z0-9_]+)\s*\(", text)))
    assert sorted(PLAIN_SIGNATURES) == declared
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    assert all(hasattr(lib, n) for n in declared)
