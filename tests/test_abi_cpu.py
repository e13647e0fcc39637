"""CPU-side checks of the engine library: it loads without a GPU and exports
every entry point include/clue_abi.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

import cluestering_alpaka3_amd as clue

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "clue_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clue_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = clue.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), f"{name} declared in clue_abi.h but not exported"
    # the Python binding types every declared function
    assert sorted(clue.exported_symbols()) == names


def test_no_device_reports_runtime_error_not_crash():
    if clue.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(clue.ClueError):
        clue.Plan(0)


def test_status_translation():
    # invalid parameters are rejected before any device work (detail/Clusterer.hpp:32-35)
    p = clue.make_params(-1.0, 1.0)
    assert p.dc == -1.0
    assert clue.STATUS[1] == "CLUE_ERR_INVALID_ARGUMENT"
    assert issubclass(clue.InvalidArgument, ValueError)
    assert issubclass(clue.OutOfRange, IndexError)


def test_autotune_candidates_without_gpu():
    # a host-only query: the points-per-tile values the autotuner may return
    c = clue.autotune_candidates()
    assert c == tuple(sorted(c)) and 16 in c and 128 in c and all(v > 0 for v in c)
