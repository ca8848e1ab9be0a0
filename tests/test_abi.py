"""CPU-side checks of the C ABI: the library loads, exports every entry point
``include/ptx.h`` declares, and rejects bad arguments before touching a GPU."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    names = set()
    for hdr in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(ptx_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib():
    from ptychox_amd import _lib

    return _lib


def test_library_exports_every_declared_symbol(lib):
    handle = lib.lib()
    declared = declared_symbols()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(handle, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared surface
    assert set(lib.SIGNATURES) == declared


def test_dynamic_symbols_are_the_c_abi_only():
    so = ROOT / "ptychox_amd" / "libptychox_hip.so"
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    ours = {s for s in exported if s.startswith("ptx_")}
    assert ours == declared_symbols()


def test_version(lib):
    h = lib.lib()
    assert h.ptx_abi_version() == 1
    assert b"gfx950" in h.ptx_version()


def test_unsupported_sizes_are_rejected_without_gpu(lib):
    h = lib.lib()
    st = h.ptx_fft2d(ctypes.c_void_p(1), 1, 12, 16, 0, None)
    assert st == 2 and b"unsupported" in h.ptx_last_error()
    st = h.ptx_fft2d(ctypes.c_void_p(1), 1, 16, 4096, 0, None)
    assert st == 2
    st = h.ptx_base_forward(None, 64, 64, None, 24, 24, None, 1, None, None, None)
    assert st == 2
    st = h.ptx_base_forward(None, 16, 16, None, 32, 32, None, 1, None, None, None)
    assert st == 3  # probe larger than the sample
    st = h.ptx_base_forward(None, 64, 64, None, 16, 16, None, 1, None, None, None)
    assert st == 1 and b"null" in h.ptx_last_error()


def test_stream_copy_rejects_bad_arguments_without_gpu(lib):
    h = lib.lib()
    assert h.ptx_stream_copy(None, None, 0, None) == 0  # nothing to copy
    st = h.ptx_stream_copy(ctypes.c_void_p(16), ctypes.c_void_p(4096), 24, None)
    assert st == 1 and b"multiple of 16" in h.ptx_last_error()
    st = h.ptx_stream_copy(ctypes.c_void_p(8), ctypes.c_void_p(4096), 32, None)
    assert st == 1 and b"aligned" in h.ptx_last_error()
    st = h.ptx_stream_copy(ctypes.c_void_p(4096), ctypes.c_void_p(4096 + 32), 64, None)
    assert st == 1 and b"overlap" in h.ptx_last_error()


def test_workspace_queries_of_an_empty_batch(lib):
    # no views: every query answers (a zero-view plan once divided by the view count)
    for name, args in (("ptx_base_backward_workspace", (256, 256, 64, 64, 0, 0)),
                       ("ptx_base_step_workspace", (256, 256, 64, 64, 0, 0)),
                       ("ptx_fpm_step_workspace", (256, 256, 64, 64, 0, 0)),
                       ("ptx_nearfield_step_workspace", (256, 256, 64, 64, 0, 0)),
                       ("ptx_coded_workspace", (64, 64, 0, 1, 0)),
                       ("ptx_coded_workspace", (64, 64, 0, 2, 1))):
        assert lib.query(name, *args) >= 0, name


def test_workspace_queries(lib):
    n = lib.query("ptx_base_backward_workspace", 2048, 2048, 512, 512, 64, 0)
    assert n >= 64 * 512 * 512 * 8
    # the default (sorted segments) needs the tile lists on top of the atomic mode's buffers
    n_atomic = lib.query("ptx_base_backward_workspace", 2048, 2048, 512, 512, 64, lib.PTX_FLAG_ATOMIC_SCATTER)
    assert n > n_atomic
    assert lib.query("ptx_base_backward_workspace", 2048, 2048, 512, 512, 64, lib.PTX_FLAG_DETERMINISTIC) == n
    assert lib.query("ptx_base_step_workspace", 2048, 2048, 512, 512, 64, 0) > n
    assert lib.query("ptx_coded_workspace", 1024, 1024, 49, 1, 1) > lib.query("ptx_coded_workspace", 1024, 1024, 49, 1, 0)
    assert lib.query("ptx_coded_workspace", 1024, 1024, 49, 3, 0) == 0  # 1024 not divisible by 3
    assert lib.query("ptx_base_backward_workspace", 2048, 2048, 500, 512, 64, 0) == 0


def test_call_raises_with_message(lib):
    with pytest.raises(ValueError, match="null pointer"):
        lib.call("ptx_base_forward", None, 64, 64, None, 16, 16, None, 1, None, None, None)


def test_adam_argument_checks(lib):
    # every rejection happens on the host before anything is launched
    fake = ctypes.c_void_p(256)
    args = (fake, fake, fake, fake)
    with pytest.raises(ValueError, match="unknown flags"):
        lib.call("ptx_adam_step", *args, 8, 0.9, 0.999, fake, 6, None)
    with pytest.raises(ValueError, match="complex pairs"):
        lib.call("ptx_adam_step", *args, 7, 0.9, 0.999, fake, lib.PTX_ADAM_PROJECT_REAL_NONNEG, None)
    with pytest.raises(ValueError, match="betas"):
        lib.call("ptx_adam_step", *args, 8, 1.0, 0.999, fake, 0, None)
    lib.call("ptx_adam_step", *args, 0, 0.9, 0.999, fake, 0, None)  # n = 0: nothing to do


def test_fused_adam_step_argument_checks(lib):
    # rejected on the host before anything is launched
    fake = ctypes.c_void_p(256)

    def call(adam_s, adam_p, flags=0, count=8):
        refs = [None if a is None else ctypes.c_void_p(ctypes.addressof(a)) for a in (adam_s, adam_p)]
        lib.call("ptx_base_train_step", fake, 256, 256, fake, 64, 64, fake, fake, count, fake, 4, 1.0, fake, None,
                 fake, fake, *refs, fake, 1 << 30, flags, None)

    good = lib.PtxAdam(256, 256, 256, 0.9, 0.999)
    with pytest.raises(ValueError, match="betas"):
        call(lib.PtxAdam(256, 256, 256, 1.0, 0.999), None)
    with pytest.raises(ValueError, match="null Adam state"):
        call(good, lib.PtxAdam(256, 0, 256, 0.9, 0.999))
    with pytest.raises(ValueError, match="ACCUMULATE"):
        call(good, good, flags=lib.PTX_FLAG_ACCUMULATE)
    with pytest.raises(ValueError, match="table of"):
        call(good, good, count=0)


def test_view_table_checks(lib):
    # an empty (or 2^31-entry) table for a non-empty batch is rejected on the
    # host, before anything is launched
    fake = ctypes.c_void_p(256)
    for count in (0, -1, 1 << 31):
        with pytest.raises(ValueError, match="table of"):
            lib.call("ptx_base_step_indexed", fake, 256, 256, fake, 64, 64, fake, fake, count, fake, 4, 1.0,
                     fake, None, fake, fake, fake, 1 << 30, 0, None)
    with pytest.raises(ValueError, match="empty position table"):
        lib.call("ptx_gather_positions", fake, fake, 0, 4, fake, fake, None)
    with pytest.raises(ValueError, match="invalid arguments"):
        lib.call("ptx_gather_positions", fake, fake, -1, 4, fake, fake, None)
    lib.call("ptx_gather_positions", fake, fake, 0, 0, fake, fake, None)  # empty batch: nothing to do


def test_noalloc_step_rejects_oversize_without_gpu(lib):
    # the scratch bound is checked on the host before the scratch array is looked up
    fake = ctypes.c_void_p(256)
    with pytest.raises(lib.NativeLibraryError, match="static maximum"):
        lib.call("ptx_base_step_noalloc", fake, 2048, 2048, fake, 512, 512, fake, 128, fake, 1.0, fake, None,
                 fake, fake, 0, None)
    assert lib.lib().ptx_base_step_noalloc(None, 64, 64, None, 24, 24, None, 1, None, 1.0, None, None, None, None,
                                           0, None) == 2  # unsupported probe side


def test_timing_hook_arms_and_disarms(lib):
    lib.call("ptx_timing_events", None, 0)
    assert lib.lib().ptx_timing_count() == 0
    with pytest.raises(ValueError):
        lib.call("ptx_timing_events", None, 4)


def test_tuning_table_roundtrip(lib):
    h = lib.lib()
    assert h.ptx_tune_get(0, 512) in (8, 16)
    lib.call("ptx_tune_set", 2, 512, 16)
    assert h.ptx_tune_get(2, 512) == 16
    lib.call("ptx_tune_set", 2, 512, 8)
    assert h.ptx_tune_get(2, 512) == 8
    lib.call("ptx_tune_set", 2, 512, 0)  # back to the default
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set", 7, 512, 8)
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set", 0, 512, 4)
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set", 0, 8, 16)
    from ptychox_amd import autotune

    t = autotune.table(autotune.keys_for_scan((512, 512)))
    assert set(t) == {"row:512", "row_loop:512", "col:512", "row:512@groups", "row_loop:512@groups",
                      "col:512@groups"}
    autotune.apply(t)
    assert autotune.table(autotune.keys_for_scan((512, 512))) == t


def test_builtin_schedule_is_radix_8_above_16_points(lib):
    """With no table entry every pass kind runs 8 elements per thread above
    16 points, including the view-looping scatter rows (kind 1), whose
    radix-8 form the autotuner chose on every measured workload."""
    h = lib.lib()
    for kind in range(6):
        for n in (32, 64, 128, 256, 512, 1024, 2048):
            lib.call("ptx_tune_set", kind, n, 0)  # no entry: the built-in choice
            assert h.ptx_tune_get(kind, n) == 8, (kind, n)
        assert h.ptx_tune_get(kind, 16) == 16
        assert h.ptx_tune_get(kind, 8) == 8
    # the scatter rows' built-in work-group target is 512: the step's workspace
    # (its probe-partial slices follow from the target) matches an explicit 512
    # and differs from 1024
    args = (2048, 2048, 512, 512, 64, 0)
    sizes = {}
    for g in (0, 512, 1024):
        lib.call("ptx_tune_set_groups", 1, 512, g)
        sizes[g] = lib.query("ptx_base_step_workspace", *args)
    lib.call("ptx_tune_set_groups", 1, 512, 0)
    assert sizes[0] == sizes[512] != sizes[1024]


def test_group_table_roundtrip(lib):
    h = lib.lib()
    assert h.ptx_tune_get_groups(1, 512) == 0  # built-in default
    lib.call("ptx_tune_set_groups", 1, 512, 64)
    assert h.ptx_tune_get_groups(1, 512) == 64
    # the scatter's work-group target fixes its probe-partial slices, hence the workspace
    before = lib.query("ptx_base_step_workspace", 2048, 2048, 512, 512, 64, 0)
    lib.call("ptx_tune_set_groups", 1, 512, 4096)
    after = lib.query("ptx_base_step_workspace", 2048, 2048, 512, 512, 64, 0)
    assert after > before
    lib.call("ptx_tune_set_groups", 1, 512, 0)
    assert h.ptx_tune_get_groups(1, 512) == 0
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set_groups", 6, 512, 64)
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set_groups", 0, 512, -1)
    with pytest.raises(ValueError):
        lib.call("ptx_tune_set_groups", 0, 512, 1 << 20)
    with pytest.raises(lib.NativeLibraryError, match="unsupported"):
        lib.call("ptx_tune_set_groups", 0, 12, 64)
