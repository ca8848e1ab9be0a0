"""Autotuner for the FFT pass schedule: radix and work-group schedule.

Every FFT pass of libptychox_hip.so is compiled for 8 and for 16 complex
elements per thread (radix-8 or radix-16 butterflies in registers, 64 or 32
threads per 512-point sequence).  Which one is faster depends on the pass kind
(plain row / column pass, or a pass that loops over views to accumulate a
shared gradient) and the transform length.  The passes that walk a chunk of
views per work-group also take a work-group target (``ptx_tune_set_groups``):
fewer, longer-walking work-groups reuse their resident operands more, more
work-groups fill the chip better.  :func:`autotune` times a real workload step
for each candidate, one table entry at a time (coordinate descent, median of
repeated hipEvent timings), and leaves the fastest settings in the library's
process-wide tables.  Results can be saved to JSON and re-applied
(``PTYCHOX_TUNING=<file>`` applies a file at import).
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Callable, Iterable

import torch

from . import _lib

PASS_ROW, PASS_ROW_LOOP, PASS_COL, PASS_COL_LOOP, PASS_VIEW_ROWS, PASS_LOSS_ROWS = 0, 1, 2, 3, 4, 5
KIND_NAMES = {PASS_ROW: "row", PASS_ROW_LOOP: "row_loop", PASS_COL: "col", PASS_COL_LOOP: "col_loop",
              PASS_VIEW_ROWS: "view_rows", PASS_LOSS_ROWS: "loss_rows"}
CANDIDATES = (8, 16)
# work-group targets tried per pass kind (0: the built-in default)
GROUP_CANDIDATES = {PASS_ROW: (0, 256, 512, 2048), PASS_ROW_LOOP: (0, 1024, 2048), PASS_COL: (0, 128, 512),
                    PASS_COL_LOOP: (0, 128, 512), PASS_VIEW_ROWS: (0, 1024, 4096), PASS_LOSS_ROWS: (0,)}
GROUPS_SUFFIX = "@groups"


def get(kind: int, n: int) -> int:
    return int(_lib.lib().ptx_tune_get(kind, n))


def set_entry(kind: int, n: int, elems: int) -> None:
    _lib.call("ptx_tune_set", kind, n, elems)


def get_groups(kind: int, n: int) -> int:
    """The work-group target of (kind, n); 0 means the built-in default."""
    return int(_lib.lib().ptx_tune_get_groups(kind, n))


def set_groups(kind: int, n: int, groups: int) -> None:
    _lib.call("ptx_tune_set_groups", kind, n, groups)


def table(keys: Iterable[tuple[int, int]]) -> dict[str, int]:
    """``{"row:512": 8, "row:512@groups": 0, ...}``: elements per thread and
    work-group target (0 = default) of every key."""
    out = {}
    for k, n in keys:
        out[f"{KIND_NAMES[k]}:{n}"] = get(k, n)
        out[f"{KIND_NAMES[k]}:{n}{GROUPS_SUFFIX}"] = get_groups(k, n)
    return out


def apply(entries: dict[str, int]) -> None:
    """Apply ``{"row:512": 8, "row:512@groups": 1024, ...}`` (the format
    :func:`table` returns; entries without ``@groups`` are elements per thread)."""
    inv = {v: k for k, v in KIND_NAMES.items()}
    for key, value in entries.items():
        groups = key.endswith(GROUPS_SUFFIX)
        kind, n = key[: -len(GROUPS_SUFFIX)].split(":") if groups else key.split(":")
        (set_groups if groups else set_entry)(inv[kind], int(n), int(value))


def save(path, entries: dict[str, int]) -> None:
    Path(path).write_text(json.dumps(entries, indent=1, sort_keys=True))


def load(path) -> dict[str, int]:
    entries = json.loads(Path(path).read_text())
    apply(entries)
    return entries


def time_ms(run: Callable[[], None], reps: int = 10, warmup: int = 2) -> float:
    """Median GPU time of ``run()`` over ``reps`` repetitions (hipEvents)."""
    for _ in range(warmup):
        run()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def autotune(run: Callable[[], None], keys: Iterable[tuple[int, int]], reps: int = 10,
             rounds: int = 1, groups: bool = True) -> dict:
    """Pick, for each ``(kind, n)`` in ``keys``, the elements-per-thread setting
    and (``groups``) the work-group target that minimise the time of ``run()``;
    the tables keep the winners.  ``run`` must size its workspace per call (the
    step classes do): a work-group target can change the workspace size.

    Returns ``{"table": {...}, "before_ms": t0, "after_ms": t1, "trials": [...]}``.
    """
    keys = [(int(k), int(n)) for k, n in keys if int(n) > 8]
    before = time_ms(run, reps)
    trials = []
    for _ in range(rounds):
        for kind, n in keys:
            best, best_t = get(kind, n), None
            for e in CANDIDATES:
                set_entry(kind, n, e)
                t = time_ms(run, reps)
                trials.append({"key": f"{KIND_NAMES[kind]}:{n}", "elems": e, "ms": round(t, 5)})
                if best_t is None or t < best_t:
                    best, best_t = e, t
            set_entry(kind, n, best)
            if groups:
                best, best_t = get_groups(kind, n), None
                for g in GROUP_CANDIDATES[kind]:
                    set_groups(kind, n, g)
                    t = time_ms(run, reps)
                    trials.append({"key": f"{KIND_NAMES[kind]}:{n}{GROUPS_SUFFIX}", "groups": g, "ms": round(t, 5)})
                    if best_t is None or t < best_t:
                        best, best_t = g, t
                set_groups(kind, n, best)
    after = time_ms(run, reps)
    return {"table": table(keys), "before_ms": round(before, 5), "after_ms": round(after, 5), "trials": trials}


def keys_for_scan(probe_shape) -> list[tuple[int, int]]:
    """Table entries a conventional / near-field step uses."""
    h, w = probe_shape
    return [(PASS_ROW, w), (PASS_ROW_LOOP, w), (PASS_COL, h)]


def keys_for_fpm(object_shape, pupil_shape) -> list[tuple[int, int]]:
    H, W = object_shape
    h, w = pupil_shape
    return sorted({(PASS_ROW, w), (PASS_ROW, W), (PASS_ROW_LOOP, w), (PASS_COL, h), (PASS_COL, H)})


def keys_for_coded(shape) -> list[tuple[int, int]]:
    H, W = shape
    return [(PASS_ROW, W), (PASS_LOSS_ROWS, W), (PASS_VIEW_ROWS, W), (PASS_COL, H), (PASS_COL_LOOP, H)]


_env = os.environ.get("PTYCHOX_TUNING")
if _env and Path(_env).exists():
    load(_env)
