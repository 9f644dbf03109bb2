"""Autotuner of the traversal schedule.

The phase-1 contraction splits its K range (SNPs) over ``splits`` workgroup
layers whose int64 partial sums meet atomically, and the phase-1 missing-call
finalize walks the missing-call records in SNP ranges sized to stay resident
in each XCD's L2.  The library picks both by heuristics (splits until the
grid holds 16 waves of resident workgroups with >= 128 chunks per split;
8 MiB ranges for one bin, 64 MiB for several).  ``tune_phase1`` times the candidates on the device
with HIP events around the kernels themselves (``gpe_profile_*``), keeps the
fastest in the library's schedule table for that shape key, and records the
choice in a JSON cache so later processes on the same device and shape start
from it (``apply_cached``).

Shape key: (rows = samples, length = SNPs of the traversal, columns = bins x
probes rounded up to 32 per bin, limbs), as the library looks it up.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import torch

from . import engine
from ._native import c_i32, call

SCHED_P1_SPLITS = 0
SCHED_P1_FINALIZE_KIB = 1
SCHED_P1_CPB = 2
SCHED_P2_CPB = 3
SCHED_P2_MISS = 4
SCHED_P2_MISS_PARTS = 5
SCHED_P1_MISS = 6
_KNOBS = {"p1_splits": SCHED_P1_SPLITS, "p1_finalize_kib": SCHED_P1_FINALIZE_KIB, "p1_cpb": SCHED_P1_CPB,
          "p2_cpb": SCHED_P2_CPB, "p2_miss": SCHED_P2_MISS, "p2_miss_parts": SCHED_P2_MISS_PARTS,
          "p1_miss": SCHED_P1_MISS}


def cache_path() -> str:
    return os.environ.get("SUMMIT_AMD_SCHEDULE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "summit_amd", "schedule.json"))


def _device_key() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|cu{p.multi_processor_count}"


def shape_key(n: int, length: int, n_bins: int, v_count: int, limbs: int) -> tuple:
    return (int(n), int(length), int(n_bins) * ((int(v_count) + 31) // 32 * 32), int(limbs))


def set_value(knob: str, key: tuple, value: int) -> None:
    call("gpe_schedule_set", _KNOBS[knob], key[0], key[1], key[2], key[3], int(value))


def get_value(knob: str, key: tuple) -> int:
    v = c_i32()
    call("gpe_schedule_get", _KNOBS[knob], key[0], key[1], key[2], key[3], ctypes.byref(v))
    return int(v.value)


def clear() -> None:
    call("gpe_schedule_clear")


def _load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, entry_key: str, entry: dict) -> None:
    data = _load_cache(path)
    data[entry_key] = entry
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the choice still holds for this process


def _entry_key(key: tuple) -> str:
    return f"{_device_key()}|n={key[0]}|len={key[1]}|cols={key[2]}|limbs={key[3]}"


def apply_cached(n: int, length: int, n_bins: int, v_count: int, limbs: int = 5, path: str | None = None) -> dict | None:
    """Load a cached choice for this device and shape into the library; returns it (or None)."""
    key = shape_key(n, length, n_bins, v_count, limbs)
    entry = _load_cache(path or cache_path()).get(_entry_key(key))
    if not entry:
        return None
    for knob, value in entry.get("choice", {}).items():
        if knob in _KNOBS:
            set_value(knob, key, value)
    return entry


def _time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, kind) -> float:
    """Median milliseconds of one kernel kind (or the sum of a tuple of kinds)
    over reps phase-1 traversals."""
    kinds = kind if isinstance(kind, tuple) else (kind,)
    times = []
    for r in range(reps):
        engine.profile_enable(True)
        engine.profile_reset()
        engine.phase1_xz(g, 0, L, 1000, annot, inv, r * v_count, v_count, seed, xz, limbs=limbs)
        ms = sum(engine.profile_read(k)[0] for k in kinds)
        engine.profile_enable(False)
        times.append(ms)
    times.sort()
    return times[len(times) // 2]


def tune_phase1(g: engine.Genotype, v_count: int, n_bins: int = 1, limbs: int = 5, seed: int = 1, reps: int = 3,
                split_candidates=(2, 3, 4, 6, 8, 12, 16, 24, 32),
                range_kib_candidates=(1024, 2048, 4096, 8192, 16384, 32768, 65536),
                path: str | None = None, store: bool = True) -> dict:
    """Time phase-1 schedules over every SNP of the prepared handle ``g`` for
    ``n_bins`` bins of ``v_count`` probes; keep and cache the fastest."""
    if g.prepared_with is None:
        raise RuntimeError("genotype handle is not prepared")
    dev = engine.default_device()
    N, L = g.n, g.m
    key = shape_key(N, L, n_bins, v_count, limbs)
    annot = torch.ones((L, n_bins), dtype=torch.float64, device=dev)
    inv = engine.resvar_inv(g, 0, L)
    xz = torch.zeros((n_bins * v_count, N), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    for knob in _KNOBS:
        set_value(knob, key, 0)
    engine.phase1_xz(g, 0, L, 1000, annot, inv, 0, v_count, seed, xz, limbs=limbs)  # warm-up, workspaces
    torch.cuda.synchronize()
    heuristic = {"p1_contract": round(_time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, "p1_contract"), 4)}
    if g.n_missing() > 0:
        heuristic["p1_finalize"] = round(_time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps,
                                                      "p1_finalize"), 4)
    split_ms = {}
    for s in split_candidates:
        set_value("p1_splits", key, s)
        split_ms[int(s)] = round(_time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, "p1_contract"), 4)
    best_split = min(split_ms, key=split_ms.get)
    set_value("p1_splits", key, best_split)
    range_ms = {}
    if g.n_missing() > 0:
        for kib in range_kib_candidates:
            set_value("p1_finalize_kib", key, kib)
            range_ms[int(kib)] = round(_time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, "p1_finalize"), 4)
        best_range = min(range_ms, key=range_ms.get)
        set_value("p1_finalize_kib", key, best_range)
    else:
        best_range = None
    # missing calls: per-call records (best range) against the missing-flag plane, each
    # timed as contraction + missing-call term
    method_ms = {}
    best_method = None
    if g.n_missing() > 0:
        for method in (1, 2):
            set_value("p1_miss", key, method)
            engine.phase1_xz(g, 0, L, 1000, annot, inv, 0, v_count, seed, xz, limbs=limbs)  # workspaces
            method_ms[method] = round(
                _time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, ("p1_contract", "p1_finalize")), 4)
        best_method = min(method_ms, key=method_ms.get)
        set_value("p1_miss", key, best_method)
    # chunks per LDS slot and barrier of the phase-1 contraction
    cpb_ms = {}
    for cpb in (1, 2):
        set_value("p1_cpb", key, cpb)
        cpb_ms[cpb] = round(_time_phase1(g, L, annot, inv, v_count, seed, xz, limbs, reps, "p1_contract"), 4)
    best_cpb = min(cpb_ms, key=cpb_ms.get)
    set_value("p1_cpb", key, best_cpb)
    torch.cuda.synchronize()
    choice = {"p1_splits": best_split, "p1_cpb": best_cpb}
    if best_range is not None:
        choice["p1_finalize_kib"] = best_range
    if best_method is not None:
        choice["p1_miss"] = best_method
    entry = {"choice": choice, "heuristic_ms": heuristic, "p1_contract_ms": split_ms, "p1_finalize_ms": range_ms,
             "p1_miss_ms": method_ms, "p1_cpb_ms": cpb_ms,
             "tune_seconds": round(time.perf_counter() - t0, 3)}
    if store:
        _store_cache(path or cache_path(), _entry_key(key), entry)
    return entry


def p2_key(n_snps: int, n: int, n_bins: int, v_count: int, limbs: int) -> tuple:
    """Phase-2 shape key: (rows = SNPs, length = samples, columns, limbs)."""
    return (int(n_snps), int(n), int(n_bins) * ((int(v_count) + 31) // 32 * 32), int(limbs))


def apply_cached_phase2(n_snps: int, n: int, n_bins: int, v_count: int, limbs: int = 5,
                        path: str | None = None) -> dict | None:
    key = p2_key(n_snps, n, n_bins, v_count, limbs)
    entry = _load_cache(path or cache_path()).get("p2|" + _entry_key(key))
    if not entry:
        return None
    for knob, value in entry.get("choice", {}).items():
        if knob in _KNOBS:
            set_value(knob, key, value)
    return entry


def tune_phase2(g: engine.Genotype, v_count: int, n_bins: int = 1, limbs: int = 5, seed: int = 1, reps: int = 3,
                path: str | None = None, store: bool = True) -> dict:
    """Time the phase-2 contraction with one and two chunks per barrier over
    every SNP of ``g`` against ``n_bins`` x ``v_count`` columns, and, with
    missing calls, both missing-call methods (missing-flag MFMA plane, or sums
    gathered over per-SNP lists: identical results); keep the fastest."""
    if g.prepared_with is None:
        raise RuntimeError("genotype handle is not prepared")
    dev = engine.default_device()
    N, M = g.n, g.m
    key = p2_key(M, N, n_bins, v_count, limbs)
    inv = engine.resvar_inv(g, 0, M)
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    xz = torch.randn((n_bins * v_count, N), generator=gen, dtype=torch.float64).to(dev)
    meansq = torch.zeros((M, n_bins), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    set_value("p2_cpb", key, 0)
    set_value("p2_miss", key, 0)
    methods = (1, 2) if g.n_missing() > 0 else (0,)
    ms = {}
    for method in methods:
        set_value("p2_miss", key, method)
        engine.phase2_accum(g, 0, M, inv, v_count, xz, meansq, N, limbs=limbs)  # warm-up, workspaces
        for cpb in (1, 2):
            set_value("p2_cpb", key, cpb)
            times = []
            for _ in range(reps):
                engine.profile_enable(True)
                engine.profile_reset()
                engine.phase2_accum(g, 0, M, inv, v_count, xz, meansq, N, limbs=limbs)
                times.append(engine.profile_read("p2_contract")[0] + engine.profile_read("p2_miss")[0])
                engine.profile_enable(False)
            times.sort()
            ms[(method, cpb)] = round(times[len(times) // 2], 4)
    best_method, best_cpb = min(ms, key=ms.get)
    set_value("p2_cpb", key, best_cpb)
    set_value("p2_miss", key, best_method)
    choice = {"p2_cpb": best_cpb}
    if best_method:
        choice["p2_miss"] = best_method
    entry = {"choice": choice, "p2_cpb_ms": {cpb: t for (mth, cpb), t in ms.items() if mth == best_method},
             "p2_ms": {f"miss{mth}_cpb{cpb}": t for (mth, cpb), t in ms.items()},
             "tune_seconds": round(time.perf_counter() - t0, 3)}
    if store:
        _store_cache(path or cache_path(), "p2|" + _entry_key(key), entry)
    return entry
