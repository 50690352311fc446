"""Autotuner front end with a repo-local JSON cache.

The tuning itself runs in the HIP library (rplsh_bf_autotune): the k' factor
is the smallest whose recall@k against the plain path on a query sample meets
the target; the filter grid size and the last-level survivor target are
picked by measured batch time. Results are cached per (precision, d, log2 n,
k, batch size) so a later process installs them without re-measuring.
"""
from __future__ import annotations

import json
import math
import os

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json")

_PREC_NAME = {0: "fp32", 1: "bf16", 2: "fp8", "fp32": "fp32", "plain": "fp32", "bf16": "bf16",
              "fp8": "fp8"}


def cache_key(index, k: int, batch_nq: int) -> str:
    n = max(1, index.num_data)
    prec = _PREC_NAME[index.config.precision]
    return f"gfx950|{prec}|d{index.dim}|n2^{round(math.log2(n))}|k{k}|b{batch_nq}"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_cache(cache: dict, path: str = DEFAULT_CACHE) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(index, sample, k: int, batch_nq: int, target_recall: float = 0.999,
             cache_path: str | None = DEFAULT_CACHE, refresh: bool = False) -> dict:
    """Installs a tuning for (k, batch_nq) on index, from the cache when
    present (unless refresh), else by measuring; returns the entry."""
    cache = load_cache(cache_path) if cache_path else {}
    key = cache_key(index, k, batch_nq)
    if not refresh and key in cache:
        index.set_tuning(cache[key])
        return cache[key]
    entry = index.autotune(sample, k, batch_nq, target_recall)
    if cache_path:
        cache = load_cache(cache_path)
        cache[key] = entry
        save_cache(cache, cache_path)
    return entry
