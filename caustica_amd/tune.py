"""Autotuner for the thread-layout variants of the k-space passes.

Each tunable pass (2 zu, 3 zp) has two or three instantiated thread layouts. For a
padded shape, ``autotune`` adds a throwaway member (zero state, the caller's plan)
to the session, times every variant of every pass over a few steps with the
library's event timing (variants interleaved over ``rounds`` rounds, the best round of
each kept), keeps the faster one, and removes the member. Choices are
cached per (shape, library version) in a JSON file (``$CAUSTICA_AMD_TUNE_CACHE`` or
``~/.cache/caustica_amd/autotune.json``); an unwritable cache only costs re-tuning.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

from caustica_amd import _native as N

PASSES = {2: "zu", 3: "zp"}
VARIANTS = {2: (0, 1, 2), 3: (0, 1)}
_memo: dict = {}


def _cache_path() -> Path:
    env = os.environ.get("CAUSTICA_AMD_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "caustica_amd" / "autotune.json"


def _key(padded3) -> str:
    return f"v{N.lib().cst_version()}:{'x'.join(map(str, padded3))}"


def _load() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except Exception:  # noqa: BLE001 - no cache yet / unreadable
        return {}


def _store(d: dict) -> None:
    try:
        p = _cache_path()
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(d, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass


def apply(sess, choice: dict) -> None:
    for p, v in choice.items():
        sess.set_variant(int(p), int(v))


def autotune(sess, plan, steps: int = 6, use_cache: bool = True, rounds: int = 2) -> dict:
    """Pick the fastest variant of each tunable pass for the session's shape."""
    key = _key(sess.padded3)
    if use_cache:
        if key in _memo:
            apply(sess, _memo[key])
            return _memo[key]
        disk = _load()
        if key in disk:
            choice = {int(k): int(v) for k, v in disk[key].items()}
            _memo[key] = choice
            apply(sess, choice)
            return choice
    slot = sess.add(plan, max_periods=2 + (2 + 2 * rounds * (steps + 1)) * len(PASSES) // max(plan.spp, 1))
    choice = {}
    try:
        sess.advance([slot], 2)  # first-launch costs out of the way
        for p, name in PASSES.items():
            best = {}
            # interleaved rounds, best per-launch time of each variant: robust to drift
            for _ in range(rounds):
                for v in VARIANTS[p]:
                    sess.set_variant(p, v)
                    sess.advance([slot], 1)
                    sess.set_profiling(True)
                    sess.advance([slot], steps)
                    ms, cnt = sess.kernel_times()[name]
                    sess.set_profiling(False)
                    t = ms / max(cnt, 1)
                    best[v] = min(best.get(v, t), t)
            choice[p] = min(best, key=best.get)
            sess.set_variant(p, choice[p])
    finally:
        sess.remove(slot)
    _memo[key] = choice
    if use_cache:
        d = _load()
        d[key] = choice
        _store(d)
    return choice
