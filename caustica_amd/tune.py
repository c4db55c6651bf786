"""Autotuner for the radix plans, tile shapes and thread layouts of the k-space passes.

Tunable passes and their variants (cst_session_set_variant):
  0 the y passes (k_ygrad, k_ydiv, k_ypass): column-tile width 0 (default), 1 (half),
    2 (twice);
  1 the x passes (k_xgrad, k_xdiv): the same tile widths;
  2 the velocity z pass (k_zu): thread layout 0..2, + 4 for the second radix plan of the
    length-nz/2 transform (the stages' lengths swapped);
  3 the pressure z pass (k_zp): layout 0..1, + 4 for the second radix plan.
For a padded shape, ``autotune`` adds a throwaway member (zero state, the caller's plan)
to the session, times every variant of every pass over a few steps with the library's
event timing (variants interleaved over ``rounds`` rounds, the best round of each kept),
keeps the fastest, and removes the member. Variants that map to the same kernel (a
second radix plan where both stages have the same length; tile widths at lengths below
64) are skipped. Choices are cached per (shape, library version) in a JSON file
(``$CAUSTICA_AMD_TUNE_CACHE`` or ``~/.cache/caustica_amd/autotune.json``); an unwritable
cache only costs re-tuning.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

from caustica_amd import _native as N

PASSES = {0: ("y3", "y1"), 1: ("xgrad", "xdiv"), 2: ("zu",), 3: ("zp",)}
VARIANTS = {0: (0, 1, 2), 1: (0, 1, 2), 2: (0, 1, 2, 4, 5, 6), 3: (0, 1, 4, 5)}


def _plan_a(n: int) -> int:
    for a in range(1, n + 1):
        if n % a == 0 and a * a >= n:
            return a
    return n


def variants_for(pass_: int, padded3) -> tuple:
    """The distinct variants of a pass at this shape."""
    nx, ny, nz = padded3
    vs = VARIANTS[pass_]
    if pass_ in (0, 1):
        length = ny if pass_ == 0 else nx
        return vs if length >= 64 else (0,)
    m = nz // 2
    if _plan_a(m) ** 2 == m:
        return tuple(v for v in vs if v < 4)
    return vs


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
    n_var = sum(len(variants_for(p, sess.padded3)) for p in PASSES)
    slot = sess.add(plan, max_periods=2 + (2 + rounds * (steps + 1) * n_var) // max(plan.spp, 1))
    choice = {}
    try:
        sess.advance([slot], 2)  # first-launch costs out of the way
        for p, names in PASSES.items():
            vs = variants_for(p, sess.padded3)
            if len(vs) == 1:
                choice[p] = vs[0]
                sess.set_variant(p, vs[0])
                continue
            best = {}
            # interleaved rounds, best per-step time of each variant: robust to drift
            for _ in range(rounds):
                for v in vs:
                    sess.set_variant(p, v)
                    sess.advance([slot], 1)
                    sess.set_profiling(True)
                    sess.advance([slot], steps)
                    kt = sess.kernel_times()
                    sess.set_profiling(False)
                    t = sum(kt[n][0] for n in names) / steps
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
