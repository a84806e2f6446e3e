"""Step indices at which a rollout saves a frame (utils/rollout_schedules.py:20-90).

Both schedules return ascending step indices on the grid 0..num_steps; the
strides for ``System.trajectory_rollout`` are their differences,
``np.diff(steps, prepend=0)`` (a leading 0 stride records the initial state).
"""

from __future__ import annotations

import numpy as np


def make_save_steps_linear(*, num_steps: int, save_freq: int, include_step0: bool = True) -> np.ndarray:
    """Every ``save_freq``-th step up to ``num_steps``, starting at step 0 or
    at ``save_freq`` (rollout_schedules.py:20-36)."""
    num_steps = int(num_steps)
    save_freq = int(save_freq)
    if num_steps < 0:
        raise ValueError("num_steps must be >= 0")
    if save_freq < 1:
        raise ValueError("save_freq must be >= 1")
    first = 0 if include_step0 else save_freq
    return np.arange(first, num_steps + 1, save_freq, dtype=int) if first <= num_steps else np.zeros((0,), int)


def make_save_steps_pseudolog(
    *,
    num_steps: int,
    reset_save_decade: int,
    min_save_decade: int,
    decade: int = 10,
    include_step0: bool = True,
    cap: int | None = None,
) -> np.ndarray:
    """Pseudo-logarithmic schedule (rollout_schedules.py:39-90).

    The step grid is cut into blocks of ``reset_save_decade`` steps.  An
    offset o in (0, reset] from a block's start is saved when it is a
    multiple of the spacing of its decade: f0 for o <= f0 * decade, and
    f0 * decade**k for f0 * decade**k < o <= f0 * decade**(k + 1), with f0 =
    ``min_save_decade``.  Step 0 is saved when ``include_step0``; a block's
    own start is the previous block's offset ``reset``.  ``cap`` keeps the
    first ``cap`` steps (at least one, as the reference returns after its
    first saved step when ``cap`` <= 1).
    """
    num_steps = int(num_steps)
    reset = int(reset_save_decade)
    f0 = int(min_save_decade)
    decade = int(decade)
    if num_steps < 0:
        raise ValueError("num_steps must be >= 0")
    if reset < 1:
        raise ValueError("reset_save_decade must be >= 1")
    if f0 < 1:
        raise ValueError("min_save_decade must be >= 1")
    if decade < 2:
        raise ValueError("decade must be >= 2")
    t = np.arange(1, num_steps + 1, dtype=np.int64)
    off = (t - 1) % reset + 1  # in 1..reset
    spacing = np.full(off.shape, f0, dtype=np.int64)
    upper = f0 * decade  # upper end of the current decade
    while True:
        above = off > upper
        if not above.any():
            break
        spacing = np.where(above, upper, spacing)
        upper *= decade
    keep = t[off % spacing == 0]
    if include_step0:
        keep = np.concatenate([np.zeros(1, np.int64), keep])
    if cap is not None:
        keep = keep[: max(int(cap), 1)]
    return keep.astype(int)


def strides_from_steps(steps) -> np.ndarray:
    """Per-frame strides for ``System.trajectory_rollout(strides=...)``."""
    s = np.asarray(steps, dtype=np.int64).reshape(-1)
    if s.size and (s[0] < 0 or np.any(np.diff(s) < 0)):
        raise ValueError("save steps must be ascending and non-negative")
    return np.diff(s, prepend=0)


__all__ = ["make_save_steps_linear", "make_save_steps_pseudolog", "strides_from_steps"]
