"""Rebuild-cadence autotuner of one NeighborList system.

The skin of a NeighborList (neighbor_list.py:137-286) is a performance knob:
the list holds every pair within cutoff + skin and is rebuilt once a sphere
has moved skin / 2 (neighbor_list.py:82-85).  A small skin rebuilds often; a
large one lists more pairs that every step tests, and past 64 listed slots a
row takes the wider form of the step kernel.  The skin does not change which
pairs touch, so every candidate gives the same forces up to the order of the
sums (the list's cell grid, hence its traversal order, follows cutoff +
skin).
"""

from __future__ import annotations

import time
from typing import Callable, Sequence

import torch


def autotune_skin_system(make_pair: Callable[[float], tuple], skins: Sequence[float], probe_steps: int = 200,
                         warm_steps: int = 20) -> tuple:
    """For every candidate skin, builds ``make_pair(skin)`` -> (state,
    system), steps ``warm_steps`` untimed (first list build) and
    ``probe_steps`` timed with ``System.step``, and keeps the skin with the
    highest particle-steps/s.  Returns (best skin, [(skin, particle-steps/s,
    rebuilds, max_neighbors), ...])."""
    from .system import System

    table = []
    for skin in skins:
        st, sy = make_pair(float(skin))
        st, sy = System.step(st, sy, n=warm_steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st2, sy2 = System.step(st, sy, n=probe_steps)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        stats = sy2.__dict__.get("_last_stats")
        table.append((float(skin), st.N * probe_steps / secs, int(stats.rebuilds) if stats is not None else -1,
                      int(sy.collider.max_neighbors)))
        del st, sy, st2, sy2
        torch.cuda.empty_cache()
    best = max(table, key=lambda r: r[1])[0]
    return best, table


__all__ = ["autotune_skin_system"]
