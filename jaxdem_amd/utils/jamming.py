"""Jamming protocols (utils/jamming.py:28-793; utils/packing_utils.py:21-113),
driving the GPU ``minimize``.

The protocol (O'Hern et al., PRE 68, 011306 (2003)): minimise; while the
minimum has no overlaps (potential energy per particle <= ``pe_tol``),
compress affinely by ``packing_fraction_increment``; once a compression
jams, bisect between the last unjammed and the first jammed packing
fraction, restarting each trial from the last unjammed state, until the two
agree to ``packing_fraction_tolerance``.  Every minimisation is one
``jdem_minimize`` call on the device; the bisection bookkeeping is host
arithmetic in float32, as the reference's jitted loop computes it.
Spheres only: a body is a single sphere, so the affine rescale moves each
sphere centre (packing_utils.py:91-102 with one sphere per group).
"""

from __future__ import annotations

from typing import Any, NamedTuple

import numpy as np

from .packing_utils import compute_packing_fraction, compute_particle_volume, scale_to_packing_fraction

F32 = np.float32


class JamResult(NamedTuple):
    """Result of ``bisection_jam`` (jamming.py:28-50), in the reference's
    field order."""

    unjammed_state: Any
    unjammed_system: Any
    jammed_state: Any
    jammed_system: Any
    packing_fraction: float
    potential_energy: float


def bisection_jam(
    state,
    system,
    n_minimization_steps: int = 1000000,
    pe_tol: float = 1e-16,
    pe_diff_tol: float = 1e-16,
    n_jamming_steps: int = 10000,
    packing_fraction_tolerance: float = 1e-10,
    packing_fraction_increment: float = 1e-3,
    verbose: bool = True,
) -> JamResult:
    """Nearest jammed state by bisection with state reversion
    (jamming.py:53-267).  Returns a ``JamResult``; ``packing_fraction`` is
    the packing fraction of the last trial, as in the reference."""
    from ..minimizers import minimize

    def relax(s, y):
        return minimize(s, y, max_steps=n_minimization_steps, pe_tol=pe_tol, pe_diff_tol=pe_diff_tol)

    state, system, n_steps, final_pe = relax(state, system)
    is_jammed = final_pe > pe_tol
    if verbose:
        if is_jammed:
            print(f"Warning: Initial state is already jammed (PE={final_pe} > tol={pe_tol}). Skipping.")
        print(f"Initial minimization took {n_steps} steps.")
    pf = compute_packing_fraction(state, system)
    pf_low, pf_high = pf, F32(-1.0)
    cur_state, cur_system = state, system
    last_state, last_system = state, system
    inc = F32(packing_fraction_increment)
    i = 0
    while i < n_jamming_steps and not is_jammed:
        s, y, n_steps, final_pe = relax(cur_state, cur_system)
        if final_pe > pe_tol:  # jammed: revert to the last unjammed state, bisect below this fraction
            new_pf_high = pf
            new_pf_low = pf_low
            new_pf = F32((new_pf_high + pf_low) / F32(2.0))
            new_state, new_system = last_state, last_system
        else:  # unjammed: keep it, then bisect (a jammed fraction is known) or compress further
            last_state, last_system = s, y
            new_pf_low = pf
            new_pf_high = pf_high
            new_pf = F32((pf_high + new_pf_low) / F32(2.0)) if pf_high > 0 else F32(new_pf_low + inc)
            new_state, new_system = s, y
        ratio = F32(new_pf_high / new_pf_low)
        is_jammed = bool(abs(float(ratio) - 1.0) < packing_fraction_tolerance and new_pf_high > 0)
        if verbose:
            print(f"Step: {i + 1} -  phi={pf}, PE={final_pe} after {n_steps} steps")
        cur_state, cur_system = scale_to_packing_fraction(new_state, new_system, new_pf)
        pf, pf_low, pf_high = new_pf, new_pf_low, new_pf_high
        i += 1
    last_jammed_pf = pf_high if pf_high > 0 else pf
    js, jy = scale_to_packing_fraction(last_state, last_system, last_jammed_pf)
    js, jy, _, final_pe = relax(js, jy)
    return JamResult(last_state, last_system, js, jy, float(pf), float(final_pe))


def pe_band_jam(
    state,
    system,
    n_minimization_steps: int = 1_000_000,
    pe_tol: float = 1e-16,
    pe_diff_tol: float = 1e-16,
    pe_band_factor: float = 2.0,
    packing_fraction_increment: float = 1e-3,
    n_jamming_steps: int = 10_000,
    verbose: bool = True,
) -> JamResult:
    """Jammed state in a potential-energy band (jamming.py:585-793): minimise;
    below the band [pe_tol, pe_band_factor * pe_tol] compress by the current
    increment, above it expand, inside it stop.  The increment halves at every
    reversal of direction.  There is no reversion: each trial rescales the
    state just minimised.  ``unjammed_state`` is the last state seen below the
    band (the input if none); the result's packing fraction is that of the
    final, re-minimised state."""
    from ..minimizers import minimize

    def relax(s, y):
        return minimize(s, y, max_steps=n_minimization_steps, pe_tol=pe_tol, pe_diff_tol=pe_diff_tol)

    pe_lo = pe_tol
    pe_hi = pe_band_factor * pe_tol
    pf = compute_packing_fraction(state, system)
    inc = F32(packing_fraction_increment)
    prev_dir = 0
    last_state, last_system = state, system
    i = 0
    done = False
    while i < n_jamming_steps and not done:
        state, system, n_steps, pe = relax(state, system)
        below, above = pe < pe_lo, pe > pe_hi
        done = not (below or above)
        direction = 1 if below else (-1 if above else 0)
        if prev_dir != 0 and direction != 0 and direction != prev_dir:
            inc = F32(F32(0.5) * inc)
        new_pf = F32(pf + F32(direction) * inc)
        if below:
            last_state, last_system = state, system
        if verbose:
            print(f"Step: {i + 1} - phi={pf}, increment={inc}, PE/N={pe} after {n_steps} steps")
        if not done:
            prev_dir = direction
            state, system = scale_to_packing_fraction(state, system, new_pf)
            pf = new_pf
        i += 1
    # relaxed on return (a no-op after an in-band exit; the last rescale is
    # still unminimised when the step cap ends the loop)
    state, system, _, final_pe = relax(state, system)
    return JamResult(last_state, last_system, state, system, float(compute_packing_fraction(state, system)),
                     float(final_pe))


def pressure_bisection_jam(
    state,
    system,
    *,
    n_minimization_steps: int = 1_000_000,
    pe_tol: float = 1e-16,
    pe_diff_tol: float = 1e-16,
    pressure_threshold: float = 1e-7,
    pressure_band_factor: float = 1.01,
    growth_rate: float = 1.001,
    fine_growth_rate: float = 1.000001,
    length_ratio_tolerance: float = 1e-14,
    n_jamming_steps: int = 10_000,
    verbose: bool = True,
) -> JamResult:
    """Jammed state in a pressure band (jamming.py:270-583).  The control
    variable is the box length L = prod(box)^(1/dim), in float64 on the host;
    the pressure is the GPU contact virial's (contacts.py:89-164).  A coarse
    phase divides L by ``growth_rate`` while the pressure is below
    [P_lo, P_hi] = [pressure_threshold, pressure_band_factor * P_lo] and
    stops at the first over-compression; a fine phase bisects in L (growing by
    ``fine_growth_rate`` when no over-compressed bound is known) until the
    pressure is in the band or |L_hi / L_lo - 1| < ``length_ratio_tolerance``.
    Every trial box rescales the last under-compressed state."""
    from ..minimizers import minimize
    from .contacts import compute_pressure

    def relax(s, y):
        return minimize(s, y, max_steps=n_minimization_steps, pe_tol=pe_tol, pe_diff_tol=pe_diff_tol)

    p_lo = float(pressure_threshold)
    p_hi = float(pressure_band_factor) * p_lo
    dim = int(state.dim)
    volume = float(compute_particle_volume(state))

    def length_of(y) -> float:
        return float(np.prod(np.asarray(y.domain.box_size, F32), dtype=F32)) ** (1.0 / dim)

    def pf_for_length(length: float) -> float:
        return volume / (length**dim)

    state, system, _, final_pe = relax(state, system)
    pressure = compute_pressure(state, system)
    if pressure > p_hi:
        if verbose:
            print(f"Warning: Initial state is already over-compressed (P={pressure:.5e} > P_hi={p_hi:.5e}). Skipping.")
        return JamResult(state, system, state, system, float(compute_packing_fraction(state, system)),
                         float(final_pe))
    length = length_of(system)
    length_hi = -1.0  # largest under-compressed box seen
    length_lo = -1.0  # smallest over-compressed box seen
    last_state, last_system = state, system
    iteration = 0

    def do_step(state, system, rate, break_on_over, check_convergence):
        nonlocal last_state, last_system, length, length_hi, length_lo, final_pe
        state, system, _, pe = relax(state, system)
        p = compute_pressure(state, system)
        final_pe = float(pe)
        if verbose:
            print(f"Step {iteration}: L={length:.8e}, phi={pf_for_length(length):.8e}, P={p:.5e}, PE={final_pe:.5e}")
        status = "continue"
        if p < p_lo:  # under-compressed: compress (bisect once a bound is known)
            last_state, last_system = state, system
            length_hi = length
            if length_lo > 0.0:
                length = 0.5 * (length_hi + length_lo)
                length_lo = -1.0
            else:
                length /= rate
        elif p > p_hi:  # over-compressed: record the bound and bisect
            length_lo = length
            length = 0.5 * (length_hi + length_lo)
            if break_on_over:
                status = "break"
        else:
            status = "success"
        if check_convergence and length_hi > 0.0 and length_lo > 0.0 and abs(length_hi / length_lo - 1.0) < \
                length_ratio_tolerance:
            status = "converged"
        if status in ("continue", "break"):
            state, system = scale_to_packing_fraction(last_state, last_system, pf_for_length(length))
        return state, system, status

    status = "continue"
    while iteration < n_jamming_steps and status == "continue":
        state, system, status = do_step(state, system, growth_rate, True, False)
        iteration += 1
    success = status == "success"
    if not success:
        status = "continue"
        while iteration < n_jamming_steps and status == "continue":
            state, system, status = do_step(state, system, fine_growth_rate, False, True)
            iteration += 1
        success = status == "success"
    if verbose and not success:
        print("Warning: pressure band not reached; returning the marginally jammed bracket bound.")
    jammed_length = length if success else (length_lo if length_lo > 0.0 else length)
    js, jy = scale_to_packing_fraction(last_state, last_system, pf_for_length(jammed_length))
    js, jy, _, final_pe = relax(js, jy)
    return JamResult(last_state, last_system, js, jy, float(compute_packing_fraction(js, jy)), float(final_pe))


__all__ = [
    "JamResult",
    "bisection_jam",
    "pe_band_jam",
    "pressure_bisection_jam",
    "compute_packing_fraction",
    "compute_particle_volume",
    "scale_to_packing_fraction",
]
