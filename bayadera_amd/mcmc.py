"""The MCMC API over the engine's samplers.

Mirrors ``src/clojure/uncomplicate/bayadera/mcmc.clj``: default step counts,
annealing schedules and ``mix!``, the adaptive tuning of the stretch
parameter a.  Host-side control; every step runs through the sampler's
C-ABI calls."""
from __future__ import annotations

import math


def init_position(samp, seed_or_position, limits=None):
    return samp.init_position(seed_or_position, limits)


def acc_rate(samp, a: float = 2.0) -> float:
    """mcmc.clj:21-25"""
    return float(samp.acc_rate(a))


def burn_in(samp, steps: int | None = None, a: float = 2.0):
    """mcmc.clj:27-33: 256 * dimension steps by default."""
    if steps is None:
        steps = 256 * samp.model().dimension
    return samp.burn_in(int(steps), a)


def run_sampler(samp, steps: int | None = None, a: float = 2.0):
    """mcmc.clj:35-41: 64 * dimension steps by default."""
    if steps is None:
        steps = 64 * samp.model().dimension
    return samp.run_sampler(int(steps), a)


def sqrt_n(temp: float):
    """mcmc.clj:43-46"""
    return lambda i: math.sqrt(temp - i)


def pow_n(power: float):
    """mcmc.clj:48-51"""
    return lambda temp: (lambda i: math.pow(temp - i, power))


def minus_n(temp: float):
    """mcmc.clj:52-54"""
    return lambda i: temp - i


def anneal(samp, steps: int | None = None, a: float = 2.0, schedule=minus_n):
    """mcmc.clj:56-64: temperatures schedule(steps)(i), i = 0 .. steps-1."""
    if steps is None:
        steps = 256 * samp.model().dimension
    return samp.anneal(schedule(float(steps)), int(steps), a)


def mix(samp, step: int = 64, dimension_power: float = 0.8, cooling_schedule=minus_n,
        a: float = 2.0, min_acc_rate: float = 0.2, max_acc_rate: float = 0.5) -> dict:
    """mix! (mcmc.clj:66-101): anneal, then adjust a until the acceptance
    rate lies in [min, max] (at most `step` rounds), then burn in."""
    dimension = samp.model().dimension
    target = (max_acc_rate + min_acc_rate) / 2.0
    n = int(step * math.pow(dimension, dimension_power))
    samp.anneal(cooling_schedule(float(n)), n, a)
    i = 0
    while True:
        rate = samp.acc_rate(a)
        if step < i:
            break
        if rate < min_acc_rate:
            a = 1.0 + (a - 1.0) * (rate / target)
        elif max_acc_rate < rate:
            a = a * (rate / target)
        else:
            break
        i += 1
    samp.burn_in(n, a)
    samp.burn_in(step, 2.0)
    return {"a": a, "acc-rate": samp.acc_rate(a), "acc-rate-2.0": samp.acc_rate(2.0)}
