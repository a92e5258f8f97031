"""Sampled / exact CP-ALS sweeps over an MI355X-resident ALS handle.

Contract kept from the reference driver (als.py:59-107): each sweep updates
every mode once with ``renormalize=True`` and ``update_sampler=True``; the
exact fit is evaluated before the first sweep and after every
``epoch_length``-th sweep; the run stops early once more than four fits exist
and the best of the newest three does not exceed the best of the older ones by
``stop_tolerance``; the result is a dict with the keys ``iterations``,
``fits``, ``update_times`` and ``fit_computation_times``.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from .als import ALS


def stalled(fits, stop_tolerance) -> bool:
    """True when the newest three fits add less than ``stop_tolerance`` to the
    best fit seen before them (needs at least five fits)."""
    if len(fits) <= 4:
        return False
    return max(fits[-3:]) < max(fits[:-3]) + stop_tolerance


@dataclass
class _Trajectory:
    iterations: list = field(default_factory=list)
    fits: list = field(default_factory=list)
    update_times: list = field(default_factory=list)
    fit_computation_times: list = field(default_factory=list)

    def as_dict(self):
        return {"iterations": self.iterations, "fits": self.fits, "update_times": self.update_times,
                "fit_computation_times": self.fit_computation_times}


def als_prod(lhs, rhs, J, method, max_iter, stop_tolerance, epoch_length=5, verbose=False,
             precision="fp64", seed=None, host_sync=True):
    """Run up to ``max_iter`` CP-ALS sweeps of ``lhs`` (a ``PyLowRank``) against
    ``rhs`` (a ``PySparseTensor``).

    ``method``: "exact" (``execute_exact_als_update``) or a sampler name of
    ``initialize_ds_als`` ("efficient", "larsen_kolda", "larsen_kolda_hybrid")
    for ``execute_ds_als_update`` with J samples per update.  ``precision``
    and ``seed`` configure the sampler (its built-in uniforms start at
    ``seed`` when given).  With ``host_sync=False`` the factors stay on the
    device between updates and are copied back once at the end.
    """
    handle = ALS(lhs.ten, rhs.ten, host_sync=host_sync)
    exact = method == "exact"
    if not exact:
        handle.initialize_ds_als(J, method, precision=precision, seed=seed)
    update = handle.execute_exact_als_update if exact else handle.execute_ds_als_update

    record = _Trajectory()

    def measure_fit(sweeps_done):
        t0 = time.time()
        value = lhs.compute_exact_fit(rhs)
        record.iterations.append(sweeps_done)
        record.fits.append(value)
        return value, time.time() - t0

    fit0, _ = measure_fit(0)
    if verbose:
        print(f"initial fit {fit0}")
    for sweep in range(1, max_iter + 1):
        for mode in range(lhs.N):
            t0 = time.time()
            update(mode, True, True)
            record.update_times.append(time.time() - t0)
        if sweep % epoch_length:
            continue
        value, spent = measure_fit(sweep)
        record.fit_computation_times.append(spent)
        if verbose:
            print(f"sweep {sweep}: fit {value}")
        if stalled(record.fits, stop_tolerance):
            break
    if not host_sync:
        handle.sync_to_host()
    return record.as_dict()
