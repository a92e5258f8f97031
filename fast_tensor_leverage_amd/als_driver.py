"""als_prod: the sampled CP-ALS sweep loop of the reference (als.py:59-107)."""
from __future__ import annotations

import time

import numpy as np

from .als import ALS


def als_prod(lhs, rhs, J, method, max_iter, stop_tolerance, epoch_length=5, verbose=False,
             precision="fp64", seed=None, host_sync=True):
    """Sweeps of ``execute_ds_als_update(j, True, True)`` (method
    "efficient" or "larsen_kolda", the sampler of initialize_ds_als) or
    ``execute_exact_als_update(j, True, True)`` (method "exact") over the modes, an exact fit every ``epoch_length`` sweeps,
    stopped when the best of the last three fits does not beat the best before
    them by ``stop_tolerance`` (als.py:59-107).  Returns the reference's record
    (iterations, fits, update_times, fit_computation_times)."""
    def verb_print(arg):
        if verbose:
            print(arg)

    als = ALS(lhs.ten, rhs.ten, host_sync=host_sync)
    if method != "exact":
        als.initialize_ds_als(J, method, precision=precision, seed=seed)

    iterations, fits, update_times, fit_computation_times = [0], [lhs.compute_exact_fit(rhs)], [], []
    verb_print(f"Before ALS:\tFit: {fits[-1]}")
    for i in range(max_iter):
        verb_print(f"Starting Iteration {i + 1}.")
        for j in range(lhs.N):
            start = time.time()
            if method == "exact":
                als.execute_exact_als_update(j, True, True)
            else:
                als.execute_ds_als_update(j, True, True)
            update_times.append(time.time() - start)
        if (i + 1) % epoch_length == 0:
            start = time.time()
            iterations.append(i + 1)
            fits.append(lhs.compute_exact_fit(rhs))
            verb_print(f"Iteration: {i + 1}\tFit: {fits[-1]}")
            fit_computation_times.append(time.time() - start)
            if len(fits) > 4 and np.max(fits[-3:]) < np.max(fits[:-3]) + stop_tolerance:
                break
    if not host_sync:
        als.sync_to_host()
    return {"iterations": iterations, "fits": fits, "update_times": update_times,
            "fit_computation_times": fit_computation_times}
