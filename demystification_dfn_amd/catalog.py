"""Event catalog and seismic moment from the device monitor's records.

The reference does this in MATLAB on its saved output: an event starts when max V
on a fault reaches V_threshold while no event is going on, and ends at the first
saved time at which max V on that same fault is back at or below the threshold
(wrapper/microseismicity.m:447-472); the seismic moment of a fault is
sum(mu * slip * ds) with slip = cumtrapz(t, V) (wrapper/network.m:474-484).  The
device gathers the inputs (dfn_monitor, dfn_moment); this module applies the
catalog rule to them, with the reference's quirks kept: one event at a time, the
faults scanned in order from `first_fault` (the reference script starts at its
second fault), the start test with >= and the end test with <=, and an event
still going on at the last record left out.
"""
from __future__ import annotations

import numpy as np


def event_catalog(times, maxV, threshold=1e-3, first_fault=2):
    """Events as (fault, start record, end record, t_start, t_end); fault numbers
    and record indices 1-based as in the reference; maxV[i, f] is max |V| on fault
    f + 1 at record i."""
    maxV = np.asarray(maxV, dtype=np.float64)
    times = np.asarray(times, dtype=np.float64)
    nrec, nfault = maxV.shape
    catalog = []
    on = False
    cur = first_fault
    start = 0
    for i in range(nrec):
        for f in range(first_fault, nfault + 1):
            if maxV[i, f - 1] >= threshold and not on:
                on = True
                cur = f
                start = i
        if on and maxV[i, cur - 1] <= threshold:
            on = False
            catalog.append((cur, start + 1, i + 1, float(times[start]), float(times[i])))
    return catalog


def moment(mu, slip, ds):
    """M0 = sum(mu * slip * ds) over a fault's elements, summed in element order."""
    m = 0.0
    for s, d in zip(np.asarray(slip, np.float64), np.asarray(ds, np.float64)):
        m = m + mu * s * d
    return m
