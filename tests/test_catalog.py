"""The event-catalog rule (demystification_dfn_amd/catalog.py) on hand-built
max-V series: start at >= threshold with no event going on, end on the same
fault at <= threshold, one event at a time, faults scanned from first_fault, an
unfinished event left out (wrapper/microseismicity.m:447-472)."""
import numpy as np

from demystification_dfn_amd.catalog import event_catalog, moment


def test_known_answer_catalog():
    t = np.arange(10) * 10.0
    v = np.full((10, 3), 1e-9)
    v[2:5, 1] = 1e-2        # fault 2: records 3..5 above, back down at record 6
    v[3:8, 2] = 1e-1        # fault 3 overlaps: ignored while fault 2's event goes on, then still high
    v[8, 2] = 1e-3          # exactly at the threshold: ends the event (<=)
    v[0:2, 0] = 1.0         # fault 1 is not scanned with first_fault = 2
    cat = event_catalog(t, v, 1e-3, first_fault=2)
    # fault 2: starts at record 3, ends at record 6; fault 3 starts at record 6
    # (fault 2 ended first within that record? no: starts are tested before ends,
    # so at record 6 fault 2 is still the event in progress), then fault 3 from 7 to 9
    assert cat == [(2, 3, 6, 20.0, 50.0), (3, 7, 9, 60.0, 80.0)]


def test_first_fault_and_unfinished_event():
    t = np.arange(6) * 1.0
    v = np.full((6, 2), 1e-9)
    v[1:3, 0] = 5e-3
    v[4:, 1] = 5e-3         # still going on at the last record
    assert event_catalog(t, v, 1e-3, first_fault=1) == [(1, 2, 4, 1.0, 3.0)]
    assert event_catalog(t, v, 1e-3, first_fault=2) == []


def test_start_and_end_in_one_record_at_the_threshold():
    t = np.arange(3) * 1.0
    v = np.full((3, 1), 1e-9)
    v[1, 0] = 1e-3          # >= starts and <= ends in the same record
    assert event_catalog(t, v, 1e-3, first_fault=1) == [(1, 2, 2, 1.0, 1.0)]


def test_moment_sum():
    assert moment(3e10, [1.0, 2.0, 0.5], [10.0, 10.0, 20.0]) == 3e10 * 10 + 3e10 * 20 + 3e10 * 10
