"""Comparison helpers that report per-column errors."""

import numpy as np

FIELD_NAMES = ["action", "energy", "dir_x", "dir_y", "dir_z", "edep", "nsec",
               "s0_pid", "s0_e", "s0_dx", "s0_dy", "s0_dz",
               "s1_pid", "s1_e", "s1_dx", "s1_dy", "s1_dz"]


def column_errors(gpu, cpu):
    """Max relative difference per named column (relative to max(|cpu|, 1e-300))."""
    rel = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-300)
    rel[(gpu == cpu)] = 0
    return {FIELD_NAMES[i]: float(rel[:, i].max()) for i in range(gpu.shape[1])}


def assert_records_close(gpu, cpu, rtol, atol=0):
    ok = np.isclose(gpu, cpu, rtol=rtol, atol=atol)
    if not ok.all():
        bad = {k: v for k, v in column_errors(gpu, cpu).items() if v > rtol}
        rows = np.where(~ok.all(axis=1))[0]
        raise AssertionError(f"{len(rows)} rows differ; worst columns {bad}; "
                             f"first row {rows[0]}: gpu={gpu[rows[0]].tolist()} "
                             f"cpu={cpu[rows[0]].tolist()}")
