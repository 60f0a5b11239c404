"""Field comparison metrics.

Restates py_utils/compare_fields_from_twoplotfiles.py:39-46 without yt:
err = sqrt(mean((f1 - f2)^2)), rel = err / mean(|f1|) on the level-0 grid,
and also reports ||f1 - f2||_2 / ||f1||_2.
"""
import numpy as np


def field_errors(f1, f2):
    f1 = np.asarray(f1, dtype=np.float64)
    f2 = np.asarray(f2, dtype=np.float64)
    d = f1 - f2
    err = float(np.sqrt(np.mean(d * d)))
    avg = float(np.mean(np.abs(f1)))
    n1 = float(np.linalg.norm(f1.ravel()))
    l2 = float(np.linalg.norm(d.ravel())) / n1 if n1 > 0 else float(np.linalg.norm(d.ravel()))
    rel = err / avg if avg > 0 else err
    return {"abs": err, "rel_rms_over_meanabs": rel, "rel_l2": l2}


def compare_states(ref, test, names):
    """Per-field error dict for two (NVAR, nz, ny, nx) states."""
    return {n: field_errors(ref[q], test[q]) for q, n in enumerate(names)}
