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


# an E component whose own norm is below this fraction of the whole field
# vector's norm has no scale of its own: the FP64 rounding of the potential
# (~1e-15 of the vector) reaches it amplified by the inverse of that
# fraction, ~1e-11 at the threshold
E_NEGLIGIBLE = 1e-4


def state_errors(ref, test, ns, dim):
    """Per-field relative L2 errors of two states (NVAR, nz, ny, nx):
    ||ref_q - test_q||_2 / ||ref_q||_2 for every component q.

    An E-field component whose own norm is negligible (below E_NEGLIGIBLE
    of the norm of the whole field vector, e.g. E_y of a symmetric
    discharge at t ~ 0) is normalised by the vector's norm instead: its
    own relative error would only measure rounding noise of the potential.
    Fields that are identically zero in the reference must be identically
    zero in `test`.  Returns {component index: rel_l2}."""
    ref = np.asarray(ref, dtype=np.float64)
    test = np.asarray(test, dtype=np.float64)
    efx = ns + 3
    evec = float(np.sqrt(sum(np.sum(ref[efx + d] ** 2) for d in range(dim))))
    out = {}
    for q in range(ref.shape[0]):
        d = float(np.linalg.norm((ref[q] - test[q]).ravel()))
        nrm = float(np.linalg.norm(ref[q].ravel()))
        if efx <= q < efx + dim and nrm < E_NEGLIGIBLE * evec:
            nrm = evec
        if nrm == 0.0:
            out[q] = 0.0 if d == 0.0 else float("inf")
        else:
            out[q] = d / nrm
    return out


def worst_rms_over_meanabs(ref, test, ns, dim):
    """max over fields of the reference script's metric, rms(f1 - f2) /
    mean|f1|, with (as in state_errors) a negligible E component measured
    against the mean magnitude of the whole field vector; fields equal in
    both states are skipped.  Returns (value, component index)."""
    ref = np.asarray(ref, dtype=np.float64)
    test = np.asarray(test, dtype=np.float64)
    efx = ns + 3
    emag = np.sqrt(sum(ref[efx + d] ** 2 for d in range(dim)))
    evec = float(np.linalg.norm(emag.ravel()))
    worst, which = 0.0, None
    for q in range(ref.shape[0]):
        e = field_errors(ref[q], test[q])
        if e["abs"] == 0.0:
            continue
        v = e["rel_rms_over_meanabs"]
        if efx <= q < efx + dim and float(np.linalg.norm(ref[q].ravel())) < E_NEGLIGIBLE * evec:
            v = e["abs"] / float(np.mean(emag))
        if v > worst:
            worst, which = v, q
    return worst, which
