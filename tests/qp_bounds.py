"""QP-energy tolerances pinned by the reference's own outputs.

The Pade + QP fixed point amplifies round-off in Sigma_c(i w) by many orders
of magnitude on some states.  The reference itself shows it: its two CPU
backends (ReferenceBackend, Gauss-Jordan, and BlasLapackBackend, MKL
zgetrf/zgetrs/zgemm; the ref_<case>.npz and ref_<case>_blas.npz fixtures,
both written by the reference's own code) agree on Sigma_c to ~1e-16
relative, yet their QP energies differ by up to 0.3 Ha on the synthetic
molecules.  A state is "reference-resolved" at a tolerance when the two
reference backends agree on it within a tenth of that tolerance: a third
correct implementation then lands within the tolerance even when its
round-off moves the state ten times further than the backends' does.

FP64 parity rule (north star: QP energies within 1e-8 Ha of the BLAS/LAPACK
backend): every reference-resolved state within 1e-8 Ha of both reference
backends; every other state within 10 x the largest backend-to-backend
spread of the case.  A non-finite QP energy of a computed state fails.

Mixed-form rule (QP within 1 meV of FP64): every state within 1 meV of the
FP64 form, except the states pinned in qp_mev_exceptions.json for that form
and case, so that any growth of the list fails.
"""
from __future__ import annotations

import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HA_MEV = 27.211386245988 * 1e3
QP_TOL_HA = 1e-8


def ref(name: str, backend: str = "reference") -> dict:
    suffix = "" if backend == "reference" else "_blas"
    return dict(np.load(os.path.join(GOLD, f"ref_{name}{suffix}.npz")))


def backend_spread(name: str) -> np.ndarray:
    """|QP(ReferenceBackend) - QP(BlasLapackBackend)| per state (Ha)."""
    return np.abs(ref(name)["qp"] - ref(name, "blas")["qp"])


def fp64_qp_violations(qp: np.ndarray, name: str, tol: float = QP_TOL_HA, factor: float = 10.0):
    """States breaking the FP64 parity rule, and a one-line summary."""
    a, b = ref(name)["qp"], ref(name, "blas")["qp"]
    spread = np.abs(a - b)
    resolved = spread <= 0.1 * tol
    da, db = np.abs(qp - a), np.abs(qp - b)
    d = np.maximum(da, db)
    bound = np.where(resolved, tol, max(factor * spread.max(), tol))
    bad = [int(j) for j in range(len(qp)) if not (d[j] <= bound[j])]
    summary = (f"QP vs reference: {int(np.sum(da <= tol))}/{len(qp)} states within {tol:g} Ha "
               f"(reference backends agree on {int(resolved.sum())}); max {np.nanmax(da):.2e} Ha, "
               f"reference backend spread max {spread.max():.2e} Ha")
    return bad, summary


def mev_exceptions(form: str, case: str):
    """Pinned states allowed beyond 1 meV for `form` on `case` (None: not pinned)."""
    with open(os.path.join(GOLD, "qp_mev_exceptions.json")) as fh:
        table = json.load(fh)["allowed"]
    return table.get(f"{form}:{case}")


def mixed_qp_violations(qp_mix: np.ndarray, qp_fp64: np.ndarray, form: str, case: str):
    """States beyond 1 meV of FP64 that are not pinned, and a summary."""
    mev = np.abs(qp_mix - qp_fp64) * HA_MEV
    allowed = mev_exceptions(form, case)
    if allowed is None:
        raise AssertionError(f"no pinned 1 meV exceptions for {form}:{case}; states beyond 1 meV: "
                             f"{[int(j) for j in np.nonzero(~(mev <= 1.0))[0]]}")
    bad = [int(j) for j in range(len(mev)) if not (mev[j] <= 1.0) and j not in allowed]
    summary = (f"QP vs FP64: {int(np.sum(mev <= 1.0))}/{len(mev)} within 1 meV, "
               f"{len(allowed)} pinned exceptions; max {np.nanmax(mev):.3g} meV")
    return bad, summary
