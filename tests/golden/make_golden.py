"""Writes the committed golden fixtures under tests/golden/.

Runs only in a container that holds the reference tree (/root/reference) and
its HDF5 tools; the GPU runs and the test-suite read the .npz files it writes.

1. ``h2o_input.npz``: the H2O input of the reference regression tests
   (regression_tests/*/gw_input.h5, identical in every test directory),
   extracted with h5dump: mo_energy, eri_mo, vxc_mo, nocc, fermi_energy.
2. ``h2o_serial_fixtures.npz``: the data files the reference's own regression
   tests compare against (regression_tests/h2o_serial): E_c_before_Pade.out
   (Sigma_c(i w) of state 5 on the 200-point grid), E_c.out (real-axis Pade
   continuation, 19836 points) and the gw.out energy-summary row of orbital 5,
   plus E_c_before_Pade.out of the other committed backends (spread estimate).
3. ``ref_*.npz``: outputs of the reference's own path (oracle/_ref/gw_ref,
   built by oracle/Makefile from the reference sources, ReferenceBackend) for
   all states of H2O, of the seeded synthetic molecule nmo=24/nocc=5, of
   small synthetic cases at eta = 0 and eta != 0, and of an ill-conditioned
   one (the nmo=30 molecule with its HOMO-LUMO gap set to 1e-4 Ha, at eta = 0
   and eta = 1e-3).
4. ``ref_*_blas.npz``: the same cases through the reference's BLAS/LAPACK
   backend (oracle/_ref/gw_ref_blas: BlasLapackBackend, zgetrf/zgetrs/zgemm
   from MKL), selected by the reference's own backend factory.  The spread
   between the two reference backends is what a correct FP64 implementation
   may differ by.

Usage: python tests/golden/make_golden.py [--only NAME ...]
"""
from __future__ import annotations

import argparse
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from minigw_amd.synthetic import GwInput, synthetic_input, with_gap  # noqa: E402

REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"
GW_REF = os.path.join(REPO, "oracle", "_ref", "gw_ref")
GW_REF_BLAS = os.path.join(REPO, "oracle", "_ref", "gw_ref_blas")


def _h5_attr(path: str, name: str) -> str:
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-a", name, path],
                         check=True, capture_output=True, text=True).stdout
    m = re.search(r"\(0\):\s*([-+0-9.eE]+)", txt)
    if not m:
        raise RuntimeError(f"attribute {name} not found")
    return m.group(1)


def _h5_dataset(path: str, name: str, shape) -> np.ndarray:
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "d.bin")
        subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path],
                       check=True, capture_output=True)
        return np.fromfile(out, dtype="<f8").reshape(shape)


def read_h2o_input() -> GwInput:
    path = os.path.join(REF, "regression_tests", "h2o_serial", "gw_input.h5")
    mo = _h5_dataset(path, "/mo_energy", (-1,))
    nmo = mo.shape[0]
    eri = _h5_dataset(path, "/eri_mo", (nmo, nmo, nmo, nmo))
    vxc = _h5_dataset(path, "/vxc_mo", (nmo, nmo))
    nocc = int(_h5_attr(path, "/nocc"))
    fermi = float(_h5_attr(path, "/fermi_energy"))
    return GwInput(mo_energy=mo, eri=eri, vxc=vxc, nocc=nocc, fermi_energy=fermi)


def _parse_complex_lines(path: str) -> np.ndarray:
    vals = []
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if line:
                re_s, im_s = line.strip("()").split(",")
                vals.append(complex(float(re_s), float(im_s)))
    return np.array(vals, dtype=np.complex128)


def _parse_summary_row(path: str, orbital: int) -> np.ndarray:
    """KS, Sigma_x, Sigma_c, Vxc, QP (eV) of one orbital in a gw.out summary."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    start = next(i for i, l in enumerate(lines) if "G0W0 Calculation Summary" in l)
    for l in lines[start:]:
        parts = l.split()
        if parts and parts[0] == str(orbital):
            return np.array([float(x) for x in parts[1:6]])
    raise RuntimeError("summary row not found")


def write_h2o_fixtures() -> None:
    inp = read_h2o_input()
    np.savez_compressed(os.path.join(HERE, "h2o_input.npz"), mo_energy=inp.mo_energy,
                        eri=inp.eri, vxc=inp.vxc, nocc=inp.nocc,
                        fermi_energy=inp.fermi_energy)
    d = os.path.join(REF, "regression_tests")
    fx = {
        "e_c_before_pade": _parse_complex_lines(os.path.join(d, "h2o_serial", "E_c_before_Pade.out")),
        "e_c": np.loadtxt(os.path.join(d, "h2o_serial", "E_c.out")),
        "summary_row_orbital5_ev": _parse_summary_row(os.path.join(d, "h2o_serial", "gw.out"), 5),
    }
    for b in ("h2o_blas_lapack", "h2o_mpi", "h2o_cublas_serial", "h2o_cublas_mpi", "h2o_cosma"):
        fx["e_c_before_pade_" + b] = _parse_complex_lines(os.path.join(d, b, "E_c_before_Pade.out"))
    np.savez_compressed(os.path.join(HERE, "h2o_serial_fixtures.npz"), **fx)


def write_reference_input(path: str, inp: GwInput, nfreq: int, npar: int, state: int,
                          eta: float, threads: int) -> None:
    """The raw input file oracle/ref_driver.cpp reads (full ERI, kdim = 0)."""
    with open(path, "wb") as fh:
        np.array([inp.nmo, inp.nocc, nfreq, npar, state, threads, 0], dtype="<i8").tofile(fh)
        np.array([inp.fermi_energy, eta], dtype="<f8").tofile(fh)
        np.ascontiguousarray(inp.mo_energy, dtype="<f8").tofile(fh)
        np.ascontiguousarray(inp.vxc, dtype="<f8").tofile(fh)
        np.ascontiguousarray(inp.eri, dtype="<f8").tofile(fh)


def run_reference(inp: GwInput, nfreq: int, npar: int, state: int, eta: float,
                  threads: int = 8, backend: str = "reference") -> dict:
    """Run oracle/_ref/gw_ref (backend "reference") or oracle/_ref/gw_ref_blas
    (backend "blas-lapack", frequencies OpenMP-parallel, one LAPACK thread
    each) on one input; returns its outputs."""
    exe = GW_REF if backend == "reference" else GW_REF_BLAS
    if not os.path.exists(exe):
        raise RuntimeError("build oracle/_ref first: make -C oracle")
    with tempfile.TemporaryDirectory() as td:
        ip, op = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_reference_input(ip, inp, nfreq, npar, state, eta, threads)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), MKL_NUM_THREADS="1",
                   MKL_THREADING_LAYER="GNU")
        cmd = [exe, ip, op] if backend == "reference" else [exe, ip, op, backend]
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        raw = np.fromfile(op, dtype="<f8")
    nmo = int(raw[:2].view("<i8")[0])
    f = int(raw[:2].view("<i8")[1])
    o = 2
    omega = raw[o:o + f]; o += f
    weight = raw[o:o + f]; o += f
    sx = raw[o:o + nmo]; o += nmo
    sc = raw[o:o + 2 * nmo * f].view(np.complex128).reshape(nmo, f); o += 2 * nmo * f
    qp = raw[o:o + nmo]
    return dict(omega=omega, weight=weight, sigma_x_diag=sx, sigma_c=sc, qp=qp,
                nfreq=f, npar=npar, eta=eta, state=state, backend=backend)


def _save_case(name: str, inp: GwInput, res: dict, with_input: bool) -> None:
    payload = dict(res)
    payload.update(nmo=inp.nmo, nocc=inp.nocc, fermi_energy=inp.fermi_energy)
    if with_input:
        payload.update(mo_energy=inp.mo_energy, eri=inp.eri, vxc=inp.vxc)
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **payload)


# name -> (input factory, nfreq, npar, state, eta, store the input in the fixture)
CASES = {
    "h2o_all": (lambda: read_h2o_input(), 200, 16, -1, 0.0, False),
    "h2o_eta": (lambda: read_h2o_input(), 40, 16, -1, 0.01, False),
    "syn_nmo24": (lambda: synthetic_input(24, 5, 20240601), 200, 16, -1, 0.0, False),
    "syn_small": (lambda: synthetic_input(10, 3, 7), 48, 16, -1, 0.0, True),
    "syn_small_eta": (lambda: synthetic_input(10, 3, 7), 48, 12, -1, 0.05, True),
    "syn_odd": (lambda: synthetic_input(13, 4, 11), 32, 8, -1, 0.0, True),
    # ill-conditioned: the nmo=30 molecule with its HOMO-LUMO gap set to 1e-4 Ha
    "syn_gap": (lambda: with_gap(synthetic_input(30, 6, 47), 1e-4), 64, 16, -1, 0.0, False),
    "syn_gap_eta": (lambda: with_gap(synthetic_input(30, 6, 47), 1e-4), 48, 16, -1, 1e-3, False),
}


# cases also run through the BLAS/LAPACK backend (ref_<name>_blas.npz)
BLAS_CASES = ("h2o_all", "syn_nmo24", "syn_small", "syn_odd", "h2o_eta", "syn_small_eta", "syn_gap",
              "syn_gap_eta")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    names = args.only or ["h2o"] + list(CASES) + [c + "_blas" for c in BLAS_CASES]
    if "h2o" in names:
        write_h2o_fixtures()
        print("wrote h2o fixtures", flush=True)
    for name in names:
        base = name[:-5] if name.endswith("_blas") else name
        if base not in CASES:
            continue
        fac, nfreq, npar, state, eta, with_input = CASES[base]
        inp = fac()
        t = time.time()
        backend = "blas-lapack" if name.endswith("_blas") else "reference"
        res = run_reference(inp, nfreq, npar, state, eta, backend=backend)
        _save_case(name, inp, res, with_input and backend == "reference")
        print(f"wrote ref_{name}.npz in {time.time() - t:.1f} s", flush=True)


if __name__ == "__main__":
    main()
