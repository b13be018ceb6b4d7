"""Seeded synthetic closed-shell G0W0 inputs.

The generator fixes every random draw so that the same seed gives the same
molecule on any machine (no ``numpy.random`` / ``std::normal_distribution``,
whose algorithms are implementation-defined):

* PRNG: SplitMix64.  State ``s_i = seed + i * 0x9E3779B97F4A7C15`` (i = 1, 2, ...),
  output ``mix(s_i)`` with the standard SplitMix64 finaliser.  Because the state
  is a counter, draw ``i`` is computed directly and the generator vectorises.
* uniform: ``(u64 >> 11) * 2**-53`` in [0, 1).
* normal: Box-Muller on consecutive uniform pairs (u1, u2):
  ``sqrt(-2 ln(1 - u1)) * cos(2 pi u2)`` (one normal per pair).

Molecule layout (mirrors the reference's ``gw_input.h5`` contents,
src/hdf5_input.cpp:192-212, and the PySCF preparation's Fermi level,
pyscf_prep/pyscf_g0w0_prep.py:108-110):

* ``mo_energy``: nocc occupied energies U[-1.2, -0.3] Ha then nvir virtual
  energies U[0.05, 3.0] Ha, each group sorted ascending;
* ``fermi_energy`` = (e_HOMO + e_LUMO) / 2;
* ``eri[p,q,r,s] = sum_K B[pq,K] B[rs,K]`` (exactly 8-fold symmetric, see
  ``eri_from_factor``) with B symmetric in (p,q),
  ``B ~ N(0, 0.15**2)`` and K = 3*nmo: 8-fold symmetric and positive
  semidefinite as a (pq, rs) matrix; C-order float64;
* ``vxc``: symmetric, N(0, 0.05**2).

Draw order: occupied uniforms, virtual uniforms, then the B factor over the
upper triangle p <= q (row-major, K fastest), then vxc over p <= q.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GOLDEN_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter-based SplitMix64 stream (state advances by the golden gamma)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0  # number of u64 draws consumed

    def u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            return _mix(self.seed + idx * GOLDEN_GAMMA)

    def uniform(self, n: int) -> np.ndarray:
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        u = self.uniform(2 * n)
        u1, u2 = u[0::2], u[1::2]
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class GwInput:
    """Same fields as the reference's ``GwInput`` (include/gw/gw.hpp:17-23)."""

    mo_energy: np.ndarray      # (nmo,)
    eri: np.ndarray            # (nmo, nmo, nmo, nmo) C-order float64
    vxc: np.ndarray            # (nmo, nmo)
    nocc: int
    fermi_energy: float

    @property
    def nmo(self) -> int:
        return int(self.mo_energy.shape[0])

    @property
    def nvir(self) -> int:
        return self.nmo - self.nocc

    @property
    def nph(self) -> int:
        return self.nocc * self.nvir


def _sym_from_upper(nmo: int, upper_vals: np.ndarray, trailing: tuple = ()) -> np.ndarray:
    out = np.zeros((nmo, nmo) + trailing, dtype=np.float64)
    iu = np.triu_indices(nmo)
    out[iu] = upper_vals.reshape((len(iu[0]),) + trailing)
    out[(iu[1], iu[0])] = out[iu]
    return out


def synthetic_factor(nmo: int, nocc: int, seed: int):
    """Return (mo_energy, fermi, B (nmo*nmo, K), vxc) for the seeded molecule."""
    if not (0 < nocc < nmo):
        raise ValueError("synthetic_factor: need 0 < nocc < nmo")
    rng = SplitMix64(seed)
    nvir = nmo - nocc
    occ = np.sort(-1.2 + 0.9 * rng.uniform(nocc))
    vir = np.sort(0.05 + 2.95 * rng.uniform(nvir))
    mo_energy = np.concatenate([occ, vir])
    fermi = 0.5 * (occ[-1] + vir[0])
    kdim = 3 * nmo
    ntri = nmo * (nmo + 1) // 2
    b = _sym_from_upper(nmo, 0.15 * rng.normal(ntri * kdim), (kdim,)).reshape(nmo * nmo, kdim)
    vxc = _sym_from_upper(nmo, 0.05 * rng.normal(ntri))
    return mo_energy, float(fermi), b, vxc


def eri_from_factor(b: np.ndarray, nmo: int) -> np.ndarray:
    """eri[p,q,r,s] = sum_K B[pq,K] B[rs,K], exactly 8-fold symmetric.

    The product is formed once per unordered pair: G = Bu Bu^T over the
    nmo(nmo+1)/2 rows p <= q, its upper triangle mirrored, then expanded, so
    that eri[p,q,r,s] == eri[q,p,r,s] == eri[r,s,p,q] bit for bit.
    """
    iu = np.triu_indices(nmo)
    bu = b.reshape(nmo, nmo, -1)[iu]
    g = bu @ bu.T
    g = np.triu(g) + np.triu(g, 1).T
    tri = np.zeros((nmo, nmo), dtype=np.int64)
    tri[iu] = np.arange(len(iu[0]))
    tri[(iu[1], iu[0])] = tri[iu]
    t = tri.ravel()
    return np.ascontiguousarray(g[t][:, t].reshape(nmo, nmo, nmo, nmo))


def synthetic_input(nmo: int, nocc: int, seed: int) -> GwInput:
    """The full molecule, ERI formed on the host (see eri_from_factor)."""
    mo_energy, fermi, b, vxc = synthetic_factor(nmo, nocc, seed)
    return GwInput(mo_energy=mo_energy, eri=eri_from_factor(b, nmo), vxc=vxc,
                   nocc=nocc, fermi_energy=fermi)


def with_gap(inp: GwInput, gap: float) -> GwInput:
    """The same molecule with its HOMO-LUMO gap set to `gap` Ha: every
    virtual level moves down by the same amount, the Fermi level to mid-gap.
    A small gap makes the dielectric ill-conditioned at small frequencies
    (S^2 = 4 D / (w^2 + D^2) reaches 4 / gap)."""
    mo = np.array(inp.mo_energy, dtype=np.float64)
    mo[inp.nocc:] -= (mo[inp.nocc] - mo[inp.nocc - 1]) - gap
    return GwInput(mo_energy=mo, eri=inp.eri, vxc=inp.vxc, nocc=inp.nocc,
                   fermi_energy=0.5 * (mo[inp.nocc - 1] + mo[inp.nocc]))


# Named workloads (nmo, nocc, seed).  Seeds are fixed here.
CONFIGS = {
    "nmo24": (24, 5, 20240601),
    "nmo120": (120, 20, 20240602),
    "nmo200": (200, 40, 20240603),
}
BATCH_SEED = 20240604
SERVE_SEED = 20240605


def batch_molecules(count: int = 32, lo: int = 24, hi: int = 120, seed: int = BATCH_SEED):
    """(nmo, nocc, seed) for the uneven-nmo batch: nmo log-uniform in [lo, hi]."""
    rng = SplitMix64(seed)
    u = rng.uniform(count)
    out = []
    for m, uu in enumerate(u):
        nmo = int(round(math.exp(math.log(lo) + uu * (math.log(hi) - math.log(lo)))))
        nocc = max(1, int(round(nmo / 5)))
        out.append((nmo, nocc, seed + 1 + m))
    return out


def serve_jobs(count: int, rate: float, lo: int = 24, hi: int = 200, seed: int = SERVE_SEED):
    """(arrival_s, nmo, nocc, nfreq, seed) for a Poisson job stream at `rate`
    jobs/s: exponential inter-arrival times, nmo log-uniform in [lo, hi],
    nocc = round(nmo / 5), nfreq uniform in 100..200, per-job seeds."""
    rng = SplitMix64(seed)
    u = rng.uniform(3 * count)
    out = []
    t_arr = 0.0
    for i in range(count):
        t_arr += -math.log(1.0 - u[3 * i]) / rate
        nmo = int(round(math.exp(math.log(lo) + u[3 * i + 1] * (math.log(hi) - math.log(lo)))))
        nfreq = 100 + int(u[3 * i + 2] * 101)
        out.append((t_arr, nmo, max(1, int(round(nmo / 5))), nfreq, seed + 1 + i))
    return out
