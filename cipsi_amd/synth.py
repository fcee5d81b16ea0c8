"""Seeded synthetic inputs for the ijkl AO->MO transform.

Everything the workloads need is generated here from a seed:

* AO two-electron integrals with an exact low-rank structure,
  (ij|kl) = sum_P B^P_ij B^P_kl, B^P symmetric, B^P_ij ~ N(0,1) exp(-|i-j|/d).
  B is rounded to a 2^-17 grid with |B| < 8, so every product is a multiple of
  2^-34 below 2^6 and any sum of up to 2^12 of them is exact in float64: the
  packed integrals are bit-identical whatever the summation order (numpy BLAS,
  torch on the GPU, or a loop), on every machine.
* MO coefficients C: the Q factor of a seeded Gaussian matrix (signs fixed so
  that diag(R) > 0), optionally with leading frozen columns.
* A symmetric one-electron matrix h, a PSD trace-1 1-RDM gamma and a PSD trace-1
  2-RDM Gamma on the MO pair space.

Layouts follow the reference (0-based here):
  pair0(a, b) = a(a+1)/2 + b, a >= b            (CIPSI_sources/ijkl/itijkl.f:38)
  packed AO integrals V[pair0(ij, kl)], ij >= kl (CIPSI_sources/ijkl/transd.f:369-371)
"""
from __future__ import annotations

import numpy as np

B_GRID_BITS = 17          # B values are integers * 2^-17
B_LIMIT = 8.0             # |B| < 8  -> |B * 2^17| < 2^20


def n_pairs(n: int) -> int:
    return n * (n + 1) // 2


def pair0(a, b):
    return a * (a + 1) // 2 + b


def pair_index_arrays(n: int):
    """(a, b) for every packed pair index, a >= b, in packed order."""
    a, b = np.tril_indices(n)
    return a.astype(np.int64), b.astype(np.int64)


def eri_factors(n_ao: int, rank: int, decay: float, seed: int, grid: bool = True) -> np.ndarray:
    """B[P, pair0(i,j)] for P < rank: z * exp(-|i-j|/decay) rounded to the 2^-17 grid.

    grid=False keeps the full 53-bit mantissas (the same z and envelope, unrounded): the
    integrals are then ordinary floating-point sums, not exact, and they carry low-order
    bits below any fixed grid, as integrals from a real integral code do."""
    if grid and rank > 4096:
        raise ValueError("rank > 4096 would break the exactness of the integral sums")
    rng = np.random.default_rng(seed)
    a, b = pair_index_arrays(n_ao)
    env = np.exp(-np.abs(a - b) / float(decay))
    z = rng.standard_normal((rank, a.size))
    z = np.clip(z, -7.99, 7.99)
    if not grid:
        return z * env
    B = np.rint(z * env * 2.0**B_GRID_BITS) * 2.0**-B_GRID_BITS
    return B


def eri_packed(B: np.ndarray, row_block: int = 2048) -> np.ndarray:
    """Packed lower triangle of B^T B (P_A(P_A+1)/2 doubles), exact."""
    P = B.shape[1]
    out = np.empty(n_pairs(P), dtype=np.float64)
    Bt = np.ascontiguousarray(B.T)
    for r0 in range(0, P, row_block):
        r1 = min(P, r0 + row_block)
        blk = Bt[r0:r1] @ B[:, :r1]                     # rows r0..r1-1, cols 0..r1-1
        rows = np.arange(r0, r1)[:, None]
        cols = np.arange(r1)[None, :]
        out[pair0(r0, 0):pair0(r1 - 1, r1 - 1) + 1] = blk[cols <= rows]
    return out


def eri_packed_torch(B: np.ndarray, device, row_block: int = 2048):
    """Same as eri_packed, computed with torch on `device` (exact, so identical bits, for
    grid factors; for full-mantissa factors an ordinary rounded product)."""
    import torch
    Bd = torch.as_tensor(B, dtype=torch.float64, device=device)
    P = B.shape[1]
    out = torch.empty(n_pairs(P), dtype=torch.float64, device=device)
    for r0 in range(0, P, row_block):
        r1 = min(P, r0 + row_block)
        blk = Bd[:, r0:r1].T @ Bd[:, :r1]
        rows = torch.arange(r0, r1, device=device)[:, None]
        cols = torch.arange(r1, device=device)[None, :]
        out[pair0(r0, 0):pair0(r1 - 1, r1 - 1) + 1] = blk[cols <= rows]
    del Bd
    return out


def orthonormal(n: int, seed: int) -> np.ndarray:
    """Square orthonormal matrix (Fortran order): Q of a seeded Gaussian, diag(R) > 0."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diag(r))[None, :]
    return np.asfortranarray(q)


def symmetric_packed(n: int, seed: int, scale: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) * scale
    a = 0.5 * (a + a.T)
    return a[np.tril_indices(n)].copy()


def psd_trace1(n: int, seed: int, rank: int | None = None) -> np.ndarray:
    """Symmetric PSD n x n matrix with unit trace (full, row-major)."""
    rng = np.random.default_rng(seed)
    r = n if rank is None else min(rank, n)
    g = rng.standard_normal((n, r))
    a = g @ g.T
    return a / np.trace(a)


def pair_density(n_mo: int, seed: int, rank: int = 32, row_block: int = 2048):
    """(gamma, Gamma_packed): seeded one- and two-particle densities, both PSD and
    trace-normalised.  gamma is M x M (row-major, trace 1).  Gamma lives on the MO pair
    supermatrix (P_M x P_M, Gamma_pqrs = Gamma[pair(p,q), pair(r,s)]) as F F^T for a
    Gaussian P_M x rank factor F, scaled so that the 4-index trace
    sum_pq Gamma_pqpq = sum_a m(a) Gamma_aa is 1 (m(a) = 2 for p != q); it is returned
    packed like the MO stream (P_M(P_M+1)/2 doubles)."""
    pm = n_pairs(n_mo)
    gamma = psd_trace1(n_mo, seed)
    rng = np.random.default_rng(seed + 1)
    F = rng.standard_normal((pm, min(rank, pm)))
    a, b = pair_index_arrays(n_mo)
    mult = np.where(a == b, 1.0, 2.0)
    scale = 1.0 / float(np.sum(mult * np.einsum("ij,ij->i", F, F)))
    out = np.empty(n_pairs(pm))
    for r0 in range(0, pm, row_block):
        r1 = min(pm, r0 + row_block)
        blk = (F[r0:r1] @ F[:r1].T) * scale
        keep = np.arange(r1)[None, :] <= np.arange(r0, r1)[:, None]
        out[pair0(r0, 0):pair0(r1 - 1, r1 - 1) + 1] = blk[keep]
    return gamma, out


def packed_lower(a: np.ndarray) -> np.ndarray:
    return a[np.tril_indices(a.shape[0])].copy()


def unpack_sym(packed: np.ndarray, n: int) -> np.ndarray:
    a = np.zeros((n, n))
    i, j = np.tril_indices(n)
    a[i, j] = packed
    a[j, i] = packed
    return a


# Workloads: sizes, ranks, decay lengths and seeds of the five standard cases.
CONFIGS = {
    0: dict(name="cpu_n40", n_ao=40, n_mo=40, rank=120, decay=8.0, frozen=0, seed=2024),
    1: dict(name="fp64_n120_m110", n_ao=120, n_mo=110, rank=360, decay=15.0, frozen=10, seed=2025),
    2: dict(name="precision_n240", n_ao=240, n_mo=240, rank=720, decay=20.0, frozen=0, seed=2026),
    4: dict(name="adjoint_n160_m100", n_ao=160, n_mo=100, rank=480, decay=8.0, frozen=0, seed=2028),
}


def make_case(n_ao: int, n_mo: int, rank: int, decay: float, seed: int, frozen: int = 0,
              with_eri: bool = True):
    """Inputs of one transform: B factors, packed AO integrals, full and active C."""
    if n_mo + frozen > n_ao:
        raise ValueError("n_mo + frozen must not exceed n_ao")
    B = eri_factors(n_ao, rank, decay, seed)
    c_full = orthonormal(n_ao, seed + 1)
    c = np.asfortranarray(c_full[:, frozen:frozen + n_mo])
    case = dict(n_ao=n_ao, n_mo=n_mo, rank=rank, decay=decay, seed=seed, frozen=frozen,
                B=B, c_full=c_full, c=c)
    if with_eri:
        case["ao_packed"] = eri_packed(B)
    return case


def batch_scan(n_geom: int = 64, n_min: int = 60, n_max: int = 180, n_frozen: int = 8,
               decay: float = 10.0, seed: int = 2027):
    """Sizes and seeds of a potential-curve scan: N_AO ~ U{n_min..n_max}, N_MO = N_AO - 8,
    rank 3 N_AO, one seed per geometry."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_geom)
    return [dict(n_ao=int(n), n_mo=int(n) - n_frozen, rank=3 * int(n), decay=decay,
                 seed=seed * 1000 + g, frozen=0) for g, n in enumerate(sizes)]


def mo_reference_closed_form(B: np.ndarray, c: np.ndarray) -> np.ndarray:
    """(pq|rs)_MO = sum_P Bt^P_pq Bt^P_rs with Bt^P = C^T B^P C: an independent route
    to the packed MO integrals (different summation order from the loop nests)."""
    n = c.shape[0]
    m = c.shape[1]
    a, b = pair_index_arrays(n)
    pm_a, pm_b = pair_index_arrays(m)
    Bt = np.empty((B.shape[0], pm_a.size))
    for P in range(B.shape[0]):
        bp = np.zeros((n, n))
        bp[a, b] = B[P]
        bp[b, a] = B[P]
        t = c.T @ bp @ c
        Bt[P] = t[pm_a, pm_b]
    return eri_packed_float(Bt)


def eri_packed_float(F: np.ndarray) -> np.ndarray:
    """Packed lower triangle of F^T F for arbitrary float factors (not exact)."""
    P = F.shape[1]
    full = F.T @ F
    return full[np.tril_indices(P)].copy()
