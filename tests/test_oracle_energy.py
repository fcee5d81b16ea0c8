"""Pins the oracle's energy and gradient (oracle/ijkl_oracle.c or_energy, or_energy_grad).

E(C) = sum_pq h^MO_pq gamma_pq + 1/2 sum_pqrs Gamma_pqrs (pq|rs)_MO is checked against a
dense einsum over full 4-index tensors, and dE/dC against (a) the dense analytic
gradient 2 h C gamma + 2 sum Gamma_pqrs C_nu q C_lam r C_sig s (mu nu|lam sig) and
(b) central finite differences of the oracle's own energy.  Planting a mistake in the
einsum (a factor dropped, a transposed C) makes these checks fail.
"""
import numpy as np

import oracle
from cipsi_amd import synth


def _full4(packed, n):
    """Full (pq|rs) from a packed pair supermatrix of n orbitals."""
    pp = synth.unpack_sym(packed, synth.n_pairs(n))
    idx = np.zeros((n, n), dtype=np.int64)
    a, b = np.tril_indices(n)
    idx[a, b] = synth.pair0(a, b)
    idx[b, a] = synth.pair0(a, b)
    return pp[idx[:, :, None, None], idx[None, None, :, :]]


def _case(n, m, seed):
    case = synth.make_case(n, m, 2 * n, 3.0, seed)
    rng = np.random.default_rng(seed + 7)
    # a general (non-orthonormal) C: the energy is a polynomial in C either way
    c = np.asfortranarray(case["c_full"][:, :m] + 0.1 * rng.standard_normal((n, m)))
    h = synth.symmetric_packed(n, seed + 3)
    gamma, Gam = synth.pair_density(m, seed + 5)
    return case["ao_packed"], h, c, gamma, Gam


def _dense_energy_grad(V, h, c, gamma, Gam):
    n, m = c.shape
    g_ao = _full4(V, n)
    hfull = synth.unpack_sym(h, n)
    G4 = _full4(Gam, m)
    g_mo = np.einsum("ijkl,ip,jq,kr,ls->pqrs", g_ao, c, c, c, c, optimize=True)
    e = np.sum((c.T @ hfull @ c) * gamma) + 0.5 * np.sum(G4 * g_mo)
    grad = 2 * hfull @ c @ gamma + 2 * np.einsum("pqrs,ijkl,jq,kr,ls->ip", G4, g_ao, c, c, c,
                                                 optimize=True)
    return e, grad


def test_energy_matches_dense_einsum():
    V, h, c, gamma, Gam = _case(7, 5, 11)
    mo = oracle.ijkl(V, c)
    hmo = oracle.tmono(h, c)
    e_dense, _ = _dense_energy_grad(V, h, c, gamma, Gam)
    e = oracle.energy(mo, hmo, gamma, Gam)
    assert abs(e - e_dense) <= 1e-12 * max(1.0, abs(e_dense))


def test_grad_matches_dense_einsum():
    V, h, c, gamma, Gam = _case(7, 5, 12)
    e, g = oracle.energy_grad(V, h, c, gamma, Gam)
    e_dense, g_dense = _dense_energy_grad(V, h, c, gamma, Gam)
    assert abs(e - e_dense) <= 1e-12 * max(1.0, abs(e_dense))
    assert np.max(np.abs(g - g_dense)) <= 1e-12 * np.max(np.abs(g_dense))


def test_two_electron_grad_matches_dense_einsum():
    V, h, c, gamma, Gam = _case(7, 6, 15)
    h, gamma = np.zeros_like(h), np.zeros_like(gamma)
    e, g = oracle.energy_grad(V, h, c, gamma, Gam)
    e_dense, g_dense = _dense_energy_grad(V, h, c, gamma, Gam)
    assert abs(e - e_dense) <= 1e-12 * max(1.0, abs(e_dense))
    assert np.max(np.abs(g - g_dense)) <= 1e-12 * np.max(np.abs(g_dense))


def test_grad_matches_central_differences():
    V, h, c, gamma, Gam = _case(6, 4, 13)
    _, g = oracle.energy_grad(V, h, c, gamma, Gam)
    step = 1e-6
    fd = np.zeros_like(g)
    for mu in range(c.shape[0]):
        for p in range(c.shape[1]):
            cp, cm = c.copy(order="F"), c.copy(order="F")
            cp[mu, p] += step
            cm[mu, p] -= step
            fd[mu, p] = (oracle.energy_grad(V, h, cp, gamma, Gam)[0]
                         - oracle.energy_grad(V, h, cm, gamma, Gam)[0]) / (2 * step)
    assert np.max(np.abs(fd - g)) <= 1e-8 * np.max(np.abs(g))


def test_planted_mistakes_fail():
    V, h, c, gamma, Gam = _case(6, 4, 14)
    _, g = oracle.energy_grad(V, h, c, gamma, Gam)
    n, m = c.shape
    hfull = synth.unpack_sym(h, n)
    G4, g_ao = _full4(Gam, m), _full4(V, n)
    two_e = np.einsum("pqrs,ijkl,jq,kr,ls->ip", G4, g_ao, c, c, c, optimize=True)
    good = 2 * hfull @ c @ gamma + 2 * two_e
    assert np.max(np.abs(g - good)) <= 1e-12 * np.max(np.abs(good))
    for bad in (2 * hfull @ c @ gamma + two_e,                     # factor 2 dropped
                hfull @ c @ gamma + 2 * two_e,                     # one-electron factor
                2 * hfull @ c @ gamma.T[::-1, ::-1] + 2 * two_e):  # permuted density
        assert np.max(np.abs(g - bad)) > 1e-6 * np.max(np.abs(good))
