"""Pins the oracle's Slater-Condon elements (or_hdig: cip/hdig.f:40-80, or_hntd:
cip/hntd.f:36-168) against second quantization done by brute force.

Each determinant is D = prod_t a+_{p_t} a_{h_t} |ref> on the closed-shell reference, with
its holes h_t and particles p_t taken from cip's arrays.  The excitation pairs act on
distinct spin-orbitals, so they commute; the sign convention of |ref> cancels between bra
and ket.  H = sum h_pq a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q is applied to D_b
operator by operator on occupation bit strings, with the fermion signs counted
explicitly.  This is an independent route to:
  hdig(D)    = <D|H|D> - <ref|H|ref>
  hntd(a, b) = <D_a|H|D_b>   (a != b; zero beyond two differing spin-orbitals).
Planted mistakes (a dropped exchange term, a flipped crossing sign) must fail.
"""
import itertools

import numpy as np
import pytest

import oracle
from cipsi_amd import synth


def _full4(mo, m):
    pp = synth.unpack_sym(mo, synth.n_pairs(m))
    idx = np.zeros((m, m), dtype=np.int64)
    a, b = np.tril_indices(m)
    idx[a, b] = synth.pair0(a, b)
    idx[b, a] = synth.pair0(a, b)
    return pp[idx[:, :, None, None], idx[None, None, :, :]]


def _apply(op, s, state):
    """a_s (op = 0) or a+_s (op = 1) on (sign, bits); None if it vanishes."""
    sign, bits = state
    occ = (bits >> s) & 1
    if occ == op:
        return None
    below = bin(bits & ((1 << s) - 1)).count("1")
    return (-sign if below & 1 else sign), bits ^ (1 << s)


class Brute:
    """Spin-orbital s (0-based) = orbital s % m, spin s // m (cip numbers it s + 1)."""

    def __init__(self, m, noc, h, g):
        self.m, self.noc, self.h, self.g = m, noc, h, g
        st = (1, 0)
        for s in [o for o in range(noc)] + [m + o for o in range(noc)]:
            st = _apply(1, s, st)
        self.ref = st

    def det(self, holes, parts):
        st = self.ref
        for h0, p0 in zip(holes, parts):
            st = _apply(0, h0, st)
            st = _apply(1, p0, st)
        return st

    def h_on(self, st):
        m, out = self.m, {}

        def add(state, coef):
            if state is not None and coef != 0.0:
                out[state[1]] = out.get(state[1], 0.0) + coef * state[0]

        for sp in range(2):
            for p, q in itertools.product(range(m), repeat=2):
                s1 = _apply(0, q + sp * m, st)
                if s1 is not None:
                    add(_apply(1, p + sp * m, s1), self.h[p, q])
        for sp, tp in itertools.product(range(2), repeat=2):
            for p, q, r, s in itertools.product(range(m), repeat=4):
                v = self.g[p, q, r, s]
                x = _apply(0, q + sp * m, st)
                if x is None:
                    continue
                x = _apply(0, s + tp * m, x)
                if x is None:
                    continue
                x = _apply(1, r + tp * m, x)
                if x is None:
                    continue
                add(_apply(1, p + sp * m, x), 0.5 * v)
        return out

    def elem(self, bra, hket):
        return bra[0] * hket.get(bra[1], 0.0)


def _case(m, noc, seed, ndet):
    case = synth.make_case(m, m, 2 * m, 3.0, seed)
    mo = oracle.ijkl(case["ao_packed"], case["c_full"])
    hmo = oracle.tmono(synth.symmetric_packed(m, seed + 1), case["c_full"])
    e, f, j, k = oracle.fock(mo, hmo, m, noc)
    rng = np.random.default_rng(seed + 2)
    occ = [o for o in range(noc)] + [m + o for o in range(noc)]          # 0-based spin-orbitals
    vir = [o for o in range(noc, m)] + [m + o for o in range(noc, m)]
    ne, nd, trou, part, lists = [], [], [], [], []
    def particles_for(hs, taken):
        # each particle has its hole's spin: cip's determinants keep S_z (the two-
        # difference branch of hntd.f checks only spin pairs that S_z conservation allows)
        ps = []
        for h0 in hs:
            cand = [x for x in vir if x // m == h0 // m and x not in ps and x not in taken]
            ps.append(cand[int(rng.integers(len(cand)))])
        return ps

    for d in range(ndet):
        n_ex = 0 if d == 0 else int(rng.integers(1, 4))
        hs = list(rng.choice(occ, n_ex, replace=False))
        ps = particles_for(hs, [])
        ne.append(n_ex)
        nd.append(len(trou))
        trou += [x + 1 for x in hs]
        part += [x + 1 for x in ps]
        lists.append((hs, ps))
    # single and double excitations of each determinant, so that most pairs interact
    for d in range(1, ndet):
        hs, ps = lists[d]
        if not hs:
            continue
        hs2, ps2 = list(hs), list(ps)
        free = [x for x in vir if x not in ps2 and x // m == ps2[-1] // m]
        if free:
            ps2[-1] = free[int(rng.integers(len(free)))]
        if len(hs2) > 1 and rng.random() < 0.5:
            free_o = [x for x in occ if x not in hs2 and x // m == hs2[0] // m]
            if free_o:
                hs2[0] = free_o[int(rng.integers(len(free_o)))]
        ne.append(len(hs2))
        nd.append(len(trou))
        trou += [x + 1 for x in hs2]
        part += [x + 1 for x in ps2]
        lists.append((hs2, ps2))
    iorb = [o + 1 for o in range(m)] * 2
    ispin = [1] * m + [2] * m
    dets = oracle.Determinants(ne, nd, trou, part, iorb, ispin)
    hfull = synth.unpack_sym(hmo, m)
    brute = Brute(m, noc, hfull, _full4(mo, m))
    return dict(mo=mo, f=f, j=j, k=k, dets=dets, lists=lists, brute=brute)


@pytest.mark.parametrize("m,noc,seed", [(5, 2, 500), (6, 3, 510)])
def test_hdig_and_hntd_match_second_quantization(m, noc, seed):
    c = _case(m, noc, seed, 12)
    br, lists = c["brute"], c["lists"]
    states = [br.det(h, p) for h, p in lists]
    hstates = [br.h_on(st) for st in states]
    e_ref = br.elem(br.ref, br.h_on(br.ref))
    hd = oracle.hdig(c["dets"], c["f"], c["j"], c["k"])
    for d, st in enumerate(states):
        assert abs(hd[d] - (br.elem(st, hstates[d]) - e_ref)) <= 1e-11, d
    n = len(states)
    pairs = np.array([(a, b) for a in range(n) for b in range(n) if a != b])
    ho = oracle.hntd(c["dets"], pairs, m, c["f"], c["mo"])
    nonzero = 0
    for (a, b), v in zip(pairs, ho):
        if states[a][1] == states[b][1]:
            # the same determinant listed twice (its excitations in another order):
            # hntd returns 0, the diagonal is hdig's (hntd.f:173-179)
            assert v == 0.0, (a, b)
            continue
        want = br.elem(states[a], hstates[b])
        assert abs(v - want) <= 1e-11, (a, b, lists[a], lists[b], v, want)
        nonzero += abs(want) > 1e-9
    assert nonzero >= 20                      # singles and doubles are exercised


def test_planted_mistakes_fail():
    """A wrong sign or a dropped exchange in the brute-force route is detected."""
    m, noc = 5, 2
    c = _case(m, noc, 520, 10)
    br, lists = c["brute"], c["lists"]
    states = [br.det(h, p) for h, p in lists]
    n = len(states)
    pairs = np.array([(a, b) for a in range(n) for b in range(n) if a != b])
    pairs = np.array([(a, b) for a, b in pairs if states[a][1] != states[b][1]])
    ho = oracle.hntd(c["dets"], pairs, m, c["f"], c["mo"])
    want = np.array([br.elem(states[a], br.h_on(states[b])) for a, b in pairs])
    assert np.max(np.abs(ho - want)) <= 1e-11
    assert np.max(np.abs(-ho - want)) > 1e-3                    # flipped crossing sign
    g = _full4(c["mo"], m)
    br2 = Brute(m, noc, br.h, g.transpose(0, 3, 2, 1))           # exchange in place of Coulomb
    want2 = np.array([br2.elem(states[a], br2.h_on(states[b])) for a, b in pairs])
    assert np.max(np.abs(ho - want2)) > 1e-3
