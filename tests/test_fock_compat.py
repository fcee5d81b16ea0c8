"""The unmodified reference `fock` program (CIPSI_sources/fock, built with flang into
oracle/_ref/fock_ref) consumes our f50/f25 files exactly as it consumes ijkl's.

fock reads f25 and f50 (fock.f:25-71), builds the closed-shell reference energy and
the J/K/Fock matrices (fock.f:80-146) and writes them to its own output file
(fock.f:147-152).  Run on (a) the reference ijkl's files and (b) files written by
cipsi_amd.formats from the oracle's packed stream, it must give the same energy and
matrices.  (The flang runtime stops fock afterwards in stkener.f:5-8, which writes
after reading to end of file; the energy and matrices are complete by then.)

Only where oracle/_ref/fock_ref and ijkl_ref were built (the reference sources are needed).
"""
import os
import re
import resource
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from cipsi_amd import formats, synth
from tests.conftest import ROOT

FOCK = os.path.join(ROOT, "oracle", "_ref", "fock_ref")
IJKL = os.path.join(ROOT, "oracle", "_ref", "ijkl_ref")
pytestmark = pytest.mark.skipif(not (os.path.exists(FOCK) and os.path.exists(IJKL)),
                                reason="reference programs not built here")


def _unlimited_stack():
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_fock(wd):
    res = subprocess.run([FOCK], input=b"&fokfil /\n", cwd=wd, capture_output=True,
                         preexec_fn=_unlimited_stack, timeout=120)
    out = res.stdout.decode(errors="replace")
    e_el = float(re.search(r"energie electronique\s+(-?[0-9.]+)", out).group(1))
    e_tot = float(re.search(r"energie totale\s+(-?[0-9.]+)", out).group(1))
    recs = formats.read_records(os.path.join(wd, "ijkl"))
    return e_el, e_tot, recs[:6]


def test_fock_on_our_files_matches_reference_files():
    n, noc = 16, 4
    case = synth.make_case(n, n, 32, 4.0, 77)
    V, c = case["ao_packed"], case["c_full"]
    h = synth.symmetric_packed(n, 78)
    xx, ix = formats.labels_from_packed(V, n)
    with tempfile.TemporaryDirectory() as ref_wd, tempfile.TemporaryDirectory() as our_wd:
        formats.write_info(os.path.join(ref_wd, "info"), n, c, h, enuc=1.25, n_elec=2 * noc)
        nrec = formats.write_pqrs(os.path.join(ref_wd, "pqrs"), xx, ix)
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec, nvecf=n)
        r = subprocess.run([IJKL], input=stdin.encode(), cwd=ref_wd, capture_output=True)
        assert r.returncode == 0
        # our files: f25 as the Fortran host writes it (byte-equal to the reference, checked
        # on the GPU), f50 from the oracle's stream through cipsi_amd.formats.write_f50
        shutil.copy(os.path.join(ref_wd, "f25"), os.path.join(our_wd, "f25"))
        mo = oracle.ijkl(V, c)
        hmo = oracle.tmono(h, c)
        formats.write_f50(os.path.join(our_wd, "f50"), mo, hmo)
        ref_f50 = formats.read_records(os.path.join(ref_wd, "f50"))
        our_f50 = formats.read_records(os.path.join(our_wd, "f50"))
        assert [len(x) for x in ref_f50] == [len(x) for x in our_f50]
        e_ref = run_fock(ref_wd)
        e_our = run_fock(our_wd)
    assert e_ref[0] == e_our[0] and e_ref[1] == e_our[1]
    assert abs(e_ref[1] - e_ref[0] - 1.25) < 1e-7            # enuc added (fock.f:101)
    for a, b in zip(e_ref[2], e_our[2]):                      # header, its, h, J, K, F
        assert len(a) == len(b)
        if len(a) % 8 == 0 and len(a) > 8:
            np.testing.assert_allclose(np.frombuffer(a, "<f8"), np.frombuffer(b, "<f8"),
                                       rtol=0, atol=1e-12)
        else:
            assert a == b


def test_fock_energy_matches_closed_shell_formula():
    """fock's E = sum_i 2 h_ii + sum_i (ii|ii) + sum_{i>k} 2[2(ii|kk) - (ik|ik)] + enuc
    (fock.f:80-101), evaluated from our packed stream, equals what fock prints."""
    n, noc = 12, 3
    case = synth.make_case(n, n, 24, 3.0, 91)
    V, c = case["ao_packed"], case["c_full"]
    h = synth.symmetric_packed(n, 92)
    xx, ix = formats.labels_from_packed(V, n)
    with tempfile.TemporaryDirectory() as wd:
        formats.write_info(os.path.join(wd, "info"), n, c, h, enuc=0.5, n_elec=2 * noc)
        nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), xx, ix)
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec, nvecf=n)
        assert subprocess.run([IJKL], input=stdin.encode(), cwd=wd, capture_output=True).returncode == 0
        e_el, e_tot, _ = run_fock(wd)
    mo = oracle.ijkl(V, c)
    hmo = oracle.tmono(h, c)

    def g(i, j, k, l):
        a, b = synth.pair0(max(i, j), min(i, j)), synth.pair0(max(k, l), min(k, l))
        return mo[synth.pair0(max(a, b), min(a, b))]

    e = 0.0
    for i in range(noc):
        e += 2 * hmo[synth.pair0(i, i)] + g(i, i, i, i)
        for k in range(i):
            e += 2 * (2 * g(i, i, k, k) - g(i, k, i, k))
    assert abs(e - e_el) < 1e-7
    assert abs(e + 0.5 - e_tot) < 1e-7
