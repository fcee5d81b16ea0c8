"""The Fortran host (cipsi_amd/fortran/ijkl_gpu) on molecules with symmetry, without a
GPU: istop = 1 stops after the MO preparation and the ijom file (ijkl.f:328-329), so
otodon / rotat2 / vector of the host (through cipsi_ijkl_sym_prepare, and
cipsi_ijkl_sym_linear + cipsi_ijkl_sym_prepare_so for C-inf-v / D-inf-h) are checked
against the reference's ijom bytes; an MO that mixes irreps stops the host with
rotat2's own error marker and tag, as the flang-built reference does on the same files."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cipsi_amd import formats
from tests.conftest import ROOT, load_golden
from tests.test_oracle_sym import LIN, SYM, regenerate

HOST = os.path.join(ROOT, "cipsi_amd", "fortran", "ijkl_gpu")
REF = os.path.join(ROOT, "oracle", "_ref", "ijkl_ref")
pytestmark = pytest.mark.skipif(not os.path.exists(HOST), reason="ijkl_gpu not built")


def write_inputs(wd, case, c_square, h_packed, noc):
    formats.write_info_molecule(os.path.join(wd, "info"), case["mol"], c_square, h_packed,
                                case["s_packed"], enuc=1.5, n_elec=2 * noc)
    return formats.write_pqrs(os.path.join(wd, "pqrs"), case["xx"], case["ix"])


def run(exe, wd, stdin):
    return subprocess.run([exe], input=stdin.encode(), cwd=wd, capture_output=True, timeout=120)


@pytest.mark.parametrize("name", SYM + LIN)
def test_host_prepare_only_matches_reference_ijom(name):
    g = load_golden(name)
    case = regenerate(g)
    with tempfile.TemporaryDirectory() as wd:
        nrec = write_inputs(wd, case, g["c_square"], g["h_packed"], int(g["noc"]))
        res = run(HOST, wd, formats.ijfil_namelist() + formats.ijklin_namelist(nrec, istop=1))
        log = res.stdout.decode(errors="replace")
        assert res.returncode == 0, log[-2000:] + res.stderr.decode(errors="replace")
        ijom = formats.read_records(os.path.join(wd, "ijom"))
    assert np.array_equal(np.frombuffer(ijom[0], "<f8"), g["ijom_c"])
    assert np.array_equal(np.frombuffer(ijom[1], "<f8"), g["ijom_c"])
    assert np.array_equal(np.frombuffer(ijom[2], "<f8"), g["ijom_s"])


def tag_of(log):
    lines = log.splitlines()
    k = max(i for i, ln in enumerate(lines) if "ERROR in ijkl" in ln)
    for ln in lines[k + 1:]:
        t = ln.strip()
        if t.lstrip("-").isdigit():
            return int(t)
    return None


@pytest.mark.skipif(not os.path.exists(REF), reason="reference build absent")
def test_host_stops_with_rotat2_tag_like_the_reference():
    g = load_golden("sym_c2v_n19")
    case = regenerate(g)
    irr = case["mo_irrep"]
    a = int(np.flatnonzero(irr == 0)[0])
    b = int(np.flatnonzero(irr != 0)[0])
    c = np.asfortranarray(g["c_square"]).copy()
    c[:, a] = (c[:, a] + c[:, b]) / np.sqrt(2.0)
    tags = []
    for exe in (HOST, REF):
        with tempfile.TemporaryDirectory() as wd:
            nrec = write_inputs(wd, case, c, g["h_packed"], int(g["noc"]))
            res = run(exe, wd, formats.ijfil_namelist() + formats.ijklin_namelist(nrec))
            log = res.stdout.decode(errors="replace")
            assert "ERROR in ijkl" in log, log[-1500:]
            tags.append(tag_of(log))
    assert tags[0] == tags[1] == 1
