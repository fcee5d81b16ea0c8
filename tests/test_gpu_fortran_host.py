"""The Fortran host program (cipsi_amd/fortran/ijkl_gpu) as a drop-in for the
reference ijkl executable: same stdin namelists and input files, outputs compared
with the reference's own f50 / f25 / ijom (golden fixtures)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cipsi_amd import formats, synth
from tests.conftest import ROOT, load_golden

pytestmark = pytest.mark.gpu
HOST = os.path.join(ROOT, "cipsi_amd", "fortran", "ijkl_gpu")


@pytest.mark.parametrize("name", ["tiny_n12_frozen2", "sparse_n10_shuffled", "cfg0_n40"])
def test_fortran_host_matches_reference_files(name):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(HOST):
        pytest.fail("ijkl_gpu is not built")
    print(f"[fortran host] {name}", flush=True)
    g = load_golden(name)
    n, frozen = int(g["n_ao"]), int(g["frozen"])
    case = synth.make_case(n, int(g["n_mo"]), int(g["rank"]), float(g["decay"]), int(g["seed"]),
                           frozen=frozen)
    xx, ix = formats.labels_from_packed(case["ao_packed"], n, cutoff=float(g["cutoff"]))
    if bool(g["shuffle"]):
        perm = np.random.default_rng(int(g["seed"]) + 3).permutation(xx.size)
        xx, ix = xx[perm], ix[perm]
    with tempfile.TemporaryDirectory() as wd:
        formats.write_info(os.path.join(wd, "info"), n, g["c_full"], g["h_packed"], enuc=1.25,
                           n_elec=2)
        nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), xx, ix)
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec, nvecf=n, frozen=frozen)
        res = subprocess.run([HOST], input=stdin.encode(), cwd=wd, capture_output=True, timeout=120)
        log = res.stdout.decode(errors="replace")
        assert res.returncode == 0 and "fin ijkl" in log, log[-2000:] + res.stderr.decode()
        f50 = formats.read_f50(os.path.join(wd, "f50"), n)
        with open(os.path.join(wd, "f25"), "rb") as f:
            f25 = f.read()
        ijom = formats.read_records(os.path.join(wd, "ijom"))
    assert f50["n_records"] == int(g["n_records"])
    assert f25 == g["f25_bytes"].tobytes()
    np.testing.assert_array_equal(np.frombuffer(ijom[0], "<f8"), g["ijom_c"])
    np.testing.assert_array_equal(np.frombuffer(ijom[2], "<f8"), g["ijom_s"])
    assert np.max(np.abs(f50["hmo"] - g["hmo"])) <= 1e-12
    if "mo" in g:
        assert np.max(np.abs(f50["mo"] - g["mo"])) <= 1e-12
        assert np.max(np.abs(f50["tail"] - g["tail"])) <= 1e-12
    else:
        assert np.max(np.abs(f50["mo"][g["sample_idx"]] - g["sample_val"])) <= 1e-12
        nr = int(g["n_mo_records"])
        assert np.max(np.abs(f50["mo"][(nr - 1) * formats.LSIZE:] - g["last_record"])) <= 1e-12
        assert np.max(np.abs(f50["tail"] - g["tail"])) <= 1e-12


def _run_host(g, wd, **keys):
    n, frozen = int(g["n_ao"]), int(g["frozen"])
    case = synth.make_case(n, int(g["n_mo"]), int(g["rank"]), float(g["decay"]), int(g["seed"]),
                           frozen=frozen)
    xx, ix = formats.labels_from_packed(case["ao_packed"], n, cutoff=float(g["cutoff"]))
    formats.write_info(os.path.join(wd, "info"), n, g["c_full"], g["h_packed"], enuc=1.25,
                       n_elec=2)
    nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), xx, ix)
    stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec, nvecf=n, frozen=frozen,
                                                               **keys)
    return subprocess.run([HOST], input=stdin.encode(), cwd=wd, capture_output=True, timeout=120)


def test_fortran_host_split_precision():
    """precision=1, nslice=6 through the namelist: the f50 stream stays within 1e-10 of
    the reference's (the split path's tolerance); an unbuilt slice count is reported with
    the Autocip error marker and tag 9 (Autocip_13/calcul.f90:169-176)."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(HOST):
        pytest.fail("ijkl_gpu is not built")
    print("[fortran host] split precision", flush=True)
    g = load_golden("cfg0_n40")
    n = int(g["n_ao"])
    with tempfile.TemporaryDirectory() as wd:
        res = _run_host(g, wd, precision=1, nslice=6)
        log = res.stdout.decode(errors="replace")
        assert res.returncode == 0 and "fin ijkl" in log, log[-2000:]
        f50 = formats.read_f50(os.path.join(wd, "f50"), n)
    assert np.max(np.abs(f50["mo"][g["sample_idx"]] - g["sample_val"])) <= 1e-10
    nr = int(g["n_mo_records"])
    assert np.max(np.abs(f50["mo"][(nr - 1) * formats.LSIZE:] - g["last_record"])) <= 1e-10
    with tempfile.TemporaryDirectory() as wd:
        res = _run_host(g, wd, precision=1, nslice=9)
        log = res.stdout.decode(errors="replace")
        assert " !! ERROR in ijkl" in log, log[-2000:]
        after = log[log.index(" !! ERROR in ijkl"):].splitlines()
        assert "9" in [x.strip() for x in after[1:]], log[-2000:]


@pytest.mark.parametrize("name", ["sym_c2v_n19", "sym_c2v_deg_n19", "sym_d2h_n20", "sym_d2h_n24",
                                  "sym_c2v_n30", "sym_d2h_n38", "lin_cinfv_n14", "lin_dinfh_n18",
                                  "lin_cinfv_sp_n9", "lin_dinfh_sp_n10", "lin_dinfh_s_n4",
                                  "lin_cinfv_n28", "lin_dinfh_n36"])
def test_fortran_host_symmetry_matches_reference_files(name):
    """Molecules with point-group symmetry (C2v, D2h, and the linear C-inf-v, D-inf-h with
    their symmetry transformations): f25 (header, product table, isydeg records, block
    directory) and ijom byte for byte, the f50 stream of every block and the h_MO record
    within 1e-12 of the flang-built reference's files."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(HOST):
        pytest.fail("ijkl_gpu is not built")
    from tests.test_oracle_sym import regenerate
    print(f"[fortran host sym] {name}", flush=True)
    g = load_golden(name)
    case = regenerate(g)
    noc = int(g["noc"])
    with tempfile.TemporaryDirectory() as wd:
        formats.write_info_molecule(os.path.join(wd, "info"), case["mol"], g["c_square"],
                                    g["h_packed"], case["s_packed"], enuc=1.5, n_elec=2 * noc)
        nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), case["xx"], case["ix"])
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec)
        res = subprocess.run([HOST], input=stdin.encode(), cwd=wd, capture_output=True, timeout=120)
        log = res.stdout.decode(errors="replace")
        assert res.returncode == 0 and "fin ijkl" in log, log[-2000:] + res.stderr.decode()
        with open(os.path.join(wd, "f25"), "rb") as f:
            f25 = f.read()
        ijom = formats.read_records(os.path.join(wd, "ijom"))
        recs = formats.read_records(os.path.join(wd, "f50"))
    assert f25 == g["f25_bytes"].tobytes()
    np.testing.assert_array_equal(np.frombuffer(ijom[0], "<f8"), g["ijom_c"])
    np.testing.assert_array_equal(np.frombuffer(ijom[2], "<f8"), g["ijom_s"])
    assert len(recs) - 1 == int(g["n_f50_records"])
    stream = np.concatenate([np.frombuffer(r, "<f8") for r in recs[:-1]])
    assert stream.size == g["f50_stream"].size            # full records, stale tail included
    d = float(np.max(np.abs(stream - g["f50_stream"])))
    print(f"[fortran host sym] {name}: max|f50 - reference| {d:.2e}", flush=True)
    assert d <= 1e-12
    assert np.max(np.abs(np.frombuffer(recs[-1], "<f8") - g["hmo"])) <= 1e-12
