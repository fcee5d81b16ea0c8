"""Generate the golden fixtures of the ijkl transform from the reference program itself.

Runs only on a machine that has the reference sources under /root/reference (it
builds oracle/_ref/ijkl_ref from them with `make -C oracle`).  For each case it
writes the reference's own input files (info, pqrs, namelists on stdin), runs the
flang-built reference ijkl, parses f50 / f25 / ijom and stores the inputs that are
not regenerable bit-for-bit (C) plus the outputs as .npz fixtures next to this script.

The AO integrals are regenerated from the seed by cipsi_amd.synth (exact by
construction, so identical on every machine); a SHA-256 of them is stored and
checked when the fixture is used.

Symmetric cases (point groups C2v and D2h, and the linear groups C-inf-v and D-inf-h,
synth_sym) store the reference's whole f50
stream (all symmetry blocks, in its record order), f25 and ijom bytes; their inputs C
and h are stored too, the AO integrals and the skeleton label stream are regenerated
from the seed (exact by construction, SHA-256 checked).

Usage:  python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cipsi_amd import formats, synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ijkl_ref")

# name -> (n_ao, n_mo_active, frozen, rank, decay, seed, cutoff, shuffle, n_samples)
CASES = {
    "tiny_n12_frozen2": (12, 10, 2, 30, 4.0, 11, 0.0, False, 0),
    "sparse_n10_shuffled": (10, 10, 0, 24, 3.0, 12, 0.05, True, 0),
    "cfg0_n40": (40, 40, 0, 120, 8.0, 2024, 0.0, False, 8192),
    "cfg1_n120_frozen10": (120, 110, 10, 360, 15.0, 2025, 0.0, False, 8192),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_reference(workdir, n_ao, c_square, h_packed, xx, ix, frozen, enuc=1.25):
    info = os.path.join(workdir, "info")
    formats.write_info(info, n_ao, c_square, h_packed, enuc=enuc, n_elec=2)
    nrec = formats.write_pqrs(os.path.join(workdir, "pqrs"), xx, ix)
    stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec, nvecf=n_ao, frozen=frozen)
    t0 = time.time()
    res = subprocess.run([REF_BIN], input=stdin.encode(), cwd=workdir, capture_output=True)
    dt = time.time() - t0
    log = res.stdout.decode(errors="replace") + res.stderr.decode(errors="replace")
    if res.returncode != 0 or "fin ijkl" not in log:
        raise RuntimeError("reference ijkl failed:\n" + log[-3000:])
    return nrec, dt, log


def make_case(name):
    n_ao, n_mo, frozen, rank, decay, seed, cutoff, shuffle, n_samples = CASES[name]
    case = synth.make_case(n_ao, n_mo, rank, decay, seed, frozen=frozen)
    V = case["ao_packed"]
    c_full = case["c_full"]
    h_packed = synth.symmetric_packed(n_ao, seed + 7)
    xx, ix = formats.labels_from_packed(V, n_ao, cutoff=cutoff)
    if shuffle:
        perm = np.random.default_rng(seed + 3).permutation(xx.size)
        xx, ix = xx[perm], ix[perm]
    with tempfile.TemporaryDirectory(prefix="ijkl_ref_") as wd:
        nrec, dt, log = run_reference(wd, n_ao, c_full, h_packed, xx, ix, frozen)
        m_ref = n_ao                   # nvec = nvecf = N: frozen columns compacted, stale tail kept
        f50 = formats.read_f50(os.path.join(wd, "f50"), m_ref)
        f25 = formats.read_f25(os.path.join(wd, "f25"))
        ijom = formats.read_records(os.path.join(wd, "ijom"))
        with open(os.path.join(wd, "f25"), "rb") as f:
            f25_bytes = f.read()
    pm = n_mo * (n_mo + 1) // 2
    n_active = pm * (pm + 1) // 2
    mo = f50["mo"]
    active = mo[:n_active]                 # leading prefix = transform of the active columns
    out = dict(
        n_ao=n_ao, n_mo=n_mo, frozen=frozen, rank=rank, decay=decay, seed=seed, cutoff=cutoff,
        shuffle=shuffle, nrec=nrec, ref_seconds=dt,
        c_full=c_full, h_packed=h_packed, ao_sha=sha(V), labels_sha=sha(ix), xx_sha=sha(xx),
        hmo=f50["hmo"], n_records=f50["n_records"], n_mo_records=f50["n_mo_records"],
        tail=f50["tail"], f25_bytes=np.frombuffer(f25_bytes, dtype=np.uint8),
        ijom_c=np.frombuffer(ijom[0], "<f8"), ijom_s=np.frombuffer(ijom[2], "<f8"),
        full_sum=np.sum(mo), full_sumsq=np.sum(mo * mo), active_sum=np.sum(active),
        active_sumsq=np.sum(active * active), active_absmax=np.max(np.abs(active)),
    )
    if n_samples == 0:
        out["mo"] = mo
    else:
        rng = np.random.default_rng(seed + 99)
        idx = np.unique(np.concatenate([
            rng.integers(0, n_active, size=n_samples),
            np.arange(min(512, n_active)), np.arange(n_active - min(512, n_active), n_active)]))
        out["sample_idx"] = idx.astype(np.int64)
        out["sample_val"] = active[idx]
        # all records' boundaries around the stale tail, full last record
        out["last_record"] = mo[(f50["n_mo_records"] - 1) * formats.LSIZE:]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: NREC={nrec} ref {dt:.1f}s f50 records={f50['n_records']} -> {path} "
          f"({os.path.getsize(path) / 1024:.0f} KiB)")


# Point-group symmetry (synth_sym): name -> (group, shells of atom 1, shells of atom 2
# (None: homonuclear), rank per irrep, seed, occupied MOs, degenerate pair)
SYM_CASES = {
    "sym_c2v_n19": ("C2v", [1, 1, 2, 3], [1, 2, 3], 2, 5, 3, False),
    "sym_c2v_deg_n19": ("C2v", [1, 1, 2, 3], [1, 2, 3], 2, 13, 3, True),
    "sym_d2h_n20": ("D2h", [1, 1, 2, 3], None, 2, 5, 3, False),
    "sym_d2h_n24": ("D2h", [1, 2, 2, 3], None, 2, 11, 5, False),
    "sym_c2v_n30": ("C2v", [1, 1, 2, 2, 3, 3], [1, 2, 2, 3], 3, 17, 6, False),
    "sym_d2h_n38": ("D2h", [1, 1, 1, 2, 2, 3, 3], None, 2, 19, 7, False),
}


def make_sym_case(name):
    from cipsi_amd import synth_sym
    grp, sa, sb, rank, seed, noc, deg = SYM_CASES[name]
    case = synth_sym.make_sym_case(grp, sa, sb, rank_per_irrep=rank, seed=seed,
                                   degenerate_pair=deg)
    mol = case["mol"]
    with tempfile.TemporaryDirectory(prefix="ijkl_ref_sym_") as wd:
        formats.write_info_molecule(os.path.join(wd, "info"), mol, case["c_square"],
                                    case["h_packed"], case["s_packed"], enuc=1.5,
                                    n_elec=2 * noc)
        nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), case["xx"], case["ix"])
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec)
        t0 = time.time()
        res = subprocess.run([REF_BIN], input=stdin.encode(), cwd=wd, capture_output=True)
        dt = time.time() - t0
        log = res.stdout.decode(errors="replace") + res.stderr.decode(errors="replace")
        if res.returncode != 0 or "fin ijkl" not in log:
            raise RuntimeError("reference ijkl failed:\n" + log[-3000:])
        recs = formats.read_records(os.path.join(wd, "f50"))
        with open(os.path.join(wd, "f25"), "rb") as f:
            f25_bytes = f.read()
        ijom = formats.read_records(os.path.join(wd, "ijom"))
    stream = np.concatenate([np.frombuffer(r, "<f8") for r in recs[:-1]])
    out = dict(group=grp, shells_a=np.array(sa), shells_b=np.array(sb if sb else []),
               rank=rank, seed=seed, noc=noc, degenerate=deg, nrec=nrec, ref_seconds=dt,
               c_square=case["c_square"], h_packed=case["h_packed"],
               ao_sha=sha(case["ao_packed"]), xx_sha=sha(case["xx"]), labels_sha=sha(case["ix"]),
               f50_stream=stream, n_f50_records=len(recs) - 1, hmo=np.frombuffer(recs[-1], "<f8"),
               f25_bytes=np.frombuffer(f25_bytes, dtype=np.uint8),
               ijom_c=np.frombuffer(ijom[0], "<f8"), ijom_s=np.frombuffer(ijom[2], "<f8"))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: N={mol.n_ao} NREC={nrec} ref {dt:.2f}s f50 records={len(recs) - 1} -> {path} "
          f"({os.path.getsize(path) / 1024:.0f} KiB)")


# Linear molecules (synth_sym.make_linear_case): name -> (group, shells of atom 1, shells
# of atom 2 (None: homonuclear), rank, seed, occupied MOs).  The info file names CINFV /
# DINFH and carries the operations of C1 / Cs, as pshf writes them.
LIN_CASES = {
    "lin_cinfv_n14": ("Cinfv", [1, 2, 3, 1], [1, 2], 2, 5, 3),
    "lin_dinfh_n18": ("Dinfh", [1, 2, 3], None, 2, 5, 3),
    "lin_cinfv_sp_n9": ("Cinfv", [1, 1, 2], [1, 2], 2, 7, 2),
    "lin_dinfh_sp_n10": ("Dinfh", [1, 1, 2], None, 2, 9, 3),
    "lin_dinfh_s_n4": ("Dinfh", [1, 1], None, 2, 3, 1),
    "lin_cinfv_n28": ("Cinfv", [1, 1, 2, 2, 3, 3], [1, 1, 2, 3], 2, 17, 6),
    "lin_dinfh_n36": ("Dinfh", [1, 1, 2, 2, 3, 3], None, 2, 19, 7),
}


def make_lin_case(name):
    from cipsi_amd import synth_sym
    grp, sa, sb, rank, seed, noc = LIN_CASES[name]
    case = synth_sym.make_linear_case(grp, sa, sb, rank=rank, seed=seed)
    mol = case["mol"]
    with tempfile.TemporaryDirectory(prefix="ijkl_ref_lin_") as wd:
        formats.write_info_molecule(os.path.join(wd, "info"), mol, case["c_square"],
                                    case["h_packed"], case["s_packed"], enuc=1.5,
                                    n_elec=2 * noc)
        nrec = formats.write_pqrs(os.path.join(wd, "pqrs"), case["xx"], case["ix"])
        stdin = formats.ijfil_namelist() + formats.ijklin_namelist(nrec)
        t0 = time.time()
        res = subprocess.run([REF_BIN], input=stdin.encode(), cwd=wd, capture_output=True)
        dt = time.time() - t0
        log = res.stdout.decode(errors="replace") + res.stderr.decode(errors="replace")
        if res.returncode != 0 or "fin ijkl" not in log:
            raise RuntimeError("reference ijkl failed:\n" + log[-3000:])
        recs = formats.read_records(os.path.join(wd, "f50"))
        with open(os.path.join(wd, "f25"), "rb") as f:
            f25_bytes = f.read()
        ijom = formats.read_records(os.path.join(wd, "ijom"))
    stream = np.concatenate([np.frombuffer(r, "<f8") for r in recs[:-1]])
    out = dict(group=grp, shells_a=np.array(sa), shells_b=np.array(sb if sb else []),
               rank=rank, seed=seed, noc=noc, nrec=nrec, ref_seconds=dt,
               c_square=case["c_square"], h_packed=case["h_packed"],
               ao_sha=sha(case["ao_packed"]), xx_sha=sha(case["xx"]), labels_sha=sha(case["ix"]),
               f50_stream=stream, n_f50_records=len(recs) - 1, hmo=np.frombuffer(recs[-1], "<f8"),
               f25_bytes=np.frombuffer(f25_bytes, dtype=np.uint8),
               ijom_c=np.frombuffer(ijom[0], "<f8"), ijom_s=np.frombuffer(ijom[2], "<f8"))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: N={mol.n_ao} NREC={nrec} ref {dt:.2f}s f50 records={len(recs) - 1} -> {path} "
          f"({os.path.getsize(path) / 1024:.0f} KiB)")


def main(argv):
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"], check=True)
    names = argv or list(CASES) + list(SYM_CASES) + list(LIN_CASES)
    for n in names:
        if n in SYM_CASES:
            make_sym_case(n)
        elif n in LIN_CASES:
            make_lin_case(n)
        else:
            make_case(n)


if __name__ == "__main__":
    main(sys.argv[1:])
