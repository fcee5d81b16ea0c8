"""CPU-side checks of the C ABI: the library loads and exports every symbol that
include/cipsi_ijkl.h declares (no compute call: there is no GPU here)."""
import os
import re

import pytest

from cipsi_amd import ijkl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cipsi_ijkl.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cipsi_ijkl_\w+)\s*\(", txt)))


def test_header_lists_match_python_exports():
    assert sorted(ijkl.EXPORTS) == header_symbols()


@pytest.mark.skipif(not os.path.exists(ijkl.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ijkl.load_library()
    for name in header_symbols():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(ijkl.LIB_PATH), reason="library not built")
def test_strerror_and_workspace_without_gpu():
    lib = ijkl.load_library()
    assert lib.cipsi_ijkl_strerror(0) == b"ok"
    assert b"label" in lib.cipsi_ijkl_strerror(-2)
    # 240 / 240 FP64: symmetric pair matrix + H + padded C
    pa = 240 * 241 // 2
    # C padded to 256 rows, in both orders (first and second product)
    assert ijkl.workspace_bytes(240, 240) == 8 * (pa * pa + pa * pa) + 2 * 8 * 256 * 240
    assert ijkl.workspace_bytes(10, 11) == -1          # n_mo > n_ao rejected
    assert ijkl.workspace_bytes(300, 10) == -1         # beyond the supported N_AO


@pytest.mark.skipif(not os.path.exists(ijkl.LIB_PATH), reason="library not built")
def test_create_without_gpu_reports_nodevice():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ijkl.IjklError) as e:
        ijkl.Context(0)
    assert e.value.status == -8


@pytest.mark.skipif(not os.path.exists(ijkl.LIB_PATH), reason="library not built")
def test_every_context_entry_point_rejects_a_null_context():
    """Each function of the header whose first parameter is the context returns
    CIPSI_ERR_ARG for a NULL context (pointers NULL, numbers 0) before touching a device,
    so a caller's mistake is a status, never a crash."""
    import ctypes
    txt = open(os.path.join(ROOT, "include", "cipsi_ijkl.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(cipsi_ijkl_\w+)\s*\(\s*cipsi_ijkl_ctx\s*\*", txt)))
    assert len(names) > 20
    lib = ijkl.load_library()
    for name in names:
        fn = getattr(lib, name)
        args = []
        for t in fn.argtypes:
            if t in (ctypes.c_double, ctypes.c_float):
                args.append(0.0)
            elif t in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_long,
                       ctypes.c_longlong, ctypes.c_uint, ctypes.c_size_t):
                args.append(0)
            else:
                args.append(None)
        assert fn(*args) == -1, name
