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
    assert ijkl.workspace_bytes(240, 240) == 8 * (pa * pa + pa * pa) + 8 * 256 * 240   # C padded to 256 rows
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
