"""The C-ABI library builds, loads without a GPU, exports every symbol that
include/reno_hip.h declares, and refuses host tensors (no CPU fallback)."""
import os
import re

import numpy as np
import pytest
import torch

from renormalizer_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "reno_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(reno_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    lib = native.load_library()
    declared = _declared()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in reno_hip.h but not exported"
    assert set(declared) == set(native.EXPORTS)


def test_library_is_current():
    """The shipped .so must be newer than every source it is built from."""
    so = native.LIB_PATH
    csrc = os.path.join(ROOT, "renormalizer_amd", "csrc")
    srcs = [os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hip", ".h"))]
    srcs.append(os.path.join(ROOT, "include", "reno_hip.h"))
    assert os.path.getmtime(so) >= max(os.path.getmtime(s) for s in srcs), "rebuild libreno_hip.so"


def test_status_strings_and_queries():
    lib = native.load_library()
    assert lib.reno_version() >= 10000
    assert lib.reno_status_string(native.RENO_EQN).decode() == "invalid quantum number"
    d = native.HopDesc(1, native.PREC_FP64, 256, 10, 256, 3, 3, 1, 1, 1)
    ws = lib.reno_hop_workspace_bytes(d)
    # permuted MPO + two stage intermediates of 256*3*10*256 complex128
    assert ws >= 2 * 256 * 3 * 10 * 256 * 16
    assert lib.reno_krylov_workspace_bytes(d, 20) > ws + 20 * 256 * 10 * 256 * 16
    bad = native.HopDesc(2, native.PREC_FP64, 1, 1, 1, 1, 1, 1, 1, 1)
    assert lib.reno_hop_workspace_bytes(bad) == 0


def test_no_host_fallback():
    L = torch.zeros((2, 1, 2), dtype=torch.complex128)
    W = torch.zeros((1, 2, 2, 1), dtype=torch.float64)
    c = torch.zeros((2, 2, 2), dtype=torch.complex128)
    with pytest.raises(native.RenoNativeError):
        native.hop_apply(L, W, L, c)
    with pytest.raises(native.RenoNativeError):
        native.expm_krylov(L, W, L, c, -0.01j)


def test_invalid_arguments_rejected_without_launch():
    lib = native.load_library()
    d = native.HopDesc(1, 7, 4, 2, 4, 1, 1, 1, 1, 1)
    st = lib.reno_hop_apply(d, 1, 2, 3, 1 << 20, None)
    assert st == native.RENO_EINVAL
    assert b"precision" in lib.reno_last_error()
    st = lib.reno_block_qr(0, 4, 4, 1, 0, None, 1, 1, 0, 1, 1 << 20, None)
    assert st == native.RENO_EQN
