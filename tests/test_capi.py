"""The C-ABI library loads on a CPU-only host and exports every function that
include/dvc_accel.h declares (no compute calls: there is no GPU here); the
ladder data it ships derive the reference's codes."""
import ctypes
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "dvc_accel.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dvc_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    from dvc_amd import _lib
    names = declared_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(_lib.lib, n), f"{n} declared in include/dvc_accel.h but not exported"
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature in dvc_amd/_lib.py"


def test_prepare_data_host_only():
    from dvc_amd import _lib
    _lib.prepare_data()  # loads both packs, checks the cached inverses, no GPU needed
    for n in (1584, 6336):
        assert (ROOT / "dvc_amd" / "data" / f"reg{n}.invh").stat().st_size == n * 4 * ((n + 31) // 32)


def test_ctx_create_fails_cleanly_without_gpu():
    from dvc_amd import _lib
    import os
    if os.path.exists("/dev/kfd"):
        return  # a GPU host: covered by the gpu tests
    try:
        _lib.Context()
    except _lib.DvcError as e:
        assert e.rc in (_lib.DVC_EDEVICE,)
    else:
        raise AssertionError("context creation succeeded without a GPU")


def test_bad_arguments_rejected_before_device_use():
    from dvc_amd import _lib
    rc = _lib.lib.dvc_ldpca_decode(None, 1584, None, None, 0, -1, None, None)
    assert rc == _lib.DVC_EINVAL
    rc = _lib.lib.dvc_decode_wz_frames(None, 1, None, 1)
    assert rc == _lib.DVC_EINVAL
    assert b"NULL" in _lib.lib.dvc_last_error()
