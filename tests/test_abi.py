"""The C-ABI library loads on a machine without a GPU and exports every
function that include/*.h declares; no compute calls are made here."""
import ctypes
import re
from pathlib import Path

import ebi_amd as E

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(ebi_\w+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(E._LIB_PATH))
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_context_creation_fails_cleanly_without_gpu_or_succeeds_with_one():
    h = ctypes.c_void_p()
    rc = E.lib().ebi_hip_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        E.lib().ebi_hip_ctx_destroy(h)
    else:
        assert rc in (E.HIP, E.INVALID_ARGUMENT)


def test_encode_rejects_bad_config():
    import numpy as np
    import pytest

    with pytest.raises(E.EbiError):
        E.encode(np.arange(10.0), E.CompressorConfig(9, 1, 8192), E.ChunkOption.record_count(4))  # Zstd
    with pytest.raises(E.EbiError):
        E.encode(np.arange(10.0), E.CompressorConfig.gorilla(), E.ChunkOption.record_count(0))
