"""The C ABI library loads and exports every function include/*.h declares (no GPU calls)."""
import os
import re

import numpy as np

from biofem_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bf_\w+|bf_FInternalAct|bf_GlobalGhost|bf_UpdateElementThickness|bf_FEDestroy)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = L.lib()
    names = declared("biofem_hip.h") + declared("biofem_fe.h") + declared("biofem_io.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_the_abi():
    names = set(declared("biofem_hip.h")) | set(declared("biofem_io.h"))
    assert names <= set(L.SIGNATURES), names - set(L.SIGNATURES)


def test_version_and_argument_errors_without_device():
    lib = L.lib()
    assert lib.bf_version().startswith(b"biofem-hip")
    assert lib.bf_loop_tables(2, L.dp(np.zeros(64))) == L.BF_ERR_ARG
    assert lib.bf_ctx_destroy(None) == L.BF_OK
    assert lib.bf_force_eval(None, None, None, None, None, 0, 0) == L.BF_ERR_ARG
