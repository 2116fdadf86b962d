"""The C-ABI library loads and exports every function include/pibronic_hip.h declares
(CPU only: no call that touches a device)."""
import ctypes
import os
import re

from conftest import ROOT
from pibronic_amd import _native

HEADER = os.path.join(ROOT, "include", "pibronic_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pib_[a-z_0-9]+)\s*\(", text)))


def _defines():
    out = {}
    for m in re.finditer(r"#define\s+(PIB_[A-Z_0-9]+)\s+\(?(-?\d+)\)?", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), f"{n} declared in the header but not exported"


def test_binding_covers_header():
    assert set(_declared()) == set(_native.SIGNATURES)


def test_abi_version_and_constants():
    lib = _native.load()
    d = _defines()
    assert lib.pib_abi_version() == d["PIB_ABI_VERSION"] == _native.ABI_VERSION
    pairs = {"PIB_OK": _native.PIB_OK, "PIB_ERR_ARG": _native.PIB_ERR_ARG,
             "PIB_ERR_HIP": _native.PIB_ERR_HIP, "PIB_ERR_STATE": _native.PIB_ERR_STATE,
             "PIB_ERR_UNSUPPORTED": _native.PIB_ERR_UNSUPPORTED,
             "PIB_SET_RHO": _native.SET_RHO, "PIB_SET_VIB": _native.SET_VIB,
             "PIB_SET_VIB_PLUS": _native.SET_VIB_PLUS, "PIB_SET_VIB_MINUS": _native.SET_VIB_MINUS,
             "PIB_EVAL_PM": _native.EVAL_PM, "PIB_EVAL_NOSCALE": _native.EVAL_NOSCALE,
             "PIB_OUT_RHO": _native.OUT_RHO, "PIB_OUT_G": _native.OUT_G, "PIB_OUT_GP": _native.OUT_GP,
             "PIB_OUT_GM": _native.OUT_GM, "PIB_OUT_LOGS": _native.OUT_LOGS,
             "PIB_NUM_OUT": _native.NUM_OUT, "PIB_NUM_SUMS": _native.NUM_SUMS,
             "PIB_SUM_R": _native.SUM_R, "PIB_SUM_R2": _native.SUM_R2, "PIB_SUM_ABSR": _native.SUM_ABSR,
             "PIB_SUM_D1": _native.SUM_D1, "PIB_SUM_D2": _native.SUM_D2, "PIB_SUM_RM": _native.SUM_RM}
    for k, v in pairs.items():
        assert d[k] == v, k


def test_null_context_is_rejected():
    lib = _native.load()
    assert lib.pib_set_beads(None, 12, 0.1) == _native.PIB_ERR_ARG
    assert lib.pib_eval_block(None, None, 0, 0, None, 0, None, None) == _native.PIB_ERR_ARG
    assert lib.pib_ctx_destroy(None) == _native.PIB_OK
    assert lib.pib_last_error(None) == b"null context"
