"""The C-ABI library loads and exports every symbol include/esrrf.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re

import numpy as np

from conftest import REPO


def declared():
    txt = open(os.path.join(REPO, "include", "esrrf.h")).read()
    return sorted(set(re.findall(r"\b(esrrf_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_binding_agree(pkg):
    assert sorted(pkg._capi.EXPORTS) == declared()


def test_library_exports_every_declared_symbol(pkg):
    L = ctypes.CDLL(pkg._capi.LIB_PATH)
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_library_has_gfx950_code_object(pkg):
    blob = open(pkg._capi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_host_only_entry_points(pkg):
    # pure host computations of the ABI (no device touched)
    k = pkg._capi.bic_constants(1.5)
    assert abs(k["radius"] - 2.0) < 1e-7
    # Float.toString(2*sigma) + "1" (LiveSRRF_CL.java:318), read as a float
    assert k["tSO"] == float(np.float32(1.27442651))
    assert pkg._capi.bic_tso_text(1.5) == "1.27442651"
    f = pkg._capi.fht_constants(1.5, 5)
    assert f["radius"] == 10.0
