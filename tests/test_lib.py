"""CPU checks of the native library: it builds, loads, exports every symbol
declared in include/*.h, and its host-side argument checks answer without a
GPU."""
import ctypes
import glob
import os
import re

import pytest

import gmm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        syms.update(re.findall(r"\b(gmm(?:hip|host)_[a-z0-9_]+)\s*\(", src))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "gmmhip_accumulate" in syms and "gmmhip_set_model" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    L = gmm_amd.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_lists_every_export():
    assert set(declared_symbols()) <= set(gmm_amd.EXPORTS) | set(getattr(gmm_amd, "EXTRA_EXPORTS", []))


def test_default_opts():
    o = gmm_amd.default_opts()
    assert o.min_loglike == pytest.approx(-1e20, rel=1e-6)
    assert o.min_gamma == pytest.approx(1e-4, rel=1e-6)
    assert o.min_var == pytest.approx(1e-6, rel=1e-6)
    assert o.precision == gmm_amd.PREC_FP32


def test_create_without_device_reports_nodev():
    if gmm_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = gmm_amd.lib().gmmhip_create(0, None, ctypes.byref(h))
    assert rc == gmm_amd.ERR_NODEV
    assert not h.value


def test_null_handle_calls_are_rejected():
    L = gmm_amd.lib()
    assert L.gmmhip_set_model(None, 1, 1, None, None, None) == gmm_amd.ERR_ARG
    assert L.gmmhip_accumulate(None, 0, None, None, None, None, None) == gmm_amd.ERR_ARG
    assert L.gmmhip_destroy(None) == gmm_amd.OK
