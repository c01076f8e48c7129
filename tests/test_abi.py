"""The C-ABI library loads and exports every symbol include/*.h declares;
host-side behaviour that needs no GPU (status codes, the coh CLI)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

import gpu_bitruss_amd as gb
from gpu_bitruss_amd import _lib
from tests import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("bitruss_hip.h", "bt_graph.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+\*?(bt_[a-z_0-9]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20, names
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings():
    L = _lib.lib()
    for code in range(0, 13):  # BT_OK .. BT_ERR_SCRATCH
        assert L.bt_status_str(code) and L.bt_status_str(code) != b"unknown status", code
    assert L.bt_status_str(12345) == b"unknown status"
    assert (_lib.BT_PARTIAL, _lib.BT_ERR_BUSY, _lib.BT_ERR_SCRATCH) == (10, 11, 12)


def test_options_are_checked():
    """Context options reject out-of-range values before any device work
    (the context handle is checked first: no device is needed)."""
    L = _lib.lib()
    for opt in (_lib.OPT_PEEL_MAX_LEVEL, _lib.OPT_PEEL_C2CAP, _lib.OPT_PEEL_WINDOW_EDGES, _lib.OPT_PEEL_BIG_MIDDLES):
        assert L.bt_ctx_set_option(None, opt, 1) == 1


@pytest.mark.skipif(gb.device_count() > 0, reason="checks the no-device path")
def test_no_device_is_reported():
    ctx = C.c_void_p()
    assert _lib.lib().bt_ctx_create(0, C.byref(ctx)) == 5  # BT_ERR_NO_DEVICE


def test_invalid_arguments_are_rejected():
    L = _lib.lib()
    assert L.bt_ctx_create(0, None) == 1
    assert L.bt_graph_upload(None, None, None) == 1
    assert L.bt_msp(None, None, None, None) == 1
    assert L.bt_graph_check(None) == 1


def run_coh(*args):
    return subprocess.run([_lib.COH_PATH, *args], capture_output=True, text=True, timeout=60)


def test_coh_cli(tmp_path):
    assert run_coh("--help").returncode == 0
    v = run_coh("--version")
    assert v.returncode == 0 and v.stdout.strip() == "1.3.1"
    assert run_coh("--bogus").returncode == 1
    txt = tmp_path / "aw.graph"
    txt.write_text("".join(f"{u} {v}\n" for u, v in G.AW_EDGES))
    b = tmp_path / "aw.bin"
    r = run_coh("--graph", str(txt), "--bin", str(b))
    assert r.returncode == 0, r.stdout
    h = gb.Graph(str(b), False)
    assert h.m == 10
    # non-.bin input is refused (graph.cpp:271-274)
    r = run_coh("--gpu", "--graph", str(txt))
    assert r.returncode == 1 and "binary file" in r.stdout
    # the CPU algorithms are not served by the product
    r = run_coh("--cpu", "--graph", str(b))
    assert r.returncode == 1
    if gb.device_count() == 0:
        r = run_coh("--gpu", "--graph", str(b))
        assert r.returncode == 1 and "no gpu devices" in r.stdout
        r = run_coh("--device", "3", "--graph", str(b))
        assert r.returncode == 1 and "not available" in r.stdout
