"""The drop-in command line on the GPU: `coh --gpu --algo X --graph F.bin`
(the reference's main.cu:96-129 dispatch, including the fallback of an
unknown algorithm to msp), its per-edge dumps against the oracle, its log
lines against the reference's wording (bfc_vpp.cu:276,289,366,
bitruss_hidx.cu:182,255, bitruss_msp.cu:445,554), the JSON summary, and
`--device_info` (utility.h:127-146)."""
from __future__ import annotations

import json
import subprocess
import sys

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from gpu_bitruss_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-cli] {msg}", flush=True)
    sys.stdout.flush()


def coh(*args):
    r = subprocess.run([_lib.COH_PATH, *args], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        say(f"coh {' '.join(args)}: rc {r.returncode}; stderr: {r.stderr.strip()}")
    return r.returncode, r.stdout


def summary(out: str) -> dict:
    lines = [l for l in out.splitlines() if l.startswith("{")]
    assert len(lines) == 1, out
    return json.loads(lines[0])


@pytest.mark.gpu
def test_coh_gpu_algorithms(tmp_path, dev):
    # coh is another process on the same GPU: this process gives back the
    # device memory its context keeps cached from earlier tests first
    dev.trim()
    g = gb.Graph.generate("small_uniform")
    f = tmp_path / "small_uniform.bin"
    g.graph_to_bin(str(f))
    osup, ototal, osmax, _ = oracle.ebfc(g, 8)
    ohidx, orounds, _ = oracle.hidx(g, 8)
    omsp, olevels, _ = oracle.msp(g, 8)
    omspp, opplevels, _ = oracle.mspp(g, 8)
    expect = {"ebfc": osup, "hidx": ohidx, "msp": omsp, "mspp": omspp, "peel": ohidx}
    wording = {
        "bfc": ["butterfly counting with vertex priority on gpu", f"btf number: {ototal:,}"],
        "ebfc": [f"the max edge support is {osmax:,}", f"total butterfly counting with vertex priority on gpu: {ototal:,}"],
        "hidx": ["running bitruss decomposition based on h-index on GPU",
                 f"k-bitruss max value {int(ohidx.max())}, with {orounds} rounds"],
        "msp": ["running bitruss decomposition online peeling", f"bitruss msp with {olevels:,} levels"],
        "mspp": ["running bitruss decomposition online peeling (baseline)",
                 f"(baseline) bitruss msp with {opplevels:,} levels"],
        "peel": ["running bitruss decomposition by peeling", f"k-bitruss max value {int(ohidx.max())} by peeling"],
    }
    for algo in ("bfc", "ebfc", "hidx", "msp", "mspp", "peel", "no_such_algo"):
        say(f"coh --gpu --algo {algo} on {f.name} (m={g.m})")
        dump = tmp_path / f"{algo}.u64"
        rc, out = coh("--gpu", "--algo", algo, "--graph", str(f), "--dump", str(dump), "--json")
        assert rc == 0, out
        s = summary(out)
        used = "msp" if algo == "no_such_algo" else algo  # main.cu:126-127
        assert s["algo"] == used and s["m"] == g.m and s["n"] == g.n
        for w in wording[used]:
            # %' groups digits by the locale; the box's C locale prints none
            assert w in out or w.replace(",", "") in out, (algo, w, out)
        if used in ("bfc", "ebfc"):
            assert s["total"] == ototal
        if used == "ebfc":
            assert s["support_max"] == osmax
        if used == "hidx":
            assert s["iters"] == orounds
        if used == "msp":
            assert s["iters"] == olevels
        if used == "mspp":
            assert s["iters"] == opplevels
        if used in expect:
            got = np.fromfile(str(dump), dtype="<u8")
            assert np.array_equal(got, expect[used]), algo
        else:
            assert not dump.exists()


@pytest.mark.gpu
def test_coh_device_info():
    rc, out = coh("--device_info")
    say(out.strip().splitlines()[2] if out.count("\n") > 2 else out)
    assert rc == 0, out
    rows = {l.split("|")[0].strip(): l.split("|")[1].strip() for l in out.splitlines() if "|" in l}
    assert rows["Device Number"] == "0"
    assert rows["Architecture"].startswith("gfx950"), rows
    assert int(rows["Number of CUs"]) > 0 and int(rows["Wavefront size"]) == 64
    assert int(rows["Total global memory (bytes)"]) > 100 << 30
