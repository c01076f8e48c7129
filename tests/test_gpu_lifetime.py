"""Handles that outlive their context, through the C-ABI: destroying a
context releases the device memory of the graphs and batches still made on
it and detaches them, so that freeing them afterwards only deletes the
handle (no use of the destroyed context), and the device stays usable."""
from __future__ import annotations

import numpy as np
import pytest

import gpu_bitruss_amd as gb
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.gpu
def test_free_after_context_destroyed(dev):
    g = gb.Graph.generate("mix_graph", 0)
    want, wlevels, _ = oracle.msp(g, 8)
    d = gb.Device(dev.device)
    dg = d.upload(g)
    b = d.upload_batch([g, g])
    res, levels = dg.msp()
    assert np.array_equal(res, want) and levels == wlevels
    outs, _ = b.run("msp")
    assert np.array_equal(outs[1], want)
    d.close()
    dg.free()
    b.free()
    with gb.Device(dev.device) as d2:
        dg2 = d2.upload(g)
        res2, _ = dg2.msp()
        dg2.free()
    assert np.array_equal(res2, want)
