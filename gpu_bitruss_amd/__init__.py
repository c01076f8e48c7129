"""MI355X-native bitruss hot path: butterfly counting (bfc), per-edge
butterfly support (ebfc), h-index bitruss (hidx) and GBiD multi-stage
peeling (msp), as hand-written HIP kernels for gfx950 behind a C-ABI
(include/bitruss_hip.h, include/bt_graph.h).
"""
from .graph import Graph, WORKLOADS
from .device import (
    Batch,
    Device,
    DeviceGraph,
    Server,
    bfc_evpp,
    bfc_vpp,
    bitruss_msp,
    bitruss_mspp,
    bitruss_peeling,
    default_device,
    device_count,
    g_bitruss_hindex,
    tune_candidates,
    tune_class,
    tune_default,
)
from ._lib import BitrussError, LIB_PATH, COH_PATH

__all__ = [
    "Graph", "WORKLOADS", "Batch", "Device", "Server", "DeviceGraph", "bfc_vpp", "bfc_evpp", "g_bitruss_hindex",
    "bitruss_msp", "bitruss_mspp", "bitruss_peeling", "default_device", "device_count", "BitrussError", "LIB_PATH", "COH_PATH",
    "tune_candidates", "tune_class", "tune_default",
]
