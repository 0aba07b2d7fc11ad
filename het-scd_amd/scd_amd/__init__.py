"""scd_amd: MI355X-native SCD community detection (host side).

The compute path is the gfx950 HIP engine in libscd_hip.so (include/scd_hip.h);
this package binds it and the host library (binary format, generators,
convert) with ctypes.  There is no CPU fallback: Engine raises ScdError when
the HIP library is missing or no device is present.
"""
from ._lib import INVALID_LABEL, ScdError  # noqa: F401
from .engine import BatchEngine, Engine, device_count  # noqa: F401
from .server import Server  # noqa: F401
from .graph import (Arena, Edges, Graph, compact_labels, compress_count, from_edge_pairs,  # noqa: F401
                    planted_batch, planted_partition, planted_random, read_labels, rmat, write_labels)
