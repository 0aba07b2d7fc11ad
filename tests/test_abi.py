"""The C ABI libraries load and export every function include/*.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

import scd_amd
from scd_amd import _lib

INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(scd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("scd_host.h", "libscd_host.so"), ("scd_hip.h", "libscd_hip.so"),
                                        ("scd_serve.h", "libscd_serve.so")])
def test_library_exports_header(header, lib):
    names = declared(header)
    assert len(names) >= 8
    L = ctypes.CDLL(os.path.join(_lib.LIB_DIR, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_dropin_exports_reference_names():
    path = os.path.join(_lib.LIB_DIR, "libscd_dropin.so")
    L = ctypes.CDLL(path)
    # the six functions of src/cuda.hpp:23-28 with C++ linkage
    for sym in ("_Z9cuda_initiib", "_Z19cuda_transfer_graphPK5Graphd", "_Z21cuda_labels_to_devicePj",
                "_Z19cuda_improve_labelsPd", "_Z17cuda_store_labelsv", "_Z23cuda_labels_from_devicePj"):
        assert hasattr(L, sym), sym


def test_refapi_exports_reference_names():
    L = ctypes.CDLL(os.path.join(_lib.LIB_DIR, "libscd_refapi.so"))
    # Graph::read (src/graph.hpp:41) and the functions of src/scd.hpp:33-36
    # and src/wcc.hpp:24, C++ linkage
    for sym in ("_ZN5Graph4readEPKc", "_Z24create_initial_partitionPK5Graph", "_Z15compress_labelsPK5GraphPjS2_",
                "_Z13collect_statsPK5GraphP9PartitionPKb", "_Z17improve_partitionPK5GraphP9PartitiondPKb",
                "_Z11compute_wccPK5GraphdPjPiPKb"):
        assert hasattr(L, sym), sym


def test_init_without_device_fails_cleanly():
    if scd_amd.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(scd_amd.ScdError) as ei:
        scd_amd.Engine(device=0)
    assert ei.value.code == -4  # SCD_EDEVICE, like cuda_init returning false


@pytest.mark.parametrize("size", [0, -64, 1025])
def test_init_rejects_bad_group_size(size):
    with pytest.raises(scd_amd.ScdError) as ei:
        scd_amd.Engine(device=0, work_group_size=size)
    assert ei.value.code == -1


def test_server_without_device_fails_cleanly():
    if scd_amd.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(scd_amd.ScdError):
        scd_amd.Server(device=0)


def test_server_footprint_grows_with_the_graph():
    f = scd_amd.Server.footprint
    assert f(1000, 5000) < f(2000, 5000) < f(2000, 10000)
    # at least the resident CSR and the arena's adjacency copies
    assert f(1000, 5000) >= 4 * (1001 + 10000) + 4 * 10000

