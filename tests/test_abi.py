"""The C-ABI library loads without a GPU and exports every entry point its
header declares; the wire structs have the reference's sizes."""
import ctypes
import os
import re

from changa_amd import build, gpu

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                   "changa_gpu.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:void|int)\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_expected_surface():
    names = declared()
    for n in ["DataManagerTransferLocalTree", "TreePieceCellListDataTransferLocal",
              "TreePiecePartListDataTransferLocal", "EwaldHost", "TransferParticleVarsBack",
              "changaBuildMoments"]:
        assert n in names
    assert set(names) == set(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(build.GPU_LIB), "build first: python -m changa_amd.build"
    lib = ctypes.CDLL(build.GPU_LIB)
    for n in declared():
        assert hasattr(lib, n), n
    gpu.lib()  # argtypes set up without touching a device


def test_wire_struct_sizes():
    assert ctypes.sizeof(gpu.MomcData) == 128
    assert ctypes.sizeof(gpu.EwtData) == 20
    assert ctypes.sizeof(gpu.CudaRequest) > 0
    assert ctypes.sizeof(gpu.ChangaGravityLaunch) == 8 * 8 + 4 * 4 + 3 * 8
