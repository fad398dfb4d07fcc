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
    names = re.findall(r"^\s*(?:void|int|size_t)\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)
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
    assert ctypes.sizeof(gpu.ChangaGravityLaunch) == 8 * 8 + 4 * 4 + 4 * 8 + 2 * 4


def test_header_compiles_as_c_and_links(tmp_path):
    """A plain-C caller (what a ChaNGa-side shim would be) compiles against the
    header and links every entry point it takes the address of."""
    import subprocess
    names = declared()
    src = tmp_path / "caller.c"
    body = "\n".join(f"  (void *)&{n}," for n in names)
    src.write_text('#include "changa_gpu.h"\n#include <stdio.h>\n'
                   f"void *table[] = {{\n{body}\n}};\n"
                   "int main(void) { printf(\"%d\\n\", (int)(sizeof(table)/sizeof(table[0]))); "
                   "return sizeof(CudaMultipoleMoments) != 108; }\n")
    inc = os.path.join(os.path.dirname(HDR))
    libdir = os.path.dirname(build.GPU_LIB)
    exe = tmp_path / "caller"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + inc, str(src), "-o", str(exe),
                    "-L" + libdir, "-lchanga_gpu", "-Wl,-rpath," + libdir,
                    "-Wl,--unresolved-symbols=ignore-in-shared-libs"], check=True)


def test_ctypes_mirrors_match_the_header(tmp_path):
    """Every struct the Python harness mirrors has the header's size and
    field offsets (a C program compiled against the header prints them)."""
    import subprocess
    structs = {name: cls for name, cls in vars(gpu).items()
               if isinstance(cls, type) and issubclass(cls, ctypes.Structure)
               and name.startswith("Changa")}
    assert {"ChangaGravityLaunch", "ChangaAdjointLaunch", "ChangaListBuild", "ChangaTreeBuild",
            "ChangaFusedLaunch"} <= set(structs)
    lines = []
    for name, cls in sorted(structs.items()):
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for f in cls._fields_:
            lines.append(f'  printf("{name}.{f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "changa_gpu.h"\n'
                   "int main(void) {\n" + "\n".join(lines) + "\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + os.path.dirname(HDR), str(src),
                    "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(out[name]) == ctypes.sizeof(cls), name
        for f in cls._fields_:
            assert int(out[f"{name}.{f[0]}"]) == getattr(cls, f[0]).offset, f"{name}.{f[0]}"


def test_ltw_layouts_match_the_header(tmp_path):
    """The GPU_LOCAL_TREE_WALK layouts: the header compiled with the define
    gives the node, particle and request records whose offsets the numpy
    records of changa_amd.ltw and the ctypes CudaRequestLTW use."""
    import subprocess
    from changa_amd import ltw
    node_fields = ["radius", "soft", "totalMass", "cm", "lesser_corner", "greater_corner",
                   "bucketStart", "bucketSize", "particleCount", "nodeArrayIndex", "children",
                   "type"]
    lines = ['  printf("node %zu\\n", sizeof(CudaMultipoleMoments));',
             '  printf("node.moments %zu\\n", offsetof(CudaMultipoleMoments, xx));',
             '  printf("part %zu\\n", sizeof(CompactPartData));',
             '  printf("part.nodeId %zu\\n", offsetof(CompactPartData, nodeId));',
             '  printf("req %zu\\n", sizeof(CudaRequest));']
    lines += [f'  printf("node.{f} %zu\\n", offsetof(CudaMultipoleMoments, {f}));'
              for f in node_fields]
    lines += [f'  printf("req.{f[0]} %zu\\n", offsetof(CudaRequest, {f[0]}));'
              for f in gpu.CudaRequestLTW._fields_]
    src = tmp_path / "ltw_layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "changa_gpu.h"\n'
                   "int main(void) {\n" + "\n".join(lines) + "\n  return 0;\n}\n")
    exe = tmp_path / "ltw_layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-DGPU_LOCAL_TREE_WALK",
                    "-I" + os.path.dirname(HDR), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], check=True, capture_output=True,
                              text=True).stdout.splitlines())
    out = {k: int(v) for k, v in out.items()}
    assert out["node"] == ltw.NODE_DTYPE.itemsize == 160
    assert out["node.moments"] == ltw.NODE_DTYPE.fields["moments"][1] == 76
    for f in node_fields:
        assert out[f"node.{f}"] == ltw.NODE_DTYPE.fields[f][1], f
    assert out["part"] == ltw.PART_DTYPE.itemsize == 24
    assert out["part.nodeId"] == ltw.PART_DTYPE.fields["nodeId"][1] == 20
    assert out["req"] == ctypes.sizeof(gpu.CudaRequestLTW)
    for f in gpu.CudaRequestLTW._fields_:
        assert out[f"req.{f[0]}"] == getattr(gpu.CudaRequestLTW, f[0]).offset, f[0]


def test_ltw_library_exports_the_boundary():
    """libchanga_gpu_ltw.so (the boundary on the GPU_LOCAL_TREE_WALK layouts)
    loads without a GPU and exports the HostCUDA / EwaldCUDA entry points."""
    assert os.path.exists(build.LTW_LIB), "build first: python -m changa_amd.build"
    lib = ctypes.CDLL(build.LTW_LIB)
    for n in gpu.LTW_EXPORTS:
        assert hasattr(lib, n), n
    assert set(gpu.LTW_EXPORTS) <= set(declared())
    gpu.lib_ltw()
