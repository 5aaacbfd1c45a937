"""CPU-side checks of the drop-in boundary: the libraries load, export every
symbol include/*.h declares, the Python mirror prepares inputs exactly as the
reference binding does, and argument errors are reported without touching a
GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from seismic_amd import _lib, correlation_lib, multicuda_lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", txt)))


@pytest.mark.parametrize("lib,header", [("libseismic_xcorr.so", "seismic_xcorr.h"),
                                         ("correlation_cuda.so", "correlation_cuda.h"),
                                         ("correlation_multicuda.so", "correlation_multicuda.h"),
                                         ("correlation_c.so", "correlation_c.h")])
def test_exports(lib, header):
    path = _lib.library_path(lib)
    assert os.path.exists(path), "build the libraries first (python -m seismic_amd.build)"
    h = ctypes.CDLL(path)
    names = _declared(header)
    assert names, header
    for n in names:
        assert hasattr(h, n), f"{lib} does not export {n}"
    for n in _lib.EXPORTS[lib]:
        assert n in names


def test_prepare_matches_reference_rules():
    ev, _ = synth.make_events(5, 1, 1500, seed=1)
    traces = [e[0] for e in ev]
    n, L, P, padded, rev, outs = correlation_lib.prepare(np.array(traces))
    assert (n, L, P) == (5, 1500, 2048)          # P = 1 << L.bit_length()
    assert np.array_equal(padded[:, :L], np.array(traces))
    assert not padded[:, L:].any()
    assert np.array_equal(rev[:, :L], np.array(traces)[:, ::-1])
    assert all(o.shape == (15,) and not o.any() for o in outs)
    # power-of-two lengths get a strictly larger period
    assert correlation_lib.prepare(np.zeros((2, 1024), np.float32))[2] == 2048


def test_prepare_accepts_trace_objects():
    class Tr:
        def __init__(self, d):
            self.data = d
    ev, _ = synth.make_events(3, 1, 700, seed=2)
    objs = np.empty(3, dtype=object)
    for i in range(3):
        objs[i] = Tr(ev[i, 0])
    n, L, P, padded, rev, outs = correlation_lib.prepare(objs)
    assert (n, L, P) == (3, 700, 1024)
    assert np.array_equal(padded[:, :700], ev[:, 0])


@pytest.mark.parametrize("shift,L,P", [(600, 700, 1024), (-1, 1024, 2048)])
def test_invalid_shift_reported(shift, L, P):
    lib = correlation_lib.initialise_library()
    n = 3
    ev = np.zeros((n, P), np.float32)
    outs = [np.full(6, 7, np.float32), np.full(6, 7, np.int32), np.full(6, 7, np.float32),
            np.full(6, 7, np.int32)]
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    lib.correlationCUDA(fp(ev), fp(ev), n, L, shift, P, 1, 0.0, fp(outs[0]), ip(outs[1]),
                        fp(outs[2]), ip(outs[3]))
    assert b"shift" in lib.seismic_last_error()
    assert all((o == 7).all() for o in outs)       # untouched on error


def test_invalid_plan_reported():
    from seismic_amd import engine
    with pytest.raises(RuntimeError, match="padded_size"):
        engine.AllPairsPlan(4, 1, 100, 1000, -10, 10)
    # supported periods: P = 128 R, R in {8, 16, 24, 32, 40, 48, 64}
    for bad in (1536, 2560, 7168, 12288, 16384, 512):
        with pytest.raises(RuntimeError, match="one of 1024, 2048, 3072, 4096, 5120, 6144, 8192"):
            engine.AllPairsPlan(4, 1, 100, bad, -10, 10)
    with pytest.raises(RuntimeError, match="lag window"):
        engine.AllPairsPlan(4, 1, 100, 1024, 10, -10)
    with pytest.raises(RuntimeError, match="n_channels"):
        engine.AllPairsPlan(4, 0, 100, 1024, -10, 10)
    # work items (row, 64-column chunk) number sum_c min(N, 64(c+1)) and are
    # int-indexed: N = 524224 gives 2,147,221,504 items, N = 524225 2,147,745,729
    with pytest.raises(RuntimeError, match="work-list limit"):
        engine.AllPairsPlan(524225, 1, 100, 1024, -10, 10)


def test_multicuda_binding_loads():
    lib = multicuda_lib.initialise_library()
    assert hasattr(lib, "correlationCUDA") and hasattr(lib, "initialiseCUDA")


def test_omp_binding_loads_and_rejects_bad_classes():
    from seismic_amd import omp_lib
    lib = omp_lib.initialize_library()
    n, P = 3, 1024
    ev = np.zeros((n, P), np.float32)
    outs = [np.full(6, 7, np.float32), np.full(6, 7, np.int32), np.full(6, 7, np.float32),
            np.full(6, 7, np.int32)]
    hist = [np.zeros(4, np.int32), np.zeros(4, np.int32)]
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    lib.correlationCPP(fp(ev), fp(ev), n, 500, 100, P, 4, 1, 0.0, 0, fp(outs[0]), ip(outs[1]),
                       fp(outs[2]), ip(outs[3]), ip(hist[0]), ip(hist[1]))
    assert b"n_classes_hist" in lib.seismic_last_error()
    assert all((o == 7).all() for o in outs) and not hist[0].any()


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror in seismic_amd._lib has the size and field offsets
    of its C struct in include/seismic_xcorr.h (compiled with the host gcc)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no host C compiler")
    mirrors = {"sxc_plan_desc": _lib.SxcPlanDesc, "sxc_stats": _lib.SxcStats,
               "sxc_timing": _lib.SxcTiming, "sxc_server_desc": _lib.SxcServerDesc,
               "sxc_match": _lib.SxcMatch, "sxc_serve_stats": _lib.SxcServerStats,
               "sxc_block_info": _lib.SxcBlockInfo}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "seismic_xcorr.h"', "int main(void) {"]
    for name, cls in mirrors.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for name, cls in mirrors.items():
        assert got[(name, "size")] == ctypes.sizeof(cls), name
        for f, _ in cls._fields_:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)


# The reference's exported signatures, as its sources declare them
# (Correlation_in_CUDA/correlation_cuda.cu:509,530-533,
# Correlation_in_MultiCUDA/correlation_multicuda.cu:19,48-51,
# Correlation_in_OMP/correlation_c.cpp:210-213).
_REFERENCE_SIGNATURES = {
    "correlation_cuda.h": [
        "void (*f0)(void) = initialiseCUDA;",
        "void (*f1)(float*, float*, int, int, int, int, int, float, float*, int*, float*, int*)"
        " = correlationCUDA;"],
    "correlation_multicuda.h": [
        "void (*f0)(int) = initialiseCUDA;",
        "void (*f1)(float*, float*, int, int, int, int, float, int, float*, int*, float*, int*)"
        " = correlationCUDA;"],
    "correlation_c.h": [
        "void (*f1)(float*, float*, int, int, int, int, int, int, float, int, float*, int*, float*,"
        " int*, int*, int*) = correlationCPP;"],
}


@pytest.mark.parametrize("header", sorted(_REFERENCE_SIGNATURES))
def test_dropin_signatures_match_reference(header, tmp_path):
    """Each drop-in header declares the reference's exact signature: taking
    the function's address into a pointer of the reference type compiles
    without a warning, as C and as C++."""
    import shutil
    import subprocess
    for comp, std, ext in (("gcc", "-std=c99", "c"), ("g++", "-std=c++17", "cpp")):
        exe = shutil.which(comp)
        if exe is None:
            pytest.skip(f"no {comp}")
        body = "\n".join(f"  {l}" for l in _REFERENCE_SIGNATURES[header])
        src = tmp_path / f"sig.{ext}"
        src.write_text(f'#include "{header}"\nint main(void) {{\n{body}\n'
                       f'  return (int)(sizeof(f1) == 0);\n}}\n')
        r = subprocess.run([exe, std, "-Wall", "-Werror", "-Wno-unused-variable", "-I",
                            os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "sig.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def _header_params(header, func):
    """Parameter C types of `func` as `header` declares it."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", header)).read(), flags=re.S)
    m = re.search(r"\b" + func + r"\s*\(([^)]*)\)\s*;", txt)
    assert m, (header, func)
    out = []
    for p in m.group(1).split(","):
        p = " ".join(p.split())
        if p == "void":
            continue
        ptr = "*" in p
        base = p.replace("*", " ").split()[0]
        out.append(base + ("*" if ptr else ""))
    return out


@pytest.mark.parametrize("module,loader,header,func", [
    ("correlation_lib", "initialise_library", "correlation_cuda.h", "correlationCUDA"),
    ("multicuda_lib", "initialise_library", "correlation_multicuda.h", "correlationCUDA"),
    ("omp_lib", "initialize_library", "correlation_c.h", "correlationCPP"),
])
def test_binding_argtypes_match_headers(module, loader, header, func):
    """The Python bindings' ctypes argtypes are the headers' parameter types."""
    import importlib
    names = {ctypes.POINTER(ctypes.c_float): "float*", ctypes.POINTER(ctypes.c_int): "int*",
             ctypes.c_int: "int", ctypes.c_float: "float"}
    lib = getattr(importlib.import_module(f"seismic_amd.{module}"), loader)()
    got = [names[t] for t in getattr(lib, func).argtypes]
    assert got == _header_params(header, func)

