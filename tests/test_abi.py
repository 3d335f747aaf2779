"""The C-ABI library loads on a host without a GPU and exports exactly the
entry points declared in include/mappraiser_gpu.h (no compute calls here)."""
import os
import re
import subprocess

from midapack_amd import lib as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mappraiser_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mappraiser_gpu_\w+)\s*\(", text)))


def test_header_matches_python_binding():
    assert declared_functions() == sorted(G.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = G.load()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mappraiser_gpu_\w+)", out))
    assert set(declared_functions()) <= exported
    # the reference's own entry points (map.h:10-18), bound unchanged by its ctypes wrapper
    for name in G.REFERENCE_EXPORTS:
        assert re.search(rf" T {name}$", out, flags=re.M), name
        assert re.search(rf"\bvoid {name}\(", open(HEADER).read()), name


def test_host_only_entry_points():
    o = G.options()
    assert o.precision == G.MGPU_FP64 and o.refresh_every == 10 and o.poll_every == 8
    assert isinstance(G.last_error(), str)
