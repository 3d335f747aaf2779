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


def test_server_refuses_concurrent_groups():
    # checked before any device call: one group on the device at a time
    for slots in (0, 2, 4):
        try:
            G.Server(max_slots=slots).close()
        except G.GPUError as e:
            assert f"max_slots={slots}" in str(e)
        else:
            raise AssertionError(f"max_slots={slots} was accepted")


def _params(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    m = re.search(rf"\bvoid {name}\((.*?)\);", text, flags=re.S)
    assert m, name
    return [re.sub(r"\s+", " ", p.strip()) for p in m.group(1).split(",")]


def test_reference_entry_point_signatures():
    """MLmap / MTmap take the reference's parameters in its order (map.h:10-18;
    MPI_Comm arrives as one integer-sized handle), so its ctypes argtypes
    (mappraiser_wrapper/__init__.py:58-122) line up one to one."""
    ml = [p.rsplit(" ", 1)[-1].lstrip("*") for p in _params("MLmap")]
    assert ml == ["comm", "outpath", "ref", "solver", "precond", "Z_2lvl", "pointing_commflag", "tol", "maxiter",
                  "enlFac", "ortho_alg", "bs_red", "nside", "data_size_proc", "nb_blocks_loc", "local_blocks_sizes",
                  "Nnz", "pix", "pixweights", "signal", "noise", "lambda", "invtt"]
    mt = _params("MTmap")
    names = [p.rsplit(" ", 1)[-1].lstrip("*") for p in mt]
    assert names == ["comm", "outpath", "ref", "solver", "precond", "Z_2lvl", "pointing_commflag", "npoly", "nhwp",
                     "delta_t", "ground", "n_sss_bins", "tol", "maxiter", "enlFac", "ortho_alg", "bs_red", "nside",
                     "sweeptstamps", "nsweeps", "az", "az_min", "az_max", "hwp_angle", "nces", "data_size_proc",
                     "nb_blocks_loc", "local_blocks_sizes", "Nnz", "pix", "pixweights", "signal", "noise",
                     "sampling_freq", "invtt"]
    assert mt[18].startswith("int **") and mt[20].startswith("double **") and mt[23].startswith("double **")
    L = G.load()
    assert len(L.MLmap.argtypes) == len(ml) and len(L.MTmap.argtypes) == len(names)
