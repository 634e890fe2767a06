"""CPU tests of the boundary: the C ABI library loads and exports every symbol
include/vite_amd.h declares, the host C++ layer exports the reference's entry points, and
the CLI rejects the flag combinations the reference rejects (main.cpp:710-769).  No compute
call is made (there is no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vite_amd.h")
LIB = os.path.join(ROOT, "vite_amd", "libvite_amd.so")
HOSTLIB = os.path.join(ROOT, "vite_amd", "libvite_host.so")
CLI = os.path.join(ROOT, "vite_amd", "bin", "vite_amd_cluster")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|uint64_t)\s+(vt_\w+)\s*\(", text, re.M)))


def dynamic_symbols(path, demangle=False):
    out = subprocess.run(["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [path],
                         capture_output=True, text=True, check=True).stdout
    return out


def test_header_declares_entry_points():
    fns = declared_functions()
    for f in ["vt_ctx_create", "vt_graph_upload", "vt_louvain_phase", "vt_sweep", "vt_rebuild", "vt_color",
              "vt_run", "vt_phase_begin", "vt_phase_step", "vt_phase_sweep", "vt_phase_end", "vt_gen_rgg"]:
        assert f in fns


def test_library_exports_every_declared_symbol():
    syms = dynamic_symbols(LIB)
    exported = set(re.findall(r"\sT\s(vt_\w+)", syms))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    import vite_amd
    assert set(vite_amd.exported_symbols()) == set(declared_functions())


def test_context_without_gpu_fails_cleanly():
    import torch  # noqa: F401  (only to report the environment; not used by the library)
    import vite_amd
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(vite_amd.VtError):
        vite_amd.Context(0)


REFERENCE_ENTRY_POINTS = [
    "distLouvainMethod(int, int, DistGraph const&, unsigned long&, unsigned long&, std::vector<long, std::allocator<long> >&, std::vector<long, std::allocator<long> >&, std::vector<long, std::allocator<long> >&, std::vector<long, std::allocator<long> >&, std::vector<long, std::allocator<long> >&, double, double, int&)",
    "distbuildNextLevelGraph(",
    "distColoringMultiHashMinMax(",
    "loadDistGraphMPIIO(",
    "generateInMemGraph(",
]


def test_host_layer_exports_reference_entry_points():
    syms = dynamic_symbols(HOSTLIB, demangle=True)
    for name in REFERENCE_ENTRY_POINTS:
        assert name in syms, name
    assert syms.count("distLouvainMethod(") == 3
    assert syms.count("distLouvainMethodWithColoring(") == 3
    assert syms.count("distLouvainMethodVertexOrder(") == 3


@pytest.mark.parametrize("args,msg", [
    (["-f", "x.bin", "-t", "7"], "early-term-type parameter is between 1-4"),
    (["-f", "x.bin", "-t", "2", "-a", "1.5"], "early-term-alpha must be between 0 and 1"),
    (["-f", "x.bin", "-c", "8", "-d", "8"], "Cannot turn on both coloring"),
    (["-f", "x.bin", "-p", "-i"], "Cannot turn on threshold-cycling"),
    ([], "Must specify a binary file name with -f"),
    (["-f", "x.bin", "-e", "1.0"], "Must specify -n"),
    (["-f", "x.bin", "-g", ""], "Must specify a ground-truth file name with -g"),
])
def test_cli_rejects_what_the_reference_rejects(args, msg):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert msg in r.stderr


def test_cli_missing_file_is_fatal(tmp_path):
    r = subprocess.run([CLI, "-f", str(tmp_path / "missing.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "Error opening file" in r.stderr


def test_cli_missing_ground_truth_is_fatal(tmp_path):
    """The ground truth is read before the graph (main.cpp:119-137), as the reference does."""
    r = subprocess.run([CLI, "-f", str(tmp_path / "g.bin"), "-g", str(tmp_path / "missing.truth")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "Error opening ground truth file" in r.stderr


def test_server_without_gpu_fails_cleanly():
    import torch  # noqa: F401
    import vite_amd
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(vite_amd.VtError):
        vite_amd.Server(slots=2)


def test_request_reservation_grows_with_size():
    import vite_amd
    a = vite_amd.request_reservation("rgg", 1, 1 << 10, 0.0)
    b = vite_amd.request_reservation("rgg", 1, 1 << 20, 0.0)
    c = vite_amd.request_reservation("ba", 1, 1 << 20, 8)
    assert 0 < a < b and 0 < c
