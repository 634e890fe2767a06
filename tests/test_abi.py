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


# ---- object layout against the reference headers --------------------------------------------

REF_DIR = "/root/reference"
MPI_INC = "/opt/conda/include"
PROBE = os.path.join(ROOT, "tests", "probes", "layout_probe.cpp")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "graphClustering_vite_amd")


def _probe(tmp_path, name, includes, header):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++11", "-O1"] + [f"-I{d}" for d in includes] +
                   [f'-DGRAPH_HEADER="{header}"', "-o", exe, PROBE], check=True, capture_output=True, text=True)
    return subprocess.run([exe], check=True, capture_output=True, text=True, timeout=60).stdout


def test_graph_layout_matches_reference_headers(tmp_path):
    """Edge, Comm, Graph and DistGraph have the same size, member offsets and inline behaviour
    through vite_graph.hpp as through the reference's distgraph.hpp (graph.hpp:77-106,
    distgraph.hpp:77-107), so objects built by code compiled from either header are read
    correctly by the other."""
    if not os.path.isfile(os.path.join(REF_DIR, "distgraph.hpp")):
        pytest.skip("reference headers not present")
    ref = _probe(tmp_path, "probe_ref", [REF_DIR, MPI_INC], "distgraph.hpp")
    ours = _probe(tmp_path, "probe_amd", [os.path.join(ROOT, "vite_amd", "host")], "vite_graph.hpp")
    assert "DistGraph size=32 totalNumVertices=0 totalNumEdges=8 localGraph=16 parts=24" in ref
    assert ours == ref


def test_graph_layout_fixed_values(tmp_path):
    """The layout vite_graph.hpp gives, stated (runs without the reference tree)."""
    ours = _probe(tmp_path, "probe_amd", [os.path.join(ROOT, "vite_amd", "host")], "vite_graph.hpp")
    assert "Edge size=16 tail=0 weight=8" in ours
    assert "Graph size=64 numVertices=0 numEdges=8 edgeListIndexes=16 edgeList=40" in ours
    assert "DistGraph size=32 totalNumVertices=0 totalNumEdges=8 localGraph=16 parts=24" in ours
    assert "copy nv=3 ne=4 tnv=3 tne=4 row1=[1,3) w2=1.50 base=0 bound=3 owner2=0 owner3=1" in ours


def test_host_layer_exports_driver_symbols():
    """Every entry point the reference driver calls (main.cpp:136-541) is exported, with the
    reference's mangled signature."""
    syms = dynamic_symbols(HOSTLIB, demangle=True)
    for name in ["loadDistGraphMPIIOBalanced(int, int, int, DistGraph*&, std::__cxx11::basic_string",
                 "gatherAllComm(int, int, int, std::vector<long, std::allocator<long> >&, std::vector<long, "
                 "std::allocator<long> > const&)",
                 "writeGraph(int, int, DistGraph*&, std::vector<long, std::allocator<long> >&, "
                 "std::__cxx11::basic_string",
                 "balanceEdges(int, std::__cxx11::basic_string",
                 "loadGroundTruthFile(", "compare_communities(", "createCommunityMPIType()",
                 "destroyEdgeMPIType()"]:
        assert name in syms, name


def _dropin():
    if not os.path.isfile(DROPIN):
        pytest.skip("oracle/_ref/graphClustering_vite_amd not built (needs the reference tree)")
    return DROPIN


def test_dropin_driver_resolves_every_symbol():
    """The reference's main.cpp, compiled unmodified against its own headers, links against
    libvite_host.so with nothing left undefined."""
    r = subprocess.run(["ldd", "-r", _dropin()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "undefined symbol" not in r.stdout + r.stderr
    assert "libvite_host.so" in r.stdout


@pytest.mark.parametrize("balanced", [False, True])
def test_dropin_driver_reads_graph(tmp_path, balanced):
    """The reference driver reads a .bin through vite_amd's loaders (-b: the edge-balanced one)
    and its own inline DistGraph::printStats (compiled from distgraph.hpp, run through MPI at
    one rank) reports the graph vite_amd built: the layouts agree at run time."""
    from oracle import pyoracle as po
    g = po.gen_rgg(11, 2000, 1.0)
    path = str(tmp_path / "g.bin")
    po.write_bin(path, g)
    r = subprocess.run([_dropin(), "-f", path, "-j"] + (["-b"] if balanced else []), capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"Number of vertices: {g.nv}\n" in r.stdout
    assert f"Number of edges: {g.ne}\n" in r.stdout
    assert "Variance: 0\n" in r.stdout
    assert ("Trying to achieve equal edge distribution" in r.stdout) == balanced
    assert "Option to just process the graph is turned on" in r.stdout
