"""GPU: the command-line driver end to end against the reference's own main().

A Vite .bin file goes in (the reference's binary graph format, distgraph.cpp:119-253); the
per-vertex community file that -o writes (<input>.communities, renumbered per level as in
main.cpp:380-425 and written as in main.cpp:543-572) and the final modularity come out.
Both must equal what the compiled reference wrote for the same file and flags
(tests/golden/reference_runs.json, "communities").  Sums follow the reference order
(VITE_AMD_SEQUENTIAL_SUMS=1), so every decision and so the partition is identical.

Each case runs through two drivers: vite_amd's own CLI, and the reference's main.cpp
compiled unmodified against its own headers and linked against libvite_host.so
(oracle/_ref/graphClustering_vite_amd, built by oracle/Makefile.ref).  The second is the
drop-in proof: the reference's driver code, its Graph / DistGraph objects and its call
sequence, with every Louvain phase, rebuild, coloring and comparison run on the device.
"""
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import graphs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "vite_amd", "bin", "vite_amd_cluster")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "graphClustering_vite_amd")
DRIVERS = {"cli": CLI, "dropin": DROPIN}
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_runs.json")))
CASES = GOLDEN["communities"]


def _build(spec):
    if spec["kind"] == "rgg":
        return po.gen_rgg(spec["seed"], spec["n"], spec["extra_pct"])
    if spec["kind"] == "ba":
        return po.gen_ba(spec["seed"], spec["n"], spec["m"])
    if spec["kind"] == "vite_rgg":
        return po.gen_vite_rgg(spec["n"])
    return graphs.graph(spec["name"])


def _driver(name):
    exe = DRIVERS[name]
    if not os.path.isfile(exe):
        pytest.skip(f"{exe} not built (the drop-in driver needs the reference tree at build time)")
    return exe


@pytest.mark.parametrize("driver", sorted(DRIVERS))
@pytest.mark.parametrize("idx", range(len(CASES)),
                         ids=[f"{c['graph']}:{' '.join(c['flags']) or 'default'}" for c in CASES])
def test_cli_community_file_matches_reference(tmp_path, idx, driver):
    exe = _driver(driver)
    case = CASES[idx]
    g = _build(GOLDEN["graphs"][case["graph"]]["spec"])
    path = str(tmp_path / (case["graph"] + ".bin"))
    po.write_bin(path, g)
    env = dict(os.environ, VITE_AMD_SEQUENTIAL_SUMS="1")
    out = subprocess.run([exe, "-f", path, "-o"] + case["flags"], capture_output=True, text=True, env=env,
                         timeout=120, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr[-2000:]
    comm = np.loadtxt(path + ".communities", dtype=np.int64, ndmin=1)
    assert len(comm) == case["nv"]
    assert len(np.unique(comm)) == case["count"]
    assert hashlib.sha256(np.ascontiguousarray(comm).tobytes()).hexdigest() == case["sha256"]
    if driver == "dropin":
        return  # the reference's driver prints Q only to six digits
    # the last Louvain call's Q; after a one-phase run (-p) the loop has already reset
    # currMod to -1, as the reference's MODS line shows (main.cpp:474-482, 526-528)
    expect = -1.0 if "-p" in case["flags"] else float.fromhex(case["final_q_hex"])
    m = re.search(r"^Final modularity: (\S+)$", out.stdout, re.M)
    assert m and float(m.group(1)) == expect


STAT_LINES = {"same_same": "Same-Same (True positive)  :", "same_diff": "Same-Diff (False negative) :",
              "diff_same": "Diff-Same (False positive) :", "nc_truth": "#communities in C1 :",
              "nc_found": "#communities in C2 :", "precision": "Precision :", "recall": "Recall    :",
              "f_score": "F-score   :", "gini_truth": "Gini coefficient, C1  :",
              "gini_found": "Gini coefficient, C2  :"}
TRUTH = GOLDEN["ground_truth"]


def _truth_vector(kind, nv):
    if kind == "factions":
        return graphs.karate_factions()
    v = np.arange(nv, dtype=np.int64)
    return v // 37 if kind == "blocks" else 3 * (v // 50) + 1


@pytest.mark.parametrize("driver", sorted(DRIVERS))
@pytest.mark.parametrize("idx", range(len(TRUTH)),
                         ids=[f"{c['graph']}:{c['truth']}:{' '.join(c['flags']) or 'default'}" for c in TRUTH])
def test_cli_ground_truth_report_matches_reference(tmp_path, idx, driver):
    """-g FILE [-z]: the comparison report (compare.cpp:274-293) equals the one the compiled
    reference printed for the same graph, truth file and flags; counts exact, scores to
    the last bit."""
    exe = _driver(driver)
    case = TRUTH[idx]
    g = _build(GOLDEN["graphs"][case["graph"]]["spec"])
    path = str(tmp_path / (case["graph"] + ".bin"))
    po.write_bin(path, g)
    truth = _truth_vector(case["truth"], g.nv)
    tpath = str(tmp_path / "truth.txt")
    one = case["one_based"]
    with open(tpath, "w") as f:
        for v, c in enumerate(truth.tolist()):
            f.write(f"{v + 1 if one else v} {c + 1 if one else c}\n")
    env = dict(os.environ, VITE_AMD_SEQUENTIAL_SUMS="1")
    out = subprocess.run([exe, "-f", path, "-g", tpath] + (["-z"] if one else []) + case["flags"],
                         capture_output=True, text=True, env=env, timeout=120, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.splitlines()
    for key, label in STAT_LINES.items():
        val = next(x for x in lines if x.startswith(label))[len(label):].split()[0]
        want = case["stats"][key]
        if isinstance(want, int):
            assert int(val) == want, key
        elif driver == "dropin":
            # the reference's driver leaves std::cout at its default six significant digits
            assert val == format(float.fromhex(want), ".6g"), key
        else:
            assert float(val).hex() == want, key
