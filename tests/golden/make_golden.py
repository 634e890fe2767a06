"""Writes tests/golden/reference_runs.json from the compiled reference (oracle/_ref).

Run in the build container after `make ref` (needs /root/reference sources only to build
oracle/_ref; this script runs the built binaries):

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py

The reference ships no tests or fixtures, so these vectors are produced here from the
reference's own code at one MPI rank and one OpenMP thread:
  * runs:    the reference main() (main.cpp:108-607) on small seeded graphs written to Vite
             .bin files, every distLouvainMethod* call's return value (17 significant
             digits, which round-trips a double), iteration count and partition hash, and
             every rebuild's CSR hash;
  * sweeps:  distExecuteLouvainIteration over all vertices from seeded states
             (targets / clusterWeight / localCupdate hashes);
  * colors:  distColoringMultiHashMinMax colorings;
  * vite_rgg: generateInMemGraph graphs (Vite's LCG RGG) as CSR hashes;
  * communities: the per-vertex community file main() writes with -o (main.cpp:380-425,
             543-572), as a hash of the int64 array, with its community count and the
             run's final Q;
  * ground_truth: main() with -g FILE [-z] (compare.cpp:58-337): the run's community file
             and the comparison statistics it printed (pair counts, precision, recall,
             F-score, Gini coefficients; floats as hex).
Graph inputs are recorded as generator specs plus a hash of the generated CSR, so the
fixture stays a few kilobytes; the hash pins the generator output the vectors were made on.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from tests import graphs  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_runs.json")


def csr_sha(g) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.offs, np.int64).tobytes())
    h.update(np.ascontiguousarray(g.edges).tobytes())
    return h.hexdigest()


def arr_sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def trace_csr_fnv(g) -> str:
    """The rebuild hash ref_harness.cpp prints: FNV-1a over offsets then tails (whole words)."""
    h = 1469598103934665603
    mask = (1 << 64) - 1
    for x in g.offs.tolist() + g.edges["tail"].tolist():
        h ^= x & mask
        h = (h * 1099511628211) & mask
    return f"{h:016x}"


GRAPH_SPECS = {
    "rgg_small": {"kind": "rgg", "seed": 21, "n": 1500, "extra_pct": 1.0},
    "rgg_plain": {"kind": "rgg", "seed": 22, "n": 1200, "extra_pct": 0.0},
    "ba_small": {"kind": "ba", "seed": 23, "n": 1500, "m": 4},
    "vite_rgg_1024": {"kind": "vite_rgg", "n": 1024},
    "hub": {"kind": "named", "name": "hub"},
    "multi": {"kind": "named", "name": "multi"},
    "grid": {"kind": "named", "name": "grid"},
    "karate": {"kind": "named", "name": "karate"},
}


def build(spec):
    if spec["kind"] == "rgg":
        return po.gen_rgg(spec["seed"], spec["n"], spec["extra_pct"])
    if spec["kind"] == "ba":
        return po.gen_ba(spec["seed"], spec["n"], spec["m"])
    if spec["kind"] == "vite_rgg":
        return po.gen_vite_rgg(spec["n"])
    return graphs.graph(spec["name"])


RUN_FLAGS = [[], ["-i"], ["-c", "8", "-t", "2", "-a", "0.25"], ["-t", "1"], ["-d", "8", "-t", "3"],
             ["-c", "8 true", "-t", "4", "-a", "0.5"], ["-p"], ["-t", "2", "-a", "0.9"], ["-i", "-c", "4"]]
# flag sets whose community file (-o: <input>.communities, main.cpp:380-425,543-572) is recorded
OUTPUT_FLAGS = [[], ["-i"], ["-c", "8", "-t", "2", "-a", "0.25"], ["-p"]]

# ground-truth comparisons (-g FILE [-z], main.cpp:119-137,574-580; compare.cpp:58-337):
# (graph, truth kind, one-based file, run flags)
TRUTH_CASES = [("karate", "factions", False, []), ("karate", "factions", True, ["-i"]),
               ("rgg_small", "blocks", False, []), ("rgg_small", "sparse_blocks", True, ["-p"]),
               ("ba_small", "blocks", False, ["-i"]), ("grid", "sparse_blocks", False, [])]

STAT_LINES = {"same_same": "Same-Same (True positive)  :", "same_diff": "Same-Diff (False negative) :",
              "diff_same": "Diff-Same (False positive) :", "nc_truth": "#communities in C1 :",
              "nc_found": "#communities in C2 :", "precision": "Precision :", "recall": "Recall    :",
              "f_score": "F-score   :", "gini_truth": "Gini coefficient, C1  :",
              "gini_found": "Gini coefficient, C2  :"}


def truth_vector(kind, nv):
    """A ground truth: the karate club's factions, blocks of 37 consecutive vertices, or
    blocks of 50 with ids 3b + 1 (ids with no members in between)."""
    if kind == "factions":
        return graphs.karate_factions()
    v = np.arange(nv, dtype=np.int64)
    return v // 37 if kind == "blocks" else 3 * (v // 50) + 1


def parse_stats(text):
    out = {}
    for key, label in STAT_LINES.items():
        line = next(x for x in text.splitlines() if x.startswith(label))
        val = line[len(label):].split()[0]
        out[key] = int(val) if key in ("same_same", "same_diff", "diff_same", "nc_truth", "nc_found") else \
            float(val).hex()
    return out


def main():
    assert po.ref_available(), "build oracle/_ref first (make ref)"
    assert os.environ.get("OMP_NUM_THREADS") == "1", "run with OMP_NUM_THREADS=1"
    out = {"about": "Vectors from the compiled reference at 1 MPI rank, 1 OpenMP thread; see make_golden.py",
           "graphs": {}, "runs": [], "sweeps": [], "colors": [], "vite_rgg": [], "rebuilds": [],
           "communities": [], "ground_truth": []}
    tmp = tempfile.mkdtemp()
    for name, spec in GRAPH_SPECS.items():
        g = build(spec)
        out["graphs"][name] = {"spec": spec, "nv": g.nv, "ne": g.ne, "csr_sha256": csr_sha(g)}
        path = os.path.join(tmp, name + ".bin")
        po.write_bin(path, g)
        for flags in RUN_FLAGS:
            recs, _ = po.ref_trace_run(path, flags)
            calls = [{"kind": r["kind"], "nv": int(r["nv"]), "q": float(r["q"]), "q_hex": float(r["q"]).hex(),
                      "iters": int(r["iters"]), "cvect_fnv": r["cvect_fnv"]} for r in recs if r["kind"] != "rebuild"]
            rebuilds = [{"nv": int(r["nv"]), "ne": int(r["ne"]), "csr_fnv": r["csr_fnv"]}
                        for r in recs if r["kind"] == "rebuild"]
            out["runs"].append({"graph": name, "flags": flags, "calls": calls, "rebuilds": rebuilds})
        for flags in OUTPUT_FLAGS:
            recs, _ = po.ref_trace_run(path, ["-o"] + flags)
            comm = np.loadtxt(path + ".communities", dtype=np.int64, ndmin=1)
            os.remove(path + ".communities")
            out["communities"].append({"graph": name, "flags": flags, "nv": int(len(comm)),
                                       "count": int(len(np.unique(comm))), "sha256": arr_sha(comm),
                                       "final_q_hex": float([r for r in recs if r["kind"] != "rebuild"][-1]["q"]).hex()})
        # sweeps from seeded states
        for seed in [None, 7]:
            if seed is None:
                vdeg, c = po.ref_init_degrees(g)
                curr, csize, cdeg = np.arange(g.nv), np.ones(g.nv, np.int64), vdeg.copy()
            else:
                curr, csize, cdeg, vdeg, c = graphs.random_state(g, seed)
            t, cw, ds, dd = po.ref_sweep(g, curr, csize, cdeg, vdeg, c)
            out["sweeps"].append({"graph": name, "state_seed": seed, "constant_hex": float(c).hex(),
                                  "target_sha256": arr_sha(t.astype(np.int64)), "cw_sha256": arr_sha(cw),
                                  "dsize_sha256": arr_sha(ds.astype(np.int64)), "ddeg_sha256": arr_sha(dd)})
        for nhash, single in [(4, False), (2, True), (1, False)]:
            nc, col = po.ref_color(g, nhash, 70, single)
            out["colors"].append({"graph": name, "nhash": nhash, "single": single, "next_color": int(nc),
                                  "colors_sha256": arr_sha(col.astype(np.int64))})
        q, cv, it = po.ref_phase(g)
        r = po.ref_rebuild(g, cv)
        out["rebuilds"].append({"graph": name, "cvect_sha256": arr_sha(cv.astype(np.int64)), "nv": r.nv,
                                "ne": r.ne, "csr_sha256": csr_sha(r)})
    for name, kind, one_based, flags in TRUTH_CASES:
        g = build(GRAPH_SPECS[name])
        path = os.path.join(tmp, name + ".bin")
        truth = truth_vector(kind, g.nv)
        tpath = os.path.join(tmp, f"{name}.{kind}.truth")
        with open(tpath, "w") as f:
            for v, c in enumerate(truth.tolist()):
                f.write(f"{v + 1 if one_based else v} {c + 1 if one_based else c}\n")
        _, text = po.ref_trace_run(path, ["-o", "-g", tpath] + (["-z"] if one_based else []) + flags)
        comm = np.loadtxt(path + ".communities", dtype=np.int64, ndmin=1)
        os.remove(path + ".communities")
        out["ground_truth"].append({"graph": name, "truth": kind, "one_based": one_based, "flags": flags,
                                    "communities": " ".join(map(str, comm.tolist())), "stats": parse_stats(text)})
    for n in [256, 1000, 4096]:
        g = po.ref_gen_rgg(n)
        out["vite_rgg"].append({"n": n, "nv": g.nv, "ne": g.ne, "csr_sha256": csr_sha(g)})
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", OUT, len(out["runs"]), "runs")


if __name__ == "__main__":
    main()
