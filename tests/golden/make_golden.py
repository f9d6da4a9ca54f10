"""Writes tests/golden/golden.json: the reference's own outputs on the fixture
graphs and on the generated benchmark graphs.

Runs in the build container only: it executes oracle/_ref/ref_driver, the
harness compiled from the reference's headers (oracle/Makefile), which exists
only where the reference tree is present.  What it records:

* triangle counts (mxmm_msa and mxmm_mca paths) of the four graphs in
  tests/golden/graphs and of the config-1 Erdos-Renyi graph;
* per-k k-truss sizes and iteration counts (the reference's k_truss from
  scratch for every k) and a SHA-256 + histogram of the per-entry truss
  numbers, with the full int32 arrays saved as .npy for the small graphs;
* the reference's float32 batched BC (sources 0..batch-1) for the small
  graphs, saved as .npy;
* SHA-256 of the generated CSR arrays, so a GPU host can check it
  regenerated the same inputs.

Usage: python tests/golden/make_golden.py [--skip-rmat]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from graphalgo_amd import graphs as G  # noqa: E402
import oracle as O  # noqa: E402

GRAPHS = HERE / "graphs"
SMALL = ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz", "cage3.mtx"]
ER_CFG = dict(n=1000, m=9990, seed=1)
RMAT16_CFG = dict(scale=16, edge_factor=16, seed=1)
RMAT18_CFG = dict(scale=18, edge_factor=16, seed=1)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_graph_file(g: G.CSR, tmp: Path, name: str) -> str:
    p = tmp / f"{name}.csr"
    g.write_csr(str(p))
    return str(p)


def truss_record(g: G.CSR, path: str, save_as: str | None) -> dict:
    truss, info = O.ref_truss(path, g.nnz)
    rec = {
        "per_k": info["per_k"],           # [k, nnz of the k-truss, iterations]
        "truss_sha256": sha(truss.astype(np.int32)),
        "truss_hist": np.bincount(truss).tolist(),
        "kmax": int(truss.max()) if len(truss) else 2,
        "ref_ms_all_k": info["ms"],
    }
    if save_as:
        np.save(HERE / save_as, truss.astype(np.int32))
        rec["truss_npy"] = save_as
    return rec


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-rmat", action="store_true")
    args = ap.parse_args()
    if not O.ref_available():
        O.build()
    assert O.ref_available(), "oracle/_ref/ref_driver not built (reference tree absent?)"
    out: dict = {"graphs": {}, "generated": {}}
    with tempfile.TemporaryDirectory() as td:
        tmp = Path(td)
        for name in SMALL:
            g = G.read_mtx(str(GRAPHS / name))
            gs = G.symm_from_lower(g)
            stem = name.split(".")[0]
            src = str(GRAPHS / name)
            if name.endswith(".gz"):   # the reference reads plain text only
                src = ref_graph_file(g, tmp, stem)
            rec = {
                "n": g.n, "nnz_read": g.nnz, "nnz_symm": gs.nnz,
                "csr_sha256": gs.sha256(),
                "tc_msa": O.ref_run("tc", src, "msa")["triangles"],
                "tc_mca": O.ref_run("tc", src, "mca")["triangles"],
            }
            rec["truss"] = truss_record(gs, ref_graph_file(gs, tmp, stem + "_symm"), f"truss_{stem}.npy")
            if name != "cage3.mtx":
                batch = 16
                bc = O.ref_bc(src, batch, g.n)
                np.save(HERE / f"bc_ref_f32_{stem}.npy", bc)
                rec["bc_ref_f32"] = {"batch": batch, "npy": f"bc_ref_f32_{stem}.npy"}
            out["graphs"][name] = rec
            print(name, rec["tc_msa"], rec["truss"]["kmax"], flush=True)

        er = G.erdos_renyi(ER_CFG["n"], ER_CFG["m"], ER_CFG["seed"])
        p = ref_graph_file(er, tmp, "er")
        out["generated"]["er_1000_9990_s1"] = {
            "cfg": ER_CFG, "n": er.n, "nnz": er.nnz, "csr_sha256": er.sha256(),
            "tc": O.ref_run("tc", p, "msa")["triangles"],
            "wedges_LL": er.lower().wedges(),
        }
        print("er", out["generated"]["er_1000_9990_s1"]["tc"], flush=True)

        if not args.skip_rmat:
            r16 = G.rmat(**RMAT16_CFG)
            p = ref_graph_file(r16, tmp, "rmat16")
            rec = {"cfg": RMAT16_CFG, "n": r16.n, "nnz": r16.nnz, "csr_sha256": r16.sha256(),
                   "tc": O.ref_run("tc", p, "msa")["triangles"], "wedges_AA": r16.wedges()}
            rec["truss"] = truss_record(r16, p, None)
            out["generated"]["rmat16_ef16_s1"] = rec
            print("rmat16", rec["tc"], rec["truss"]["kmax"], rec["truss"]["ref_ms_all_k"], flush=True)
            r18 = G.rmat(**RMAT18_CFG)
            p = ref_graph_file(r18, tmp, "rmat18")
            out["generated"]["rmat18_ef16_s1"] = {
                "cfg": RMAT18_CFG, "n": r18.n, "nnz": r18.nnz, "csr_sha256": r18.sha256(),
                "tc": O.ref_run("tc", p, "msa")["triangles"],
            }
            print("rmat18", out["generated"]["rmat18_ef16_s1"]["tc"], flush=True)
    (HERE / "golden.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "golden.json")


if __name__ == "__main__":
    main()
