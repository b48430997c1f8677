"""End-to-end run of a benchmark setup on the GPU (ultrahigh_ann_amd.benchmark_report):
NPY dataset in, schema-6 JSON + CSV out, with the reference's report layout
and figures that agree with the engine's own answers."""

import json

import numpy as np
import pytest

import ultrahigh_ann_amd as uh
from ultrahigh_ann_amd import benchmark_report as br
from test_benchmark_report import check_layout

pytestmark = pytest.mark.gpu


def _dataset(root, r=60, d=300, b=40, seed=5):
    g = np.random.default_rng(seed)
    classes = g.random((6, d), dtype=np.float32) * 3
    rl = (np.arange(r) % 6).astype(np.uint16)
    ql = g.integers(0, 6, b).astype(np.uint16)
    refs = (classes[rl] + g.random((r, d), dtype=np.float32)).astype(np.float32)
    q = (classes[ql] + g.random((b, d), dtype=np.float32)).astype(np.float32)
    root.mkdir(parents=True, exist_ok=True)
    np.save(root / "reference_vectors.npy", refs)
    np.save(root / "queries.npy", q)
    np.save(root / "reference_labels.npy", rl)
    np.save(root / "query_labels.npy", ql)
    return refs, q, rl, ql


def _setup(tmp_path, distance, runs, extra=""):
    p = tmp_path / f"{distance}.tsv"
    p.write_text("ultrahigh_ann_benchmark_setup_v2\n"
                 f"dataset\tdata\njson_output\tout/{distance}.json\ncsv_output\tout/{distance}.csv\n"
                 f"reference\texact_direct\ndistance\t{distance}\nbatch_sizes\t8,64\nwarmups\t1\ntrials\t2\n"
                 + extra + runs)
    return p


@pytest.mark.parametrize("diagnostics_mode", ["full_distance_table", "selected_distances"])
def test_l2_setup_end_to_end(tmp_path, oracle_mod, diagnostics_mode):
    refs, q, rl, ql = _dataset(tmp_path / "data")
    p = _setup(tmp_path, "l2",
               "run\texact_direct\texact\tbackend=cuda\tstrategy=direct\n"
               "run\texact_gemm\texact\tbackend=cuda\tstrategy=gemm\n"
               "run\tflat_t16\tflat\tbackend=cuda\tstrategy=direct\trepetitions=16\tseed=7\n"
               "run\tflat_t16_gemm\tflat\tbackend=cuda\tstrategy=gemm\trepetitions=16\tseed=7\n"
               "run\tuniform_t16\tuniform\tbackend=cuda\tstrategy=direct\trepetitions=16\tseed=7\n"
               "run\thier_p16\thierarchical\tbackend=cuda\tstrategy=direct\trepetitions=16\tseed=7\t"
               "projection_dimension=16\n"
               "run\thier_p16_gemm\thierarchical\tbackend=cuda\trepetitions=16\tseed=7\tprojection_dimension=16\n",
               f"diagnostics\t{diagnostics_mode}\n")
    report = br.run_setup(p, echo=lambda *_: None)
    check_layout(report)
    on_disk = json.loads((tmp_path / "out" / "l2.json").read_text())
    assert on_disk["runs"] == json.loads(json.dumps(report["runs"]))
    runs = {r["name"]: r for r in report["runs"]}
    assert [r["name"] for r in report["runs"]][0] == "exact_direct"
    assert [m["batch_size"] for m in runs["exact_direct"]["measurements"]] == [8, 40]  # 64 clamped to 40 queries
    exp, _ = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    # the exact GPU paths agree with the reference's binary32 argmin and with each other
    assert runs["exact_gemm"]["measurements"][0]["exact_neighbor_agreement"] == 1.0
    exact = uh.HipExactL2Index(refs)
    idx, _ = exact.query_batch(q, 40, exact.make_query_workspace(40))
    np.testing.assert_array_equal(idx, exp)
    # the flat direct run is the engine's flat index for the same probabilities and seed
    probs = uh.importance_probabilities(uh.METRIC_L2, refs)
    flat = uh.HipFlatL2AnnIndex(refs, probs, 16, uh.Rng(7))
    fidx, _ = flat.query_batch(q, 40, flat.make_query_workspace(40))
    m = runs["flat_t16"]["measurements"][1]
    assert m["exact_neighbor_agreement_count"] == int((fidx == exp).sum())
    assert m["correct"] == int((rl[fidx] == ql).sum())
    assert report["sampling_probabilities"]["sampling_mass"] == uh.sampling_mass(probs)
    a = m["approximation"]
    assert a["distance_optimal_count"] + a["non_optimal_count"] == 40
    if diagnostics_mode == "full_distance_table":
        assert report["diagnostic_execution"]["query_geometry"]["query_count"] == 40
        assert a["returned_neighbor_rank"]["count"] == 40
    csv_rows = (tmp_path / "out" / "l2.csv").read_text().splitlines()
    assert len(csv_rows) == 1 + sum(len(mm["trial_ms"]) for r in report["runs"] for mm in r["measurements"])


def test_l1_setup_and_probability_file(tmp_path):
    refs, q, _, _ = _dataset(tmp_path / "data", seed=9)
    digest = uh.sha256_file(tmp_path / "data" / "reference_vectors.npy")
    probs = uh.importance_probabilities(uh.METRIC_L1, refs)
    uh.save_importance_probabilities(tmp_path / "p.uap", uh.METRIC_L1, refs.shape[0], digest, probs)
    runs = ("run\texact_direct\texact\tbackend=cuda\n"
            "run\tflat_t8\tflat\tbackend=cuda\trepetitions=8\tseed=20\n"
            "run\thier_p16\thierarchical\tbackend=cuda\tstrategy=gemm\trepetitions=8\tseed=20\tprojection_dimension=16\n")
    p = _setup(tmp_path, "l1", runs, "probabilities\tp.uap\n")
    report = br.run_setup(p, echo=lambda *_: None)
    check_layout(report)
    sp = report["sampling_probabilities"]
    assert sp["policy"] == "load" and sp["source_sha256"] == uh.sha256_file(tmp_path / "p.uap").hex()
    assert sp["reference_vectors_sha256"] == digest.hex()
    # a probability file bound to other data is refused
    uh.save_importance_probabilities(tmp_path / "p.uap", uh.METRIC_L1, refs.shape[0], bytes(32), probs)
    with pytest.raises(uh.UhannError, match="does not match the configured dataset"):
        br.run_setup(p, echo=lambda *_: None)
    # CPU runs belong to the reference's CPU path
    p2 = _setup(tmp_path, "l1", "run\texact_direct\texact\tbackend=cpu\n")
    with pytest.raises(uh.UhannError, match="backend=cpu"):
        br.run_setup(p2, echo=lambda *_: None)
