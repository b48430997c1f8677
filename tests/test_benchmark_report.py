"""CPU checks of the benchmark-setup reader and the schema-6 report writer
(ultrahigh_ann_amd.benchmark_report): the reference's setup rules
(src/benchmarks/benchmark_setup.cpp), its NPY dataset checks, and the report
layout -- every key path and value kind of the reference's own checked-in
schema-6 reports (tests/golden/schema6_report_keys.json, written from them by
tests/golden/make_schema6_keys.py). The end-to-end run on the GPU is in
tests/test_gpu_benchmark_report.py."""

import json
import struct
from pathlib import Path

import numpy as np
import pytest

import ultrahigh_ann_amd as uh
from ultrahigh_ann_amd import benchmark_report as br
from ultrahigh_ann_amd import diagnostics

GOLDEN = Path(__file__).parent / "golden" / "schema6_report_keys.json"
HEADER = "ultrahigh_ann_benchmark_setup_v2"


def _write_setup(tmp_path, body, name="setup.tsv"):
    p = tmp_path / name
    p.write_text("# comment\n" + HEADER + "\n" + body)
    return p


BASE = "dataset\tdata\njson_output\tout/r.json\ncsv_output\tout/r.csv\nreference\texact\n"


def test_setup_defaults_and_paths(tmp_path):
    p = _write_setup(tmp_path, BASE + "batch_sizes\t1,8\ntrials\t3\n"
                     "run\texact\texact\tbackend=cuda\tstrategy=direct\n"
                     "run\tflat.t8-s7\tflat\tbackend=cuda\trepetitions=8\tseed=7\ttrials=1\n"
                     "run\th\thierarchical\tbackend=cuda\trepetitions=4\tseed=2\tprojection_dimension=16\n")
    s = br.parse_setup(p)
    assert s.distance == "l2" and s.probability_policy == "sequential" and s.diagnostics == "selected_distances"
    assert s.dataset_directory == tmp_path / "data" and s.json_output == tmp_path / "out" / "r.json"
    assert s.reference_vectors_path == tmp_path / "data" / "reference_vectors.npy"
    assert [(r.name, r.strategy, r.batch_sizes, r.trials) for r in s.runs] == [
        ("exact", "direct", [1, 8], 3), ("flat.t8-s7", "gemm", [1, 8], 1), ("h", "gemm", [1, 8], 3)]
    assert s.runs[2].projection_dimension == 16
    l1 = br.parse_setup(_write_setup(tmp_path, BASE + "distance\tl1\nrun\texact\texact\tbackend=cuda\n", "l1.tsv"))
    assert l1.runs[0].strategy == "direct"  # the CUDA default for L1
    loaded = br.parse_setup(_write_setup(tmp_path, BASE + "probabilities\tp.uap\nrun\texact\texact\tbackend=cuda\n",
                                         "load.tsv"))
    assert loaded.probability_policy == "load" and loaded.probabilities_path == tmp_path / "p.uap"


@pytest.mark.parametrize("body,message", [
    ("run\texact\texact\tbackend=cuda\n", "does not define dataset"),
    (BASE + "colour\tred\nrun\texact\texact\n", "unknown directive"),
    (BASE + "run\texact\texact\nrun\texact\texact\n", "duplicate run name"),
    (BASE + "run\texact\texact\tbackend=cpu\tstrategy=direct\n", "CPU runs cannot use direct or gemm"),
    (BASE + "distance\tl1\nrun\texact\texact\tbackend=cuda\tstrategy=gemm\n", "must be direct"),
    (BASE + "run\texact\tflat\tbackend=cuda\trepetitions=4\tseed=1\n", "must use the exact index"),
    (BASE + "run\texact\texact\tbackend=cuda\trepetitions=4\n", "exact runs do not accept"),
    (BASE + "run\texact\texact\nrun\tf\tflat\tbackend=cuda\tseed=1\n", "require repetitions and seed"),
    (BASE + "run\texact\texact\nrun\th\thierarchical\tbackend=cuda\trepetitions=2\tseed=1\n", "projection_dimension"),
    (BASE + "batch_sizes\t4,4\nrun\texact\texact\n", "unique"),
    (BASE + "trials\t0\nrun\texact\texact\n", "positive"),
    (BASE + "run\tbad name!\texact\n", "run name may contain"),
    (BASE + "probability_policy\tload\nrun\texact\texact\n", "requires probabilities"),
    (BASE + "run\texact\texact\tcolour=red\n", "unsupported run parameter"),
])
def test_setup_errors(tmp_path, body, message):
    with pytest.raises(br.SetupError, match=message):
        br.parse_setup(_write_setup(tmp_path, body) if body.startswith(("dataset", "run")) else
                       _write_setup(tmp_path, body))


def test_setup_header_required(tmp_path):
    p = tmp_path / "x.tsv"
    p.write_text("ultrahigh_ann_benchmark_setup_v1\n" + BASE)
    with pytest.raises(br.SetupError, match="expected ultrahigh_ann_benchmark_setup_v2"):
        br.parse_setup(p)


def _npy_bytes(arr, fortran=False):
    import io

    buf = io.BytesIO()
    np.lib.format.write_array(buf, np.asfortranarray(arr) if fortran else arr)
    return buf.getvalue()


def test_npy_checks(tmp_path):
    good = np.arange(12, dtype=np.float32).reshape(3, 4)
    p = tmp_path / "m.npy"
    p.write_bytes(_npy_bytes(good))
    np.testing.assert_array_equal(br.load_npy(p, "<f4", 2), good)
    cases = {
        "expected dtype <f4": _npy_bytes(good.astype(np.float64)),
        "Fortran-order": _npy_bytes(good, fortran=True),
        "two-dimensional": _npy_bytes(good.reshape(-1)),
        "trailing data": _npy_bytes(good) + b"\0",
        "truncated": _npy_bytes(good)[:-3],
    }
    for message, data in cases.items():
        p.write_bytes(data)
        with pytest.raises(uh.UhannInvalidArgument, match=message):
            br.load_npy(p, "<f4", 2)
    labels = tmp_path / "l.npy"
    labels.write_bytes(_npy_bytes(np.array([1, 2, 3], np.uint16)))
    assert br.load_npy(labels, "<u2", 1).tolist() == [1, 2, 3]


def _paths(v, prefix, out):
    kind = ("null" if v is None else "boolean" if isinstance(v, bool) else "number" if isinstance(v, (int, float))
            else "string" if isinstance(v, str) else "array" if isinstance(v, list) else "object")
    out.setdefault(prefix, set()).add(kind)
    if isinstance(v, dict):
        for k, x in v.items():
            _paths(x, f"{prefix}.{k}" if prefix else k, out)
    elif isinstance(v, list):
        for x in v:
            _paths(x, prefix + "[]", out)
    return out


# Fields the reference's writer leaves null when labels are absent, when no
# query is ratio-eligible, for a reference run, or when Git or host details
# are unavailable; and the probability GPU block, null in the checked-in
# reports (loaded probabilities).
_MAY_BE_NULL = {"dataset.reference_labels_file", "dataset.query_labels_file", "dataset.reference_labels_sha256",
                "dataset.query_labels_sha256", "runs[].measurements[].correct", "runs[].measurements[].accuracy",
                "runs[].measurements[].label_agreement_with_exact", "provenance.source.git.commit",
                "provenance.source.git.dirty", "provenance.host.processor_model",
                "provenance.host.physical_memory_bytes"}


def check_layout(report: dict) -> None:
    """Every key path of the report exists with a value kind the reference's
    reports use, and every path of theirs exists here (below non-null
    nodes)."""
    golden = {k: set(v) for k, v in json.loads(GOLDEN.read_text())["paths"].items()}
    mine = _paths(report, "", {})
    mine.pop("")
    for path, kinds in mine.items():
        if path == "sampling_probabilities.gpu" or path.startswith("sampling_probabilities.gpu."):
            continue  # an object when the probabilities were computed (the reference's GPU policies)
        assert path in golden, path
        extra = kinds - golden[path] - ({"null"} if path in _MAY_BE_NULL or path.endswith(
            (".mean", ".median", ".percentile_95", ".percentile_99", ".maximum", "_mean", ".mean_distance_ratio"))
            else set())
        assert not extra, (path, kinds, golden[path])
    null_nodes = [p for p, k in mine.items() if k == {"null"}]
    empty_arrays = [p for p, k in mine.items() if "array" in k and p + "[]" not in mine]
    for path in golden:
        if path in mine or any(path.startswith(n + ".") or path.startswith(n + "[]") for n in null_nodes):
            continue
        if any(path.startswith(a + "[]") for a in empty_arrays):
            continue
        raise AssertionError(f"missing {path}")


def test_report_layout_matches_reference_reports(tmp_path):
    """A report built from recorded executions (no GPU) has the reference's
    layout, for both diagnostic modes."""
    p = _write_setup(tmp_path, BASE + "diagnostics\tfull_distance_table\n"
                     "run\texact\texact\tbackend=cuda\tstrategy=direct\n"
                     "run\tf\tflat\tbackend=cuda\trepetitions=8\tseed=7\n")
    s = br.parse_setup(p)
    g = np.random.default_rng(3)
    r, d, b = 12, 5, 9
    refs = g.random((r, d), dtype=np.float32)
    q = g.random((b, d), dtype=np.float32)
    labels = np.arange(r) % 3
    data = br.Dataset(refs, q, labels, g.integers(0, 3, b), True, b"\1" * 32, b"\2" * 32, b"\3" * 32, b"\4" * 32,
                      1.0, 2.0)
    table = np.sqrt(((q.astype(np.float64)[:, None, :] - refs.astype(np.float64)[None]) ** 2).sum(-1))
    exact = table.argmin(1)
    approx = np.where(g.random(b) < 0.5, exact, (exact + 1) % r)
    ex = [br.Execution(s.runs[0], 3.0, {"index_payload_bytes": 10, "query_workspace_payload_bytes": 4,
                                        "unique_coordinates": 0, "sampled_multiplicity": 0},
                       [br.Measurement(4, [1.0, 1.5], exact, 4)], 0, "dev"),
          br.Execution(s.runs[1], 3.0, {"index_payload_bytes": 10, "query_workspace_payload_bytes": 4,
                                        "unique_coordinates": 3, "sampled_multiplicity": 8},
                       [br.Measurement(4, [1.0], approx, 4), br.Measurement(9, [0.5], approx, 4)], 0, "dev")]
    br.add_quality(data, b, exact, ex)
    t = diagnostics.ExactDistanceTable.from_table(table)
    for m in ex[1].measurements:
        m.approximation = t.evaluate(m.predictions)
    diag = {"build_ms": 1.0, "evaluation_ms": 1.0, "payload_bytes": t.payload_bytes(), "geometry": t.query_geometry()}
    probs = {"probabilities": np.full(d, 0.5), "mass": 2.5, "build_ms": 1.0, "load_ms": 0.0, "source": None,
             "source_sha256": None, "refs_sha256": b"\1" * 32, "gpu": None}
    report = br.build_report(uh, s, data, b, probs, diag, ex)
    check_layout(report)
    assert report["schema_version"] == 6 and report["runs"][0]["reference"] is True
    m = report["runs"][1]["measurements"][0]
    assert m["exact_neighbor_agreement_count"] == int((approx == exact).sum())
    assert m["correct"] == int((labels[approx] == data.query_labels).sum())
    assert m["median_ms"] == 1.0 and report["runs"][0]["measurements"][0]["median_ms"] == 1.25
    # selected-distance mode: no geometry, no rank, no buckets
    for mm in ex[1].measurements:
        mm.approximation = diagnostics.evaluate_selected_distances(t, exact, mm.predictions)
    diag = {"build_ms": 0.0, "evaluation_ms": 1.0, "payload_bytes": 0, "geometry": None}
    report = br.build_report(uh, s, data, b, probs, diag, ex)
    check_layout(report)
    rows = br.csv_text(report).splitlines()
    assert rows[0].startswith("run_name,index,backend") and len(rows) == 1 + 2 + 1 + 1
    assert all(len(line.split(",")) >= 40 for line in rows[1:])
