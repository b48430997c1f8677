"""Writes schema6_report_keys.json: the key paths (lists collapsed to "[]")
and JSON value kinds of the reference's checked-in schema-6 benchmark reports
(assets/*.json of the reference checkout), the layout a report written by
ultrahigh_ann_amd.benchmark_report must reproduce. Data only: no code of the
reference is copied or run.

    python3 tests/golden/make_schema6_keys.py /root/reference
"""

import json
import sys
from pathlib import Path


def kind(v):
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "boolean"
    if isinstance(v, (int, float)):
        return "number"
    if isinstance(v, str):
        return "string"
    if isinstance(v, list):
        return "array"
    return "object"


def walk(v, prefix, out):
    out.setdefault(prefix, set()).add(kind(v))
    if isinstance(v, dict):
        for k, x in v.items():
            walk(x, f"{prefix}.{k}" if prefix else k, out)
    elif isinstance(v, list):
        for x in v:
            walk(x, prefix + "[]", out)


def main():
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = {}
    names = sorted(p.name for p in (root / "assets").glob("*.json"))
    for name in names:
        walk(json.loads((root / "assets" / name).read_text()), "", out)
    keys = {k: sorted(v) for k, v in sorted(out.items())
            if k and not k.startswith("provenance.report_format_update")}
    dest = Path(__file__).with_name("schema6_report_keys.json")
    dest.write_text(json.dumps({"source_reports": names, "paths": keys}, indent=1) + "\n")
    print(f"{len(keys)} key paths from {len(names)} reports -> {dest}")


if __name__ == "__main__":
    main()
