"""Writes tests/golden/benchmark_schema.json from the JSONL files the
reference publishes (test/baselines/*.jsonl, test/results/benchmark-1080p-
2.1.0/*.jsonl in the reference tree): data only, no reference code runs.

The fixture holds, per record kind (metadata, latency frame, latency
summary, throughput summary), every key set those files use (their schema-2
records come from two versions of the harness), the stage and CPU-slot names
of their profile maps, every case key they contain, and the case, median_ms,
fps, threads and stage_median_us of each summary (the published baselines a
``--baseline`` comparison reads).  Run in a container that holds the
reference tree:

    python3 tests/golden/make_benchmark_schema.py [/root/reference]
"""
import glob
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def main():
    files = sorted(glob.glob(os.path.join(REF, "test", "baselines", "*.jsonl")) +
                   glob.glob(os.path.join(REF, "test", "results", "*", "*.jsonl")))
    key_sets = {}
    stages, slots, cases, summaries = set(), set(), [], []
    for path in files:
        rel = os.path.relpath(path, REF)
        with open(path, encoding="utf-8") as f:
            for line in f:
                r = json.loads(line)
                kind = r["record_type"]
                if kind == "summary":
                    kind += "/" + r.get("mode", "latency")
                ks = sorted(r)
                if ks not in key_sets.setdefault(kind, []):
                    key_sets[kind].append(ks)
                for name in ("stage_us", "stage_median_us"):
                    stages.update(r.get(name, {}))
                for name in ("cpu_profile_us", "cpu_profile_median_us"):
                    slots.update(r.get(name, {}))
                if "case" in r and r["case"] not in cases:
                    cases.append(r["case"])
                if r["record_type"] == "summary":
                    summaries.append({"file": rel, "case": r["case"], "median_ms": r.get("median_ms"),
                                      "fps": r.get("fps"), "threads": r.get("threads"),
                                      "stage_median_us": r.get("stage_median_us", {})})
    out = {"files": [os.path.relpath(p, REF) for p in files], "key_sets": key_sets,
           "stages": sorted(stages), "cpu_slots": sorted(slots), "cases": cases, "summaries": summaries}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "benchmark_schema.json")
    with open(dst, "w", encoding="utf-8") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {dst}: {len(cases)} cases, {len(summaries)} summaries from {len(files)} files")


if __name__ == "__main__":
    main()
