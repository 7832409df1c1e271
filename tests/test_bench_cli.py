"""bench.py's command line on the CPU: more than one GPU is refused with a
message and exit status 2, before anything touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=60,
                          env=dict(os.environ, **(env or {})))


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_distributed_launch_is_refused():
    r = _run([], env={"WORLD_SIZE": "4"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _stub_result():
    """A full result dict shaped like main()'s, with long strings and a big
    autotune table, as a real run produces."""
    sec = {"value": 1.2345678e9, "ms_per_step": 0.123456, "p50_step_ms": 0.12,
           "p99_step_ms": 0.13, "misses": 0, "violations": 0,
           "roofline": {"atomic_frac": 0.5}, "workload": "x" * 300}
    row = {"pairs_per_s": 2.9e9, "ms_per_step": 0.8, "p50_step_ms": 0.8, "p99_step_ms": 0.8,
           "misses": 0, "violations": 0, "pairs_per_step": 2269305.0, "path": "batched",
           "form": "workgroup", "segment_bytes": 67108864, "smallest": 128, "biggest": 4096,
           "threads": 1024, "in_grid": True}
    best = {k: row[k] for k in ("form", "segment_bytes", "smallest", "biggest", "threads")}
    cpu = {"value": 1.1e8, "unit": "pairs/s", "cores": 16, "online_cores": 192,
           "kind": "glibc malloc/free", "sample": "y" * 200}
    return {
        "metric": "m" * 90, "value": 4.0e9, "unit": "pairs/s", "n_gpus": 1, "steps": 20,
        "warmup": 5, "ms_per_step": 22.8, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic " * 8,
        "config": {"workload": "w" * 140, "path": "batched", "allocator": "Gallatin<>",
                   "pool_bytes": 8 << 30, "threads": 1 << 24, "rounds": 10, "slots0": 128,
                   "slots_min": 32, "xcd_partitions": 8, "flags": 0},
        "p50_step_ms": 22.8, "p99_step_ms": 23.0, "pairs_per_step": 9.2e7, "misses": 0,
        "violations": 0, "batched_over_plain": 1.07,
        "roofline": {"bound": "hbm", "achieved": 4013.2, "peak": 8000.0, "unit": "GB/s",
                     "frac": 0.5, "traffic": 9.7e10, "kernel": "k_churn_batched<1024>",
                     "algorithmic_bytes_per_launch": 9.1e10},
        "plain": dict(sec, kernel="k_churn"),
        "atomic_roofline": {"achieved": 8e9, "peak": 2.6e10, "unit": "atomics/s",
                            "frac": 0.31, "per_pair": 2, "ceilings": {"a": 1.0, "b": 2.0}},
        "hbm_ceilings": {"copy_gbs": 6000.0, "scatter_256_gbs": 4800.0,
                         "scatter_2048_gbs": 4880.0, "traffic_gbs": 4280.0,
                         "traffic_frac_of_copy": 0.72, "traffic_frac_of_scatter_2048": 0.88},
        "tree_search": {"workload": "t" * 100, "searches_per_s": 3e9},
        "plain_small": dict(sec, coherent_payload=dict(sec), partitioned_lazy_boot=dict(sec)),
        "churn_one_word": {"workload": "w" * 150, "value": 1.3e10,
                           "batched_over_plain": 1.6,
                           "batched": dict(sec, kernel="k" * 30),
                           "plain": dict(sec, kernel="k" * 12),
                           "roofline": {"atomic_frac": 0.99}},
        "block_multiseg": dict(sec, pairs_per_step=1300.0, misses_per_step=2800.0,
                               fresh_after_drain=True, steps=20),
        "batched_zipf": dict(sec, best=best, plain_best=2.6e9, batched_over_plain=1.1,
                             outside_grid=dict(best, value=3.0e9, ms_per_step=0.75),
                             per_batch_launch={f: dict(pairs_per_s=4.0e8, ms_per_step=5.7,
                                                       misses=0, violations=0)
                                               for f in ("plain", "batched")},
                             persistent_batched_over_per_batch_launch_plain=7.7,
                             table=[dict(row) for _ in range(60)], grid="g" * 100),
        "serving": {"p50_step_ms": 2.1, "p99_step_ms": 2.5, "mean_step_ms": 2.2,
                    "table_ops_per_s": 9.7e8, "words_per_s": 1.4e10,
                    "allocator_ops_per_s": 3.4e7, "live_sessions": 197000,
                    "session_words_bad": 0, "table_keys": 161000000, "refused_sessions": 1234,
                    "words_refused": 0, "chunk_failures": 0, "allocator_errors": 0,
                    "gb_live": 170.0, "counters": {str(i): i for i in range(16)},
                    "workload": "s" * 300, "roofline": {"hbm_frac": 0.1}},
        "cpu_baseline": {"value": 2.2e6, "unit": "pairs/s", "cores": 1, "kind": "port",
                         "sample": "z" * 160},
        "cpu_glibc": cpu, "cpu_glibc_plain_small": cpu, "cpu_glibc_batched_zipf": cpu,
        "cpu_glibc_block_multiseg": cpu,
        "cpu_glibc_serving": dict(cpu, unit="ms per step (p50)", p99_step_ms=60.0,
                                  live_sessions=197000),
    }


def test_headline_line_is_compact_and_last(capsys):
    import importlib.util
    import json
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    out = _stub_result()
    bench.emit(out)
    lines = capsys.readouterr().out.strip().splitlines()
    last = lines[-1]
    assert len(last.encode()) < 4096, len(last)
    head = json.loads(last)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline", "plain", "batched_over_plain", "workloads"):
        assert k in head, k
    assert head["config"]["workload"] == out["config"]["workload"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel"):
        assert k in head["roofline"], k
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in head["cpu_baseline"], k
    assert set(head["workloads"]) == {"plain_small", "churn_one_word", "block_multiseg",
                                      "batched_zipf", "serving"}
    # every earlier line is a section of its own, also valid JSON
    assert all("section" in json.loads(x) for x in lines[:-1])
