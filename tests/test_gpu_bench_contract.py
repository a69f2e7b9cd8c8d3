"""bench.py's output contract on the GPU: one JSON line per run with the
metric, the roofline object (achieved / peak / frac over the launch-stream
kernel time) and the CPU baseline, at a reduced size so it runs in seconds."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
        "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config")


def _run(args, timeout=240):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                       capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    return r


def _json(r):
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert r.returncode == 0, r.stderr[-2000:]
    assert len(lines) == 1, r.stdout[-2000:]
    return json.loads(lines[0])


def test_uniform_line_has_roofline_and_cpu_baseline():
    j = _json(_run(["--n", "1024", "--systems", "2048", "--steps", "3", "--warmup", "1",
                    "--cpu-seconds", "0.2", "--served-requests", "4"]))
    for k in KEYS:
        assert k in j, k
    assert j["n_gpus"] == 1 and j["unit"] == "unknowns/s" and j["value"] > 0
    rf = j["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-9
    cb = j["cpu_baseline"]
    assert cb["kind"] in ("port", "reference") and cb["cores"] >= 1 and cb["value"] > 0
    assert cb["gpu_max_rel_err_vs_fp64_thomas"] <= 1e-4
    assert "FP64" in cb["sample"] and cb["t_precision"]["value"] > 0   # FP64 oracle + FP32
    assert "workload" in j["config"]
    sv = j["served"]                   # the served half of the metric
    assert sv["requests"] == 4 and sv["unknowns_per_s"] > 0
    assert sv["plain_per_size_unknowns_per_s"] > 0
    assert sv["latency_ms_p50"] > 0 and sv["latency_ms_p99"] >= sv["latency_ms_p50"]
    assert sv["max_rel_residual"] <= 1e-12
    sc = j["stream_ceiling"]           # plain stream of the same bytes
    assert sc["us"] > 0 and sc["bytes"] == 5 * 1024 * 2048 * 4



def test_adi_and_pivoting_lines():
    j = _json(_run(["--workload", "adi", "--steps", "2", "--warmup", "1",
                    "--cpu-seconds", "0.2"]))
    for k in KEYS:
        assert k in j, k
    assert j["dtype"] == "f64" and j["value"] > 0 and j["roofline"]["frac"] > 0
    assert j["cpu_baseline"]["gpu_max_rel_err_vs_oracle_step"] <= 1e-12
    by_grid = j["cpu_baseline"]["gpu_max_rel_err_vs_oracle_step_by_grid"]
    assert set(by_grid) == {"2048", "4096"} and max(by_grid.values()) <= 1e-12
    j = _json(_run(["--workload", "pivoting", "--n", "512", "--systems", "4096",
                    "--steps", "2", "--warmup", "1", "--cpu-seconds", "0.2"]))
    for k in KEYS:
        assert k in j, k
    cb = j["cpu_baseline"]
    assert cb["gpu_bitwise_equal_to_oracle"] and cb["gpu_info_equal_to_oracle"]


def test_gauss_sweep_line_reports_the_separator_form():
    j = _json(_run(["--workload", "gauss-sweep", "--steps", "5", "--warmup", "1",
                    "--cpu-seconds", "0.2"]))
    for k in KEYS:
        assert k in j, k
    assert [r["depth"] for r in j["sweep"] if r["in_sweep"]] == list(range(2, 9))
    for r in j["sweep"]:      # every depth against FP64 Thomas at full size
        assert r["max_rel_err_vs_fp64_thomas"] <= 1e-4 and r["reference_rms_failures"] == 0
    sf = j["separator_form"]
    assert sf["family"] == 3 and sf["ms"] > 0
    assert sf["max_rel_err_vs_fp64_thomas"] <= 1e-4 and sf["reference_rms_failures"] == 0
    assert j["cpu_baseline"]["gpu_max_rel_err_vs_fp64_thomas"] <= 1e-4
