"""bench.py's output contract (no GPU): the whole result goes to the detail
file and the last stdout line is one JSON object of at most 2 KB with the
headline metric, its roofline, the CPU baseline and the time-to-RMSE
figures, however large the tuner tables and curves of the legs are."""
import importlib.util
import io
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _tuner_table(n):
    return [{"unroll": 4, "waves_per_eu": 4 + k % 3, "scalar_bounds": True, "groups": 4, "pixel_queue": False,
             "wave_prio": True, "ms": 30.0 + k * 0.123456789,
             "options": "-ffp-contract=off -DHT_STRICT_FP=1 -DHT_CULL_UNROLL=4 -DHT_WAVES_PER_EU=4"} for k in range(n)]


def stub_result(bench):
    """A full result as main() assembles it, with every leg and tables as
    large as a real run's."""
    tuning = {"best": _tuner_table(1)[0], "table": _tuner_table(21)}
    curve = [(16 * k, 0.1 / (k + 1)) for k in range(1, 40)]
    return {
        "metric": bench.METRIC, "value": 639.123456789, "unit": "Mpaths/s", "n_gpus": 1, "steps": 20, "warmup": 5,
        "ms_per_step": 30.7654321, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "hyperbolic_cells", "width": 640, "height": 480, "spp": 64, "hops": 8,
                   "primitives": 192, "kernel": "render_passes", "maps": "x" * 300},
        "roofline": {"bound": "valu", "achieved": 11.9, "peak": 157.3, "unit": "TFLOP/s", "frac": 0.0757,
                     "frac_nonpacked": 0.151, "kernel": "render_passes", "kernel_ms": 30.04,
                     "flops_per_path": 18234.7, "fp32_flops_per_path": 4745.0, "fp64_flops_per_path": 13490.0,
                     "flops_sample_paths": 19200, "flops_note": "n" * 400, "plain_flops_per_path": 225327.0,
                     "plain_equivalent_tflops": 147.0, "traffic": 13221888, "algorithmic_bytes": 12306560},
        "render_tuning": tuning,
        "plain": {"kernel": "k", "value": 48.7, "unit": "Mpaths/s", "ms_per_frame": 403.0},
        "dropin_per_pass": {"kernel": "k", "value": 346.0, "unit": "Mpaths/s", "ms_per_frame": 56.8},
        "fast_fp": {"kernel": "k", "value": 717.0, "unit": "Mpaths/s", "kernel_ms": 27.4},
        "time_to_rmse": {"reached": True, "target": 0.01, "spp": 236, "chosen_s": 0.107, "plain_s": 1.50,
                         "speedup": 14.0, "curve": curve, "oracle_spp": 4096, "oracle_noise_rmse": 0.002},
        "fp64_error": {"spp": 64, "rmse_fp32_vs_fp64_same_seeds": 0.0069},
        "euclidean_320x240": {"chosen": {"value": 4200.0}, "plain": {"value": 2300.0}, "time_to_rmse": {}},
        "radiance_cache_1080p": {"bit_exact_forms": {"reached": True, "spp": 168, "chosen_s": 0.28, "plain_s": 4.8,
                                                     "speedup": 17.1},
                                 "render_tuning": tuning, "cached": {"reached": True, "spp": 280, "seconds": 0.63,
                                                                     "curve": curve, "speedup_vs_chosen": 0.43}},
        "batch_32_jobs": {"graph_replay": {"value": 1000.0}, "eager": {"value": 990.0},
                          "one_job_per_launch": {"value": 350.0}, "bit_identical_to_single_jobs": True,
                          "tuner_table": _tuner_table(120)},
        "wavefront_cells": {"megakernel": {}, "wavefront": {}, "bit_identical": True,
                            "wavefront_over_megakernel": 2.13},
        "sdf_1080p": {"value": 277.2, "render_tuning": tuning},
        "cpu_baseline": {"value": 0.25, "unit": "Mpaths/s", "cores": 16, "kind": "port",
                         "sample": "CPU oracle (FP32, OpenMP, 16 threads) on 3010560 paths of the same frame",
                         "host": {"os_cpu_count": 16}},
    }


def test_last_line_is_a_small_parseable_summary(tmp_path):
    bench = _bench()
    out = stub_result(bench)
    buf = io.StringIO()
    detail = tmp_path / "detail.json"
    bench.emit(out, str(detail), stream=buf)
    last = buf.getvalue().strip().splitlines()[-1]
    assert len(last.encode()) <= 2048, len(last.encode())
    s = json.loads(last)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline", "time_to_rmse"):
        assert key in s, key
    assert s["metric"] == bench.METRIC
    for key in ("bound", "achieved", "peak", "unit", "frac", "frac_nonpacked", "kernel_ms", "flops_per_path",
                "traffic", "algorithmic_bytes"):
        assert key in s["roofline"], key
    assert {"spp", "chosen_s", "plain_s", "speedup"} <= set(s["time_to_rmse"])
    assert {"value", "unit", "cores", "kind", "sample"} <= set(s["cpu_baseline"])
    assert s["legs"]["batch_32_jobs"]["bit_identical"] is True
    assert abs(s["value"] - 639.12) < 0.01
    # the detail file keeps everything, tables included
    full = json.loads(detail.read_text())
    assert len(full["batch_32_jobs"]["tuner_table"]) == 120
    assert full["roofline"]["flops_note"] == "n" * 400


def test_summary_of_failed_legs_stays_small():
    bench = _bench()
    out = stub_result(bench)
    for leg in ("euclidean_320x240", "radiance_cache_1080p", "batch_32_jobs", "wavefront_cells", "sdf_1080p"):
        out[leg] = {"error": "RuntimeError: " + "e" * 5000}
    s = bench.summary(out)
    assert len(json.dumps(s).encode()) <= 2048
    assert s["legs"]["sdf_1080p"]["error"].startswith("RuntimeError")


def test_bias_variance_fit_recovers_known_terms():
    import numpy as np

    bench = _bench()
    b, v = 0.006, 0.02
    curve = [(n, float(np.sqrt(b * b + v / n))) for n in (8, 16, 32, 64, 128, 256)]
    fit = bench.bias_variance(curve)
    assert abs(fit["floor_rmse"] - b) < 1e-9 and abs(fit["var_per_spp"] - v) < 1e-9
    assert abs(bench.spp_for(fit, 0.01) - v / (1e-4 - b * b)) < 1e-6
    # a floor above the target: no number of samples reaches it
    assert bench.spp_for(bench.bias_variance([(n, float(np.sqrt(0.012 ** 2 + v / n))) for n in (8, 64, 512)]),
                         0.01) is None
    assert bench.bias_variance([(4, 0.1)]) is None
