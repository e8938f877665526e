"""Host logic of the render tuner and the program options: option strings,
the batch tile unit, and that every option macro is one the device sources
read (no GPU)."""
import pathlib
import re

from hypertrace_amd.processing import STRICT_FP
from hypertrace_amd.tuning import CANDIDATES, _knobs, candidate_options, options

DEVICE = pathlib.Path(__file__).resolve().parents[1] / "hypertrace_amd" / "csrc" / "device" / "ht"


def test_candidates_distinct():
    opts = [candidate_options(c) for c in CANDIDATES]
    assert len(set(opts)) == len(opts)
    assert all(o.startswith(STRICT_FP) for o in opts)


def test_knob_defaults():
    assert _knobs((4, 5)) == (4, 5, False, 0, False, False, False, False, False, 0, 0, False)
    assert candidate_options((4, 4, True, 4, False, True, True)).endswith(
        "-DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_WAVE_PRIO=1 -DHT_GROUP_QUEUE=1")
    assert options(4, 5) == candidate_options((4, 5, False, 0, False))
    assert candidate_options((4, 5, True, 0, False, True)).endswith("-DHT_SCALAR_BOUNDS=1 -DHT_WAVE_PRIO=1")
    assert "-DHT_PIXEL_QUEUE=1" in options(4, 4, pixel_queue=True)
    assert "-DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=8" in options(2, 6, scalar_bounds=True, groups=8)


def test_knob_macros_are_read_by_the_device_sources():
    # a misspelt define would silently build the default program
    source = "\n".join(p.read_text() for p in DEVICE.glob("*.hh"))
    names = set()
    for c in CANDIDATES:
        names |= set(re.findall(r"-D(HT_[A-Z0-9_]+)=", candidate_options(c)))
    assert {"HT_CULL_UNROLL", "HT_WAVES_PER_EU", "HT_SCALAR_BOUNDS", "HT_CULL_GROUPS", "HT_PIXEL_QUEUE",
            "HT_WAVE_PRIO", "HT_GROUP_QUEUE", "HT_HIT_ORDER", "HT_HELPERS", "HT_SPARSE_L1", "HT_PIXEL_SPREAD",
            "HT_PIXEL_CLASSES"} <= names
    for name in names:
        assert re.search(rf"#\s*(ifdef|ifndef|if|elif)\b[^\n]*\b{name}\b", source), name


def test_program_option_strings():
    from hypertrace_amd.processing import NO_CULL_FP, program_options

    assert program_options() == STRICT_FP
    assert program_options(no_cull=True) == NO_CULL_FP
    assert program_options(eps32=1e-5).endswith("-DHT_EPS32=1e-05f")
    assert "-DHT_HY_DIST_FP64=1" in program_options(hy_distance_fp64=True)
    assert "-DHT_RR_DEPTH=3" in program_options(rr_depth=3)


def test_batch_tile_unit():
    from hypertrace_amd.batch import OPTIONS, SCALAR_BOUNDS, WAVE_QUEUE, tile_unit

    assert tile_unit(None, 256) == 256
    assert tile_unit(SCALAR_BOUNDS, 512) == 512
    assert tile_unit(WAVE_QUEUE, 256) == 64
    assert any(opts == WAVE_QUEUE for opts, _ in OPTIONS)


def test_option_macros_are_read_by_the_device_sources():
    from hypertrace_amd.batch import WAVE_QUEUE
    from hypertrace_amd.processing import program_options

    source = "\n".join(p.read_text() for p in DEVICE.glob("*.hh"))
    opts = " ".join([program_options(rr_depth=3, no_cull=True, scalar_bounds=True, hy_distance_fp64=True,
                                     eps32=1e-5), WAVE_QUEUE])
    names = set(re.findall(r"-D(HT_[A-Z0-9_]+)=", opts))
    assert {"HT_EPS32", "HT_BATCH_WAVE_QUEUE", "HT_HY_DIST_FP64", "HT_RR_DEPTH", "HT_NO_CULL"} <= names
    for name in names:
        assert re.search(rf"\b{name}\b", source) and re.search(rf"#\s*(ifdef|ifndef|if|elif)\b[^\n]*\b{name}\b",
                                                                 source), name


def test_bench_fixed_knobs_and_gpu_count_guard():
    """bench.py's --render-knobs parser gives the tuner's option strings, and
    --gpus above 1 exits with a message (one GPU only)."""
    import importlib.util
    import subprocess
    import sys

    root = pathlib.Path(__file__).resolve().parents[1]
    spec = importlib.util.spec_from_file_location("_bench", root / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    k = bench.fixed_knobs("4,6,sb,g4,q")["best"]
    assert (k["unroll"], k["waves_per_eu"], k["scalar_bounds"], k["groups"], k["pixel_queue"]) == (4, 6, True, 4, True)
    assert k["options"] == candidate_options((4, 6, True, 4, True))
    assert bench.fixed_knobs("4,5,sb,p")["best"]["options"] == candidate_options((4, 5, True, 0, False, True))
    assert bench.fixed_knobs("2,5")["best"]["options"] == options(2, 5)
    assert bench.fixed_knobs("4,4,sb,g4,p,gq")["best"]["options"] == candidate_options((4, 4, True, 4, False, True, True))
    assert bench.fixed_knobs("4,3,sb,g4,p,gq,ho")["best"]["options"] == candidate_options(
        (4, 3, True, 4, False, True, True, True))
    assert bench.count_mode(bench.fixed_knobs("4,3,sb,g4,p,gq,ho")["best"]) == 2004
    assert bench.count_mode(bench.fixed_knobs("4,4,sb,g4,p,gq")["best"]) == 1004
    k = bench.fixed_knobs("4,4,sb,g4,p,gq,ho,hl,s32")["best"]
    assert k["sparse"] == 32 and k["options"] == candidate_options((4, 4, True, 4, False, True, True, True, True, 32))
    assert k["options"].endswith("-DHT_HELPERS=1 -DHT_SPARSE_L1=32") and bench.count_mode(k) == 2004
    assert bench.fixed_knobs("4,4,sb,g4,q,p,gq,ho,hl,s32,sp1")["best"]["options"].endswith("-DHT_PIXEL_SPREAD=1")
    k = bench.fixed_knobs("4,4,sb,g4,q,p,gq,ho,hl,s32,sp1,pc")["best"]
    assert k["classes"] and k["options"] == candidate_options((4, 4, True, 4, True, True, True, True, True, 32, 1, True))
    assert k["options"].endswith("-DHT_PIXEL_SPREAD=1 -DHT_PIXEL_CLASSES=1")
    k = bench.fixed_knobs("4,4,sb,g4,q,p,gq,ho,hl,s32,sp64")["best"]
    assert (k["sparse"], k["spread"], k["pixel_queue"]) == (32, 64, True)
    assert k["options"] == candidate_options((4, 4, True, 4, True, True, True, True, True, 32, 64))
    assert k["options"].endswith("-DHT_SPARSE_L1=32 -DHT_PIXEL_SPREAD=64") and bench.count_mode(k) == 2004
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_pick_prefers_spill_free_programs():
    from hypertrace_amd.tuning import pick

    # a spill-free candidate within the tolerance wins over a faster spilling one
    table = [{"ms": 10.0, "scratch": 24}, {"ms": 10.1, "scratch": 0}, {"ms": 12.0, "scratch": 0}]
    assert pick(table) is table[1]
    # beyond the tolerance the spilling one wins
    table = [{"ms": 10.0, "scratch": 24}, {"ms": 11.0, "scratch": 0}]
    assert pick(table) is table[0]
    assert pick(table, tolerance=0.2) is table[1]
    # every candidate spills: the fastest
    spilling = [{"ms": 10.0, "scratch": 24}, {"ms": 9.0, "scratch": 100}]
    assert pick(spilling) is spilling[1]
