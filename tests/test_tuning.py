"""Host logic of the render tuner: candidate option strings (no GPU)."""
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
    assert _knobs((4, 5)) == (4, 5, False, 0, False)
    assert options(4, 5) == candidate_options((4, 5, False, 0, False))
    assert "-DHT_PIXEL_QUEUE=1" in options(4, 4, pixel_queue=True)
    assert "-DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=8" in options(2, 6, scalar_bounds=True, groups=8)


def test_knob_macros_are_read_by_the_device_sources():
    # a misspelt define would silently build the default program
    source = "\n".join(p.read_text() for p in DEVICE.glob("*.hh"))
    names = set()
    for c in CANDIDATES:
        names |= set(re.findall(r"-D(HT_[A-Z0-9_]+)=", candidate_options(c)))
    assert {"HT_CULL_UNROLL", "HT_WAVES_PER_EU", "HT_SCALAR_BOUNDS", "HT_CULL_GROUPS", "HT_PIXEL_QUEUE"} <= names
    for name in names:
        assert re.search(rf"#\s*(ifdef|ifndef|if|elif)\b[^\n]*\b{name}\b", source), name
