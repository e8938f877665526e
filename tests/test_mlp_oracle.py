"""Pins of the radiance-cache MLP oracle (oracle/mlp.py) by what the
mathematics fixes: trigonometric identities of the encoding, finite
differences of the loss for the backward pass, the first Adam step's closed
form, and exact BF16 rounding cases."""
import numpy as np

from oracle import mlp as om
from hypertrace_amd.radiance_cache import init_params, unpack_params


def _params(seed=1, scale=1.0):
    p = {k: np.asarray(v, np.float64) * (scale if k.startswith("W") else 1.0)
         for k, v in unpack_params(init_params(seed)).items()}
    rng = np.random.default_rng(seed + 100)
    for l in range(6):
        p[f"b{l}"] = rng.normal(0.0, 0.05, p[f"b{l}"].shape)
    return p


def test_bf16_rounding_exact_cases():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    cases = {
        one: one,
        one + ulp: one + ulp,
        one + ulp / 2: one,  # tie to even (mantissa 0)
        one + 3 * ulp / 2: one + 2 * ulp,  # tie to even (mantissa 2)
        one + ulp / 2 + 2.0 ** -20: one + ulp,
        -3.0: -3.0,
        0.0: 0.0,
    }
    for x, want in cases.items():
        assert om.bf16(np.float32(x)) == want, (x, om.bf16(np.float32(x)), want)


def test_encoding_identities():
    rng = np.random.default_rng(3)
    feat = rng.uniform(-1, 1, (50, 12))
    e = om.encode(feat, emulate=False)
    for f in range(12):
        for b in range(6):
            s, c = e[:, 12 * f + 2 * b], e[:, 12 * f + 2 * b + 1]
            assert np.allclose(s * s + c * c, 1.0, atol=1e-14)
            if b + 1 < 6:
                # double angle: band b+1 = 2x the angle of band b
                assert np.allclose(e[:, 12 * f + 2 * (b + 1)], 2 * s * c, atol=1e-13)
                assert np.allclose(e[:, 12 * f + 2 * (b + 1) + 1], c * c - s * s, atol=1e-13)
        assert np.allclose(e[:, 12 * f], np.sin(np.pi * feat[:, f]), atol=1e-15)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = _params(2)
    feat = rng.uniform(-1, 1, (8, 12))
    target = rng.uniform(0, 2, (8, 3))
    fwd = om.forward(p, feat, emulate=False)
    den0 = fwd["out"] ** 2 + om.LOSS_EPS  # the loss holds the denominator fixed
    g = om.backward(p, fwd, target, emulate=False)
    eps = 1e-6
    for name in om.NAMES:
        for _ in range(3):
            v = rng.normal(size=p[name].shape)
            plus = {k: (p[k] + eps * v if k == name else p[k]) for k in p}
            minus = {k: (p[k] - eps * v if k == name else p[k]) for k in p}
            fd = (om.loss(om.forward(plus, feat, False)["out"], target, den0)
                  - om.loss(om.forward(minus, feat, False)["out"], target, den0)) / (2 * eps)
            an = float(np.sum(g[name] * v))
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (name, fd, an)


def test_first_adam_step_closed_form():
    """With zero moments, step 1 moves each parameter by lr * g / (|g| + eps)."""
    rng = np.random.default_rng(7)
    p = {k: rng.normal(size=(4, 3)) for k in om.NAMES}
    g = {k: rng.normal(size=(4, 3)) for k in om.NAMES}
    m = {k: np.zeros((4, 3)) for k in om.NAMES}
    v = {k: np.zeros((4, 3)) for k in om.NAMES}
    before = {k: p[k].copy() for k in om.NAMES}
    om.adam(p, m, v, g, 1, 1e-3)
    for k in om.NAMES:
        want = before[k] - 1e-3 * g[k] / (np.abs(g[k]) + 1e-8)
        assert np.allclose(p[k], want, rtol=0, atol=1e-12)


def test_emulated_forward_close_to_exact():
    """BF16 rounding changes the prediction by about its precision, not more."""
    rng = np.random.default_rng(9)
    p = _params(4)
    feat = rng.uniform(-1, 1, (256, 12))
    a = om.forward(p, feat, True)["out"]
    b = om.forward(p, feat, False)["out"]
    scale = np.abs(b).mean()
    assert np.abs(a - b).mean() < 0.05 * scale
