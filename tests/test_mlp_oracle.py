"""Pins for the CPU restatement of the closure network (oracle/mlp_oracle.py): the analytic
gradient of the float64 network against central finite differences, and the bf16 network
against the float64 one."""
import numpy as np

from oracle import mlp_oracle as M


def _random_params(rng, scale=0.3):
    L = M.layout()
    p = rng.normal(0.0, scale, L["total"])
    # padding stays zero, as in the library
    w1 = p[L["w1"][0]:L["w1"][0] + L["w1"][1]].reshape(M.IN_PAD, M.HID)
    w1[195:] = 0.0
    w4 = p[L["w4"][0]:L["w4"][0] + L["w4"][1]].reshape(M.HID, M.OUT_PAD)
    w4[:, 3:] = 0.0
    p[L["b4"][0] + 3:L["b4"][0] + L["b4"][1]] = 0.0
    return p


def test_finite_difference_gradient():
    rng = np.random.default_rng(0)
    p = _random_params(rng)
    X = np.zeros((6, M.IN_PAD))
    X[:, :195] = rng.normal(0, 1, (6, 195))
    T = rng.normal(0, 1, (6, 3))
    _, loss, g = M.forward_backward(p, X, T, bf16=False)
    L = M.layout()
    picks = []
    for k in ("w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        off, n = L[k]
        valid = np.arange(n)
        if k == "w1":
            valid = valid[valid // M.HID < 195]
        elif k == "w4":
            valid = valid[valid % M.OUT_PAD < 3]
        elif k == "b4":
            valid = valid[:3]
        picks += list(off + rng.choice(valid, min(6, valid.size), replace=False))
    eps = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (M.forward_backward(pp, X, T, bf16=False)[1] - M.forward_backward(pm, X, T, bf16=False)[1]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-9, (i, num, g[i])


def test_planted_sign_error_fails_finite_differences():
    """The check above notices a flipped sign in the SiLU derivative."""
    rng = np.random.default_rng(1)
    p = _random_params(rng)
    X = np.zeros((4, M.IN_PAD))
    X[:, :195] = rng.normal(0, 1, (4, 195))
    T = rng.normal(0, 1, (4, 3))
    orig = M.dsilu
    try:
        M.dsilu = lambda z: -orig(z)
        _, _, g = M.forward_backward(p, X, T, bf16=False)
    finally:
        M.dsilu = orig
    i = M.layout()["w1"][0] + 5
    eps = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += eps
    pm[i] -= eps
    num = (M.forward_backward(pp, X, T, bf16=False)[1] - M.forward_backward(pm, X, T, bf16=False)[1]) / (2 * eps)
    assert abs(num - g[i]) > 1e-6 * max(1.0, abs(num))


def test_bf16_network_tracks_float64():
    rng = np.random.default_rng(2)
    p = _random_params(rng, 0.1)
    X = np.zeros((64, M.IN_PAD))
    X[:, :195] = M.bf16_round(rng.normal(0, 1, (64, 195)))
    T = rng.normal(0, 0.5, (64, 3))
    y16, l16, g16 = M.forward_backward(p, X, T, bf16=True)
    y64, l64, g64 = M.forward_backward(p, X, T, bf16=False)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 2e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 65504.0], np.float32)
    got = M.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1.0)
    assert got[1] == 1.0
    assert got[2] == 1.0078125
    assert abs(got[4] + 3.140625) < 1e-12
