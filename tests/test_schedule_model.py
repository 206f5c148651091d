"""Learned schedule, host side: feature vector, forward pass, model file.

The forward pass is checked against a numpy float64 restatement of the same
network (2x64 ReLU MLP); the feature vector against log2(1 + x) in float64.
"""
import numpy as np
import pytest

import wedge_amd as W

APP_BFS, APP_CC, APP_SSSP = W.APP_BFS, W.APP_CC, W.APP_SSSP


def _random_params(seed=0, scale=0.3, lo=-1e30, hi=1e30, margins=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    I = W.MLP_IN
    p = rng.normal(0.0, scale, W.MLP_PARAMS).astype(np.float32)
    p[-4:-1] = margins
    p[-1] = 1e30  # no size bound
    p[:I] = lo                                         # clamp range
    p[I:2 * I] = hi
    p[2 * I:3 * I] = rng.normal(0.0, 1.0, I)           # mean
    p[3 * I:4 * I] = rng.uniform(0.2, 2.0, I)          # inv_std
    return p


def _split(p):
    I, H, A = W.MLP_IN, W.MLP_HIDDEN, W.MLP_ACTIONS
    o = 0
    out = {}
    for name, n in (("lo", I), ("hi", I), ("mean", I), ("istd", I), ("W1", I * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * A),
                    ("b3", A), ("margin", 3), ("lv_max", 1)):
        out[name] = p[o:o + n].astype(np.float64)
        o += n
    assert o == W.MLP_PARAMS
    out["W1"] = out["W1"].reshape(I, H)
    out["W2"] = out["W2"].reshape(H, H)
    out["W3"] = out["W3"].reshape(H, A)
    return out


def _forward_np(p, x):
    q = _split(p)
    z = (np.clip(x.astype(np.float64), q["lo"], q["hi"]) - q["mean"]) * q["istd"]
    h1 = np.maximum(z @ q["W1"] + q["b1"], 0.0)
    h2 = np.maximum(h1 @ q["W2"] + q["b2"], 0.0)
    return h2 @ q["W3"] + q["b3"]


def test_feature_vector_matches_definition():
    V, E = 1 << 20, 16_000_000
    for app, it, fv, fe, unv in ((APP_BFS, 3, 1234, 56789, 900_000), (APP_SSSP, 0, 1, 7, 0), (APP_CC, 17, V, E, 0),
                                 (APP_BFS, 2**40, 0, 0, 0)):
        x = W.mlp_features(app, it, V, E, fv, fe, unv)
        L = lambda v: np.log2(1.0 + float(v))  # noqa: E731
        u = 0.0 if app == APP_CC else 1.0
        want = [L(fv) - L(V), L(fe) - L(E), u * (L(unv) - L(V)), L(E) - L(V), L(fe) - L(fv), L(it) - L(V) / 2,
                app == APP_BFS, app == APP_CC, app == APP_SSSP, L(fe) - L(V), L(fv) - L(E), u * (L(unv) - L(E))]
        assert np.allclose(x, np.array(want, np.float64), atol=1.6e-2, rtol=0)
    # log2(1 + x) is exact where the mantissa fraction is 0 and at the 1/2 point
    x = W.mlp_features(APP_BFS, 0, 2**20 - 1, 1, 2**30 - 1, 3 * 2**29 - 1, 0)
    assert x[0] == 10.0
    assert abs(x[4] - (np.log2(1.5 * 2**30) - 30.0)) < 1e-6


def test_work_estimate_matches_definition():
    V, E = 25_165_824, 62_902_008
    for fe in (0, 1, 10_000, 5_000_000, E):
        b = W.mlp_base(V, E, fe)
        want = [np.log2(1.0 + E + 2**18)] + [np.log2(1.0 + min(E, fe << k) + (V >> 5) + 2**18) for k in (3, 4, 5, 6, 8)]
        assert np.allclose(b, want, atol=8e-3, rtol=0)


def test_forward_matches_numpy_and_picks_argmin():
    for lo, hi in ((-1e30, 1e30), (-1.0, 1.5)):  # with and without clamping
        p = _random_params(1, lo=lo, hi=hi)
        W.mlp_set(p)
        rng = np.random.default_rng(2)
        for _ in range(50):
            x = rng.normal(0.0, 3.0, W.MLP_IN).astype(np.float32)
            base = W.mlp_base(int(rng.integers(1, 2**26)), int(rng.integers(1, 2**30)), int(rng.integers(0, 2**24)))
            a, y = W.mlp_predict(x, APP_BFS, base)
            ref = _forward_np(p, x) + base
            assert np.allclose(y, ref, rtol=1e-4, atol=1e-4)
            assert a == int(np.argmin(y))  # any action for BFS, lowest index on ties
            for app in (APP_CC, APP_SSSP):  # dense or the default granularity only
                a2, _ = W.mlp_predict(x, app, base)
                assert a2 == (0 if y[0] <= y[2] else 2)
            _, y0 = W.mlp_predict(x, APP_BFS)  # no base: the network output alone
            assert np.allclose(y0, _forward_np(p, x), rtol=1e-4, atol=1e-4)
    W.mlp_unload()


def test_plain_choice_stands_within_the_margin():
    p = _random_params(5, margins=(0.0, 0.75, 1.5))
    W.mlp_set(p)
    rng = np.random.default_rng(6)
    for _ in range(200):
        x = rng.normal(0.0, 3.0, W.MLP_IN).astype(np.float32)
        for app, k in ((APP_BFS, 0), (APP_CC, 1), (APP_SSSP, 2)):
            allowed = list(range(W.MLP_ACTIONS)) if app == APP_BFS else [0, 2]
            for plain in (-1, 0, 2):
                a, y = W.mlp_predict(x, app, None, plain)
                best = min(allowed, key=lambda q: (y[q], q))
                if plain < 0 or y[plain] - y[best] > p[-4 + k]:
                    assert a == best
                else:
                    assert a == plain
    with pytest.raises(W.WedgeError):
        W.mlp_predict(np.zeros(W.MLP_IN, np.float32), APP_BFS, None, W.MLP_ACTIONS)
    W.mlp_unload()


def test_model_file_round_trip_is_exact(tmp_path):
    p = _random_params(3)
    W.mlp_set(p)
    path = str(tmp_path / "m.txt")
    W.mlp_save(path)
    W.mlp_unload()
    assert not W.mlp_loaded()
    W.mlp_load(path)
    assert W.mlp_loaded()
    assert np.array_equal(W.mlp_get(), p)  # bit-exact FP32 round trip
    W.mlp_unload()


def test_model_file_errors(tmp_path):
    with pytest.raises(W.WedgeError) as e:
        W.mlp_load(str(tmp_path / "missing.txt"))
    assert e.value.code == W.WG_E_IO
    bad = tmp_path / "bad.txt"
    bad.write_text("wedge-schedule-mlp 1\nin 12 hidden 32 actions 6\n")
    with pytest.raises(W.WedgeError) as e:
        W.mlp_load(str(bad))
    assert e.value.code == W.WG_E_FORMAT
    p = _random_params(4)
    W.mlp_set(p)
    short = str(tmp_path / "short.txt")
    W.mlp_save(short)
    W.mlp_unload()
    text = open(short).read().splitlines()
    with open(short, "w") as f:
        f.write("\n".join(text[:-1]) + "\n")  # one parameter missing
    with pytest.raises(W.WedgeError) as e:
        W.mlp_load(short)
    assert e.value.code == W.WG_E_FORMAT
    with pytest.raises(W.WedgeError):
        W.mlp_set(p[:-1])
    with pytest.raises(W.WedgeError):
        W.mlp_predict(np.zeros(W.MLP_IN, np.float32))  # nothing loaded


def test_actions_map_to_modes():
    assert W.mlp_action(0) == (W.MODE_DENSE, 4)
    assert [W.mlp_action(a)[1] for a in range(1, W.MLP_ACTIONS)] == [3, 4, 5, 6, 8]
    assert all(W.mlp_action(a)[0] == W.MODE_FRONTIER for a in range(1, W.MLP_ACTIONS))
    with pytest.raises(W.WedgeError):
        W.mlp_action(W.MLP_ACTIONS)


def test_default_model_file_loads():
    """The committed model (trained on engine traces) loads and has the right shape."""
    import os

    if not os.path.exists(W.DEFAULT_MODEL_PATH):
        pytest.skip("no trained model committed yet")
    W.mlp_load(W.DEFAULT_MODEL_PATH)
    p = W.mlp_get()
    assert p.shape == (W.MLP_PARAMS,) and np.all(np.isfinite(p))
    W.mlp_unload()
