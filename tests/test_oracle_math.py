"""The oracle's deterministic log / exp / pow (det math mode).

These replace glibc's log/pow of the reference (LA/Tensor2.C:1185-1194,
FE/IsoHardElastoplastic.C:175-183) by functions built from IEEE +,-,*,/ only,
shared in algorithm with the GPU kernels.  Pinned here against libm (numpy)
by ulp distance, and on the exact special cases the reference's linear
hardening relies on (pow(x, 1) = x, pow(x, 0) = 1).
"""
import numpy as np

import oracle_lib as O


def ulp_dist(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return np.abs(ia - ib)


def test_det_log_within_one_ulp():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        1.0 + rng.uniform(-0.2, 0.2, 20000),          # stretches of the polar decomposition
        1.0 + rng.uniform(-1e-6, 1e-6, 5000),
        np.exp(rng.uniform(-700, 700, 20000)),
        rng.uniform(1e-10, 10.0, 5000),
        [1.0, 2.0, 0.5, np.sqrt(2.0), 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53, 5e-324, 1e308],
    ])
    got = np.array([L.dnlo_det_log(float(x)) for x in xs])
    ref = np.log(xs)
    d = ulp_dist(got, ref)
    assert d.max() <= 1, (xs[np.argmax(d)], got[np.argmax(d)], ref[np.argmax(d)])
    assert L.dnlo_det_log(1.0) == 0.0


def test_det_exp_within_two_ulp():
    L = O.lib()
    rng = np.random.default_rng(2)
    ys = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-1, 1, 20000), [0.0, 1.0, -1.0]])
    got = np.array([L.dnlo_det_exp(float(y)) for y in ys])
    ref = np.exp(ys)
    assert ulp_dist(got, ref).max() <= 2
    assert L.dnlo_det_exp(0.0) == 1.0


def test_det_pow_special_cases_and_accuracy():
    L = O.lib()
    rng = np.random.default_rng(3)
    for x in rng.uniform(1e-10, 5.0, 1000):
        assert L.dnlo_det_pow(x, 1.0) == x       # linear hardening: B evp^1
        assert L.dnlo_det_pow(x, 0.0) == 1.0     # its derivative: n B evp^0
    xs = rng.uniform(1e-10, 5.0, 20000)
    for y in (0.31, -0.69, 0.5, 2.0):
        got = np.array([L.dnlo_det_pow(float(x), y) for x in xs])
        ref = np.power(xs, y)
        # exp(y log x): error grows with |y log x| (|y log x| <= 16 here)
        assert np.max(np.abs(got - ref) / ref) < 2e-14
