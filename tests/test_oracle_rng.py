"""Pins Philox4x32-10, u01 and the uniform direct sampler of the oracle."""
import numpy as np
import pytest

import oracle as O


def test_philox_known_answers(goldens):
    for v in goldens["philox_kat"]["vectors"]:
        ctr = [int(h, 16) for h in v["ctr"]]
        key = [int(h, 16) for h in v["key"]]
        out = ["%08x" % w for w in O.philox4x32_10(ctr, key)]
        assert out == v["out"]


def test_philox_planted_round_count_fails(goldens):
    # a plausible mistake (one Philox round too few, i.e. nine key bumps
    # applied) must not reproduce the known answers
    v = goldens["philox_kat"]["vectors"][2]
    ctr = np.array([int(h, 16) for h in v["ctr"]], np.uint64)
    key = np.array([int(h, 16) for h in v["key"]], np.uint64)
    M = 0xFFFFFFFF

    def rounds(c, k, r):
        c, k = [int(x) for x in c], [int(x) for x in k]
        for _ in range(r):
            p0, p1 = 0xD2511F53 * c[0], 0xCD9E8D57 * c[2]
            c = [(p1 >> 32) ^ c[1] ^ k[0], p1 & M, (p0 >> 32) ^ c[3] ^ k[1], p0 & M]
            k = [(k[0] + 0x9E3779B9) & M, (k[1] + 0xBB67AE85) & M]
        return ["%08x" % w for w in c]

    assert rounds(ctr, key, 10) == v["out"]
    assert rounds(ctr, key, 9) != v["out"]


def test_u01_closed_form():
    # uniform-sampler.cl:8-10: (0.5 + (i >> 9)) * 2^-23, open interval (0, 1)
    for i in [0, 1, 511, 512, 0x12345678, 0xFFFFFFFF]:
        expect = np.float32((0.5 + (i >> 9)) * 2.0 ** -23)
        assert O.u01(i) == expect
    assert 0.0 < O.u01(0) < O.u01(0xFFFFFFFF) < 1.0
    assert O.u01(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)


def test_uniform_direct_sampler_golden(goldens):
    g = goldens["G1_uniform_direct_sampler"]
    x = O.uniform_sample(g["seed"], g["params"][0], g["params"][1], g["n"])
    assert x[:4].tolist() == g["first4"]
    assert x[9996:10000].tolist() == g["at9996"]
    assert float(x.max()) == g["max"]
    assert float(x.min()) == g["min"]
    # the reference's mean is a float reduction on the device; 1e-7 relative
    assert float(x.astype(np.float64).mean()) == pytest.approx(g["mean"], rel=1e-7)


def test_uniform_sampler_contraction_is_pinned(goldens):
    # the golden discriminates u*(hi-lo)+lo rounded twice from the fused form
    g = goldens["G1_uniform_direct_sampler"]
    r = O.philox4x32_10([0, 0xF00DCAFE, 0xDEADBEEF, 0xBEEFF00D], [g["seed"], 0xDECAFAAA])
    u = np.array([O.u01(int(w)) for w in r], np.float32)
    lo, hi = np.float32(g["params"][0]), np.float32(g["params"][1])
    unfused = (u * (hi - lo) + lo).astype(np.float32)
    assert unfused.tolist() != g["first4"]


KINDS = {"gaussian": 1, "exponential": 2, "erlang": 3}


@pytest.mark.parametrize("name", sorted(KINDS))
def test_direct_samplers_goldens(goldens, name):
    g = goldens["G2_direct_samplers"][name]
    x = O.direct_sample(KINDS[name], 123, g["params"], 10000)
    # the reference's native_sin/cos/log differ from the C library's in the
    # last bits: 2e-6 relative
    assert np.allclose(x[:4], g["first4"], rtol=2e-6, atol=0)
    assert np.allclose(x[9996:], g["at9996"], rtol=2e-6, atol=0)
    assert float(x.max()) == pytest.approx(g["max"], rel=2e-6)
    assert float(x.min()) == pytest.approx(g["min"], rel=2e-6)
    assert float(x.astype(np.float64).mean()) == pytest.approx(g["mean"], rel=1e-6)


def test_direct_samplers_moments():
    # closed forms: N(100, 3^2), Exp(4) mean 1/4, Erlang(2, 3) mean 3/2 var 3/4
    x = O.direct_sample(1, 5, [100.0, 3.0], 400000).astype(np.float64)
    assert abs(x.mean() - 100) < 0.03 and abs(x.std() - 3) < 0.02
    x = O.direct_sample(2, 6, [4.0], 400000).astype(np.float64)
    assert abs(x.mean() - 0.25) < 0.002
    x = O.direct_sample(3, 7, [2.0, 3.0], 400000).astype(np.float64)
    assert abs(x.mean() - 1.5) < 0.01 and abs(x.var() - 0.75) < 0.01
