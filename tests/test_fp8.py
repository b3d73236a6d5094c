"""fp8 e4m3 (OCP) encoding and per-column quantisation on the host."""
import numpy as np

from bayadera_amd import models


def test_e4m3_known_values():
    v = models.FP8_E4M3_VALUES
    assert v[0x38] == 1.0 and v[0x30] == 0.5 and v[0xC0] == -2.0
    assert v[0x7E] == 448.0 and v[0x01] == 2.0 ** -9 and v[0x08] == 2.0 ** -6
    assert np.isnan(v[0x7F]) and np.isnan(v[0xFF])
    assert v[0x80] == 0.0  # negative zero


def test_e4m3_encoding_rounds_to_nearest_even_and_saturates():
    enc = models.to_fp8_e4m3
    assert enc(1.0) == 0x38 and enc(-1.0) == 0xB8 and enc(0.0) == 0x00
    assert enc(1000.0) == 0x7E and enc(-1e9) == 0xFE
    # 2^-10 lies halfway between 0 (code 0) and 2^-9 (code 1): to the even code
    assert enc(2.0 ** -10) == 0x00
    # 1.5 * 2^-9 is halfway between codes 1 and 2: to code 2
    assert enc(1.5 * 2.0 ** -9) == 0x02
    # 1 + 1/16 is halfway between 1 (0x38) and 1.125 (0x39): to 0x38
    assert enc(1.0 + 1.0 / 16) == 0x38
    assert enc(1.0 + 3.0 / 16) == 0x3A  # halfway 1.125 / 1.25: to the even code 0x3a
    assert enc(1.07) == 0x39


def test_e4m3_round_trip_of_every_finite_code():
    codes = np.array([c for c in range(256) if c not in (0x7F, 0xFF, 0x80)], np.uint8)
    back = models.to_fp8_e4m3(models.from_fp8_e4m3(codes))
    assert np.array_equal(back, codes)


def test_column_quantisation():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((500, 8)) / np.sqrt(8)
    X[:, 3] = 0.0
    q, sc = models.quantize_fp8_columns(X)
    deq = models.from_fp8_e4m3(q).astype(np.float64) * sc[None, :]
    assert sc[3] == 1.0 and np.all(q[:, 3] == 0)
    big = np.abs(X).argmax(axis=0)
    for j in (0, 1, 2, 4):
        assert abs(models.from_fp8_e4m3(q[big[j], j])) == 448.0
    # relative error of e4m3 is at most 2^-4 for normal values
    nz = np.abs(X) > 2.0 ** -6 * sc[None, :]
    assert np.all(np.abs(deq - X)[nz] <= np.abs(X)[nz] * 2.0 ** -4 + 1e-12)


def test_e4m3_encoding_is_nearest_by_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(3000) * np.exp(rng.uniform(-12, 7, 3000))
    vals = models.FP8_E4M3_VALUES.astype(np.float64)
    fin = vals[np.isfinite(vals)]
    a = np.clip(x, -448.0, 448.0)
    got = vals[models.to_fp8_e4m3(x)]
    best = np.min(np.abs(fin[None, :] - a[:, None]), axis=1)
    assert np.array_equal(np.abs(got - a), best)
