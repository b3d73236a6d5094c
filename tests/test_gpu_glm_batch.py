"""The batched GLM runner (fp8 data, uneven problems) on the GPU against the
fp64 oracle and against single-posterior GLM samplers."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


def _problems(seed, shapes, pow2_scales=False):
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for N, d in shapes:
        X = rng.standard_normal((N, d)) / np.sqrt(d)
        b = rng.standard_normal(d)
        p = 1.0 / (1.0 + np.exp(-(X @ b)))
        Xs.append(X)
        ys.append((rng.random(N) < p).astype(np.float32))
    m = models.glm_logistic_batch_model(Xs, ys, prior_sd=5.0)
    if pow2_scales:
        # scales that are powers of two make scale * fp8 exact in bf16
        for p in range(m.problems):
            s = np.exp2(np.ceil(np.log2(np.asarray(m.scales[p], np.float64)))).astype(np.float32)
            deq = m.dequantized(p)
            m.X[p] = models.to_fp8_e4m3(deq / s[None, :])
            m.scales[p] = s
    return m


@pytest.mark.parametrize("f16", [False, True])
def test_device_fp8_decoding_matches_the_host_table(fac, f16):
    """Both device decodings (fp32 path; fp16 matrix-core path, scaled by
    2^-8 on the device) give every e4m3 code's exact value."""
    codes = np.arange(256, dtype=np.uint8)
    dev = fac.fp8_decode(codes, f16=f16)
    ref = models.FP8_E4M3_VALUES
    assert np.array_equal(np.isnan(dev), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(dev[ok], ref[ok])
    assert np.array_equal(np.signbit(dev[ok]), np.signbit(ref[ok]))


SHAPES = [(1000, 8), (2049, 16), (4100, 32), (20011, 64), (16, 64), (300, 3), (7000, 40)]


@pytest.mark.parametrize("prec,tol", [(_lib.BH_GLM_FP64, 1e-12), (_lib.BH_GLM_FP32, 2e-6),
                                      (_lib.BH_GLM_BF16X3, 2e-6), (_lib.BH_GLM_FP16X2, 2e-6)])
def test_batch_initial_log_posteriors_vs_oracle(fac, prec, tol):
    # the oracle holds X in float32, where scale * fp8 is exact only for
    # power-of-two scales: those make the fp64 comparison exact to 1e-12
    m = _problems(3, SHAPES, pow2_scales=(prec == _lib.BH_GLM_FP64))
    bt = fac.glm_batch(m, 512).set_precision(prec)
    lim = np.tile(np.array([-0.8, 0.8], np.float32), (m.problems, 64, 1))
    bt.init_position(11, lim)
    xs, lp = bt.state()
    for p in range(m.problems):
        ref = O.glm_logpost(m.dequantized(p), m.y[p], 5.0, xs[p][:96])
        err = np.abs(lp[p, :96] - ref) / (np.abs(ref) + 1.0)
        assert err.max() < tol, (p, err.max())
    bt.release()


def test_batch_init_uses_init_walkers_streams(fac):
    m = _problems(4, [(100, 8), (100, 32)])
    bt = fac.glm_batch(m, 512)
    lim = np.tile(np.array([-1.0, 2.0], np.float32), (2, 64, 1))
    bt.init_position(21, lim)
    xs, _ = bt.state()
    for p, d in enumerate(m.dims):
        ref = O.init_walkers(21 + p, lim[p, :d], int(d), 512)
        assert np.array_equal(xs[p].reshape(-1), ref.reshape(-1))
    bt.release()


@pytest.mark.parametrize("d", [32, 64])
def test_batch_problem_equals_single_glm_sampler(fac, d):
    """fp64 likelihoods, power-of-two column scales: problem p of the batch
    takes the same moves as a single GLM sampler on its data with init seed
    seed + p and bare seed seed0 + 2p."""
    m = _problems(5, [(3000, d), (5000, d)], pow2_scales=True)
    W = 1024
    bt = fac.glm_batch(m, W).set_precision(_lib.BH_GLM_FP64)
    lim = np.tile(np.array([-0.5, 0.5], np.float32), (2, 64, 1))
    bt.init_position(31, lim)
    bt.burn_in(40, 6, 2.0)
    xs, _ = bt.state()
    for p in range(2):
        Xb = models.to_bf16(m.dequantized(p).astype(np.float32))
        assert np.array_equal(models.from_bf16(Xb).reshape(-1, d), m.dequantized(p))
        sf = fac.mcmc_factory(models.glm_logistic_model(Xb, m.y[p], prior_sd=5.0))
        s = sf.create_sampler(40 + 2 * p, W, np.zeros(1, np.float32))
        s.set_glm_precision(_lib.BH_GLM_FP64).set_partner(_lib.BH_PARTNER_WALKER)
        s.init(40 + 2 * p)
        s.init_position(31 + p, lim[p, :d].reshape(-1))
        s.burn_in(6, 2.0)
        single = s.state()[0]
        same = np.all(single == xs[p], axis=1).mean()
        assert same > 0.999, (p, same)
        s.release()
        sf.release()
    bt.release()


@pytest.mark.parametrize("prec", [_lib.BH_GLM_BF16X3, _lib.BH_GLM_FP16X2])
def test_batch_run_accounting_and_series(fac, prec):
    m = _problems(6, SHAPES)
    bt = fac.glm_batch(m, 512).set_precision(prec)
    lim = np.tile(np.array([-0.3, 0.3], np.float32), (m.problems, 64, 1))
    bt.init_position(7, lim)
    bt.burn_in(1, 30, 2.0)
    n = 60
    series = np.zeros(n * int(m.dims.sum()), np.float32)
    res = bt.run(2, n, 2.0, series=series)
    acc = res["acceptance-rate"]
    assert np.all((acc > 0.0) & (acc < 1.0))
    for p, blk in enumerate(bt.series_blocks(series, n)):
        d = int(m.dims[p])
        assert blk.shape == (n, d)
        assert np.allclose(blk.astype(np.float64).mean(0), res["mean"][p, :d], rtol=1e-5,
                           atol=1e-6)
        assert np.all(res["sd"][p, :d] > 0) and np.all(res["mean"][p, d:] == 0)
    bt.release()


def test_batch_walker_rule(fac):
    m = _problems(7, [(100, 8)])
    with pytest.raises(_lib.BayaderaError):
        fac.glm_batch(m, 768)


def test_batch_matrix_core_precisions_agree_with_fp64_on_moves(fac):
    """From the same state, a few bare steps in fp16x2 and in bf16x3 take
    (nearly) the moves of the fp64 batch: the contractions are fp32-accurate,
    so only walkers within ~1e-4 of an accept threshold can differ."""
    m = _problems(8, [(3000, 32), (8000, 64), (600, 10)])
    W = 1024
    out = {}
    for prec in (_lib.BH_GLM_FP64, _lib.BH_GLM_BF16X3, _lib.BH_GLM_FP16X2):
        bt = fac.glm_batch(m, W).set_precision(prec)
        bt.init_position(41, np.tile(np.array([-0.4, 0.4], np.float32), (m.problems, 64, 1)))
        bt.burn_in(50, 4, 2.0)
        out[prec] = bt.state()[0].copy()
        bt.release()
    for prec in (_lib.BH_GLM_BF16X3, _lib.BH_GLM_FP16X2):
        for p in range(m.problems):
            same = np.all(out[prec][p] == out[_lib.BH_GLM_FP64][p], axis=1).mean()
            assert same > 0.99, (prec, p, same)


TILE_SHAPES = [(1000, 8), (2049, 16), (4100, 32), (20011, 64), (300, 3), (7000, 40), (9000, 24)]


def test_batch_tile_shapes_give_the_same_log_posteriors(fac):
    """The contraction's values do not depend on the tile shape: 8 column
    blocks per wave give the 4-block kernel's log posteriors bit for bit at
    the same unit rows.  Other unit rows move the boundaries of the
    epilogue's runs (up to four 16-row blocks share one fp32 logarithm of
    their factors' product) and the order of the units' double sums: the
    log posteriors agree to fp32 rounding of the run sums (~1e-9 relative
    at these sizes)."""
    m = _problems(5, TILE_SHAPES)
    bt = fac.glm_batch(m, 2048).set_precision(_lib.BH_GLM_FP16X2)
    lim = np.tile(np.array([-0.8, 0.8], np.float32), (m.problems, 64, 1))
    bt.init_position(12, lim)
    out = {}
    for ur, cb in ((2048, 4), (2048, 8), (1024, 4), (4096, 8)):
        print(f"tiles: stride-32 unit rows {ur}, column blocks {cb}", flush=True)
        bt.set_tiles(32, ur, cb).set_tiles(64, ur, 4)
        out[(ur, cb)] = bt.log_density(_lib.BH_GLM_FP16X2)
    assert np.array_equal(out[(2048, 4)], out[(2048, 8)])
    for k in ((1024, 4), (4096, 8)):
        assert np.allclose(out[k], out[(2048, 4)], rtol=1e-8, atol=0)
    with pytest.raises(_lib.ArgumentError):
        bt.set_tiles(64, 2048, 8)
    with pytest.raises(_lib.ArgumentError):
        bt.set_tiles(32, 100, 4)
    bt.release()


def test_batch_autotune_then_log_posteriors_vs_oracle(fac):
    """The autotuner keeps a valid shape per row stride, and the log
    posteriors it recomputes over the chosen units match the oracle's."""
    m = _problems(6, TILE_SHAPES)
    bt = fac.glm_batch(m, 2048).set_precision(_lib.BH_GLM_FP16X2)
    lim = np.tile(np.array([-0.8, 0.8], np.float32), (m.problems, 64, 1))
    bt.init_position(13, lim)
    tuned = bt.autotune(2)
    print("autotune", tuned, flush=True)
    for s in ("32", "64"):
        assert tuned[s]["unit_rows"] in (1024, 2048, 4096, 8192, 16384)
        assert tuned[s]["col_blocks"] in ((4, 8) if s == "32" else (4,))
        assert tuned[s]["ms"] > 0.0
    xs, lp = bt.state()
    for p in range(m.problems):
        ref = O.glm_logpost(m.dequantized(p), m.y[p], 5.0, xs[p][:64])
        err = np.abs(lp[p, :64] - ref) / (np.abs(ref) + 1.0)
        assert err.max() < 2e-6, (p, err.max())
    # and it keeps sampling: a few accounted steps after the rebuild
    res = bt.run(31, 4, 2.0)
    assert np.all(res["acceptance-rate"] > 0.0)
    bt.release()
