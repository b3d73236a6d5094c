"""Stretch sampler through the C-ABI on the GPU, against the reference's
goldens and the oracle."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models

pytestmark = pytest.mark.gpu

W_TEST = 44 * 256  # amd_gcn_test.clj:162
STEPWISE, RESIDENT = _lib.BH_EXEC_STEPWISE, _lib.BH_EXEC_RESIDENT


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


@pytest.fixture(scope="module")
def uniform_sf(fac):
    return fac.mcmc_factory(models.uniform_model())


@pytest.fixture(scope="module")
def gaussian_sf(fac):
    return fac.mcmc_factory(models.gaussian_model())


def _oracle(model_id, walkers, params, seed, params_len=2, dim=1, partner=O.PARTNER_FLAT):
    s = O.Stretch(model_id, dim, walkers, params, seed, params_len=params_len, partner=partner,
                  threads=4)
    s.init(seed)
    return s


def test_uniform_direct_sampler_golden(fac, goldens):
    g = goldens["G1_uniform_direct_sampler"]
    ds = fac.direct_sampler_engine(models.uniform_model())
    x = ds.sample(g["seed"], g["params"], g["n"])
    assert x[:4].tolist() == g["first4"]
    assert x[9996:].tolist() == g["at9996"]
    assert float(x.max()) == g["max"] and float(x.min()) == g["min"]
    assert np.array_equal(x, O.uniform_sample(g["seed"], *g["params"], g["n"]))


def test_uniform_stretch_goldens(uniform_sf, goldens):
    s = uniform_sf.create_sampler(123, W_TEST, [-1, 2])
    s.init(123)
    s.init_position(123, [-1, 2])
    ref = _oracle(O.UNIFORM_LOGPDF, W_TEST, [-1, 2], 123)
    ref.init_position(123, [-1, 2])
    for row in goldens["G3_uniform_stretch_sample"]["rows"]:
        x = s.sample()
        assert x[:4, 0].tolist() == row
        assert np.array_equal(x, ref.sample())


def test_manual_bare_pairs_and_accu_golden(uniform_sf, goldens):
    # amd_gcn_test.clj:210-254 through move-bare!, init-move! and move!
    rows = goldens["G3_uniform_stretch_sample"]["rows"]
    g = goldens["G5_stretch_accu"]
    s = uniform_sf.create_sampler(123, W_TEST, [-1, 2])
    s.init(123)
    s.init_position(123, [-1, 2])
    s.move_bare()
    assert s.state()[0][:4, 0].tolist() == rows[0]
    s.move_bare()
    assert s.state()[0][:4, 0].tolist() == rows[1]
    xs, lf, c = s.state()
    c[1] = g["seed_odd"] - 2  # init-move! adds 2 to the move seed (amd_gcn.clj:330)
    s.set_state(counters=c)
    s.init_move(g["a"], 1)
    s.move()
    assert s.state()[0][:4, 0].tolist() == g["xs_first4"]
    accept, means = s.read_accounting(1)
    assert means[:10].tolist() == np.float32(g["means_acc_first10"]).tolist()
    assert accept[:10].tolist() == g["accept_first10"]
    assert float(O.tree_sum_f32(means)) == pytest.approx(g["sum_acc"], abs=1e-3)


def test_gaussian_stretch_goldens(gaussian_sf, goldens):
    g = goldens["G6_gaussian_stretch_sample"]
    s = gaussian_sf.create_sampler(g["seed"], g["walkers"], g["params"])
    s.init(g["seed"])
    s.init_position(g["seed"], g["limits"])
    for row in g["rows"]:
        assert s.sample()[:4, 0].tolist() == row


def test_gaussian_burn_in_vs_oracle(gaussian_sf, goldens):
    g = goldens["G7_gaussian_burn_in"]
    s = gaussian_sf.create_sampler(123, W_TEST, g["params"])
    s.init(g["init_seed"])
    s.init_position(g["position_seed"], g["limits"])
    s.burn_in(g["burn_in"], g["a"])
    x = s.sample()[:, 0]
    ref = _oracle(O.GAUSSIAN_MCMC_LOGPDF, W_TEST, g["params"], 123)
    ref.init(g["init_seed"])
    ref.init_position(g["position_seed"], g["limits"])
    ref.burn_in(g["burn_in"], g["a"])
    r = ref.sample()[:, 0]
    # exp() in the accept test differs from the oracle's in the last bits;
    # a flipped decision decouples that walker's trajectory
    same = np.mean(x == r)
    assert same > 0.99, same
    assert float(x.astype(np.float64).mean()) == pytest.approx(g["mean"], abs=0.03)
    assert float(x.astype(np.float64).std()) == pytest.approx(g["sd"], abs=0.03)


def test_engine_statistics(fac, goldens):
    # amd_gcn_test.clj:293-312
    g = goldens["G8_engine_statistics"]
    sf = fac.mcmc_factory(models.gaussian_model())
    s = sf.create_sampler(g["seed"], g["walkers"], g["params"])
    s.init(g["seed"])
    s.init_position(g["seed"], g["limits"])
    s.burn_in(g["burn_in"], g["a"])
    assert s.acc_rate(g["a"]) == pytest.approx(g["acc_rate"], abs=g["acc_rate_tol"])
    res = s.run_sampler(g["run_steps"], g["a"])
    assert float(res["autocorrelation"].tau[0]) == pytest.approx(g["tau"], abs=g["tau_tol"])
    assert s.info()["iteration-counter"] == g["burn_in"] + 1 + g["run_steps"]


@pytest.mark.parametrize("exec_mode", [STEPWISE, RESIDENT])
def test_uniform_run_sampler_exact_vs_oracle(uniform_sf, exec_mode):
    """Uniform model: accept decisions are exact, so accept counts, per-step
    means and their acor match the oracle bit-for-bit."""
    W = 1024
    s = uniform_sf.create_sampler(7, W, [-1, 2]).set_exec(exec_mode)
    s.init(7)
    s.init_position(9, [-1, 2])
    s.burn_in(30, 2.0)
    res = s.run_sampler(200, 2.0)
    ref = _oracle(O.UNIFORM_LOGPDF, W, [-1, 2], 7)
    ref.init_position(9, [-1, 2])
    ref.burn_in(30, 2.0)
    rr = ref.run_sampler(200, 2.0)
    assert np.array_equal(res["means"], rr["means"])
    assert res["acceptance-rate"] == rr["acceptance_rate"]
    a, b = res["autocorrelation"], rr["autocorrelation"]
    assert np.array_equal(a.tau, b.tau) and np.array_equal(a.sigma, b.sigma)
    assert np.array_equal(a.mean, b.mean) and a.lag == b.lag
    assert np.array_equal(s.state()[0].reshape(-1), ref.xs)


def _mvn_params(dim, seed=1001):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(dim).astype(np.float32)
    A = rng.standard_normal((dim, dim))
    P = np.linalg.inv(A @ A.T + np.eye(dim)).astype(np.float32)
    return np.concatenate([mu, P.reshape(-1)]).astype(np.float32)


# DIM 4: one 16-byte piece per record, a record per thread; DIM 9 and 16:
# records staged through LDS (float pieces, 16-byte pieces)
@pytest.mark.parametrize("dim,walkers", [(4, 1024), (9, 512), (16, 512)])
@pytest.mark.parametrize("partner", [_lib.BH_PARTNER_FLAT, _lib.BH_PARTNER_WALKER])
def test_resident_equals_stepwise_mvn(fac, partner, dim, walkers):
    """Correlated Gaussian: the resident kernel (one record per thread, LDS
    resident ensemble) and the two-kernels-per-step path produce identical
    trajectories, accounting and samples."""
    params = _mvn_params(dim)
    sf = fac.mcmc_factory(models.mvn_model(dim))
    out = []
    for mode in (STEPWISE, RESIDENT):
        s = sf.create_sampler(123, walkers, params).set_exec(mode).set_partner(partner)
        s.init(123)
        s.init_position(123, np.tile([-5.0, 5.0], dim))
        s.burn_in(64, 2.0)
        r = s.run_sampler(256, 2.0)
        smp = s.sample(3000)
        out.append((s.state()[0], r["means"], r["acceptance-rate"], smp))
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(x), np.asarray(y))


@pytest.mark.parametrize("dim", [4, 9, 16])
@pytest.mark.parametrize("partner", [O.PARTNER_FLAT, O.PARTNER_WALKER])
def test_mvn_steps_vs_oracle(fac, partner, dim):
    params = _mvn_params(dim)
    sf = fac.mcmc_factory(models.mvn_model(dim))
    s = sf.create_sampler(5, 2048, params).set_partner(partner).set_exec(STEPWISE)
    s.init(5)
    s.init_position(6, np.tile([-5.0, 5.0], dim))
    ref = O.Stretch(O.MVN_LOGPDF, dim, 2048, params, 5, params_len=dim + dim * dim,
                    partner=partner, threads=4)
    ref.init(5)
    ref.init_position(6, np.tile(np.array([[-5.0, 5.0]], np.float32), (dim, 1)))
    assert np.array_equal(s.state()[0].reshape(-1), ref.xs)
    s.burn_in(5, 2.0)
    ref.burn_in(5, 2.0)
    x = s.state()[0].reshape(-1, dim)
    r = ref.xs.reshape(-1, dim)
    same = np.all(x == r, axis=1).mean()
    assert same > 0.995, same
    close = np.abs(x - r) <= 1e-5 * np.maximum(1.0, np.abs(r))
    assert np.all(close[np.all(x == r, axis=1)])


# DIM 32: records of eight 16-byte pieces; DIM 33: 33 float pieces
@pytest.mark.parametrize("dim", [32, 33])
@pytest.mark.parametrize("partner", [O.PARTNER_FLAT, O.PARTNER_WALKER])
def test_diag_gaussian_steps_vs_oracle(fac, partner, dim):
    rng = np.random.default_rng(dim)
    mu = rng.standard_normal(dim).astype(np.float32)
    w = rng.uniform(0.25, 4.0, dim).astype(np.float32)
    params = np.concatenate([mu, w])
    sf = fac.mcmc_factory(models.diag_gaussian_model(dim))
    s = sf.create_sampler(9, 2048, params).set_partner(partner).set_exec(STEPWISE)
    s.init(9)
    s.init_position(10, np.tile([-3.0, 3.0], dim))
    ref = O.Stretch(O.DIAG_GAUSS_LOGPDF, dim, 2048, params, 9, params_len=2 * dim,
                    partner=partner, threads=4)
    ref.init(9)
    ref.init_position(10, np.tile(np.array([[-3.0, 3.0]], np.float32), (dim, 1)))
    s.burn_in(6, 2.0)
    ref.burn_in(6, 2.0)
    x = s.state()[0].reshape(-1, dim)
    r = ref.xs.reshape(-1, dim)
    same = np.all(x == r, axis=1).mean()
    assert same > 0.995, same
    r_run = s.run_sampler(64, 2.0)
    ref_run = ref.run_sampler(64, 2.0)
    assert abs(r_run["acceptance-rate"] - ref_run["acceptance_rate"]) < 0.01


def test_staged_move_large_ensemble_vs_oracle(fac):
    """The staged move kernels over 2^18 walkers (512 work-groups per half)
    of the 16-D correlated Gaussian: the same moves as the oracle for all
    but the few walkers whose accept test sits on the last bit of exp."""
    dim, W = 16, 1 << 18
    params = _mvn_params(dim, 7)
    sf = fac.mcmc_factory(models.mvn_model(dim))
    s = sf.create_sampler(21, W, params).set_partner(_lib.BH_PARTNER_WALKER).set_exec(STEPWISE)
    s.init(21)
    s.init_position(22, np.tile([-5.0, 5.0], dim))
    ref = O.Stretch(O.MVN_LOGPDF, dim, W, params, 21, params_len=dim + dim * dim,
                    partner=O.PARTNER_WALKER, threads=8)
    ref.init(21)
    ref.init_position(22, np.tile(np.array([[-5.0, 5.0]], np.float32), (dim, 1)))
    s.burn_in(3, 2.0)
    ref.burn_in(3, 2.0)
    x = s.state()[0].reshape(-1, dim)
    r = ref.xs.reshape(-1, dim)
    assert np.all(x == r, axis=1).mean() > 0.999


def test_anneal_vs_oracle(gaussian_sf):
    sched = lambda i: 10.0 - i  # noqa: E731  (mcmc.clj minus-n)
    s = gaussian_sf.create_sampler(11, 1024, [0.0, 1.0])
    s.init(11)
    s.init_position(12, [-3, 3])
    s.anneal(sched, 9, 2.0)
    s.move_bare()  # keeps the last temperature, as the reference does
    ref = _oracle(O.GAUSSIAN_MCMC_LOGPDF, 1024, [0.0, 1.0], 11)
    ref.init_position(12, [-3, 3])
    ref.anneal(sched, 9, 2.0)
    ref.move_bare()
    assert np.mean(s.state()[0].reshape(-1) == ref.xs) > 0.99


def test_histogram_and_moments_vs_oracle(gaussian_sf):
    s = gaussian_sf.create_sampler(3, 2048, [1.0, 2.0])
    s.init(3)
    s.init_position(4, [-5, 5])
    ref = _oracle(O.GAUSSIAN_MCMC_LOGPDF, 2048, [1.0, 2.0], 3)
    ref.init_position(4, [-5, 5])
    h = s.histogram(3)
    hr = ref.histogram(3)
    assert np.array_equal(h.limits, hr.limits)
    # identical trajectories -> identical counts -> identical pdf and ranks
    if np.array_equal(s.state()[0].reshape(-1), ref.xs):
        assert np.array_equal(h.pdf, hr.pdf)
        assert np.array_equal(h.bin_ranks, hr.bin_ranks)
    m, v = s.mean(), s.variance()
    xs = s.state()[0].astype(np.float64)
    assert np.allclose(m, xs.mean(0), rtol=1e-6, atol=1e-7)
    assert np.allclose(v, xs.var(0), rtol=1e-5)
    assert s.info()["iteration-counter"] == 2


def test_walker_count_rule(uniform_sf):
    with pytest.raises(_lib.ArgumentError, match="multiple of 512"):
        uniform_sf.create_sampler(1, 768, [-1, 2])


def test_run_sampler_too_short(uniform_sf):
    s = uniform_sf.create_sampler(1, 512, [-1, 2])
    s.init_position(1, [-1, 2])
    with pytest.raises(_lib.ArgumentError, match="must not be less than 50"):
        s.run_sampler(20, 2.0)


def test_state_resume_reproduces_trajectory(gaussian_sf):
    a = gaussian_sf.create_sampler(21, 1024, [0.0, 1.0])
    a.init(21)
    a.init_position(22, [-2, 2])
    a.burn_in(10, 2.0)
    xs, lf, c = a.state()
    b = gaussian_sf.create_sampler(99, 1024, [0.0, 1.0])
    b.set_state(xs, lf, c)
    a.burn_in(10, 2.0)
    b.burn_in(10, 2.0)
    assert np.array_equal(a.state()[0], b.state()[0])
    c2 = gaussian_sf.create_sampler(5, 1024, [0.0, 1.0])
    c2.init_position(a)
    assert np.array_equal(c2.state()[0], a.state()[0])


def test_device_tensors_in_place(fac, gaussian_sf):
    torch = pytest.importorskip("torch")
    s = gaussian_sf.create_sampler(31, 1024, [0.0, 1.0])
    s.init_position(32, [-2, 2])
    out = torch.empty((2048, 1), dtype=torch.float32, device="cuda")
    s.sample(2048, out=out)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all().item()
    m = fac.dataset_engine().data_mean(out)
    assert abs(float(m[0]) - float(out.double().mean())) < 1e-5


@pytest.mark.parametrize("name,kind", [("gaussian", 1), ("exponential", 2), ("erlang", 3)])
def test_direct_samplers_vs_oracle_and_goldens(fac, goldens, name, kind):
    g = goldens["G2_direct_samplers"][name]
    model = {"gaussian": models.gaussian_model, "exponential": models.exponential_model,
             "erlang": models.erlang_model}[name]()
    x = fac.direct_sampler_engine(model).sample(123, g["params"], 10000)
    r = O.direct_sample(kind, 123, g["params"], 10000)
    # sin/cos/log of the device and of the C library differ by an ulp; near
    # a zero of sin that is an absolute error of ~1e-7 * scale
    scale = float(g["params"][1]) if name == "gaussian" else 1.0
    assert np.allclose(x, r, rtol=2e-6, atol=2e-6 * scale)
    assert np.allclose(x[:4], g["first4"], rtol=2e-6, atol=0)
    assert np.allclose(x[9996:], g["at9996"], rtol=2e-6, atol=0)


def test_linreg_user_model_vs_oracle(fac):
    """The linear-regression posterior as a user model in the reference's
    dialect (the plugin path, 17 dimensions, one thread per walker over the
    rows): the same moves as the oracle's restatement, up to accept tests
    that sit on the last bit of the model's exp."""
    rng = np.random.default_rng(31)
    N, d = 3000, 16
    X = rng.standard_normal((N, d))
    y = X @ rng.standard_normal(d) + 0.5 * rng.standard_normal(N)
    rows = np.concatenate([X, y[:, None]], axis=1).astype(np.float32)
    params = np.concatenate([rows.reshape(-1), [10.0, 1.0]]).astype(np.float32)
    lik = models.likelihood_model(models.LINREG_SOURCE, name="linreg")
    prior = models.distribution_model(models.LINREG_SOURCE, name="linreg_prior",
                                      dimension=d + 1, params_size=2)
    sf = fac.mcmc_factory(models.posterior_model(prior, "linreg_post", lik))
    W = 1024
    limits = np.tile(np.array([[-2.0, 2.0]], np.float32), (d + 1, 1))
    s = sf.create_sampler(41, W, params).set_partner(_lib.BH_PARTNER_WALKER)
    s.init(41)
    s.init_position(42, limits.reshape(-1))
    ref = O.Stretch(O.LINREG_POST, d + 1, W, params, 41, data_len=N * (d + 1), params_len=2,
                    partner=O.PARTNER_WALKER, threads=4)
    ref.init(41)
    ref.init_position(42, limits)
    x0, lf0 = s.state()[0].reshape(-1, d + 1), s.state()[1]
    assert np.array_equal(x0.reshape(-1), ref.xs)
    assert np.allclose(lf0, ref.logfn, rtol=2e-5)
    s.burn_in(4, 2.0)
    ref.burn_in(4, 2.0)
    x = s.state()[0].reshape(-1, d + 1)
    r = ref.xs.reshape(-1, d + 1)
    assert np.all(x == r, axis=1).mean() > 0.99
