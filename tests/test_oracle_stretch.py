"""Pins init_walkers, the stretch move and its accounting in the oracle."""
import numpy as np
import pytest

import oracle as O

W_TEST = 44 * 256  # amd_gcn_test.clj:162


def _uniform(seed=123):
    s = O.Stretch(O.UNIFORM_LOGPDF, 1, W_TEST, [-1, 2], seed, params_len=2)
    s.init(seed)
    s.init_position(seed, [-1, 2])
    return s


def test_uniform_stretch_sample_goldens(goldens):
    s = _uniform()
    for row in goldens["G3_uniform_stretch_sample"]["rows"]:
        assert s.sample()[:4, 0].tolist() == row


def test_manual_bare_half_steps_and_accu(goldens):
    # amd_gcn_test.clj:210-254: two bare pairs (steps 0 and 1), then one accu
    # pair with seeds 123/124, constants 1111/2222, step 0
    rows = goldens["G3_uniform_stretch_sample"]["rows"]
    s = _uniform()
    s.move_bare()
    assert s.xs[:4].tolist() == rows[0]
    s.move_bare()
    assert s.xs[:4].tolist() == rows[1]
    g = goldens["G5_stretch_accu"]
    groups = s.K // 256
    means_acc = np.zeros(groups, np.float32)
    s.move_seed = g["seed_odd"] - 2  # init-move! adds 2 (amd_gcn.clj:330)
    s.init_move(means_acc, g["a"])
    s.move()
    assert s.xs[:4].tolist() == g["xs_first4"]
    assert means_acc[:10].tolist() == np.float32(g["means_acc_first10"]).tolist()
    assert s.accept[:10].tolist() == g["accept_first10"]
    # sum_means_vertical over the 22 groups; the reference then sums on the
    # device with a tree of its own choice, hence a 2-ulp tolerance
    assert float(O.tree_sum_f32(means_acc)) == pytest.approx(g["sum_acc"], abs=2 * 4.9e-4)


def test_gaussian_stretch_sample_goldens(goldens):
    g = goldens["G6_gaussian_stretch_sample"]
    s = O.Stretch(O.GAUSSIAN_MCMC_LOGPDF, 1, g["walkers"], g["params"], g["seed"], params_len=2)
    s.init(g["seed"])
    s.init_position(g["seed"], g["limits"])
    for row in g["rows"]:
        assert s.sample()[:4, 0].tolist() == row


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_planted_contraction_mistakes_fail_goldens(goldens, variant):
    """Planting an un-fused z (bit 0) or Y (bit 1) in the oracle must break
    the uniform or Gaussian goldens: the contraction choice is pinned."""
    try:
        O.set_variant(variant)
        s = _uniform()
        got = [s.sample()[:4, 0].tolist() for _ in range(4)]
        g = goldens["G6_gaussian_stretch_sample"]
        t = O.Stretch(O.GAUSSIAN_MCMC_LOGPDF, 1, g["walkers"], g["params"], g["seed"], params_len=2)
        t.init_position(g["seed"], g["limits"])
        got2 = [t.sample()[:4, 0].tolist() for _ in range(3)]
    finally:
        O.set_variant(0)
    assert got != goldens["G3_uniform_stretch_sample"]["rows"] or got2 != g["rows"]


def test_walker_count_rule():
    # amd_gcn.clj:516, 566-567
    with pytest.raises(ValueError, match="multiple of 512"):
        O.Stretch(O.UNIFORM_LOGPDF, 1, 768, [-1, 2], 1, params_len=2)
    with pytest.raises(ValueError):
        O.Stretch(O.UNIFORM_LOGPDF, 1, 256, [-1, 2], 1, params_len=2)


def test_init_walkers_layout_and_limits():
    dim, W = 3, 512
    lim = np.array([[-1, 1], [10, 20], [-5, -4]], np.float32)
    xs = O.init_walkers(7, lim, dim, W).reshape(W, dim)
    for d in range(dim):
        assert (xs[:, d] > lim[d, 0]).all() and (xs[:, d] < lim[d, 1]).all()
    # coordinate i of the flat array takes limits row i % dim (stretch.cl:143-146)
    r = O.philox4x32_10([0, 0xF00DCAFE, 0xDEADBEEF, 0xBEEFF00D], [7, 0xDECAFAAA])
    u = np.float32(O.u01(int(r[1])))
    expect = np.float32(np.float64(u) * 20.0 + np.float64(np.float32((np.float32(1) - u) * np.float32(10))))
    assert xs[0, 1] == expect


def test_partner_modes_agree_for_dim1():
    a = O.Stretch(O.GAUSSIAN_MCMC_LOGPDF, 1, 1024, [0, 1], 5, params_len=2, partner=O.PARTNER_FLAT)
    b = O.Stretch(O.GAUSSIAN_MCMC_LOGPDF, 1, 1024, [0, 1], 5, params_len=2, partner=O.PARTNER_WALKER)
    for s in (a, b):
        s.init_position(5, [-3, 3])
        s.burn_in(20, 2.0)
    assert np.array_equal(a.xs, b.xs)


def test_gaussian_burn_in_statistics(goldens):
    # amd_gcn_test.clj:274-291: trajectory-dependent values, checked loosely;
    # the exact path depends on the device's native_exp in the accept test
    g = goldens["G7_gaussian_burn_in"]
    s = O.Stretch(O.GAUSSIAN_MCMC_LOGPDF, 1, W_TEST, g["params"], 123, params_len=2)
    s.init(g["init_seed"])
    s.init_position(g["position_seed"], g["limits"])
    s.burn_in(g["burn_in"], g["a"])
    x = s.sample()[:, 0].astype(np.float64)
    assert x.mean() == pytest.approx(g["mean"], abs=0.03)
    assert x.std() == pytest.approx(g["sd"], abs=0.03)


def test_mvn_closed_form_posterior_walker_mode():
    """4-D correlated Gaussian: ensemble moments converge to mu and diag(Sigma)
    (the model's closed form) with walker-aligned partners."""
    rng = np.random.default_rng(1001)
    mu = rng.standard_normal(4).astype(np.float32)
    A = rng.standard_normal((4, 4))
    Sigma = A @ A.T + np.eye(4)
    P = np.linalg.inv(Sigma).astype(np.float32)
    params = np.concatenate([mu, P.reshape(-1)]).astype(np.float32)
    s = O.Stretch(O.MVN_LOGPDF, 4, 1024, params, 123, params_len=20,
                  partner=O.PARTNER_WALKER, threads=4)
    s.init_position(123, np.tile(np.array([[-5, 5]], np.float32), (4, 1)))
    s.burn_in(400, 2.0)
    draws = np.concatenate([s.sample() for _ in range(200)]).astype(np.float64)
    sd = np.sqrt(np.diag(Sigma))
    assert np.all(np.abs(draws.mean(0) - mu) < 0.25 * sd)
    assert np.all(np.abs(draws.std(0) / sd - 1) < 0.15)


def test_diag_gaussian_model_closed_form():
    """Independent Gaussian coordinates: log density 0 at the mean, -1/2 one
    standard deviation away along any axis, and -sum (x-mu)^2/(2 sigma^2)
    at random points."""
    dim = 32
    rng = np.random.default_rng(7)
    mu = rng.standard_normal(dim).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, dim).astype(np.float32)
    w = (1.0 / (sd.astype(np.float64) ** 2)).astype(np.float32)
    params = np.concatenate([mu, w])
    assert O.logfn_many(O.DIAG_GAUSS_LOGPDF, params, mu, dim)[0] == 0.0
    axes = np.tile(mu, (dim, 1)) + np.diag(sd)
    lp = O.logfn_many(O.DIAG_GAUSS_LOGPDF, params, axes, dim)
    assert np.allclose(lp, -0.5, rtol=1e-5)
    x = (mu + 3 * rng.standard_normal((100, dim))).astype(np.float32)
    want = -0.5 * np.sum(w.astype(np.float64) * (x.astype(np.float64) - mu) ** 2, axis=1)
    got = O.logfn_many(O.DIAG_GAUSS_LOGPDF, params, x, dim)
    assert np.allclose(got, want, rtol=1e-5)


def test_hbb_model_against_its_definition():
    """Hierarchical beta-binomial on logits: the oracle's density equals,
    up to a constant, binomial likelihoods x Beta(a, b) priors on theta_g x
    the logit Jacobian theta (1 - theta) x N(0, s0^2) priors on log a and
    log b, built from scipy's distributions in float64 (differences between
    points, so the dropped constants cancel)."""
    from scipy import special, stats
    rng = np.random.default_rng(11)
    G, s0 = 12, 10.0
    n = np.round(np.exp(rng.uniform(np.log(10), np.log(1e4), G)))
    y = rng.binomial(n.astype(np.int64), rng.beta(2, 5, G)).astype(np.float64)
    params = np.concatenate([n, y, [s0]]).astype(np.float32)

    def ref(x):
        la, lb, t = x[0], x[1], x[2:]
        a, b = np.exp(la), np.exp(lb)
        theta = special.expit(t)
        lp = stats.binom.logpmf(y, n, theta).sum()
        lp += stats.beta.logpdf(theta, a, b).sum()
        lp += np.sum(np.log(theta) + np.log1p(-theta))
        return lp + stats.norm.logpdf(la, 0, s0) + stats.norm.logpdf(lb, 0, s0)

    t0 = np.log((y + 0.5) / (n - y + 0.5))
    xs = np.array([np.concatenate([[rng.uniform(0, 1.5), rng.uniform(0.5, 2.0)],
                                   t0 + 0.2 * rng.standard_normal(G)]) for _ in range(20)],
                  np.float32)
    got = O.logfn_many(O.HBB_LOGPDF, params, xs, G + 2, data_len=2 * G,
                       params_len=1).astype(np.float64)
    want = np.array([ref(x.astype(np.float64)) for x in xs])
    assert np.allclose(got - got[0], want - want[0], rtol=1e-5, atol=0.01)


def test_linreg_posterior_against_its_definition():
    """Gaussian linear regression with normal priors: the oracle's density
    equals, up to a constant, scipy's normal log-densities of the residuals
    (sd sigma) plus N(0, s_beta^2) on each coefficient and N(0, s_sigma^2)
    on log sigma, in float64 (differences between points)."""
    from scipy import stats
    rng = np.random.default_rng(12)
    N, d, sb, ssig = 400, 6, 10.0, 1.0
    X = rng.standard_normal((N, d))
    y = X @ rng.standard_normal(d) + 0.5 * rng.standard_normal(N)
    rows = np.concatenate([X, y[:, None]], axis=1).astype(np.float32)
    params = np.concatenate([rows.reshape(-1), [sb, ssig]]).astype(np.float32)
    Xf, yf = rows[:, :d].astype(np.float64), rows[:, d].astype(np.float64)

    def ref(x):
        beta, ls = x[:d], x[d]
        lp = stats.norm.logpdf(yf, Xf @ beta, np.exp(ls)).sum()
        return lp + stats.norm.logpdf(beta, 0, sb).sum() + stats.norm.logpdf(ls, 0, ssig)

    xs = np.concatenate([rng.standard_normal((20, d)), rng.uniform(-1.5, 0.5, (20, 1))],
                        axis=1).astype(np.float32)
    got = O.logfn_many(O.LINREG_POST, params, xs, d + 1, data_len=N * (d + 1),
                       params_len=2).astype(np.float64)
    want = np.array([ref(x.astype(np.float64)) for x in xs])
    assert np.allclose(got - got[0], want - want[0], rtol=1e-5, atol=0.05)


# --------------------------------------------------------------------------
# The logistic-regression posterior (the checker of the GLM samplers)
# --------------------------------------------------------------------------
def _logistic_problem(seed=21, N=600, d=8):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((N, d)) / np.sqrt(d)).astype(np.float32)
    beta_true = rng.standard_normal(d)
    y = (rng.uniform(size=N) < 1.0 / (1.0 + np.exp(-(X.astype(np.float64) @ beta_true))))
    pts = (beta_true + 0.5 * rng.standard_normal((16, d))).astype(np.float32)
    return X, y.astype(np.float32), pts


def _logistic_scipy(X, y, s, pts):
    """Bernoulli log-likelihood of sigmoid(X beta) plus N(0, s^2) log-priors,
    from scipy's distributions in float64 (normalising constants included;
    compare differences between points)."""
    from scipy import special, stats
    Xd = X.astype(np.float64)
    out = []
    for b in pts.astype(np.float64):
        out.append(stats.bernoulli.logpmf(y.astype(np.int64), special.expit(Xd @ b)).sum()
                   + stats.norm.logpdf(b, 0.0, s).sum())
    return np.array(out)


def test_glm_logpost_against_scipy():
    X, y, pts = _logistic_problem()
    got = O.glm_logpost(X, y, 5.0, pts)
    want = _logistic_scipy(X, y, 5.0, pts)
    assert np.allclose(got - got[0], want - want[0], rtol=0, atol=1e-9)


def test_glm_logpost_closed_forms():
    X, y, pts = _logistic_problem(N=512)
    # beta = 0: every row contributes -ln 2 and the prior nothing
    zero = O.glm_logpost(X, y, 5.0, np.zeros((1, X.shape[1]), np.float32))[0]
    assert zero == pytest.approx(-X.shape[0] * np.log(2.0), rel=1e-14)
    # label flip: (y, X) -> (1 - y, -X) leaves every row's term unchanged
    a = O.glm_logpost(X, y, 5.0, pts)
    b = O.glm_logpost(-X, 1.0 - y, 5.0, pts)
    assert np.allclose(a, b, rtol=1e-13, atol=0)
    # zero data: only the prior -|beta|^2 / (2 s^2) remains beside -N ln 2
    Z = np.zeros_like(X)
    got = O.glm_logpost(Z, y, 2.0, pts)
    want = -X.shape[0] * np.log(2.0) - 0.125 * np.sum(pts.astype(np.float64) ** 2, axis=1)
    assert np.allclose(got, want, rtol=1e-14, atol=0)


@pytest.mark.parametrize("variant", [4, 8])
def test_glm_logpost_planted_mistakes_fail(variant):
    """A dropped prior (bit 2) or a sign-flipped linear predictor (bit 3)
    planted in the oracle must fail the scipy comparison above."""
    X, y, pts = _logistic_problem()
    want = _logistic_scipy(X, y, 5.0, pts)
    O.set_variant(variant)
    try:
        got = O.glm_logpost(X, y, 5.0, pts)
    finally:
        O.set_variant(0)
    assert not np.allclose(got - got[0], want - want[0], rtol=0, atol=1e-3)


def test_glm_stretch_zero_data_moves_as_gaussian_stretch():
    """With every row of X zero the logistic posterior is N(0, s^2 I): the GLM
    sampler (log posteriors kept in double) takes the moves of the plain
    sampler on the independent-Gaussian model from the same seeds (its float
    log densities can flip a decision only at a rounding tie)."""
    d, W, s, seed = 3, 1024, 1.0, 77
    Z = np.zeros((16, d), np.float32)
    y = np.ones(16, np.float32)
    g = O.GlmStretch(Z, y, s, W, seed, partner=O.PARTNER_WALKER)
    g.init(seed)
    lim = np.tile([-2.0, 2.0], d)
    g.init_position(seed, lim)
    params = np.concatenate([np.zeros(d), np.full(d, 1.0 / s ** 2)]).astype(np.float32)
    p = O.Stretch(O.DIAG_GAUSS_LOGPDF, d, W, params, seed, params_len=2 * d,
                  partner=O.PARTNER_WALKER)
    p.init(seed)
    p.init_position(seed, lim)
    assert np.array_equal(g.xs, p.xs)
    for _ in range(20):
        g.move_bare()
        p.move_bare()
    same = np.all(g.xs.reshape(W, d) == p.xs.reshape(W, d), axis=1)
    assert same.mean() >= 0.999


def test_glm_stretch_zero_data_moments():
    """The zero-data GLM posterior N(0, s^2 I): means within 5 MC standard
    errors of 0 and sds within 5% of s over a long run."""
    d, W, s, seed = 2, 512, 1.0, 91
    Z = np.zeros((4, d), np.float32)
    g = O.GlmStretch(Z, np.zeros(4, np.float32), s, W, seed, partner=O.PARTNER_WALKER,
                     threads=4)
    g.init(seed)
    g.init_position(seed, np.tile([-1.0, 1.0], d))
    g.burn_in(200, 2.0)
    draws = np.concatenate([g.sample(W) for _ in range(400)])
    m, sd = draws.mean(0), draws.std(0)
    # ~20 steps of autocorrelation: effective draws ~ W * 400 / 20
    mcse = s / np.sqrt(W * 400 / 20)
    assert np.all(np.abs(m) < 5 * mcse), (m, mcse)
    assert np.all(np.abs(sd - s) < 0.05 * s), sd


def _logistic_rows(X, y, s):
    rows = np.concatenate([X, y[:, None]], axis=1).astype(np.float32)
    return np.concatenate([rows.reshape(-1), [s]]).astype(np.float32)


def test_logistic_user_model_against_glm_logpost():
    """The logistic posterior as a reference-style fp32 user model (row loop
    inside LOGFN) agrees with the fp64 GLM posterior, up to the prior's
    constant and fp32 rounding."""
    X, y, pts = _logistic_problem(N=600, d=8)
    params = _logistic_rows(X, y, 5.0)
    got = O.logfn_many(O.LOGISTIC_POST, params, pts, 8, data_len=600 * 9,
                       params_len=1).astype(np.float64)
    want = O.glm_logpost(X, y, 5.0, pts)
    assert np.allclose(got, want, rtol=2e-6, atol=1e-3)


def test_logistic_blocked_half_equals_per_walker_half():
    """The walker-blocked half-step (the timed CPU baseline) takes exactly
    the per-walker half-step's decisions, positions and log posteriors, with
    a partial last block."""
    X, y, _ = _logistic_problem(N=300, d=6)
    params = _logistic_rows(X, y, 5.0)
    K, d = 200, 6  # one full block of 128 walkers and a partial one
    rng = np.random.default_rng(4)
    xs = (0.3 * rng.standard_normal(2 * K * d)).astype(np.float32)
    lf = O.logfn_many(O.LOGISTIC_POST, params, xs.reshape(2 * K, d), d, data_len=300 * 7,
                      params_len=1).astype(np.float32)
    out = []
    for blocked in (False, True):
        x, l = xs.copy(), lf.copy()
        acc = np.zeros(K, np.uint8)
        L = O.lib()
        if blocked:
            n = L.bho_logistic_half_blocked(9, 3333, 2, K, d, x.ctypes.data + 4 * K * d,
                                            x.ctypes.data, l.ctypes.data, 2.0, 1.0,
                                            O.PARTNER_WALKER, 300 * 7, params, acc.ctypes.data,
                                            3)
        else:
            n = L.bho_stretch_half(9, 3333, 2, K, d, x.ctypes.data + 4 * K * d, x.ctypes.data,
                                   l.ctypes.data, 2.0, 1.0, O.PARTNER_WALKER, O.LOGISTIC_POST,
                                   300 * 7, 1, params, acc.ctypes.data, 3)
        out.append((n, x, l, acc))
    assert out[0][0] == out[1][0] and 0 < out[0][0] < K
    for a, b in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(a, b)
