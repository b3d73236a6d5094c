"""CPU oracle for the stretch-sampler hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's OpenCL
kernels (``bh_oracle.c``) and the host orchestration of
``src/clojure/uncomplicate/bayadera/internal/device/amd_gcn.clj`` on the CPU.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or, for ``cpu_baseline``,
as the timed CPU restatement).  The HIP engine in ``bayadera_amd`` never
imports it.

Pins (each checked by ``tests/test_oracle_*.py``):

* Philox4x32-10 -- Random123 known-answer vectors (published with Random123;
  cross-checked against rocRAND's MIT-licensed host implementation) and the
  reference's uniform direct-sampler golden (``amd_gcn_test.clj:40-54``).
* u01, init_walkers, stretch_move (DIM=1) -- exact goldens
  ``amd_gcn_test.clj:196-272``.
* stretch_move_accu accounting -- exact accept counts and means accumulators
  ``amd_gcn_test.clj:233-254``.
* acor -- exact fixture goldens ``amd_gcn_test.clj:314-353``.
* histogram / min-max / uint_to_real / bitonic ranks -- brute force on small
  inputs and the invariants sum(counts) = n, integral(pdf) = 1.
* mean / variance -- closed forms and numpy float64.
* Stretch moves at DIM > 1 -- the reference reads out of bounds there
  (``amd-gcn-mcmc-stretch.cl:57-65``); the oracle defines the read as a wrap
  inside the complementary half.  Pinned by the DIM=1 reduction (identical)
  and by the closed-form posterior of a Gaussian target (statistical).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbh_oracle.so")
_lib = None

# model ids of bh_oracle.c
UNIFORM_LOGPDF = 0
GAUSSIAN_MCMC_LOGPDF = 1
GAUSSIAN_LOGPDF = 2
BETA_MCMC_LOGPDF = 3
BETA_LOGPDF = 4
BINOMIAL_LOGLIK = 5
BETA_BINOMIAL_POST = 6
MVN_LOGPDF = 7
GAUSSIAN_LOGLIK = 8
DIAG_GAUSS_LOGPDF = 9
HBB_LOGPDF = 10
LINREG_POST = 11
LOGISTIC_POST = 12

PARTNER_FLAT = 0
PARTNER_WALKER = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_c_u32 = ctypes.c_uint32
_c_f32 = ctypes.c_float


def build() -> str:
    """Compile ``libbh_oracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "bh_oracle.c"))
        ):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.bho_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.bho_u01.argtypes = [_c_u32]
        L.bho_u01.restype = ctypes.c_float
        L.bho_uniform_sample.argtypes = [_c_u32, _c_f32, _c_f32, _c_u32, _f32p]
        L.bho_init_walkers.argtypes = [_c_u32, _f32p, _c_u32, _c_u32, _f32p]
        L.bho_model_logfn.argtypes = [ctypes.c_int, _c_u32, _c_u32, _f32p, _c_u32, _f32p]
        L.bho_model_logfn.restype = ctypes.c_float
        L.bho_stretch_half.argtypes = [
            _c_u32, _c_u32, _c_u32, _c_u32, _c_u32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, _c_f32, _c_f32, ctypes.c_int, ctypes.c_int, _c_u32, _c_u32,
            _f32p, ctypes.c_void_p, ctypes.c_int,
        ]
        L.bho_stretch_half.restype = _c_u32
        L.bho_logfn.argtypes = [ctypes.c_int, _c_u32, _c_u32, _f32p, _c_u32, _c_u32,
                                ctypes.c_void_p, ctypes.c_void_p]
        L.bho_tree_sum_f32.argtypes = [_f32p, _c_u32]
        L.bho_tree_sum_f32.restype = ctypes.c_float
        L.bho_tree_sum_f64.argtypes = [_f64p, _c_u32]
        L.bho_tree_sum_f64.restype = ctypes.c_double
        L.bho_stretch_accu.argtypes = [_c_u32, _c_u32, _c_u32, _c_u32, ctypes.c_void_p,
                                       _u8p, _u32p, _f32p]
        L.bho_reduce_rows_f32.argtypes = [_f32p, _c_u32, _c_u32, _c_u32, _f32p]
        L.bho_min_max.argtypes = [_f32p, _c_u32, _c_u32, _c_u32, _f32p]
        L.bho_hist_bin.argtypes = [_c_f32, _c_f32, _c_f32, _c_u32]
        L.bho_hist_bin.restype = _c_u32
        L.bho_histogram_add.argtypes = [_f32p, _f32p, _c_u32, _c_u32, _c_u32, _c_u32, _u32p]
        L.bho_uint_to_real.argtypes = [_c_f32, _f32p, _u32p, _c_u32, _c_u32, _f32p]
        L.bho_bitonic_ranks.argtypes = [_f32p, _c_u32, _f32p]
        L.bho_mean_var_f64.argtypes = [_f32p, _c_u32, ctypes.c_uint64, _c_u32, _f64p,
                                       ctypes.c_void_p]
        L.bho_acor.argtypes = [_f32p, _c_u32, _c_u32, _c_u32, _f32p, _f32p, _f32p,
                               ctypes.POINTER(_c_u32)]
        L.bho_acor.restype = ctypes.c_int
        L.bho_set_variant.argtypes = [ctypes.c_int]
        L.bho_direct_sample.argtypes = [ctypes.c_int, _c_u32, _c_f32, _c_f32, _c_u32, _f32p]
        L.bho_glm_logpost.argtypes = [_f32p, _f32p, ctypes.c_uint64, _c_u32, ctypes.c_double,
                                      _f32p]
        L.bho_glm_logpost.restype = ctypes.c_double
        L.bho_glm_stretch_half.argtypes = [
            _c_u32, _c_u32, _c_u32, _c_u32, _c_u32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, _c_f32, _c_f32, ctypes.c_int, _f32p, _f32p, ctypes.c_uint64,
            ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        L.bho_glm_stretch_half.restype = _c_u32
        _declare_blocked(L)
        _lib = L
    return _lib


def _declare_blocked(L):
    L.bho_logistic_half_blocked.argtypes = [
        _c_u32, _c_u32, _c_u32, _c_u32, _c_u32, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p, _c_f32, _c_f32, ctypes.c_int, _c_u32, _f32p, ctypes.c_void_p,
        ctypes.c_int]
    L.bho_logistic_half_blocked.restype = _c_u32


_fast = None


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


def fast_lib():
    """The timed CPU-baseline build of the same source (-O3, vector
    expf/logf; AVX-512 where the host has it).  Only bench.py's cpu_baseline
    leg times it; parity checks use lib()."""
    global _fast
    if _fast is None:
        name = "libbh_oracle_fast_v4.so" if _has_avx512() else "libbh_oracle_fast_v3.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path) or (
            os.path.getmtime(path) < os.path.getmtime(os.path.join(_HERE, "bh_oracle.c"))
        ):
            build()
        L = ctypes.CDLL(path)
        _declare_blocked(L)
        _fast = (L, name)
    return _fast[0]


def fast_lib_name() -> str:
    fast_lib()
    return _fast[1]


def set_variant(flags: int) -> None:
    """Planted-mistake switch (tests only): bit 0 un-fuses z, bit 1 un-fuses Y,
    bit 2 drops the logistic posterior's prior, bit 3 flips the sign of its
    linear predictor."""
    lib().bho_set_variant(int(flags))


def _ptr(a: np.ndarray, offset_elems: int = 0) -> int:
    return a.ctypes.data + offset_elems * a.itemsize


# --------------------------------------------------------------------------
# Elementary functions
# --------------------------------------------------------------------------
def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().bho_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def u01(i: int) -> float:
    return lib().bho_u01(int(i))


def uniform_sample(seed: int, lower: float, upper: float, n: int) -> np.ndarray:
    """uniform-sampler.cl:19-34 for n (a multiple of 4) samples."""
    if n % 4:
        raise ValueError("n must be a multiple of 4")
    out = np.zeros(n, np.float32)
    lib().bho_uniform_sample(seed, np.float32(lower), np.float32(upper), n // 4, out)
    return out


DIRECT_UNIFORM, DIRECT_GAUSSIAN, DIRECT_EXPONENTIAL, DIRECT_ERLANG = 0, 1, 2, 3


def direct_sample(kind: int, seed: int, params, n: int) -> np.ndarray:
    """The direct samplers of rng/NAME-sampler.cl for n (a multiple of 4) draws."""
    if n % 4:
        raise ValueError("n must be a multiple of 4")
    p = list(np.asarray(params, np.float32)) + [np.float32(0)]
    out = np.zeros(n, np.float32)
    lib().bho_direct_sample(kind, seed & 0xFFFFFFFF, p[0], p[1], n // 4, out)
    return out


def init_walkers(seed: int, limits: np.ndarray, dim: int, walkers: int) -> np.ndarray:
    """stretch.cl:138-159; returns xs[walker*dim + d] (AoS)."""
    lim = np.ascontiguousarray(limits, np.float32).reshape(-1)
    xs = np.zeros(dim * walkers, np.float32)
    lib().bho_init_walkers(seed, lim, dim, walkers, xs)
    return xs


def model_logfn(model: int, params: np.ndarray, x: np.ndarray, data_len: int = 0,
                params_len: int = 0) -> float:
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    return lib().bho_model_logfn(model, data_len, params_len, p, xx.size, xx)


def logfn_many(model: int, params: np.ndarray, xs: np.ndarray, dim: int, data_len: int = 0,
               params_len: int = 0) -> np.ndarray:
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(xs, np.float32).reshape(-1)
    n = xx.size // dim
    res = np.zeros(n, np.float32)
    lib().bho_logfn(model, data_len, params_len, p, dim, n, _ptr(xx), _ptr(res))
    return res


def tree_sum_f32(v) -> np.float32:
    v = np.ascontiguousarray(v, np.float32)
    return np.float32(lib().bho_tree_sum_f32(v, v.size))


def reduce_rows_f32(a: np.ndarray, local_n: int) -> np.ndarray:
    """clojurecl enq-reduce! over the columns of an m x n column-major block.
    ``a`` has shape (n, m) in C order (= m x n column-major)."""
    a = np.ascontiguousarray(a, np.float32)
    n, m = a.shape
    out = np.zeros(m, np.float32)
    lib().bho_reduce_rows_f32(a.reshape(-1), m, n, local_n, out)
    return out


# --------------------------------------------------------------------------
# Dataset estimates
# --------------------------------------------------------------------------
def min_max(data: np.ndarray) -> np.ndarray:
    """data shape (n, m) C order == m x n column-major; returns limits (m, 2)."""
    a = np.ascontiguousarray(data, np.float32)
    n, m = a.shape
    lim = np.zeros(2 * m, np.float32)
    lib().bho_min_max(a.reshape(-1), m, n, m, lim)
    return lim.reshape(m, 2)


def histogram_counts(limits: np.ndarray, data: np.ndarray, bins: int = 256,
                     counts: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(data, np.float32)
    n, m = a.shape
    if counts is None:
        counts = np.zeros(m * bins, np.uint32)
    lib().bho_histogram_add(np.ascontiguousarray(limits, np.float32).reshape(-1),
                            a.reshape(-1), m, n, m, bins, counts)
    return counts


def uint_to_real(alpha: float, limits: np.ndarray, counts: np.ndarray, bins: int = 256) -> np.ndarray:
    lim = np.ascontiguousarray(limits, np.float32).reshape(-1)
    m = lim.size // 2
    pdf = np.zeros(m * bins, np.float32)
    lib().bho_uint_to_real(np.float32(alpha), lim, np.ascontiguousarray(counts, np.uint32),
                           m, bins, pdf)
    return pdf


def bitonic_ranks(values: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros_like(v)
    lib().bho_bitonic_ranks(v, v.size, out)
    return out


@dataclass
class Histogram:
    """protocols.clj:13-18: limits 2 x DIM, pdf bins x DIM, bin-ranks bins x DIM
    (column-major; here arrays of shape (DIM, 2) and (DIM, bins))."""
    limits: np.ndarray
    pdf: np.ndarray
    bin_ranks: np.ndarray


def histogram(data: np.ndarray, bins: int = 256) -> Histogram:
    """GCNDatasetEngine.histogram, amd_gcn.clj:182-214."""
    a = np.ascontiguousarray(data, np.float32)
    n, m = a.shape
    lim = min_max(a)
    counts = histogram_counts(lim, a, bins)
    alpha = np.float32(bins / n)
    pdf = uint_to_real(alpha, lim, counts, bins).reshape(m, bins)
    ranks = np.stack([bitonic_ranks(pdf[r]) for r in range(m)])
    return Histogram(lim, pdf, ranks)


def mean_var_f64(data: np.ndarray):
    a = np.ascontiguousarray(data, np.float32)
    n, m = a.shape
    mean = np.zeros(m, np.float64)
    var = np.zeros(m, np.float64)
    lib().bho_mean_var_f64(a.reshape(-1), m, n, m, mean, _ptr(var))
    return mean, var


# --------------------------------------------------------------------------
# Autocorrelation
# --------------------------------------------------------------------------
@dataclass
class Autocorrelation:
    """protocols.clj:111-116."""
    tau: np.ndarray
    mean: np.ndarray
    sigma: np.ndarray
    steps: int
    lag: int


def acor(series: np.ndarray, wgs: int = 256) -> Autocorrelation:
    """GCNAcorEngine.acor (amd_gcn.clj:216-267).  ``series`` has shape
    (n, dim) in C order (= dim x n column-major).  Raises ValueError when the
    series is shorter than 5*lag, as the reference does (amd_gcn.clj:264)."""
    s = np.array(series, np.float32, order="C", copy=True)
    n, dim = s.shape
    tau = np.zeros(dim, np.float32)
    mean = np.zeros(dim, np.float32)
    sigma = np.zeros(dim, np.float32)
    lag = _c_u32(0)
    rc = lib().bho_acor(s.reshape(-1), dim, n, wgs, tau, mean, sigma, ctypes.byref(lag))
    if rc != 0:
        raise ValueError(
            "The autocorrelation time is too long relative to the variance. "
            f"Number of steps ({n}) must not be less than {lag.value * 5}.")
    return Autocorrelation(tau, mean, sigma, n, lag.value)


# --------------------------------------------------------------------------
# The stretch sampler, host orchestration of GCNStretch (amd_gcn.clj:271-567)
# --------------------------------------------------------------------------
class Stretch:
    """CPU restatement of ``GCNStretch``.

    State layout follows the reference: ``xs[walker*DIM + d]`` (AoS) with the
    two half-ensembles contiguous; ``logfn[walker]`` (the reference places
    the second half's logfn values at offset DIM*K, amd_gcn.clj:527-532,
    which the ``logfn`` kernel never writes when DIM > 1; here they sit at
    offset K, which is the same place for DIM = 1).
    """

    def __init__(self, model: int, dim: int, walkers: int, params, seed: int,
                 data_len: int = 0, params_len: int = 0, wgs: int = 256,
                 partner: int = PARTNER_FLAT, threads: int = 1):
        walkers = int(walkers)
        if not (2 * wgs <= walkers and walkers % (2 * wgs) == 0):
            # amd_gcn.clj:516, 566-567
            raise ValueError(f"Number of walkers ({walkers}) must be a multiple of {2 * wgs}.")
        self.model, self.dim, self.W, self.K, self.wgs = model, dim, walkers, walkers // 2, wgs
        self.params = np.ascontiguousarray(params, np.float32)
        self.data_len, self.params_len = data_len, params_len
        self.partner, self.threads = partner, threads
        self.xs = np.zeros(dim * walkers, np.float32)
        self.logfn = np.zeros(walkers, np.float32)
        self.iteration_counter = 0
        self.move_seed = 0
        self.move_counter = 0
        self.init(seed)

    # MCMC/init!  amd_gcn.clj:378-387
    def init(self, seed: int):
        self.bare_seed = seed & 0xFFFFFFFF
        self.bare_a = np.float32(2.0)
        self.beta = np.float32(1.0)
        self.move_bare_counter = 0
        self.move_seed = seed & 0xFFFFFFFF
        return self

    # MCMC/init-position!  amd_gcn.clj:388-402
    def init_position(self, seed=None, limits=None, position: "Stretch | None" = None):
        if position is not None:
            self.xs[:] = position.xs
        else:
            self.xs[:] = init_walkers(seed, np.asarray(limits, np.float32), self.dim, self.W)
        self.logfn[:] = logfn_many(self.model, self.params, self.xs, self.dim, self.data_len,
                                   self.params_len)
        self.iteration_counter = 0
        return self

    def _half(self, seed, odd_or_even, step, first: bool, a, beta, accepted=None):
        K, D = self.K, self.dim
        xo = 0 if first else K * D
        so = K * D if first else 0
        lo = 0 if first else K
        return lib().bho_stretch_half(
            seed & 0xFFFFFFFF, odd_or_even, step & 0xFFFFFFFF, K, D, _ptr(self.xs, so),
            _ptr(self.xs, xo), _ptr(self.logfn, lo), np.float32(a), np.float32(beta),
            self.partner, self.model, self.data_len, self.params_len, self.params,
            None if accepted is None else _ptr(accepted), self.threads)

    # MCMCStretch/move-bare!  amd_gcn.clj:347-353 (constants 3333/4444, :547-550)
    def move_bare(self):
        c = self.move_bare_counter
        self._half(self.bare_seed, 3333, c, True, self.bare_a, self.beta)
        self._half(self.bare_seed + 1, 4444, c, False, self.bare_a, self.beta)
        self.move_bare_counter += 1
        return self

    # MCMCStretch/set-temperature!  amd_gcn.clj:354-358
    def set_temperature(self, t: float):
        self.beta = np.float32(1.0 / t)
        return self

    # MCMCStretch/init-move!  amd_gcn.clj:329-339
    def init_move(self, means_acc: np.ndarray, a: float):
        self.move_seed = (self.move_seed + 2) & 0xFFFFFFFF
        self.move_counter = 0
        self.accept = np.zeros(self.K // self.wgs, np.uint32)
        self.means_acc = means_acc
        self.accu_a = np.float32(a)
        return self

    # MCMCStretch/move!  amd_gcn.clj:340-346 (constants 1111/2222, :543-546)
    def move(self):
        K, D = self.K, self.dim
        acc = np.zeros(K, np.uint8)
        c = self.move_counter
        self._half(self.move_seed, 1111, c, True, self.accu_a, 1.0, acc)
        lib().bho_stretch_accu(K, D, self.wgs, c, _ptr(self.xs, 0), acc, self.accept,
                               self.means_acc)
        self._half(self.move_seed + 1, 2222, c, False, self.accu_a, 1.0, acc)
        lib().bho_stretch_accu(K, D, self.wgs, c, _ptr(self.xs, K * D), acc, self.accept,
                               self.means_acc)
        self.move_counter += 1
        return self

    # RandomSampler/sample!  amd_gcn.clj:360-376; returns (n, DIM) C order
    def sample(self, n: int | None = None) -> np.ndarray:
        n = self.W if n is None else int(n)
        out = np.zeros(n * self.dim, np.float32)
        self.set_temperature(1.0)
        avail = self.W * self.dim
        ofst, req = 0, n * self.dim
        while True:
            self.move_bare()
            self.iteration_counter += 1
            if req <= avail:
                out[ofst:ofst + req] = self.xs[:req]
                break
            out[ofst:ofst + avail] = self.xs
            ofst += avail
            req -= avail
        return out.reshape(n, self.dim)

    # MCMC/burn-in!  amd_gcn.clj:403-411
    def burn_in(self, n: int, a: float):
        self.bare_a = np.float32(a)
        self.set_temperature(1.0)
        for _ in range(n):
            self.move_bare()
        self.iteration_counter += n
        return self

    # MCMC/anneal!  amd_gcn.clj:412-420
    def anneal(self, schedule, n: int, a: float):
        self.bare_a = np.float32(a)
        for i in range(n):
            self.set_temperature(schedule(i))
            self.move_bare()
        self.iteration_counter += n
        return self

    def _means_count(self):
        return self.K // self.wgs

    # MCMC/run-sampler!  amd_gcn.clj:421-443
    def run_sampler(self, n: int, a: float):
        D, groups = self.dim, self._means_count()
        means_acc = np.zeros(D * groups * n, np.float32)
        self.init_move(means_acc, a)
        for _ in range(n):
            self.move()
        self.iteration_counter += n
        # sum_means_vertical (+ sum_reduction_horizontal) over the groups of
        # each (step, dim) row, local size min(groups, WGS); then scal!
        rows = means_acc.reshape(n * D, groups)  # row r = step*D + d
        local_m = min(groups, self.wgs)
        sums = reduce_rows_f32(np.ascontiguousarray(rows.T), local_m)
        alpha = np.float32(0.5 / (self.wgs * groups))
        means = (sums * alpha).astype(np.float32).reshape(n, D)
        acc_rate = float(int(self.accept.astype(np.uint64).sum())) / (self.W * n)
        return {"acceptance_rate": acc_rate, "a": float(np.float32(a)),
                "autocorrelation": acor(means, self.wgs), "means": means}

    # MCMC/acc-rate!  amd_gcn.clj:444-453
    def acc_rate(self, a: float) -> float:
        means_acc = np.zeros(self.dim * self._means_count(), np.float32)
        self.init_move(means_acc, a)
        self.move()
        self.iteration_counter += 1
        return float(int(self.accept.astype(np.uint64).sum())) / self.W

    # EstimateEngine/histogram!  amd_gcn.clj:457-487
    def histogram(self, cycles: int = 1, bins: int = 256) -> Histogram:
        D, W = self.dim, self.W
        cur = self.xs.reshape(W, D)
        lim = min_max(cur)
        counts = histogram_counts(lim, cur, bins)
        self.set_temperature(1.0)
        for _ in range(cycles - 1):
            self.move_bare()
            histogram_counts(lim, self.xs.reshape(W, D), bins, counts)
        self.iteration_counter += cycles - 1
        alpha = np.float32(bins / (cycles * W))
        pdf = uint_to_real(alpha, lim, counts, bins).reshape(D, bins)
        ranks = np.stack([bitonic_ranks(pdf[r]) for r in range(D)])
        return Histogram(lim, pdf, ranks)

    # Location/mean, Spread/variance (amd_gcn.clj:489-505), FP64 accumulation
    def mean(self) -> np.ndarray:
        return mean_var_f64(self.xs.reshape(self.W, self.dim))[0]

    def variance(self) -> np.ndarray:
        return mean_var_f64(self.xs.reshape(self.W, self.dim))[1]

    def sd(self) -> np.ndarray:
        return np.sqrt(self.variance())


# --------------------------------------------------------------------------
# GLM (logistic regression) posterior, all in double
# --------------------------------------------------------------------------
def glm_logpost(Xd: np.ndarray, y: np.ndarray, prior_sd: float, beta: np.ndarray) -> np.ndarray:
    """Log posteriors of the rows of beta (n, d); Xd holds the bf16 data as
    float32 (N, d), y the labels."""
    X = np.ascontiguousarray(Xd, np.float32)
    yy = np.ascontiguousarray(y, np.float32)
    B = np.ascontiguousarray(beta, np.float32).reshape(-1, X.shape[1])
    pp = 1.0 / (float(prior_sd) ** 2)
    return np.array([lib().bho_glm_logpost(X.reshape(-1), yy, X.shape[0], X.shape[1], pp, b)
                     for b in B])


class GlmStretch(Stretch):
    """The engine's GLM sampler restated: the moves of ``Stretch`` with the
    logistic log posterior computed in double and kept in double."""

    def __init__(self, Xd, y, prior_sd, walkers, seed, wgs=256, partner=PARTNER_FLAT,
                 threads=1):
        self.Xd = np.ascontiguousarray(Xd, np.float32)
        self.yv = np.ascontiguousarray(y, np.float32)
        self.prior_prec = 1.0 / (float(prior_sd) ** 2)
        super().__init__(-1, self.Xd.shape[1], walkers, np.zeros(1, np.float32), seed, wgs=wgs,
                         partner=partner, threads=threads)
        self.lp = np.zeros(walkers, np.float64)

    def init_position(self, seed=None, limits=None, position=None):
        if position is not None:
            self.xs[:] = position.xs
        else:
            self.xs[:] = init_walkers(seed, np.asarray(limits, np.float32), self.dim, self.W)
        self.lp = glm_logpost(self.Xd, self.yv, 1.0 / np.sqrt(self.prior_prec),
                              self.xs.reshape(self.W, self.dim))
        self.iteration_counter = 0
        return self

    def _half(self, seed, odd_or_even, step, first, a, beta, accepted=None):
        K, D = self.K, self.dim
        xo = 0 if first else K * D
        so = K * D if first else 0
        lo = 0 if first else K
        return lib().bho_glm_stretch_half(
            seed & 0xFFFFFFFF, odd_or_even, step & 0xFFFFFFFF, K, D, _ptr(self.xs, so),
            _ptr(self.xs, xo), _ptr(self.lp, lo), np.float32(a), np.float32(beta), self.partner,
            self.Xd.reshape(-1), self.yv, self.Xd.shape[0], self.prior_prec,
            None if accepted is None else _ptr(accepted), self.threads)
