"""The HIP engine behind the reference's protocols.

One class per engine type of ``src/clojure/uncomplicate/bayadera/internal/
device/amd_gcn.clj``, with the protocol methods of ``internal/protocols.clj``
under their Python names (``burn-in!`` -> ``burn_in`` and so on).  Every
method is a call through the C-ABI of ``libbayadera_hip.so``.

Arrays: a DIM x n column-major matrix of the reference is a numpy array of
shape (n, DIM) in C order (the same memory), or a CUDA torch tensor of that
shape, which the engine reads in place.  Results come back as numpy arrays
(the reference's ``native!``), or are written into a CUDA tensor passed as
``out``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import call
from .models import DistributionModel, GlmModel, LikelihoodModel

WGS_DEFAULT = 256


def _source(model) -> bytes:
    return "\n".join(model.source).encode()


class _Arr:
    """A float32 (or other dtype) array as a pointer the C-ABI can take."""

    def __init__(self, x, dtype=np.float32):
        if hasattr(x, "data_ptr") and hasattr(x, "is_cuda"):
            import torch
            want = {np.float32: torch.float32, np.float64: torch.float64,
                    np.uint32: torch.int32}[dtype]
            if x.dtype != want or not x.is_contiguous():
                raise ValueError(f"device tensors must be contiguous {want}")
            self.keep = x
            self.ptr = x.data_ptr()
            self.size = x.numel()
        else:
            a = np.ascontiguousarray(x, dtype=dtype)
            self.keep = a
            self.ptr = a.ctypes.data
            self.size = a.size


def _out(out, shape, dtype=np.float32):
    """The caller's ``out`` (written in place by the C side) or a new array.
    ``out`` must already have the dtype, be contiguous and hold at least the
    elements the call writes: a conversion would write into a temporary copy
    and a short buffer would be overrun."""
    if out is None:
        a = np.zeros(shape, dtype)
        return a, _Arr(a, dtype)
    need = int(np.prod(shape))
    if hasattr(out, "data_ptr") and hasattr(out, "is_cuda"):
        import torch
        want = {np.float32: torch.float32, np.float64: torch.float64,
                np.uint32: torch.int32}[dtype]
        ok_dtype, contiguous, size = out.dtype == want, out.is_contiguous(), out.numel()
    else:
        if not isinstance(out, np.ndarray):
            raise ValueError("out must be a numpy array or a torch tensor")
        ok_dtype = out.dtype == np.dtype(dtype)
        contiguous, size = out.flags.c_contiguous, out.size
    if not ok_dtype:
        raise ValueError(f"out has dtype {out.dtype}, needs {np.dtype(dtype).name}")
    if not contiguous:
        raise ValueError("out must be contiguous")
    if size < need:
        raise ValueError(f"out holds {size} elements, the call writes {need}")
    return out, _Arr(out, dtype)


@dataclass
class Histogram:
    """protocols.clj:13-18.  limits (DIM, 2); pdf and bin_ranks (DIM, WGS)
    (the reference's 2 x DIM and WGS x DIM column-major matrices)."""
    limits: np.ndarray
    pdf: np.ndarray
    bin_ranks: np.ndarray


@dataclass
class Autocorrelation:
    """protocols.clj:111-116."""
    tau: np.ndarray
    mean: np.ndarray
    sigma: np.ndarray
    steps: int
    lag: int


class _Handle:
    _destroy = ""

    def __init__(self):
        self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def release(self):
        if self._h and self._h.value:
            getattr(_lib.load(), self._destroy)(self._h)
            self._h = ctypes.c_void_p()
        return True

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()


def _build_log():
    return ctypes.create_string_buffer(16384)


# ---------------------------------------------------------------------------
class HipBayaderaFactory(_Handle):
    """GCNBayaderaFactory (amd_gcn.clj:635-678): the EngineFactory."""
    _destroy = "bh_ctx_destroy"

    def __init__(self, device: int = 0, wgs: int = WGS_DEFAULT):
        super().__init__()
        call("bh_ctx_create", int(device), int(wgs), ctypes.byref(self._h))
        cu = ctypes.c_int()
        w = ctypes.c_uint32()
        call("bh_ctx_info", self._h, ctypes.byref(cu), ctypes.byref(w))
        self.device, self.compute_units, self.wgs = device, cu.value, w.value
        self._dataset = HipDatasetEngine(self)
        self._acor = HipAcorEngine(self)

    # EngineFactory
    def distribution_engine(self, model: DistributionModel) -> "HipDistributionEngine":
        return HipDistributionEngine(self, model)

    def likelihood_engine(self, model: LikelihoodModel) -> "HipLikelihoodEngine":
        return HipLikelihoodEngine(self, model)

    def dataset_engine(self) -> "HipDatasetEngine":
        return self._dataset

    def acor_engine(self) -> "HipAcorEngine":
        return self._acor

    def direct_sampler_engine(self, model: DistributionModel) -> "HipDirectSamplerEngine":
        return HipDirectSamplerEngine(self, model)

    def batch(self, model: DistributionModel, params, dims, hyper_len, walkers: int) -> "HipBatch":
        return HipBatch(self, model, params, dims, hyper_len, walkers)

    def glm_batch(self, model, walkers: int) -> "HipGlmBatch":
        return HipGlmBatch(self, model, walkers)

    def fp8_decode(self, codes, f16: bool = False) -> np.ndarray:
        """The device's e4m3 decoding of a byte array (length a multiple of
        4): the fp32 path's, or with f16 the fp16 matrix-core path's."""
        c = np.ascontiguousarray(codes, np.uint8).reshape(-1)
        out = np.zeros(c.size, np.float32)
        call("bh_fp8_decode_f16" if f16 else "bh_fp8_decode", self.handle, c.ctypes.data,
             out.ctypes.data, c.size)
        return out

    def mcmc_factory(self, model) -> "HipStretchFactory":
        if isinstance(model, GlmModel):
            return HipGlmStretchFactory(self, model)
        return HipStretchFactory(self, model)

    def processing_elements(self) -> int:
        return self.compute_units * self.wgs

    def device_rates(self) -> dict:
        """Measured rates: v_exp_f32 and v_fma_f32 lane-ops per second,
        v_mfma_f32_16x16x32_f16 flops per second, random 16-byte gathers
        per second."""
        e, f, m, g = (ctypes.c_double() for _ in range(4))
        call("bh_device_rates", self.handle, ctypes.byref(e), ctypes.byref(f), ctypes.byref(m),
             ctypes.byref(g))
        return {"exp2_per_s": e.value, "fma_per_s": f.value, "mfma_f16_flops_per_s": m.value,
                "gathers_per_s": g.value}

    def device_copy_rate(self, nbytes: int = 1 << 30) -> dict:
        """Read + written GB/s of the engine's own streaming copy kernel of
        nbytes each way, with plain and with non-temporal accesses."""
        g, gn = ctypes.c_double(), ctypes.c_double()
        call("bh_device_copy_rate", self.handle, int(nbytes), ctypes.byref(g), ctypes.byref(gn))
        return {"copy_gbps": g.value, "copy_nt_gbps": gn.value}

    def device_gather_rate(self, nbytes: int = 1 << 30, nontemporal: bool = False) -> float:
        """Random 16-byte gathers per second from a buffer of nbytes."""
        r = ctypes.c_double()
        call("bh_device_gather_rate", self.handle, int(nbytes), int(bool(nontemporal)),
             ctypes.byref(r))
        return r.value


class HipDirectSamplerEngine:
    """GCNDirectSamplerEngine (amd_gcn.clj:49-62) for the library models with
    direct samplers (rng/NAME-sampler.cl): uniform, gaussian, exponential, erlang."""
    KINDS = {"uniform": 0, "gaussian": 1, "exponential": 2, "erlang": 3}

    def __init__(self, factory: HipBayaderaFactory, model: DistributionModel):
        if model.name not in self.KINDS:
            raise NotImplementedError(f"no direct sampler for {model.name}")
        self.factory, self.kind = factory, self.KINDS[model.name]

    def sample(self, seed: int, params, n: int, out=None):
        if n % 4:
            raise ValueError("GCN direct sampler supports only matrices with ncols and offset "
                             "that are multiple of 4.")
        p = _Arr(np.asarray(params, np.float32).reshape(-1))
        res, o = _out(out, (n,))
        call("bh_direct_sample", self.factory.handle, self.kind, int(seed) & 0xFFFFFFFF, p.ptr,
             p.size, o.ptr, n)
        return res


class HipDistributionEngine(_Handle):
    """GCNDistributionEngine (amd_gcn.clj:66-93): DensityEngine."""
    _destroy = "bh_dist_engine_destroy"

    def __init__(self, factory: HipBayaderaFactory, model: DistributionModel):
        super().__init__()
        self.factory, self.dist_model = factory, model
        log = _build_log()
        call("bh_dist_engine_create", factory.handle, _source(model), model.logpdf.encode(),
             int(model.params_size), ctypes.byref(self._h), log, len(log))

    def model(self):
        return self.dist_model

    def _density(self, fn, params, x, out):
        p = _Arr(params)
        xa = _Arr(x)
        n, dim = _shape(x)
        res, o = _out(out, (n,))
        call(fn, self._h, p.ptr, p.size, xa.ptr, dim, n, o.ptr)
        return res

    def log_density(self, params, x, out=None):
        return self._density("bh_log_density", params, x, out)

    def density(self, params, x, out=None):
        return self._density("bh_density", params, x, out)


class HipLikelihoodEngine(_Handle):
    """GCNLikelihoodEngine (amd_gcn.clj:97-133): DensityEngine + LikelihoodEngine."""
    _destroy = "bh_lik_engine_destroy"

    def __init__(self, factory: HipBayaderaFactory, model: LikelihoodModel):
        super().__init__()
        self.factory, self.lik_model = factory, model
        log = _build_log()
        call("bh_lik_engine_create", factory.handle, _source(model), model.loglik.encode(),
             ctypes.byref(self._h), log, len(log))

    def model(self):
        return self.lik_model

    def _run(self, fn, data, x, out):
        d = _Arr(data)
        xa = _Arr(x)
        n, dim = _shape(x)
        res, o = _out(out, (n,))
        call(fn, self._h, d.ptr, d.size, xa.ptr, dim, n, o.ptr)
        return res

    def log_density(self, data, x, out=None):
        return self._run("bh_loglik", data, x, out)

    def density(self, data, x, out=None):
        return self._run("bh_lik", data, x, out)

    def evidence(self, data, x) -> float:
        d = _Arr(data)
        xa = _Arr(x)
        n, dim = _shape(x)
        r = ctypes.c_double()
        call("bh_evidence", self._h, d.ptr, d.size, xa.ptr, dim, n, ctypes.byref(r))
        return r.value


def _shape(x):
    s = tuple(x.shape)
    if len(s) == 1:
        return s[0], 1
    if len(s) != 2:
        raise ValueError("expected a (n, DIM) array")
    return s[0], s[1]


class HipDatasetEngine:
    """GCNDatasetEngine (amd_gcn.clj:137-214): DatasetEngine + EstimateEngine."""

    def __init__(self, factory: HipBayaderaFactory):
        self.factory = factory

    def data_mean(self, data, out=None):
        return self._moment("bh_dataset_mean", data, out)

    def data_variance(self, data, out=None):
        return self._moment("bh_dataset_variance", data, out)

    def _moment(self, fn, data, out):
        a = _Arr(data)
        n, m = _shape(data)
        res, o = _out(out, (m,))
        call(fn, self.factory.handle, a.ptr, m, n, m, o.ptr)
        return res

    def histogram(self, data) -> Histogram:
        a = _Arr(data)
        n, m = _shape(data)
        bins = self.factory.wgs
        lim = np.zeros((m, 2), np.float32)
        pdf = np.zeros((m, bins), np.float32)
        ranks = np.zeros((m, bins), np.float32)
        call("bh_dataset_histogram", self.factory.handle, a.ptr, m, n, m, lim.ctypes.data,
             pdf.ctypes.data, ranks.ctypes.data)
        return Histogram(lim, pdf, ranks)


class HipAcorEngine:
    """GCNAcorEngine (amd_gcn.clj:216-267)."""

    def __init__(self, factory: HipBayaderaFactory):
        self.factory = factory

    def acor(self, series) -> Autocorrelation:
        a = _Arr(series)
        n, dim = _shape(series)
        tau = np.zeros(dim, np.float32)
        mean = np.zeros(dim, np.float32)
        sigma = np.zeros(dim, np.float32)
        lag = ctypes.c_uint32()
        call("bh_acor", self.factory.handle, a.ptr, dim, n, tau.ctypes.data, mean.ctypes.data,
             sigma.ctypes.data, ctypes.byref(lag))
        return Autocorrelation(tau, mean, sigma, n, lag.value)

    def acor_batch(self, series, dims, n: int) -> list:
        """acor of many series in one launch: series holds series p as an
        (n, dims[p]) block after those of series 0 .. p-1 (a batch's
        recorded run); one Autocorrelation per series."""
        a = _Arr(series)
        dv = np.ascontiguousarray(dims, np.uint32)
        total = int(dv.sum())
        tau = np.zeros(total, np.float32)
        mean = np.zeros(total, np.float32)
        sigma = np.zeros(total, np.float32)
        lag = ctypes.c_uint32()
        call("bh_acor_batch", self.factory.handle, a.ptr, len(dv), dv.ctypes.data, int(n),
             tau.ctypes.data, mean.ctypes.data, sigma.ctypes.data, ctypes.byref(lag))
        offs = np.concatenate([[0], np.cumsum(dv.astype(np.int64))])
        return [Autocorrelation(tau[offs[p]:offs[p + 1]], mean[offs[p]:offs[p + 1]],
                                sigma[offs[p]:offs[p + 1]], int(n), lag.value)
                for p in range(len(dv))]


# ---------------------------------------------------------------------------
class HipStretchFactory(_Handle):
    """GCNStretchFactory / gcn-stretch-factory (amd_gcn.clj:509-567, 623-631):
    the model's mcmc-logpdf compiled with DIM and WGS into the engine kernels."""
    _destroy = "bh_stretch_factory_destroy"

    def __init__(self, factory: HipBayaderaFactory, model: DistributionModel):
        super().__init__()
        self.factory, self.dist_model = factory, model
        log = _build_log()
        call("bh_stretch_factory_create", factory.handle, _source(model),
             model.mcmc_logpdf.encode(), int(model.dimension), int(model.params_size),
             ctypes.byref(self._h), log, len(log))

    def create_sampler(self, seed: int, walkers: int, params) -> "HipStretch":
        return HipStretch(self, seed, walkers, params)


class HipGlmStretchFactory(HipStretchFactory):
    """A stretch factory for a GLM posterior declared by form (GlmModel): the
    likelihood of a whole half-ensemble runs as one X.B contraction."""

    def __init__(self, factory: HipBayaderaFactory, model: GlmModel):
        _Handle.__init__(self)
        self.factory, self.dist_model = factory, model
        n, d = model.X.shape
        log = _build_log()
        y = _Arr(model.y)
        if model.kind == "gaussian":
            X = _Arr(model.X)
            call("bh_glm_gaussian_factory_create", factory.handle, X.ptr, y.ptr, int(n), int(d),
                 float(model.prior_sd), float(model.log_sigma_sd), ctypes.byref(self._h), log,
                 len(log))
            return
        kind = {"logistic": _lib.BH_GLM_LOGISTIC}[model.kind]
        X = _Arr(model.X, np.uint16) if not hasattr(model.X, "data_ptr") else _Arr(model.X)
        call("bh_glm_stretch_factory_create", factory.handle, kind, X.ptr, y.ptr, int(n), int(d),
             float(model.prior_sd), ctypes.byref(self._h), log, len(log))

    def log_density(self, beta, precision: int = _lib.BH_GLM_FP64) -> np.ndarray:
        """Log posterior (likelihood + prior) of the rows of beta (n, d)."""
        b = _Arr(beta)
        n = b.size // self.dist_model.dimension
        out = np.zeros(n, np.float64)
        call("bh_glm_log_density", self._h, b.ptr, n, int(precision), out.ctypes.data)
        return out


def _shape_size(x) -> int:
    return int(x.numel()) if hasattr(x, "numel") else int(np.asarray(x).size)


class HipBatch(_Handle):
    """Many independent posteriors of one model in one launch (no reference
    counterpart).  ``params``: one array per problem ([data, hyperparameters]
    as create-sampler takes them), ``dims`` their dimensions (<= the model's
    dimension), ``hyper_len`` each problem's hyperparameter count."""
    _destroy = "bh_batch_destroy"

    def __init__(self, factory: HipBayaderaFactory, model: DistributionModel, params, dims,
                 hyper_len, walkers: int):
        super().__init__()
        self.factory, self.dist_model = factory, model
        self.problems = len(params)
        self.dims = np.ascontiguousarray(dims, np.uint32)
        self.walkers = int(walkers)
        self.max_dim = int(model.dimension)
        offs = np.zeros(self.problems + 1, np.uint64)
        offs[1:] = np.cumsum([len(p) for p in params])
        flat = np.concatenate([np.asarray(p, np.float32) for p in params]).astype(np.float32)
        hl = np.broadcast_to(np.asarray(hyper_len, np.uint32), (self.problems,)).copy()
        log = _build_log()
        call("bh_batch_create", factory.handle, _source(model), model.mcmc_logpdf.encode(),
             self.max_dim, self.problems, self.dims.ctypes.data, offs.ctypes.data,
             hl.ctypes.data, flat.ctypes.data, self.walkers, ctypes.byref(self._h), log, len(log))

    def set_partner(self, mode: int):
        call("bh_batch_set_partner", self._h, int(mode))
        return self

    def init_position(self, seed: int, limits):
        lim = _Arr(np.asarray(limits, np.float32).reshape(self.problems, self.max_dim, 2))
        call("bh_batch_init_position", self._h, int(seed) & 0xFFFFFFFF, lim.ptr)
        return self

    def burn_in(self, seed0: int, n: int, a: float):
        call("bh_batch_burn_in", self._h, int(seed0) & 0xFFFFFFFF, int(n), float(a))
        return self

    def set_state(self, where: int):
        """Walkers of every launch class in LDS (_lib.BH_BATCH_STATE_LDS) or
        in place in global memory (_lib.BH_BATCH_STATE_GLOBAL) during a call."""
        call("bh_batch_set_state", self._h, int(where))
        return self

    def class_states(self) -> list:
        n = ctypes.c_uint32(0)
        call("bh_batch_class_states", self._h, ctypes.byref(n), None)
        st = (ctypes.c_int * max(1, n.value))()
        call("bh_batch_class_states", self._h, ctypes.byref(n), st)
        return [int(st[i]) for i in range(n.value)]

    def autotune(self, steps: int = 20) -> dict:
        """Times the candidate launch shapes from the current state and keeps
        the fastest (the state is restored)."""
        nb, nc, ms = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        call("bh_batch_autotune", self._h, int(steps), ctypes.byref(nb), ctypes.byref(nc),
             ctypes.byref(ms))
        return {"threads_per_problem": nb.value, "launch_classes": nc.value, "ms": ms.value,
                "class_states": ["global" if s == _lib.BH_BATCH_STATE_GLOBAL else "lds"
                                 for s in self.class_states()]}

    def run(self, seed0: int, n: int, a: float, series=None) -> dict:
        """n accounted steps.  ``series``: optional (device or host) float
        array of n * sum(dims) values that receives each problem's
        per-step ensemble means, problem b's (n, dims[b]) block after those
        of problems 0 .. b-1 (see ``series_blocks``)."""
        acc = np.zeros(self.problems, np.float64)
        mean = np.zeros((self.problems, self.max_dim), np.float64)
        sd = np.zeros((self.problems, self.max_dim), np.float64)
        if series is None:
            call("bh_batch_run", self._h, int(seed0) & 0xFFFFFFFF, int(n), float(a),
                 acc.ctypes.data, mean.ctypes.data, sd.ctypes.data)
        else:
            total = int(n) * int(self.dims.astype(np.int64).sum())
            if _shape_size(series) < total:
                raise ValueError(f"series holds fewer than {total} values")
            ser = _Arr(series)
            call("bh_batch_run_series", self._h, int(seed0) & 0xFFFFFFFF, int(n), float(a),
                 acc.ctypes.data, mean.ctypes.data, sd.ctypes.data, ser.ptr)
        return {"acceptance-rate": acc, "mean": mean, "sd": sd}

    def series_blocks(self, series, n: int):
        """Problem b's (n, dims[b]) view of a recorded series."""
        offs = np.concatenate([[0], np.cumsum(self.dims.astype(np.int64) * int(n))])
        return [series[offs[b]:offs[b + 1]].reshape(int(n), int(d))
                for b, d in enumerate(self.dims)]

    def state(self):
        xs = np.zeros((self.problems, self.walkers * self.max_dim), np.float32)
        lf = np.zeros((self.problems, self.walkers), np.float32)
        call("bh_batch_state_get", self._h, xs.ctypes.data, lf.ctypes.data)
        # problem b's positions are packed with its own dimension
        return [xs[b, :self.walkers * d].reshape(self.walkers, d)
                for b, d in enumerate(self.dims)], lf

    def last_gpu_ms(self) -> float:
        ms = ctypes.c_double()
        call("bh_batch_last_gpu_ms", self._h, ctypes.byref(ms))
        return ms.value


class HipGlmBatch(_Handle):
    """Many independent logistic-regression posteriors (GlmBatchModel) in
    lockstep, their likelihoods as one grouped contraction over fp8 data (no
    reference counterpart)."""
    _destroy = "bh_glm_batch_destroy"

    def __init__(self, factory: HipBayaderaFactory, model, walkers: int):
        super().__init__()
        self.factory, self.dist_model = factory, model
        self.problems = model.problems
        self.walkers = int(walkers)
        self.dims = model.dims
        rows = model.rows
        X = np.concatenate([np.ascontiguousarray(x, np.uint8).reshape(-1) for x in model.X])
        sc = np.concatenate([np.asarray(s, np.float32) for s in model.scales])
        y = np.concatenate(model.y).astype(np.float32)
        kind = {"logistic": _lib.BH_GLM_LOGISTIC}[model.kind]
        call("bh_glm_batch_create", factory.handle, kind, self.problems, rows.ctypes.data,
             self.dims.ctypes.data, X.ctypes.data, sc.ctypes.data, y.ctypes.data,
             float(model.prior_sd), self.walkers, ctypes.byref(self._h))

    def set_precision(self, precision: int):
        call("bh_glm_batch_set_precision", self._h, int(precision))
        return self

    def set_tiles(self, row_stride: int, unit_rows: int, col_blocks: int):
        """Tile shape of the contraction for the problems of one row stride
        (32 or 64): rows per unit and 16-point column blocks per wave."""
        call("bh_glm_batch_set_tiles", self._h, int(row_stride), int(unit_rows),
             int(col_blocks))
        return self

    def autotune(self, reps: int = 3) -> dict:
        """Times every candidate tile shape per row stride on the current
        positions and keeps the fastest; returns the choices and their
        milliseconds per contraction launch."""
        ur = np.zeros(2, np.uint32)
        cb = np.zeros(2, np.uint32)
        ms = np.zeros(2, np.float32)
        call("bh_glm_batch_autotune", self._h, int(reps), ur.ctypes.data, cb.ctypes.data,
             ms.ctypes.data)
        return {str(s): {"unit_rows": int(ur[i]), "col_blocks": int(cb[i]),
                         "ms": float(ms[i])} for i, s in enumerate((32, 64))}

    def init_position(self, seed: int, limits):
        """limits: (problems, 64, 2) or (problems, d, 2) with d <= 64."""
        lim = np.zeros((self.problems, 64, 2), np.float32)
        la = np.asarray(limits, np.float32)
        lim[:, :la.shape[1]] = la
        call("bh_glm_batch_init_position", self._h, int(seed) & 0xFFFFFFFF, lim.ctypes.data)
        return self

    def burn_in(self, seed0: int, n: int, a: float):
        call("bh_glm_batch_burn_in", self._h, int(seed0) & 0xFFFFFFFF, int(n), float(a))
        return self

    def run(self, seed0: int, n: int, a: float, series=None) -> dict:
        acc = np.zeros(self.problems, np.float64)
        mean = np.zeros((self.problems, 64), np.float64)
        sd = np.zeros((self.problems, 64), np.float64)
        ptr = None
        if series is not None:
            total = int(n) * int(self.dims.astype(np.int64).sum())
            if _shape_size(series) < total:
                raise ValueError(f"series holds fewer than {total} values")
            ser = _Arr(series)
            ptr = ser.ptr
        call("bh_glm_batch_run", self._h, int(seed0) & 0xFFFFFFFF, int(n), float(a),
             acc.ctypes.data, mean.ctypes.data, sd.ctypes.data, ptr)
        return {"acceptance-rate": acc, "mean": mean, "sd": sd}

    def series_blocks(self, series, n: int):
        offs = np.concatenate([[0], np.cumsum(self.dims.astype(np.int64) * int(n))])
        return [series[offs[b]:offs[b + 1]].reshape(int(n), int(d))
                for b, d in enumerate(self.dims)]

    def state(self):
        """(positions per problem (walkers, d_p), log posteriors (problems,
        walkers) in double)."""
        xs = np.zeros((self.problems, self.walkers, 64), np.float32)
        lp = np.zeros((self.problems, self.walkers), np.float64)
        call("bh_glm_batch_state_get", self._h, xs.ctypes.data, lp.ctypes.data)
        return [xs[b, :, :d] for b, d in enumerate(self.dims)], lp

    def log_density(self, precision: int) -> np.ndarray:
        """Log posteriors of the current positions in `precision`
        (problems, walkers)."""
        lp = np.zeros((self.problems, self.walkers), np.float64)
        call("bh_glm_batch_log_density", self._h, int(precision), lp.ctypes.data)
        return lp

    def last_gpu_ms(self) -> float:
        ms = ctypes.c_double()
        call("bh_glm_batch_last_gpu_ms", self._h, ctypes.byref(ms))
        return ms.value


class HipStretch(_Handle):
    """GCNStretch (amd_gcn.clj:271-507): MCMC, MCMCStretch, RandomSampler,
    EstimateEngine, Location, Spread."""
    _destroy = "bh_stretch_destroy"

    def __init__(self, sf: HipStretchFactory, seed: int, walkers: int, params):
        super().__init__()
        self.sf = sf
        self.dim = sf.dist_model.dimension
        self.walkers = int(walkers)
        self.wgs = sf.factory.wgs
        p = _Arr(params)
        self._params = p
        call("bh_stretch_create", sf.handle, int(seed) & 0xFFFFFFFF, self.walkers, p.ptr, p.size,
             ctypes.byref(self._h))

    def model(self):
        return self.sf.dist_model

    # engine options (no reference counterpart: the reference has one mode)
    def set_partner(self, mode: int):
        call("bh_stretch_set_partner", self._h, int(mode))
        return self

    def set_exec(self, mode: int):
        call("bh_stretch_set_exec", self._h, int(mode))
        return self

    def set_glm_precision(self, precision: int):
        call("bh_stretch_set_glm_precision", self._h, int(precision))
        return self

    def set_glm_guard(self, guard: float):
        """Guarded accept test of a reduced-precision GLM sampler: walkers
        within ``guard`` (log units) of their accept threshold are decided
        on fp64 log posteriors (0 turns it off)."""
        call("bh_stretch_set_glm_guard", self._h, float(guard))
        return self

    def set_kernel_timing(self, on: bool = True, every: int = 1):
        """HIP events around every ``every``-th launch of the dominant kernel
        (the first 1024 such launches after this call)."""
        call("bh_stretch_set_kernel_timing", self._h, int(every) if on else 0)
        return self

    def kernel_timing(self) -> dict:
        avg, n = ctypes.c_double(), ctypes.c_uint32()
        call("bh_stretch_kernel_timing", self._h, ctypes.byref(avg), ctypes.byref(n))
        return {"avg_ms": avg.value, "launches": n.value}

    def glm_guard_stats(self) -> dict:
        listed, overflow = ctypes.c_uint64(), ctypes.c_uint64()
        call("bh_stretch_glm_guard_stats", self._h, ctypes.byref(listed), ctypes.byref(overflow))
        return {"listed": listed.value, "overflow": overflow.value}

    def set_timing(self, on: bool = True):
        call("bh_stretch_set_timing", self._h, int(bool(on)))
        return self

    def last_gpu_ms(self) -> float:
        ms = ctypes.c_double()
        call("bh_stretch_last_gpu_ms", self._h, ctypes.byref(ms))
        return ms.value

    # Info
    def info(self) -> dict:
        w = ctypes.c_uint32()
        it = ctypes.c_uint64()
        call("bh_stretch_info", self._h, ctypes.byref(w), ctypes.byref(it))
        return {"walker-count": w.value, "iteration-counter": it.value}

    # MCMC
    def init(self, seed: int):
        call("bh_stretch_init", self._h, int(seed) & 0xFFFFFFFF)
        return self

    def init_position(self, seed_or_position, limits=None):
        if isinstance(seed_or_position, HipStretch):
            call("bh_stretch_init_position_from", self._h, seed_or_position.handle)
        else:
            lim = _Arr(np.asarray(limits, np.float32).reshape(-1))
            if lim.size != 2 * self.dim:
                raise ValueError(f"limits must be 2 x {self.dim}")
            call("bh_stretch_init_position", self._h, int(seed_or_position) & 0xFFFFFFFF, lim.ptr)
        return self

    def burn_in(self, n: int, a: float):
        call("bh_stretch_burn_in", self._h, int(n), float(a))
        return self

    def anneal(self, schedule, n: int, a: float):
        temps = np.array([schedule(i) for i in range(int(n))], np.float64)
        call("bh_stretch_anneal", self._h, temps.ctypes.data, int(n), float(a))
        return self

    def acc_rate(self, a: float) -> float:
        r = ctypes.c_double()
        call("bh_stretch_acc_rate", self._h, float(a), ctypes.byref(r))
        return r.value

    def run_sampler(self, n: int, a: float, means_out=None) -> dict:
        D = self.dim
        tau = np.zeros(D, np.float32)
        mean = np.zeros(D, np.float32)
        sigma = np.zeros(D, np.float32)
        lag = ctypes.c_uint32()
        acc = ctypes.c_double()
        means, m = _out(means_out, (int(n), D))
        rc = _lib.load().bh_stretch_run(self._h, int(n), float(a), ctypes.byref(acc), m.ptr,
                                        tau.ctypes.data, mean.ctypes.data, sigma.ctypes.data,
                                        ctypes.byref(lag))
        _lib.check(rc, "run-sampler!")
        return {"acceptance-rate": acc.value, "a": float(np.float32(a)),
                "autocorrelation": Autocorrelation(tau, mean, sigma, int(n), lag.value),
                "means": means}

    # MCMCStretch
    def init_move(self, a: float, steps: int = 1):
        call("bh_stretch_init_move", self._h, float(a), int(steps))
        return self

    def move(self):
        call("bh_stretch_move", self._h)
        return self

    def move_bare(self):
        call("bh_stretch_move_bare", self._h)
        return self

    def set_temperature(self, t: float):
        call("bh_stretch_set_temperature", self._h, float(t))
        return self

    def read_accounting(self, steps: int):
        groups = self.walkers // 2 // self.wgs
        acc = np.zeros(groups, np.uint32)
        means = np.zeros(self.dim * groups * steps, np.float32)
        call("bh_stretch_read_accounting", self._h, acc.ctypes.data, means.ctypes.data, int(steps))
        return acc, means

    # RandomSampler
    def sample(self, n: int | None = None, out=None):
        n = self.walkers if n is None else int(n)
        res, o = _out(out, (n, self.dim))
        call("bh_stretch_sample", self._h, o.ptr, n)
        return res

    # EstimateEngine
    def histogram(self, cycles: int = 1) -> Histogram:
        bins = self.wgs
        lim = np.zeros((self.dim, 2), np.float32)
        pdf = np.zeros((self.dim, bins), np.float32)
        ranks = np.zeros((self.dim, bins), np.float32)
        call("bh_stretch_histogram", self._h, int(cycles), lim.ctypes.data, pdf.ctypes.data,
             ranks.ctypes.data)
        return Histogram(lim, pdf, ranks)

    # Location / Spread
    def mean(self) -> np.ndarray:
        out = np.zeros(self.dim, np.float32)
        call("bh_stretch_mean", self._h, out.ctypes.data)
        return out

    def variance(self) -> np.ndarray:
        out = np.zeros(self.dim, np.float32)
        call("bh_stretch_variance", self._h, out.ctypes.data)
        return out

    def sd(self) -> np.ndarray:
        return np.sqrt(self.variance())

    # state for resume
    def state(self):
        xs = np.zeros((self.walkers, self.dim), np.float32)
        lf = np.zeros(self.walkers, np.float32)
        c = np.zeros(6, np.uint32)
        call("bh_stretch_state_get", self._h, xs.ctypes.data, lf.ctypes.data, c.ctypes.data)
        return xs, lf, c

    def set_state(self, xs=None, logfn=None, counters=None):
        x = None if xs is None else _Arr(xs)
        lf = None if logfn is None else _Arr(logfn)
        c = None if counters is None else _Arr(counters, np.uint32)
        call("bh_stretch_state_set", self._h, None if x is None else x.ptr,
             None if lf is None else lf.ptr, None if c is None else c.ptr)
        return self
