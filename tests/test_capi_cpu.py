"""The C-ABI library on the CPU: it loads, exports what the header declares,
and compiles every model source with hipRTC for gfx950 (no device needed)."""
import ctypes

import pytest

from bayadera_amd import _lib, models


def test_library_exports_every_declared_function():
    L = _lib.load()
    declared = _lib.header_functions()
    assert len(declared) >= 40
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.SIGNATURES)
    assert L.bh_version() == 100


def _compile(source, logfn=None, logpdf=None, loglik=None, dim=1):
    log = ctypes.create_string_buffer(16384)
    rc = _lib.load().bh_rtc_compile_check(
        source.encode(), logfn and logfn.encode(), logpdf and logpdf.encode(),
        loglik and loglik.encode(), dim, 256, b"gfx950", log, len(log))
    return rc, log.value.decode(errors="replace")


@pytest.mark.parametrize("name", sorted(models.distribution_models()))
def test_library_distribution_sources_compile(name):
    m = models.distribution_models()[name]
    rc, log = _compile("\n".join(m.source), m.mcmc_logpdf, m.logpdf, None, m.dimension)
    assert rc == 0, log + _lib.last_error()


@pytest.mark.parametrize("name", sorted(models.likelihood_models()))
def test_library_likelihood_sources_compile(name):
    m = models.likelihood_models()[name]
    rc, log = _compile("\n".join(m.source), None, None, m.loglik)
    assert rc == 0, log + _lib.last_error()


def test_posterior_model_compiles():
    post = models.beta_model().posterior_model("beta_binomial",
                                               models.likelihood_models()["binomial"])
    assert post.params_size == 3 and post.dimension == 1
    rc, log = _compile("\n".join(post.source), post.mcmc_logpdf, post.logpdf)
    assert rc == 0, log + _lib.last_error()


def test_benchmark_model_sources_compile():
    rc, log = _compile(models.MVN_SOURCE, "mvn_logpdf", "mvn_logpdf", dim=4)
    assert rc == 0, log
    lin = models.likelihood_model(models.LINREG_SOURCE, name="linreg")
    prior = models.distribution_model(models.LINREG_SOURCE, name="linreg_prior", dimension=17,
                                      params_size=2)
    post = prior.posterior_model("linreg_post", lin)
    rc, log = _compile("\n".join(post.source), post.mcmc_logpdf, post.logpdf, dim=17)
    assert rc == 0, log + _lib.last_error()
    rc, log = _compile(models.LOGISTIC_SOURCE, None, "gaussian_iid_logpdf", "logistic_loglik")
    assert rc == 0, log


def test_reference_dialect_user_source_compiles():
    # a user model in the reference's OpenCL C99 dialect (distribution.cl:1-4)
    src = """
    inline REAL my_logpdf(const uint data_len, const uint params_len, const REAL* params,
                          const uint dim, const REAL* x) {
        return pown(x[0] - params[0], 2) * -0.5f + native_log(params[1]) - lgamma(2.5f)
            + (isnan(x[0]) ? NAN : 0.0f);
    }
    """
    rc, log = _compile(src, "my_logpdf", "my_logpdf")
    assert rc == 0, log + _lib.last_error()


def test_broken_source_reports_compiler_log():
    rc, log = _compile("inline REAL broken(const uint a) { return undefined_symbol; }",
                       "broken", None)
    assert rc == _lib.BH_ERR_RTC
    assert "undefined_symbol" in log
    assert "undefined_symbol" in _lib.last_error()


def test_invalid_names_rejected():
    rc, _ = _compile(models.GAUSSIAN_SOURCE, "gaussian_logpdf; evil()", None)
    assert rc == _lib.BH_ERR_ARG


def test_context_creation_fails_loudly_or_succeeds():
    from bayadera_amd import engine
    try:
        f = engine.HipBayaderaFactory(0, 256)
    except _lib.BayaderaError as e:
        assert e.code == _lib.BH_ERR_HIP and "hip" in str(e).lower()
    else:
        assert f.processing_elements() == f.compute_units * 256
        f.release()


def test_bad_wgs_rejected():
    h = ctypes.c_void_p()
    assert _lib.load().bh_ctx_create(0, 100, ctypes.byref(h)) == _lib.BH_ERR_ARG
    assert "power of two" in _lib.last_error()


@pytest.mark.parametrize("defines,source", [
    (["-DWGS=256", "-DDIM=64", "-DBH_GLM=1"], ""),
    (["-DWGS=256", "-DDIM=32", "-DBH_GLM=1"], ""),
    (["-DWGS=256", "-DDIM=17", "-DBH_GLM=1", "-DBH_GLM_KIND=2"], ""),
    (["-DWGS=256", "-DDIM=52", "-DBH_BATCH=1", "-DLOGFN=hbb_logpdf"], models.HBB_SOURCE),
    (["-DWGS=256", "-DDIM=6", "-DBH_BATCH=1", "-DLOGFN=mvn_logpdf"], models.MVN_SOURCE),
    # the move kernels' LDS-staged records: 16-byte pieces, float pieces
    (["-DWGS=256", "-DDIM=16", "-DLOGFN=mvn_logpdf"], models.MVN_SOURCE),
    (["-DWGS=256", "-DDIM=33", "-DLOGFN=diag_gaussian_logpdf"], models.DIAG_GAUSS_SOURCE),
])
def test_engine_variants_compile(tmp_path, defines, source):
    """The GLM and batched variants of the run-time-compiled engine source
    build for gfx950 (hipcc on the same text hipRTC receives)."""
    import os
    import shutil
    import subprocess
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc not available")
    csrc = os.path.join(os.path.dirname(_lib.HERE), "bayadera_amd", "csrc")
    text = ('#include "bh_rtc_prelude.h"\n#pragma clang force_cuda_host_device begin\n' + source
            + "\n#pragma clang force_cuda_host_device end\n"
            + open(os.path.join(csrc, "bh_rtc_engine.hip")).read())
    src = tmp_path / "engine.hip"
    src.write_text(text)
    res = subprocess.run([hipcc, "--offload-arch=gfx950", "--cuda-device-only", "-O1",
                          "-std=c++17", "-ffp-contract=on", "-I", csrc, *defines, "-c", str(src),
                          "-o", str(tmp_path / "engine.o")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]


def test_batch_state_calls_reject_null():
    lib = _lib.load()
    assert lib.bh_batch_set_state(None, _lib.BH_BATCH_STATE_GLOBAL) == _lib.BH_ERR_ARG
    n = ctypes.c_uint32(0)
    assert lib.bh_batch_class_states(None, ctypes.byref(n), None) == _lib.BH_ERR_ARG


def test_acor_batch_rejects_null():
    lib = _lib.load()
    dims = (ctypes.c_uint32 * 1)(4)
    assert lib.bh_acor_batch(None, None, 1, dims, 100, None, None, None, None) == _lib.BH_ERR_ARG


def test_glm_batch_tile_calls_reject_null():
    lib = _lib.load()
    assert lib.bh_glm_batch_set_tiles(None, 32, 2048, 4) == _lib.BH_ERR_ARG
    assert lib.bh_glm_batch_autotune(None, 3, None, None, None) == _lib.BH_ERR_ARG
