"""The C++ scheduler and request server (saige-gpu_amd/csrc/fit_engine.cpp)
and the C++ drivers (fit_core.cpp) on the CPU, over tests/native/
engine_mock.cpp: a stand-in for the block-PCG slot engine's C-ABI calls with
a dense FP64 K (one Jacobi-preconditioned CG iteration per slot and step,
the reference's stopping rule, FG.cpp:2696-2792).  Both are compiled here
with g++ into one shared library.

Checks: a fit's result is bitwise the same alone, in a batch and through the
server with staggered arrivals; tau agrees with the FP64 oracle driver on
the same K and probes; an injected engine failure ends every fit with an
error and no hang; admission waits while the reported free memory is below
the reserve; every trace buffer goes back to the context when a batch ends
and the next batch takes it again.  tests/native/engine_mock_main.cpp runs
the same scenarios from several threads under the host sanitizers."""
import ctypes
import os
import subprocess
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from saige_gpu import _lib, api, plink, synth
from saige_gpu import driver as D

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "saige-gpu_amd", "csrc")
INC = os.path.join(HERE, "..", "include")
N, M = 240, 600


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("enginemock")
    vm, so = str(tmp / "vmath.o"), str(tmp / "libenginemock.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffast-math", "-fopenmp-simd", "-c", "-o", vm,
                    os.path.join(CSRC, "vmath.cpp")], check=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-I", CSRC, "-I", INC, "-o", so,
                    os.path.join(HERE, "native", "engine_mock.cpp"), os.path.join(CSRC, "fit_core.cpp"),
                    os.path.join(CSRC, "fit_engine.cpp"), vm, "-lmvec", "-lm"], check=True)
    L = ctypes.CDLL(so)
    for name, (res, args) in _lib.SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is not None:
            fn.restype, fn.argtypes = res, args
    L.mock_create.restype = ctypes.c_void_p
    L.mock_create.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.mock_destroy.argtypes = [ctypes.c_void_p]
    L.mock_fail_step.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.mock_set_free_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.mock_steps.restype = L.mock_live_allocs.restype = L.mock_parked.restype = ctypes.c_int64
    L.mock_steps.argtypes = L.mock_live_allocs.argtypes = L.mock_parked.argtypes = [ctypes.c_void_p]
    return L


@pytest.fixture(scope="module")
def cohort():
    c = synth.small_cohort(N, M, maf_lo=0.05, maf_hi=0.5, miss_rate=0.0, seed=synth.config_seed(1), ncausal=150)
    og = O.setgeno(c["bed"], N, M, min_maf=0.01, max_missing=0.15)
    K = O.grm_f64(og) / og["Mqc"]
    X, _, _ = plink.covariate_transform(c["X"])
    g = synth.polygenic(c["Zc"], 2.0, c["seed"])
    rng = np.random.default_rng(5)
    sub = np.sort(rng.choice(N, 180, replace=False))
    Xs, _, _ = plink.covariate_transform(c["X"][sub])
    yb = synth.binary_phenotype(c["X"], g, 0.3, c["seed"])
    jobs = [
        {"trait": "binary", "y": yb, "X": X},
        {"trait": "binary", "y": synth.binary_phenotype(c["X"], g, 0.15, c["seed"] + 1), "X": X},
        {"trait": "quantitative", "X": X,
         "y": synth.quantitative_phenotype(c["X"], synth.polygenic(c["Zc"], 0.5, 3), 0.5, c["seed"])},
        {"trait": "binary", "y": yb[sub], "X": Xs, "subset": sub},
    ]
    return np.asfortranarray(K), jobs


def _ctx(L, K):
    return ctypes.c_void_p(L.mock_create(K.ctypes.data_as(ctypes.c_void_p), N))


def _batch(L, ctx, jobs, opt):
    specs = (_lib.FitSpec * len(jobs))()
    results = (_lib.FitResult * len(jobs))()
    keep = [api._fit_spec(j, specs[i]) for i, j in enumerate(jobs)]
    outs = [api._fit_result(j, results[i]) for i, j in enumerate(jobs)]
    o = api._fit_options(L, opt)
    rc = L.saige_fit_null_batch(ctx, len(jobs), specs, ctypes.byref(o), 16, results)
    del keep
    return rc, [api._result_dict(results[i], outs[i], j["trait"]) for i, j in enumerate(jobs)]


def _same(a, b):
    for k in ("theta", "coefficients", "fitted_values", "linear_predictors"):
        assert np.array_equal(a[k], b[k]), k
    assert a["iterations"] == b["iterations"]


OPT = D.Options(nrun=10, tolPCG=1e-6)


def test_batch_alone_mixed_and_fp64_oracle(lib, cohort):
    K, jobs = cohort
    ctx = _ctx(lib, K)
    try:
        rc, mixed = _batch(lib, ctx, jobs, OPT)
        assert rc == _lib.SAIGE_OK and all(r["status"] == 0 for r in mixed)
        parked = lib.mock_parked(ctx)
        assert parked > 0 and lib.mock_live_allocs(ctx) == 0     # trace buffers back on the context
        for j, r in zip(jobs, mixed):
            rc1, alone = _batch(lib, ctx, [j], OPT)
            assert rc1 == _lib.SAIGE_OK
            _same(alone[0], r)
        assert lib.mock_live_allocs(ctx) == 0 and lib.mock_parked(ctx) == parked   # reused, none made
        # tau against the FP64 oracle driver on the same K and probes
        grm = O.DenseGRM(K, 1)
        kw = dict(nrun=OPT.nrun, tolPCG=OPT.tolPCG, maxiterPCG=OPT.maxiterPCG, cv_cutoff=OPT.traceCVcutoff,
                  tol=OPT.tol, maxiter=OPT.maxiter)
        for j, r in zip(jobs[:3], mixed[:3]):
            f = O.glmmkin_binary if j["trait"] == "binary" else O.glmmkin_quantitative
            ref = f(grm, j["y"], j["X"], D.RademacherProbes(N, 200), **kw)
            assert r["iterations"] == ref["iterations"]
            assert np.allclose(r["theta"], ref["theta"], rtol=1e-3, atol=1e-6), (r["theta"], ref["theta"])
        assert mixed[0]["theta"][1] > 0.01
    finally:
        lib.mock_destroy(ctx)


def _serve(L, ctx, jobs, opt, delays, reserve=0, poll_wait=-1):
    o = api._fit_options(L, opt)
    h = ctypes.c_void_p()
    assert L.saige_server_start(ctx, ctypes.byref(o), 4, reserve, ctypes.byref(h)) == _lib.SAIGE_OK
    tickets, keep = [], []
    for j, d in zip(jobs, delays):
        time.sleep(d)
        spec = _lib.FitSpec()
        keep.append(api._fit_spec(j, spec))
        t = ctypes.c_int64(0)
        rc = L.saige_submit(h, ctypes.byref(spec), ctypes.byref(t))
        # after an engine failure the server has stopped and refuses requests
        assert rc in (_lib.SAIGE_OK, _lib.SAIGE_E_STATE)
        tickets.append(t.value if rc == _lib.SAIGE_OK else None)
    del keep                                        # the server holds its own copies
    out = []
    for j, t in zip(jobs, tickets):
        if t is None:
            out.append({"status": _lib.SAIGE_E_STATE})
            continue
        res = _lib.FitResult()
        o2 = api._fit_result(j, res)
        times = np.zeros(2)
        rc = L.saige_poll(h, t, poll_wait, ctypes.byref(res), times.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == _lib.SAIGE_OK
        r = api._result_dict(res, o2, j["trait"])
        r["t_admit"], r["t_done"] = times
        out.append(r)
    return h, out


def test_server_staggered_arrivals_bitwise_equal_to_batch(lib, cohort):
    K, jobs = cohort
    ctx = _ctx(lib, K)
    try:
        _, ref = _batch(lib, ctx, jobs, OPT)
        h, got = _serve(lib, ctx, jobs, OPT, [0.0, 0.01, 0.03, 0.0])
        assert lib.saige_server_stop(h) == _lib.SAIGE_OK
        assert len(got) == len(ref)
        for a, b in zip(got, ref):
            assert a["status"] == 0
            _same(a, b)
            assert 0 <= a["t_admit"] <= a["t_done"]
        assert lib.mock_live_allocs(ctx) == 0
    finally:
        lib.mock_destroy(ctx)


def test_engine_failure_ends_every_fit(lib, cohort):
    K, jobs = cohort
    ctx = _ctx(lib, K)
    try:
        lib.mock_fail_step(ctx, 5)
        rc, res = _batch(lib, ctx, jobs, OPT)
        assert rc == _lib.SAIGE_E_DEVICE
        assert all(r["status"] != 0 for r in res)
        assert b"injected" in lib.saige_last_error()
        # the server: every request ends with an error, stop reports it
        lib.mock_fail_step(ctx, lib.mock_steps(ctx) + 3)
        h, got = _serve(lib, ctx, jobs, OPT, [0.0] * len(jobs))
        assert all(r["status"] != 0 for r in got)
        assert lib.saige_server_stop(h) == _lib.SAIGE_E_DEVICE
        assert lib.mock_live_allocs(ctx) == 0
    finally:
        lib.mock_destroy(ctx)


def test_admission_waits_for_free_memory(lib, cohort):
    K, jobs = cohort
    ctx = _ctx(lib, K)
    try:
        lib.mock_set_free_bytes(ctx, 1 << 20)
        timer = threading.Timer(0.3, lambda: lib.mock_set_free_bytes(ctx, 1 << 40))
        timer.start()
        h, got = _serve(lib, ctx, jobs[:2], OPT, [0.0, 0.0], reserve=1 << 30)
        timer.join()
        assert lib.saige_server_stop(h) == _lib.SAIGE_OK
        for r in got:
            assert r["status"] == 0
            assert r["t_admit"] >= 0.25          # held in the queue until memory was reported free
    finally:
        lib.mock_destroy(ctx)


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_scheduler_under_host_sanitizers(tmp_path, san):
    """engine_mock_main.cpp (batch, alone vs mix, server with submitting and
    polling threads, injected failure; the fits' loops in pieces on the host
    thread pool) built with g++ -fsanitize=... on the host code only, run to
    completion with no report."""
    exe = str(tmp_path / "mock_main")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", f"-fsanitize={san}", "-fno-omit-frame-pointer",
           "-I", CSRC, "-I", INC, "-o", exe, os.path.join(HERE, "native", "engine_mock_main.cpp"),
           os.path.join(HERE, "native", "engine_mock.cpp"), os.path.join(CSRC, "fit_core.cpp"),
           os.path.join(CSRC, "fit_engine.cpp"), os.path.join(CSRC, "vmath.cpp"), "-lmvec", "-lm"]
    subprocess.run(cmd, check=True)
    # SAIGE_HOST_PIECE_MIN=64: the fits' O(n) loops run as pieces on the
    # host thread pool even at this size, so the sanitizers watch it too
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1",
               SAIGE_HOST_PIECE_MIN="64")
    p = subprocess.run([exe], env=env, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-6000:]
    assert "all scenarios passed" in p.stdout
