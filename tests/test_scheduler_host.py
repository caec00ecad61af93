"""The batch scheduler's host logic (saige_gpu.driver.BatchScheduler) on the
CPU: slot and system bookkeeping, sample-subset compaction, product slots,
lowest-slot-first packing and the overlap of host work with the in-flight
iteration.  A numpy stand-in for the slot engine (same ABI calls as
saige_gpu.api.Session, FP64 PCG per column with the stopping rule of
FG.cpp:2696-2792) runs under the scheduler; every fit must come out bitwise
equal to the same coroutine run one column at a time on the same engine."""
import collections

import numpy as np
import pytest

from fixtures import random_bed
from oracle import oracle as O
from saige_gpu import driver as D
from saige_gpu import plink


class FakeEngine:
    """Slot engine semantics in FP64 numpy: a column's arithmetic depends only
    on its own right-hand side and system, as on the device."""

    def __init__(self, K, kdiag):
        self.K = np.asarray(K, np.float64)
        self.kdiag = np.asarray(kdiag, np.float64)
        self.N = self.K.shape[0]
        self.calls = collections.Counter()
        self.width = 0

    # the Session surface the scheduler uses
    def getNnomissingOut(self):
        return self.N

    def device_memory(self):
        return 1 << 40, 1 << 40

    def engine_system_bytes(self):
        return 9 * self.N

    def engine_reserve(self, width, n_systems):
        self.width = width
        self.systems = [None] * n_systems
        self.slots = [None] * width

    def engine_set_system(self, s, w, tau, mask=None):
        self.calls["set_system"] += 1
        idx = np.arange(self.N) if mask is None else np.nonzero(np.asarray(mask))[0]
        w = np.asarray(w, np.float32).astype(np.float64)[idx]
        tau = np.asarray(tau, np.float32).astype(np.float64)
        d = np.maximum(tau[1] * self.kdiag[idx] + tau[0] / w, 1e-4)
        self.systems[s] = {"idx": idx, "w": w, "tau": tau, "minv": 1.0 / d, "K": self.K[np.ix_(idx, idx)]}

    def _sigma(self, sy, p):
        if sy["tau"][1] == 0:
            return sy["tau"][0] * (p / sy["w"])
        return sy["tau"][0] * (p / sy["w"]) + sy["tau"][1] * (sy["K"] @ p)

    def _add(self, slots, systems, B, kind, maxiter=0, tol=0.0):
        B = np.asarray(B, np.float32)
        for j, (slot, s) in enumerate(zip(slots, systems)):
            assert self.slots[slot] is None, "slot reused while occupied"
            sy = self.systems[s]
            b = B[:, j].astype(np.float64)[sy["idx"]]
            c = {"sys": s, "kind": kind, "it": 0, "maxiter": maxiter, "tol": tol}
            if kind == "solve":
                c.update(x=np.zeros_like(b), r=b.copy(), z=sy["minv"] * b)
                c["p"] = c["z"].copy()
                c["sumr2"] = c["r"] @ c["r"]
                c["done"] = not (c["sumr2"] > tol and maxiter > 0)
            else:
                c.update(b=b, done=False)
            self.slots[slot] = c

    def engine_add_batch(self, slots, systems, B, maxiter, tol):
        self.calls["add"] += 1
        self._add(slots, systems, B, "solve", maxiter, tol)

    def engine_add_products(self, slots, systems, B):
        self.calls["add_products"] += 1
        self._add(slots, systems, B, "product")

    def engine_step(self, n_iter, width=64):
        self.calls["step"] += 1
        for _ in range(n_iter):
            for c in self.slots:
                if c is None or c["done"]:
                    continue
                sy = self.systems[c["sys"]]
                if c["kind"] == "product":
                    c["x"] = (sy["K"] @ c["b"]) if sy["tau"] is not None else None
                    c["it"], c["done"] = 1, True
                    continue
                c["it"] += 1
                Ap = self._sigma(sy, c["p"])
                a = (c["r"] @ c["z"]) / (c["p"] @ Ap)
                c["x"] = c["x"] + a * c["p"]
                r1 = c["r"] - a * Ap
                z1 = sy["minv"] * r1
                bet = (z1 @ r1) / (c["z"] @ c["r"])
                c["p"] = z1 + bet * c["p"]
                c["z"], c["r"] = z1, r1
                c["sumr2"] = c["r"] @ c["r"]
                c["done"] = not (c["sumr2"] > c["tol"] and c["it"] < c["maxiter"])
        done = np.array([0 if c is None else int(c["done"]) for c in self.slots] + [0] * (width - self.width),
                        np.uint8)
        iters = np.array([0 if c is None else c["it"] for c in self.slots] + [0] * (width - self.width), np.int32)
        return done, iters

    def engine_take_batch(self, slots):
        self.calls["take"] += 1
        X = np.zeros((self.N, len(slots)), np.float32, order="F")
        for j, slot in enumerate(slots):
            c = self.slots[slot]
            assert c is not None and c["done"], "take of a slot that is not done"
            X[self.systems[c["sys"]]["idx"], j] = c["x"]
            self.slots[slot] = None
        return X


class _DevBuf:
    """A stand-in device buffer: a column-major FP32 array with an address."""
    _next = 1

    def __init__(self, nbytes, N):
        self.value = _DevBuf._next
        _DevBuf._next += 1
        self.a = np.zeros((N, max(1, nbytes // (4 * N))), np.float32, order="F")


class FakeDeviceEngine(FakeEngine):
    """FakeEngine plus the device-resident trace calls of saige_gpu.api.Session
    (buffers, probes generated from the probe hash, FP64 column dots)."""

    def alloc_device(self, nbytes):
        self.calls["alloc"] += 1
        return _DevBuf(nbytes, self.N)

    def free_device(self, buf):
        self.calls["free"] += 1

    def copy_host_to_device(self, buf, arr, col0=0):
        a = np.asarray(arr, np.float32).reshape(self.N, -1)
        buf.a[:, col0:col0 + a.shape[1]] = a

    def probes_device(self, buf, col0, seed, i0, k):
        for j in range(k):
            buf.a[:, col0 + j] = D.probe_vector(seed, i0 + j, self.N)

    def copy_bytes_to_device(self, dst, arr):
        dst.rows = np.array(arr, np.int64)

    def probes_device_rows(self, buf, col0, seed, i0, k, rows_buf, n):
        self.calls["probes_rows"] += 1
        rows = rows_buf.rows[:n]
        for j in range(k):
            buf.a[:, col0 + j] = 0
            buf.a[rows, col0 + j] = D.probe_vector(seed, i0 + j, n)

    def dots_device(self, A, k, Q, m):
        a, q = A.a.astype(np.float64), Q.a.astype(np.float64)
        return np.array([[np.dot(a[:, j], q[:, r]) for r in range(m)] for j in range(k)])

    def engine_add_batch_device(self, slots, systems, buf, col0, maxiter, tol):
        self._add(slots, systems, buf.a[:, col0:col0 + len(slots)], "solve", maxiter, tol)

    def engine_add_products_device(self, slots, systems, buf, col0):
        self._add(slots, systems, buf.a[:, col0:col0 + len(slots)], "product")

    def engine_take_batch_ex(self, slots, download=False, X_dev=None, xcols=None, P1=None, P2=None, pcols=None,
                             Q=None, m=0):
        self.calls["take_ex"] += 1
        X = self.engine_take_batch(slots)
        xf = X.astype(np.float64)
        dots = []
        for j in range(len(slots)):
            row = []
            pc = pcols[j] if pcols is not None else j
            for P in (P1, P2):
                if P is not None:
                    row.append(np.dot(xf[:, j], P.a[:, pc].astype(np.float64)))
            for r in range(m):
                row.append(np.dot(xf[:, j], Q.a[:, r].astype(np.float64)))
            dots.append(row)
            if X_dev is not None:
                X_dev.a[:, xcols[j] if xcols is not None else j] = X[:, j]
        return (X if download else None), np.array(dots).reshape(len(slots), -1)


class OneColumnExecutor:
    """Runs a coroutine with every column alone on the engine (reference)."""

    def __init__(self, eng, opt, subset=None):
        self.eng, self.opt, self.subset = eng, opt, subset

    def _col(self, w, tau, b, kind):
        N = self.eng.N
        mask = None
        wf, bf = w, b
        if self.subset is not None:
            mask = np.zeros(N, np.uint8)
            mask[self.subset] = 1
            wf = np.ones(N)
            wf[self.subset] = w
            bf = np.zeros(N)
            bf[self.subset] = b
        self.eng.engine_reserve(1, 1)
        self.eng.engine_set_system(0, wf, np.asarray(tau, np.float32), mask)
        B = np.asarray(bf, np.float32)[:, None]
        if kind == "solve":
            self.eng.engine_add_batch([0], [0], B, self.opt.maxiterPCG, self.opt.tolPCG)
        else:
            self.eng.engine_add_products([0], [0], B)
        while not self.eng.engine_step(1, 1)[0][0]:
            pass
        x = self.eng.engine_take_batch([0])[:, 0]
        return x if self.subset is None else x[self.subset]

    def run(self, gen):
        value, w, tau = None, None, None
        while True:
            try:
                req = gen.send(value)
            except StopIteration as e:
                return e.value
            if req[0] == "solve":
                _, w, tau, B = req
                value = self._solve(w, tau, B)
            elif req[0] == "matvec":
                value = self._matvec(req[1], w)
            else:
                w, tau = req[1], req[2]
                value = D.trace_round_host(req, lambda B: self._matvec(B, w), self._solve)

    def _solve(self, w, tau, B):
        v = np.column_stack([self._col(w, tau, B[:, j], "solve") for j in range(B.shape[1])])
        return np.asfortranarray(v, dtype=np.float64)   # the scheduler's result layout

    def _matvec(self, B, w):
        ww = w if w is not None else np.ones(B.shape[0])
        v = np.column_stack([self._col(ww, (1.0, 0.0), B[:, j], "product") for j in range(B.shape[1])])
        return np.asfortranarray(v, dtype=np.float64)


@pytest.fixture(scope="module")
def cohort():
    N, M = 160, 500
    body, _, _ = random_bed(N, M, 0.005, seed=41, maf_lo=0.05, maf_hi=0.5)
    g = O.setgeno(body, N, M, min_maf=0.01)
    S = O.grm_f64(g)
    K = S / g["Mqc"]
    kdiag = O.diag_std(g).astype(np.float64) / g["Mqc"]
    rng = np.random.default_rng(5)
    X = np.column_stack([np.ones(N), rng.normal(0, 1, N)])
    Xt, _, _ = plink.covariate_transform(X)
    gz = O.std_matrix(g).astype(np.float64)[:, :60] @ rng.normal(0, np.sqrt(0.6 / 60), 60)
    jobs = []
    for k, prev in enumerate([-1.0, -0.3, 0.2]):
        y = (rng.random(N) < 1 / (1 + np.exp(-(prev + gz)))).astype(float)
        jobs.append({"trait": "binary", "y": y, "X": Xt, "tag": f"b{k}"})
    jobs.append({"trait": "quantitative", "y": gz + rng.normal(0, 0.8, N), "X": Xt, "tag": "q"})
    sub = np.sort(rng.choice(N, 110, replace=False))
    Xs, _, _ = plink.covariate_transform(X[sub])
    ys = (rng.random(110) < 1 / (1 + np.exp(-(-0.5 + gz[sub])))).astype(float)
    jobs.append({"trait": "binary", "y": ys, "X": Xs, "tag": "sub", "subset": sub})
    return K, kdiag, jobs


def test_scheduler_equals_one_column_at_a_time(cohort):
    """Without device-resident trace rounds (products and solves only) the
    scheduler's results are bitwise those of one column at a time."""
    K, kdiag, jobs = cohort
    opt = D.Options(nrun=6, traceCVcutoff=1.0, maxiter=6, tolPCG=1e-8, maxiterPCG=300)
    eng = FakeEngine(K, kdiag)
    handles, sch = D.fit_batched(eng, jobs, opt, width=64, n_systems=3)
    assert not sch.device_trace
    assert all(h.error is None for h in handles)
    # columns of several fits shared iterations, and host work overlapped an iteration
    assert sch.n_iterations < sum(h.iters for h in handles)
    assert sch.n_overlapped > 0
    for j, h in zip(jobs, handles):
        probes = D.RademacherProbes(len(j["y"]), 200)
        ref = OneColumnExecutor(FakeEngine(K, kdiag), opt, j.get("subset")).run(
            D.fit_co(j["trait"], j["y"], j["X"], probes, opt, lib=True))
        assert np.array_equal(h.result["theta"], ref["theta"]), j["tag"]
        assert np.array_equal(h.result["coefficients"], ref["coefficients"]), j["tag"]
        assert np.array_equal(h.result["linear_predictors"], ref["linear_predictors"]), j["tag"]
        assert h.result["iterations"] == ref["iterations"]


def test_device_trace_rounds_match(cohort):
    """With device-resident trace rounds (probes generated in place, K u
    stashed, the Sigma^-1 u solves returning only FP64 dots) every fit matches
    the one-column run: the same iterations, tau to FP64 rounding (the
    device keeps Sigma^-1 X in FP32 for the h and g dots)."""
    K, kdiag, jobs = cohort
    opt = D.Options(nrun=6, traceCVcutoff=1.0, maxiter=6, tolPCG=1e-8, maxiterPCG=300)
    eng = FakeDeviceEngine(K, kdiag)
    handles, sch = D.fit_batched(eng, jobs, opt, width=64, n_systems=3)
    assert sch.device_trace
    assert all(h.error is None for h in handles), [h.error for h in handles]
    assert eng.calls["take_ex"] > 0 and eng.calls["alloc"] > 0 and eng.calls["free"] == 0
    assert eng.calls["probes_rows"] > 0                  # the subset fit's probes made in place
    # the buffers stay on the engine for the next scheduler, and free once
    n_alloc = eng.calls["alloc"]
    D.fit_batched(eng, jobs[:1], opt, width=64, n_systems=3)
    assert eng.calls["alloc"] == n_alloc and eng.calls["free"] == 0
    D.release_trace_buffers(eng)
    assert eng.calls["alloc"] == eng.calls["free"]
    for j, h in zip(jobs, handles):
        probes = D.RademacherProbes(len(j["y"]), 200)
        ref = OneColumnExecutor(FakeEngine(K, kdiag), opt, j.get("subset")).run(
            D.fit_co(j["trait"], j["y"], j["X"], probes, opt, lib=True))
        assert h.result["iterations"] == ref["iterations"], j["tag"]
        assert np.allclose(h.result["theta"], ref["theta"], rtol=1e-6, atol=1e-9), j["tag"]
        assert np.allclose(h.result["coefficients"], ref["coefficients"], rtol=1e-6, atol=1e-9), j["tag"]


def test_probe_hash_is_a_fair_sign():
    """probe_vector: entries +-1, half of each over many rows, independent
    across probes and seeds, fixed for a (seed, i)."""
    a = D.probe_vector(200, 0, 200_000)
    b = D.probe_vector(200, 1, 200_000)
    c = D.probe_vector(201, 0, 200_000)
    assert set(np.unique(a)) == {-1.0, 1.0}
    assert abs(a.mean()) < 0.01 and abs((a * b).mean()) < 0.01 and abs((a * c).mean()) < 0.01
    assert np.array_equal(a, D.probe_vector(200, 0, 200_000))
    assert np.array_equal(a[:1000], D.probe_vector(200, 0, 1000))


def test_scheduler_packs_low_slots_and_caps_systems(cohort):
    K, kdiag, jobs = cohort
    opt = D.Options(nrun=4, traceCVcutoff=1.0, maxiter=3, tolPCG=1e-6, maxiterPCG=200)
    eng = FakeEngine(K, kdiag)
    seen = []
    orig = eng.engine_add_batch

    def spy(slots, systems, B, maxiter, tol):
        seen.append(list(slots))
        assert max(systems) < 2              # n_systems caps concurrent fits
        return orig(slots, systems, B, maxiter, tol)

    eng.engine_add_batch = spy
    handles, sch = D.fit_batched(eng, jobs[:3], opt, width=64, n_systems=2)
    assert all(h.error is None for h in handles)
    # the first request of the first fit takes the lowest slots
    assert seen[0] == list(range(len(seen[0])))
    # with at most two fits of (1 + p) or nrun + 1 columns, no slot above 31 is ever used
    assert max(max(s) for s in seen) < 32


def test_trace_buffers_survive_schedulers_without_trace_rounds(cohort):
    """A scheduler that runs no trace round (variance ratios, LOCO, a
    caller's own solves) leaves the parked buffers on the session, closed or
    not, so release_trace_buffers frees every buffer fit_batched made."""
    K, kdiag, jobs = cohort
    opt = D.Options(nrun=4, traceCVcutoff=1.0, maxiter=3, tolPCG=1e-6, maxiterPCG=200)
    eng = FakeDeviceEngine(K, kdiag)
    handles, _ = D.fit_batched(eng, jobs[:2], opt, width=64, n_systems=3)
    assert all(h.error is None for h in handles) and eng.calls["alloc"] > 0
    fit = handles[0].result

    def solves():
        w = fit["fitted_values"] * (1 - fit["fitted_values"])
        x = yield ("solve", w, fit["theta"], np.ones((len(w), 2)))
        return x

    sch = D.BatchScheduler(eng, opt, width=64, n_systems=1)    # never closed
    h = sch.submit(solves())
    sch.run()
    assert h.error is None
    with D.BatchScheduler(eng, opt, width=64, n_systems=1) as sch2:
        h2 = sch2.submit(solves())
        sch2.run()
    assert np.array_equal(h.result, h2.result)
    D.release_trace_buffers(eng)
    assert eng.calls["alloc"] == eng.calls["free"]


def test_trace_round_beyond_device_buffers_runs_on_host_path(cohort):
    """A fit with more covariates than the device trace buffers hold (16 > 14)
    runs beside a small fit: its trace rounds take the products-and-solves
    path, the small fit keeps the device path, and both match one column at
    a time."""
    K, kdiag, jobs = cohort
    N = K.shape[0]
    rng = np.random.default_rng(12)
    X16, _, _ = plink.covariate_transform(np.column_stack([np.ones(N), rng.standard_normal((N, 15))]))
    assert X16.shape[1] == 16 > D.BatchScheduler.TRACE_MAX_COVARIATES
    big = {"trait": "binary", "y": jobs[0]["y"], "X": X16, "tag": "p16"}
    opt = D.Options(nrun=6, traceCVcutoff=1.0, maxiter=4, tolPCG=1e-8, maxiterPCG=300)
    eng = FakeDeviceEngine(K, kdiag)
    handles, sch = D.fit_batched(eng, [jobs[0], big], opt, width=64, n_systems=3)
    assert all(h.error is None for h in handles), [h.error for h in handles]
    assert eng.calls["take_ex"] > 0
    for j, h in zip([jobs[0], big], handles):
        probes = D.RademacherProbes(N, 200)
        ref = OneColumnExecutor(FakeEngine(K, kdiag), opt).run(
            D.fit_co(j["trait"], j["y"], j["X"], probes, opt, lib=True))
        assert h.result["iterations"] == ref["iterations"], j["tag"]
        assert np.allclose(h.result["theta"], ref["theta"], rtol=1e-6, atol=1e-9), j["tag"]
    D.release_trace_buffers(eng)
