"""Host-side mirror of the reference's Rcpp exports for SAIGE step 1's hot
path (src/SAIGE/src/SAIGE_fitGLMM_fast.cpp, FG.cpp below), over libsaigehip.

A Session owns one device context; it plays the role of the reference's
process-global `genoClass geno` (FG.cpp:1188).  Method names follow the
Rcpp export names, argument meaning and order follow their signatures, and
failures raise (the reference raises R errors through BEGIN_RCPP/END_RCPP).
Vectors are numpy float32 of length Nnomissing.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, cols_f32, f32, ptr


class Session:
    def __init__(self, device=0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        check(self._L.saige_create(int(device), ctypes.byref(h)))
        self.ctx = h
        self.device = device

    def close(self):
        if self.ctx:
            # device buffers a batch scheduler parked here for the next one
            park = self.__dict__.pop("_trace_buffers", None)
            if park is not None:
                for bufs in park[1].values():
                    for b in bufs.values():
                        self.free_device(b)
            self._L.saige_destroy(self.ctx)
            self.ctx = None

    closeGenoFile_plink = close  # FG.cpp:1191-1209

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ setters
    def setminMAFforGRM(self, minMAFforGRM):                      # FG.cpp:4923
        check(self._L.saige_set_min_maf_for_grm(self.ctx, float(minMAFforGRM)))

    def setmaxMissingRateforGRM(self, maxMissingforGRM):          # FG.cpp:4928
        check(self._L.saige_set_max_missing_rate_for_grm(self.ctx, float(maxMissingforGRM)))

    def setminMAC_VarianceRatio(self, t_minMACVarRatio, t_maxMACVarRatio,
                                t_isVarianceRatioinGeno):          # FG.cpp:4970
        check(self._L.saige_set_min_mac_variance_ratio(self.ctx, float(t_minMACVarRatio),
                                                       float(t_maxMACVarRatio),
                                                       int(bool(t_isVarianceRatioinGeno))))

    def set_vr_random_markers(self, idx):
        """The common-marker pool g_randMarkerIndforVR (FG.cpp:866-871)."""
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        check(self._L.saige_set_vr_random_markers(self.ctx, ptr(idx), len(idx)))

    # ------------------------------------------------------------ ingest
    def setgeno(self, bedFile, bimFile, famFile, subSampleInGeno, indicatorGenoSamplesWithPheno=None,
                memoryChunk=2.0, isDiagofKinSetAsOne=False):
        """setgeno (FG.cpp:2267-2275).  memoryChunk is accepted and unused, as in
        the reference (numMarkersofEachArray = 1, FG.cpp:824)."""
        sub = np.ascontiguousarray(subSampleInGeno, dtype=np.int32)
        ind = None
        if indicatorGenoSamplesWithPheno is not None:
            ind = np.ascontiguousarray(np.asarray(indicatorGenoSamplesWithPheno).astype(np.uint8))
        check(self._L.saige_setgeno(self.ctx, bedFile.encode(), bimFile.encode(), famFile.encode(),
                                    ptr(sub), len(sub), ptr(ind), int(bool(isDiagofKinSetAsOne))))

    def setgeno_bytes(self, bed_body, N, M, subSampleInGeno=None, indicator=None,
                      isDiagofKinSetAsOne=False):
        """setgeno from an in-memory .bed body (magic bytes stripped)."""
        body = np.ascontiguousarray(bed_body, dtype=np.uint8)
        assert body.size == M * ((N + 3) // 4)
        sub = np.arange(1, N + 1, dtype=np.int32) if subSampleInGeno is None else \
            np.ascontiguousarray(subSampleInGeno, dtype=np.int32)
        ind = None if indicator is None else np.ascontiguousarray(np.asarray(indicator).astype(np.uint8))
        check(self._L.saige_setgeno_host(self.ctx, ptr(body), int(N), int(M), ptr(sub), len(sub), ptr(ind),
                                         int(bool(isDiagofKinSetAsOne))))

    def setgeno_device(self, dev_ptr, N, M, subSampleInGeno=None, indicator=None,
                       isDiagofKinSetAsOne=False):
        sub = np.arange(1, N + 1, dtype=np.int32) if subSampleInGeno is None else \
            np.ascontiguousarray(subSampleInGeno, dtype=np.int32)
        ind = None if indicator is None else np.ascontiguousarray(np.asarray(indicator).astype(np.uint8))
        check(self._L.saige_setgeno_device(self.ctx, dev_ptr, int(N), int(M), ptr(sub), len(sub), ptr(ind),
                                           int(bool(isDiagofKinSetAsOne))))

    # ------------------------------------------------------------ getters
    def getNnomissingOut(self):
        return int(self._L.saige_get_nnomissing(self.ctx))

    def gettotalMarker(self):
        return int(self._L.saige_get_total_markers(self.ctx))

    def getMsub_MAFge_minMAFtoConstructGRM(self):
        return int(self._L.saige_get_mqc(self.ctx))

    def n_vr_markers(self):
        return int(self._L.saige_get_n_vr_markers(self.ctx))

    def getAlleleFreqVec(self):
        out = np.zeros(self.getMsub_MAFge_minMAFtoConstructGRM(), np.float32)
        check(self._L.saige_get_allele_freq(self.ctx, ptr(out)))
        return out

    def getMACVec(self):
        out = np.zeros(self.getMsub_MAFge_minMAFtoConstructGRM(), np.int32)
        check(self._L.saige_get_mac(self.ctx, ptr(out)))
        return out

    def getInvStdVec(self):
        out = np.zeros(self.getMsub_MAFge_minMAFtoConstructGRM(), np.float32)
        check(self._L.saige_get_inv_std(self.ctx, ptr(out)))
        return out

    def getQCdMarkerIndex(self):
        out = np.zeros(self.gettotalMarker(), np.uint8)
        check(self._L.saige_get_qc_marker_index(self.ctx, ptr(out)))
        return out.astype(bool)

    def getMACVec_forVarRatio(self):
        out = np.zeros(self.n_vr_markers(), np.int32)
        check(self._L.saige_get_mac_vr(self.ctx, ptr(out)))
        return out

    def getIndexVec_forVarRatio(self):
        out = np.zeros(self.n_vr_markers(), np.int32)
        check(self._L.saige_get_index_vr(self.ctx, ptr(out)))
        return out

    def marker_stats(self):
        M = self.gettotalMarker()
        out = {k: np.zeros(M, t) for k, t in [("allele_count", np.int32), ("num_missing", np.int32),
                                              ("mac", np.int32), ("fill", np.int32),
                                              ("freq", np.float32), ("missing_rate", np.float32),
                                              ("inv_std", np.float32), ("flags", np.uint8)]}
        check(self._L.saige_get_marker_stats(self.ctx, *[ptr(out[k]) for k in
                                                         ["allele_count", "num_missing", "mac", "fill",
                                                          "freq", "missing_rate", "inv_std", "flags"]]))
        return out

    def Get_OneSNP_Geno(self, SNPIdx):                              # FG.cpp:2281
        out = np.zeros(self.getNnomissingOut(), np.int32)
        check(self._L.saige_get_one_snp_geno(self.ctx, int(SNPIdx), ptr(out)))
        return out

    def Get_OneSNP_Geno_forVarRatio(self, SNPIdx):                  # FG.cpp:2291
        out = np.zeros(self.getNnomissingOut(), np.int32)
        check(self._L.saige_get_one_snp_geno_vr(self.ctx, int(SNPIdx), ptr(out)))
        return out

    def Get_OneSNP_StdGeno(self, SNPIdx):                           # FG.cpp:582-662
        out = np.zeros(self.getNnomissingOut(), np.float32)
        check(self._L.saige_get_one_snp_std_geno(self.ctx, int(SNPIdx), ptr(out)))
        return out

    def Get_Diagof_StdGeno(self):                                   # FG.cpp:665-704
        out = np.zeros(self.getNnomissingOut(), np.float32)
        check(self._L.saige_get_diag_std(self.ctx, ptr(out)))
        return out

    def get_DiagofKin(self):                                        # FG.cpp:4358-4375
        out = np.zeros(self.getNnomissingOut(), np.float32)
        check(self._L.saige_get_diag_of_kin(self.ctx, ptr(out)))
        return out

    get_GRMdiagVec = get_DiagofKin                                  # FG.cpp:4913-4920

    # ------------------------------------------------------------ GRM
    _GRM_PRECISIONS = {"fp32": _lib.GRM_FP32, "bf16": _lib.GRM_BF16, "bf16_panel": _lib.GRM_BF16_PANEL,
                       "bf16_split": _lib.GRM_BF16_SPLIT}

    def build_grm(self, precision="fp32"):
        check(self._L.saige_build_grm(self.ctx, self._GRM_PRECISIONS[precision]))
        return float(self._L.saige_last_grm_seconds(self.ctx))

    def grm_precision(self):
        """The precision name the resident GRM was built with (None if none)."""
        p = int(self._L.saige_grm_precision(self.ctx))
        return next((k for k, v in self._GRM_PRECISIONS.items() if v == p), None)

    def grm_block(self, r0, c0, nr, nc):
        out = np.zeros((nr, nc), np.float32)
        check(self._L.saige_get_grm_block(self.ctx, int(r0), int(c0), int(nr), int(nc), ptr(out)))
        return out

    def getCrossprodMatAndKin(self, bVec):                          # FG.cpp:1953-1981
        b = f32(bVec)
        out = np.zeros_like(b)
        check(self._L.saige_crossprod_mat_and_kin(self.ctx, ptr(b), ptr(out)))
        return out

    parallelCrossProd = getCrossprodMatAndKin                      # FG.cpp:1669-1708

    def grm_matvec(self, B):
        """K B for an N x k block (one GEMM launch per 64 columns)."""
        B = np.asarray(B)
        if B.ndim == 1:
            return self.getCrossprodMatAndKin(B)
        Bc = cols_f32(B)
        out = np.zeros(Bc.shape, np.float32, order="F")
        check(self._L.saige_grm_matvec(self.ctx, ptr(Bc), Bc.shape[1], ptr(out)))
        return out

    def getDiagOfSigma(self, wVec, tauVec):                         # FG.cpp:2322-2365
        w = f32(wVec)
        tau = f32(tauVec)
        out = np.zeros_like(w)
        check(self._L.saige_get_diag_of_sigma(self.ctx, ptr(w), ptr(tau), ptr(out)))
        return out

    def getCrossprod(self, bVec, wVec, tauVec):                     # FG.cpp:2397-2425
        b, w, tau = f32(bVec), f32(wVec), f32(tauVec)
        out = np.zeros_like(b)
        check(self._L.saige_get_crossprod(self.ctx, ptr(b), ptr(w), ptr(tau), ptr(out)))
        return out

    # ------------------------------------------------------------ PCG
    def getPCG1ofSigmaAndVector(self, wVec, tauVec, bVec, maxiterPCG=500, tolPCG=1e-5,
                                return_iters=False):                # FG.cpp:2593-2809
        w, tau, b = f32(wVec), f32(tauVec), f32(bVec)
        x = np.zeros_like(b)
        it = ctypes.c_int32(0)
        rc = self._L.saige_pcg1(self.ctx, ptr(w), ptr(tau), ptr(b), int(maxiterPCG), float(tolPCG),
                                ptr(x), ctypes.byref(it))
        check(rc, allow_notconv=True)   # the reference warns and returns x (FG.cpp:2794-2797)
        return (x, int(it.value)) if return_iters else x

    # ------------------------------------------------------------ LOCO
    def setStartEndIndex(self, startIndex, endIndex, chromIndex):   # FG.cpp:3067-3082, 4934-4958
        """Leave GRM markers [startIndex, endIndex] (QC'd-list indices,
        inclusive) out of K: builds S_loco and diag_loco on the device."""
        check(self._L.saige_set_start_end_index(self.ctx, int(startIndex), int(endIndex), int(chromIndex)))

    def loco_range(self):
        s, e, m = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        ch = ctypes.c_int(0)
        check(self._L.saige_loco_range(self.ctx, ctypes.byref(s), ctypes.byref(e), ctypes.byref(ch),
                                       ctypes.byref(m)))
        return int(s.value), int(e.value), int(ch.value), int(m.value)

    def getCrossprodMatAndKin_LOCO(self, bVec):                     # FG.cpp:1989-2000
        b = f32(bVec)
        out = np.zeros_like(b)
        check(self._L.saige_crossprod_mat_and_kin_loco(self.ctx, ptr(b), ptr(out)))
        return out

    def getDiagOfSigma_LOCO(self, wVec, tauVec):                    # FG.cpp:2368-2392
        w, tau = f32(wVec), f32(tauVec)
        out = np.zeros_like(w)
        check(self._L.saige_get_diag_of_sigma_loco(self.ctx, ptr(w), ptr(tau), ptr(out)))
        return out

    def getCrossprod_LOCO(self, bVec, wVec, tauVec):                # FG.cpp:2430-2460
        b, w, tau = f32(bVec), f32(wVec), f32(tauVec)
        out = np.zeros_like(b)
        check(self._L.saige_get_crossprod_loco(self.ctx, ptr(b), ptr(w), ptr(tau), ptr(out)))
        return out

    def getPCG1ofSigmaAndVector_LOCO(self, wVec, tauVec, bVec, maxiterPCG=500, tolPCG=1e-5,
                                     return_iters=False):           # FG.cpp:2943-3030
        w, tau, b = f32(wVec), f32(tauVec), f32(bVec)
        x = np.zeros_like(b)
        it = ctypes.c_int32(0)
        rc = self._L.saige_pcg1_loco(self.ctx, ptr(w), ptr(tau), ptr(b), int(maxiterPCG), float(tolPCG),
                                     ptr(x), ctypes.byref(it))
        check(rc, allow_notconv=True)
        return (x, int(it.value)) if return_iters else x

    # ------------------------------------------------------- autotuner
    def autotune(self, reps=3):
        """Time every block-GEMM shape on the resident GRM and keep the
        fastest.  Returns {"rt", "pd", "nsplit", "ms", "table": [(rt, pd,
        nsplit, ms)]}."""
        n = self._L.saige_autotune_candidates(None, None, None)
        rt = np.zeros(n, np.int32)
        pd = np.zeros(n, np.int32)
        ns = np.zeros(n, np.int32)
        self._L.saige_autotune_candidates(ptr(rt), ptr(pd), ptr(ns))
        ms = np.zeros(n + 8, np.float64)
        check(self._L.saige_autotune(self.ctx, int(reps), ptr(ms)))
        cfg = self.gemm_config()
        # shapes a mode does not use read -1 and are left out
        shapes = [(1, 1), (2, 1), (1, 3), (2, 3)]
        cfg["table"] = [(int(a), int(b), int(c), float(t)) for a, b, c, t in zip(rt, pd, ns, ms[:n]) if t >= 0]
        cfg["half_table"] = [(a, b, cfg["nsplit"], float(t)) for (a, b), t in zip(shapes, ms[n:n + 4]) if t >= 0]
        cfg["quarter_table"] = [(a, b, cfg["nsplit"], float(t)) for (a, b), t in zip(shapes, ms[n + 4:]) if t >= 0]
        return cfg

    def set_gemm_config(self, rt, pd, nsplit):
        check(self._L.saige_set_gemm_config(self.ctx, int(rt), int(pd), int(nsplit)))

    GEMM_MODES = ("fp32", "bf16x3", "bf16x3_tri")

    def set_gemm_mode(self, mode):
        """Block-GEMM arithmetic: "fp32" (FP32 MFMA), "bf16x3" (FP32
        operands split exactly into three BF16 pieces, six piece products on
        the BF16 MFMA, FP32 accumulation) or "bf16x3_tri" (the same products
        read from the GRM's lower triangle only, each block used for S_IJ P_J
        and S_IJ^T P_I).  Clears tuned times."""
        check(self._L.saige_set_gemm_mode(self.ctx, self.GEMM_MODES.index(mode)))

    GEMV_MODES = ("full", "tri")

    def set_gemv_mode(self, mode):
        """Single-column GRM products: "tri" (the default; the GRM's lower
        triangle only, each element used for s_ij x_j and s_ij x_i) or
        "full" (all of S)."""
        check(self._L.saige_set_gemv_mode(self.ctx, self.GEMV_MODES.index(mode)))

    def gemv_mode(self):
        m = ctypes.c_int(0)
        check(self._L.saige_get_gemv_mode(self.ctx, ctypes.byref(m)))
        return self.GEMV_MODES[m.value]

    def gemm_mode(self):
        m = ctypes.c_int(0)
        check(self._L.saige_get_gemm_mode(self.ctx, ctypes.byref(m)))
        return self.GEMM_MODES[m.value]

    def gemm_config(self):
        rt, pd, ns, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        check(self._L.saige_get_gemm_config(self.ctx, ctypes.byref(rt), ctypes.byref(pd), ctypes.byref(ns),
                                            ctypes.byref(ms)))
        hr, hp, hm = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        check(self._L.saige_get_gemm_config_half(self.ctx, ctypes.byref(hr), ctypes.byref(hp), ctypes.byref(hm)))
        qr, qp, qm = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        check(self._L.saige_get_gemm_config_quarter(self.ctx, ctypes.byref(qr), ctypes.byref(qp), ctypes.byref(qm)))
        return {"rt": int(rt.value), "pd": int(pd.value), "nsplit": int(ns.value), "ms": float(ms.value),
                "mode": self.gemm_mode(),
                "half": {"rt": int(hr.value), "pd": int(hp.value), "ms": float(hm.value)},
                "quarter": {"rt": int(qr.value), "pd": int(qp.value), "ms": float(qm.value)}}

    def engine_use_loco(self, on):
        """Point the block engine at the LOCO GRM (on) or the full one."""
        check(self._L.saige_engine_use_loco(self.ctx, 1 if on else 0))

    def pcg_block(self, wVec, tauVec, B, maxiterPCG=500, tolPCG=1e-5):
        """Block PCG: k right-hand sides, same Sigma, one GEMM per iteration."""
        w, tau = f32(wVec), f32(tauVec)
        Bc = cols_f32(B if np.ndim(B) == 2 else np.asarray(B)[:, None])
        k = Bc.shape[1]
        X = np.zeros(Bc.shape, np.float32, order="F")
        iters = np.zeros(k, np.int32)
        rc = self._L.saige_pcg_block(self.ctx, ptr(w), ptr(tau), ptr(Bc), k, int(maxiterPCG), float(tolPCG),
                                     ptr(X), ptr(iters))
        check(rc, allow_notconv=True)
        return X, iters

    def getSigma_X(self, wVec, tauVec, Xmat, maxiterPCG=500, tolPCG=1e-5):   # FG.cpp:3347-3366
        X = np.asarray(Xmat, np.float32)
        out = np.zeros(X.shape, np.float32)
        for i in range(X.shape[1]):
            out[:, i] = self.getPCG1ofSigmaAndVector(wVec, tauVec, X[:, i], maxiterPCG, tolPCG)
        return out

    def getSigma_G(self, wVec, tauVec, Gvec, maxiterPCG=500, tolPCG=1e-5):   # FG.cpp:3394-3398
        return self.getPCG1ofSigmaAndVector(wVec, tauVec, Gvec, maxiterPCG, tolPCG)

    # ------------------------------------------------------ engine (batched)
    def engine_reserve(self, width, n_systems):
        check(self._L.saige_engine_reserve(self.ctx, int(width), int(n_systems)))

    def engine_set_system(self, s, w, tau, mask=None):
        w = f32(w)
        tau = f32(tau)
        m = None if mask is None else np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
        check(self._L.saige_engine_set_system(self.ctx, int(s), ptr(w), ptr(tau), ptr(m)))

    def engine_add(self, slot, s, b, maxiter, tol):
        b = f32(b)
        check(self._L.saige_engine_add(self.ctx, int(slot), int(s), ptr(b), int(maxiter), float(tol)))

    def engine_add_batch(self, slots, systems, B, maxiter, tol):
        """Start len(slots) solves: column j of B (N x k) in slot slots[j]."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        sy = np.ascontiguousarray(systems, dtype=np.int32)
        Bc = cols_f32(B)
        assert Bc.shape[1] == len(sl)
        check(self._L.saige_engine_add_batch(self.ctx, len(sl), ptr(sl), ptr(sy), ptr(Bc), int(maxiter),
                                             float(tol)))

    def engine_add_products(self, slots, systems, B):
        """Queue GRM products K b (column j of B in slot slots[j])."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        sy = np.ascontiguousarray(systems, dtype=np.int32)
        Bc = cols_f32(B)
        check(self._L.saige_engine_add_products(self.ctx, len(sl), ptr(sl), ptr(sy), ptr(Bc)))

    def engine_take_batch(self, slots):
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        X = np.zeros((self.getNnomissingOut(), len(sl)), np.float32, order="F")
        check(self._L.saige_engine_take_batch(self.ctx, len(sl), ptr(sl), ptr(X)))
        return X

    ENGINE_WIDTH = 64      # the block engine's slots (saige_engine_reserve always makes 64)

    def engine_step(self, n_iter, width=64):
        """One or more engine iterations; returns the done flags and iteration
        counts of the first `width` slots.  The library writes all 64."""
        done = np.zeros(self.ENGINE_WIDTH, np.uint8)
        iters = np.zeros(self.ENGINE_WIDTH, np.int32)
        check(self._L.saige_engine_step(self.ctx, int(n_iter), ptr(done), ptr(iters)))
        return done[:width], iters[:width]

    # ------------------------------------------ device-resident columns
    # Device buffers are column-major with leading dimension N (Nnomissing);
    # (buf, col) addresses column col of buffer buf.
    def _col_ptr(self, buf, col=0):
        if buf is None:
            return None
        return ctypes.c_void_p(buf.value + 4 * int(col) * self.getNnomissingOut())

    def engine_add_batch_device(self, slots, systems, buf, col0, maxiter, tol):
        """Start solves whose right-hand sides are columns col0.. of buf."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        sy = np.ascontiguousarray(systems, dtype=np.int32)
        check(self._L.saige_engine_add_batch_device(self.ctx, len(sl), ptr(sl), ptr(sy), self._col_ptr(buf, col0),
                                                    int(maxiter), float(tol)))

    def engine_add_products_device(self, slots, systems, buf, col0):
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        sy = np.ascontiguousarray(systems, dtype=np.int32)
        check(self._L.saige_engine_add_products_device(self.ctx, len(sl), ptr(sl), ptr(sy),
                                                       self._col_ptr(buf, col0)))

    def engine_take_batch_ex(self, slots, download=False, X_dev=None, xcols=None, P1=None, P2=None, pcols=None,
                             Q=None, m=0):
        """Take finished slots; optionally download them (N x k), store them in
        columns xcols of device buffer X_dev, and return FP64 dots with
        P1[:, pcols[j]], P2[:, pcols[j]] and the m columns of Q (k x nd)."""
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        k = len(sl)
        N = self.getNnomissingOut()
        X = np.zeros((N, k), np.float32, order="F") if download else None
        xc = None if xcols is None else np.ascontiguousarray(xcols, dtype=np.int32)
        pc = None if pcols is None else np.ascontiguousarray(pcols, dtype=np.int32)
        nd = (P1 is not None) + (P2 is not None) + int(m)
        dots = np.zeros((k, max(nd, 1)), np.float64)
        check(self._L.saige_engine_take_batch_ex(self.ctx, k, ptr(sl), ptr(X), X_dev, N, ptr(xc), P1, P2, N,
                                                 ptr(pc), Q, int(m), N, ptr(dots) if nd else None))
        return X, dots[:, :nd]

    def dots_device(self, A, k, Q, m):
        """FP64 dots of the first k columns of device buffer A with the m
        columns of Q: k x m."""
        out = np.zeros((k, m), np.float64)
        N = self.getNnomissingOut()
        check(self._L.saige_dots_device(self.ctx, A, N, int(k), Q, N, int(m), ptr(out)))
        return out

    def dots_device_noalloc(self, A, k, Q, m):
        """dots_device through saige_dots_device_noalloc (fixed device
        scratch, no allocation; at most 262,144 samples)."""
        out = np.zeros((k, m), np.float64)
        N = self.getNnomissingOut()
        check(self._L.saige_dots_device_noalloc(self.ctx, A, N, int(k), Q, N, int(m), ptr(out)))
        return out

    def probes_device(self, buf, col0, seed, i0, k):
        """Rademacher probes i0 .. i0+k-1 into columns col0.. of buf."""
        check(self._L.saige_probes_device(self.ctx, self._col_ptr(buf, col0), self.getNnomissingOut(), int(seed),
                                          int(i0), int(k)))

    def probes_device_rows(self, buf, col0, seed, i0, k, rows_buf, n):
        """Probes i0 .. i0+k-1 of a fit on n samples into columns col0.. of
        buf, entry i at row rows_buf[i] (a device array of n int64), zero
        elsewhere."""
        check(self._L.saige_probes_device_rows(self.ctx, self._col_ptr(buf, col0), self.getNnomissingOut(),
                                               int(seed), int(i0), int(k), rows_buf, int(n)))

    def copy_host_to_device(self, buf, arr, col0=0):
        """Copy a host FP32 array (N or N x k, column-major) to columns col0.."""
        a = np.asarray(arr, np.float32)
        a = np.asfortranarray(a.reshape(a.shape[0], -1))
        check(self._L.saige_copy_host_to_device(self.ctx, self._col_ptr(buf, col0), ptr(a), a.nbytes))

    def engine_take(self, slot):
        x = np.zeros(self.getNnomissingOut(), np.float32)
        check(self._L.saige_engine_take(self.ctx, int(slot), ptr(x)))
        return x

    LAUNCH_KINDS = ("gemv", "gemm_half", "gemm_full", "gemm_quarter")

    def fit_null_batch(self, jobs, opt=None, max_concurrent=16):
        """Null fits through the library's C++ drivers (saige_fit_null_batch):
        jobs are dicts with trait ("binary" / "quantitative"), y (n), X
        (n x p, covariate-transformed), and optionally subset (ascending
        indices into the N samples) and offset.  opt: driver.Options-like
        (maxiter, tol, nrun, tolPCG, maxiterPCG, traceCVcutoff) or None for
        the defaults.  Returns one dict per job (its "status" is 0 on
        success)."""
        o = _fit_options(self._L, opt)
        n_fits = len(jobs)
        specs = (_lib.FitSpec * n_fits)()
        results = (_lib.FitResult * n_fits)()
        keep, outs = [], []
        for i, j in enumerate(jobs):
            keep.append(_fit_spec(j, specs[i]))
            outs.append(_fit_result(j, results[i]))
        check(self._L.saige_fit_null_batch(self.ctx, n_fits, specs, ctypes.byref(o), int(max_concurrent), results))
        return [_result_dict(results[i], o_, jobs[i]["trait"]) for i, o_ in enumerate(outs)]

    def server(self, opt=None, max_concurrent=16, hbm_reserve_bytes=1 << 30):
        """A request server on this context (saige_server_start); see Server."""
        return Server(self, opt, max_concurrent, hbm_reserve_bytes)

    def engine_stats(self, kind=None):
        """Device seconds and launch count of the GRM product launches: for one
        kind ("gemv", "gemm_half", "gemm_full", "gemm_quarter") or summed over
        all."""
        kinds = self.LAUNCH_KINDS if kind is None else (kind,)
        tot_s, tot_n = 0.0, 0
        for k in kinds:
            s = ctypes.c_double(0.0)
            n = ctypes.c_int64(0)
            check(self._L.saige_engine_stats(self.ctx, self.LAUNCH_KINDS.index(k), ctypes.byref(s),
                                             ctypes.byref(n)))
            tot_s += float(s.value)
            tot_n += int(n.value)
        return tot_s, tot_n

    def device_memory(self):
        f = ctypes.c_int64(0)
        t = ctypes.c_int64(0)
        check(self._L.saige_device_memory(self.ctx, ctypes.byref(f), ctypes.byref(t)))
        return int(f.value), int(t.value)

    def engine_system_bytes(self):
        return int(self._L.saige_engine_system_bytes(self.ctx))

    # ------------------------------------------------------ device helpers
    def alloc_device(self, nbytes):
        p = ctypes.c_void_p()
        check(self._L.saige_alloc_device(self.ctx, int(nbytes), ctypes.byref(p)))
        return p

    def free_device(self, p):
        check(self._L.saige_free_device(self.ctx, p))

    def synth_bed_device(self, dev, N, M, maf_lo, maf_hi, miss_rate, pair_frac, seed):
        check(self._L.saige_synth_bed_device(self.ctx, dev, int(N), int(M), float(maf_lo), float(maf_hi),
                                             float(miss_rate), float(pair_frac), int(seed)))

    def copy_bytes_to_device(self, dst, arr):
        """Copy a host array's bytes, as they are, to device memory dst."""
        a = np.ascontiguousarray(arr)
        check(self._L.saige_copy_host_to_device(self.ctx, dst, ptr(a), a.nbytes))

    def copy_device(self, dst, src, nbytes):
        check(self._L.saige_copy_device(self.ctx, dst, src, int(nbytes)))

    def copy_device_to_host(self, dev, nbytes, offset=0):
        out = np.zeros(int(nbytes), np.uint8)
        src = ctypes.c_void_p(dev.value + int(offset))
        check(self._L.saige_copy_device_to_host(self.ctx, src, int(nbytes), ptr(out)))
        return out


def _fit_options(L, opt):
    o = _lib.FitOptions()
    L.saige_fit_options_default(ctypes.byref(o))
    if opt is not None:
        o.maxiter, o.tol, o.nrun = int(opt.maxiter), float(opt.tol), int(opt.nrun)
        o.tol_pcg, o.maxiter_pcg, o.trace_cv_cutoff = float(opt.tolPCG), int(opt.maxiterPCG), float(opt.traceCVcutoff)
    return o


def _fit_spec(job, spec):
    """Fill a saige_fit_spec from a job dict; returns the arrays it points at."""
    pd = ctypes.POINTER(ctypes.c_double)
    y = np.ascontiguousarray(job["y"], np.float64)
    X = np.asfortranarray(job["X"], np.float64)
    n, p = X.shape
    sub = None if job.get("subset") is None else np.ascontiguousarray(job["subset"], np.int64)
    off = None if job.get("offset") is None else np.ascontiguousarray(job["offset"], np.float64)
    spec.trait = 0 if job["trait"] == "binary" else 1
    spec.n, spec.p = n, p
    spec.subset = None if sub is None else sub.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    spec.y, spec.X = y.ctypes.data_as(pd), X.ctypes.data_as(pd)
    spec.offset = None if off is None else off.ctypes.data_as(pd)
    return (y, X, sub, off)


def _fit_result(job, res):
    """Output arrays for one fit, wired into a saige_fit_result."""
    pd = ctypes.POINTER(ctypes.c_double)
    n, p = np.shape(job["X"])
    o = {"coefficients": np.zeros(p), "linear_predictors": np.zeros(n), "fitted_values": np.zeros(n),
         "Y": np.zeros(n), "cov": np.zeros((p, p), order="F")}
    res.coefficients = o["coefficients"].ctypes.data_as(pd)
    res.linear_predictors = o["linear_predictors"].ctypes.data_as(pd)
    res.fitted_values = o["fitted_values"].ctypes.data_as(pd)
    res.working_y = o["Y"].ctypes.data_as(pd)
    res.cov = o["cov"].ctypes.data_as(pd)
    return o


def _result_dict(r, o, trait):
    o.update({"theta": np.array([r.theta[0], r.theta[1]]), "converged": bool(r.converged),
              "iterations": int(r.iterations), "status": int(r.status), "traitType": trait})
    return o


class Server:
    """Null-fit requests served on a session's resident GRM (saige_server_*,
    include/saige_hip.h): submit() queues a job dict (as fit_null_batch
    takes) and returns its ticket at once; poll() returns the result dict
    once the request has finished (with "t_admit" and "t_done", seconds from
    submission), or None.  The server's scheduler admits requests while an
    engine system and HBM are free and batches their PCG iterations with
    those already under way."""

    def __init__(self, sess, opt=None, max_concurrent=16, hbm_reserve_bytes=1 << 30):
        self._L = sess._L
        self.sess = sess
        o = _fit_options(self._L, opt)
        h = ctypes.c_void_p()
        check(self._L.saige_server_start(sess.ctx, ctypes.byref(o), int(max_concurrent), int(hbm_reserve_bytes),
                                         ctypes.byref(h)))
        self.h = h
        self._jobs = {}

    def submit(self, job):
        spec = _lib.FitSpec()
        keep = _fit_spec(job, spec)
        t = ctypes.c_int64(0)
        check(self._L.saige_submit(self.h, ctypes.byref(spec), ctypes.byref(t)))
        del keep                                  # the server holds its own copy
        self._jobs[t.value] = job
        return t.value

    def poll(self, ticket, wait_ms=0):
        job = self._jobs[ticket]
        res = _lib.FitResult()
        o = _fit_result(job, res)
        times = np.zeros(2)
        rc = check(self._L.saige_poll(self.h, int(ticket), int(wait_ms), ctypes.byref(res),
                                      times.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), allow_pending=True)
        if rc == _lib.SAIGE_PENDING:
            return None
        del self._jobs[ticket]
        out = _result_dict(res, o, job["trait"])
        out["t_admit"], out["t_done"] = float(times[0]), float(times[1])
        return out

    def close(self):
        if self.h is not None:
            h, self.h = self.h, None
            check(self._L.saige_server_stop(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def glm_start_native(trait, y, X, offset=None):
    """glm() start values (FG.R:1119, 1370) by the library's C++ IRLS:
    returns {"coef", "eta", "mu", "offset"} as saige_gpu.driver.glm_start.
    The library call holds no Python lock, so several fits' start values
    can run at once on threads."""
    L = _lib.load()
    y = np.ascontiguousarray(y, np.float64)
    X = np.asfortranarray(X, np.float64)
    n, p = X.shape
    off = None if offset is None else np.ascontiguousarray(offset, np.float64)
    coef, eta, mu = np.zeros(p), np.zeros(n), np.zeros(n)
    pd = ctypes.POINTER(ctypes.c_double)
    t = {"binary": 0, "quantitative": 1}[trait]
    check(L.saige_glm_start(t, n, y.ctypes.data_as(pd), X.ctypes.data_as(pd), p,
                            None if off is None else off.ctypes.data_as(pd), coef.ctypes.data_as(pd),
                            eta.ctypes.data_as(pd), mu.ctypes.data_as(pd)))
    return {"coef": coef, "eta": eta, "mu": mu, "offset": np.zeros(n) if off is None else off}


def glm_working_native(trait, y, eta, offset=None):
    """Get_Coef's working response (FG.R:9-12, 24-27) by the library's C++:
    returns (mu, Y, W) as saige_gpu.driver.get_coef_co computes them."""
    L = _lib.load()
    y = np.ascontiguousarray(y, np.float64)
    eta = np.ascontiguousarray(eta, np.float64)
    off = None if offset is None else np.ascontiguousarray(offset, np.float64)
    n = len(y)
    mu, Y, W = np.empty(n), np.empty(n), np.empty(n)
    pd = ctypes.POINTER(ctypes.c_double)
    t = {"binary": 0, "quantitative": 1}[trait]
    check(L.saige_glm_working(t, n, y.ctypes.data_as(pd), eta.ctypes.data_as(pd),
                              None if off is None else off.ctypes.data_as(pd), mu.ctypes.data_as(pd),
                              Y.ctypes.data_as(pd), W.ctypes.data_as(pd)))
    return mu, Y, W


def coefficients_native(Y, X, w, tau0, sol):
    """getCoefficients' algebra after its solves (FG.cpp:3167-3199) by the
    library's C++: sol = Sigma^-1 [Y, X] (n x (1+p)); X must be column-major
    FP64 (it is not copied).  Returns (cov, alpha, eta)."""
    L = _lib.load()
    Y = np.ascontiguousarray(Y, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    sol = np.asfortranarray(sol, np.float64)
    assert X.dtype == np.float64 and X.flags["F_CONTIGUOUS"]
    n, p = X.shape
    assert sol.shape == (n, p + 1) and len(Y) == n and len(w) == n
    cov, alpha, eta = np.zeros((p, p), order="F"), np.zeros(p), np.empty(n)
    pd = ctypes.POINTER(ctypes.c_double)
    check(L.saige_coefficients(n, p, Y.ctypes.data_as(pd), X.ctypes.data_as(pd), w.ctypes.data_as(pd),
                               float(tau0), sol.ctypes.data_as(pd), cov.ctypes.data_as(pd), alpha.ctypes.data_as(pd),
                               eta.ctypes.data_as(pd)))
    return cov, alpha, eta


def covariate_transform_native(X):
    """Covariate_Transform (FG.R:1612-1647) by the library's C++ code:
    returns (X_new, R, kept_columns) as saige_gpu.plink.covariate_transform."""
    L = _lib.load()
    X = np.asfortranarray(X, np.float64)
    n, p = X.shape
    Xo, R = np.zeros(n * p), np.zeros(p * p)
    kept, pk = np.zeros(p, np.int32), ctypes.c_int(0)
    pd = ctypes.POINTER(ctypes.c_double)
    check(L.saige_covariate_transform(X.ctypes.data_as(pd), n, p, Xo.ctypes.data_as(pd), R.ctypes.data_as(pd),
                                      kept.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(pk)))
    k = pk.value
    return Xo[:n * k].reshape((n, k), order="F"), R[:k * k].reshape((k, k), order="F"), kept[:k].copy()
