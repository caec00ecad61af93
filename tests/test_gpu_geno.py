"""GPU parity of genotype ingest: decode, counts, missingness, QC,
imputation, repacking, standardisation and the diagonal -- bit-exact against
the oracle (FG.cpp:326-704, 739-1024) and the reference's fixtures."""
import os

import numpy as np
import pytest

from fixtures import load_acount, load_bed, random_bed
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STAT_KEYS = ["allele_count", "num_missing", "mac", "fill", "freq", "missing_rate"]


def _compare_stats(sess, og):
    st = sess.marker_stats()
    for k in STAT_KEYS:
        a, b = st[k], og["stats"][k]
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), k
    qc = og["stats"]["pass_qc"].astype(bool)
    assert np.array_equal(st["flags"] & 1, og["stats"]["pass_qc"])
    assert np.array_equal((st["flags"] >> 1) & 1, og["stats"]["pass_vr"])
    assert np.array_equal(st["inv_std"][qc].view(np.uint32), og["stats"]["inv_std"][qc].view(np.uint32))
    assert sess.getMsub_MAFge_minMAFtoConstructGRM() == og["Mqc"]
    assert np.array_equal(sess.getAlleleFreqVec().view(np.uint32), og["freq"].view(np.uint32))
    assert np.array_equal(sess.getMACVec(), og["mac"])
    assert np.array_equal(sess.getQCdMarkerIndex(), qc)


def test_fixture_counts_and_stats(session):
    body, ids = load_bed("random1000", 1000)
    print("case: random1000 fixture N=1000 M=1000", flush=True)
    session.setminMAFforGRM(0.0)
    session.setmaxMissingRateforGRM(1.0)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, 1000, 1000)
    og = O.setgeno(body, 1000, 1000)
    _compare_stats(session, og)
    acount = load_acount()
    st = session.marker_stats()
    assert all(st["allele_count"][j] == acount[m][0] for j, m in enumerate(ids))


@pytest.mark.parametrize("N,M,miss,subset", [(1003, 517, 0.01, False), (777, 300, 0.2, True),
                                             (4099, 65, 0.05, True)])
def test_synthetic_bit_exact(session, N, M, miss, subset):
    print(f"case: synthetic N={N} M={M} miss={miss} subset={subset}", flush=True)
    body, G, mm = random_bed(N, M, miss, seed=N + M, maf_lo=0.0, maf_hi=0.5)
    rng = np.random.default_rng(N)
    if subset:
        sub = rng.permutation(np.sort(rng.choice(N, N - N // 3, replace=False)) + 1).astype(np.int32)
    else:
        sub = np.arange(1, N + 1, dtype=np.int32)
    session.setminMAFforGRM(0.02)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(20, -1, False)
    session.setgeno_bytes(body, N, M, sub)
    og = O.setgeno(body, N, M, subidx=sub, min_maf=0.02, max_missing=0.15)
    _compare_stats(session, og)
    Nnm = len(sub)
    assert session.getNnomissingOut() == Nnm
    for j in list(range(0, og["Mqc"], max(1, og["Mqc"] // 7))) + [og["Mqc"] - 1]:
        assert np.array_equal(session.Get_OneSNP_Geno(j), O.get_geno(og["packed"][j], Nnm))
        z = session.Get_OneSNP_StdGeno(j)
        zo = O.std_geno(og["packed"][j], Nnm, og["freq"][j], og["inv_std"][j])
        assert np.array_equal(z.view(np.uint32), zo.view(np.uint32))
    d = session.Get_Diagof_StdGeno()
    assert np.array_equal(d.view(np.uint32), O.diag_std(og).view(np.uint32))
    dk = session.get_DiagofKin()
    assert np.array_equal(dk, O.diag_std(og) / np.float32(og["Mqc"]))


def test_variance_ratio_markers(session):
    N, M = 600, 200
    print("case: VR marker selection", flush=True)
    body, G, mm = random_bed(N, M, 0.0, seed=3, maf_lo=0.0, maf_hi=0.5)
    rand = np.arange(0, M, 4, dtype=np.int32)
    flags = np.zeros(M, np.uint8)
    flags[rand] = 1
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setminMAC_VarianceRatio(10, 30, True)
    session.set_vr_random_markers(rand)
    session.setgeno_bytes(body, N, M)
    og = O.setgeno(body, N, M, min_maf=0.01, max_missing=0.15, vr_mode=1, min_mac_vr=10, max_mac_vr=30,
                   randset=flags)
    _compare_stats(session, og)
    assert np.array_equal(session.getIndexVec_forVarRatio(), og["index_vr"])
    assert np.array_equal(session.getMACVec_forVarRatio(), og["mac_vr"])
    for k in range(0, len(og["index_vr"]), 5):
        assert np.array_equal(session.Get_OneSNP_Geno_forVarRatio(k), O.get_geno(og["packed_vr"][k], N))
    session.setminMAC_VarianceRatio(20, -1, False)
    session.set_vr_random_markers(np.zeros(0, np.int32))


def test_setgeno_from_files(session, tmp_path):
    from saige_gpu import plink
    N, M = 301, 77
    print("case: setgeno from .bed/.bim/.fam", flush=True)
    rng = np.random.default_rng(5)
    G = rng.binomial(2, rng.uniform(0.05, 0.5, M)[None, :], (N, M))
    miss = rng.random((N, M)) < 0.02
    prefix = str(tmp_path / "t")
    plink.write_bed(prefix, G, miss)
    sub = np.arange(2, N + 1, dtype=np.int32)
    session.setminMAFforGRM(0.01)
    session.setmaxMissingRateforGRM(0.15)
    session.setgeno(prefix + ".bed", prefix + ".bim", prefix + ".fam", sub)
    og = O.setgeno(plink.pack_bed(G, miss), N, M, subidx=sub, min_maf=0.01, max_missing=0.15)
    _compare_stats(session, og)
    with pytest.raises(Exception):
        session.setgeno(prefix + ".bed", prefix + ".bim", os.path.join(str(tmp_path), "absent.fam"), sub)


def test_errors_are_reported(session):
    from saige_gpu import Session, SaigeError
    fresh = Session(0)
    with pytest.raises(SaigeError) as e:
        fresh.getCrossprodMatAndKin(np.zeros(4, np.float32))
    assert "setgeno" in str(e.value)
    body, _, _ = random_bed(16, 4, 0.0, seed=1)
    with pytest.raises(SaigeError):
        fresh.setgeno_bytes(body, 16, 4, np.array([0, 1, 2], np.int32))   # indices are 1-based
    fresh.close()
