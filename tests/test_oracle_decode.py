"""Pins the oracle's genotype decode / QC / imputation (restating
FG.cpp:326-579, 739-1024) against the reference's own fixtures and a
brute-force decoder written from the PLINK-1 specification."""
import numpy as np
import pytest

from fixtures import brute_decode, load_acount, load_bed, load_frq, load_vmiss, random_bed
from oracle import oracle as O


@pytest.mark.parametrize("name,M", [("random1000", 1000), ("forcate_vr", 2000)])
def test_allele_counts_match_acount(name, M):
    body, ids = load_bed(name, 1000)
    assert len(ids) == M
    g = O.setgeno(body, 1000, M)
    acount = load_acount()
    vmiss = load_vmiss()
    st = g["stats"]
    checked = 0
    for j, mid in enumerate(ids):
        alt, obs = acount[mid]
        assert obs == 2000
        # ALT is the .bim A1 allele, so ALT_CTS is the reference's alleleCount
        assert st["allele_count"][j] == alt, mid
        assert st["num_missing"][j] == vmiss[mid] == 0
        checked += 1
    assert checked == M


def test_maf_matches_frq():
    body, ids = load_bed("grm1000x10k", 1000)
    g = O.setgeno(body, 1000, len(ids))
    frq = load_frq()
    maf = np.array([frq[i] for i in ids])
    freq = g["stats"]["freq"].astype(np.float64)
    # .frq MAF is the A1 frequency printed with 4 significant decimals
    assert np.max(np.abs(freq - maf)) <= 5e-5 + 1e-7
    # QC at minMAF = 0.01 from exact integer counts
    g2 = O.setgeno(body, 1000, len(ids), min_maf=0.01)
    ac = g["stats"]["allele_count"]
    mac_expect = np.minimum(ac, 2000 - ac)
    assert np.array_equal(g2["stats"]["mac"], mac_expect)
    # with N = 1000 and no missing calls, FP32 maf >= 0.01f  <=>  MAC >= 20
    n_fail = int(np.sum(mac_expect < 20))
    assert g2["Mqc"] == len(ids) - n_fail
    assert n_fail == 350


def _expected_stats(G, miss, subidx, indicator, min_maf, max_miss):
    """FG.cpp:424-495 evaluated with numpy float32 scalars."""
    f32 = np.float32
    Nnm = len(subidx)
    out = []
    for j in range(G.shape[1]):
        sel = indicator.astype(bool)
        m = int(np.sum(miss[sel, j]))
        ac = int(np.sum(np.where(miss[sel, j], 0, G[sel, j])))
        with np.errstate(invalid="ignore", divide="ignore"):
            alt0 = f32(ac) / f32(2 * (Nnm - m))
        twice = f32(2) * alt0
        fill = 0 if np.isnan(twice) else int(np.floor(float(twice) + 0.5))
        if m > 0:
            ac = ac + fill * m
        freq = f32(ac) / f32(2 * Nnm)
        mr = f32(m) / f32(Nnm)
        one_m = f32(1) - freq
        maf = one_m if one_m < freq else freq
        mac = min(ac, 2 * Nnm - ac)
        qc = bool(maf >= f32(min_maf) and mr <= f32(max_miss))
        sd = np.sqrt(f32(2) * freq * one_m)
        inv = f32(0) if sd == 0 else f32(1) / sd
        out.append((ac, m, mac, fill, freq, mr, inv, qc))
    return out


@pytest.mark.parametrize("N,M,miss_rate,subset", [(37, 23, 0.1, False), (64, 40, 0.3, True),
                                                  (101, 17, 0.02, True)])
def test_decode_brute_force(N, M, miss_rate, subset):
    body, G, miss = random_bed(N, M, miss_rate, seed=N * 7 + M)
    # independent spec decoder agrees with the generator
    Gb, mb = brute_decode(body, N, M)
    assert np.array_equal(mb, miss)
    assert np.array_equal(np.where(miss, 0, Gb), np.where(miss, 0, G))
    rng = np.random.default_rng(N)
    if subset:
        subidx = np.sort(rng.choice(N, N - N // 4, replace=False)).astype(np.int32) + 1
        subidx = rng.permutation(subidx).astype(np.int32)   # any order of the .fam rows
    else:
        subidx = np.arange(1, N + 1, dtype=np.int32)
    ind = np.zeros(N, np.uint8)
    ind[subidx - 1] = 1
    g = O.setgeno(body, N, M, subidx=subidx, indicator=ind, min_maf=0.05, max_missing=0.2)
    exp = _expected_stats(Gb, mb, subidx, ind, 0.05, 0.2)
    st = g["stats"]
    for j, (ac, m, mac, fill, freq, mr, inv, qc) in enumerate(exp):
        assert st["allele_count"][j] == ac
        assert st["num_missing"][j] == m
        assert st["mac"][j] == mac
        assert st["fill"][j] == fill
        assert st["freq"][j].tobytes() == np.float32(freq).tobytes()
        assert st["missing_rate"][j].tobytes() == np.float32(mr).tobytes()
        assert bool(st["pass_qc"][j]) == qc
        if qc:
            assert st["inv_std"][j].tobytes() == np.float32(inv).tobytes()
    # repacked genotypes: subset order, missing calls imputed with fill
    qc_idx = np.nonzero(st["pass_qc"])[0]
    assert g["Mqc"] == len(qc_idx)
    for k, j in enumerate(qc_idx):
        expect = np.where(mb[subidx - 1, j], exp[j][3], Gb[subidx - 1, j])
        assert np.array_equal(O.get_geno(g["packed"][k], len(subidx)), expect)


def test_all_missing_marker_fails_qc():
    N, M = 12, 3
    G = np.ones((N, M), np.int64)
    miss = np.zeros((N, M), bool)
    miss[:, 1] = True
    from saige_gpu.plink import pack_bed
    g = O.setgeno(pack_bed(G, miss), N, M, max_missing=0.5)
    assert g["stats"]["num_missing"][1] == N
    assert not g["stats"]["pass_qc"][1]
    assert g["stats"]["pass_qc"][0] and g["stats"]["pass_qc"][2]


def test_variance_ratio_selection():
    """VR markers (FG.cpp:496-547): categorical MAC window, the random pool
    above it, and their exclusion from the GRM set."""
    N, M = 200, 60
    body, G, miss = random_bed(N, M, 0.0, seed=5, maf_lo=0.0, maf_hi=0.5)
    base = O.setgeno(body, N, M)
    mac = base["stats"]["mac"]
    rand = np.zeros(M, np.uint8)
    rand[::3] = 1
    g = O.setgeno(body, N, M, vr_mode=1, min_mac_vr=5, max_mac_vr=20, randset=rand)
    vr = g["stats"]["pass_vr"].astype(bool)
    expect = ((mac >= 5) & (mac < 20)) | ((mac >= 20) & rand.astype(bool))
    assert np.array_equal(vr, expect)
    assert not np.any(vr & g["stats"]["pass_qc"].astype(bool))
    g2 = O.setgeno(body, N, M, vr_mode=2, min_mac_vr=20, randset=rand)
    assert np.array_equal(g2["stats"]["pass_vr"].astype(bool), (mac >= 20) & rand.astype(bool))
    # the VR genotypes are repacked like the QC'd ones
    for k, j in enumerate(np.nonzero(vr)[0]):
        assert np.array_equal(O.get_geno(g["packed_vr"][k], N), G[:, j])
