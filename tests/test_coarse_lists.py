"""The coarse schedule's task lists, evaluated on the host in FP64 (no GPU):
merging a launch into the run before it re-associates a segment over a
skeleton block x_P = sum_s A_s B_s into segments (A_c A_s) B_s over composite
blocks (A_s itself a composite when P joined the run earlier).  The lists of every schedule, run launch by launch in numpy, must
give the oracle's evaluation (symGen/HGEMM_gen.h:117-238 and the HSS form
symGen/HGEMM_v1.h:117-405) in both directions; the coarse lists must do it
in fewer launches while still writing every skeleton block."""
import numpy as np
import pytest

from conftest import make_problem, rel
from matrox_ru_amd import hgemm, host
from oracle import oracle

CASES = {
    "h2b_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0),
    "hss_small": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.0, h=1.0),
    "h2b_ragged": dict(n=3001, d=5, leaf=100, rank=70, tol=1e-7, budget=0.3, h=1.2),
}


def _op(buf, off, trans, lda, m, k):
    """op(A) (m x k) of a stored block: column-major with leading dimension
    lda, or its transpose (A stored k x m)."""
    i = np.arange(m)[:, None]
    j = np.arange(k)[None, :]
    return buf[off + i * lda + j] if trans else buf[off + j * lda + i]


def run_lists(hm, L, TW, nrhs):
    """Evaluate the described lists in FP64: launches in order, each task
    C = beta C + sum over segments of op(A) B."""
    mats = [np.asarray(hm.D, np.float64), np.asarray(hm.B, np.float64), np.asarray(hm.VT, np.float64)]
    sk = int(max(hm.skelrow[v] + hm.slen[v] for v in range(len(hm.slen))))
    vec = [TW.copy(), np.zeros(sk * nrhs), np.zeros(sk * nrhs), np.zeros(hm.N * nrhs)]
    comp = {}

    def composite(i):
        # op(A1) op(A2); A2 a stored block or (buffer 3) another composite
        if i not in comp:
            c = [int(x) for x in L["comps"][i]]
            a1 = _op(mats[c[2]], c[0], c[3], c[4], c[8], c[9])
            a2 = composite(c[1]) if c[5] == 3 else _op(mats[c[5]], c[1], c[6], c[7], c[9], c[10])
            assert a2.shape == (c[9], c[10])
            comp[i] = a1 @ a2
        return comp[i]

    def a_of(s, m):
        a_off, _, a_buf, a_trans, lda, k = (int(x) for x in s[:6])
        if a_buf == 3:
            return composite(a_off)
        return _op(mats[a_buf], a_off, a_trans, lda, m, k)

    for li in range(len(L["launches"]) - 1):
        for t in L["ltasks"][L["launches"][li]:L["launches"][li + 1]]:
            c_row, m, c_buf, sb, se, beta = (int(x) for x in L["tasks"][t])
            acc = np.zeros((m, nrhs))
            for s in L["segs"][sb:se]:
                k, b_buf, b_row, b_ld = int(s[5]), int(s[6]), int(s[1]), int(s[7])
                B = vec[b_buf][b_row * nrhs:(b_row + b_ld) * nrhs].reshape(nrhs, b_ld).T[:k]
                acc += a_of(s, m) @ B
            C = vec[c_buf][c_row * nrhs:(c_row + m) * nrhs].reshape(nrhs, m).T
            C[...] = acc + (C if beta else 0.0)
    return vec


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("direction", ["forward", "adjoint"])
def test_coarse_lists_match_oracle(case, direction):
    X, hm = make_problem(**CASES[case])
    nrhs = 3
    TW = host.normal((hm.N * nrhs,), seed=41)
    ref = oracle.hgemm(hm if direction == "forward" else hm.transposed(), TW, nrhs)[0]
    out = {}
    for sc in ("fused", "split", "coarse", "coarse_split"):
        L = hgemm.lists_describe(hm, direction, sc)
        vec = run_lists(hm, L, TW, nrhs)
        assert rel(vec[3], ref) <= 1e-12, sc
        out[sc] = (L, vec)
    assert len(out["coarse_split"][0]["launches"]) < len(out["split"][0]["launches"])
    Lf, vf = out["fused"]
    Lc, vc = out["coarse"]
    assert len(Lc["launches"]) < len(Lf["launches"]) and len(Lc["comps"]) > 0
    assert len(Lf["comps"]) == 0
    # every skeleton block is still written (a task whose segments all
    # vanished, over blocks that are zero, writes zeros), with the same values
    for buf in (1, 2):
        wf = {(int(Lf["tasks"][t][0]), int(Lf["tasks"][t][1])) for t in Lf["ltasks"] if Lf["tasks"][t][2] == buf}
        wc = {(int(Lc["tasks"][t][0]), int(Lc["tasks"][t][1])) for t in Lc["ltasks"] if Lc["tasks"][t][2] == buf}
        assert wf == wc
        assert rel(vc[buf], vf[buf]) <= 1e-12


def test_coarse_lists_keep_merges_cheap():
    """A merge adds at most 4.5 MFLOP per right-hand side column: the coarse
    lists' work per column exceeds the fused lists' by no more than that per
    merged launch; runs of more than two launches (composites of
    composites) occur."""
    X, hm = make_problem(**CASES["h2b_small"])

    def work(L):
        f = 0.0
        for t in L["tasks"]:
            for s in L["segs"][t[3]:t[4]]:
                f += 2.0 * t[1] * s[5]
        return f

    Lf = hgemm.lists_describe(hm, "forward", "fused")
    Lc = hgemm.lists_describe(hm, "forward", "coarse")
    merged = len(Lf["launches"]) - len(Lc["launches"])
    assert merged >= 1
    # (it can also be less: a composite is narrower than the block it
    # replaces when the producer's segments are shorter than its rows)
    assert work(Lc) <= work(Lf) + 4.5e6 * merged
    assert any(int(c[5]) == 3 for c in Lc["comps"])
