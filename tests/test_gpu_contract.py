"""The FP64 MFMA contraction kernels against numpy einsum (GPU)."""

import zlib

import numpy as np
import pytest

from cc4s_amd import api
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

# (A indices, B indices, C indices) covering every operand orientation the
# CCSD terms use: K- or M/N-contiguous operands, transposed results, empty K
# (outer product), empty N (vector result), large K with split-K.
CASES = [
    ("ac", "cbij", "abij"),
    ("ki", "abkj", "abij"),
    ("Gai", "Gbj", "abij"),
    ("akic", "cbkj", "abij"),
    ("bkci", "ackj", "abij"),
    ("klcd", "adkl", "ac"),
    ("kldc", "cdil", "ki"),
    ("Gkd", "dk", "G"),
    ("Gac", "G", "ac"),
    ("Gki", "ak", "Gai"),
    ("dclk", "adil", "akic"),
    ("ai", "bj", "abij"),
    ("cxdy", "cdij", "xyij"),
    ("Gxc", "Gyd", "cxdy"),
    ("klij", "abkl", "abij"),
]
LENS = {"a": 13, "b": 13, "c": 13, "d": 13, "i": 5, "j": 5, "k": 5, "l": 5,
        "G": 37, "x": 7, "y": 6}


# longer contracted lengths: every case below has K > 256 (the complex kernel
# for K <= 256 is a 64x64-tile variant; longer K takes the 128x64 tile)
LENS_LONG = {"a": 21, "b": 19, "c": 23, "d": 17, "i": 6, "j": 5, "k": 12, "l": 16,
             "G": 300, "x": 9, "y": 6}
CASES_LONG = [
    ("Gai", "Gbj", "abij"),
    ("akic", "cbkj", "abij"),
    ("klcd", "adkl", "ac"),
    ("kldc", "cdil", "ki"),
    ("cxdy", "cdij", "xyij"),
    ("Gxc", "Gyd", "cxdy"),
    ("dclk", "adil", "akic"),
]


def _rand(idx, cplx, rng, lens=None):
    shape = tuple((lens or LENS)[c] for c in idx)
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("case", CASES, ids=["%s,%s->%s" % c for c in CASES])
def test_contract_matches_einsum(case, cplx):
    ia, ib, ic = case
    # seeded from a stable digest of the case (str hashing is salted per process)
    rng = np.random.default_rng(zlib.crc32(",".join(case).encode()))
    A, B, C = (_rand(i, cplx, rng) for i in (ia, ib, ic))
    for alpha, beta in ((1.0, 0.0), (-0.5, 1.0)):
        ref = alpha * np.einsum(f"{ia},{ib}->{ic}", A, B) + beta * C
        out = api.contract(alpha, A, ia, B, ib, beta, C, ic)
        assert rel_err(out, ref) < 1e-13, (case, alpha, beta)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("case", CASES_LONG, ids=["%s,%s->%s" % c for c in CASES_LONG])
def test_contract_long_k_matches_einsum(case, cplx):
    ia, ib, ic = case
    rng = np.random.default_rng(zlib.crc32(("long," + ",".join(case)).encode()))
    A, B, C = (_rand(i, cplx, rng, LENS_LONG) for i in (ia, ib, ic))
    K = int(np.prod([LENS_LONG[c] for c in ia if c not in ic]))
    assert K > 256
    for alpha, beta in ((1.0, 0.0), (-0.5, 1.0)):
        ref = alpha * np.einsum(f"{ia},{ib}->{ic}", A, B) + beta * C
        out = api.contract(alpha, A, ia, B, ib, beta, C, ic)
        assert rel_err(out, ref) < 1e-13, (case, alpha, beta)


@pytest.mark.parametrize("conj_a,conj_b", [(True, False), (False, True), (True, True)])
def test_contract_conjugation(conj_a, conj_b):
    rng = np.random.default_rng(3)
    A = _rand("Gai", True, rng)
    B = _rand("Gbj", True, rng)
    C = np.zeros((13, 13, 5, 5), complex)
    a = np.conj(A) if conj_a else A
    b = np.conj(B) if conj_b else B
    ref = np.einsum("Gai,Gbj->abij", a, b)
    out = api.contract(1.0, A, "Gai", B, "Gbj", 0.0, C, "abij",
                       conj_a=conj_a, conj_b=conj_b)
    assert rel_err(out, ref) < 1e-13


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (129, 65, 17), (300, 200, 1000),
                                   (64, 64, 20000)])
def test_contract_ragged_and_split_k(M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    C = np.zeros((M, N), complex)
    out = api.contract(1.0, A, "mk", B, "kn", 0.0, C, "mn")
    assert rel_err(out, A @ B) < 1e-13
    out = api.contract(1.0, A.real.copy(), "mk", B.real.copy(), "kn", 0.0,
                       C.real.copy(), "mn")
    assert rel_err(out, A.real @ B.real) < 1e-13


def test_contract_rejects_bad_indices():
    A = np.ones((3, 4))
    with pytest.raises(api.CcsdError, match="only in one operand"):
        api.contract(1.0, A, "ab", A, "cd", 0.0, np.ones((3, 3)), "ac")
