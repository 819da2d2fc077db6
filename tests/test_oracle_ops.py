"""Pins oracle.ctf_ops (the CTF operator semantics) with brute-force loops.

The brute force below enumerates every assignment of every label and applies
the CTF rule directly (contracted / summed / broadcast / diagonal), sharing no
code with the einsum-based oracle."""

import itertools

import numpy as np
import pytest

from oracle import ctf_ops as ops


def brute_mult(alpha, A, ia, B, ib, beta, C, ic):
    labels = sorted(set(ia) | set(ib) | set(ic))
    dims = {}
    for T, idx in ((A, ia), (B, ib), (C, ic)):
        for k, c in enumerate(idx):
            dims[c] = T.shape[k]
    out = C.copy()
    touched = set()
    acc = {}
    for vals in itertools.product(*[range(dims[c]) for c in labels]):
        env = dict(zip(labels, vals))
        ci = tuple(env[c] for c in ic)
        a = A[tuple(env[c] for c in ia)]
        b = B[tuple(env[c] for c in ib)]
        acc[ci] = acc.get(ci, 0.0) + a * b
        touched.add(ci)
    # a label only in C is part of the output key, so each broadcast value
    # receives the same sum over the non-C labels
    for ci in touched:
        out[ci] = beta * C[ci] + alpha * acc[ci]
    return out


def brute_sum(alpha, A, ia, beta, B, ib):
    labels = sorted(set(ia) | set(ib))
    dims = {}
    for T, idx in ((A, ia), (B, ib)):
        for k, c in enumerate(idx):
            dims[c] = T.shape[k]
    acc = {}
    for vals in itertools.product(*[range(dims[c]) for c in labels]):
        env = dict(zip(labels, vals))
        bi = tuple(env[c] for c in ib)
        acc[bi] = acc.get(bi, 0.0) + A[tuple(env[c] for c in ia)]
    out = B.copy()
    for bi, s in acc.items():
        out[bi] = beta * B[bi] + alpha * s
    return out


def rnd(rng, *shape):
    return rng.standard_normal(shape)


# every contraction pattern issued by rhfccsd.cxx:231-309 (with tiny dims)
RHF_PATTERNS = [
    ("ai", "bj", "abij"), ("efmn", "fn", "me"), ("efmn", "efin", "mi"), ("me", "ei", "mi"),
    ("finm", "fn", "mi"), ("fimn", "fn", "mi"), ("efmn", "afmn", "ae"), ("me", "am", "ae"),
    ("efam", "fm", "ae"), ("feam", "fm", "ae"), ("efmn", "efij", "mnij"), ("ejmn", "ei", "mnij"),
    ("efmn", "fj", "mnej"), ("efam", "efij", "amij"), ("amej", "ei", "amij"), ("amie", "ej", "amij"),
    ("efmn", "afin", "amie"), ("feam", "fi", "amie"), ("efam", "fi", "amie"), ("mnei", "an", "amie"),
    ("nmei", "an", "amie"), ("efnm", "afni", "amei"), ("efam", "fi", "amei"), ("nmei", "an", "amei"),
    ("ae", "ei", "ai"), ("mi", "am", "ai"), ("amie", "em", "ai"), ("me", "aeim", "ai"),
    ("efam", "efim", "ai"), ("mnei", "eamn", "ai"), ("abej", "ei", "abij"), ("amij", "bm", "abij"),
    ("ae", "ebij", "abij"), ("mi", "abmj", "abij"), ("abef", "efij", "abij"), ("mnij", "abmn", "abij"),
    ("amie", "ebmj", "abij"), ("amei", "ebjm", "abij"), ("amej", "ebim", "abij"),
]


@pytest.mark.parametrize("ia,ib,ic", RHF_PATTERNS)
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_contract_rhf_patterns(ia, ib, ic, beta):
    rng = np.random.default_rng(hash((ia, ib, ic)) & 0xFFFF)
    dim = {c: (2 if c in "ijmn" else 3) for c in "abefijmn"}
    A = rnd(rng, *[dim[c] for c in ia])
    B = rnd(rng, *[dim[c] for c in ib])
    C = rnd(rng, *[dim[c] for c in ic])
    ref = brute_mult(-0.5, A, ia, B, ib, beta, C, ic)
    ops.contract(-0.5, A, ia, B, ib, beta, C, ic)
    np.testing.assert_allclose(C, ref, rtol=1e-13, atol=1e-13)


def test_contract_summed_and_diagonal_labels():
    rng = np.random.default_rng(3)
    A = rnd(rng, 3, 4, 3)      # "aka": diagonal over a, k summed
    B = rnd(rng, 3, 2)         # "ab"
    C = rnd(rng, 2)
    ref = brute_mult(1.5, A, "aka", B, "ab", 1.0, C, "b")
    ops.contract(1.5, A, "aka", B, "ab", 1.0, C, "b")
    np.testing.assert_allclose(C, ref, rtol=1e-13)


def test_contract_broadcast_label():
    rng = np.random.default_rng(4)
    A = rnd(rng, 3, 2)
    B = rnd(rng, 2)
    C = rnd(rng, 3, 4)
    ops.contract(2.0, A, "ak", B, "k", 0.0, C, "ab")
    expect = 2.0 * (A @ B)
    for b in range(4):
        np.testing.assert_allclose(C[:, b], expect, rtol=1e-13)


@pytest.mark.parametrize("ia,ib", [("abij", "abij"), ("abji", "abij"), ("baji", "abij"),
                                   ("aijb", "aijb"), ("aibj", "aijb"), ("em", "me"), ("nmji", "mnij")])
def test_sum_permutations(ia, ib):
    rng = np.random.default_rng(5)
    dim = {c: (2 if c in "ijmn" else 3) for c in "abefijmn"}
    A = rnd(rng, *[dim[c] for c in ia])
    B = rnd(rng, *[dim[c] for c in ib])
    ref = brute_sum(-1.0, A, ia, 2.0, B, ib)
    ops.tsum(-1.0, A, ia, 2.0, B, ib)
    np.testing.assert_allclose(B, ref, rtol=1e-14)


def test_sum_aliased_snapshot():
    """Z2["abij"] += Z2["baji"] (rhfccsd.cxx:309) reads the pre-update Z2."""
    rng = np.random.default_rng(6)
    Z = rnd(rng, 3, 3, 2, 2)
    snap = Z.copy()
    ops.tsum(1.0, Z, "baji", 1.0, Z, "abij")
    np.testing.assert_allclose(Z, snap + snap.transpose(1, 0, 3, 2), rtol=1e-15)
    # result symmetric under (ai)<->(bj), exactly
    assert np.array_equal(Z, Z.transpose(1, 0, 3, 2))
    W = rnd(rng, 2, 2, 2, 2)
    snapw = W.copy()
    ops.tsum(1.0, W, "nmji", 1.0, W, "mnij")
    np.testing.assert_allclose(W, snapw + snapw.transpose(1, 0, 3, 2), rtol=1e-15)


def test_sum_reduce_and_broadcast():
    rng = np.random.default_rng(7)
    A = rnd(rng, 3, 4)
    B = rnd(rng, 3, 5)
    ops.tsum(1.0, A, "ak", 0.0, B, "ab")
    for b in range(5):
        np.testing.assert_allclose(B[:, b], A.sum(axis=1), rtol=1e-14)


def test_weight_closed_form_and_threshold():
    rng = np.random.default_rng(8)
    X = rnd(rng, 3, 2)
    dA = np.array([-1.0, -2.0, -3.0])
    dI = np.array([0.5, 1.0 + 5e-5])   # 1+5e-5-1 -> |den| < 1e-4 -> zero
    X0 = X.copy()
    ops.weight(X, [dA, dI])
    for a in range(3):
        for i in range(2):
            den = dA[a] + dI[i]
            expect = 0.0 if abs(den) < 1e-4 else X0[a, i] / den
            assert X[a, i] == expect
    assert X[0, 1] == 0.0
    # shift argument
    Y = np.ones((2,))
    ops.weight(Y, [np.array([1.0, 2.0])], shift=1.0)
    np.testing.assert_array_equal(Y, [0.5, 1.0 / 3.0])


def test_norms_and_dot():
    X = np.array([[1.0, -4.0], [2.0, 0.5]])
    assert ops.norm(X, 0) == 4.0
    assert ops.norm(X, 1) == 7.5
    assert ops.norm(X, 2) == pytest.approx(np.sqrt(1 + 16 + 4 + 0.25), rel=1e-15)
    Y = np.array([[2.0, 1.0], [3.0, 1.0]])
    assert ops.dot(X, "ab", Y, "ab") == 2.0 - 4.0 + 6.0 + 0.5
    assert ops.dot(X, "ab", Y, "ba") == 2.0 - 12.0 + 2.0 + 0.5


def test_div_and_invert_skip_tiny_divisors():
    A = np.array([1.0, 2.0, 3.0, 4.0])
    B = np.array([2.0, 0.0, 1e-310, -4.0])
    C = np.array([10.0, 20.0, 30.0, 40.0])
    ops.div(3.0, A, B, 0.5, C)
    np.testing.assert_array_equal(C, [5.0 + 1.5, 20.0, 30.0, 20.0 - 3.0])
    D = np.array([1.0, 1.0, 1.0, 1.0])
    ops.invert(2.0, B, 1.0, D)
    np.testing.assert_array_equal(D, [2.0, 1.0, 1.0, 0.5])


def test_keys_are_column_major():
    X = np.arange(24, dtype=float).reshape(2, 3, 4)
    # key = i0 + 2*i1 + 6*i2
    assert ops.read_keys(X, [0 + 2 * 1 + 6 * 3])[0] == X[0, 1, 3]
    flat = ops.to_flat(X)
    assert flat[1] == X[1, 0, 0] and flat[2] == X[0, 1, 0]
    np.testing.assert_array_equal(ops.from_flat(flat, X.shape), X)
    ops.write_keys(X, [1], [100.0], alpha=2.0, beta=1.0)
    assert X[1, 0, 0] == 12.0 + 200.0
