"""GPU parity of the tensor operator surface (aq_contract / aq_sum / aq_scale /
aq_fill / aq_dot / aq_weight / aq_norm / key I/O / aq_slice) against the
oracle's CTF semantics on the same seeded inputs.

Tolerances: FP64 contractions 1e-12 relative to the largest reference entry;
permuting sums, weight, fill and I/O bit-exact; FP32 contractions 2e-6 and
BF16x3 contractions 5e-5 relative to the largest reference entry."""

import numpy as np
import pytest

import aquarius_amd as aq
from oracle import ctf_ops as ops

from test_oracle_ops import RHF_PATTERNS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = aq.Context()
    yield c
    c.close()


def dims_for(o, v):
    return {c: (o if c in "ijklmn" else v) for c in "abcdefghijklmnopqrstuvwxyzQ"}


def rnd(rng, shape):
    return rng.standard_normal(shape)


def check_contract(ctx, ia, ib, ic, dim, alpha=-0.75, beta=1.0, seed=0, tol=1e-12, prec="fp64"):
    rng = np.random.default_rng(seed)
    A = rnd(rng, [dim[c] for c in ia])
    B = rnd(rng, [dim[c] for c in ib])
    Cm = rnd(rng, [dim[c] for c in ic])
    ref = Cm.copy()
    ops.contract(alpha, A, ia, B, ib, beta, ref, ic)
    tA, tB, tC = ctx.from_numpy(A), ctx.from_numpy(B), ctx.from_numpy(Cm)
    ctx.set_precision(prec, 0.0)
    try:
        ctx.contract(alpha, tA, ia, tB, ib, beta, tC, ic)
        got = tC.numpy()
    finally:
        ctx.set_precision("fp64")
        for t in (tA, tB, tC):
            t.free()
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert err <= tol, f"{ia},{ib}->{ic} [{prec}] rel err {err:.3e}"
    return err


@pytest.mark.parametrize("ia,ib,ic", RHF_PATTERNS)
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_contract_rhf_patterns_fp64(ctx, ia, ib, ic, beta):
    check_contract(ctx, ia, ib, ic, dims_for(3, 5), beta=beta, seed=len(ia + ib + ic))
    check_contract(ctx, ia, ib, ic, dims_for(5, 19), beta=beta, seed=7)


@pytest.mark.parametrize("ia,ib,ic,o,v", [
    ("abef", "efij", "abij", 6, 40),      # ladder shape: split-K path
    ("amie", "ebmj", "abij", 8, 30),      # ring: permuted operands and output
    ("efmn", "efin", "mi", 6, 30),        # tiny output, huge K: split-K
    ("efmn", "fn", "me", 8, 40),          # N = 1
    ("me", "aeim", "ai", 8, 40),          # M and N small
    ("efam", "efij", "amij", 6, 36),      # o^3 v^3 W term
])
def test_contract_gemm_paths(ctx, ia, ib, ic, o, v):
    check_contract(ctx, ia, ib, ic, dims_for(o, v), seed=3)


def test_contract_large_tile_path(ctx):
    # 2100 x 2100 output -> the 128x128 tile configuration, K = 72
    dim = {"a": 2100, "b": 2100, "k": 72}
    check_contract(ctx, "ak", "kb", "ab", dim, beta=0.5, seed=1)
    check_contract(ctx, "ka", "bk", "ab", dim, beta=0.0, seed=2)   # both operands transposed
    check_contract(ctx, "ak", "kb", "ba", dim, beta=1.0, seed=3)   # swapped output


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
@pytest.mark.parametrize("ia,ib", [("mk", "kn"), ("km", "kn"), ("mk", "nk"), ("km", "nk")])
def test_contract_one_short_side_tiles(ctx, prec, tol, ia, ib):
    """128 x 64 blocks (N = 48 against 33000 rows) and 64 x 128 blocks (M = 40
    against 33001 columns): every tile runs past N (or M), so every tile takes
    the complete-k-tile path with clamped operand rows and skipped MFMA blocks;
    all four operand layouts, a partial last k-tile (K = 72, 37)."""
    print(f"case: one short side {prec} {ia},{ib}", flush=True)
    check_contract(ctx, ia, ib, "mn", {"m": 33000, "n": 48, "k": 72}, seed=31, tol=tol, prec=prec)
    check_contract(ctx, ia, ib, "mn", {"m": 40, "n": 33001, "k": 37}, seed=32, tol=tol, prec=prec)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 5e-5)])
@pytest.mark.parametrize("ia,ib,ic,dim", [
    # ring: GEMM rows (a,i), columns (b,j) written into C[abij] through the output map
    ("amie", "ebmj", "abij", {"a": 30, "b": 30, "e": 30, "m": 8, "i": 8, "j": 8}),
    # a batch label z beside two row labels
    ("aekz", "ebz", "abkz", {"a": 40, "b": 24, "e": 32, "k": 6, "z": 3}),
    # rows (a,m) around the column b in C; 7 columns and K = 3000: split-K partials reduced through the map
    ("akm", "kb", "abm", {"a": 20, "k": 3000, "m": 5, "b": 7}),
    # three row digits
    ("aeij", "eb", "aibj", {"a": 36, "e": 40, "i": 5, "j": 6, "b": 9}),
])
def test_contract_mapped_output(ctx, prec, tol, ia, ib, ic, dim):
    """Contractions whose C is a permutation of the GEMM output [rows cols
    batch]: the GEMM epilogue (or the split-K reduction) writes C in place
    through per-label strides, with no temporary and no permuting pass."""
    print(f"case: mapped output {prec} {ia},{ib}->{ic}", flush=True)
    ctx.enable_timing(True)
    ctx.timing_reset()
    try:
        check_contract(ctx, ia, ib, ic, dim, beta=0.5, seed=41, tol=tol, prec=prec)
        tim = ctx.timing()
    finally:
        ctx.enable_timing(False)
    assert tim.get("strided_sum", {"launches": 0})["launches"] == 0, tim


def test_contract_batch_and_special_labels(ctx):
    dim = {"a": 7, "b": 5, "k": 9, "z": 3, "x": 4}
    check_contract(ctx, "akz", "kbz", "abz", dim, seed=4)          # batch label z
    check_contract(ctx, "azk", "zkb", "bza", dim, seed=5)          # batch, permuted
    check_contract(ctx, "akx", "kb", "ab", dim, seed=6)            # x summed on A only
    check_contract(ctx, "ak", "kb", "abx", dim, seed=7)            # x broadcast into C
    check_contract(ctx, "aka", "kb", "ab", {"a": 6, "k": 5, "b": 4}, seed=8)   # diagonal of A
    check_contract(ctx, "ai", "bj", "abij", {"a": 9, "b": 9, "i": 4, "j": 4}, seed=9)   # outer product


def test_contract_into_aliased_output(ctx):
    rng = np.random.default_rng(11)
    W = rnd(rng, (4, 4, 4, 4))
    T = rnd(rng, (4, 4))
    ref = W.copy()
    ops.contract(1.0, ref.copy(), "mnij", T, "jk", 1.0, ref, "mnik")
    tW, tT = ctx.from_numpy(W), ctx.from_numpy(T)
    ctx.contract(1.0, tW, "mnij", tT, "jk", 1.0, tW, "mnik")
    np.testing.assert_allclose(tW.numpy(), ref, rtol=1e-13, atol=1e-13)
    tW.free()
    tT.free()


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 5e-5)])
@pytest.mark.parametrize("ia,ib,ic,o,v", [
    ("abef", "efij", "abij", 6, 40),
    ("amie", "ebmj", "abij", 8, 30),
    ("efam", "efij", "amij", 5, 33),
    ("efmn", "efin", "mi", 6, 30),
    ("efmn", "fn", "me", 7, 41),
])
def test_contract_reduced_precision(ctx, prec, tol, ia, ib, ic, o, v):
    check_contract(ctx, ia, ib, ic, dims_for(o, v), beta=0.0, seed=21, tol=tol, prec=prec)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("ia,ib,ic,dim", [
    ("ak", "kb", "ab", {"a": 300, "b": 270, "k": 200}),
    ("amie", "ebmj", "abij", {"a": 40, "b": 40, "e": 40, "m": 8, "i": 8, "j": 8}),
    ("ak", "kb", "ab", {"a": 520, "b": 390, "k": 1100}),
])
def test_bf16x3_tile_variants(ctx, variant, ia, ib, ic, dim):
    """Every BF16x3 kernel (128x128, 256x128, 128x256, 256x256 blocks; 128x128
    loading two k-tiles ahead) on ragged edges, odd and even k-tile counts,
    with and without split-K."""
    print(f"case: bf16x3 variant {variant} {ia},{ib}->{ic} {dim}", flush=True)
    ctx.set_gemm_variant("bf16x3", variant)
    try:
        check_contract(ctx, ia, ib, ic, dim, beta=0.5, seed=9, tol=5e-5, prec="bf16x3")
    finally:
        ctx.set_gemm_variant("bf16x3", -1)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_gemm_variants_and_autotuner_are_bitwise_identical(ctx, prec):
    """Kernel variants (and hence the autotuner's choice) never change a
    result: the split-K partition and the per-element summation order do not
    depend on the variant.  Shapes: a 128x128-tiled GEMM and a split-K one."""
    print(f"case: variants/autotune {prec}", flush=True)
    rng = np.random.default_rng(3)
    for dim in ({"a": 2048, "b": 2048, "k": 96}, {"a": 300, "b": 260, "k": 4096}):
        A = ctx.from_numpy(rnd(rng, (dim["a"], dim["k"])))
        B = ctx.from_numpy(rnd(rng, (dim["k"], dim["b"])))
        outs = []
        ctx.set_precision(prec, 0.0)
        try:
            for v in ((0, 1, 2, 3, 4, -1) if prec == "bf16x3" else (0, 1, 2, -1)):
                ctx.set_gemm_variant(prec, v)
                ctx.set_autotune(v == -1)
                Cc = ctx.tensor((dim["a"], dim["b"]), zero=True)
                ctx.contract(1.0, A, "ak", B, "kb", 0.0, Cc, "ab")
                outs.append(Cc.numpy())
                Cc.free()
        finally:
            ctx.set_precision("fp64")
            ctx.set_gemm_variant(prec, -1)
            ctx.set_autotune(False)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
        A.free()
        B.free()
    table = aq.autotune_table()
    code = aq.PRECISIONS[prec]
    mine = [e for e in table if e["precision"] == code and e["m_bucket"] == 22 and e["n_bucket"] == 22]
    assert mine and mine[0]["variant"] in (0, 1, 2, 3, 4)
    assert mine[0]["ms"][mine[0]["variant"]] == min(m for m in mine[0]["ms"] if m >= 0)


def test_bf16x3_is_more_accurate_than_bf16_alone(ctx):
    """Error of the split product stays near 2^-17 per term, far below the
    2^-9 of single BF16 rounding."""
    err = check_contract(ctx, "ak", "kb", "ab", {"a": 300, "b": 260, "k": 512}, beta=0.0, seed=5, tol=5e-5,
                         prec="bf16x3")
    assert err < 2e-5


@pytest.mark.parametrize("ia,ib", [("abij", "abij"), ("abji", "abij"), ("baji", "abij"), ("aijb", "aijb"),
                                   ("aibj", "aijb"), ("em", "me"), ("nmji", "mnij"), ("ejmn", "mnej")])
def test_sum_permutations_bit_exact(ctx, ia, ib):
    rng = np.random.default_rng(5)
    dim = dims_for(7, 37)
    A = rnd(rng, [dim[c] for c in ia])
    B = rnd(rng, [dim[c] for c in ib])
    for alpha, beta in ((1.0, 0.0), (-1.0, 1.0), (2.0, 0.5)):
        ref = B.copy()
        ops.tsum(alpha, A, ia, beta, ref, ib)
        tA, tB = ctx.from_numpy(A), ctx.from_numpy(B)
        ctx.sum(alpha, tA, ia, beta, tB, ib)
        assert np.array_equal(tB.numpy(), ref), (ia, ib, alpha, beta)
        tA.free()
        tB.free()


def test_sum_aliased_snapshot(ctx):
    rng = np.random.default_rng(6)
    Z = rnd(rng, (33, 33, 9, 9))
    ref = Z.copy()
    ops.tsum(1.0, ref, "baji", 1.0, ref, "abij")
    tZ = ctx.from_numpy(Z)
    ctx.sum(1.0, tZ, "baji", 1.0, tZ, "abij")
    got = tZ.numpy()
    assert np.array_equal(got, ref)
    assert np.array_equal(got, got.transpose(1, 0, 3, 2))
    tZ.free()


def test_sum_reduce_broadcast_and_diagonal(ctx):
    rng = np.random.default_rng(7)
    A = rnd(rng, (6, 5))
    B = rnd(rng, (6, 4))
    ref = B.copy()
    ops.tsum(1.5, A, "ak", 1.0, ref, "ab")
    tA, tB = ctx.from_numpy(A), ctx.from_numpy(B)
    ctx.sum(1.5, tA, "ak", 1.0, tB, "ab")
    np.testing.assert_allclose(tB.numpy(), ref, rtol=1e-14)
    D = rnd(rng, (5, 5))
    refd = D.copy()
    ops.scale(3.0, refd, "ii")
    tD = ctx.from_numpy(D)
    ctx.scale(tD, 3.0, "ii")
    assert np.array_equal(tD.numpy(), refd)
    for t in (tA, tB, tD):
        t.free()


def test_weight_bit_exact_including_threshold(ctx):
    rng = np.random.default_rng(8)
    o, v = 5, 11
    X = rnd(rng, (v, v, o, o))
    dA = -rng.uniform(0.2, 5.0, v)
    dI = rng.uniform(-1.5, -0.3, o)
    dA[3] = -(dI[1] + dI[1] + dA[2] - 5e-5) - dA[2] + 0.0   # forces one |den| < 1e-4
    ref = X.copy()
    ops.weight(ref, [dA, dA, dI, dI])
    tX = ctx.from_numpy(X)
    ctx.weight(tX, [dA, dA, dI, dI])
    got = tX.numpy()
    assert np.array_equal(got, ref)
    Y = rnd(rng, (v, o))
    refy = Y.copy()
    ops.weight(refy, [dA, dI], shift=0.25)
    tY = ctx.from_numpy(Y)
    ctx.weight(tY, [dA, dI], shift=0.25)
    assert np.array_equal(tY.numpy(), refy)
    tX.free()
    tY.free()


def test_dot_norm_fill_io_slice(ctx):
    rng = np.random.default_rng(9)
    A = rnd(rng, (13, 7, 5))
    B = rnd(rng, (7, 13, 5))
    tA, tB = ctx.from_numpy(A), ctx.from_numpy(B)
    assert ctx.dot(tA, "abc", tA, "abc") == pytest.approx(ops.dot(A, "abc", A, "abc"), rel=1e-13)
    assert ctx.dot(tA, "abc", tB, "bac") == pytest.approx(ops.dot(A, "abc", B, "bac"), rel=1e-12)
    assert ctx.dot(tA, "abc", tB, "bad") == pytest.approx(ops.dot(A, "abc", B, "bad"), rel=1e-12)
    for p in (0, 1, 2):
        assert ctx.norm(tA, p) == pytest.approx(ops.norm(A, p), rel=1e-13)
    # whole-tensor I/O is bit exact; keys are column-major
    assert np.array_equal(tA.numpy(), A)
    keys = np.array([0, 1, 13, 13 * 7, 13 * 7 * 5 - 1])
    np.testing.assert_array_equal(tA.read(keys), ops.read_keys(A, keys))
    ref = A.copy()
    ops.write_keys(ref, keys, [1.0, 2.0, 3.0, 4.0, 5.0], alpha=2.0, beta=0.5)
    tA.write(keys, [1.0, 2.0, 3.0, 4.0, 5.0], alpha=2.0, beta=0.5)
    assert np.array_equal(tA.numpy(), ref)
    ctx.fill(tA, 0.25, 2.0)
    ops.fill(ref, 0.25, 2.0)
    assert np.array_equal(tA.numpy(), ref)
    # slice: B[1:4, 2:7, 0:3] = 0.5*B + 2*A[2:5, 0:5, 1:4]   (A viewed [a,b,c])
    Cm = rnd(rng, (6, 9, 4))
    tC = ctx.from_numpy(Cm)
    ctx.slice(2.0, tA, [2, 0, 1], 0.5, tC, [1, 2, 0], [3, 5, 3])
    refc = Cm.copy()
    refc[1:4, 2:7, 0:3] = 0.5 * refc[1:4, 2:7, 0:3] + 2.0 * ref[2:5, 0:5, 1:4]
    np.testing.assert_allclose(tC.numpy(), refc, rtol=1e-15)
    for t in (tA, tB, tC):
        t.free()


def test_div_invert_bit_exact(ctx):
    rng = np.random.default_rng(12)
    A = rnd(rng, (7, 5, 3))
    B = rnd(rng, (7, 5, 3))
    B[0, 0, 0] = 0.0
    B[1, 2, 0] = 1e-310
    Cm = rnd(rng, (7, 5, 3))
    ref = Cm.copy()
    ops.div(1.5, A, B, 0.25, ref)
    tA, tB, tC = ctx.from_numpy(A), ctx.from_numpy(B), ctx.from_numpy(Cm)
    ctx.div(1.5, tA, tB, 0.25, tC)
    assert np.array_equal(tC.numpy(), ref)
    ops.invert(-2.0, B, 1.0, ref)
    ctx.invert(-2.0, tB, 1.0, tC)
    assert np.array_equal(tC.numpy(), ref)
    for t in (tA, tB, tC):
        t.free()


def test_operand_cache_tracks_relabelling_and_writes(ctx):
    """A constant tensor reached with different label names, and then
    overwritten, must always be read through the right permutation/version."""
    rng = np.random.default_rng(10)
    o, v = 4, 9
    V = rnd(rng, (v, v, o, o))
    T = rnd(rng, (v, o))
    tV, tT = ctx.from_numpy(V), ctx.from_numpy(T)
    for ia, ib, ic in (("efmn", "fn", "me"), ("efnm", "fn", "me"), ("efmn", "fm", "ne"), ("efmn", "fn", "me")):
        ref = np.zeros((o, v)) if ic == "me" else np.zeros((o, v))
        ops.contract(1.0, V, ia, T, ib, 0.0, ref, ic)
        out = ctx.tensor((o, v))
        ctx.contract(1.0, tV, ia, tT, ib, 0.0, out, ic)
        np.testing.assert_allclose(out.numpy(), ref, rtol=1e-12, atol=1e-13)
        out.free()
    V2 = rnd(rng, V.shape)
    tV.write_all(V2)
    ref = np.zeros((o, v))
    ops.contract(1.0, V2, "efnm", T, "fn", 0.0, ref, "me")
    out = ctx.tensor((o, v))
    ctx.contract(1.0, tV, "efnm", tT, "fn", 0.0, out, "me")
    np.testing.assert_allclose(out.numpy(), ref, rtol=1e-12, atol=1e-13)
    for t in (tV, tT, out):
        t.free()


def test_invalid_arguments_raise(ctx):
    t = ctx.tensor((3, 4))
    u = ctx.tensor((4, 5))
    with pytest.raises(aq.AqError) as e:
        ctx.contract(1.0, t, "abc", u, "bc", 0.0, t, "ac")
    assert e.value.code == 1 and "invalid ndim" in str(e.value)
    with pytest.raises(aq.AqError):
        ctx.contract(1.0, t, "ab", u, "ab", 0.0, t, "ab")   # length mismatch on a/b
    with pytest.raises(aq.AqError):
        ctx.norm(t, 3)
    t.free()
    u.free()


def test_resize_gives_fresh_storage(ctx):
    """CTFTensor::resize (ctf_tensor.cxx:189-201): new lengths, the old
    contents gone (cleared when asked), no cached copy of the old contents
    reused by a later contraction."""
    rng = np.random.default_rng(31)
    A = ctx.from_numpy(rng.standard_normal((12, 7)))
    B = ctx.from_numpy(rng.standard_normal((12, 5)))
    C1 = ctx.tensor((7, 5))
    ctx.contract(1.0, A, "km", B, "kn", 0.0, C1, "mn")     # caches a copy of A
    A.resize((12, 9), zero=True)
    assert A.numpy().shape == (12, 9) and not A.numpy().any()
    a2 = rng.standard_normal((12, 9))
    A.write_all(a2)
    C2 = ctx.tensor((9, 5))
    ctx.contract(1.0, A, "km", B, "kn", 0.0, C2, "mn")
    ref = a2.T @ B.numpy()
    np.testing.assert_allclose(C2.numpy(), ref, rtol=1e-12, atol=1e-12)
    A.resize((3,), zero=True)
    assert A.numpy().tolist() == [0.0, 0.0, 0.0]
    for t in (A, B, C1, C2):
        t.free()
