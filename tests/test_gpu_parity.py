"""GPU parity: every product the library computes, through the C ABI, against
the CPU oracle on the same seeded inputs.  Integer work, so the bar is
bit-exact equality of every output word.

Raw products are compared modulo the output modulus 2^W (W = EQ for
MatrixVectorMul and polymul, EP for InnerProd): that is the part of the
reference's uint16 results that any later step reads.  Rounded outputs are
compared exactly; the closed forms used for the expected rounding are the ones
tests/test_oracle.py pins against the reference's own forms for all 2^16
inputs.
"""
import numpy as np
import pytest

import oracle as O
import tmvp_amd as T

pytestmark = pytest.mark.gpu

N = 256
PS_ALL = ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"]
# the MFMA form with each int8 tile (16x16x64 by default, 32x32x32 via TMVP_TILE_32) and the VALU form
BACKENDS = [T.MFMA, T.MFMA32, T.VALU]


@pytest.fixture(scope="module")
def ctx():
    c = T.Context(0)
    yield c
    c.close()


def secret(rng, shape, ps):
    p = O.params(ps)
    v = rng.integers(-(p["mu"] // 2), p["mu"] // 2 + 1, shape)
    if p["sable"]:  # Sable stores secrets mod q (SB/src/fips202.cu:646-649)
        return (v % (1 << p["eq"])).astype(np.uint16)
    return v.astype(np.int16).view(np.uint16)


def public(rng, shape, bits=13):
    return rng.integers(0, 1 << bits, shape, dtype=np.uint16)


def q2p(ps, x):
    p = O.params(ps)
    x = x.astype(np.int64)
    return (((x + p["h1"]) % (1 << p["eq"])) >> (p["eq"] - p["ep"])).astype(np.uint16)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("batch", [1, 5, 64, 131])
def test_polymul_vs_oracle(ctx, backend, batch):
    rng = np.random.default_rng(100 + batch)
    a = public(rng, (batch, N))
    s = secret(rng, (batch, N), "saber")
    got = T.run_polymul(ctx, a, s, 13, backend)
    want = O.polymul_batch(a, s, nthreads=8) & 8191
    assert np.array_equal(got, want)


@pytest.mark.parametrize("backend", BACKENDS)
def test_polymul_unreduced_public_and_closed_forms(ctx, backend):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 65536, (16, N), dtype=np.uint16)      # high bits must be ignored
    s = np.zeros((16, N), np.uint16)
    for i in range(16):
        s[i, (17 * i) % N] = 1 if i % 2 == 0 else 0xFFFF       # +x^k and -x^k
    got = T.run_polymul(ctx, a, s, 13, backend)
    want = O.polymul_batch(a, s) & 8191
    assert np.array_equal(got, want)
    # closed form: a * x^k is a rotated by k with the k wrapped coefficients
    # negated (x^256 = -1); -x^k negates the whole rotation
    for i in range(16):
        k = (17 * i) % N
        rot = np.concatenate([-a[i, N - k:].astype(np.int64), a[i, :N - k].astype(np.int64)])
        if i % 2:
            rot = -rot
        assert np.array_equal(got[i], (rot % 8192).astype(np.uint16)), f"instance {i}, k = {k}"


@pytest.mark.parametrize("eq_bits", [8, 11, 14])
def test_polymul_moduli_agree_across_backends(ctx, eq_bits):
    rng = np.random.default_rng(eq_bits)
    a = rng.integers(0, 65536, (33, N), dtype=np.uint16)
    s = rng.integers(-60, 61, (33, N)).astype(np.int16).view(np.uint16)
    want = O.polymul_batch(a, s) & ((1 << eq_bits) - 1)
    for backend in BACKENDS:
        assert np.array_equal(T.run_polymul(ctx, a, s, eq_bits, backend), want)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ps", PS_ALL)
@pytest.mark.parametrize("transpose", [False, True])
def test_matvec_vs_oracle(ctx, backend, ps, transpose):
    p = O.params(ps)
    l = p["l"]
    rng = np.random.default_rng(2 * PS_ALL.index(ps) + int(transpose))
    batch = 37
    A = public(rng, (batch, l * l * N), p["eq"])
    s = secret(rng, (batch, l * N), ps)
    raw = O.matvec_batch(l, A, s, transpose=transpose, method=1, nthreads=8).reshape(batch, l, N)
    got = T.run_matvec(ctx, A, s, ps, backend, transpose=transpose)
    assert np.array_equal(got, raw & ((1 << p["eq"]) - 1))
    got = T.run_matvec(ctx, A, s, ps, backend, transpose=transpose, round_to_p=True)
    assert np.array_equal(got, q2p(ps, raw))


@pytest.mark.parametrize("backend", [T.MFMA, T.MFMA32])
@pytest.mark.parametrize("ps", PS_ALL)
def test_matvec_random_large_batches_vs_oracle(ctx, backend, ps):
    """Seeded random batch sizes past one round of the persistent grid (each
    wave walks several instances, the last round partly filled), both
    transposes, raw and rounded, against the oracle."""
    p = O.params(ps)
    l = p["l"]
    rng = np.random.default_rng(300 + PS_ALL.index(ps))
    for batch in [int(x) for x in rng.integers(4097, 20000, 2)]:
        A = public(rng, (batch, l * l * N), p["eq"])
        s = secret(rng, (batch, l * N), ps)
        for transpose in (False, True):
            raw = O.matvec_batch(l, A, s, transpose=transpose, method=1, nthreads=8).reshape(batch, l, N)
            got = T.run_matvec(ctx, A, s, ps, backend, transpose=transpose)
            assert np.array_equal(got, raw & ((1 << p["eq"]) - 1)), (batch, transpose)
            got = T.run_matvec(ctx, A, s, ps, backend, transpose=transpose, round_to_p=True)
            assert np.array_equal(got, q2p(ps, raw)), (batch, transpose)


@pytest.mark.parametrize("backend", BACKENDS)
def test_matvec_accumulate(ctx, backend):
    rng = np.random.default_rng(11)
    batch = 9
    A = public(rng, (batch, 9 * N))
    s = secret(rng, (batch, 3 * N), "saber")
    init = rng.integers(0, 65536, (batch, 3, N), dtype=np.uint16)
    raw = O.matvec_batch(3, A, s, out=init)
    got = T.run_matvec(ctx, A, s, "saber", backend, out_init=init)
    assert np.array_equal(got, raw & 8191)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ps", PS_ALL)
def test_innerprod_vs_oracle(ctx, backend, ps):
    p = O.params(ps)
    l, ep, et = p["l"], p["ep"], p["et"]
    rng = np.random.default_rng(200 + p["l"] + 10 * p["sable"])
    batch = 41
    b = public(rng, (batch, l * N), ep)
    s = secret(rng, (batch, l * N), ps)
    raw = O.innerprod_batch(l, b, s, nthreads=8).astype(np.int64)
    got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_RAW)
    assert np.array_equal(got, (raw % (1 << ep)).astype(np.uint16))
    # encaps epilogue (post_process2 / post_process3+4)
    mbits = rng.integers(0, 2, (batch, N), dtype=np.uint16)
    got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_ENC, aux=mbits)
    want = ((raw + p["h1"] - (mbits.astype(np.int64) << (ep - 1))) % (1 << ep)) >> (ep - et)
    assert np.array_equal(got, want.astype(np.uint16))
    # decaps epilogue (post_process3 / post_process7)
    cm = rng.integers(0, 1 << et, (batch, N), dtype=np.uint16)
    got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_DEC, aux=cm)
    want = ((raw + p["h2"] - (cm.astype(np.int64) << (ep - et))) % (1 << ep)) >> (ep - 1)
    assert np.array_equal(got, want.astype(np.uint16))


@pytest.mark.parametrize("ps", PS_ALL)
def test_innerprod_and_polymul_random_large_batches(ctx, ps):
    """InnerProd with each epilogue, and polymul, at seeded random batch sizes
    past one round of the persistent grid (MFMA, both tiles) against the
    oracle."""
    p = O.params(ps)
    l, ep, et = p["l"], p["ep"], p["et"]
    rng = np.random.default_rng(400 + PS_ALL.index(ps))
    batch = int(rng.integers(8193, 30000))
    b = public(rng, (batch, l * N), ep)
    s = secret(rng, (batch, l * N), ps)
    raw = O.innerprod_batch(l, b, s, nthreads=8).astype(np.int64)
    mbits = rng.integers(0, 2, (batch, N), dtype=np.uint16)
    cm = rng.integers(0, 1 << et, (batch, N), dtype=np.uint16)
    want_enc = ((raw + p["h1"] - (mbits.astype(np.int64) << (ep - 1))) % (1 << ep)) >> (ep - et)
    want_dec = ((raw + p["h2"] - (cm.astype(np.int64) << (ep - et))) % (1 << ep)) >> (ep - 1)
    for backend in (T.MFMA, T.MFMA32):
        got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_RAW)
        assert np.array_equal(got, (raw % (1 << ep)).astype(np.uint16)), (backend, batch)
        got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_ENC, aux=mbits)
        assert np.array_equal(got, want_enc.astype(np.uint16)), (backend, batch)
        got = T.run_innerprod(ctx, b, s, ps, backend, T.IP_DEC, aux=cm)
        assert np.array_equal(got, want_dec.astype(np.uint16)), (backend, batch)
    a = public(rng, (batch, N), p["eq"])
    s1 = s.reshape(batch, l, N)[:, 0]
    want = O.polymul_batch(a, s1, method=1, nthreads=8) & ((1 << p["eq"]) - 1)
    for backend in (T.MFMA, T.MFMA32):
        assert np.array_equal(T.run_polymul(ctx, a, s1, p["eq"], backend), want), (backend, batch)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ps", PS_ALL)
def test_enc_dec_epilogues_decoding_window_on_gpu(ctx, backend, ps):
    """The fused encryption / decryption roundings of tmvp_innerprod over every
    v in Z_p and m in {0, 1}: bit-exact against the oracle's post_process2 /
    post_process3 forms (S/src/poly.cu:391-403; SB/src/pack.cu:290-335, 436-442),
    and, from the mathematics, decoding recovers m for every noise |e| <= W =
    2^(EP-2) - 2^(EP-ET-1) and fails for some v at e = +-(W + 1).  The secret is
    the polynomial 1, so the InnerProd passes b through to the epilogue."""
    p = O.params(ps)
    l, ep, et = p["l"], p["ep"], p["et"]
    P, W = 1 << ep, (1 << (ep - 2)) - (1 << (ep - et - 1))
    vals = np.concatenate([np.arange(P), np.arange(P)])
    ms = np.concatenate([np.zeros(P, np.int64), np.ones(P, np.int64)])

    def through_epilogue(x, aux, epilogue):
        x = np.asarray(x).reshape(-1, N)
        batch = x.shape[0]
        b = np.zeros((batch, l * N), np.uint16)
        b[:, :N] = x
        s = np.zeros((batch, l * N), np.uint16)
        s[:, 0] = 1
        got = T.run_innerprod(ctx, b, s, ps, backend, epilogue, aux=np.asarray(aux, np.uint16).reshape(batch, N))
        return got.reshape(-1).astype(np.int64)

    c = through_epilogue(vals, ms, T.IP_ENC)
    assert np.array_equal(c, O.round_enc_v(ps, vals, ms).astype(np.int64))
    es = np.arange(-W - 1, W + 2)
    x = (vals[None, :] + es[:, None]) % P
    aux = np.broadcast_to(c, x.shape)
    m = through_epilogue(x, aux, T.IP_DEC).reshape(x.shape)
    dec_table = np.stack([O.round_dec_m(ps, np.arange(P), np.full(P, cc)) for cc in range(1 << et)]).astype(np.int64)
    assert np.array_equal(m, dec_table[aux, x])
    fails = np.count_nonzero(m != ms[None, :], axis=1)
    assert not fails[1:-1].any(), f"decoding fails inside the window at e = {es[1:-1][fails[1:-1] > 0]}"
    assert fails[0] > 0 and fails[-1] > 0


# ---------------------------------------------------------------- reference test vectors
@pytest.mark.parametrize("backend", BACKENDS)
def test_saber_tv_public_key_on_gpu(ctx, backend, saber_tv):
    """pk (TMVP-Saber/SABER-GPU/include/tv.h:8) = pack10(round(A^T s)), A from tv.h:4."""
    s = O.unpack_bits(saber_tv["sk_tv"][:1248], 768, 13)
    got = T.run_matvec(ctx, saber_tv["A"], s, "saber", backend, transpose=True, round_to_p=True)
    assert O.pack_bits(got.reshape(-1), 10) == saber_tv["pk"][:960]


@pytest.mark.parametrize("backend", BACKENDS)
def test_sable_tv_public_key_on_gpu(ctx, backend, sable_tv):
    """pk_tv (TMVP_SABLE/include/tv.h:15) = pack9(round(A^T s)); s from the 2-bit sk."""
    pk = sable_tv["pk_tv"]
    A = O.gen_matrix("sable2", pk[864:896])
    codes = O.unpack_bits(sable_tv["sk_tv"][:192], 768, 2)
    s = ((codes ^ 2).astype(np.int32) - 2).astype(np.uint16)
    got = T.run_matvec(ctx, A, s, "sable2", backend, transpose=True, round_to_p=True)
    assert O.pack_bits(got.reshape(-1), 9) == pk[:864]


@pytest.mark.parametrize("backend", BACKENDS)
def test_saber_tv_encaps_core_on_gpu(ctx, backend, saber_tv):
    """The reference's encaps inputs (buf_tv, pk) through GPU MatVec + InnerProd give the
    oracle's ciphertext, which decrypts to m_tv."""
    pk = saber_tv["pk"]
    m = O.sha3_256(saber_tv["buf_tv"][:32])
    kr = O.sha3_512(m + O.sha3_256(pk))
    sp = O.gen_secret("saber", kr[32:])
    assert np.array_equal(sp, saber_tv["sp"])
    bp = T.run_matvec(ctx, saber_tv["A"], sp, "saber", backend, round_to_p=True).reshape(-1)
    b = O.unpack_bits(pk, 768, 10)
    mbits = O.unpack_bits(m, 256, 1)
    vp = T.run_innerprod(ctx, b, sp, "saber", backend, T.IP_ENC, aux=mbits).reshape(-1)
    ct = O.pack_bits(bp, 10) + O.pack_bits(vp, 4)
    assert ct == O.indcpa_enc("saber", m, kr[32:], pk)
    assert O.indcpa_dec("saber", saber_tv["sk_tv"][:1248], ct) == m


# ---------------------------------------------------------------- full size + edge cases
def test_saber_matvec_full_batch_vs_oracle(ctx):
    """B = 65536 (the headline shape): MFMA == VALU bit for bit, and both == oracle."""
    rng = np.random.default_rng(65536)
    batch = 65536
    A = public(rng, (batch, 9 * N))
    s = secret(rng, (batch, 3 * N), "saber")
    mf = T.run_matvec(ctx, A, s, "saber", T.MFMA, round_to_p=True)
    va = T.run_matvec(ctx, A, s, "saber", T.VALU, round_to_p=True)
    assert np.array_equal(mf, va)
    assert np.array_equal(T.run_matvec(ctx, A, s, "saber", T.MFMA32, round_to_p=True), mf)
    raw = O.matvec_batch(3, A, s, method=1, nthreads=16)
    assert np.array_equal(mf, q2p("saber", raw))


def test_polymul_full_batch_backends_agree(ctx):
    rng = np.random.default_rng(3)
    batch = 65536
    a = public(rng, (batch, N))
    s = secret(rng, (batch, N), "saber")
    mf = T.run_polymul(ctx, a, s, 13, T.MFMA)
    va = T.run_polymul(ctx, a, s, 13, T.VALU)
    assert np.array_equal(mf, va)
    idx = rng.choice(batch, 512, replace=False)
    assert np.array_equal(mf[idx], O.polymul_batch(a[idx], s[idx], nthreads=8) & 8191)


def test_zero_batch_is_a_no_op(ctx):
    lib = T.load_library()
    assert lib.tmvp_polymul(ctx.handle, T.MFMA, None, None, None, 0, 13, 0, None) == 0
    assert lib.tmvp_matvec(ctx.handle, 1, T.VALU, 0, None, None, None, 0, 1, 0, None) == 0


def test_argument_errors(ctx):
    lib = T.load_library()
    buf = T.DeviceBuffer(4096)
    assert lib.tmvp_polymul(ctx.handle, T.MFMA, buf.ptr, buf.ptr, buf.ptr, 1, 15, 0, None) == -4
    assert lib.tmvp_polymul(ctx.handle, T.MFMA, buf.ptr, buf.ptr, buf.ptr, 1, 7, 0, None) == -4
    assert lib.tmvp_polymul(ctx.handle, 7, buf.ptr, buf.ptr, buf.ptr, 1, 13, 0, None) == -1
    assert lib.tmvp_polymul(ctx.handle, T.MFMA, buf.ptr, buf.ptr, buf.ptr, 1, 17, 0, None) == -1
    assert lib.tmvp_matvec(ctx.handle, 9, T.MFMA, 0, buf.ptr, buf.ptr, buf.ptr, 1, 0, 0, None) == -1
    assert lib.tmvp_matvec(ctx.handle, 1, T.MFMA, 0, buf.ptr, buf.ptr, buf.ptr, 1, 1, T.ACCUMULATE, None) == -1
    assert lib.tmvp_innerprod(ctx.handle, 1, T.MFMA, T.IP_ENC, buf.ptr, buf.ptr, buf.ptr, None, 1, 0, None) == -1
    both = T.TILE_16 | T.TILE_32  # two tiles at once, or an unknown flag
    assert lib.tmvp_polymul(ctx.handle, T.MFMA, buf.ptr, buf.ptr, buf.ptr, 1, 13, both, None) == -1
    assert lib.tmvp_matvec(ctx.handle, 1, T.MFMA, 0, buf.ptr, buf.ptr, buf.ptr, 1, 0, both, None) == -1
    assert lib.tmvp_innerprod(ctx.handle, 1, T.MFMA, 0, buf.ptr, buf.ptr, buf.ptr, None, 1, 0x40, None) == -1
    assert lib.tmvp_innerprod(ctx.handle, 1, T.MFMA, 5, buf.ptr, buf.ptr, buf.ptr, None, 1, 0, None) == -1
    T.device_synchronize()


def test_out_of_range_secret_is_reported(ctx):
    rng = np.random.default_rng(5)
    a = public(rng, (4, N))
    s = np.zeros((4, N), np.uint16)
    s[2, 9] = 300                       # not an int8 value mod 2^13
    da, ds_, dc = T.DeviceBuffer.from_numpy(a), T.DeviceBuffer.from_numpy(s), T.DeviceBuffer(a.nbytes)
    ctx.polymul(dc, da, ds_, 4, 13, T.MFMA)
    assert ctx.status() == -5
    assert ctx.status() == 0            # the flag is cleared once reported
    # the VALU form is exact for any uint16 secret
    got = T.run_polymul(ctx, a, s, 13, T.VALU)
    assert np.array_equal(got, O.polymul_batch(a, s) & 8191)


def test_repeatable(ctx):
    """Same inputs twice -> identical bits (no races in the LDS staging)."""
    rng = np.random.default_rng(9)
    A = public(rng, (999, 16 * N))
    s = secret(rng, (999, 4 * N), "firesaber")
    r1 = T.run_matvec(ctx, A, s, "firesaber", T.MFMA)
    r2 = T.run_matvec(ctx, A, s, "firesaber", T.MFMA)
    assert np.array_equal(r1, r2)
