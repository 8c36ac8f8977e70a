"""Pins the CPU oracle (oracle/tmvp_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or
index, a transposed operand, a wrong rounding constant) fails it:

* FIPS-202 against Python's hashlib (an independent implementation).
* GenMatrix / GenSecret / SHA3 against the reference's test vectors
  (TMVP-Saber/SABER-GPU/include/tv.h, TMVP_SABLE/include/tv.h).
* MatrixVectorMul (transposed) + q->p rounding against the public keys in
  those headers: pk = round(A^T s) with s unpacked from sk.
* The non-transposed MatrixVectorMul, InnerProd and the encaps/decaps rounding
  through the IND-CPA round trip on the reference's own key pair.
* The polynomial product against numpy's linear convolution folded
  negacyclically, against closed forms (x^k rotates with sign, 1 is the
  identity), and Toom-4 against schoolbook mod 2^13.
"""
import hashlib

import numpy as np
import pytest

import oracle as O

N = 256


def rng(seed=0):
    return np.random.default_rng(seed)


def negacyclic_numpy(a, s):
    full = np.convolve(a.astype(np.int64), s.astype(np.int64))
    full = np.concatenate([full, [0]])
    return ((full[:N] - full[N:]) % 65536).astype(np.uint16)


# ---------------------------------------------------------------- FIPS 202
@pytest.mark.parametrize("n", [0, 1, 31, 32, 71, 72, 73, 135, 136, 137, 167, 168, 169, 500])
def test_fips202_vs_hashlib(n):
    data = bytes((7 * i + 3) & 255 for i in range(n))
    assert O.sha3_256(data) == hashlib.sha3_256(data).digest()
    assert O.sha3_512(data) == hashlib.sha3_512(data).digest()
    for outlen in (1, 168, 169, 3744):
        assert O.shake128(data, outlen) == hashlib.shake_128(data).digest(outlen)


# ---------------------------------------------------------------- golden vectors (Saber)
def test_saber_message_is_hash_of_buf(saber_tv):
    assert O.sha3_256(saber_tv["buf_tv"][:32]) == bytes(saber_tv["m_tv"].astype(np.uint8))


def test_saber_genmatrix_matches_tv(saber_tv):
    pk = saber_tv["pk"]
    assert np.array_equal(O.gen_matrix("saber", pk[960:992]), saber_tv["A"])


def test_saber_gensecret_matches_tv(saber_tv):
    m = O.sha3_256(saber_tv["buf_tv"][:32])
    kr = O.sha3_512(m + O.sha3_256(saber_tv["pk"]))
    assert np.array_equal(O.gen_secret("saber", kr[32:]), saber_tv["sp"])


def test_saber_sk_layout(saber_tv):
    sk, pk = saber_tv["sk_tv"], saber_tv["pk"]
    assert sk[1248:2240] == pk
    assert sk[2240:2272] == O.sha3_256(pk)
    s = O.unpack_bits(sk[:1248], 768, 13).astype(np.int32)
    s = np.where(s >= 4096, s - 8192, s)
    assert s.min() >= -4 and s.max() <= 4


def test_saber_pk_is_rounded_transposed_matvec(saber_tv):
    sk, pk, A = saber_tv["sk_tv"], saber_tv["pk"], saber_tv["A"]
    s = O.unpack_bits(sk[:1248], 768, 13)
    for method in (0, 1):
        b = O.matvec_batch(3, A, s, transpose=True, method=method).reshape(-1)
        assert O.pack_bits(O.round_q2p("saber", b), 10) == pk[:960]
    # the non-transposed product must NOT reproduce pk (guards a transposed operand)
    b = O.matvec_batch(3, A, s, transpose=False).reshape(-1)
    assert O.pack_bits(O.round_q2p("saber", b), 10) != pk[:960]


def test_saber_indcpa_round_trip(saber_tv):
    m = O.sha3_256(saber_tv["buf_tv"][:32])
    kr = O.sha3_512(m + O.sha3_256(saber_tv["pk"]))
    ct = O.indcpa_enc("saber", m, kr[32:], saber_tv["pk"])
    assert O.indcpa_dec("saber", saber_tv["sk_tv"][:1248], ct) == m


def test_saber_kem_round_trip_on_tv_keys(saber_tv):
    ct, ss = O.kem_enc("saber", saber_tv["buf_tv"][:32], saber_tv["pk"])
    ss2, fail = O.kem_dec("saber", ct, saber_tv["sk_tv"])
    assert not fail and ss == ss2
    bad = bytearray(ct)
    bad[5] ^= 1
    ss3, fail = O.kem_dec("saber", bytes(bad), saber_tv["sk_tv"])
    assert fail and ss3 != ss


# ---------------------------------------------------------------- golden vectors (Sable type 2)
def sable_secret(sk):
    codes = O.unpack_bits(sk[:192], 768, 2)
    return ((codes ^ 2).astype(np.int32) - 2).astype(np.uint16)


def test_sable_hashes_match_tv(sable_tv):
    m = O.sha3_256(sable_tv["buf_tv"][:32])
    assert m == bytes(sable_tv["m_tv"].astype(np.uint8))
    assert O.sha3_512(m + O.sha3_256(sable_tv["pk_tv"])) == sable_tv["kr_tv"]


def test_sable_sk_layout(sable_tv):
    sk, pk = sable_tv["sk_tv"], sable_tv["pk_tv"]
    assert sk[192:1088] == pk
    assert sk[1088:1120] == O.sha3_256(pk)
    s = sable_secret(sk).astype(np.int16)
    assert set(np.unique(s)) <= {-1, 0, 1}


def test_sable_pk_is_rounded_transposed_matvec(sable_tv):
    pk = sable_tv["pk_tv"]
    A = O.gen_matrix("sable2", pk[864:896])
    s = sable_secret(sable_tv["sk_tv"])
    for method in (0, 1):
        b = O.matvec_batch(3, A, s, transpose=True, method=method).reshape(-1)
        assert O.pack_bits(O.round_q2p("sable2", b), 9) == pk[:864]


def test_sable_indcpa_and_kem_round_trip(sable_tv):
    m = O.sha3_256(sable_tv["buf_tv"][:32])
    ct = O.indcpa_enc("sable2", m, sable_tv["kr_tv"][32:], sable_tv["pk_tv"])
    assert O.indcpa_dec("sable2", sable_tv["sk_tv"][:192], ct) == m
    ct, ss = O.kem_enc("sable2", sable_tv["buf_tv"][:32], sable_tv["pk_tv"])
    ss2, fail = O.kem_dec("sable2", ct, sable_tv["sk_tv"])
    assert not fail and ss == ss2


# ---------------------------------------------------------------- polynomial product
def test_polymul_vs_numpy_convolution():
    r = rng(1)
    for _ in range(8):
        a = r.integers(0, 65536, N, dtype=np.uint16)
        s = r.integers(0, 65536, N, dtype=np.uint16)
        assert np.array_equal(O.polymul(a, s), negacyclic_numpy(a, s))


def test_polymul_closed_forms():
    r = rng(2)
    a = r.integers(0, 8192, N, dtype=np.uint16)
    one = np.zeros(N, np.uint16)
    one[0] = 1
    assert np.array_equal(O.polymul(a, one), a)
    for k in (1, 17, 128, 255):
        xk = np.zeros(N, np.uint16)
        xk[k] = 1
        expect = np.concatenate([(-a[N - k:].astype(np.int64)) % 65536, a[:N - k]]).astype(np.uint16)
        assert np.array_equal(O.polymul(a, xk), expect)
    # x^255 * x = x^256 = -1
    x255 = np.zeros(N, np.uint16)
    x255[255] = 1
    x1 = np.zeros(N, np.uint16)
    x1[1] = 1
    expect = np.zeros(N, np.uint16)
    expect[0] = 65535
    assert np.array_equal(O.polymul(x255, x1), expect)


def test_polymul_commutative_and_linear():
    r = rng(3)
    a, b = (r.integers(0, 65536, N, dtype=np.uint16) for _ in range(2))
    s = r.integers(-5, 6, N).astype(np.int16).view(np.uint16)
    assert np.array_equal(O.polymul(a, s), O.polymul(s, a))
    lhs = O.polymul((a.astype(np.uint32) + b) .astype(np.uint16), s)
    rhs = (O.polymul(a, s).astype(np.uint32) + O.polymul(b, s)).astype(np.uint16)
    assert np.array_equal(lhs, rhs)


def test_toom4_equals_schoolbook_mod_2_13():
    r = rng(4)
    a = r.integers(0, 65536, (64, N), dtype=np.uint16)
    s = r.integers(0, 65536, (64, N), dtype=np.uint16)
    c0 = O.polymul_batch(a, s, method=0)
    c1 = O.polymul_batch(a, s, method=1)
    assert np.array_equal(c0 & 8191, c1 & 8191)


def test_batch_threads_consistent():
    r = rng(5)
    A = r.integers(0, 8192, (33, 9, N), dtype=np.uint16)
    s = r.integers(-4, 5, (33, 3, N)).astype(np.int16).view(np.uint16)
    one = O.matvec_batch(3, A, s, nthreads=1)
    many = O.matvec_batch(3, A, s, nthreads=4)
    assert np.array_equal(one, many)


# ---------------------------------------------------------------- rounding
@pytest.mark.parametrize("ps", ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"])
def test_rounding_family_forms_equal_closed_form(ps):
    """Both families' reference forms equal the mod-2^k closed forms, for every uint16 input."""
    p = O.params(ps)
    eq, ep, et, h1, h2 = p["eq"], p["ep"], p["et"], p["h1"], p["h2"]
    x = np.arange(65536, dtype=np.int64)
    assert np.array_equal(O.round_q2p(ps, x), ((x + h1) % (1 << eq)) >> (eq - ep))
    for bit in (0, 1):
        v = O.round_enc_v(ps, x, np.full_like(x, bit))
        assert np.array_equal(v, ((x + h1 - (bit << (ep - 1))) % (1 << ep)) >> (ep - et))
    c = x % (1 << et)
    m = O.round_dec_m(ps, x, c)
    assert np.array_equal(m, ((x + h2 - (c << (ep - et))) % (1 << ep)) >> (ep - 1))


ALL_SETS = ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"]
# Decryption noise tolerance |e| <= 2^(EP-2) - 2^(EP-ET-1) per set (derived below).
DECODE_WINDOW = {"lightsaber": 192, "saber": 224, "firesaber": 248, "sable1": 96, "sable2": 120, "sable3": 192}


def rounding_tables(ps):
    """enc[m][v] = or_round_enc_v(v, m) for v in Z_p, and dec[c][x] = or_round_dec_m(x, c)
    for every x in Z_p and every wire value c in [0, 2^ET)."""
    p = O.params(ps)
    P, T = 1 << p["ep"], 1 << p["et"]
    v = np.arange(P)
    enc = np.stack([O.round_enc_v(ps, v, np.full(P, m)) for m in (0, 1)]).astype(np.int64)
    dec = np.stack([O.round_dec_m(ps, v, np.full(P, c)) for c in range(T)]).astype(np.int64)
    return enc, dec


def decode_failures(ps, enc, dec, e):
    """Messages bits that decode wrongly when decryption sees v + e instead of v,
    over every v in Z_p and m in {0, 1}."""
    P = 1 << O.params(ps)["ep"]
    v = np.arange(P)
    bad = 0
    for m in (0, 1):
        c = enc[m]
        got = dec[c, (v + e) % P]
        bad += int(np.count_nonzero(got != m))
    return bad


@pytest.mark.parametrize("ps", ALL_SETS)
def test_encaps_decaps_rounding_decoding_window(ps):
    """Pins the encryption rounding (post_process2, S/src/poly.cu:391-395; SB/src/pack.cu:290-335)
    and the decryption rounding (post_process3, S/src/poly.cu:398-403; SB/src/pack.cu:436-442)
    from what they are for, not from their formula.

    Encryption sends c = floor((v - m 2^(EP-1) + h1) / 2^(EP-ET)) mod 2^ET, so
    c 2^(EP-ET) = v - m 2^(EP-1) + h1 - r with r in [0, 2^(EP-ET)) depending on v.
    Decryption sees v + e and computes x = v + e + h2 - c 2^(EP-ET)
    = m 2^(EP-1) + e + (h2 - h1) + r (mod p), and m is recovered exactly when
    0 <= e + (h2 - h1) + r < 2^(EP-1) for every r.  With h2 - h1 = 2^(EP-2) - 2^(EP-ET-1)
    that is the symmetric window |e| <= W = 2^(EP-2) - 2^(EP-ET-1): every (v, m)
    decodes for |e| <= W and some (v, m) fails at e = W + 1 and at e = -(W + 1).
    A dropped h1, an h2 off by a step, a wrong shift or a sign slip in either
    rounding moves or narrows the window and fails here."""
    p = O.params(ps)
    W = (1 << (p["ep"] - 2)) - (1 << (p["ep"] - p["et"] - 1))
    assert W == DECODE_WINDOW[ps]
    enc, dec = rounding_tables(ps)
    assert enc.min() >= 0 and enc.max() < (1 << p["et"])
    assert set(np.unique(dec)) <= {0, 1}
    for e in range(-W, W + 1):
        assert decode_failures(ps, enc, dec, e) == 0, f"{ps}: noise {e} inside the window fails"
    assert decode_failures(ps, enc, dec, W + 1) > 0, f"{ps}: window wider than W on the + side"
    assert decode_failures(ps, enc, dec, -(W + 1)) > 0, f"{ps}: window wider than W on the - side"


def test_decoding_window_catches_planted_rounding_slips():
    """The window check above rejects the plausible slips: h1 dropped from the
    encryption rounding, h2 one half-step off, and the two message halves swapped."""
    ps = "saber"
    p = O.params(ps)
    P, T = 1 << p["ep"], 1 << p["et"]
    ep, et, h1, h2 = p["ep"], p["et"], p["h1"], p["h2"]
    v = np.arange(P)
    W = DECODE_WINDOW[ps]

    def window_ok(enc_f, dec_f):
        enc = np.stack([enc_f(v, m) for m in (0, 1)])
        dec = np.stack([dec_f(v, c) for c in range(T)])
        inside = all(decode_failures(ps, enc, dec, e) == 0 for e in range(-W, W + 1))
        edges = decode_failures(ps, enc, dec, W + 1) > 0 and decode_failures(ps, enc, dec, -W - 1) > 0
        return inside and edges

    def enc_ok(x, m, hh=h1):
        return ((x - (m << (ep - 1)) + hh) % P) >> (ep - et)

    def dec_ok(x, c, hh=h2):
        return ((x + hh - (c << (ep - et))) % P) >> (ep - 1)

    enc_tab, dec_tab = rounding_tables(ps)
    assert window_ok(lambda x, m: enc_tab[m][x], lambda x, c: dec_tab[c][x])   # the oracle itself
    assert not window_ok(lambda x, m: enc_ok(x, m, 0), dec_ok)                 # h1 dropped
    assert not window_ok(enc_ok, lambda x, c: dec_ok(x, c, h2 + (1 << (ep - et - 1))))
    assert not window_ok(enc_ok, lambda x, c: dec_ok(x, c, h2 - (1 << (ep - et - 1))))
    assert not window_ok(lambda x, m: enc_ok(x, 1 - m), dec_ok)                # message halves swapped


def test_params_table_matches_headers():
    # byte sizes in S/include/params.h and SB/include/SABLE_params.h (Saber L=3, Sable type 2)
    s = O.params("saber")
    assert (s["pk_bytes"], s["sk_bytes"], s["ct_bytes"], s["h1"], s["h2"]) == (992, 2304, 1088, 4, 228)
    t = O.params("sable2")
    assert (t["pk_bytes"], t["sk_bytes"], t["ct_bytes"], t["h1"], t["h2"]) == (896, 1152, 1024, 2, 122)
    assert O.params("lightsaber")["h2"] == 196 and O.params("firesaber")["h2"] == 252
    assert O.params("sable1")["h2"] == 98 and O.params("sable3")["h2"] == 193


@pytest.mark.parametrize("ps", ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"])
def test_kem_round_trip_all_sets(ps):
    r = rng(6)
    for _ in range(3):
        seed_a, seed_s, z, rnd = (bytes(r.integers(0, 256, 32, dtype=np.uint8)) for _ in range(4))
        pk, sk = O.kem_keypair(ps, seed_a, seed_s, z)
        ct, ss = O.kem_enc(ps, rnd, pk)
        ss2, fail = O.kem_dec(ps, ct, sk)
        assert not fail and ss == ss2


@pytest.mark.parametrize("ps", ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"])
def test_cbd_range_and_distribution(ps):
    p = O.params(ps)
    r = rng(7)
    buf = r.integers(0, 256, p["mu"] * N // 8, dtype=np.uint8)
    s = O.cbd(ps, buf).astype(np.int64)
    q = 1 << p["eq"]
    s = np.where(s >= 32768, s - 65536, s)   # Saber: two's complement
    s = np.where(s >= q // 2, s - q, s)      # Sable: mod q
    assert s.min() >= -p["mu"] // 2 and s.max() <= p["mu"] // 2
    # brute-force definition on the first coefficients
    bits = np.unpackbits(buf, bitorder="little").astype(np.int64)
    mu = p["mu"]
    for k in range(8):
        chunk = bits[k * mu:(k + 1) * mu]
        assert s[k] == chunk[: mu // 2].sum() - chunk[mu // 2:].sum()
