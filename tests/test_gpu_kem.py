"""GPU parity of the batched KEM (key generation, encapsulation,
decapsulation) through the C ABI, against the CPU oracle on the same seeded
inputs.  Byte work, so the bar is byte-exact equality of every pk, sk, ct and
shared secret, and of the rejection flag.

The oracle's KEM is pinned against the reference's test vectors (tv.h public
keys, pre-keys and ciphertext seeds) in tests/test_oracle.py.  At the full
serving batch the check is the round-trip property (decapsulation recovers
the encapsulated secret for every instance, a tampered ciphertext is
rejected) plus an oracle spot check of a sample of instances.
"""
import numpy as np
import pytest

import oracle as O
import tmvp_amd as T

pytestmark = pytest.mark.gpu

PS_ALL = ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"]
BACKENDS = [T.MFMA, T.VALU]


@pytest.fixture(scope="module")
def ctx():
    c = T.Context(0)
    yield c
    c.close()


def seeds_for(rng, batch):
    return rng.integers(0, 256, (batch, 96), dtype=np.uint8)


def tamper(ct, rng, rows):
    bad = ct.copy()
    for r in rows:
        pos = int(rng.integers(0, ct.shape[1]))
        bad[r, pos] ^= np.uint8(1 << int(rng.integers(0, 8)))
    return bad


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ps", PS_ALL)
def test_kem_keypair_vs_oracle(ctx, ps, backend):
    rng = np.random.default_rng(10 + PS_ALL.index(ps))
    seeds = seeds_for(rng, 67)
    pk, sk = T.run_kem_keypair(ctx, seeds, ps, backend)
    want_pk, want_sk = O.kem_keypair_batch(ps, seeds, nthreads=8)
    assert np.array_equal(pk, want_pk)
    assert np.array_equal(sk, want_sk)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ps", PS_ALL)
def test_kem_encaps_decaps_vs_oracle(ctx, ps, backend):
    rng = np.random.default_rng(20 + PS_ALL.index(ps))
    batch = 67
    pk, sk = O.kem_keypair_batch(ps, seeds_for(rng, batch), nthreads=8)
    rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
    ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, backend)
    want_ct, want_ss = O.kem_enc_batch(ps, rnd, pk, nthreads=8)
    assert np.array_equal(ct, want_ct)
    assert np.array_equal(ss, want_ss)

    # decapsulation of honest ciphertexts, and of a tampered third of them
    bad_rows = list(range(0, batch, 3))
    cts = tamper(ct, rng, bad_rows)
    ss_dec, fail = T.run_kem_decaps(ctx, cts, sk, ps, backend)
    want_dec = O.kem_dec_batch(ps, cts, sk, nthreads=8)
    assert np.array_equal(ss_dec, want_dec)
    expect_fail = np.zeros(batch, np.uint8)
    expect_fail[bad_rows] = 1
    assert np.array_equal(fail, expect_fail)
    good = np.setdiff1d(np.arange(batch), bad_rows)
    assert np.array_equal(ss_dec[good], ss[good])
    assert not np.any(np.all(ss_dec[bad_rows] == ss[bad_rows], axis=1))


@pytest.mark.parametrize("ps", ["saber", "sable2"])
def test_kem_tv_public_key(ctx, ps, saber_tv, sable_tv):
    """Encapsulation under the reference's own test-vector public key."""
    tv = saber_tv if ps == "saber" else sable_tv
    pk = np.frombuffer(tv["pk"] if ps == "saber" else tv["pk_tv"], np.uint8)[None, :]
    rnd = np.arange(32, dtype=np.uint8)[None, :]
    for backend in BACKENDS:
        ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, backend)
        want_ct, want_ss = O.kem_enc(ps, rnd[0], pk[0])
        assert ct[0].tobytes() == want_ct and ss[0].tobytes() == want_ss


def test_kem_backends_agree_full_batch(ctx):
    """Saber at the serving batch: both backends produce the same bytes, every
    honest ciphertext decapsulates to its secret, every tampered one is
    rejected, and a sample agrees with the oracle."""
    ps, batch = "saber", 16384
    rng = np.random.default_rng(99)
    seeds = seeds_for(rng, batch)
    pk, sk = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
    rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
    ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, T.MFMA)
    ct_v, ss_v = T.run_kem_encaps(ctx, pk, rnd, ps, T.VALU)
    assert np.array_equal(ct, ct_v) and np.array_equal(ss, ss_v)
    ss_dec, fail = T.run_kem_decaps(ctx, ct, sk, ps, T.MFMA)
    assert not fail.any()
    assert np.array_equal(ss_dec, ss)
    bad_rows = list(range(5, batch, 97))
    ss_bad, fail_bad = T.run_kem_decaps(ctx, tamper(ct, rng, bad_rows), sk, ps, T.MFMA)
    assert np.array_equal(np.flatnonzero(fail_bad), np.array(bad_rows))
    sample = rng.choice(batch, 48, replace=False)
    want_pk, want_sk = O.kem_keypair_batch(ps, seeds[sample], nthreads=8)
    assert np.array_equal(pk[sample], want_pk) and np.array_equal(sk[sample], want_sk)
    want_ct, want_ss = O.kem_enc_batch(ps, rnd[sample], want_pk, nthreads=8)
    assert np.array_equal(ct[sample], want_ct) and np.array_equal(ss[sample], want_ss)


@pytest.mark.parametrize("ps", ["saber", "firesaber", "sable1"])
def test_kem_one_launch_batch_edges(ctx, ps):
    """The MFMA backend runs key generation, encapsulation and decapsulation in
    one launch of 64 instances per workgroup, split 32 / 32 between two groups
    of sponge waves: batches that leave a workgroup partly filled, one group
    empty or a group with a single instance give the oracle's bytes, and a
    tampered third of the ciphertexts is rejected."""
    rng = np.random.default_rng(21 + PS_ALL.index(ps))
    for batch in (1, 31, 32, 33, 64, 65, 97, 130):
        seeds = seeds_for(rng, batch)
        pk, sk = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
        want_pk, want_sk = O.kem_keypair_batch(ps, seeds, nthreads=8)
        assert np.array_equal(pk, want_pk), (ps, batch)
        assert np.array_equal(sk, want_sk), (ps, batch)
        rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
        ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, T.MFMA)
        want_ct, want_ss = O.kem_enc_batch(ps, rnd, want_pk, nthreads=8)
        assert np.array_equal(ct, want_ct), (ps, batch)
        assert np.array_equal(ss, want_ss), (ps, batch)
        bad_rows = list(range(batch - 1, -1, -3))
        cts = tamper(ct, rng, bad_rows)
        dec, fail = T.run_kem_decaps(ctx, cts, sk, ps, T.MFMA)
        assert np.array_equal(dec, O.kem_dec_batch(ps, cts, want_sk, nthreads=8)), (ps, batch)
        expect_fail = np.zeros(batch, np.uint8)
        expect_fail[bad_rows] = 1
        assert np.array_equal(fail, expect_fail), (ps, batch)


@pytest.mark.parametrize("ps", PS_ALL)
def test_kem_one_launch_random_batches(ctx, ps):
    """Seeded random batch sizes up to 2500 (partly filled workgroups, claims
    of one or two instances at a product job's end): every operation of the
    one launch gives the oracle's bytes, a repeated key generation gives the
    same bytes, and every second ciphertext, tampered, is rejected."""
    rng = np.random.default_rng(60 + PS_ALL.index(ps))
    for batch in [int(x) for x in rng.integers(1, 2500, 4)]:
        seeds = seeds_for(rng, batch)
        pk, sk = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
        again = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
        want_pk, want_sk = O.kem_keypair_batch(ps, seeds, nthreads=8)
        assert np.array_equal(pk, want_pk) and np.array_equal(sk, want_sk), (ps, batch)
        assert np.array_equal(again[0], pk) and np.array_equal(again[1], sk), (ps, batch)
        rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
        ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, T.MFMA)
        want_ct, want_ss = O.kem_enc_batch(ps, rnd, want_pk, nthreads=8)
        assert np.array_equal(ct, want_ct) and np.array_equal(ss, want_ss), (ps, batch)
        bad_rows = list(range(0, batch, 2))
        cts = tamper(ct, rng, bad_rows)
        dec, fail = T.run_kem_decaps(ctx, cts, sk, ps, T.MFMA)
        assert np.array_equal(dec, O.kem_dec_batch(ps, cts, want_sk, nthreads=8)), (ps, batch)
        expect_fail = np.zeros(batch, np.uint8)
        expect_fail[bad_rows] = 1
        assert np.array_equal(fail, expect_fail), (ps, batch)


@pytest.mark.parametrize("ps", PS_ALL)
def test_kem_one_launch_equals_chain(ctx, ps):
    """Every parameter set at 700 instances (eleven workgroups of the one
    launch, the last partly filled): the MFMA backend's one-launch kernels and
    the VALU backend's kernel chains give the same keys, ciphertexts, shared
    keys and rejections."""
    rng = np.random.default_rng(40 + PS_ALL.index(ps))
    batch = 700
    seeds = seeds_for(rng, batch)
    pk, sk = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
    pk2, sk2 = T.run_kem_keypair(ctx, seeds, ps, T.VALU)
    assert np.array_equal(pk, pk2) and np.array_equal(sk, sk2)
    rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
    ct, ss = T.run_kem_encaps(ctx, pk, rnd, ps, T.MFMA)
    ct2, ss2 = T.run_kem_encaps(ctx, pk, rnd, ps, T.VALU)
    assert np.array_equal(ct, ct2) and np.array_equal(ss, ss2)
    bad_rows = list(range(3, batch, 7))
    cts = tamper(ct, rng, bad_rows)
    dec, fail = T.run_kem_decaps(ctx, cts, sk, ps, T.MFMA)
    dec2, fail2 = T.run_kem_decaps(ctx, cts, sk, ps, T.VALU)
    assert np.array_equal(dec, dec2) and np.array_equal(fail, fail2)
    expect_fail = np.zeros(batch, np.uint8)
    expect_fail[bad_rows] = 1
    assert np.array_equal(fail, expect_fail)
    good = np.setdiff1d(np.arange(batch), bad_rows)
    assert np.array_equal(dec[good], ss[good])
    sample = rng.choice(batch, 12, replace=False)
    assert np.array_equal(dec[sample], O.kem_dec_batch(ps, cts[sample], sk[sample], nthreads=8))


def test_kem_workspace_regrowth(ctx):
    """Batches that grow and shrink on one stream reuse and regrow the
    context's workspace without changing results."""
    ps = "firesaber"
    rng = np.random.default_rng(5)
    for batch in (3, 700, 40, 1500, 1):
        seeds = seeds_for(rng, batch)
        pk, sk = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
        want_pk, want_sk = O.kem_keypair_batch(ps, seeds, nthreads=8)
        assert np.array_equal(pk, want_pk) and np.array_equal(sk, want_sk)


def test_kem_argument_errors(ctx):
    buf = T.DeviceBuffer(4096)
    lib = T.load_library()
    h = ctx.handle
    assert lib.tmvp_kem_keypair(h, 6, T.MFMA, buf.ptr, buf.ptr, buf.ptr, 1, None) == -1
    assert lib.tmvp_kem_keypair(h, 1, 7, buf.ptr, buf.ptr, buf.ptr, 1, None) == -1
    assert lib.tmvp_kem_keypair(h, 1, T.MFMA, None, buf.ptr, buf.ptr, 1, None) == -1
    assert lib.tmvp_kem_encaps(h, 1, T.MFMA, buf.ptr, buf.ptr, buf.ptr, None, 1, None) == -1
    assert lib.tmvp_kem_decaps(h, 1, T.MFMA, buf.ptr, None, buf.ptr, None, 1, None) == -1
    assert lib.tmvp_kem_decaps(h, 1, T.MFMA, buf.ptr, buf.ptr, buf.ptr, None, -1, None) == -1
    assert lib.tmvp_kem_encaps(h, 1, T.MFMA, None, None, None, None, 0, None) == 0
    assert lib.tmvp_kem_encaps(None, 1, T.MFMA, buf.ptr, buf.ptr, buf.ptr, buf.ptr, 1, None) == -1
    buf.free()
