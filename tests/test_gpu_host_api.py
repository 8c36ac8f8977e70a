"""GPU checks of the reference-named host entry points (saber_enc,
saber_dec, SABLE_enc, SABLE_dec): host buffers in, host buffers out, bytes
equal to the CPU oracle's KEM for the same keys and randomness, for every
reference mode."""
import ctypes

import numpy as np
import pytest

import oracle as O
import tmvp_amd as T

pytestmark = pytest.mark.gpu


def _buf(arr):
    return arr.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture()
def lib():
    lib = T.load_library()
    yield lib
    assert lib.tmvp_host_release() == 0


@pytest.mark.parametrize("family,ps,batch", [("saber", "saber", 3), ("saber", "lightsaber", 5),
                                             ("sable", "sable2", 7), ("sable", "sable3", 2)])
def test_reference_host_functions_match_oracle(lib, family, ps, batch):
    rng = np.random.default_rng(batch)
    pid = T.PARAM_SET_NAMES[ps]
    if family == "saber":
        assert lib.tmvp_host_configure(0, batch, pid, 1, T.SABLE2) == 0
    else:
        assert lib.tmvp_host_configure(0, 1, T.SABER, batch, pid) == 0
    pk, sk = O.kem_keypair_batch(ps, rng.integers(0, 256, (batch, 96), dtype=np.uint8), nthreads=4)
    p = O.params(ps)
    for mode in (0, 1, 2):
        rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
        want_ct, want_ss = O.kem_enc_batch(ps, rnd, pk, nthreads=4)
        ct = np.zeros((batch, p["ct_bytes"]), np.uint8)
        k = np.zeros((batch, 32), np.uint8)
        k2 = np.zeros((batch, 32), np.uint8)
        assert lib.tmvp_host_set_randomness(_buf(rnd), batch) == 0
        if family == "saber":
            lib.saber_enc(mode, _buf(k), _buf(pk), _buf(ct))
        else:
            lib.SABLE_enc(_buf(k), _buf(pk), _buf(ct), mode)
        assert lib.tmvp_host_status() == 0
        assert np.array_equal(ct, want_ct) and np.array_equal(k, want_ss)
        if family == "saber":
            lib.saber_dec(mode, _buf(ct), _buf(pk), _buf(sk), _buf(k2))
        else:
            lib.SABLE_dec(_buf(ct), _buf(pk), _buf(sk), _buf(k2), mode)
        assert lib.tmvp_host_status() == 0
        assert np.array_equal(k2, want_ss)


def test_reference_host_functions_fresh_randomness(lib):
    """Without caller randomness each encapsulation draws its own: two calls
    give different ciphertexts, and each decapsulates to its secret."""
    batch = 4
    assert lib.tmvp_host_configure(0, batch, T.SABER, 1, T.SABLE2) == 0
    pk, sk = O.kem_keypair_batch("saber", np.arange(batch * 96, dtype=np.uint32).astype(np.uint8).reshape(batch, 96))
    outs = []
    for _ in range(2):
        ct = np.zeros((batch, 1088), np.uint8)
        k = np.zeros((batch, 32), np.uint8)
        k2 = np.zeros((batch, 32), np.uint8)
        lib.saber_enc(0, _buf(k), _buf(pk), _buf(ct))
        assert lib.tmvp_host_status() == 0
        lib.saber_dec(0, _buf(ct), _buf(pk), _buf(sk), _buf(k2))
        assert lib.tmvp_host_status() == 0
        assert np.array_equal(k, k2)
        outs.append(ct)
    assert not np.array_equal(outs[0], outs[1])


def test_randomness_for_another_batch_is_an_error(lib):
    """Randomness set for a batch other than the configured one is refused (an
    error status, nothing written), not silently replaced by fresh randomness;
    it is dropped, so the next encapsulation draws its own."""
    batch = 4
    assert lib.tmvp_host_configure(0, batch, T.SABER, 1, T.SABLE2) == 0
    pk, sk = O.kem_keypair_batch("saber", np.arange(batch * 96, dtype=np.uint32).astype(np.uint8).reshape(batch, 96))
    rnd = np.ones((batch - 1, 32), np.uint8)
    assert lib.tmvp_host_set_randomness(_buf(rnd), batch - 1) == 0
    ct = np.zeros((batch, 1088), np.uint8)
    k = np.zeros((batch, 32), np.uint8)
    lib.saber_enc(0, _buf(k), _buf(pk), _buf(ct))
    assert lib.tmvp_host_status() == T.EARG
    assert not ct.any() and not k.any()
    lib.saber_enc(0, _buf(k), _buf(pk), _buf(ct))
    assert lib.tmvp_host_status() == 0
    k2 = np.zeros((batch, 32), np.uint8)
    lib.saber_dec(0, _buf(ct), _buf(pk), _buf(sk), _buf(k2))
    assert lib.tmvp_host_status() == 0 and np.array_equal(k, k2)
