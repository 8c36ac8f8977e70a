"""GPU check of tmvp_kem_encaps_noalloc, the encapsulation whose scratch is
the library's fixed device array instead of an allocated workspace: through
the C ABI, two calls in a row (the second smaller, so it runs over what the
first left in the array) give exactly the bytes of tmvp_kem_encaps on the same
inputs, for every parameter set and both product backends; a batch beyond the
array's limit is refused before anything runs."""
import numpy as np
import pytest

import tmvp_amd as T

pytestmark = pytest.mark.gpu

# Largest batch the 32 MiB array holds, per matrix rank L (include/tmvp.h).
LIMIT = {2: 5318, 3: 3224, 4: 2161}


@pytest.fixture(scope="module")
def ctx():
    c = T.Context(0)
    yield c
    c.close()


def noalloc(ctx, ps, backend, ct, ss, pk, rnd, batch, stream):
    lib = T.load_library()
    return lib.tmvp_kem_encaps_noalloc(ctx.handle, T.PARAM_SET_NAMES[ps], backend, ct.ptr, ss.ptr, pk.ptr, rnd.ptr,
                                       int(batch), stream.handle)


@pytest.mark.parametrize("backend", [T.MFMA, T.VALU], ids=["mfma", "valu"])
@pytest.mark.parametrize("ps", ["lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3"])
def test_noalloc_encaps_equals_encaps(ctx, ps, backend):
    p = T.get_params(ps)
    rng = np.random.default_rng(300 + T.PARAM_SET_NAMES[ps] + 10 * backend)
    big = LIMIT[p["l"]]
    pk, _ = T.run_kem_keypair(ctx, rng.integers(0, 256, (big, 96), dtype=np.uint8), ps, T.MFMA)
    st = T.Stream()
    try:
        for batch in (big, 97):  # a full array, then a smaller batch over its leftovers
            rnd = rng.integers(0, 256, (batch, 32), dtype=np.uint8)
            dpk, drnd = T.DeviceBuffer.from_numpy(pk[:batch]), T.DeviceBuffer.from_numpy(rnd)
            ct, ss = T.DeviceBuffer(batch * p["ct_bytes"]), T.DeviceBuffer(batch * 32)
            assert noalloc(ctx, ps, backend, ct, ss, dpk, drnd, batch, st) == 0
            st.synchronize()
            ctx.check(st)
            want_ct, want_ss = T.run_kem_encaps(ctx, pk[:batch], rnd, ps, backend)
            assert np.array_equal(ct.download(np.uint8, (batch, -1)), want_ct), (ps, batch)
            assert np.array_equal(ss.download(np.uint8, (batch, -1)), want_ss), (ps, batch)
            for b in (dpk, drnd, ct, ss):
                b.free()
    finally:
        st.destroy()


def test_noalloc_refuses_a_batch_beyond_the_array(ctx):
    st = T.Stream()
    one = [T.DeviceBuffer(64) for _ in range(4)]
    try:
        for ps, l in (("saber", 3), ("firesaber", 4), ("lightsaber", 2)):
            rc = noalloc(ctx, ps, T.MFMA, one[0], one[1], one[2], one[3], LIMIT[l] + 1, st)
            assert rc == T.ENOMEM, ps
        assert noalloc(ctx, "saber", T.MFMA, one[0], one[1], one[2], one[3], 0, st) == 0
        assert noalloc(ctx, "saber", 7, one[0], one[1], one[2], one[3], 1, st) == T.EARG
    finally:
        st.destroy()
        for b in one:
            b.free()
