"""GPU checks of the serving loop (tmvp_server): requests of mixed parameter
sets and operations, submitted together so that they share batches and
overlap on the device, give byte for byte what each gives alone (through the
server with nothing else queued, and through the plain batched KEM calls),
and a sample agrees with the CPU oracle."""
import numpy as np
import pytest

import oracle as O
import tmvp_amd as T

pytestmark = pytest.mark.gpu

SETS = ("saber", "sable2")
NKEYS = 600


@pytest.fixture(scope="module")
def ctx():
    c = T.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def keys(ctx):
    out = {}
    for i, ps in enumerate(SETS):
        seeds = np.random.default_rng(40 + i).integers(0, 256, (NKEYS, 96), dtype=np.uint8)
        out[ps] = T.run_kem_keypair(ctx, seeds, ps, T.MFMA)
    return out


def make_requests(ctx, keys, rng, n_req, max_count):
    """Random encaps / decaps requests; decaps inputs are honest ciphertexts
    with every fifth instance tampered."""
    reqs = []
    for r in range(n_req):
        ps = SETS[r % 2]
        op = T.SRV_ENCAPS if (r // 2) % 2 == 0 else T.SRV_DECAPS
        n = int(rng.integers(1, max_count + 1))
        ids = rng.integers(0, NKEYS, n).astype(np.int32)
        pk, sk = keys[ps][0][ids], keys[ps][1][ids]
        if op == T.SRV_ENCAPS:
            data = rng.integers(0, 256, (n, 32), dtype=np.uint8)
            want = T.run_kem_encaps(ctx, pk, data, ps, T.MFMA)
        else:
            ct, _ = T.run_kem_encaps(ctx, pk, rng.integers(0, 256, (n, 32), dtype=np.uint8), ps, T.MFMA)
            for i in range(0, n, 5):
                ct[i, int(rng.integers(0, ct.shape[1]))] ^= 0x10
            data = ct
            want = T.run_kem_decaps(ctx, ct, sk, ps, T.MFMA)
        reqs.append((op, ps, ids, data, want))
    return reqs


def run_all(srv, reqs):
    outs, done = {}, {}
    for k, (op, ps, ids, data, _) in enumerate(reqs):
        rid, o = srv.submit(op, ps, ids, data, tag=k)
        outs[k] = o
    while len(done) < len(reqs):
        for c in srv.poll(64, timeout_us=200000):
            done[c["tag"]] = c
    return outs, done


def test_serving_mix_equals_alone(ctx, keys):
    rng = np.random.default_rng(5)
    reqs = make_requests(ctx, keys, rng, 24, 300)
    srv = T.Server(ctx, max_batch=2048, slots=3)
    try:
        for ps in SETS:
            assert srv.add_keys(ps, *keys[ps]) == 0
        outs, done = run_all(srv, reqs)
        assert all(c["status"] == 0 for c in done.values())
        assert max(c["batch_requests"] for c in done.values()) > 1  # requests shared batches
        for k, (op, ps, ids, data, want) in enumerate(reqs):
            for got, exp in zip(outs[k], want):
                assert np.array_equal(got, exp), (k, op, ps)
        # the same requests alone
        for k in range(0, len(reqs), 5):
            op, ps, ids, data, want = reqs[k]
            _, o = srv.submit(op, ps, ids, data, tag=1000 + k)
            srv.drain()
            c = srv.poll(4, timeout_us=100000)
            assert len(c) == 1 and c[0]["batch_requests"] == 1 and c[0]["status"] == 0
            for got, exp in zip(o, outs[k]):
                assert np.array_equal(got, exp), k
        st = srv.stats()
        assert st["requests_done"] == len(reqs) + len(range(0, len(reqs), 5))
    finally:
        srv.close()
    # oracle sample: first encapsulation and first decapsulation request
    for op in (T.SRV_ENCAPS, T.SRV_DECAPS):
        k = next(i for i, r in enumerate(reqs) if r[0] == op)
        _, ps, ids, data, _ = reqs[k]
        if op == T.SRV_ENCAPS:
            ct, ss = O.kem_enc_batch(ps, data, keys[ps][0][ids], nthreads=8)
            assert np.array_equal(outs[k][0], ct) and np.array_equal(outs[k][1], ss)
        else:
            ss = O.kem_dec_batch(ps, data, keys[ps][1][ids], nthreads=8)
            assert np.array_equal(outs[k][0], ss)
            want_fail = np.zeros(len(ids), np.uint8)
            want_fail[::5] = 1
            assert np.array_equal(outs[k][1], want_fail)


def test_serving_batches_overlap(ctx, keys):
    """Large requests of both sets keep several slots busy at once."""
    rng = np.random.default_rng(6)
    srv = T.Server(ctx, max_batch=4096, slots=4)
    try:
        for ps in SETS:
            srv.add_keys(ps, *keys[ps])
        reqs = []
        for r in range(16):
            ps = SETS[r % 2]
            ids = rng.integers(0, NKEYS, 4096).astype(np.int32)
            reqs.append((T.SRV_ENCAPS, ps, ids, rng.integers(0, 256, (4096, 32), dtype=np.uint8), None))
        outs, done = run_all(srv, reqs)
        assert all(c["status"] == 0 for c in done.values())
        st = srv.stats()
        assert st["max_concurrent"] >= 2
        k = 3
        ct, ss = T.run_kem_encaps(ctx, keys[reqs[k][1]][0][reqs[k][2]], reqs[k][3], reqs[k][1], T.MFMA)
        assert np.array_equal(outs[k][0], ct) and np.array_equal(outs[k][1], ss)
    finally:
        srv.close()


def test_serving_errors(ctx, keys):
    with pytest.raises(T.TmvpError) as e:
        T.Server(ctx, hbm_budget=1 << 20, max_batch=16384)
    assert e.value.code == -3
    srv = T.Server(ctx, max_batch=64, slots=1)
    try:
        srv.add_keys("saber", *keys["saber"])
        srv.add_keys("sable2", keys["sable2"][0])  # encapsulation-only keys
        rnd = np.zeros((2, 32), np.uint8)
        for bad in ([0, NKEYS], [-1, 0]):
            with pytest.raises(T.TmvpError):
                srv.submit(T.SRV_ENCAPS, "saber", np.array(bad), rnd)
        with pytest.raises(T.TmvpError):
            srv.submit(T.SRV_ENCAPS, "saber", np.zeros(65, np.int32), np.zeros((65, 32), np.uint8))
        with pytest.raises(T.TmvpError):
            srv.submit(T.SRV_DECAPS, "sable2", np.zeros(1, np.int32), np.zeros((1, 1024), np.uint8))
        with pytest.raises(T.TmvpError):
            srv.submit(T.SRV_ENCAPS, "firesaber", np.zeros(1, np.int32), rnd[:1])  # no keys of that set
        rid, (ct, ss) = srv.submit(T.SRV_ENCAPS, "sable2", np.array([3, 4], np.int32), rnd)
        srv.drain()
        c = srv.poll(4, timeout_us=100000)
        assert c[0]["id"] == rid and c[0]["status"] == 0
        want_ct, want_ss = O.kem_enc_batch("sable2", rnd, keys["sable2"][0][[3, 4]])
        assert np.array_equal(ct, want_ct) and np.array_equal(ss, want_ss)
    finally:
        srv.close()


def test_serving_pinned_buffers(ctx, keys):
    """Requests whose buffers are page-locked go by DMA straight to and from
    them; results are the same as through the staging path."""
    rng = np.random.default_rng(12)
    reqs = make_requests(ctx, keys, rng, 12, 500)
    srv = T.Server(ctx, max_batch=1024, slots=2)
    try:
        for ps in SETS:
            srv.add_keys(ps, *keys[ps])
        outs = {}
        for k, (op, ps, ids, data, want) in enumerate(reqs):
            pdata = T.pinned_array(data.shape)
            pdata[:] = data
            o = (T.pinned_array(want[0].shape), T.pinned_array(want[1].shape))
            _, outs[k] = srv.submit(op, ps, ids, pdata if k % 3 else data, tag=k, out=o if k % 2 else None)
        done = {}
        while len(done) < len(reqs):
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
    finally:
        srv.close()
    for k, (op, ps, ids, data, want) in enumerate(reqs):
        for got, exp in zip(outs[k], want):
            assert np.array_equal(got, exp), (k, op, ps)


def test_serving_pinned_between_pageable(ctx, keys):
    """Within one batch, a request whose input is page-locked sits between two
    whose inputs are pageable: each request's input reaches the device whole,
    whichever memory it lives in."""
    rng = np.random.default_rng(21)
    reqs = make_requests(ctx, keys, rng, 12, 200)
    srv = T.Server(ctx, max_batch=4096, slots=1)
    try:
        for ps in SETS:
            srv.add_keys(ps, *keys[ps])
        # requests k, k + 4, k + 8 share a (set, operation): pageable, pinned, pageable
        inputs = []
        for k, (op, ps, ids, data, want) in enumerate(reqs):
            if (k // 4) % 2 == 1:
                pdata = T.pinned_array(data.shape)
                pdata[:] = data
                inputs.append(pdata)
            else:
                inputs.append(data.copy())
        # a large request first keeps the one slot busy while the others queue,
        # so each (set, operation) forms one batch: pageable, pinned, pageable
        blocker_ids = rng.integers(0, NKEYS, 4000).astype(np.int32)
        srv.submit(T.SRV_ENCAPS, "saber", blocker_ids, rng.integers(0, 256, (4000, 32), dtype=np.uint8), tag=99)
        outs = {}
        for k, (op, ps, ids, data, want) in enumerate(reqs):
            _, outs[k] = srv.submit(op, ps, ids, inputs[k], tag=k)
        done = {}
        while len(done) < len(reqs) + 1:
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
        assert max(c["batch_requests"] for c in done.values()) >= 2
    finally:
        srv.close()
    for k, (op, ps, ids, data, want) in enumerate(reqs):
        for got, exp in zip(outs[k], want):
            assert np.array_equal(got, exp), (k, op, ps)


@pytest.mark.parametrize("offset", [0, 4, 1])
def test_serving_pinned_input_alignment(ctx, keys, offset):
    """Page-locked inputs at 16-byte, 4-byte and odd addresses (the copy-in
    kernel's three access widths) give the plain batched call's results."""
    rng = np.random.default_rng(30 + offset)
    srv = T.Server(ctx, max_batch=2048, slots=2)
    try:
        for ps in SETS:
            srv.add_keys(ps, *keys[ps])
        reqs = []
        for ps in SETS:
            n = 97
            ids = rng.integers(0, NKEYS, n).astype(np.int32)
            rnd = rng.integers(0, 256, (n, 32), dtype=np.uint8)
            ct, ss = T.run_kem_encaps(ctx, keys[ps][0][ids], rnd, ps, T.MFMA)
            for op, data, want in ((T.SRV_ENCAPS, rnd, (ct, ss)),
                                   (T.SRV_DECAPS, ct, T.run_kem_decaps(ctx, ct, keys[ps][1][ids], ps, T.MFMA))):
                buf = T.pinned_array((data.size + 16,))
                view = buf[offset:offset + data.size].reshape(data.shape)
                view[:] = data
                reqs.append((op, ps, ids, view, want))
        outs = {}
        for k, (op, ps, ids, view, _) in enumerate(reqs):
            _, outs[k] = srv.submit(op, ps, ids, view, tag=k)
        done = {}
        while len(done) < len(reqs):
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
    finally:
        srv.close()
    for k, (op, ps, ids, view, want) in enumerate(reqs):
        for got, exp in zip(outs[k], want):
            assert np.array_equal(got, exp), (k, op, ps, offset)


@pytest.mark.parametrize("backend", [T.MFMA, T.VALU], ids=["mfma", "valu"])
def test_serving_all_parameter_sets(ctx, backend):
    """Keyed encapsulation and decapsulation through the server for every
    parameter set, on either product backend (the MFMA one reads A at the key
    id, the VALU one a gathered copy), give the plain batched calls' bytes."""
    rng = np.random.default_rng(77)
    sets = ("lightsaber", "saber", "firesaber", "sable1", "sable2", "sable3")
    srv = T.Server(ctx, max_batch=1024, slots=2, backend=backend)
    try:
        want, outs = {}, {}
        for i, ps in enumerate(sets):
            pk, sk = T.run_kem_keypair(ctx, rng.integers(0, 256, (40, 96), dtype=np.uint8), ps, T.MFMA)
            srv.add_keys(ps, pk, sk)
            ids = rng.integers(0, 40, 37).astype(np.int32)
            rnd = rng.integers(0, 256, (37, 32), dtype=np.uint8)
            ct, ss = T.run_kem_encaps(ctx, pk[ids], rnd, ps, T.MFMA)
            bad = ct.copy()
            bad[::4, 5] ^= 0x40
            want[(ps, 0)] = (ct, ss)
            want[(ps, 1)] = T.run_kem_decaps(ctx, bad, sk[ids], ps, T.MFMA)
            _, outs[(ps, 0)] = srv.submit(T.SRV_ENCAPS, ps, ids, rnd, tag=2 * i)
            _, outs[(ps, 1)] = srv.submit(T.SRV_DECAPS, ps, ids, bad, tag=2 * i + 1)
        done = {}
        while len(done) < 2 * len(sets):
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
    finally:
        srv.close()
    for key, exp in want.items():
        for got, e in zip(outs[key], exp):
            assert np.array_equal(got, e), key
    for ps in sets:
        assert outs[(ps, 1)][1][::4].all() and not outs[(ps, 1)][1][1::4].any(), ps


def test_serving_key_ids_copied_at_submit(ctx, keys):
    """The server checks and copies a request's key ids at submit: overwriting
    the caller's id array right after submit (with ids that would be out of
    range) changes nothing, however late the request's batch is launched."""
    rng = np.random.default_rng(90)
    srv = T.Server(ctx, max_batch=2048, slots=2)
    try:
        srv.add_keys("saber", *keys["saber"])
        reqs = []
        for k in range(24):
            n = int(rng.integers(200, 2048))
            ids = rng.integers(0, NKEYS, n).astype(np.int32)
            keep = ids.copy()
            rnd = rng.integers(0, 256, (n, 32), dtype=np.uint8)
            _, out = srv.submit(T.SRV_ENCAPS, "saber", ids, rnd, tag=k)
            ids[:] = -1  # the caller reuses its id buffer at once
            reqs.append((keep, rnd, out))
        done = {}
        while len(done) < len(reqs):
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
    finally:
        srv.close()
    for k in (0, 11, 23):
        keep, rnd, (ct, ss) = reqs[k]
        want_ct, want_ss = T.run_kem_encaps(ctx, keys["saber"][0][keep], rnd, "saber", T.MFMA)
        assert np.array_equal(ct, want_ct) and np.array_equal(ss, want_ss), k


def test_serving_add_keys_rejects_mismatched_sk(ctx, keys):
    """An sk that does not hold its own pk is refused: keyed decapsulation
    would take A from one key and b, H(pk), z from the other."""
    pk, sk = keys["saber"]
    srv = T.Server(ctx, max_batch=64, slots=1)
    try:
        with pytest.raises(T.TmvpError) as e:
            srv.add_keys("saber", pk[:8], sk[1:9])
        assert e.value.code == T.EARG
        assert srv.add_keys("saber", pk[:8], sk[:8]) == 0  # nothing was added by the refused call
    finally:
        srv.close()


def test_serving_store_growth_keeps_keys(ctx):
    """Keys added in several calls, each growing the store (the keys already
    there are copied to the new allocation), serve the same bytes as the plain
    batched calls, low and high key ids alike."""
    rng = np.random.default_rng(91)
    total = 2300
    pk, sk = T.run_kem_keypair(ctx, rng.integers(0, 256, (total, 96), dtype=np.uint8), "saber", T.MFMA)
    srv = T.Server(ctx, max_batch=1024, slots=2)
    try:
        first = [srv.add_keys("saber", pk[a:b], sk[a:b]) for a, b in ((0, 700), (700, 1400), (1400, total))]
        assert first == [0, 700, 1400]
        ids = np.concatenate([rng.integers(0, 1024, 100), rng.integers(1024, total, 300)]).astype(np.int32)
        rnd = rng.integers(0, 256, (ids.size, 32), dtype=np.uint8)
        want_ct, want_ss = T.run_kem_encaps(ctx, pk[ids], rnd, "saber", T.MFMA)
        _, (ct, ss) = srv.submit(T.SRV_ENCAPS, "saber", ids, rnd, tag=0)
        _, (ss2, fail) = srv.submit(T.SRV_DECAPS, "saber", ids, want_ct, tag=1)
        done = {}
        while len(done) < 2:
            for c in srv.poll(8, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
    finally:
        srv.close()
    assert np.array_equal(ct, want_ct) and np.array_equal(ss, want_ss)
    assert np.array_equal(ss2, want_ss) and not fail.any()


def test_serving_admission_by_hbm_budget(ctx, keys):
    """The HBM budget bounds the device bytes of the key stores plus the
    batches in flight, each batch reserving what its set, operation and size
    need: with a budget of the stores plus ONE max_batch batch, smaller
    batches of the four queues still run side by side; requests beyond what
    fits wait (none fails) and give the plain calls' bytes; keys that would
    leave less than one max_batch batch are refused."""
    rng = np.random.default_rng(92)
    sizing = T.Server(ctx, max_batch=4096, slots=4)
    try:
        for ps in SETS:
            sizing.add_keys(ps, *keys[ps])
        st0 = sizing.stats()
    finally:
        sizing.close()
    budget = st0["store_bytes"] + st0["slot_bytes"]
    srv = T.Server(ctx, hbm_budget=budget, max_batch=4096, slots=4)
    reqs = []
    try:
        for ps in SETS:
            srv.add_keys(ps, *keys[ps])
        with pytest.raises(T.TmvpError) as e:  # more keys would not leave one batch
            srv.add_keys("saber", *keys["saber"])
        assert e.value.code == T.ENOMEM
        outs, done = {}, {}
        for k in range(24):
            ps = SETS[k % 2]
            op = T.SRV_ENCAPS if (k // 2) % 2 == 0 else T.SRV_DECAPS
            n = int(rng.integers(600, 1400))
            ids = rng.integers(0, NKEYS, n).astype(np.int32)
            if op == T.SRV_ENCAPS:
                data = rng.integers(0, 256, (n, 32), dtype=np.uint8)
            else:
                data, _ = T.run_kem_encaps(ctx, keys[ps][0][ids], rng.integers(0, 256, (n, 32), dtype=np.uint8), ps,
                                           T.MFMA)
            reqs.append((op, ps, ids, data))
        for k, (op, ps, ids, data) in enumerate(reqs):
            _, outs[k] = srv.submit(op, ps, ids, data, tag=k)
        while len(done) < len(reqs):
            for c in srv.poll(64, timeout_us=200000):
                done[c["tag"]] = c
        assert all(c["status"] == 0 for c in done.values())
        st = srv.stats()
    finally:
        srv.close()
    assert st["store_bytes"] == st0["store_bytes"] and st["hbm_budget"] == budget
    assert st["max_inflight_bytes"] <= budget - st["store_bytes"]
    assert st["max_concurrent"] >= 2  # a fixed slot count from this budget would allow one
    for k in (0, 1, 2, 3, 22, 23):
        op, ps, ids, data = reqs[k]
        if op == T.SRV_ENCAPS:
            want = T.run_kem_encaps(ctx, keys[ps][0][ids], data, ps, T.MFMA)
        else:
            want = T.run_kem_decaps(ctx, data, keys[ps][1][ids], ps, T.MFMA)
        for got, exp in zip(outs[k], want):
            assert np.array_equal(got, exp), (k, op, ps)
