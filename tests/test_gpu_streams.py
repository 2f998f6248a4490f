"""The free-standing C-ABI calls on several streams and from several threads:
back-to-back products alternating between two streams with no
synchronisation between them (temporary buffers must not move between streams
while work is pending), and two threads calling at once, each on its own
stream; every result against the oracle.  Destroying a stream returns the
calling thread's cached temporaries of that stream."""
import ctypes
import threading

import numpy as np
import pytest

import oracle as O
import piranha_amd as P

pytestmark = pytest.mark.gpu

FP = 13


def new_stream():
    s = ctypes.c_void_p()
    P.check(P.lib().pir_stream_create(ctypes.byref(s)))
    return s.value


def cases(rng, count, same_shape=False):
    out = []
    for i in range(count):
        if same_shape:
            M, N, K = 64, 96, 160
        else:
            M, N, K = 64 + 8 * (i % 5), 96 + 16 * (i % 3), 160 + 32 * (i % 4)
        a = O.init_list(rng.uniform(-4, 4, M * K), FP)
        b = O.init_list(rng.uniform(-4, 4, K * N), FP)
        out.append((M, N, K, a, b, O.matmul(a, b, M, N, K, False, False, False, FP)))
    return out


def upload(cs):
    bufs = [(P.to_dev(c[3]), P.to_dev(c[4]), P.Buffer((3, 2, c[0] * c[1]), np.uint64)) for c in cs]
    P.synchronize()
    return bufs


def launch(case, bufs, stream):
    M, N, K = case[:3]
    da, db, dc = bufs
    P.check(P.lib().pir_rss_matmul(da.ptr, db.ptr, dc.ptr, M, N, K, 0, 0, 0, FP, P.ALGO_LIMB, stream))


def test_alternating_streams(rng):
    # one shape throughout: every call asks for temporaries of the same sizes,
    # so a cache keyed by size alone would hand the block the previous call on
    # the other stream had just released to a call whose stream has not yet
    # waited for that work
    streams = [new_stream(), new_stream()]
    try:
        cs = cases(rng, 16, same_shape=True)
        live = upload(cs)
        for i, (c, b) in enumerate(zip(cs, live)):  # back to back, no synchronisation
            launch(c, b, streams[i % 2])
        for s in streams:
            P.check(P.lib().pir_stream_synchronize(s))
        for c, (_, _, dc) in zip(cs, live):
            assert np.array_equal(dc.numpy(), c[5])
    finally:
        for s in streams:
            P.check(P.lib().pir_stream_destroy(s))


def test_two_threads(rng):
    work = [cases(rng, 10), cases(rng, 10)]
    errors = []

    def run(cs):
        s = new_stream()
        try:
            live = upload(cs)
            for c, b in zip(cs, live):
                launch(c, b, s)
            P.check(P.lib().pir_stream_synchronize(s))
            for c, (_, _, dc) in zip(cs, live):
                if not np.array_equal(dc.numpy(), c[5]):
                    errors.append("mismatch")
        except Exception as e:  # reported below
            errors.append(repr(e))
        finally:
            P.check(P.lib().pir_stream_destroy(s))

    threads = [threading.Thread(target=run, args=(w,)) for w in work]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads)
    assert errors == []


def test_stream_destroy_releases_cache(rng):
    L = P.lib()
    P.check(L.pir_release_cached())
    s = new_stream()
    c = cases(rng, 1)[0]
    b = upload([c])[0]
    launch(c, b, s)
    P.check(L.pir_stream_synchronize(s))
    assert np.array_equal(b[2].numpy(), c[5])
    assert L.pir_cached_bytes() > 0  # the product's temporaries, cached for stream s
    P.check(L.pir_stream_destroy(s))
    assert L.pir_cached_bytes() == 0
