"""The run boundary of the C ABI on the GPU: history-key ranges (shards of
one run), a caller's stream, device-pointer outputs, and problems built by
mc_problem_create.

A history's stream is fixed by its key alone (RN_init_particle,
Random.cpp:196-204), so runs over disjoint key ranges are shards of one run:
their per-history records are the full run's, bit for bit, and their moment
sums add up to the full run's to rounding (the sums themselves are FP64
additions in a launch-dependent order)."""
import ctypes as C

import numpy as np
import pytest

from mc_testlib import HipRuntime, decks, mc, oracle_run, shield_xs_dir

pytestmark = pytest.mark.gpu

FIELDS = ("rng", "flights", "x", "y", "z", "E", "score")


def same_records(a, b, what):
    for f in FIELDS:
        bad = np.nonzero(a[f] != b[f])[0]
        assert bad.size == 0, f"{what}: {f} differs at {bad[:8].tolist()}"


def test_two_half_ranges_are_the_full_run():
    n = 12000
    p = mc.Problem.from_string(decks.shield_deck(n), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    full = eng.forward(n_histories=n, trace=n)
    lo = eng.forward(n_histories=n // 2, key_offset=0, trace=n // 2)
    hi = eng.forward(n_histories=n // 2, key_offset=n // 2, trace=n // 2)
    same_records(np.concatenate([lo.trace, hi.trace]), full.trace, "halves")
    assert lo.flights + hi.flights == full.flights
    np.testing.assert_allclose(lo.sum + hi.sum, full.sum, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(lo.sq + hi.sq, full.sq, rtol=1e-12, atol=1e-300)


def test_gradient_shards_add_up():
    n = 6000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    go = np.random.default_rng(4).normal(size=p.n_bins)
    full = eng.backward(go, n_histories=n)
    a = eng.backward(go, n_histories=n // 3, key_offset=0)
    b = eng.backward(go, n_histories=n - n // 3, key_offset=n // 3)
    # grad = gsum / N per run: recombine the shards' sums
    np.testing.assert_allclose((a.grad * (n // 3) + b.grad * (n - n // 3)) / n, full.grad, rtol=1e-9, atol=1e-13)


def test_batched_key_offsets_match_oracle():
    """Three problems of one engine on caller-chosen key ranges (one above
    2^32, one overlapping another problem's range) against the oracle run
    on the same keys."""
    n = 2000
    probs = [mc.Problem.from_string(decks.slab_deck(2.0, c, 6, n)) for c in (0.3, 0.6, 0.9)]
    eng = mc.Engine(probs, bank_capacity=1 << 11)
    offs = [5 * 10 ** 9, 0, 1000]
    g = eng.forward(key_offset=offs, trace=n)
    for p in range(3):
        o = oracle_run(probs[p], n, key_offset=offs[p], trace=n)
        same_records(g.trace[p * n:(p + 1) * n], o.trace, f"problem {p}")
        np.testing.assert_allclose(g.sum[eng.bin_range(p)], o.sum, rtol=1e-12, atol=1e-300)


def test_device_pointers_on_a_caller_stream():
    """Outputs written to device buffers by a run issued on a caller's HIP
    stream: the same records, bit for bit, and the same moments to rounding
    as the host-pointer run; gradient weights read from and gradients
    written to device memory likewise."""
    n = 8000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    host = eng.forward(n_histories=n, trace=n)
    hip = HipRuntime()
    nb, npar = p.n_bins, p.n_params
    bufs = [hip.malloc(8 * nb) for _ in range(4)]
    for b in bufs:
        hip.fill(b, 0xAB, 8 * nb)
    rec = hip.malloc(n * mc.TRACE_DTYPE.itemsize)
    st = hip.stream()
    try:
        r = eng.forward_device(*bufs, n_histories=n, stream=st, trace_ptr=rec, trace=n)
        assert r["flights"] == host.flights
        got = [hip.download(b, np.empty(nb)) for b in bufs]
        # two runs add their histories' FP64 moments in different orders
        for g, want in zip(got, (host.sum, host.sq, host.mean, host.uncer)):
            np.testing.assert_allclose(g, want, rtol=1e-12, atol=1e-300)
        # the device's mean and uncertainty are the host formulas on its own sums, bit for bit
        dsum, dsq, dmean, duncer = got
        mean = dsum / n
        assert (dmean == mean).all()
        assert (duncer == np.sqrt((dsq / n - mean * mean) / (n - 1.0))).all()
        tr = hip.download(rec, np.empty(n, dtype=mc.TRACE_DTYPE))
        same_records(tr, host.trace, "device trace")
        # backward with device-side weights and outputs
        go = np.random.default_rng(9).normal(size=nb)
        hb = eng.backward(go, n_histories=n)
        gd = hip.upload(go)
        g, gs = hip.malloc(8 * npar), hip.malloc(8 * npar)
        eng.backward_device(gd, g, gs, n_histories=n, stream=st)
        np.testing.assert_allclose(hip.download(g, np.empty(npar)), hb.grad, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(hip.download(gs, np.empty(npar)), hb.grad_sigma, rtol=1e-8)
        for b in (gd, g, gs):
            hip.free(b)
    finally:
        hip.stream_destroy(st)
        for b in bufs + [rec]:
            hip.free(b)


def test_problem_from_flat_arrays_runs_on_the_gpu():
    n = 5000
    p = mc.Problem.from_string(decks.box_deck(n), shield_xs_dir())
    q = mc.Problem.from_desc(p.desc_ptr)
    del p                                  # the copy owns its arrays
    g = mc.Engine(q, bank_capacity=1 << 12).forward(trace=n)
    o = oracle_run(q, n, trace=n)
    same_records(g.trace, o.trace, "mc_problem_create")


def test_engine_outlives_its_problems():
    n = 3000
    p = mc.Problem.from_string(decks.slab_deck(2.0, 0.7, 4, n))
    o = oracle_run(p, n, trace=n)
    h = C.c_void_p()
    assert mc.lib.mc_engine_create((C.c_void_p * 1)(p._h.value), 1, None, C.byref(h)) == 0
    mc.lib.mc_problem_destroy(p._h)        # the engine copied what it needs
    p._h = None
    tr = np.zeros(n, dtype=mc.TRACE_DTYPE)
    run = mc.RunCfg()
    run.trace = tr.ctypes.data
    run.trace_count = n
    out = mc.TallyOut()
    assert mc.lib.mc_forward(h, C.byref(run), C.byref(out)) == 0, mc.lib.mc_last_error()
    mc.lib.mc_engine_destroy(h)
    same_records(tr, o.trace, "after mc_problem_destroy")
