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

from mc_testlib import decks, mc, oracle_run, shield_xs_dir

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
    """Outputs written to torch device tensors by a run issued on a torch
    stream: the same values, bit for bit, as the host-pointer run; gradient
    weights read from and gradients written to device memory likewise."""
    torch = pytest.importorskip("torch")
    n = 8000
    p = mc.Problem.from_string(decks.shield_deck(n, params="24"), shield_xs_dir())
    eng = mc.Engine(p, bank_capacity=1 << 12)
    host = eng.forward(n_histories=n, trace=n)
    nb = p.n_bins
    dev = torch.device("cuda:0")
    s, q, m, u = (torch.full((nb,), -7.0, dtype=torch.float64, device=dev) for _ in range(4))
    rec = torch.zeros(n * mc.TRACE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    st = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()               # the tensors' fills are done before the run writes them
    r = eng.forward_device(s.data_ptr(), q.data_ptr(), m.data_ptr(), u.data_ptr(), n_histories=n,
                           stream=st.cuda_stream, trace_ptr=rec.data_ptr(), trace=n)
    torch.cuda.synchronize()
    assert r["flights"] == host.flights
    assert (s.cpu().numpy() == host.sum).all() and (q.cpu().numpy() == host.sq).all()
    assert (m.cpu().numpy() == host.mean).all() and (u.cpu().numpy() == host.uncer).all()
    same_records(np.frombuffer(rec.cpu().numpy().tobytes(), dtype=mc.TRACE_DTYPE), host.trace, "device trace")
    # backward with device-side weights and outputs
    go = np.random.default_rng(9).normal(size=nb)
    hb = eng.backward(go, n_histories=n)
    gd = torch.from_numpy(go).to(dev)
    g, gs = (torch.zeros(p.n_params, dtype=torch.float64, device=dev) for _ in range(2))
    torch.cuda.synchronize()
    eng.backward_device(gd.data_ptr(), g.data_ptr(), gs.data_ptr(), n_histories=n, stream=st.cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_allclose(g.cpu().numpy(), hb.grad, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(gs.cpu().numpy(), hb.grad_sigma, rtol=1e-8)


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
