"""The particle comb (population_control.cpp:55-84, applied after each
particle's walk, handler.cpp:24-28) in the oracle: its selection pinned by
known answers worked out by hand from the reference's loop, by the
systematic comb's unbiasedness (entry i is selected with probability
w_i / w_avg when no entry outweighs a tooth, and the combed bank keeps the
total weight), by planted mistakes each of those checks rejects, and by whole
runs of a splitting slab with the comb against the deterministic solution.
"""
import ctypes as C

import numpy as np
import pytest

import slab_analytic as sa
from mc_testlib import importance_slab, mc, oracle, oracle_run
from test_oracle_pinning import mutant

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


def bind(lib):
    lib.mco_comb_plan.argtypes = [DP, C.c_int, C.c_int, C.c_double, IP, DP, IP]
    lib.mco_comb_plan.restype = C.c_int
    return lib


bind(oracle)


def plan(lib, w, teeth, u):
    w = np.ascontiguousarray(w, dtype=np.float64)
    src = np.zeros(teeth, dtype=np.int32)
    wout = np.zeros(teeth)
    fill = np.zeros(teeth, dtype=np.int32)
    j = lib.mco_comb_plan(w.ctypes.data_as(DP), w.size, teeth, u, src.ctypes.data_as(IP),
                          wout.ctypes.data_as(DP), fill.ctypes.data_as(IP))
    return j, src, wout, fill


def test_comb_known_answers():
    # four equal weights combed to two teeth, first tooth at 0.25 * 2 = 0.5:
    # sums 1 > 0.5 (entry 0, next tooth 2.5), 2, 3 > 2.5 (entry 2, next 4.5), 4
    j, src, wout, fill = plan(oracle, [1.0, 1.0, 1.0, 1.0], 2, 0.25)
    assert j == 2 and src.tolist() == [0, 2] and wout.tolist() == [2.0, 2.0] and fill.tolist() == [0, 0]
    # an entry heavier than a tooth is selected once: weights 3, 1 to four
    # teeth (w_avg 1, first tooth 0.5) select entries 0 and 1 at weight 1; the
    # last two entries stay copies of entry 0 at its own weight 3
    # (bank_tmp(comb_teeth, Pbank[0]) in the reference)
    j, src, wout, fill = plan(oracle, [3.0, 1.0], 4, 0.5)
    assert j == 2 and src.tolist() == [0, 1, 0, 0] and wout.tolist() == [1.0, 1.0, 3.0, 3.0]
    assert fill.tolist() == [0, 0, 1, 2]
    # a first tooth past the first entry's weight skips it: w_avg 2, teeth at
    # 0.8 and 2.8; sums 0.5, 2.0 > 0.8 (entry 1), 3.0 > 2.8 (entry 2), 4.0
    j, src, wout, fill = plan(oracle, [0.5, 1.5, 1.0, 1.0], 2, 0.4)
    assert j == 2 and src.tolist() == [1, 2] and wout.tolist() == [2.0, 2.0]


def comb_failures(lib, trials=20000):
    """Unbiasedness of the systematic comb on weights none of which exceeds
    a tooth: selection frequency of entry i = w_i / w_avg, every comb keeps
    the total weight and selects exactly `teeth` entries in bank order."""
    rng = np.random.default_rng(11)
    w = rng.uniform(0.2, 1.0, 24)
    teeth = 12
    w_avg = w.sum() / teeth
    if w.max() > w_avg:
        raise AssertionError("test weights must not exceed a tooth")
    count = np.zeros(w.size)
    bad = []
    for u in (np.arange(trials) + 0.5) / trials:
        j, src, wout, fill = plan(lib, w, teeth, u)
        if j != teeth or fill.any():
            bad.append(f"u={u}: {j} selected")
            break
        if not np.all(np.diff(src) > 0):
            bad.append("selection out of bank order")
            break
        if abs(wout.sum() - w.sum()) > 1e-12 * w.sum():
            bad.append("total weight changed")
            break
        count[src] += 1
    p = count / trials
    if np.max(np.abs(p - w / w_avg)) > 2e-3:
        bad.append(f"selection frequency off by {np.max(np.abs(p - w / w_avg)):.3g}")
    return bad


def test_comb_is_unbiased():
    assert comb_failures(oracle) == []


@pytest.mark.parametrize("old,new", [
    # the tooth does not advance after a selection
    ("            tooth += w_avg;\n            ++j;", "            ++j;"),
    # the first tooth is u, not u * w_avg
    ("    double tooth = u * w_avg;", "    double tooth = u;"),
    # the selected copy keeps its own weight
    ("            wout[j] = w_avg;\n            fill[j] = 0;", "            wout[j] = w[i];\n            fill[j] = 0;"),
    # the average over the bank's size instead of the teeth
    ("    const double w_avg = W / (double)teeth;\n    double tooth = u * w_avg;",
     "    const double w_avg = W / (double)n;\n    double tooth = u * w_avg;"),
])
def test_comb_checks_reject_planted_mistakes(old, new):
    m = bind(mutant(old, new))
    try:
        bad = comb_failures(m, trials=4000)
    except Exception as e:          # a mistake that breaks the plan's shape fails too
        bad = [repr(e)]
    assert bad != []


def comb_slab(n, bank_max, teeth, L=4.0, c=0.5, nb=20, ratio=1.5):
    deck = importance_slab(L, c, nb, n, ratio)
    return deck.replace("</simulation>", "</simulation>\n<population_control>\n"
                        f'    <particle_comb bank_max="{bank_max}" teeth="{teeth}"/>\n</population_control>', 1)


def test_comb_loader_and_desc():
    p = mc.Problem.from_string(comb_slab(100, 8, 4))
    assert (p.head.comb_bank_max, p.head.comb_teeth) == (8, 4)
    q = mc.Problem.from_desc(p.desc_ptr)
    assert (q.head.comb_bank_max, q.head.comb_teeth) == (8, 4)
    for bad in ('bank_max="0" teeth="4"', 'bank_max="8" teeth="0"'):
        deck = comb_slab(100, 8, 4).replace('bank_max="8" teeth="4"', bad)
        with pytest.raises(mc.MCError) as e:
            mc.Problem.from_string(deck)
        assert e.value.status == mc.MC_ERR_INPUT


def run_vs_slab(deck, n):
    """z scores of T, R and the flux bins against the deterministic slab."""
    a = sa.slab(4.0, 0.5, n_bins=20)
    r = oracle_run(mc.Problem.from_string(deck), n)
    mean = r.sum / n
    sig = np.sqrt(np.maximum(r.sq / n - mean ** 2, 0.0) / (n - 1))
    z = np.concatenate([[(mean[0] - a["T"]) / sig[0], (mean[1] - a["R"]) / sig[1]], (mean[2:] - a["flux"]) / sig[2:]])
    return r, z


def test_comb_runs_change_histories_and_stay_unbiased():
    """Importance x1.5 per bin through a 4 cm c = 0.5 slab with a comb of one
    tooth once a history's bank holds 8 copies (a comb of one tooth keeps
    one copy, copy i with probability w_i / W, at weight W: no copy outweighs
    the tooth, so the comb is unbiased): fewer flights than without the comb,
    and T, R and the 20 flux bins within 3.5 sigma of the deterministic
    solution."""
    n = 40000
    plain = oracle_run(mc.Problem.from_string(importance_slab(4.0, 0.5, 20, n, 1.5)), n)
    r, z = run_vs_slab(comb_slab(n, 8, 1), n)
    assert r.flights < 0.8 * plain.flights
    assert np.abs(z).max() < 3.5, z


def test_comb_keeps_the_references_single_selection():
    """The reference keeps a copy at most once however many teeth its weight
    spans, and fills the rest of the combed bank with its first entry at that
    entry's own weight, which adds weight whenever a copy outweighs a tooth.
    With as many teeth as bank_max every comb keeps the bank's size, heavy
    copies are common, and the transmitted fraction comes out far above the
    deterministic one; a comb that kept such copies once per tooth would not
    (the oracle restates the reference, so this is its behaviour too)."""
    n = 20000
    _, z = run_vs_slab(comb_slab(n, 4, 4), n)
    assert z[0] > 10, z[:2]


def test_comb_refused_in_gradient_runs():
    p = mc.Problem.from_string(comb_slab(100, 8, 4))
    with pytest.raises(RuntimeError):
        oracle_run(p, 100, grad_out=np.ones(p.n_bins))
