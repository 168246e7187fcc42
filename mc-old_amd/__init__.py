"""mc-old_amd — MI355X-native fixed-source Monte Carlo transport (Python side).

Thin ctypes layer over the C ABI of libmcold.so (include/mc_api.h).  It
mirrors the reference's Simulator interface (include/simulator.h:102-150):

    sim = Simulator("examples/slab/")      # Simulator::Simulator(io_dir)
    sim.start()                            # Simulator::start()
    sim.Estimators[0].tally(0).mean        # Estimator::tally(i).mean / .uncer
    sim.report("examples/slab/")           # Simulator::report(io_dir)

and adds the engine-level calls (batched runs, likelihood-ratio gradients,
the standalone cross-section lookup) and a torch.autograd.Function.

The native library is the only compute path: importing this package on a
machine without it raises, and nothing here falls back to a CPU routine.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcold.so")

# ----------------------------------------------------------------- C structs
MC_OK, MC_ERR_ARGUMENT, MC_ERR_INPUT, MC_ERR_IO, MC_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
MC_ERR_DEVICE, MC_ERR_LOST_PARTICLE, MC_ERR_CAPACITY = 5, 6, 7
MC_RUN_BATCHED, MC_RUN_PLAIN = 0, 1


class EngineCfg(C.Structure):
    _fields_ = [("device", C.c_int32), ("block", C.c_int32), ("grid_per_cu", C.c_int32),
                ("iters_per_graph", C.c_int32), ("bank_capacity", C.c_uint64),
                ("use_graph", C.c_int32), ("profile", C.c_int32), ("drain_divisor", C.c_int32),
                ("events_per_launch", C.c_int32), ("grad_events_per_launch", C.c_int32),
                ("grad_live_divisor", C.c_int32)]


class TraceRecord(C.Structure):
    _fields_ = [("rng", C.c_uint64), ("flights", C.c_uint64), ("x", C.c_double), ("y", C.c_double),
                ("z", C.c_double), ("E", C.c_double), ("score", C.c_double)]


TRACE_DTYPE = np.dtype([("rng", "<u8"), ("flights", "<u8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
                        ("E", "<f8"), ("score", "<f8")])


class RunCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("device_ptrs", C.c_int32), ("n_histories", C.POINTER(C.c_uint64)),
                ("theta", C.POINTER(C.c_double)), ("trace", C.c_void_p), ("trace_count", C.c_uint64),
                ("key_offset", C.POINTER(C.c_uint64)), ("stream", C.c_void_p)]


class TallyOut(C.Structure):
    _fields_ = [("sum", C.POINTER(C.c_double)), ("sq", C.POINTER(C.c_double)),
                ("mean", C.POINTER(C.c_double)), ("uncer", C.POINTER(C.c_double)),
                ("flights", C.c_uint64), ("seconds", C.c_double)]


KERNEL_NAMES = ("k_refill_count", "k_refill_scatter", "k_xs_lookup", "k_flight", "k_collision", "k_drain")


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64 * len(KERNEL_NAMES)), ("items", C.c_uint64 * len(KERNEL_NAMES)),
                ("seconds", C.c_double * len(KERNEL_NAMES))]


class EngineInfo(C.Structure):
    _fields_ = [("n_problems", C.c_int32), ("n_bins", C.c_int32), ("n_params", C.c_int32),
                ("pad", C.c_int32), ("bank_capacity", C.c_uint64), ("block", C.c_int32),
                ("grid_per_cu", C.c_int32), ("iters_per_graph", C.c_int32), ("n_cu", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid_per_cu", C.c_int32), ("iters_per_graph", C.c_int32),
                ("drain_divisor", C.c_int32), ("bank_capacity", C.c_uint64), ("histories_per_second", C.c_double),
                ("events_per_launch", C.c_int32), ("pad", C.c_int32)]


class EstimatorDesc(C.Structure):
    _fields_ = [("attach_kind", C.c_int32), ("n_attach", C.c_int32), ("n_scores", C.c_int32),
                ("score", C.c_int32 * 8), ("score_kernel", C.c_int32 * 8), ("egrid_begin", C.c_int32),
                ("egrid_count", C.c_int32), ("bin_offset", C.c_int32), ("n_bins", C.c_int32)]


class ProblemDesc(C.Structure):
    """mc_problem_desc (array members as raw pointers)."""
    _fields_ = [("n_surfaces", C.c_int32), ("n_cells", C.c_int32), ("n_materials", C.c_int32),
                ("n_nuclides", C.c_int32), ("n_sources", C.c_int32), ("n_estimators", C.c_int32),
                ("n_params", C.c_int32), ("n_bins", C.c_int32)] + [
        (n, C.c_void_p) for n in ("surfaces", "cells", "cell_surf", "cell_sense", "materials", "mat_nuclide",
                                  "mat_density", "nuclides", "xs_energy", "xs_scatter", "xs_capture",
                                  "xs_total", "sources", "source_cdf", "estimators", "att_estimator",
                                  "att_index", "egrid", "params", "mat_param")] + [
        ("wr", C.c_double), ("ws", C.c_double), ("n_histories", C.c_uint64), ("seed", C.c_uint64),
        ("comb_bank_max", C.c_int32), ("comb_teeth", C.c_int32)]


EXPORTS = {
    "mc_last_error": (C.c_char_p, []),
    "mc_version": (C.c_char_p, []),
    "mc_problem_from_xml": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "mc_problem_from_xml_string": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "mc_problem_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mc_problem_get_desc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "mc_problem_estimator_name": (C.c_char_p, [C.c_void_p, C.c_int32]),
    "mc_problem_object_name": (C.c_char_p, [C.c_void_p, C.c_int32, C.c_int32]),
    "mc_problem_destroy": (None, [C.c_void_p]),
    "mc_engine_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(EngineCfg), C.POINTER(C.c_void_p)]),
    "mc_engine_info_get": (C.c_int, [C.c_void_p, C.POINTER(EngineInfo)]),
    "mc_engine_bin_offset": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mc_engine_param_offset": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "mc_engine_kernel_stats": (C.c_int, [C.c_void_p, C.POINTER(KernelStats), C.c_int32]),
    "mc_engine_copy_ring": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]),
    "mc_forward": (C.c_int, [C.c_void_p, C.POINTER(RunCfg), C.POINTER(TallyOut)]),
    "mc_backward": (C.c_int, [C.c_void_p, C.POINTER(RunCfg), C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double), C.POINTER(TallyOut)]),
    "mc_xs_lookup": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_double),
                               C.c_uint64]),
    "mc_xs_lookup_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double), C.c_uint64]),
    "mc_xs_lookup_bench": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_uint64, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "mc_gather_bench": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_uint64, C.c_uint64, C.c_int32,
                                  C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "mc_selftest_math": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64]),
    "mc_autotune": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(TuneResult)]),
    "mc_engine_destroy": (None, [C.c_void_p]),
    "mc_report_text": (C.c_int, [C.c_void_p, C.POINTER(TallyOut), C.c_uint64, C.c_char_p]),
    "mc_report_h5": (C.c_int, [C.c_void_p, C.POINTER(TallyOut), C.c_uint64, C.c_char_p]),
}


def load_library(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python __graft_entry__.py` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = load_library()


class MCError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


def _check(status: int) -> None:
    if status != MC_OK:
        raise MCError(status, (lib.mc_last_error() or b"").decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# -------------------------------------------------------------------- problem
class Problem:
    """A deck in the reference schema (Simulator::Simulator, setup.cpp:30-1069)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        d = C.c_void_p()
        _check(lib.mc_problem_get_desc(self._h, C.byref(d)))
        self.desc_ptr = d.value
        self.head = ProblemDesc.from_address(d.value)

    @classmethod
    def from_xml(cls, io_dir: str, xs_dir: str | None = None) -> "Problem":
        h = C.c_void_p()
        _check(lib.mc_problem_from_xml(io_dir.encode(), xs_dir.encode() if xs_dir else None, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_desc(cls, desc_ptr: int) -> "Problem":
        """mc_problem_create: a problem from flat arrays (an mc_problem_desc at
        desc_ptr, e.g. another problem's desc_ptr), copied."""
        h = C.c_void_p()
        _check(lib.mc_problem_create(C.c_void_p(desc_ptr), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_string(cls, xml: str, xs_dir: str | None = None) -> "Problem":
        h = C.c_void_p()
        _check(lib.mc_problem_from_xml_string(xml.encode(), xs_dir.encode() if xs_dir else None, C.byref(h)))
        return cls(h.value)

    @property
    def n_bins(self) -> int:
        return self.head.n_bins

    @property
    def n_params(self) -> int:
        return self.head.n_params

    @property
    def n_histories(self) -> int:
        return int(self.head.n_histories)

    def estimators(self) -> list[EstimatorDesc]:
        arr = (EstimatorDesc * self.head.n_estimators).from_address(self.head.estimators)
        return list(arr)

    def estimator_names(self) -> list[str]:
        out, e = [], 0
        while True:
            n = lib.mc_problem_estimator_name(self._h, e)
            if n is None:
                return out
            out.append(n.decode())
            e += 1

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "_h", None) is not None and self._h.value and lib is not None:
            lib.mc_problem_destroy(self._h)
            self._h = None


# --------------------------------------------------------------------- engine
@dataclass
class TallyResult:
    mean: np.ndarray
    uncer: np.ndarray
    sum: np.ndarray
    sq: np.ndarray
    flights: int
    seconds: float
    n_histories: np.ndarray
    trace: np.ndarray | None = None
    grad: np.ndarray | None = None
    grad_sigma: np.ndarray | None = None


class Engine:
    """Device state for one or many problems (mc_engine_create)."""

    def __init__(self, problems, device: int = 0, block: int = 0, grid_per_cu: int = 0,
                 iters_per_graph: int = 0, bank_capacity: int = 0, use_graph: bool = True,
                 profile: bool = False, drain_divisor: int = 0, events_per_launch: int = 0,
                 grad_events_per_launch: int = 0, grad_live_divisor: int = 0):
        if isinstance(problems, Problem):
            problems = [problems]
        self.problems = list(problems)
        arr = (C.c_void_p * len(self.problems))(*[p._h.value for p in self.problems])
        cfg = EngineCfg(device, block, grid_per_cu, iters_per_graph, bank_capacity, int(use_graph), int(profile),
                        int(drain_divisor), int(events_per_launch), int(grad_events_per_launch),
                        int(grad_live_divisor))
        h = C.c_void_p()
        _check(lib.mc_engine_create(arr, len(self.problems), C.byref(cfg), C.byref(h)))
        self._h = h
        info = EngineInfo()
        _check(lib.mc_engine_info_get(self._h, C.byref(info)))
        self.info = info
        self.n_bins = info.n_bins
        self.n_params = info.n_params

    def bin_range(self, p: int) -> slice:
        off, n = C.c_int32(), C.c_int32()
        _check(lib.mc_engine_bin_offset(self._h, p, C.byref(off), C.byref(n)))
        return slice(off.value, off.value + n.value)

    def param_range(self, p: int) -> slice:
        off, n = C.c_int32(), C.c_int32()
        _check(lib.mc_engine_param_offset(self._h, p, C.byref(off), C.byref(n)))
        return slice(off.value, off.value + n.value)

    def _run_cfg(self, n_histories, theta, trace, mode, key_offset=None, stream=None):
        run = RunCfg()
        run.mode = mode
        keep = []
        if key_offset is not None:
            ko = np.ascontiguousarray(np.broadcast_to(np.asarray(key_offset, dtype=np.uint64),
                                                      (len(self.problems),)))
            keep.append(ko)
            run.key_offset = ko.ctypes.data_as(C.POINTER(C.c_uint64))
        if stream is not None:
            run.stream = C.c_void_p(int(stream))
        if n_histories is not None:
            nh = np.ascontiguousarray(np.broadcast_to(np.asarray(n_histories, dtype=np.uint64),
                                                      (len(self.problems),)))
            keep.append(nh)
            run.n_histories = nh.ctypes.data_as(C.POINTER(C.c_uint64))
        else:
            nh = np.array([p.n_histories for p in self.problems], dtype=np.uint64)
        if theta is not None:
            th = np.ascontiguousarray(theta, dtype=np.float64)
            assert th.shape == (self.n_params,)
            keep.append(th)
            run.theta = _dptr(th)
        tr = None
        if trace:
            tr = np.zeros(trace * len(self.problems), dtype=TRACE_DTYPE)
            run.trace = tr.ctypes.data
            run.trace_count = trace
        return run, keep, nh, tr

    def forward(self, n_histories=None, theta=None, trace: int = 0, mode: int = MC_RUN_BATCHED,
                key_offset=None, stream=None) -> TallyResult:
        """key_offset: per problem (or one for all) first history key of the run;
        stream: a hipStream_t handle (int) to issue the run on."""
        run, keep, nh, tr = self._run_cfg(n_histories, theta, trace, mode, key_offset, stream)
        nb = max(1, self.n_bins)
        s, q, m, u = (np.zeros(nb) for _ in range(4))
        out = TallyOut(_dptr(s), _dptr(q), _dptr(m), _dptr(u), 0, 0.0)
        _check(lib.mc_forward(self._h, C.byref(run), C.byref(out)))
        return TallyResult(m[:self.n_bins], u[:self.n_bins], s[:self.n_bins], q[:self.n_bins],
                           out.flights, out.seconds, nh, tr)

    def backward(self, grad_out, n_histories=None, theta=None, mode: int = MC_RUN_BATCHED,
                 key_offset=None, stream=None) -> TallyResult:
        run, keep, nh, tr = self._run_cfg(n_histories, theta, 0, mode, key_offset, stream)
        go = np.ascontiguousarray(grad_out, dtype=np.float64)
        assert go.shape == (self.n_bins,)
        nb = max(1, self.n_bins)
        s, q, m, u = (np.zeros(nb) for _ in range(4))
        g = np.zeros(max(1, self.n_params))
        gs = np.zeros(max(1, self.n_params))
        out = TallyOut(_dptr(s), _dptr(q), _dptr(m), _dptr(u), 0, 0.0)
        _check(lib.mc_backward(self._h, C.byref(run), _dptr(go), _dptr(g), _dptr(gs), C.byref(out)))
        return TallyResult(m[:self.n_bins], u[:self.n_bins], s[:self.n_bins], q[:self.n_bins],
                           out.flights, out.seconds, nh, None, g[:self.n_params], gs[:self.n_params])

    def forward_device(self, sum_ptr: int, sq_ptr: int, mean_ptr: int, uncer_ptr: int, n_histories=None,
                       theta=None, key_offset=None, stream=None, trace_ptr: int = 0, trace: int = 0) -> dict:
        """mc_forward with device-pointer outputs (run.device_ptrs = 1): the
        moment arrays (engine n_bins doubles each, 0 to skip one) and the trace
        records (trace per problem) are written on the device."""
        run, keep, nh, _ = self._run_cfg(n_histories, theta, 0, MC_RUN_BATCHED, key_offset, stream)
        run.device_ptrs = 1
        if trace:
            run.trace = C.c_void_p(trace_ptr)
            run.trace_count = trace
        dp = C.POINTER(C.c_double)
        out = TallyOut(C.cast(C.c_void_p(sum_ptr), dp), C.cast(C.c_void_p(sq_ptr), dp),
                       C.cast(C.c_void_p(mean_ptr), dp), C.cast(C.c_void_p(uncer_ptr), dp), 0, 0.0)
        _check(lib.mc_forward(self._h, C.byref(run), C.byref(out)))
        return {"flights": out.flights, "seconds": out.seconds}

    def backward_device(self, grad_out_ptr: int, grad_ptr: int, grad_sigma_ptr: int, n_histories=None,
                        theta=None, key_offset=None, stream=None) -> dict:
        """mc_backward with device-pointer weights and gradient outputs."""
        run, keep, nh, _ = self._run_cfg(n_histories, theta, 0, MC_RUN_BATCHED, key_offset, stream)
        run.device_ptrs = 1
        out = TallyOut()
        dp = C.POINTER(C.c_double)
        _check(lib.mc_backward(self._h, C.byref(run), C.cast(C.c_void_p(grad_out_ptr), dp),
                               C.cast(C.c_void_p(grad_ptr), dp), C.cast(C.c_void_p(grad_sigma_ptr), dp),
                               C.byref(out)))
        return {"flights": out.flights, "seconds": out.seconds}

    def kernel_stats(self, reset: bool = True) -> dict:
        """Per-kernel device seconds, launches and particles processed
        (recorded when the engine was created with profile=True)."""
        k = KernelStats()
        _check(lib.mc_engine_kernel_stats(self._h, C.byref(k), int(reset)))
        return {n: dict(launches=k.launches[i], items=k.items[i], seconds=k.seconds[i])
                for i, n in enumerate(KERNEL_NAMES)}

    def copy_ring(self) -> dict:
        """The split-copy ring: capacity (entries), growths (doublings since
        the engine was made) and waits (the last run's splits and turns that
        found it without room)."""
        cap, gr, wt = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib.mc_engine_copy_ring(self._h, C.byref(cap), C.byref(gr), C.byref(wt)))
        return {"capacity": cap.value, "growths": gr.value, "waits": wt.value}

    def xs_lookup(self, energy, mat) -> np.ndarray:
        E = np.ascontiguousarray(energy, dtype=np.float64)
        M = np.ascontiguousarray(mat, dtype=np.int32)
        out = np.zeros_like(E)
        _check(lib.mc_xs_lookup(self._h, _dptr(E), M.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(out), E.size))
        return out

    def xs_bench(self, n: int, iters: int = 20, seed: int = 12345) -> tuple[float, float]:
        sec, chk = C.c_double(), C.c_double()
        _check(lib.mc_xs_lookup_bench(self._h, n, iters, seed, C.byref(sec), C.byref(chk)))
        return sec.value, chk.value

    def gather_bench(self, n: int, dependent_reads: int = 1, table_bytes: int = 0, iters: int = 20,
                     seed: int = 12345) -> tuple[float, int]:
        """(seconds per launch, lookup table bytes) of the random-gather
        ceiling kernel (mc_gather_bench)."""
        sec, tb = C.c_double(), C.c_uint64()
        _check(lib.mc_gather_bench(self._h, n, iters, seed, table_bytes, dependent_reads, C.byref(sec), C.byref(tb)))
        return sec.value, tb.value

    def autotune(self, sample_histories: int) -> TuneResult:
        r = TuneResult()
        _check(lib.mc_autotune(self._h, sample_histories, C.byref(r)))
        return r

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value and lib is not None:
            lib.mc_engine_destroy(self._h)
            self._h = None


def selftest_math(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(4 * x.size)
    _check(lib.mc_selftest_math(_dptr(x), _dptr(out), x.size))
    return out.reshape(-1, 4)


# ------------------------------------------------ reference-interface mirror
@dataclass
class Tally:
    """Estimator.h:257-269 (mean and uncertainty after end_simulation)."""
    mean: float
    uncer: float


class Estimator:
    def __init__(self, name: str, mean: np.ndarray, uncer: np.ndarray):
        self.name = name
        self._mean, self._uncer = mean, uncer

    def tally(self, i: int) -> Tally:
        return Tally(float(self._mean[i]), float(self._uncer[i]))

    def tally_size(self) -> int:
        return int(self._mean.size)


class Simulator:
    """Simulator(io_dir) / start() / report(io_dir), include/simulator.h:144-149."""

    def __init__(self, io_dir: str, xs_dir: str | None = None, **engine_kw):
        self.io_dir = io_dir
        self.problem = Problem.from_xml(io_dir, xs_dir)
        self.engine = Engine(self.problem, **engine_kw)
        self.Estimators: list[Estimator] = []
        self.Nsample = self.problem.n_histories
        self.Ntrack = 0
        self._result: TallyResult | None = None

    def start(self) -> None:
        r = self.engine.forward()
        self._result = r
        self.Ntrack = r.flights
        self.Estimators = []
        off = 0
        names = self.problem.estimator_names()
        sizes = _estimator_sizes(self.problem)
        for name, n in zip(names, sizes):
            self.Estimators.append(Estimator(name, r.mean[off:off + n], r.uncer[off:off + n]))
            off += n

    def report(self, io_dir: str | None = None) -> str:
        """<io_dir>/output.h5 in the reference's layout (report.cpp:9-50), or
        output.txt with the same content where no HDF5 library loads."""
        if self._result is None:
            raise RuntimeError("start() has not run")
        d = io_dir or self.io_dir
        r = self._result
        out = TallyOut(_dptr(r.sum), _dptr(r.sq), _dptr(r.mean), _dptr(r.uncer), r.flights, r.seconds)
        path = os.path.join(d, "output.h5")
        st = lib.mc_report_h5(self.problem._h, C.byref(out), int(self.Nsample), path.encode())
        if st == MC_ERR_UNSUPPORTED:
            path = os.path.join(d, "output.txt")
            st = lib.mc_report_text(self.problem._h, C.byref(out), int(self.Nsample), path.encode())
        _check(st)
        return path


def _estimator_sizes(p: Problem) -> list[int]:
    return [e.n_bins for e in p.estimators()]


# ------------------------------------------------------ torch.autograd bridge
def transport_function():
    """torch.autograd.Function whose forward runs mc_forward at theta and whose
    backward runs mc_backward with the incoming grad_output (likelihood-ratio
    vector-Jacobian product over the same history keys)."""
    import torch

    class TransportFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, engine, n_histories=None):
            th = theta.detach().cpu().double().numpy()
            r = engine.forward(n_histories=n_histories, theta=th)
            ctx.engine, ctx.n_histories, ctx.theta = engine, n_histories, th
            return torch.from_numpy(r.mean.copy()).to(theta.device)

        @staticmethod
        def backward(ctx, grad_output):
            go = grad_output.detach().cpu().double().numpy()
            r = ctx.engine.backward(go, n_histories=ctx.n_histories, theta=ctx.theta)
            g = torch.from_numpy(r.grad.copy()).to(grad_output.device)
            return g, None, None

    return TransportFunction
