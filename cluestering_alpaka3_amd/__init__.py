"""MI355X-native CLUE clustering engine: Python host binding.

The engine is libclue_hip.so (hand-written gfx950 HIP kernels behind the C ABI
declared in include/clue_abi.h).  This module binds that ABI with ctypes; it
allocates device memory and streams through the ABI itself, so no deep-learning
framework is involved.  There is no CPU fallback: if the library is missing the
import of the engine fails loudly.

`main_run` mirrors the reference's Python FFI entry `mainRun`
(CLUEstering/BindingModules/serial/binding_cpu.cpp:35-52, called from
CLUEstering/CLUEstering.py:668-707).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclue_hip.so")
MAX_DIM = 10

KERNEL_FLAT, KERNEL_GAUSSIAN, KERNEL_EXPONENTIAL = 0, 1, 2
METRIC_EUCLIDEAN, METRIC_WEIGHTED_EUCLIDEAN, METRIC_PERIODIC_EUCLIDEAN = 0, 1, 2
METRIC_MANHATTAN, METRIC_CHEBYSHEV, METRIC_WEIGHTED_CHEBYSHEV = 3, 4, 5
SEEDS_REFERENCE, SEEDS_UNBOUNDED = 0, 1

STATUS = {0: "CLUE_OK", 1: "CLUE_ERR_INVALID_ARGUMENT", 2: "CLUE_ERR_OUT_OF_RANGE",
          3: "CLUE_ERR_RUNTIME", 4: "CLUE_ERR_OUT_OF_MEMORY", 5: "CLUE_ERR_UNSUPPORTED"}


class ClueError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class InvalidArgument(ClueError, ValueError):
    """CLUE_ERR_INVALID_ARGUMENT (std::invalid_argument in the reference)."""


class OutOfRange(ClueError, IndexError):
    """CLUE_ERR_OUT_OF_RANGE (std::out_of_range in the reference)."""


class ClueParams(ctypes.Structure):
    _fields_ = [("dc", ctypes.c_float), ("rhoc", ctypes.c_float), ("dm", ctypes.c_float),
                ("seed_dc", ctypes.c_float), ("points_per_tile", ctypes.c_int32),
                ("seed_mode", ctypes.c_int32)]


class ClueKernelDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("params", ctypes.c_float * 3)]


class ClueMetricDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("params", ctypes.c_float * MAX_DIM)]


class ClueStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("seed_candidates", "seed_capacity", "n_seeds", "n_clusters", "n_dropped_seeds")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class ClueTuning(ctypes.Structure):
    _fields_ = [("nh_first_radius", ctypes.c_float), ("morton_bits", ctypes.c_int32),
                ("xcd_remap", ctypes.c_int32), ("pending_blocks", ctypes.c_int32),
                ("pending_group", ctypes.c_int32), ("density_order", ctypes.c_int32),
                ("small_call_points", ctypes.c_int32), ("wave_dims", ctypes.c_int32),
                ("tile_call_points", ctypes.c_int32), ("density_filter", ctypes.c_int32),
                ("fused_event_points", ctypes.c_int32), ("split_call_points", ctypes.c_int32),
                ("nh_lone_lanes", ctypes.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ClueGridInfo(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("npoints", ctypes.c_int32), ("npd", ctypes.c_int32),
                ("ntiles", ctypes.c_int32), ("minmax", ctypes.c_float * (2 * MAX_DIM)),
                ("tile_size", ctypes.c_float * MAX_DIM)]


class ClueServerConfig(ctypes.Structure):
    _fields_ = [("device_budget_bytes", ctypes.c_size_t), ("slots", ctypes.c_int32),
                ("max_batch_points", ctypes.c_int32), ("max_batch_events", ctypes.c_int32),
                ("max_dim", ctypes.c_int32), ("small_batch_points", ctypes.c_int32)]


class ClueRequest(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("n", ctypes.c_int32), ("coords", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("params", ClueParams), ("kernel", ClueKernelDesc),
                ("metric", ClueMetricDesc), ("wrapped", ctypes.c_uint8 * MAX_DIM),
                ("cluster_out", ctypes.c_void_p), ("is_seed_out", ctypes.c_void_p), ("rho_out", ctypes.c_void_p),
                ("nearest_higher_out", ctypes.c_void_p)]


class ClueRequestTiming(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_int64), ("launch_ns", ctypes.c_int64), ("done_ns", ctypes.c_int64),
                ("batch_events", ctypes.c_int32)]


class ClueServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("events_submitted", "events_done", "points_done", "batches",
                                              "max_batch_events", "max_batch_points", "slot_capacity_points",
                                              "device_bytes")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# every symbol include/clue_abi.h declares, with its ctypes signature
_P = ctypes.POINTER
_vp = ctypes.c_void_p
_SIGNATURES = {
    "clue_last_error_message": ([], ctypes.c_char_p),
    "clue_version": ([], ctypes.c_char_p),
    "clue_device_count": ([_P(ctypes.c_int32)], ctypes.c_int),
    "clue_set_device": ([ctypes.c_int32], ctypes.c_int),
    "clue_stream_create": ([ctypes.c_int32, _P(_vp)], ctypes.c_int),
    "clue_stream_destroy": ([_vp], ctypes.c_int),
    "clue_stream_synchronize": ([_vp], ctypes.c_int),
    "clue_malloc_device": ([ctypes.c_int32, ctypes.c_size_t, _P(_vp)], ctypes.c_int),
    "clue_free_device": ([_vp], ctypes.c_int),
    "clue_malloc_host": ([ctypes.c_size_t, _P(_vp)], ctypes.c_int),
    "clue_free_host": ([_vp], ctypes.c_int),
    "clue_memcpy_h2d": ([_vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "clue_memcpy_d2h": ([_vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "clue_memcpy_d2d": ([_vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "clue_memset_device": ([_vp, ctypes.c_int32, ctypes.c_size_t, _vp], ctypes.c_int),
    "clue_plan_create": ([ctypes.c_int32, _P(_vp)], ctypes.c_int),
    "clue_plan_destroy": ([_vp], ctypes.c_int),
    "clue_plan_reset_capacity": ([_vp], ctypes.c_int),
    "clue_plan_reserve": ([_vp, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
    "clue_plan_device_bytes": ([_vp, _P(ctypes.c_size_t)], ctypes.c_int),
    "clue_make_clusters": ([_vp, _vp, _P(ClueParams), ctypes.c_int32, ctypes.c_int32, _P(_vp), _vp,
                            _P(ClueKernelDesc), _P(ClueMetricDesc), _P(ctypes.c_uint8), _vp, _vp,
                            _vp, _vp], ctypes.c_int),
    "clue_make_clusters_noalloc": ([ctypes.c_int32, _vp, _P(ClueParams), ctypes.c_int32, ctypes.c_int32, _P(_vp),
                                    _vp, _P(ClueKernelDesc), _P(ClueMetricDesc), _P(ctypes.c_uint8), _vp, _vp,
                                    _vp, _vp], ctypes.c_int),
    "clue_make_clusters_batched": ([_vp, _vp, _P(ClueParams), ctypes.c_int32, ctypes.c_int32,
                                    _P(ctypes.c_int64), _P(_vp), _vp, _P(ClueKernelDesc),
                                    _P(ClueMetricDesc), _P(ctypes.c_uint8), _vp, _vp, _vp, _vp], ctypes.c_int),
    "clue_batch_graph_create": ([ctypes.c_int32, _P(ClueParams), ctypes.c_int32, ctypes.c_int32,
                                 _P(ctypes.c_int64), _P(_vp), _vp, _P(ClueKernelDesc), _P(ClueMetricDesc),
                                 _P(ctypes.c_uint8), _vp, _vp, _vp, _vp, _P(ClueTuning), _P(_vp)], ctypes.c_int),
    "clue_batch_graph_launch": ([_vp, _vp], ctypes.c_int),
    "clue_batch_graph_destroy": ([_vp], ctypes.c_int),
    "clue_autotune_batched": ([_vp, _vp, _P(ClueParams), ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_int64),
                               _P(_vp), _vp, _P(ClueKernelDesc), _P(ClueMetricDesc), _P(ctypes.c_uint8),
                               ctypes.c_int32, _P(ClueTuning), _P(ctypes.c_float)], ctypes.c_int),
    "clue_autotune_candidates": ([_P(ctypes.c_int32), ctypes.c_int32, _P(ctypes.c_int32)], ctypes.c_int),
    "clue_server_create": ([ctypes.c_int32, _P(ClueServerConfig), _P(_vp)], ctypes.c_int),
    "clue_server_submit": ([_vp, _P(ClueRequest), _P(ctypes.c_uint64)], ctypes.c_int),
    "clue_server_wait": ([_vp, ctypes.c_uint64, _P(ClueRequestTiming)], ctypes.c_int),
    "clue_server_get_stats": ([_vp, _P(ClueServerStats)], ctypes.c_int),
    "clue_server_destroy": ([_vp], ctypes.c_int),
    "clue_cluster_map": ([_vp, _vp, ctypes.c_int32, _vp, _vp, _vp, _vp], ctypes.c_int),
    "clue_cluster_centroids": ([_vp, _vp, ctypes.c_int32, _P(_vp), _vp, _vp, _vp, ctypes.c_int32, _vp],
                               ctypes.c_int),
    "clue_selftest_sqrt": ([ctypes.c_uint32, ctypes.c_uint64, _P(ctypes.c_uint64)], ctypes.c_int),
    "clue_main_run": ([ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                       _P(ctypes.c_uint8), _P(ctypes.c_float), _P(ctypes.c_int32), _P(ClueKernelDesc),
                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
    "clue_plan_last_stats": ([_vp, _P(ClueStats)], ctypes.c_int),
    "clue_plan_last_grid": ([_vp, _P(ClueGridInfo)], ctypes.c_int),
    "clue_plan_copy_tiles": ([_vp, _P(ctypes.c_int32), _P(ctypes.c_int32)], ctypes.c_int),
    "clue_plan_copy_delta": ([_vp, _P(ctypes.c_float)], ctypes.c_int),
    "clue_plan_set_tuning": ([_vp, _P(ClueTuning)], ctypes.c_int),
    "clue_plan_get_tuning": ([_vp, _P(ClueTuning)], ctypes.c_int),
    "clue_plan_set_timing": ([_vp, ctypes.c_int32], ctypes.c_int),
    "clue_plan_stage_times": ([_vp, _P(ctypes.c_float), ctypes.c_int32, _P(ctypes.c_int32)], ctypes.c_int),
    "clue_stage_name": ([ctypes.c_int32], ctypes.c_char_p),
    "clue_selftest_math": ([ctypes.c_int32, _vp, _vp, _vp, ctypes.c_int64, _vp], ctypes.c_int),
}

_lib = None


def lib():
    """Load libclue_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def autotune_candidates():
    """The points-per-tile values the autotuner tries."""
    cnt = ctypes.c_int32(0)
    check(lib().clue_autotune_candidates(None, 0, ctypes.byref(cnt)))
    out = (ctypes.c_int32 * cnt.value)()
    check(lib().clue_autotune_candidates(out, cnt.value, ctypes.byref(cnt)))
    return tuple(int(v) for v in out)


def exported_symbols():
    return list(_SIGNATURES)


def check(status):
    if status == 0:
        return
    msg = lib().clue_last_error_message().decode(errors="replace")
    if status == 1:
        raise InvalidArgument(status, msg)
    if status == 2:
        raise OutOfRange(status, msg)
    raise ClueError(status, msg)


def device_count():
    c = ctypes.c_int32(0)
    st = lib().clue_device_count(ctypes.byref(c))
    return c.value if st == 0 else 0


class Stream:
    """A HIP stream on one device (the reference's queue)."""

    def __init__(self, device=0):
        self.device = device
        h = _vp()
        check(lib().clue_stream_create(device, ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib().clue_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib().clue_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A flat device buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib().clue_malloc_device(device, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, device=0, stream=None):
        a = np.ascontiguousarray(arr)
        d = cls(a.shape, a.dtype, device)
        d.upload(a, stream)
        return d

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().clue_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes, stream.handle if stream else None))
        if stream is None:
            check(lib().clue_stream_synchronize(None))

    def download(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(lib().clue_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream.handle if stream else None))
        check(lib().clue_stream_synchronize(stream.handle if stream else None))
        return out

    def offset(self, index):
        """Device address of element `index` of the flattened array."""
        return self.ptr + int(index) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            lib().clue_free_device(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def kernel_desc(kernel):
    """kernel: ("flat", h) | ("gaussian", mean, std, amplitude) | ("exponential", avg, amplitude)."""
    k = ClueKernelDesc()
    k.type = {"flat": 0, "gaussian": 1, "exponential": 2}[kernel[0]]
    for i, v in enumerate(kernel[1:]):
        k.params[i] = v
    return k


def metric_desc(metric="euclidean", params=None):
    m = ClueMetricDesc()
    m.type = {"euclidean": 0, "weighted_euclidean": 1, "periodic_euclidean": 2, "manhattan": 3,
              "chebyshev": 4, "weighted_chebyshev": 5}[metric]
    if params is not None:
        for i, v in enumerate(params):
            m.params[i] = v
    return m


@dataclass
class Result:
    cluster: np.ndarray
    is_seed: np.ndarray
    rho: np.ndarray
    nearest_higher: np.ndarray


class Plan:
    """Persistent clustering state on one device (the Clusterer's scratch)."""

    def __init__(self, device=0):
        self.device = device
        h = _vp()
        check(lib().clue_plan_create(device, ctypes.byref(h)))
        self.handle = h.value

    def reserve(self, ndim, n):
        check(lib().clue_plan_reserve(self.handle, ndim, n))

    def reset_capacity(self):
        check(lib().clue_plan_reset_capacity(self.handle))

    def device_bytes(self):
        b = ctypes.c_size_t(0)
        check(lib().clue_plan_device_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def make_clusters_device(self, stream, params, ndim, n, coord_ptrs, weight_ptr, kernel, metric,
                             wrapped, cluster_ptr, is_seed_ptr=None, rho_ptr=None, nh_ptr=None):
        """Enqueue one clustering of device-resident points (no synchronisation)."""
        cp = (_vp * MAX_DIM)(*([int(p) for p in coord_ptrs] + [0] * (MAX_DIM - len(coord_ptrs))))
        wr = None
        if wrapped is not None:
            wr = (ctypes.c_uint8 * MAX_DIM)(*([int(v) for v in wrapped] + [0] * (MAX_DIM - len(wrapped))))
        check(lib().clue_make_clusters(self.handle, stream.handle if stream else None, ctypes.byref(params),
                                       ndim, n, cp, weight_ptr, ctypes.byref(kernel), ctypes.byref(metric),
                                       wr, cluster_ptr, is_seed_ptr, rho_ptr, nh_ptr))

    def make_clusters_batched_device(self, stream, params, ndim, event_offsets, coord_ptrs, weight_ptr,
                                     kernel, metric, wrapped, cluster_ptr, is_seed_ptr=None, rho_ptr=None,
                                     nh_ptr=None):
        """Enqueue one batched clustering: event e is points [event_offsets[e],
        event_offsets[e+1]) of the device arrays; every event is clustered as a
        fresh Clusterer would cluster it alone (labels and nearest-higher ids local
        to the event).  No synchronisation."""
        off = np.ascontiguousarray(event_offsets, dtype=np.int64)
        if off.ndim != 1 or off.size < 2:
            raise InvalidArgument(1, "event_offsets needs n_events + 1 entries")
        cp = (_vp * MAX_DIM)(*([int(p) for p in coord_ptrs] + [0] * (MAX_DIM - len(coord_ptrs))))
        wr = None
        if wrapped is not None:
            wr = (ctypes.c_uint8 * MAX_DIM)(*([int(v) for v in wrapped] + [0] * (MAX_DIM - len(wrapped))))
        check(lib().clue_make_clusters_batched(self.handle, stream.handle if stream else None,
                                               ctypes.byref(params), ndim, off.size - 1,
                                               off.ctypes.data_as(_P(ctypes.c_int64)), cp, weight_ptr,
                                               ctypes.byref(kernel), ctypes.byref(metric), wr, cluster_ptr,
                                               is_seed_ptr, rho_ptr, nh_ptr))

    def autotune_batched(self, stream, params, ndim, event_offsets, coord_ptrs, weight_ptr, kernel, metric,
                         wrapped=None, tune_points_per_tile=False):
        """Time batched calls on these events; returns (tuning dict, points_per_tile, median ms).
        The plan's own tuning is unchanged; apply the result with set_tuning(**tuning)."""
        off = np.ascontiguousarray(event_offsets, dtype=np.int64)
        cp = _coord_array(coord_ptrs)
        wr = _wrap_array(wrapped)
        prm = ClueParams.from_buffer_copy(params)
        best = ClueTuning()
        ms = ctypes.c_float(0)
        check(lib().clue_autotune_batched(self.handle, stream.handle if stream else None, ctypes.byref(prm), ndim,
                                          off.size - 1, off.ctypes.data_as(_P(ctypes.c_int64)), cp, weight_ptr,
                                          ctypes.byref(kernel), ctypes.byref(metric), wr,
                                          1 if tune_points_per_tile else 0, ctypes.byref(best), ctypes.byref(ms)))
        return best.as_dict(), prm.points_per_tile, ms.value

    def stats(self):
        s = ClueStats()
        check(lib().clue_plan_last_stats(self.handle, ctypes.byref(s)))
        return s.as_dict()

    def grid(self):
        g = ClueGridInfo()
        check(lib().clue_plan_last_grid(self.handle, ctypes.byref(g)))
        d = g.ndim
        return {"npd": g.npd, "ntiles": g.ntiles, "minmax": np.array(g.minmax[: 2 * d], np.float32),
                "tile_size": np.array(g.tile_size[:d], np.float32)}

    def tiles(self):
        g = self.grid()
        n = ClueGridInfo()
        check(lib().clue_plan_last_grid(self.handle, ctypes.byref(n)))
        off = np.empty(g["ntiles"] + 1, np.int32)
        idx = np.empty(n.npoints, np.int32)
        check(lib().clue_plan_copy_tiles(self.handle, off.ctypes.data_as(_P(ctypes.c_int32)),
                                         idx.ctypes.data_as(_P(ctypes.c_int32))))
        return off, idx

    def delta(self):
        g = ClueGridInfo()
        check(lib().clue_plan_last_grid(self.handle, ctypes.byref(g)))
        out = np.empty(g.npoints, np.float32)
        check(lib().clue_plan_copy_delta(self.handle, out.ctypes.data_as(_P(ctypes.c_float))))
        return out

    def get_tuning(self):
        t = ClueTuning()
        check(lib().clue_plan_get_tuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = ClueTuning()
        check(lib().clue_plan_get_tuning(self.handle, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        check(lib().clue_plan_set_tuning(self.handle, ctypes.byref(t)))

    def set_timing(self, enable=True):
        check(lib().clue_plan_set_timing(self.handle, 1 if enable else 0))

    def stage_times(self):
        """{stage name: milliseconds} of the last call made with timing enabled."""
        cnt = ctypes.c_int32(0)
        check(lib().clue_plan_stage_times(self.handle, None, 0, ctypes.byref(cnt)))
        ms = (ctypes.c_float * cnt.value)()
        check(lib().clue_plan_stage_times(self.handle, ms, cnt.value, ctypes.byref(cnt)))
        return {lib().clue_stage_name(k).decode(): float(ms[k]) for k in range(cnt.value)}

    def close(self):
        if self.handle:
            lib().clue_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _coord_array(coord_ptrs):
    return (_vp * MAX_DIM)(*([int(p) for p in coord_ptrs] + [0] * (MAX_DIM - len(coord_ptrs))))


def _wrap_array(wrapped):
    if wrapped is None:
        return None
    return (ctypes.c_uint8 * MAX_DIM)(*([int(v) for v in wrapped] + [0] * (MAX_DIM - len(wrapped))))


class BatchGraph:
    """A batched call captured as a HIP graph (clue_batch_graph_*): replay() re-runs
    it on the same buffers, whose values may change between replays."""

    def __init__(self, params, ndim, event_offsets, coord_ptrs, weight_ptr, kernel, metric, wrapped,
                 cluster_ptr, is_seed_ptr=None, rho_ptr=None, nh_ptr=None, tuning=None, device=0):
        off = np.ascontiguousarray(event_offsets, dtype=np.int64)
        t = None
        if tuning is not None:
            t = ClueTuning()
            for k, v in tuning.items():
                setattr(t, k, v)
        h = _vp()
        check(lib().clue_batch_graph_create(device, ctypes.byref(params), ndim, off.size - 1,
                                            off.ctypes.data_as(_P(ctypes.c_int64)), _coord_array(coord_ptrs),
                                            weight_ptr, ctypes.byref(kernel), ctypes.byref(metric),
                                            _wrap_array(wrapped), cluster_ptr, is_seed_ptr, rho_ptr, nh_ptr,
                                            ctypes.byref(t) if t is not None else None, ctypes.byref(h)))
        self.handle = h.value

    def replay(self, stream):
        check(lib().clue_batch_graph_launch(self.handle, stream.handle if stream else None))

    def close(self):
        if self.handle:
            lib().clue_batch_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Served clustering (clue_server_*): submit() queues one event and returns a
    Ticket; Ticket.result() blocks until it is done.  Requests of one class
    (dimension, parameters, kernel, metric) are batched continuously on the GPU."""

    class Ticket:
        def __init__(self, server, ticket, keep, outputs):
            self.server, self.ticket, self._keep, self.outputs = server, ticket, keep, outputs
            self.timing = None

        def result(self):
            t = ClueRequestTiming()
            try:
                check(lib().clue_server_wait(self.server.handle, self.ticket, ctypes.byref(t)))
            finally:
                self.server._outstanding.pop(self.ticket, None)
            self.timing = {"submit_ns": t.submit_ns, "launch_ns": t.launch_ns, "done_ns": t.done_ns,
                           "batch_events": t.batch_events}
            self._keep = None
            return Result(*self.outputs)

    def __init__(self, device=0, budget_bytes=0, slots=0, max_batch_points=0, max_batch_events=0, max_dim=0,
                 small_batch_points=0):
        cfg = ClueServerConfig(budget_bytes, slots, max_batch_points, max_batch_events, max_dim, small_batch_points)
        h = _vp()
        check(lib().clue_server_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._outstanding = {}  # ticket -> host arrays the server still reads or writes

    def submit(self, coords, weights, dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5),
               metric="euclidean", metric_params=None, wrapped=None, seed_mode=SEEDS_REFERENCE, outputs="all",
               cluster_out=None):
        """Queue one clustering request; returns a Ticket.  outputs="labels" returns the
        labels only; cluster_out (int32, n elements, C-contiguous) receives them in place
        of a new array."""
        c = np.ascontiguousarray(coords, dtype=np.float32)
        if c.ndim == 1:
            c = c[None, :]
        ndim, n = c.shape
        w = np.ascontiguousarray(weights, dtype=np.float32)
        r = ClueRequest()
        r.ndim, r.n = ndim, n
        r.coords, r.weights = c.ctypes.data, w.ctypes.data
        r.params = make_params(dc, rhoc, dm, seed_dc, ppb, seed_mode)
        r.kernel = kernel_desc(kernel)
        r.metric = metric_desc(metric, metric_params)
        for d, v in enumerate(wrapped or []):
            r.wrapped[d] = int(v)
        if cluster_out is not None:
            if (not isinstance(cluster_out, np.ndarray) or cluster_out.dtype != np.int32 or cluster_out.size != n
                    or not cluster_out.flags.c_contiguous or not cluster_out.flags.writeable):
                raise InvalidArgument(1, "cluster_out must be a writeable C-contiguous int32 array of n elements")
            cl = cluster_out.reshape(-1)
        else:
            cl = np.empty(n, np.int32)
        full = outputs == "all"
        seed = np.empty(n, np.int32) if full else None
        rho = np.empty(n, np.float32) if full else None
        nh = np.empty(n, np.int32) if full else None
        r.cluster_out = cl.ctypes.data
        r.is_seed_out = seed.ctypes.data if full else None
        r.rho_out = rho.ctypes.data if full else None
        r.nearest_higher_out = nh.ctypes.data if full else None
        t = ctypes.c_uint64(0)
        check(lib().clue_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)))
        self._outstanding[t.value] = (c, w, cl, seed, rho, nh)
        return Server.Ticket(self, t.value, (c, w), (cl, seed, rho, nh))

    def stats(self):
        st = ClueServerStats()
        check(lib().clue_server_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.handle:
            lib().clue_server_destroy(self.handle)  # finishes every submitted request first
            self.handle = None
            self._outstanding.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_params(dc, rhoc, dm=None, seed_dc=None, ppb=128, seed_mode=SEEDS_REFERENCE):
    """Reference defaults: dm = seed_dc = dc, pPBin = 128 (core/Clusterer.hpp:303-309)."""
    p = ClueParams()
    p.dc, p.rhoc = dc, rhoc
    p.dm = dc if dm is None else dm
    p.seed_dc = dc if seed_dc is None else seed_dc
    p.points_per_tile = ppb
    p.seed_mode = seed_mode
    return p


def cluster_points(coords, weights, dc, rhoc, dm=None, seed_dc=None, ppb=128, kernel=("flat", 0.5),
                   metric="euclidean", metric_params=None, wrapped=None, seed_mode=SEEDS_REFERENCE,
                   device=0, plan=None, stream=None):
    """Cluster host points (coords: (ndim, n) float32 SoA) on the GPU; returns a Result
    with every per-point output in original point order."""
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim == 1:
        c = c[None, :]
    ndim, n = c.shape
    if not 1 <= ndim <= MAX_DIM:
        raise InvalidArgument(1, "number of dimensions must be in [1, 10]")
    w = np.ascontiguousarray(weights, dtype=np.float32)
    own_plan = plan is None
    plan = plan or Plan(device)
    own_stream = stream is None
    stream = stream or Stream(device)
    try:
        dc_arr = DeviceArray.from_host(c, device)
        dw = DeviceArray.from_host(w, device)
        out_c = DeviceArray((n,), np.int32, device)
        out_s = DeviceArray((n,), np.int32, device)
        out_r = DeviceArray((n,), np.float32, device)
        out_nh = DeviceArray((n,), np.int32, device)
        params = make_params(dc, rhoc, dm, seed_dc, ppb, seed_mode)
        plan.make_clusters_device(stream, params, ndim, n, [dc_arr.offset(d * n) for d in range(ndim)], dw.ptr,
                                  kernel_desc(kernel), metric_desc(metric, metric_params), wrapped,
                                  out_c.ptr, out_s.ptr, out_r.ptr, out_nh.ptr)
        stream.synchronize()
        return Result(out_c.download(stream), out_s.download(stream), out_r.download(stream),
                      out_nh.download(stream))
    finally:
        if own_stream:
            stream.close()
        if own_plan:
            plan.close()


def cluster_events(coords, weights, event_offsets, dc, rhoc, dm=None, seed_dc=None, ppb=128,
                   kernel=("flat", 0.5), metric="euclidean", metric_params=None, wrapped=None,
                   seed_mode=SEEDS_REFERENCE, device=0, plan=None, stream=None):
    """Cluster a batch of independent events in one call.  coords: (ndim, n) float32
    SoA of all events back to back; event e is columns [event_offsets[e],
    event_offsets[e+1]).  Returns a Result in batch order whose labels and
    nearest-higher ids are local to each event."""
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim == 1:
        c = c[None, :]
    ndim, n = c.shape
    if not 1 <= ndim <= MAX_DIM:
        raise InvalidArgument(1, "number of dimensions must be in [1, 10]")
    w = np.ascontiguousarray(weights, dtype=np.float32)
    off = np.ascontiguousarray(event_offsets, dtype=np.int64)
    if off.size < 2 or off[0] != 0 or off[-1] != n:
        raise InvalidArgument(1, "event_offsets must run from 0 to the number of points")
    own_plan = plan is None
    plan = plan or Plan(device)
    own_stream = stream is None
    stream = stream or Stream(device)
    try:
        dc_arr = DeviceArray.from_host(c, device)
        dw = DeviceArray.from_host(w, device)
        out_c = DeviceArray((n,), np.int32, device)
        out_s = DeviceArray((n,), np.int32, device)
        out_r = DeviceArray((n,), np.float32, device)
        out_nh = DeviceArray((n,), np.int32, device)
        params = make_params(dc, rhoc, dm, seed_dc, ppb, seed_mode)
        plan.make_clusters_batched_device(stream, params, ndim, off, [dc_arr.offset(d * n) for d in range(ndim)],
                                          dw.ptr, kernel_desc(kernel), metric_desc(metric, metric_params),
                                          wrapped, out_c.ptr, out_s.ptr, out_r.ptr, out_nh.ptr)
        stream.synchronize()
        return Result(out_c.download(stream), out_s.download(stream), out_r.download(stream),
                      out_nh.download(stream))
    finally:
        if own_stream:
            stream.close()
        if own_plan:
            plan.close()


def cluster_map(labels, device=0, plan=None, stream=None):
    """Cluster -> points map of a label array, built on the GPU (clue_cluster_map):
    returns (offsets[n_clusters+1], indexes[clustered]); cluster sizes are
    np.diff(offsets)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    n = lab.size
    own_plan, own_stream = plan is None, stream is None
    plan = plan or Plan(device)
    stream = stream or Stream(device)
    try:
        dl = DeviceArray.from_host(lab, device)
        off = DeviceArray((n + 1,), np.int32, device)
        idx = DeviceArray((n,), np.int32, device)
        cnt = DeviceArray((2,), np.int32, device)
        check(lib().clue_cluster_map(plan.handle, stream.handle, n, dl.ptr, off.ptr, idx.ptr, cnt.ptr))
        counts = cnt.download(stream)
        return off.download(stream)[: counts[0] + 1], idx.download(stream)[: counts[1]]
    finally:
        if own_stream:
            stream.close()
        if own_plan:
            plan.close()


def cluster_centroids(coords, labels, device=0, plan=None, stream=None):
    """(n_clusters, ndim) centroids on the GPU (clue_cluster_map + clue_cluster_centroids),
    with the reference's serial float arithmetic."""
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim == 1:
        c = c[None, :]
    ndim, n = c.shape
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    own_plan, own_stream = plan is None, stream is None
    plan = plan or Plan(device)
    stream = stream or Stream(device)
    try:
        dc_arr = DeviceArray.from_host(c, device)
        dl = DeviceArray.from_host(lab, device)
        off = DeviceArray((n + 1,), np.int32, device)
        idx = DeviceArray((n,), np.int32, device)
        cnt = DeviceArray((2,), np.int32, device)
        out = DeviceArray((n, ndim), np.float32, device)
        check(lib().clue_cluster_map(plan.handle, stream.handle, n, dl.ptr, off.ptr, idx.ptr, cnt.ptr))
        cp = _coord_array([dc_arr.offset(d * n) for d in range(ndim)])
        check(lib().clue_cluster_centroids(plan.handle, stream.handle, ndim, cp, off.ptr, idx.ptr, cnt.ptr, n,
                                           out.ptr))
        nc = int(cnt.download(stream)[0])
        return out.download(stream)[:nc]
    finally:
        if own_stream:
            stream.close()
        if own_plan:
            plan.close()


def main_run(dc, rhoc, dm, seed_dc, ppbin, wrapped, data, results, kernel, ndim, n_points,
             block_size=256, device_id=0):
    """mainRun(dc, rhoc, dm, seed_dc, pPBin, wrapped, data, results, kernel, Ndim, n_points,
    block_size, device_id) of the reference's bindings: data is float32[(ndim+1)*n] SoA
    (coordinates then weights), results int32[n] receives the labels in place."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    if results.dtype != np.int32 or not results.flags["C_CONTIGUOUS"]:
        raise InvalidArgument(1, "results must be a contiguous int32 array")
    wr = (ctypes.c_uint8 * MAX_DIM)(*([int(v) for v in wrapped] + [0] * (MAX_DIM - len(wrapped))))
    k = kernel_desc(kernel)
    check(lib().clue_main_run(dc, rhoc, dm, seed_dc, ppbin, wr, data.ctypes.data_as(_P(ctypes.c_float)),
                              results.ctypes.data_as(_P(ctypes.c_int32)), ctypes.byref(k), ndim, n_points,
                              block_size, device_id))


def selftest_math(op, a, b=None, device=0):
    """Evaluate the engine's device arithmetic on a host array: 0 expf, 1 sqrtf, 2 a/b, 3 sqrt bound,
    4 a/b through a double reciprocal, 5 expf of non-positive arguments, 6 / 7 the largest /
    smallest float whose sqrtf equals sqrtf(a), 8 / 9 expf of arguments in [-14, 0] as the density
    pass evaluates it for unit-amplitude kernels (9: its wide table, arguments clamped to [-14, 0]),
    10 the nearest-higher searches' sqrtf of finite a >= 0, 11 their bound below which a squared
    distance has a smaller sqrtf than sqrtf(a) (finite a >= 0)."""
    a = np.ascontiguousarray(a, np.float32)
    da = DeviceArray.from_host(a, device)
    db = DeviceArray.from_host(np.ascontiguousarray(b, np.float32), device) if b is not None else None
    out = DeviceArray(a.shape, np.float32, device)
    check(lib().clue_selftest_math(op, da.ptr, db.ptr if db else None, out.ptr, a.size, None))
    return out.download()


def selftest_sqrt(first_bits, count):
    """Mismatches of the kernels' fast sqrt against IEEE sqrtf over a range of float bit patterns."""
    m = ctypes.c_uint64(0)
    check(lib().clue_selftest_sqrt(first_bits, count, ctypes.byref(m)))
    return m.value
