"""clue_make_clusters_noalloc (the entry point that allocates nothing, its
buffers carved from one static device workspace per GPU) against
clue_make_clusters on a fresh plan: every output bit-exact, on two calls in a
row (the second smaller, in another dimension, so the workspace holds the first
call's contents), with the device's free memory unchanged across the calls,
and a call beyond the workspace refused with a clear error."""
import ctypes

import numpy as np
import pytest

import cluestering_alpaka3_amd as clue
from cluestering_alpaka3_amd import datasets

pytestmark = pytest.mark.gpu


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


class Case:
    def __init__(self, ps, prm, stream):
        self.ps, self.prm = ps, prm
        D, n = ps.ndim, ps.n
        self.dc = clue.DeviceArray.from_host(ps.coords)
        self.dw = clue.DeviceArray.from_host(ps.weights)
        self.cp = (ctypes.c_void_p * clue.MAX_DIM)(*([self.dc.offset(d * n) for d in range(D)] +
                                                    [0] * (clue.MAX_DIM - D)))
        self.params = clue.make_params(prm["dc"], prm["rhoc"], prm["dm"])
        self.kd, self.md = clue.kernel_desc(prm["kernel"]), clue.metric_desc("euclidean")
        self.outs = [clue.DeviceArray((n,), t) for t in (np.int32, np.int32, np.float32, np.int32)]
        self.stream = stream

    def noalloc(self):
        ps = self.ps
        return clue.lib().clue_make_clusters_noalloc(
            0, self.stream.handle, ctypes.byref(self.params), ps.ndim, ps.n, self.cp, self.dw.ptr,
            ctypes.byref(self.kd), ctypes.byref(self.md), None, *[o.ptr for o in self.outs])

    def download(self):
        return [o.download(self.stream) for o in self.outs]

    def planned(self):
        plan = clue.Plan(0)
        try:
            plan.make_clusters_device(self.stream, self.params, self.ps.ndim, self.ps.n,
                                      [self.dc.offset(d * self.ps.n) for d in range(self.ps.ndim)], self.dw.ptr,
                                      self.kd, self.md, None, *[o.ptr for o in self.outs])
            return self.download()
        finally:
            plan.close()


def test_noalloc_equals_make_clusters():
    st = clue.Stream(0)
    first = Case(datasets.workload_scaled("2d_1m_gauss", 200000, seed=12), datasets.WORKLOADS["2d_1m_gauss"], st)
    second = Case(datasets.workload_scaled("3d_4m_exp", 60000, seed=13), datasets.WORKLOADS["3d_4m_exp"], st)
    want = []
    for c in (first, second):
        print(f"noalloc reference: n={c.ps.n} ndim={c.ps.ndim}", flush=True)
        want.append(c.planned())
    print("noalloc: warm-up call", flush=True)
    clue.check(first.noalloc())
    st.synchronize()
    free0 = device_free_bytes()
    for rep in range(2):
        for c, w in zip((first, second), want):
            print(f"noalloc: call n={c.ps.n} ndim={c.ps.ndim} rep={rep}", flush=True)
            for o in c.outs:
                clue.check(clue.lib().clue_memset_device(o.ptr, 0x7f, o.nbytes, st.handle))
            clue.check(c.noalloc())
            got = c.download()
            for a, b, name in zip(got, w, ("cluster", "is_seed", "rho", "nearest_higher")):
                np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32),
                                              err_msg=f"{name} n={c.ps.n} rep={rep}")
    st.synchronize()
    assert device_free_bytes() == free0


def test_noalloc_refuses_calls_beyond_the_workspace():
    st = clue.Stream(0)
    ps = datasets.workload_scaled("2d_1m_gauss", 800000, seed=14)
    c = Case(ps, datasets.WORKLOADS["2d_1m_gauss"], st)
    print(f"noalloc: oversized call n={ps.n}", flush=True)
    status = c.noalloc()
    assert status == 2  # CLUE_ERR_OUT_OF_RANGE
    msg = clue.lib().clue_last_error_message().decode()
    assert "workspace" in msg and "800000" in msg
