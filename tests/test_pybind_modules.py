"""The reference's Python FFI modules rebuilt over the engine: CLUE_Convolutional_Kernels
and CLUE_GPU_HIP, called the way CLUEstering.py calls a backend (CLUEstering.py:668-707)."""
import os
import sys

import numpy as np
import pytest

import oracle

PYBIND_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cluestering_alpaka3_amd",
                          "pybind")
sys.path.insert(0, PYBIND_DIR)


def test_modules_import_and_kernel_validation():
    import CLUE_Convolutional_Kernels as k
    import CLUE_GPU_HIP as hip
    k.FlatKernel(0.5)
    k.ExponentialKernel(1.0, 1.0)
    k.GaussianKernel(1.0, 2.0, 1.0)
    # the reference constructors' throwing conditions (detail/ConvolutionalKernel.hpp:10-53)
    for bad in (lambda: k.FlatKernel(0.0), lambda: k.FlatKernel(-1.0), lambda: k.GaussianKernel(0.0, 1.0, 1.0),
                lambda: k.GaussianKernel(1.0, -1.0, 1.0), lambda: k.ExponentialKernel(-1.0, 1.0),
                lambda: k.ExponentialKernel(1.0, 0.0)):
        with pytest.raises(ValueError):
            bad()
    assert hasattr(hip, "mainRun") and hasattr(hip, "listDevices")


@pytest.mark.gpu
@pytest.mark.parametrize("name,ndim,dc,rhoc,dm,kernel", [
    ("data_32768", 2, 1.3, 10.0, 1.3, ("flat", 0.5)),
    ("toyDetector_1000", 2, 4.0, 2.5, 4.0, ("exponential", 1.0, 2.0)),
    ("blob", 3, 1.0, 5.0, 2.0, ("gaussian", 1.0, 2.0, 1.0)),
])
def test_main_run_like_the_reference_python(ref_data, name, ndim, dc, rhoc, dm, kernel):
    import CLUE_Convolutional_Kernels as k
    import CLUE_GPU_HIP as hip
    c, w = ref_data(name)
    n = c.shape[1]
    data = np.ascontiguousarray(np.concatenate([c.reshape(-1), w]).astype(np.float32))
    results = np.zeros(n, np.int32)
    kobj = {"flat": lambda: k.FlatKernel(kernel[1]), "exponential": lambda: k.ExponentialKernel(*kernel[1:]),
            "gaussian": lambda: k.GaussianKernel(*kernel[1:])}[kernel[0]]()
    print(f"mainRun: {name} ndim={ndim} n={n} kernel={kernel}", flush=True)
    hip.mainRun(dc, rhoc, dm, dc, 128, [0] * ndim, data, results, kobj, ndim, n, 256, 0)
    ref = oracle.make_clusters(c, w, oracle.make_params(dc, rhoc, dm, dc, 128, kernel), with_tiles=False)
    np.testing.assert_array_equal(results, ref.cluster)
