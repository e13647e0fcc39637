"""The C++20 drop-in API (include/CLUEstering/*.hpp) driven like the reference's
doctest suite: a compiled test program (tests/cpp/test_api.cpp) runs each
make_clusters overload on the reference's datasets; its labels must equal the
oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_api")


def need_bin():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (run __graft_entry__.build())")


def write_csv(path, coords, weights):
    with open(path, "w") as f:
        f.write(",".join([f"x{d}" for d in range(coords.shape[0])] + ["weight"]) + "\n")
        for i in range(coords.shape[1]):
            f.write(",".join(repr(float(v)) for v in coords[:, i]) + f",{repr(float(weights[i]))}\n")


def test_cpp_unit_checks():
    need_bin()
    r = subprocess.run([BIN, "unit"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


def test_csv_roundtrip_is_exact(ref_data, tmp_path):
    c, w = ref_data("toyDetector_1000")
    path = tmp_path / "t.csv"
    write_csv(path, c, w)
    c2, w2 = oracle.read_csv(path, 2)
    np.testing.assert_array_equal(c, c2)
    np.testing.assert_array_equal(w, w2)


@pytest.mark.gpu
@pytest.mark.parametrize("overload", ["host", "host_queue", "host_device", "device"])
@pytest.mark.parametrize("name,ndim,dc,rhoc,dm", [("data_32768", 2, 1.3, 10.0, 1.3), ("blob", 3, 1.0, 5.0, 2.0),
                                                  ("toyDetector_5000", 2, 4.0, 2.5, 4.0)])
def test_cpp_make_clusters(ref_data, tmp_path, overload, name, ndim, dc, rhoc, dm):
    need_bin()
    c, w = ref_data(name)
    csv = tmp_path / f"{name}.csv"
    write_csv(csv, c, w)
    out = tmp_path / "labels.bin"
    print(f"case: {name} overload={overload}", flush=True)
    r = subprocess.run([BIN, "cluster", str(csv), str(ndim), str(dc), str(rhoc), str(dm), overload, str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    labels = np.fromfile(out, np.int32)
    ref = oracle.make_clusters(c, w, oracle.make_params(dc, rhoc, dm))
    np.testing.assert_array_equal(labels, ref.cluster)
    if name == "data_32768":
        assert "clusters: 20" in r.stdout  # tests/test_host_points.cpp:437-450
