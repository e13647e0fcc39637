"""Writes the committed fixtures under tests/golden/ (run in a checkout that has
the CLUEstering reference tree; the tests themselves never read that tree).

reference_data.npz   the reference's own test datasets (data/*.csv), parsed
                     with std::stof semantics (strtof), as float32 SoA arrays
                     `<name>__coords` (ndim, n) and `<name>__weights` (n,).
reference_truth.npz  cluster_ids / is_seed columns of data/truth_files/*.csv.
config1_golden.npz   the 2D 16k-point workload (10 blobs, sigma 1.5, 10% noise)
                     and the oracle's outputs for it (dc=1.5, rhoc=5, dm=3,
                     FlatKernel(0.5)) -- a regression pin of the oracle.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from cluestering_alpaka3_amd import datasets  # noqa: E402

REF = os.environ.get("CLUE_REFERENCE_DIR", "/root/reference")

DATASETS = {
    "data_1024": 2, "data_2048": 2, "data_4096": 2, "data_8192": 2, "data_16384": 2, "data_32768": 2,
    "toyDetector_1000": 2, "toyDetector_5000": 2, "toyDetector_10000": 2,
    "sissa_1000": 2, "sissa_4000": 2, "aniso_1000": 2, "moons": 2, "circles": 2,
    "opposite_angles": 2, "square": 2, "box": 3, "blob": 3, "data_dim_3": 3, "data_dim_10": 10,
}
TRUTH = ["data_1024_truth", "data_32768_truth", "toy_det_1000_truth", "toy_det_5000_truth",
         "sissa_1000_truth", "blobs_truth", "moons_1000_truth", "aniso_1000_truth"]


def main():
    arrays = {}
    for name, ndim in DATASETS.items():
        c, w = oracle.read_csv(os.path.join(REF, "data", name + ".csv"), ndim)
        arrays[name + "__coords"] = c
        arrays[name + "__weights"] = w
    np.savez_compressed(os.path.join(HERE, "reference_data.npz"), **arrays)

    truth = {}
    for name in TRUTH:
        path = os.path.join(REF, "data", "truth_files", name + ".csv")
        with open(path) as f:
            header = f.readline().strip().split(",")
        tab = np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)
        truth[name + "__cluster_ids"] = tab[:, header.index("cluster_ids")].astype(np.int32)
        if "is_seed" in header:
            truth[name + "__is_seed"] = tab[:, header.index("is_seed")].astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "reference_truth.npz"), **truth)

    ps = datasets.workload("2d_16k_flat")
    prm = datasets.WORKLOADS["2d_16k_flat"]
    r = oracle.make_clusters(ps.coords, ps.weights,
                             oracle.make_params(prm["dc"], prm["rhoc"], prm["dm"], kernel=prm["kernel"]))
    np.savez_compressed(os.path.join(HERE, "config1_golden.npz"), coords=ps.coords, weights=ps.weights,
                        rho=r.rho, nh=r.nh, is_seed=r.is_seed, cluster=r.cluster,
                        tile_offsets=r.tile_offsets, tile_indexes=r.tile_indexes)
    print("fixtures written")


if __name__ == "__main__":
    main()
