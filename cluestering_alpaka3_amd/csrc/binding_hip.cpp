// The reference's per-backend Python module (CLUEstering/BindingModules/serial/
// binding_cpu.cpp:17-240 with Run.hpp:8-28), under the name its Python layer
// imports for the HIP backend (CLUEstering.py:57): listDevices and the three
// mainRun overloads, running the MI355X engine through the C++ drop-in.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CLUEstering/CLUEstering.hpp"

namespace py = pybind11;

namespace {

void listDevices(const std::string& backend) {
  const int32_t n = clue::DevicePool::size();
  if (n == 0) {
    std::cout << "No devices found for the " << backend << " backend." << std::endl;
    return;
  }
  std::cout << backend << " devices found: \n";
  for (int32_t i = 0; i < n; ++i) std::cout << '\t' << "device " << i << ": MI355X (gfx950)\n";
}

// Run.hpp:8-28: a Clusterer on the device's queue, host points borrowing the
// caller's SoA data (coordinates then weights) and result array, temporary
// device points, make_clusters with the Euclidean metric.
template <std::size_t Ndim, typename Kernel>
void run(float dc, float rhoc, float dm, float seed_dc, int pPBin, std::vector<uint8_t>&& wrapped, float* data,
         int* results, int32_t n_points, const Kernel& kernel, std::size_t device_id, std::size_t block_size) {
  auto& queue = clue::get_queue(static_cast<uint32_t>(device_id));
  auto device = queue.getDevice();
  const clue::Dim<Ndim> dim{};
  clue::Clusterer algo(queue, dim, dc, rhoc, dm, seed_dc, pPBin);
  algo.setWrappedCoordinates(wrapped);
  clue::PointsHost<Ndim> h_points(dim, n_points, data, results);
  clue::PointsDevice d_points(device, dim, n_points);
  algo.make_clusters(queue, h_points, d_points, clue::EuclideanMetric<Ndim>{}, kernel, block_size);
}

template <typename Kernel>
void mainRun(float dc, float rhoc, float dm, float seed_dc, int pPBin, std::vector<uint8_t> wrapped,
             py::array_t<float> data, py::array_t<int> results, const Kernel& kernel, int Ndim, int32_t n_points,
             std::size_t block_size, std::size_t device_id) {
  auto rd = data.request();
  auto rr = results.request();
  if (rd.size < static_cast<py::ssize_t>(n_points) * (Ndim + 1) || rr.size < n_points)
    throw std::invalid_argument("data or results smaller than the number of points");
  auto* pd = static_cast<float*>(rd.ptr);
  auto* pr = static_cast<int*>(rr.ptr);
  py::gil_scoped_release release;
  switch (Ndim) {
    case 1: return run<1>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 2: return run<2>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 3: return run<3>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 4: return run<4>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 5: return run<5>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 6: return run<6>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 7: return run<7>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 8: return run<8>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 9: return run<9>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    case 10: return run<10>(dc, rhoc, dm, seed_dc, pPBin, std::move(wrapped), pd, pr, n_points, kernel, device_id, block_size);
    default: std::cout << "This library only works up to 10 dimensions\n";
  }
}

template <typename Kernel>
void def_main_run(py::module_& m) {
  m.def("mainRun", &mainRun<Kernel>, "mainRun", py::arg("dc"), py::arg("rhoc"), py::arg("dm"), py::arg("seed_dc"),
        py::arg("pPBin"), py::arg("wrapped"), py::arg("data"), py::arg("results"), py::arg("kernel"),
        py::arg("Ndim"), py::arg("n_points"), py::arg("block_size"), py::arg("device_id"));
}

}  // namespace

PYBIND11_MODULE(CLUE_GPU_HIP, m) {
  m.doc() = "Binding of the CLUE algorithm running on MI355X GPUs";
  py::module_::import("CLUE_Convolutional_Kernels");  // kernel classes must be registered
  m.def("listDevices", &listDevices, "List the available devices for the HIP backend");
  def_main_run<clue::FlatKernel>(m);
  def_main_run<clue::ExponentialKernel>(m);
  def_main_run<clue::GaussianKernel>(m);
}
