// The reference's kernels module (CLUEstering/BindingModules/binding_kernels.cpp:
// 8-16): the three convolutional kernels as Python classes, with the C++
// drop-in's constructors (and their std::invalid_argument -> ValueError).
#include <pybind11/pybind11.h>

#include "CLUEstering/core/ConvolutionalKernel.hpp"

PYBIND11_MODULE(CLUE_Convolutional_Kernels, m) {
  m.doc() = "Convolutional kernels of the CLUE algorithm (MI355X engine)";
  pybind11::class_<clue::FlatKernel>(m, "FlatKernel").def(pybind11::init<float>());
  pybind11::class_<clue::ExponentialKernel>(m, "ExponentialKernel").def(pybind11::init<float, float>());
  pybind11::class_<clue::GaussianKernel>(m, "GaussianKernel").def(pybind11::init<float, float, float>());
}
