// Convolutional kernels of the local density (reference:
// core/ConvolutionalKernel.hpp:14-85, detail/ConvolutionalKernel.hpp:10-65).
// The constructors keep the reference's checks; desc() hands the kernel to the
// engine, which evaluates the same formula on the GPU (exp correctly rounded).
#pragma once
#include <cmath>
#include <stdexcept>

#include "clue_abi.h"

namespace clue {

  class FlatKernel {
  public:
    FlatKernel(float flat) : m_flat{flat} {
      if (flat <= 0.f) throw std::invalid_argument("Flat kernel parameter must be positive.");
    }
    /// Host evaluation: 1 for the point itself, else the flat height.
    float operator()(float /*dist_ij*/, int point_id, int j) const { return point_id == j ? 1.f : m_flat; }
    clue_kernel_desc desc() const { return {CLUE_KERNEL_FLAT, {m_flat, 0.f, 0.f}}; }

  private:
    float m_flat;
  };

  class GaussianKernel {
  public:
    GaussianKernel(float gaus_avg, float gaus_std, float gaus_amplitude)
        : m_avg{gaus_avg}, m_std{gaus_std}, m_amplitude{gaus_amplitude} {
      if (gaus_std <= 0.f || gaus_amplitude <= 0.f || gaus_avg <= 0.f)
        throw std::invalid_argument("Gaussian kernel parameters must be positive.");
    }
    /// Same formula without the positivity check on the mean, so a kernel
    /// centred at 0 can be used (the public constructor rejects mean <= 0).
    static GaussianKernel unchecked(float gaus_avg, float gaus_std, float gaus_amplitude) {
      return GaussianKernel(gaus_avg, gaus_std, gaus_amplitude, 0);
    }
    float operator()(float dist_ij, int point_id, int j) const {
      if (point_id == j) return 1.f;
      const float a = dist_ij - m_avg;
      return m_amplitude * std::exp((-a) * a / ((2.f * m_std) * m_std));
    }
    clue_kernel_desc desc() const { return {CLUE_KERNEL_GAUSSIAN, {m_avg, m_std, m_amplitude}}; }

  private:
    GaussianKernel(float a, float s, float amp, int) : m_avg{a}, m_std{s}, m_amplitude{amp} {}
    float m_avg, m_std, m_amplitude;
  };

  class ExponentialKernel {
  public:
    ExponentialKernel(float exp_avg, float exp_amplitude) : m_avg{exp_avg}, m_amplitude{exp_amplitude} {
      if (exp_avg <= 0.f || exp_amplitude <= 0.f)
        throw std::invalid_argument("Exponential kernel parameters must be positive.");
    }
    float operator()(float dist_ij, int point_id, int j) const {
      return point_id == j ? 1.f : m_amplitude * std::exp((-m_avg) * dist_ij);
    }
    clue_kernel_desc desc() const { return {CLUE_KERNEL_EXPONENTIAL, {m_avg, m_amplitude, 0.f}}; }

  private:
    float m_avg, m_amplitude;
  };

  namespace concepts {
    template <typename K>
    concept convolutional_kernel = requires(const K& k) {
      { k.desc() } -> std::same_as<clue_kernel_desc>;
    };
  }  // namespace concepts

}  // namespace clue
