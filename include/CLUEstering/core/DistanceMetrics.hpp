// Distance metrics (reference: core/DistanceMetrics.hpp:17-258).  operator()
// evaluates on the host with the reference's operation order (squared terms
// summed as a right fold, internal/meta/accumulate.hpp:14-15); desc() hands the
// metric to the engine, which evaluates the identical expression on the GPU.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <concepts>
#include <cstddef>

#include "clue_abi.h"

namespace clue {

  template <std::size_t Ndim>
  using Point = std::array<float, Ndim + 1>;

  namespace detail {
    template <std::size_t Ndim, typename F>
    constexpr float right_fold(F&& term) {
      float s = 0.f;
      for (std::size_t k = Ndim; k-- > 0;) s = term(k) + s;
      return s;
    }
    template <std::size_t Ndim>
    clue_metric_desc make_desc(int32_t type, const std::array<float, Ndim>* params) {
      clue_metric_desc d{};
      d.type = type;
      if (params)
        for (std::size_t k = 0; k < Ndim && k < CLUE_MAX_DIM; ++k) d.params[k] = (*params)[k];
      return d;
    }
  }  // namespace detail

  template <std::size_t Ndim>
  class EuclideanMetric {
  public:
    constexpr EuclideanMetric() {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      return std::sqrt(detail::right_fold<Ndim>([&](std::size_t k) {
        const float df = lhs[k] - rhs[k];
        return df * df;
      }));
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_EUCLIDEAN, nullptr); }
  };

  template <std::size_t Ndim>
  class WeightedEuclideanMetric {
  public:
    template <std::floating_point... TValues>
      requires(sizeof...(TValues) == Ndim)
    constexpr WeightedEuclideanMetric(TValues... weights) : m_weights{static_cast<float>(weights)...} {}
    constexpr WeightedEuclideanMetric(const std::array<float, Ndim>& weights) : m_weights{weights} {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      return std::sqrt(detail::right_fold<Ndim>([&](std::size_t k) {
        const float df = lhs[k] - rhs[k];
        return (m_weights[k] * df) * df;
      }));
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_WEIGHTED_EUCLIDEAN, &m_weights); }

  private:
    std::array<float, Ndim> m_weights;
  };

  /// Periods per dimension; a period of 0 marks a non-periodic coordinate.
  template <std::size_t Ndim>
  class PeriodicEuclideanMetric {
  public:
    constexpr PeriodicEuclideanMetric(const std::array<float, Ndim>& periods) : m_periods{periods} {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      return std::sqrt(detail::right_fold<Ndim>([&](std::size_t k) {
        const float diff = std::fabs(lhs[k] - rhs[k]);
        const float per = m_periods[k] - diff;
        const float pd = per < diff ? per : diff;
        return pd * pd;
      }));
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_PERIODIC_EUCLIDEAN, &m_periods); }

  private:
    std::array<float, Ndim> m_periods;
  };

  template <std::size_t Ndim>
  class ManhattanMetric {
  public:
    constexpr ManhattanMetric() {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      return detail::right_fold<Ndim>([&](std::size_t k) { return std::fabs(lhs[k] - rhs[k]); });
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_MANHATTAN, nullptr); }
  };

  template <std::size_t Ndim>
  class ChebyshevMetric {
  public:
    constexpr ChebyshevMetric() {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      float m = 0.f;
      for (std::size_t k = 0; k < Ndim; ++k) m = std::max(m, std::fabs(lhs[k] - rhs[k]));
      return m;
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_CHEBYSHEV, nullptr); }
  };

  template <std::size_t Ndim>
  class WeightedChebyshevMetric {
  public:
    template <std::floating_point... TValues>
      requires(sizeof...(TValues) == Ndim)
    constexpr WeightedChebyshevMetric(TValues... weights) : m_weights{static_cast<float>(weights)...} {}
    constexpr WeightedChebyshevMetric(const std::array<float, Ndim>& weights) : m_weights{weights} {}
    float operator()(const Point<Ndim>& lhs, const Point<Ndim>& rhs) const {
      float m = 0.f;
      for (std::size_t k = 0; k < Ndim; ++k) m = std::max(m, m_weights[k] * std::fabs(lhs[k] - rhs[k]));
      return m;
    }
    clue_metric_desc desc() const { return detail::make_desc<Ndim>(CLUE_METRIC_WEIGHTED_CHEBYSHEV, &m_weights); }

  private:
    std::array<float, Ndim> m_weights;
  };

  namespace concepts {
    template <typename TMetric, std::size_t Ndim>
    concept distance_metric = requires(const TMetric& m) {
      { m(Point<Ndim>{}, Point<Ndim>{}) } -> std::same_as<float>;
      { m.desc() } -> std::same_as<clue_metric_desc>;
    };
  }  // namespace concepts

  namespace metrics {
    template <std::size_t Ndim>
    using Euclidean = EuclideanMetric<Ndim>;
    template <std::size_t Ndim>
    using WeightedEuclidean = WeightedEuclideanMetric<Ndim>;
    template <std::size_t Ndim>
    using PeriodicEuclidean = PeriodicEuclideanMetric<Ndim>;
    template <std::size_t Ndim>
    using Manhattan = ManhattanMetric<Ndim>;
    template <std::size_t Ndim>
    using Chebyshev = ChebyshevMetric<Ndim>;
    template <std::size_t Ndim>
    using WeightedChebyshev = WeightedChebyshevMetric<Ndim>;
  }  // namespace metrics

}  // namespace clue
