// clue::Clusterer of the drop-in API (reference: core/Clusterer.hpp:30-207,
// core/detail/Clusterer.hpp:24-217).  Parameters, wrapped coordinates and the
// grow-only clustering state (a plan of the C ABI) live here; every
// make_clusters overload runs the HIP engine and returns with the queue drained,
// as the reference waits on its queue before returning.
#pragma once
#include <algorithm>
#include <array>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <ranges>
#include <stdexcept>
#include <vector>

#include "CLUEstering/core/ConvolutionalKernel.hpp"
#include "CLUEstering/core/Device.hpp"
#include "CLUEstering/core/DistanceMetrics.hpp"
#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/data_structures/PointsDevice.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"
#include "CLUEstering/detail/Dim.hpp"

namespace clue {

  template <concepts::Queue TQueue, std::size_t Ndim>
  class Clusterer {
    using TDev = DevType<TQueue>;
    using TPointsHost = PointsHost<Ndim>;
    using TPointsDevice = PointsDevice<TDev, Ndim>;

  public:
    Clusterer(TQueue& /*queue*/, Dim<Ndim>, float dc, float rhoc, std::optional<float> dm = std::nullopt,
              std::optional<float> seed_dc = std::nullopt, int pPBin = 128) {
      setParameters(dc, rhoc, dm, seed_dc, pPBin);
    }

    void setParameters(float dc, float rhoc, std::optional<float> dm = std::nullopt,
                       std::optional<float> seed_dc = std::nullopt, int pPBin = 128) {
      m_dc = dc;
      m_dm = dm.value_or(dc);
      m_seed_dc = seed_dc.value_or(dc);
      m_rhoc = rhoc;
      m_pointsPerTile = pPBin;
      if (m_dc <= 0.f || m_rhoc < 0.f || m_dm <= 0.f || m_seed_dc <= 0.f || m_pointsPerTile <= 0)
        throw std::invalid_argument("Invalid clustering parameters. The parameters must be positive.");
    }

    /// Host points, temporary device points on the queue's device.
    template <concepts::convolutional_kernel Kernel = FlatKernel,
              concepts::distance_metric<Ndim> DistanceMetric = EuclideanMetric<Ndim>>
    void make_clusters(TQueue& queue, TPointsHost& h_points, const DistanceMetric& metric = EuclideanMetric<Ndim>{},
                       const Kernel& kernel = FlatKernel{.5f}, std::size_t block_size = 256) {
      auto device = queue.getDevice();
      TPointsDevice d_points(device, Dim<Ndim>{}, h_points.size());
      make_clusters(queue, h_points, d_points, metric, kernel, block_size);
    }

    /// Host points on device 0 with its shared queue.
    template <concepts::convolutional_kernel Kernel = FlatKernel,
              concepts::distance_metric<Ndim> DistanceMetric = EuclideanMetric<Ndim>>
    void make_clusters(TPointsHost& h_points, const DistanceMetric& metric = EuclideanMetric<Ndim>{},
                       const Kernel& kernel = FlatKernel{.5f}, std::size_t block_size = 256) {
      auto device = DevicePool::deviceAt(0u);
      auto& queue = get_queue(device);
      make_clusters(queue, h_points, metric, kernel, block_size);
    }

    /// Host and device points: copy in, cluster, copy the labels back.
    template <concepts::convolutional_kernel Kernel = FlatKernel,
              concepts::distance_metric<Ndim> DistanceMetric = EuclideanMetric<Ndim>>
    void make_clusters(TQueue& queue, TPointsHost& h_points, TPointsDevice& d_points,
                       const DistanceMetric& metric = EuclideanMetric<Ndim>{}, const Kernel& kernel = FlatKernel{.5f},
                       std::size_t block_size = 256) {
      copyToDevice(queue, d_points, h_points);
      run(queue, d_points, metric, kernel, block_size);
      // the engine leaves coordinates and weights untouched: only the labels come back
      detail::check(clue_memcpy_d2h(h_points.view().cluster_index, d_points.view().cluster_index,
                                    static_cast<std::size_t>(h_points.size()) * 4, queue.handle()));
      queue.wait();
      h_points.mark_clustered();
      d_points.mark_clustered();
    }

    /// Device points only.
    template <concepts::convolutional_kernel Kernel = FlatKernel,
              concepts::distance_metric<Ndim> DistanceMetric = EuclideanMetric<Ndim>>
    void make_clusters(TQueue& queue, TPointsDevice& d_points, const DistanceMetric& metric = EuclideanMetric<Ndim>{},
                       const Kernel& kernel = FlatKernel{.5f}, std::size_t block_size = 256) {
      run(queue, d_points, metric, kernel, block_size);
      queue.wait();
      d_points.mark_clustered();
    }

    template <std::ranges::contiguous_range TRange>
      requires std::integral<std::ranges::range_value_t<TRange>>
    void setWrappedCoordinates(const TRange& wrapped_coordinates) {
      const std::size_t n = std::min<std::size_t>(std::ranges::size(wrapped_coordinates), Ndim);
      auto it = std::ranges::begin(wrapped_coordinates);
      for (std::size_t k = 0; k < n; ++k, ++it) m_wrapped[k] = static_cast<uint8_t>(*it);
    }
    template <std::integral... TArgs>
    void setWrappedCoordinates(TArgs... wrapped_coordinates) {
      m_wrapped = {static_cast<uint8_t>(wrapped_coordinates)...};
    }

    HostAssociationMap getClusters(const TPointsHost& h_points) {
      return detail::get_clusters(h_points.clusterIndexes());
    }
    /// Cluster -> points map of device points (built on the host from the labels).
    HostAssociationMap getClusters(TQueue& queue, const TPointsDevice& d_points) {
      std::vector<int32_t> labels(static_cast<std::size_t>(d_points.size()));
      detail::check(clue_memcpy_d2h(labels.data(), d_points.view().cluster_index, labels.size() * 4, queue.handle()));
      queue.wait();
      return detail::get_clusters(labels);
    }

    /// Counters of the last call (seed candidates, capacity, seeds, clusters, dropped seeds).
    clue_stats lastStats() const {
      clue_stats s{};
      detail::check(clue_plan_last_stats(m_plan.get(), &s));
      return s;
    }

  private:
    template <typename Kernel, typename DistanceMetric>
    void run(TQueue& queue, TPointsDevice& d_points, const DistanceMetric& metric, const Kernel& kernel,
             std::size_t /*block_size*/) {
      const int32_t dev = queue.getDevice().index();
      if (!m_plan || m_plan_device != dev) {
        clue_plan_t p = nullptr;
        detail::check(clue_plan_create(dev, &p));
        m_plan = std::shared_ptr<clue_plan_s>(p, [](clue_plan_t q) { clue_plan_destroy(q); });
        m_plan_device = dev;
      }
      const clue_params prm{m_dc, m_rhoc, m_dm, m_seed_dc, m_pointsPerTile, CLUE_SEEDS_REFERENCE};
      const clue_kernel_desc kd = kernel.desc();
      const clue_metric_desc md = metric.desc();
      const float* coords[Ndim];
      for (std::size_t d = 0; d < Ndim; ++d) coords[d] = d_points.view().coords[d];
      detail::check(clue_make_clusters(m_plan.get(), queue.handle(), &prm, static_cast<int32_t>(Ndim), d_points.size(),
                                       coords, d_points.view().weight, &kd, &md, m_wrapped.data(),
                                       d_points.view().cluster_index, d_points.view().is_seed, d_points.view().rho,
                                       d_points.view().nearest_higher));
    }

    float m_dc{}, m_seed_dc{}, m_rhoc{}, m_dm{};
    int m_pointsPerTile{128};
    std::array<uint8_t, Ndim> m_wrapped{};
    std::shared_ptr<clue_plan_s> m_plan;
    int32_t m_plan_device = -1;
  };

  template <concepts::Queue Q, std::size_t N>
  Clusterer(Q&, Dim<N>, float, float, std::optional<float>, std::optional<float>, int) -> Clusterer<std::remove_cvref_t<Q>, N>;

}  // namespace clue
