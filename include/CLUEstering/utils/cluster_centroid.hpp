// Cluster centroids (reference: utils/cluster_centroid.hpp,
// utils/detail/cluster_centroid.hpp:14-65): per dimension the coordinates of a
// cluster's points summed from 0.f in ascending point order, divided by
// (float)size.  The host versions loop on the host; the device version builds
// the map and the centroids on the GPU (clue_cluster_map, clue_cluster_centroids)
// with the same arithmetic.
#pragma once
#include <array>
#include <cstddef>
#include <vector>

#include "CLUEstering/data_structures/DeviceAssociationMap.hpp"
#include "CLUEstering/data_structures/PointsDevice.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"
#include "CLUEstering/utils/get_clusters.hpp"

namespace clue {

  template <std::size_t Ndim>
  using Centroid = std::array<float, Ndim>;
  template <std::size_t Ndim>
  using Centroids = std::vector<Centroid<Ndim>>;

  template <std::size_t Ndim>
  Centroid<Ndim> cluster_centroid(const PointsHost<Ndim>& points, std::size_t cluster_id) {
    const auto labels = points.clusterIndexes();
    const auto clusters = get_clusters(points);
    Centroid<Ndim> c{};
    const auto size = clusters.count(static_cast<int32_t>(cluster_id));
    for (std::size_t dim = 0; dim < Ndim; ++dim) {
      const auto x = points.coords(dim);
      for (std::size_t i = 0; i < static_cast<std::size_t>(points.size()); ++i)
        if (labels[i] >= 0 && static_cast<std::size_t>(labels[i]) == cluster_id) c[dim] += x[i];
      c[dim] /= static_cast<float>(size);
    }
    return c;
  }

  template <std::size_t Ndim>
  Centroids<Ndim> cluster_centroids(const PointsHost<Ndim>& points) {
    const auto labels = points.clusterIndexes();
    const auto clusters = get_clusters(points);
    Centroids<Ndim> out(clusters.size());
    for (std::size_t dim = 0; dim < Ndim; ++dim) {
      const auto x = points.coords(dim);
      for (std::size_t i = 0; i < static_cast<std::size_t>(points.size()); ++i)
        if (labels[i] >= 0) out[static_cast<std::size_t>(labels[i])][dim] += x[i];
      for (std::size_t c = 0; c < out.size(); ++c)
        out[c][dim] /= static_cast<float>(clusters.count(static_cast<int32_t>(c)));
    }
    return out;
  }

  /// Centroids of clustered device points, computed on the GPU; synchronises `queue`.
  template <concepts::Queue TQueue, std::size_t Ndim>
  Centroids<Ndim> cluster_centroids(TQueue& queue, const PointsDevice<DevType<TQueue>, Ndim>& points) {
    DeviceAssociationMap map(queue, points.size(), points.view().cluster_index);
    const int32_t n = points.size();
    detail::DeviceBuffer buf(queue.getDevice().index(), static_cast<std::size_t>(n) * Ndim * 4);
    const float* cp[Ndim];
    for (std::size_t d = 0; d < Ndim; ++d) cp[d] = points.view().coords[d];
    detail::check(clue_cluster_centroids(map.plan(), queue.handle(), static_cast<int32_t>(Ndim), cp, map.offsets(),
                                         map.indexes(), map.counts(), n, static_cast<float*>(buf.ptr)));
    const std::size_t nc = map.size();  // waits for the queue
    Centroids<Ndim> out(nc);
    if (nc) {
      detail::check(clue_memcpy_d2h(out.data(), buf.ptr, nc * Ndim * 4, queue.handle()));
      queue.wait();
    }
    return out;
  }

}  // namespace clue
