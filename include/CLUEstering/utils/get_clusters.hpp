// Cluster -> points maps (reference: utils/get_clusters.hpp, detail/get_clusters.hpp:21-46).
#pragma once
#include <vector>

#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/data_structures/DeviceAssociationMap.hpp"
#include "CLUEstering/data_structures/PointsDevice.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"

namespace clue {
  /// Host map of host points.
  template <std::size_t Ndim>
  HostAssociationMap get_clusters(const PointsHost<Ndim>& points) {
    return detail::get_clusters(points.clusterIndexes());
  }
  /// Device map of device points, built on the GPU (asynchronous on `queue`).
  template <concepts::Queue TQueue, std::size_t Ndim>
  DeviceAssociationMap get_clusters(TQueue& queue, const PointsDevice<DevType<TQueue>, Ndim>& points) {
    return DeviceAssociationMap(queue, points.size(), points.view().cluster_index);
  }
}  // namespace clue
