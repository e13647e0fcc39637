// Cluster -> points maps (reference: utils/get_clusters.hpp, detail/get_clusters.hpp:21-46).
#pragma once
#include <vector>

#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/data_structures/PointsDevice.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"

namespace clue {
  template <std::size_t Ndim>
  HostAssociationMap get_clusters(const PointsHost<Ndim>& points) {
    return detail::get_clusters(points.clusterIndexes());
  }
  template <concepts::Queue TQueue, std::size_t Ndim>
  HostAssociationMap get_clusters(TQueue& queue, const PointsDevice<DevType<TQueue>, Ndim>& points) {
    std::vector<int32_t> labels(static_cast<std::size_t>(points.size()));
    detail::check(clue_memcpy_d2h(labels.data(), points.view().cluster_index, labels.size() * 4, queue.handle()));
    queue.wait();
    return detail::get_clusters(labels);
  }
}  // namespace clue
