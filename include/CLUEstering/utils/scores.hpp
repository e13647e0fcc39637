// Silhouette score (reference: utils/detail/scores.hpp:20-99): mean over the
// clustered points of clusters with at least two points of (b - a) / max(a, b).
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "CLUEstering/data_structures/PointsHost.hpp"

namespace clue {
  template <std::size_t Ndim>
  float silhouette(const PointsHost<Ndim>& points) {
    const auto labels = points.clusterIndexes();
    const auto clusters = detail::get_clusters(labels);
    auto dist = [&](int i, int j) {
      float s = 0.f;
      for (std::size_t d = 0; d < Ndim; ++d) {
        const float df = points.view().coords[d][i] - points.view().coords[d][j];
        s += df * df;
      }
      return std::sqrt(s);
    };
    std::vector<float> scores;
    for (int i = 0; i < points.size(); ++i) {
      const int c = labels[i];
      if (c < 0 || clusters.count(c) < 2) continue;
      float a = 0.f;
      for (auto it = clusters.lower_bound(c); it != clusters.upper_bound(c); ++it)
        if (*it != i) a += dist(i, *it);
      a /= static_cast<float>(clusters.count(c) - 1);
      float b = std::numeric_limits<float>::max();
      for (int k = 0; k < static_cast<int>(clusters.size()); ++k) {
        if (k == c) continue;
        float s = 0.f;
        for (auto it = clusters.lower_bound(k); it != clusters.upper_bound(k); ++it) s += dist(i, *it);
        b = std::min(b, s / static_cast<float>(clusters.count(k)));
      }
      scores.push_back((b - a) / std::max(a, b));
    }
    float total = 0.f;
    for (float s : scores) total += s;
    return total / static_cast<float>(scores.size());
  }
}  // namespace clue
