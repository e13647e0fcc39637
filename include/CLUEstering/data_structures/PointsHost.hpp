// Host points of the drop-in API (reference: data_structures/PointsHost.hpp:45-226,
// detail/PointsHost.hpp).  SoA layout [x_0][x_1]...[x_{D-1}][weight][cluster],
// owned or borrowed.
#pragma once
#include <array>
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <ranges>
#include <span>
#include <stdexcept>
#include <vector>

#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/detail/Dim.hpp"

namespace clue {

  template <std::size_t Ndim>
  struct PointsView {
    std::array<float*, Ndim> coords{};
    float* weight = nullptr;
    int* cluster_index = nullptr;
    int* is_seed = nullptr;
    float* rho = nullptr;
    int* nearest_higher = nullptr;
    int32_t n = 0;
  };

  /// Cluster properties computed lazily from the labels (ClusterProperties.hpp:19-51).
  class ClusterProperties {
  public:
    explicit ClusterProperties(std::span<const int32_t> labels)
        : m_map{detail::get_clusters(labels)}, m_sizes(m_map.size()) {
      for (std::size_t k = 0; k < m_map.size(); ++k) m_sizes[k] = m_map.count(static_cast<int32_t>(k));
    }
    const HostAssociationMap& clusters() const { return m_map; }
    const std::vector<std::size_t>& cluster_sizes() const { return m_sizes; }
    std::size_t n_clusters() const { return m_map.size(); }

  private:
    HostAssociationMap m_map;
    std::vector<std::size_t> m_sizes;
  };

  template <std::size_t Ndim>
  class PointsHost {
  public:
    class Point {
    public:
      Point(const std::array<float, Ndim>& coordinates, float weight, int cluster_index)
          : m_coordinates{coordinates}, m_weight{weight}, m_cluster_index{cluster_index} {}
      float operator[](std::size_t dim) const { return m_coordinates[dim]; }
      float weight() const { return m_weight; }
      int cluster_index() const { return m_cluster_index; }

    private:
      std::array<float, Ndim> m_coordinates;
      float m_weight;
      int m_cluster_index;
    };

    PointsHost(Dim<Ndim>, int32_t n_points) : m_size{checked(n_points)} {
      m_owned.reset(new std::byte[soa_bytes(n_points)]);
      partition(m_owned.get());
    }
    PointsHost(Dim<Ndim>, int32_t n_points, std::span<std::byte> buffer) : m_size{checked(n_points)} {
      if (buffer.size() < soa_bytes(n_points)) throw std::invalid_argument("buffer too small for the points");
      partition(buffer.data());
    }
    PointsHost(Dim<Ndim>, int32_t n_points, std::span<float> input, std::span<int> output)
        : PointsHost(Dim<Ndim>{}, n_points, input.data(), output.data()) {}
    PointsHost(Dim<Ndim>, int32_t n_points, std::span<float> coordinates, std::span<float> weights,
               std::span<int> output)
        : PointsHost(Dim<Ndim>{}, n_points, coordinates.data(), weights.data(), output.data()) {}
    /// input: SoA coordinates followed by the weights, (Ndim+1)*n floats.
    PointsHost(Dim<Ndim>, int32_t n_points, float* input, int* output) : m_size{checked(n_points)} {
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = input + d * n_points;
      m_view.weight = input + Ndim * n_points;
      m_view.cluster_index = output;
      m_view.n = n_points;
    }
    PointsHost(Dim<Ndim>, int32_t n_points, float* coordinates, float* weights, int* output)
        : m_size{checked(n_points)} {
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = coordinates + d * n_points;
      m_view.weight = weights;
      m_view.cluster_index = output;
      m_view.n = n_points;
    }
    /// One buffer per coordinate, then the weights and the labels.
    template <typename... TPtrs>
      requires(sizeof...(TPtrs) == Ndim + 2 && Ndim > 1 && (std::is_pointer_v<TPtrs> && ...))
    PointsHost(Dim<Ndim>, int32_t n_points, TPtrs... buffers) : m_size{checked(n_points)} {
      void* ptrs[] = {static_cast<void*>(buffers)...};
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = static_cast<float*>(ptrs[d]);
      m_view.weight = static_cast<float*>(ptrs[Ndim]);
      m_view.cluster_index = static_cast<int*>(ptrs[Ndim + 1]);
      m_view.n = n_points;
    }
    template <std::ranges::contiguous_range... TRanges>
      requires(sizeof...(TRanges) == Ndim + 2 && Ndim > 1)
    PointsHost(Dim<Ndim> dim, int32_t n_points, TRanges&&... buffers)
        : PointsHost(dim, n_points, std::ranges::data(buffers)...) {}

    PointsHost(const PointsHost&) = delete;
    PointsHost& operator=(const PointsHost&) = delete;
    PointsHost(PointsHost&&) = default;
    PointsHost& operator=(PointsHost&&) = default;

    int32_t size() const { return m_size; }
    std::span<float> coords(std::size_t dim) {
      if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
      return {m_view.coords[dim], static_cast<std::size_t>(m_size)};
    }
    std::span<const float> coords(std::size_t dim) const {
      if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
      return {m_view.coords[dim], static_cast<std::size_t>(m_size)};
    }
    std::span<float> weights() { return {m_view.weight, static_cast<std::size_t>(m_size)}; }
    std::span<const float> weights() const { return {m_view.weight, static_cast<std::size_t>(m_size)}; }
    std::span<int> clusterIndexes() {
      assert(m_clustered && "The points have not been clustered yet");
      return {m_view.cluster_index, static_cast<std::size_t>(m_size)};
    }
    std::span<const int> clusterIndexes() const {
      assert(m_clustered && "The points have not been clustered yet");
      return {m_view.cluster_index, static_cast<std::size_t>(m_size)};
    }
    bool clustered() const { return m_clustered; }
    const PointsView<Ndim>& view() const { return m_view; }
    PointsView<Ndim>& view() { return m_view; }

    Point operator[](std::size_t idx) const {
      if (idx >= static_cast<std::size_t>(m_size)) throw std::out_of_range("Index out of range in PointsHost::operator[]");
      std::array<float, Ndim> c;
      for (std::size_t d = 0; d < Ndim; ++d) c[d] = m_view.coords[d][idx];
      return Point(c, m_view.weight[idx], m_view.cluster_index[idx]);
    }

    /// max label + 1 (utils/detail/get_cluster_properties.hpp)
    const std::size_t& n_clusters() {
      assert(m_clustered);
      if (!m_nclusters) {
        int32_t mx = -1;
        for (int32_t i = 0; i < m_size; ++i) mx = std::max(mx, m_view.cluster_index[i]);
        m_nclusters = static_cast<std::size_t>(mx + 1);
      }
      return *m_nclusters;
    }
    const HostAssociationMap& clusters() { return cluster_properties().clusters(); }
    const std::vector<std::size_t>& cluster_sizes() { return cluster_properties().cluster_sizes(); }
    const ClusterProperties& cluster_properties() {
      assert(m_clustered);
      if (!m_props) m_props.emplace(std::span<const int32_t>(m_view.cluster_index, static_cast<std::size_t>(m_size)));
      return *m_props;
    }

    void mark_clustered() {
      m_clustered = true;
      m_nclusters.reset();
      m_props.reset();
    }

  private:
    static int32_t checked(int32_t n) {
      if (n <= 0) throw std::invalid_argument("Number of points passed to PointsHost constructor must be positive.");
      return n;
    }
    static std::size_t soa_bytes(int32_t n) { return ((Ndim + 1) * sizeof(float) + sizeof(int)) * static_cast<std::size_t>(n); }
    void partition(std::byte* b) {
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = reinterpret_cast<float*>(b + d * m_size * sizeof(float));
      m_view.weight = reinterpret_cast<float*>(b + Ndim * m_size * sizeof(float));
      m_view.cluster_index = reinterpret_cast<int*>(b + (Ndim + 1) * m_size * sizeof(float));
      m_view.n = m_size;
    }

    std::unique_ptr<std::byte[]> m_owned;
    PointsView<Ndim> m_view;
    int32_t m_size;
    bool m_clustered = false;
    std::optional<std::size_t> m_nclusters;
    std::optional<ClusterProperties> m_props;
  };

}  // namespace clue
