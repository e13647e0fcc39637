// Device points of the drop-in API (reference: data_structures/PointsDevice.hpp:
// 19-271, detail/PointsDevice.hpp).  SoA in HBM: [x_0]...[x_{D-1}][weight]
// [cluster][is_seed][rho][nearest_higher]; borrowed inputs keep only the three
// result columns in an owned buffer.  Spans returned by the accessors point to
// device memory.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <vector>

#include "CLUEstering/core/Device.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"
#include "CLUEstering/detail/Dim.hpp"

namespace clue {

  namespace detail {
    struct DeviceBuffer {
      void* ptr = nullptr;
      DeviceBuffer(int32_t device, std::size_t bytes) { check(clue_malloc_device(device, bytes, &ptr)); }
      ~DeviceBuffer() { clue_free_device(ptr); }
      DeviceBuffer(const DeviceBuffer&) = delete;
      DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    };
  }  // namespace detail

  template <typename TDev, std::size_t Ndim>
  class PointsDevice {
  public:
    PointsDevice(TDev& device, Dim<Ndim>, int32_t n_points) : m_device{device}, m_size{checked(n_points)} {
      m_owned = std::make_shared<detail::DeviceBuffer>(device.index(), soa_bytes(n_points));
      partition(static_cast<std::byte*>(m_owned->ptr));
    }
    PointsDevice(TDev& device, Dim<Ndim>, int32_t n_points, std::span<std::byte> buffer)
        : m_device{device}, m_size{checked(n_points)} {
      if (buffer.size() < soa_bytes(n_points)) throw std::invalid_argument("buffer too small for the points");
      partition(buffer.data());
    }
    PointsDevice(TDev& device, Dim<Ndim> dim, int32_t n_points, std::span<float> input, std::span<int> output)
        : PointsDevice(device, dim, n_points, input.data(), output.data()) {}
    PointsDevice(TDev& device, Dim<Ndim> dim, int32_t n_points, std::span<float> coordinates,
                 std::span<float> weights, std::span<int> output)
        : PointsDevice(device, dim, n_points, coordinates.data(), weights.data(), output.data()) {}
    /// input: device SoA coordinates followed by the weights; output: device labels.
    PointsDevice(TDev& device, Dim<Ndim>, int32_t n_points, float* input, int* output)
        : m_device{device}, m_size{checked(n_points)} {
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = input + d * n_points;
      m_view.weight = input + Ndim * n_points;
      m_view.cluster_index = output;
      own_results();
    }
    PointsDevice(TDev& device, Dim<Ndim>, int32_t n_points, float* coordinates, float* weights, int* output)
        : m_device{device}, m_size{checked(n_points)} {
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = coordinates + d * n_points;
      m_view.weight = weights;
      m_view.cluster_index = output;
      own_results();
    }
    template <typename... TPtrs>
      requires(sizeof...(TPtrs) == Ndim + 2 && Ndim > 1 && (std::is_pointer_v<TPtrs> && ...))
    PointsDevice(TDev& device, Dim<Ndim>, int32_t n_points, TPtrs... buffers)
        : m_device{device}, m_size{checked(n_points)} {
      void* ptrs[] = {static_cast<void*>(buffers)...};
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = static_cast<float*>(ptrs[d]);
      m_view.weight = static_cast<float*>(ptrs[Ndim]);
      m_view.cluster_index = static_cast<int*>(ptrs[Ndim + 1]);
      own_results();
    }

    PointsDevice(const PointsDevice&) = delete;
    PointsDevice& operator=(const PointsDevice&) = delete;
    PointsDevice(PointsDevice&&) = default;
    PointsDevice& operator=(PointsDevice&&) = default;

    int32_t size() const { return m_size; }
    const TDev& device() const { return m_device; }
    std::span<float> coords(std::size_t dim) const {
      if (dim >= Ndim) throw std::out_of_range("Dimension out of range in call to coords.");
      return {m_view.coords[dim], static_cast<std::size_t>(m_size)};
    }
    std::span<float> weights() const { return {m_view.weight, static_cast<std::size_t>(m_size)}; }
    std::span<int> clusterIndexes() const { return {m_view.cluster_index, static_cast<std::size_t>(m_size)}; }
    std::span<float> rho() const { return {m_view.rho, static_cast<std::size_t>(m_size)}; }
    std::span<int> nearestHigher() const { return {m_view.nearest_higher, static_cast<std::size_t>(m_size)}; }
    std::span<int> isSeed() const { return {m_view.is_seed, static_cast<std::size_t>(m_size)}; }
    bool clustered() const { return m_clustered; }
    const PointsView<Ndim>& view() const { return m_view; }
    PointsView<Ndim>& view() { return m_view; }

    /// Number of clusters: max label + 1 (copies the labels to the host).
    const std::size_t& n_clusters() {
      if (!m_nclusters) {
        std::vector<int32_t> labels(static_cast<std::size_t>(m_size));
        detail::check(clue_memcpy_d2h(labels.data(), m_view.cluster_index, labels.size() * 4, nullptr));
        detail::check(clue_stream_synchronize(nullptr));
        int32_t mx = -1;
        for (int32_t v : labels) mx = std::max(mx, v);
        m_nclusters = static_cast<std::size_t>(mx + 1);
      }
      return *m_nclusters;
    }

    void mark_clustered() {
      m_clustered = true;
      m_nclusters.reset();
    }

  private:
    static int32_t checked(int32_t n) {
      if (n <= 0) throw std::invalid_argument("Number of points passed to PointsDevice constructor must be positive.");
      return n;
    }
    static std::size_t soa_bytes(int32_t n) { return ((Ndim + 2) * sizeof(float) + 3 * sizeof(int)) * static_cast<std::size_t>(n); }
    void partition(std::byte* b) {
      const std::size_t col = static_cast<std::size_t>(m_size) * 4;
      for (std::size_t d = 0; d < Ndim; ++d) m_view.coords[d] = reinterpret_cast<float*>(b + d * col);
      m_view.weight = reinterpret_cast<float*>(b + Ndim * col);
      m_view.cluster_index = reinterpret_cast<int*>(b + (Ndim + 1) * col);
      m_view.is_seed = reinterpret_cast<int*>(b + (Ndim + 2) * col);
      m_view.rho = reinterpret_cast<float*>(b + (Ndim + 3) * col);
      m_view.nearest_higher = reinterpret_cast<int*>(b + (Ndim + 4) * col);
      m_view.n = m_size;
    }
    void own_results() {
      const std::size_t col = static_cast<std::size_t>(m_size) * 4;
      m_owned = std::make_shared<detail::DeviceBuffer>(m_device.index(), 3 * col);
      auto* b = static_cast<std::byte*>(m_owned->ptr);
      m_view.is_seed = reinterpret_cast<int*>(b);
      m_view.rho = reinterpret_cast<float*>(b + col);
      m_view.nearest_higher = reinterpret_cast<int*>(b + 2 * col);
      m_view.n = m_size;
    }

    TDev m_device;
    std::shared_ptr<detail::DeviceBuffer> m_owned;
    PointsView<Ndim> m_view;
    int32_t m_size;
    bool m_clustered = false;
    std::optional<std::size_t> m_nclusters;
  };

  template <typename TDev, std::size_t Ndim>
  PointsDevice(TDev&, Dim<Ndim>, int32_t) -> PointsDevice<TDev, Ndim>;
  template <typename TDev, std::size_t Ndim, typename... Rest>
  PointsDevice(TDev&, Dim<Ndim>, int32_t, Rest...) -> PointsDevice<TDev, Ndim>;

  /// PointsDevice.hpp:65-99: coordinates, weights and labels host -> device.
  template <concepts::Queue TQueue, std::size_t Ndim>
  void copyToDevice(TQueue& queue, PointsDevice<DevType<TQueue>, Ndim>& d_points, const PointsHost<Ndim>& h_points) {
    if (d_points.size() != h_points.size())
      throw std::invalid_argument("Host and device points have different sizes");
    const std::size_t bytes = static_cast<std::size_t>(h_points.size()) * 4;
    for (std::size_t d = 0; d < Ndim; ++d)
      detail::check(clue_memcpy_h2d(d_points.view().coords[d], h_points.view().coords[d], bytes, queue.handle()));
    detail::check(clue_memcpy_h2d(d_points.view().weight, h_points.view().weight, bytes, queue.handle()));
  }

  /// PointsDevice.hpp:19-57: coordinates, weights and labels device -> host.
  template <concepts::Queue TQueue, std::size_t Ndim>
  void copyToHost(TQueue& queue, PointsHost<Ndim>& h_points, const PointsDevice<DevType<TQueue>, Ndim>& d_points) {
    if (d_points.size() != h_points.size())
      throw std::invalid_argument("Host and device points have different sizes");
    const std::size_t bytes = static_cast<std::size_t>(h_points.size()) * 4;
    for (std::size_t d = 0; d < Ndim; ++d)
      detail::check(clue_memcpy_d2h(h_points.view().coords[d], d_points.view().coords[d], bytes, queue.handle()));
    detail::check(clue_memcpy_d2h(h_points.view().weight, d_points.view().weight, bytes, queue.handle()));
    detail::check(clue_memcpy_d2h(h_points.view().cluster_index, d_points.view().cluster_index, bytes, queue.handle()));
  }

  template <concepts::Queue TQueue, std::size_t Ndim>
  PointsDevice<DevType<TQueue>, Ndim> copyToDevice(TQueue& queue, const PointsHost<Ndim>& h_points) {
    auto dev = queue.getDevice();
    PointsDevice<DevType<TQueue>, Ndim> d(dev, Dim<Ndim>{}, h_points.size());
    copyToDevice(queue, d, h_points);
    return d;
  }

}  // namespace clue
