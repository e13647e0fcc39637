// Devices and queues of the drop-in API.  A Queue is a HIP stream; the
// reference's alpaka device pool and static queues (internal/alpaka/devices.hpp:
// 108-133, utils/get_queue.hpp:14-27) map onto HIP devices and streams.
#pragma once
#include <concepts>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>

#include "CLUEstering/detail/Status.hpp"

namespace clue {

  class Queue;

  class Device {
  public:
    explicit Device(int32_t index = 0) : m_index{index} {}
    int32_t index() const { return m_index; }
    Queue makeQueue() const;
    bool operator==(const Device& other) const { return m_index == other.m_index; }

  private:
    int32_t m_index;
  };

  /// A HIP stream on one device (shared ownership; the last copy destroys it).
  class Queue {
  public:
    explicit Queue(const Device& device) : m_device{device} {
      void* s = nullptr;
      detail::check(clue_stream_create(device.index(), &s));
      m_stream = std::shared_ptr<void>(s, [](void* p) { clue_stream_destroy(p); });
    }
    const Device& getDevice() const { return m_device; }
    void* handle() const { return m_stream.get(); }
    void wait() const { detail::check(clue_stream_synchronize(m_stream.get())); }

  private:
    Device m_device;
    std::shared_ptr<void> m_stream;
  };

  inline Queue Device::makeQueue() const { return Queue{*this}; }

  /// Blocks until all work enqueued on the queue has finished (alpaka::onHost::wait).
  inline void wait(const Queue& queue) { queue.wait(); }

  struct DevicePool {
    static int32_t size() {
      int32_t n = 0;
      if (clue_device_count(&n) != CLUE_OK) return 0;
      return n;
    }
    /// devices.hpp:77-88: throws std::runtime_error("No device available!") without a GPU.
    static Device deviceAt(uint32_t i) {
      const int32_t n = size();
      if (n == 0) throw std::runtime_error("No device available!");
      if (static_cast<int32_t>(i) >= n) throw std::out_of_range("Device index out of range");
      return Device{static_cast<int32_t>(i)};
    }
  };

  /// One queue per device, shared by every caller (get_queue.hpp:14-19).
  inline Queue& get_queue(const Device& device) {
    static std::mutex mutex;
    static std::map<int32_t, std::unique_ptr<Queue>> queues;
    std::lock_guard<std::mutex> lock(mutex);
    auto& q = queues[device.index()];
    if (!q) q = std::make_unique<Queue>(device);
    return *q;
  }
  inline Queue& get_queue(uint32_t device_index) { return get_queue(DevicePool::deviceAt(device_index)); }

  namespace concepts {
    template <typename T>
    concept Queue = requires(const T& q) {
      { q.getDevice() };
      { q.handle() } -> std::convertible_to<void*>;
    };
  }  // namespace concepts

  template <typename TQueue>
  using DevType = std::remove_cvref_t<decltype(std::declval<TQueue>().getDevice())>;

}  // namespace clue
