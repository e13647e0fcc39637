// Device cluster -> points map (reference: DevAssociationMap filled by
// internal/MakeAssociator.hpp:13-35).  Built on the device by clue_cluster_map:
// CSR offsets and point indexes stay in HBM; size(), count() and toHost() copy
// what they need once, after the map's queue has finished.
#pragma once
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <vector>

#include "CLUEstering/core/Device.hpp"
#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/detail/Status.hpp"

namespace clue {

  class DeviceAssociationMap {
  public:
    using key_type = int32_t;
    using mapped_type = int32_t;
    using size_type = std::size_t;

    /// Map of n device labels (negative: outlier), built on `queue` (asynchronous).
    DeviceAssociationMap(const Queue& queue, int32_t n, const int32_t* labels_dev)
        : m_queue{queue}, m_state{std::make_shared<State>(queue.getDevice().index(), n)} {
      detail::check(clue_cluster_map(m_state->plan, queue.handle(), n, labels_dev, m_state->offsets(),
                                     m_state->indexes(), m_state->counts()));
    }

    /// Number of clusters (largest label + 1).
    size_type size() const { return static_cast<size_type>(host().counts[0]); }
    /// Points in clusters (labels >= 0).
    size_type values() const { return static_cast<size_type>(host().counts[1]); }
    size_type count(key_type key) const {
      const auto& h = host();
      if (key < 0 || key >= h.counts[0]) throw std::out_of_range("AssociationMap::count key out of range");
      return static_cast<size_type>(h.offsets[key + 1] - h.offsets[key]);
    }
    /// Device arrays: offsets (size()+1 valid entries), indexes (values() valid
    /// entries, each cluster's points ascending), counts {size(), values()}.
    const int32_t* offsets() const { return m_state->offsets(); }
    const int32_t* indexes() const { return m_state->indexes(); }
    const int32_t* counts() const { return m_state->counts(); }
    clue_plan_t plan() const { return m_state->plan; }

    HostAssociationMap toHost() const {
      const auto& h = host();
      HostAssociationMap m(static_cast<size_type>(h.counts[1]), static_cast<size_type>(h.counts[0]));
      // rebuild through fill() from (cluster, point) pairs: identical CSR
      std::vector<int32_t> assoc(static_cast<size_type>(m_state->n), -1);
      for (int32_t c = 0; c < h.counts[0]; ++c)
        for (int32_t k = h.offsets[c]; k < h.offsets[c + 1]; ++k) assoc[h.idx[k]] = c;
      m.fill(assoc);
      return m;
    }

  private:
    struct State {
      int32_t n;
      clue_plan_t plan = nullptr;
      void* buf = nullptr;  // offsets (n+1) | indexes (n) | counts (2)
      State(int32_t device, int32_t n_) : n{n_} {
        if (n <= 0) throw std::invalid_argument("make_associator: elements and associations must be non-zero");
        detail::check(clue_plan_create(device, &plan));
        const clue_status st = clue_malloc_device(device, (static_cast<size_t>(n) * 2 + 3) * 4, &buf);
        if (st != CLUE_OK) {
          clue_plan_destroy(plan);
          detail::check(st);
        }
      }
      ~State() {
        clue_plan_destroy(plan);
        clue_free_device(buf);
      }
      int32_t* offsets() const { return static_cast<int32_t*>(buf); }
      int32_t* indexes() const { return offsets() + n + 1; }
      int32_t* counts() const { return indexes() + n; }
    };
    struct Host {
      int32_t counts[2];
      std::vector<int32_t> offsets, idx;
    };
    const Host& host() const {
      if (!m_host) {
        Host h;
        m_queue.wait();
        detail::check(clue_memcpy_d2h(h.counts, m_state->counts(), 8, m_queue.handle()));
        m_queue.wait();
        h.offsets.resize(static_cast<size_type>(h.counts[0]) + 1);
        h.idx.resize(static_cast<size_type>(h.counts[1]));
        detail::check(clue_memcpy_d2h(h.offsets.data(), m_state->offsets(), h.offsets.size() * 4, m_queue.handle()));
        if (!h.idx.empty())
          detail::check(clue_memcpy_d2h(h.idx.data(), m_state->indexes(), h.idx.size() * 4, m_queue.handle()));
        m_queue.wait();
        m_host = std::move(h);
      }
      return *m_host;
    }

    Queue m_queue;
    std::shared_ptr<State> m_state;
    mutable std::optional<Host> m_host;
  };

}  // namespace clue
