// Host key -> values map in CSR form (offsets[keys+1], indexes[values]), filled
// by a counting sort in ascending value order (reference:
// data_structures/AssociationMap.hpp:24-75, detail/AssociationMapBase.hpp:33-150).
#pragma once
#include <cstdint>
#include <numeric>
#include <span>
#include <stdexcept>
#include <vector>

namespace clue {

  class HostAssociationMap {
  public:
    using key_type = int32_t;
    using mapped_type = int32_t;
    using size_type = std::size_t;
    using iterator = std::vector<mapped_type>::iterator;
    using const_iterator = std::vector<mapped_type>::const_iterator;
    struct Extents {
      size_type keys{}, values{};
    };

    HostAssociationMap(size_type nelements, size_type nbins) : m_offsets(nbins + 1, 0), m_indexes(nelements) {}

    /// Counting sort of `associations` (negative keys are skipped).
    void fill(std::span<const key_type> associations) {
      std::vector<key_type> sizes(size(), 0);
      for (key_type k : associations)
        if (k >= 0) ++sizes[k];
      m_offsets[0] = 0;
      std::inclusive_scan(sizes.begin(), sizes.end(), m_offsets.begin() + 1);
      std::vector<key_type> cursor(m_offsets.begin(), m_offsets.end() - 1);
      for (size_type i = 0; i < associations.size(); ++i)
        if (associations[i] >= 0) m_indexes[cursor[associations[i]]++] = static_cast<mapped_type>(i);
    }

    size_type size() const { return m_offsets.size() - 1; }
    Extents extents() const { return {size(), m_indexes.size()}; }
    size_type count(key_type key) const {
      check(key, "count");
      return static_cast<size_type>(m_offsets[key + 1] - m_offsets[key]);
    }
    bool contains(key_type key) const {
      check(key, "contains");
      return count(key) > 0;
    }
    const_iterator lower_bound(key_type key) const {
      check(key, "lower_bound");
      return m_indexes.begin() + m_offsets[key];
    }
    const_iterator upper_bound(key_type key) const {
      check(key, "upper_bound");
      return m_indexes.begin() + m_offsets[key + 1];
    }
    std::pair<const_iterator, const_iterator> equal_range(key_type key) const {
      return {lower_bound(key), upper_bound(key)};
    }
    std::span<const mapped_type> operator[](key_type key) const {
      check(key, "operator[]");
      return {m_indexes.data() + m_offsets[key], static_cast<size_type>(m_offsets[key + 1] - m_offsets[key])};
    }
    const std::vector<key_type>& offsets() const { return m_offsets; }
    const std::vector<mapped_type>& indexes() const { return m_indexes; }

  private:
    void check(key_type key, const char* what) const {
      if (key < 0 || static_cast<size_type>(key) >= size())
        throw std::out_of_range(std::string("AssociationMap::") + what + " key out of range");
    }
    std::vector<key_type> m_offsets;
    std::vector<mapped_type> m_indexes;
  };

  namespace detail {
    /// Cluster -> points map of a label array (utils/detail/get_clusters.hpp:21-24,
    /// internal/MakeAssociator.hpp:36-42).
    inline HostAssociationMap get_clusters(std::span<const int32_t> cluster_ids) {
      int32_t clustered = 0, mx = -1;
      for (int32_t c : cluster_ids) {
        clustered += c > -1;
        mx = c > mx ? c : mx;
      }
      if (clustered == 0 || cluster_ids.empty())
        throw std::invalid_argument("make_associator: elements and associations must be non-zero");
      HostAssociationMap map(static_cast<std::size_t>(clustered), static_cast<std::size_t>(mx + 1));
      map.fill(cluster_ids);
      return map;
    }
  }  // namespace detail

}  // namespace clue
