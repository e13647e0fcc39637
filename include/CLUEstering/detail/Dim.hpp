// Dimension tag of the drop-in API (reference: detail/Dim.hpp:4-7).
#pragma once
#include <cstddef>

namespace clue {
  template <std::size_t N>
  struct Dim {
    static constexpr std::size_t value = N;
  };
}  // namespace clue
