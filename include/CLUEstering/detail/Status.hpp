// Translation of the C ABI status codes into the exception types the reference
// throws (std::invalid_argument, std::out_of_range, std::runtime_error).
#pragma once
#include <stdexcept>
#include <string>

#include "clue_abi.h"

namespace clue::detail {
  inline void check(clue_status st) {
    if (st == CLUE_OK) return;
    const std::string msg = clue_last_error_message();
    switch (st) {
      case CLUE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
      case CLUE_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
      default: throw std::runtime_error(msg);
    }
  }
}  // namespace clue::detail
