// CSV reader (reference: utils/detail/read_csv.hpp:17-45): a header line, then
// per line the coordinates and the weight, parsed with std::stof.
#pragma once
#include <algorithm>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>

#include "CLUEstering/data_structures/PointsHost.hpp"

namespace clue {
  template <std::size_t NDim>
  PointsHost<NDim> read_csv(Dim<NDim> dim, const std::string& file_path) {
    std::ifstream file(file_path);
    if (!file.is_open()) throw std::runtime_error("Could not open file: " + file_path);
    std::string line;
    std::getline(file, line);  // header
    std::vector<std::string> lines;
    while (std::getline(file, line))
      if (!line.empty() && line != "\r") lines.push_back(line);
    PointsHost<NDim> points(dim, static_cast<int32_t>(lines.size()));
    for (std::size_t i = 0; i < lines.size(); ++i) {
      std::stringstream ss(lines[i]);
      std::string value;
      for (std::size_t d = 0; d < NDim; ++d) {
        std::getline(ss, value, ',');
        points.coords(d)[i] = std::stof(value);
      }
      std::getline(ss, value);
      points.weights()[i] = std::stof(value);
    }
    return points;
  }
}  // namespace clue
