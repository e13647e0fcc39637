// Exercises the C++20 drop-in API the way the reference's doctest suite does
// (tests/test_clusterer.cpp, test_host_points.cpp, test_device_points.cpp,
// test_distance_metrics.cpp, test_convolutional_kernels.cpp).
//
//   test_api unit                         host-only checks (no GPU)
//   test_api cluster <csv> <ndim> <dc> <rhoc> <dm> <overload> <labels.bin>
//       overload: host | host_queue | host_device | device
//   test_api capacity <csv> <csv2> <dc> <rhoc> <dm> <seed_dc> <out.bin>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CLUEstering/CLUEstering.hpp"

static int g_fail = 0;
#define CHECK(x)                                                         \
  do {                                                                   \
    if (!(x)) {                                                          \
      std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #x); \
      ++g_fail;                                                          \
    }                                                                    \
  } while (0)
#define CHECK_THROWS_AS(expr, T)                                         \
  do {                                                                   \
    bool thrown_ = false;                                                \
    try {                                                                \
      (void)(expr);                                                      \
    } catch (const T&) {                                                 \
      thrown_ = true;                                                    \
    } catch (...) {                                                      \
    }                                                                    \
    CHECK(thrown_);                                                      \
  } while (0)

static void write_labels(const std::string& path, std::span<const int> labels) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(labels.data()), static_cast<std::streamsize>(labels.size() * 4));
}

static int unit() {
  // test_distance_metrics.cpp
  {
    std::array<float, 3> p1{1.f, 2.f, 0.f}, p2{4.f, 6.f, 0.f};
    CHECK(clue::metrics::Euclidean<2>()(p1, p2) == 5.f);
    CHECK(clue::metrics::WeightedEuclidean<2>(1.f, 2.f)(p1, p2) == std::sqrt(41.f));
    CHECK(clue::metrics::WeightedEuclidean<2>(std::array<float, 2>{1.f, 2.f})(p1, p2) == std::sqrt(41.f));
    std::array<float, 3> m1{-1.f, 3.f, 0.f}, m2{0.f, 6.f, 0.f};
    CHECK(clue::metrics::Manhattan<2>()(m1, m2) == 4.f);
    std::array<float, 3> c1{-1.f, 4.f, 0.f}, c2{0.f, 6.f, 0.f};
    CHECK(clue::metrics::Chebyshev<2>()(c1, c2) == 2.f);
    CHECK(clue::metrics::WeightedChebyshev<2>(1.f, 2.f)(c1, c2) == 4.f);
  }
  // test_convolutional_kernels.cpp: the throwing conditions
  CHECK_THROWS_AS(clue::FlatKernel(-1.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::FlatKernel(0.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::GaussianKernel(0.f, 1.f, 1.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::GaussianKernel(1.f, -1.f, 1.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::GaussianKernel(1.f, 1.f, 0.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::ExponentialKernel(-1.f, 1.f), std::invalid_argument);
  CHECK_THROWS_AS(clue::ExponentialKernel(1.f, 0.f), std::invalid_argument);
  CHECK(clue::GaussianKernel::unchecked(0.f, 1.f, 1.f).desc().params[0] == 0.f);
  CHECK(clue::FlatKernel(.5f)(0.f, 3, 3) == 1.f);
  CHECK(clue::FlatKernel(.5f)(0.f, 3, 4) == .5f);
  // test_host_points.cpp: constructors, accessors, throws
  {
    const int32_t n = 1000;
    clue::PointsHost<2> points(clue::Dim<2>{}, n);
    CHECK(points.size() == n);
    for (int i = 0; i < n; ++i) {
      points.coords(0)[i] = static_cast<float>(i);
      points.coords(1)[i] = static_cast<float>(2 * i);
      points.weights()[i] = 1.f;
    }
    CHECK(points[10][1] == 20.f);
    CHECK_THROWS_AS(points.coords(2), std::out_of_range);
    CHECK_THROWS_AS(points[n], std::out_of_range);
    CHECK_THROWS_AS(clue::PointsHost<2>(clue::Dim<2>{}, 0), std::invalid_argument);
    std::vector<float> input(3 * n);
    std::vector<int> output(n);
    clue::PointsHost<2> borrowed(clue::Dim<2>{}, n, input.data(), output.data());
    CHECK(borrowed.coords(1).data() == input.data() + n);
    CHECK(borrowed.weights().data() == input.data() + 2 * n);
    std::vector<float> x(n), y(n), w(n);
    clue::PointsHost<2> separate(clue::Dim<2>{}, n, x.data(), y.data(), w.data(), output.data());
    CHECK(separate.coords(1).data() == y.data());
    CHECK(separate.weights().data() == w.data());
    std::vector<std::byte> bytes(n * 4 * 4);
    clue::PointsHost<2> from_bytes(clue::Dim<2>{}, n, std::span<std::byte>(bytes));
    CHECK(reinterpret_cast<std::byte*>(from_bytes.weights().data()) == bytes.data() + 2 * n * 4);
  }
  // test_association_map.cpp: a binary map of 1000 elements
  {
    std::vector<int32_t> keys(1000);
    for (int i = 0; i < 1000; ++i) keys[i] = i % 2;
    clue::HostAssociationMap map(1000, 2);
    map.fill(keys);
    CHECK(map.size() == 2);
    CHECK(map.count(0) == 500 && map.count(1) == 500);
    CHECK(*map.lower_bound(1) == 1);
    CHECK_THROWS_AS(map.count(2), std::out_of_range);
    CHECK_THROWS_AS(map.count(-1), std::out_of_range);
  }
  // Clusterer parameter validation (detail/Clusterer.hpp:32-35)
  std::printf("unit checks: %d failures\n", g_fail);
  return g_fail ? 1 : 0;
}

template <std::size_t N>
static int cluster(const std::string& csv, float dc, float rhoc, float dm, const std::string& overload,
                   const std::string& out) {
  clue::Dim<N> dim{};
  clue::PointsHost<N> h_points = clue::read_csv(dim, csv);
  auto device = clue::DevicePool::deviceAt(0u);
  auto queue = device.makeQueue();
  clue::Clusterer algo(queue, dim, dc, rhoc, dm);
  CHECK_THROWS_AS(clue::Clusterer(queue, dim, -1.f, rhoc, dm), std::invalid_argument);
  if (overload == "host") {
    algo.make_clusters(h_points);
  } else if (overload == "host_queue") {
    algo.make_clusters(queue, h_points);
  } else if (overload == "host_device") {
    clue::PointsDevice d_points{device, dim, h_points.size()};
    algo.make_clusters(queue, h_points, d_points);
    CHECK(d_points.n_clusters() == h_points.n_clusters());
    auto dc_map = algo.getClusters(queue, d_points);
    CHECK(dc_map.size() == h_points.n_clusters());
  } else if (overload == "device") {
    clue::PointsDevice d_points{device, dim, h_points.size()};
    clue::copyToDevice(queue, d_points, h_points);
    algo.make_clusters(queue, d_points);
    std::vector<int> labels(h_points.size());
    clue::detail::check(clue_memcpy_d2h(labels.data(), d_points.clusterIndexes().data(), labels.size() * 4,
                                        queue.handle()));
    clue::wait(queue);
    write_labels(out, labels);
    // device post-processing against the host versions
    auto dmap = clue::get_clusters(queue, d_points);
    auto hmap = clue::detail::get_clusters(std::span<const int>(labels));
    CHECK(dmap.size() == hmap.size());
    const auto back = dmap.toHost();
    CHECK(back.offsets() == hmap.offsets());
    CHECK(back.indexes() == hmap.indexes());
    for (std::size_t c = 0; c < hmap.size(); ++c) CHECK(dmap.count(static_cast<int32_t>(c)) == hmap.count(static_cast<int32_t>(c)));
    h_points.mark_clustered();
    std::copy(labels.begin(), labels.end(), h_points.clusterIndexes().begin());
    h_points.mark_clustered();  // drop properties cached before the copy
    const auto hcent = clue::cluster_centroids(h_points);
    const auto dcent = clue::cluster_centroids(queue, d_points);
    CHECK(hcent.size() == dcent.size());
    for (std::size_t c = 0; c < hcent.size() && c < dcent.size(); ++c)
      CHECK(std::memcmp(hcent[c].data(), dcent[c].data(), sizeof(float) * N) == 0);
    if (!hcent.empty()) {
      const auto one = clue::cluster_centroid(h_points, 0);
      CHECK(std::memcmp(one.data(), hcent[0].data(), sizeof(float) * N) == 0);
    }
    std::printf("clusters: %zu\n", d_points.n_clusters());
    return g_fail ? 1 : 0;
  } else {
    std::fprintf(stderr, "unknown overload %s\n", overload.c_str());
    return 2;
  }
  CHECK(h_points.clustered());
  const auto& sizes = h_points.cluster_sizes();
  std::size_t total = 0;
  for (auto s : sizes) total += s;
  std::size_t clustered = 0;
  for (int v : h_points.clusterIndexes()) clustered += v >= 0;
  CHECK(total == clustered);
  CHECK(h_points.clusters().size() == h_points.n_clusters());
  write_labels(out, h_points.clusterIndexes());
  std::printf("clusters: %zu silhouette: %.4f\n", h_points.n_clusters(), clue::silhouette(h_points));
  return g_fail ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  try {
    if (mode == "unit") return unit();
    if (mode == "cluster" && argc == 9) {
      const int ndim = std::stoi(argv[3]);
      const float dc = std::stof(argv[4]), rhoc = std::stof(argv[5]), dm = std::stof(argv[6]);
      switch (ndim) {
        case 2: return cluster<2>(argv[2], dc, rhoc, dm, argv[7], argv[8]);
        case 3: return cluster<3>(argv[2], dc, rhoc, dm, argv[7], argv[8]);
        default: return 2;
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 3;
  }
  return 2;
}
