// A C++ caller of the drop-in API, written the way the reference's callers
// use cublox::OccupancyGrid: its perf test (test/test_occupancy_grid_perf.cpp:
// 54-100: a 401x401x201 grid at 0.05 m with centre binning and no recenter
// threshold, a 500,000-point spherical shell from mt19937 seed 42, one
// warm-up update and timed updates, grid.config_ read for the report) and its
// driver (src/main.cpp:437-498: logit() log-odds, setMaxRaycastRange,
// update from the sensor origin, recenter on the base position,
// fetchOccupancyAround with the OCCUPIED filter, point queries).
//
// It includes only <cublox/OccupancyGrid.hpp> and <cublox/utils.hpp> and
// links libcublox.so. Everything it feeds the grids and everything it reads
// back is written to the output directory given as argv[1], so a test can
// replay the same calls on the oracle:
//   a_cloud.f32   the shell cloud, N x 3 column-major (PointCloud storage)
//   a_occ.f32     grid A's occupancy after the updates (hash order)
//   b_cloud<k>.f32, b_meta.f32 (sensor xyz, base xyz per scan)
//   b_occ.f32, b_fetch<q>.f32 / .u8 (positions, states),
//   b_queries.f32 / .u8 (positions; isOccupied, isFree, isUnknown)
#include <cublox/OccupancyGrid.hpp>
#include <cublox/utils.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <optional>
#include <random>
#include <string>
#include <vector>

namespace {

std::string g_dir;

template <typename T>
void dump(const std::string &name, const T *p, size_t count) {
  std::ofstream f(g_dir + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), static_cast<std::streamsize>(count * sizeof(T)));
}

// Points uniform in radius on [r0, r1] and uniform on the sphere.
cublox::PointCloud shell(int n, float r0, float r1, unsigned seed) {
  cublox::PointCloud pc(n, 3);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> radius(r0, r1), cz(-1.0f, 1.0f),
      az(0.0f, 6.2831853f);
  for (int i = 0; i < n; ++i) {
    const float r = radius(gen), c = cz(gen), a = az(gen);
    const float s = std::sqrt(std::max(0.0f, 1.0f - c * c));
    pc(i, 0) = r * s * std::cos(a);
    pc(i, 1) = r * s * std::sin(a);
    pc(i, 2) = r * c;
  }
  return pc;
}

// A ring scan around `sensor`: 16 elevations x 360 azimuths, ranges from a
// wall at 6 m (x) / 7 m (y) and a ground plane 1.8 m below the sensor;
// beams past 10 m end at 11 m (free space only once clipped).
cublox::PointCloud ring(const Eigen::Vector3f &sensor) {
  const int ne = 16, na = 360;
  cublox::PointCloud pc(ne * na, 3);
  for (int e = 0; e < ne; ++e) {
    const float el = (-15.0f + 2.0f * static_cast<float>(e)) * 0.017453292f;
    for (int a = 0; a < na; ++a) {
      const float th = static_cast<float>(a) * 0.017453292f;
      const float dx = std::cos(el) * std::cos(th), dy = std::cos(el) * std::sin(th),
                  dz = std::sin(el);
      float t = 11.0f;
      if (std::fabs(dx) > 1e-6f) t = std::min(t, std::fabs(6.0f / dx));
      if (std::fabs(dy) > 1e-6f) t = std::min(t, std::fabs(7.0f / dy));
      if (dz < -1e-6f) t = std::min(t, 1.8f / -dz);
      const int i = e * na + a;
      pc(i, 0) = sensor.x() + t * dx;
      pc(i, 1) = sensor.y() + t * dy;
      pc(i, 2) = sensor.z() + t * dz;
    }
  }
  return pc;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT_DIR [timed updates]\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  const int timed = argc > 2 ? std::atoi(argv[2]) : 3;

  // ---- A: the perf test's grid and cloud -------------------------------
  cublox::OccupancyGrid grid(Eigen::Vector3i(200, 200, 100), 0.05f, true, std::nullopt,
                             Eigen::Vector3f::Zero());
  const int n_a = 500000;
  const cublox::PointCloud cloud = shell(n_a, 1.5f, 8.0f, 42);
  dump("a_cloud.f32", cloud.data(), static_cast<size_t>(n_a) * 3);
  auto once = [&] {
    const auto t0 = std::chrono::steady_clock::now();
    grid.update(cloud, Eigen::Vector3f::Zero());
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
  };
  std::printf("grid: %d x %d x %d voxels (%d) at %.2f m\n", grid.config_.map_size_i.x(),
              grid.config_.map_size_i.y(), grid.config_.map_size_i.z(),
              grid.config_.voxel_num, static_cast<double>(grid.config_.resolution));
  std::printf("warm-up update(): %.3f ms\n", once());
  double sum = 0.0, lo = std::numeric_limits<double>::max(), hi = 0.0;
  for (int k = 0; k < timed; ++k) {
    const double ms = once();
    sum += ms;
    lo = std::min(lo, ms);
    hi = std::max(hi, ms);
  }
  if (timed > 0) {
    std::printf("%d updates of %d points: mean %.3f ms, min %.3f, max %.3f\n", timed, n_a,
                sum / timed, lo, hi);
  }
  {
    std::vector<float> occ(static_cast<size_t>(grid.config_.voxel_num));
    grid.copyOccupancyToHost(occ.data());
    dump("a_occ.f32", occ.data(), occ.size());
  }

  // ---- B: the driver's call pattern -------------------------------------
  cublox::OccupancyGrid map(Eigen::Vector3i(100, 100, 25), 0.1f, false, 2.0,
                            Eigen::Vector3f::Zero());
  map.setMaxRaycastRange(10.0f);
  map.setLogOddsParams(cublox::logit(0.9f), cublox::logit(0.49f), cublox::logit(0.12f),
                       cublox::logit(0.98f), cublox::logit(0.499f), cublox::logit(0.85f));
  std::vector<float> meta;
  int moves = 0;
  for (int k = 0; k < 8; ++k) {
    const Eigen::Vector3f base(0.6f * static_cast<float>(k), 0.2f * static_cast<float>(k),
                               0.0f);
    const Eigen::Vector3f sensor = base + Eigen::Vector3f(0.1f, 0.0f, 1.8f);
    const cublox::PointCloud pc = ring(sensor);
    dump("b_cloud" + std::to_string(k) + ".f32", pc.data(),
         static_cast<size_t>(pc.rows()) * 3);
    for (int a = 0; a < 3; ++a) meta.push_back(sensor(a));
    for (int a = 0; a < 3; ++a) meta.push_back(base(a));
    map.update(pc, sensor);
    const Eigen::Vector3i before = map.getOriginIndex();
    map.recenter(base);
    moves += map.getOriginIndex() != before ? 1 : 0;
  }
  dump("b_meta.f32", meta.data(), meta.size());
  {
    std::vector<float> occ(static_cast<size_t>(map.config_.voxel_num));
    map.copyOccupancyToHost(occ.data());
    dump("b_occ.f32", occ.data(), occ.size());
  }
  const Eigen::Vector3f robot(4.2f, 1.4f, 0.0f);
  const std::optional<cublox::OccupancyGrid::VoxelState> filters[2] = {
      cublox::OccupancyGrid::VoxelState::OCCUPIED, std::nullopt};
  for (int q = 0; q < 2; ++q) {
    const auto samples = map.fetchOccupancyAround(robot, 3.0f, filters[q]);
    std::vector<float> pos;
    std::vector<unsigned char> st;
    for (const auto &s : samples) {
      pos.push_back(s.position.x());
      pos.push_back(s.position.y());
      pos.push_back(s.position.z());
      st.push_back(static_cast<unsigned char>(s.state));
    }
    dump("b_fetch" + std::to_string(q) + ".f32", pos.data(), pos.size());
    dump("b_fetch" + std::to_string(q) + ".u8", st.data(), st.size());
  }
  std::vector<float> qpos;
  std::vector<unsigned char> qres;
  for (int i = 0; i < 60; ++i) {
    const Eigen::Vector3f p(-2.0f + 0.37f * static_cast<float>(i),
                            1.4f + 0.05f * static_cast<float>(i),
                            0.05f + 0.03f * static_cast<float>(i % 7));
    for (int a = 0; a < 3; ++a) qpos.push_back(p(a));
    qres.push_back(map.isOccupied(p) ? 1 : 0);
    qres.push_back(map.isFree(p) ? 1 : 0);
    qres.push_back(map.isUnknown(p) ? 1 : 0);
  }
  dump("b_queries.f32", qpos.data(), qpos.size());
  dump("b_queries.u8", qres.data(), qres.size());
  std::printf("recenters: %d; origin %d %d %d\n", moves, map.getOriginIndex().x(),
              map.getOriginIndex().y(), map.getOriginIndex().z());
  return 0;
}
