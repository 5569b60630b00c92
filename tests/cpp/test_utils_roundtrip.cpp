// Index round trips of the drop-in host API (include/cublox/utils.hpp),
// mirroring the reference's test/test_utils_index_roundtrip.cpp:60-180:
// the same three map fixtures, seeds 42-46, 500 samples per invariant.
// Built with g++ against the Eigen subset; prints "<n> failures".
#include <cublox/Grid.hpp>
#include <cublox/utils.hpp>

#include <cstdio>
#include <random>
#include <vector>

namespace {

struct Fixture {
  Eigen::Vector3i half, map, origin;
  float res, inv;
  bool center;
};

Fixture make(const Eigen::Vector3i &half, const Eigen::Vector3i &origin,
             float res, bool center) {
  return Fixture{half, 2 * half + Eigen::Vector3i::Ones(), origin, res,
                 1.0f / res, center};
}

int g_fail = 0;
#define EXPECT_EQ_V(a, b, what)                                               \
  do {                                                                        \
    if (!((a) == (b))) {                                                      \
      ++g_fail;                                                               \
      if (g_fail < 10) std::printf("FAIL %s\n", what);                        \
    }                                                                         \
  } while (0)

Eigen::Vector3i randLocal(std::mt19937 &rng, const Fixture &f) {
  std::uniform_int_distribution<int> dx(-f.half.x(), f.half.x());
  std::uniform_int_distribution<int> dy(-f.half.y(), f.half.y());
  std::uniform_int_distribution<int> dz(-f.half.z(), f.half.z());
  const int x = dx(rng), y = dy(rng), z = dz(rng);
  return Eigen::Vector3i(x, y, z);
}

} // namespace

int main() {
  const std::vector<Fixture> fixtures = {
      make(Eigen::Vector3i(200, 200, 100), Eigen::Vector3i(17, -42, 8), 0.05f, true),
      make(Eigen::Vector3i(350, 350, 75), Eigen::Vector3i(-120, 55, 0), 0.1f, false),
      make(Eigen::Vector3i(3, 5, 7), Eigen::Vector3i(2, -1, 4), 0.25f, true),
  };
  for (const Fixture &f : fixtures) {
    {  // local -> global -> local
      std::mt19937 rng(42);
      for (int i = 0; i < 500; ++i) {
        const Eigen::Vector3i l = randLocal(rng, f);
        Eigen::Vector3i g, lb;
        cublox::localIndexToGlobalIndex(l, f.map, f.half, f.origin, g);
        cublox::globalIndexToLocalIndex(g, f.map, f.half, lb);
        EXPECT_EQ_V(l, lb, "GlobalLocalGlobal");
      }
    }
    {  // local -> hash -> local
      std::mt19937 rng(43);
      for (int i = 0; i < 500; ++i) {
        const Eigen::Vector3i l = randLocal(rng, f);
        Eigen::Vector3i lb;
        cublox::hashIdToLocalIndex(cublox::localIndexToHashId(l, f.map, f.half),
                                   f.map, f.half, lb);
        EXPECT_EQ_V(l, lb, "LocalHashLocal");
      }
    }
    {  // local -> global -> hash -> global
      std::mt19937 rng(44);
      for (int i = 0; i < 500; ++i) {
        const Eigen::Vector3i l = randLocal(rng, f);
        Eigen::Vector3i g, gb;
        cublox::localIndexToGlobalIndex(l, f.map, f.half, f.origin, g);
        const int h = cublox::globalIndexToHashId(g, f.map, f.half);
        cublox::hashIdToGlobalIndex(h, f.map, f.half, f.origin, gb);
        EXPECT_EQ_V(g, gb, "GlobalHashGlobal");
      }
    }
    {  // pos -> global -> pos -> global
      std::mt19937 rng(45);
      std::uniform_real_distribution<float> pd(-20.0f, 20.0f);
      for (int i = 0; i < 500; ++i) {
        const float px = pd(rng), py = pd(rng), pz = pd(rng);
        const Eigen::Vector3f p(px, py, pz);
        Eigen::Vector3i g, gb;
        Eigen::Vector3f pb;
        cublox::posToGlobalIndex(p, f.inv, f.center, g);
        cublox::globalIndexToPos(g, f.res, f.center, pb);
        cublox::posToGlobalIndex(pb, f.inv, f.center, gb);
        EXPECT_EQ_V(g, gb, "PosGlobalPos");
      }
    }
    {  // pos -> hash -> pos -> hash
      std::mt19937 rng(46);
      std::uniform_real_distribution<float> pd(-20.0f, 20.0f);
      for (int i = 0; i < 500; ++i) {
        const float px = pd(rng), py = pd(rng), pz = pd(rng);
        const Eigen::Vector3f p(px, py, pz);
        const int h = cublox::posToHashIndex(p, f.map, f.half, f.inv, f.center);
        Eigen::Vector3f pb;
        cublox::hashIdToPos(h, f.map, f.half, f.origin, f.res, f.center, pb);
        EXPECT_EQ_V(h, cublox::posToHashIndex(pb, f.map, f.half, f.inv, f.center),
                    "PosHashPos");
      }
    }
  }
  std::printf("%d failures\n", g_fail);
  return g_fail == 0 ? 0 : 1;
}
