// Open-loop load generator for cublox::MapServer (benchmark client; no
// reference counterpart). Scan arrivals form a Poisson process of the given
// total rate over the sessions (each arrival picks a session uniformly, so
// every session sees a Poisson stream of rate / n); fetch arrivals form an
// independent Poisson process at fetch_ratio times that rate. Each session
// replays its template scan with the pose advanced along x by `speed` per
// scan, and recenters toward the advanced base; fetches ask for the sphere
// around the current sensor position.
#include <cublox/MapServer.hpp>
#include <cublox/capi.h>
#include <cublox/ms_capi.h>

#include <chrono>
#include <cmath>
#include <exception>
#include <stdexcept>
#include <thread>
#include <vector>

namespace cublox_ffi {
void set_last_error(const char *msg);              // og_capi.cpp
cublox::MapServer &server_of(cublox_ms *h);        // ms_capi.cpp
}

namespace {

// SplitMix64: a fixed, platform-independent stream (std:: distributions
// are implementation-defined).
struct SplitMix64 {
  unsigned long long s;
  unsigned long long next() {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  double exponential(double rate) { return -std::log1p(-uniform()) / rate; }
};

}  // namespace

extern "C" {

int cublox_ms_load_run(cublox_ms *h, const int *sids, int n_sessions,
                       const float *const *xs, const float *const *ys,
                       const float *const *zs, const int *n, const float *T,
                       const float *base, const cublox_ms_load_cfg *cfg,
                       long long *submitted) {
  try {
    if (!h || !sids || n_sessions <= 0 || !xs || !ys || !zs || !n || !T || !base || !cfg ||
        cfg->scan_rate_hz <= 0.0 || cfg->duration_s <= 0.0 || cfg->fetch_ratio < 0.0) {
      throw std::invalid_argument("cublox_ms_load_run: bad arguments");
    }
    cublox::MapServer &srv = cublox_ffi::server_of(h);
    SplitMix64 rng{cfg->seed};
    std::vector<long long> scans_of(static_cast<size_t>(n_sessions), 0);
    std::vector<float> Tk(12);
    const double fetch_rate = cfg->scan_rate_hz * cfg->fetch_ratio;
    double t_scan = rng.exponential(cfg->scan_rate_hz);
    double t_fetch = fetch_rate > 0.0 ? rng.exponential(fetch_rate) : 1e300;
    const auto t0 = std::chrono::steady_clock::now();
    long long count = 0;
    while (true) {
      const bool is_scan = t_scan <= t_fetch;
      const double t = is_scan ? t_scan : t_fetch;
      if (t >= cfg->duration_s) break;
      std::this_thread::sleep_until(t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                             std::chrono::duration<double>(t)));
      const int k = static_cast<int>(rng.uniform() * n_sessions) % n_sessions;
      const float dx = cfg->speed * static_cast<float>(scans_of[static_cast<size_t>(k)]);
      if (is_scan) {
        for (int i = 0; i < 12; ++i) Tk[static_cast<size_t>(i)] = T[12 * k + i];
        Tk[3] += dx;
        const Eigen::Vector3f b(base[3 * k] + dx, base[3 * k + 1], base[3 * k + 2]);
        srv.submitScan(sids[k], xs[k], ys[k], zs[k], n[k], Tk.data(), b);
        ++scans_of[static_cast<size_t>(k)];
        t_scan += rng.exponential(cfg->scan_rate_hz);
      } else {
        const Eigen::Vector3f c(T[12 * k + 3] + dx, T[12 * k + 7], T[12 * k + 11]);
        srv.submitFetch(sids[k], c, cfg->fetch_radius);
        t_fetch += rng.exponential(fetch_rate);
      }
      ++count;
    }
    if (submitted) *submitted = count;
    return CUBLOX_OK;
  } catch (const std::exception &e) {
    cublox_ffi::set_last_error(e.what());
    return CUBLOX_ERR_RUNTIME;
  }
}

}  // extern "C"
