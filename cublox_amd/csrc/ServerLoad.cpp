// Open-loop load generator for cublox::MapServer (benchmark client; no
// reference counterpart). Scan arrivals form a Poisson process of the given
// total rate over the sessions (each arrival picks a session uniformly, so
// every session sees a Poisson stream of rate / n); fetch arrivals form an
// independent Poisson process at fetch_ratio times that rate. Each session
// replays its template scan with the pose advanced along x by `speed` per
// scan, and recenters toward the advanced base; fetches ask for a sphere
// of uniformly drawn radius around the current sensor position.
#include <cublox/MapServer.hpp>
#include <cublox/capi.h>
#include <cublox/ms_capi.h>

#include <chrono>
#include <cmath>
#include <algorithm>
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

// One client: its own Poisson streams over its own sessions.
void run_client(cublox::MapServer &srv, const std::vector<int> &mine, const int *sids,
                const float *const *xs, const float *const *ys, const float *const *zs,
                const int *n, const float *T, const float *base,
                const cublox_ms_load_cfg &cfg, double scan_rate, unsigned long long seed,
                std::chrono::steady_clock::time_point t0, long long *count,
                double *max_lag_ms) {
  SplitMix64 rng{seed};
  std::vector<long long> scans_of(mine.size(), 0);
  float Tk[12];
  const double fetch_rate = scan_rate * cfg.fetch_ratio;
  double t_scan = rng.exponential(scan_rate);
  double t_fetch = fetch_rate > 0.0 ? rng.exponential(fetch_rate) : 1e300;
  const int m = static_cast<int>(mine.size());
  while (true) {
    const bool is_scan = t_scan <= t_fetch;
    const double t = is_scan ? t_scan : t_fetch;
    if (t >= cfg.duration_s) break;
    const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                              std::chrono::duration<double>(t));
    std::this_thread::sleep_until(due);
    const double lag =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - due).count();
    if (lag > *max_lag_ms) *max_lag_ms = lag;
    const int j = static_cast<int>(rng.uniform() * m) % m;
    const int k = mine[static_cast<size_t>(j)];
    const float dx = cfg.speed * static_cast<float>(scans_of[static_cast<size_t>(j)]);
    if (is_scan) {
      for (int i = 0; i < 12; ++i) Tk[i] = T[12 * k + i];
      Tk[3] += dx;
      const Eigen::Vector3f b(base[3 * k] + dx, base[3 * k + 1], base[3 * k + 2]);
      srv.submitScan(sids[k], xs[k], ys[k], zs[k], n[k], Tk, b);
      ++scans_of[static_cast<size_t>(j)];
      t_scan += rng.exponential(scan_rate);
    } else {
      const Eigen::Vector3f c(T[12 * k + 3] + dx, T[12 * k + 7], T[12 * k + 11]);
      const float r = cfg.fetch_radius_max > cfg.fetch_radius
                          ? cfg.fetch_radius + static_cast<float>(rng.uniform()) *
                                                   (cfg.fetch_radius_max - cfg.fetch_radius)
                          : cfg.fetch_radius;
      srv.submitFetch(sids[k], c, r);
      t_fetch += rng.exponential(fetch_rate);
    }
    ++*count;
  }
}

}  // namespace

extern "C" {

int cublox_ms_load_run(cublox_ms *h, const int *sids, int n_sessions,
                       const float *const *xs, const float *const *ys,
                       const float *const *zs, const int *n, const float *T,
                       const float *base, const cublox_ms_load_cfg *cfg,
                       long long *submitted, double *max_lag_ms) {
  try {
    if (!h || !sids || n_sessions <= 0 || !xs || !ys || !zs || !n || !T || !base || !cfg ||
        cfg->scan_rate_hz <= 0.0 || cfg->duration_s <= 0.0 || cfg->fetch_ratio < 0.0) {
      throw std::invalid_argument("cublox_ms_load_run: bad arguments");
    }
    cublox::MapServer &srv = cublox_ffi::server_of(h);
    const int clients = std::max(1, std::min(cfg->clients, n_sessions));
    std::vector<std::vector<int>> mine(static_cast<size_t>(clients));
    for (int k = 0; k < n_sessions; ++k) mine[static_cast<size_t>(k % clients)].push_back(k);
    std::vector<long long> counts(static_cast<size_t>(clients), 0);
    std::vector<double> lags(static_cast<size_t>(clients), 0.0);
    std::vector<std::exception_ptr> errs(static_cast<size_t>(clients));
    std::vector<std::thread> threads;
    const auto t0 = std::chrono::steady_clock::now();
    for (int c = 0; c < clients; ++c) {
      const size_t ci = static_cast<size_t>(c);
      // Each client's share of the total rate is its share of the sessions.
      const double rate = cfg->scan_rate_hz * static_cast<double>(mine[ci].size()) / n_sessions;
      threads.emplace_back([&, ci, rate] {
        try {
          run_client(srv, mine[ci], sids, xs, ys, zs, n, T, base, *cfg, rate,
                     cfg->seed + 0x9E37ull * ci, t0, &counts[ci], &lags[ci]);
        } catch (...) {
          errs[ci] = std::current_exception();
        }
      });
    }
    for (auto &t : threads) t.join();
    for (auto &e : errs) {
      if (e) std::rethrow_exception(e);
    }
    long long total = 0;
    double lag = 0.0;
    for (size_t c = 0; c < counts.size(); ++c) {
      total += counts[c];
      lag = std::max(lag, lags[c]);
    }
    if (submitted) *submitted = total;
    if (max_lag_ms) *max_lag_ms = lag;
    return CUBLOX_OK;
  } catch (const std::exception &e) {
    cublox_ffi::set_last_error(e.what());
    return CUBLOX_ERR_RUNTIME;
  }
}

}  // extern "C"
