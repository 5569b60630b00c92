// Occupancy map server (extension; the reference serves one grid from one
// ROS node, main.cpp). Many sessions, each an OccupancyGrid of its own
// shape, share one GPU:
//
//   * admission: a session is admitted while its device bytes fit the HBM
//     budget and a session slot is free; otherwise it waits (FIFO) until
//     closed sessions free room. Requests to a waiting session queue up.
//   * upload at submission: a scan's points go to a device arena in the
//     submitting thread, while earlier ticks run.
//   * continuous batching: the moment a tick finishes on the device, a
//     worker thread forms the next from the requests queued then -- at most
//     one scan per session, in session round-robin order, within the tick's
//     point, grid, tile and segment caps -- and launches it as one GridBatch
//     tick (update, then the recenter of every scanned grid toward its base
//     position); the finished tick's completions are recorded while the
//     next one runs.
//   * fetches: a fetch request runs once every scan its session submitted
//     before it has been applied, so it sees exactly those scans.
//
// Each session's grid ends every request exactly as the same requests
// issued alone on that grid would leave it (update, recenter, fetch).
#pragma once

#include <Eigen/Dense>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <thread>
#include <vector>

#include <cublox/GridBatch.hpp>
#include <cublox/OccupancyGrid.hpp>

namespace cublox {

struct SessionSpec {
  Eigen::Vector3i half_map_size_i{Eigen::Vector3i(100, 100, 25)};
  float resolution{0.1f};
  bool origin_at_center{false};
  std::optional<double> recenter_threshold{};
  float max_range{25.0f};
  // l_hit, l_miss, l_min, l_max, l_free, l_occupied (logits).
  std::array<float, 6> log_odds{0.847f, -0.405f, -1.386f, 1.386f, -0.2f, 0.2f};
};

struct ServerLimits {
  std::size_t hbm_budget_bytes{200ull << 30};  // sessions' device bytes
  int max_sessions{256};             // admitted at once
  int max_points_per_scan{131072};
  int max_points_per_tick{1 << 20};  // bounds a tick's device time
  int max_grids_per_tick{256};
  int max_tiles_per_tick{1 << 21};   // sum of the tick's grids' tile counts
  int max_segments_per_ray{96};      // sessions above it are refused
  bool graph{true};                  // replay the tick as a HIP graph
  std::size_t point_arena_bytes{4ull << 30};  // queued scans' device points
  // When a session with queued requests waits for room, evict the least
  // recently used idle session (nothing queued or in flight): its map goes
  // to a host snapshot and comes back, bit-identical, when it is readmitted.
  bool evict_idle{false};
  // Overlap consecutive ticks while requests back up (GridBatch
  // pipelining: a tick's walk and binning beside the previous tick's folds,
  // recenters and fetches; a second workspace); one tick at a time when a
  // formed tick leaves less than half a tick of points queued. Every session's requests still take
  // effect in submission order.
  bool pipelined{false};
  // Records per hot-tile work item of the ticks (GridBatch TuneSetting
  // chunk): ticks here are small, so smaller than the batch default.
  int chunk{4096};
  // Fetch samples a tick's output arena holds from the start (13 B each,
  // device and pinned host, per arena); arenas grow past it on demand.
  long long fetch_samples_per_tick{1 << 22};
};

class MapServer {
public:
  enum class Kind { SCAN = 0, FETCH = 1 };

  struct Completion {
    long long id;
    int session;
    Kind kind;
    double submit_ms, start_ms, done_ms;  // since the server started
    long long tick;                       // tick that served it
    long long result;                     // fetch: returned samples
  };

  struct Stats {
    long long ticks{0}, scans{0}, fetches{0}, points{0};
    double device_ms{0.0};      // summed tick device time (HIP events)
    double launch_ms{0.0};      // summed host time from a tick's end on the
                                // device to the next tick's launch
    std::size_t resident_bytes{0}, workspace_bytes{0};
    int admitted{0}, waiting{0};
    long long evictions{0}, restores{0};
    std::size_t snapshot_bytes{0};  // host bytes held by evicted sessions
  };

  explicit MapServer(const ServerLimits &limits);
  ~MapServer();
  MapServer(const MapServer &) = delete;
  MapServer &operator=(const MapServer &) = delete;

  // Device bytes a session of this spec is charged against the budget.
  static std::size_t sessionBytes(const SessionSpec &spec);

  // Returns the session id. The session is admitted by the worker when it
  // fits; throws if it can never fit (shape above the limits).
  int openSession(const SessionSpec &spec);
  // The session closes after its queued requests complete; its grid is
  // freed and waiting sessions are admitted.
  void closeSession(int sid);
  bool admitted(int sid) const;
  // Moves the session's map to a host snapshot (cublox/snapshot.h) and
  // frees its device memory once none of its requests is in flight. Later
  // requests queue; the session is readmitted (its map restored exactly)
  // when it has requests and fits again.
  void evictSession(int sid);
  // The snapshot an evicted session holds (empty for a resident one).
  std::vector<uint8_t> evictedSnapshot(int sid) const;

  // Thread-safe; the points are copied to the device before the call
  // returns (it waits while the point arena is full). T is row-major 3x4
  // sensor->world.
  long long submitScan(int sid, const float *xs, const float *ys,
                       const float *zs, int n, const float T[12],
                       const Eigen::Vector3f &base);
  long long submitFetch(int sid, const Eigen::Vector3f &center, float radius,
                        std::optional<OccupancyGrid::VoxelState> filter = std::nullopt,
                        int max_results = 0);

  // Blocks until every request submitted so far to an admitted session
  // has completed (waiting sessions' requests stay queued). Not while
  // paused.
  void drain();
  // While paused the worker forms no new tick (a running one completes);
  // requests keep queueing. resume() lets the next tick take them.
  void pause();
  void resume();
  std::vector<Completion> takeCompletions();
  // Keep fetch samples for takeFetchResult (off by default).
  void keepFetchResults(bool on);
  std::vector<OccupancyGrid::OccupancySample> takeFetchResult(long long id);

  // The session's grid; use only after drain() with no request in flight.
  const OccupancyGrid &grid(int sid) const;
  Stats stats() const;
  double nowMs() const;

private:
  struct Request {
    long long id;
    Kind kind;
    double submit_ms;
    float *d_pts{nullptr};   // device: x block, y block, z block
    std::size_t arena_off{0}, arena_bytes{0};
    int n{0};
    std::array<float, 12> T{};
    Eigen::Vector3f pos{Eigen::Vector3f::Zero()};  // base or fetch centre
    float radius{0.0f};
    std::optional<OccupancyGrid::VoxelState> filter{};
    int max_results{0};
  };
  struct Session {
    SessionSpec spec;
    std::unique_ptr<OccupancyGrid> grid;
    std::size_t bytes{0};
    int tiles{0}, seg_bound{0};
    bool admitted{false}, closing{false};
    bool in_waiting{false};       // listed in the admission queue
    bool evict_requested{false};
    int busy{0};  // requests taken into a tick not yet complete
    long long last_tick{-1};      // last tick that took its requests
    std::vector<uint8_t> snapshot;  // the evicted map (empty when resident)
    std::deque<Request> queue;
  };
  struct FetchJob {
    bool live{false};       // the plan found voxels to test
    unsigned cap{0};        // output capacity (samples) of its arena range
    unsigned limit{0};      // reference cap: min(max_results, default cap)
    std::size_t off{0};     // first sample of its arena range
  };
  struct Tick {
    long long index{0};
    std::vector<std::pair<int, Request>> scans;
    std::vector<std::pair<int, Request>> fetches;
    std::vector<FetchJob> jobs;  // one per fetch, planned at enqueue
    double start_ms{0.0}, done_ms{0.0}, launched_ms{0.0};
    double device_ms{0.0};
    int buf{0};                  // fetch output arena
    void *stream{nullptr};       // the batch stream its work went to
  };
  struct FetchArena {            // fetch outputs of one tick
    float *d_pos{nullptr}, *h_pos{nullptr};
    unsigned char *d_st{nullptr}, *h_st{nullptr};
    unsigned *d_cnt{nullptr}, *h_cnt{nullptr};
    std::size_t cap{0}, cnt_cap{0};
  };

  void run();
  void admitWaiting();                       // worker thread, lock held
  void retireClosed();                       // worker thread, lock held
  void evictRequested();                     // worker thread, lock held
  void evict(int sid, Session &s);           // worker thread, lock held
  bool evictIdle();                          // worker thread, lock held
  bool idleVictim() const;                   // lock held
  bool formTick(Tick &t);                    // lock held
  bool backlog() const;                      // lock held
  void submitTick(Tick &t);
  void enqueueFetches(Tick &t);
  void waitTick(Tick &t);
  void finishTick(Tick &t);
  void growFetchArena(FetchArena &a, std::size_t samples, std::size_t fetches);
  std::size_t allocPoints(std::size_t bytes);  // blocks while full
  void releasePoints(std::size_t off, std::size_t bytes);
  bool admittable(const Session &s) const;  // lock held
  bool quiescent() const;                    // lock held

  ServerLimits lim_;
  std::unique_ptr<GridBatch> batch_;
  FetchArena farena_[3];  // one per tick alive: in flight, or being finished
  // Device arena of queued scans' points: first-fit free list (offset ->
  // size), its own lock, and the stream the submitting threads copy on.
  char *d_arena_{nullptr};
  std::map<std::size_t, std::size_t> arena_free_;
  std::mutex arena_mu_;
  std::condition_variable arena_cv_;
  bool arena_closed_{false};
  // Upload stages: pinned buffers of one largest scan's three columns, each
  // with its own stream; a submitting thread takes a free one, so uploads
  // from several client threads run side by side.
  struct UploadStage {
    void *host{nullptr};
    void *stream{nullptr};
  };
  std::vector<UploadStage> stages_;
  std::vector<int> free_stages_;
  std::mutex stage_mu_;
  std::condition_variable stage_cv_;

  mutable std::mutex mu_;
  std::condition_variable cv_work_, cv_idle_;
  std::map<int, Session> sessions_;
  std::deque<int> waiting_;
  int next_sid_{0};
  long long next_id_{0};
  long long gen_{0};            // bumped by every client state change
  bool worker_busy_{false};     // the worker holds formed or running ticks
  std::size_t resident_{0};
  int admitted_count_{0};
  int rr_{0};
  bool stop_{false};
  bool paused_{false};
  bool keep_fetch_{false};
  std::vector<Completion> done_;
  std::map<long long, std::vector<OccupancyGrid::OccupancySample>> fetch_results_;
  Stats stats_{};
  long long ticks_{0};
  std::chrono::steady_clock::time_point t0_;
  std::thread worker_;
  std::exception_ptr worker_error_;
};

}  // namespace cublox
