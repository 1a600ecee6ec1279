// server.h -- request loop over HBM-resident scenes with continuous batching.
//
// Requests (scene, camera, target size) queue on submit(). A submission
// thread folds every queued request that fits into the next batch as soon as
// one of the pipelines (ServerConfig::pipelines, 3 by default) is free, so
// requests arriving while a batch is on the device ride a following one, and
// the pipelines' batches overlap on the device (separate streams). A completion thread per pipeline waits for
// its batch, delivers the images and wakes the waiters. Each request's image
// is the frame a context would render for it alone: batching never changes
// a view's sort or blend.
//
// Memory admission (the analogue of direct_scene_preflight,
// render-wgpu/src/lib.rs:2202-2304, with the device budget in place of the
// storage-binding limit): resident scenes plus, per pipeline, the working
// set of its batch (frame_working_bytes) stay within the budget. A scene is
// admitted when the budget still holds the working set of one 1080p view of
// it on each pipeline; a batch takes queued requests in arrival order while
// its working set fits the pipeline's share.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pipeline.h"

namespace gs {

struct ServerConfig {
  uint64_t memory_budget = 0;  // bytes; 0 = 90% of the device memory free at creation
  uint32_t max_batch_views = 64;
  uint32_t max_queue = 4096;
  uint32_t pipelines = 3;  // batches in flight at once (1..8), one stream each
  // Balance batches by cost: a free pipeline takes queued requests up to its
  // share of the outstanding work (queued + running elsewhere) instead of all
  // of them (also GSPLAT_SERVER_BALANCE=0/1).
  bool balance = true;
};

struct RequestResult {
  int32_t status = 0;  // Code
  std::string error;
  uint32_t visible = 0;
  uint32_t batch_views = 0;
  float queue_ms = 0, device_ms = 0, total_ms = 0;
};

struct ServerStats {
  uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0;
  uint64_t views_in_batches = 0;
  uint64_t resident_scene_bytes = 0, memory_budget = 0;
  uint32_t scenes = 0, queued = 0;
  uint64_t overflow_rerenders = 0;  // batches rendered again after a tile-pair overflow
  uint64_t buffer_grows = 0;        // batches that reallocated their pipeline's frame buffers
};

// Pinned host allocations handed out to callers for server outputs; the
// raster writes a request's RGBA8 pixels straight into its target when the
// target lies in one of them. A block stays held while a queued or running
// request targets it: host_free refuses it (InvalidArgument) until then.
void *host_alloc(size_t bytes);
void host_free(void *p);
bool host_is_pinned(const void *p, size_t len);
// Holds the block containing [p, p + len) for one request (false, holding
// nothing, when no block contains it); host_release lets one hold go.
bool host_hold(const void *p, size_t len);
void host_release(const void *p);

class Server {
 public:
  explicit Server(const ServerConfig &cfg);
  ~Server();
  Server(const Server &) = delete;
  Server &operator=(const Server &) = delete;

  uint32_t add_scene(std::shared_ptr<DeviceScene> scene);
  void remove_scene(uint32_t id);
  // Queues one render; `out` (RGBA8, width*height*4 bytes, may be null)
  // must stay valid until wait() returns for the ticket.
  uint64_t submit(uint32_t scene_id, const GsplatCamera &camera, uint32_t width, uint32_t height, uint8_t *out,
                  uint64_t out_len);
  // Blocks until the request is done and releases its ticket.
  RequestResult wait(uint64_t ticket);
  ServerStats stats();

 private:
  struct Request {
    uint64_t ticket = 0;
    std::shared_ptr<DeviceScene> scene;
    FrameParams fp{};
    double cost = 0;  // request_cost
    uint8_t *out = nullptr;
    bool out_pinned = false;
    std::chrono::steady_clock::time_point arrival, start;
    bool done = false;
    RequestResult result;
  };
  struct Slot {
    std::unique_ptr<Pipeline> pipe;
    std::vector<std::shared_ptr<Request>> batch;
    double cost = 0;     // request_cost summed over the batch
    bool busy = false;   // a batch is assigned (being enqueued or in flight)
    bool ready = false;  // the batch is enqueued; the completer may wait on it
    std::condition_variable cv;
    std::thread completer;
  };

  void submit_loop();
  void complete_loop(Slot *slot);
  uint64_t slot_budget();  // working-set share of one pipeline (under m_)
  // Resident scene bytes plus removed scenes still referenced (under m_).
  uint64_t held_bytes();
  // Marks the request done and lets its pinned target go (under m_).
  void finish(Request &r);

  ServerConfig cfg_;
  std::mutex m_;
  std::condition_variable submit_cv_, done_cv_;
  std::deque<std::shared_ptr<Request>> queue_;
  double queued_cost_ = 0;  // request_cost summed over queue_
  std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets_;
  std::map<uint32_t, std::shared_ptr<DeviceScene>> scenes_;
  uint32_t next_scene_ = 1;
  uint64_t next_ticket_ = 1;
  uint64_t resident_bytes_ = 0;
  // Removed scenes and their bytes: counted against the budget while any
  // request or pipeline still holds them.
  std::vector<std::pair<std::weak_ptr<DeviceScene>, uint64_t>> retired_;
  ServerStats stats_{};
  bool stop_ = false;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::thread submitter_;
};

}  // namespace gs
