// server.cpp -- continuous-batching request loop.

#include "server.h"

#include <cstdlib>

#include <algorithm>

#include "renderer.h"

namespace gs {

namespace {

// Every host_alloc block: its size and the number of queued or running
// requests whose target lies in it (host_free refuses a block in use).
struct PinnedBlock {
  size_t bytes = 0;
  uint32_t users = 0;
};
std::mutex g_pinned_m;
std::map<uintptr_t, PinnedBlock> g_pinned;  // base -> block

// Block holding [p, p + len), or end() (caller holds g_pinned_m).
std::map<uintptr_t, PinnedBlock>::iterator find_block(const void *p, size_t len) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return g_pinned.end();
  --it;
  return (a >= it->first && a + len <= it->first + it->second.bytes) ? it : g_pinned.end();
}

constexpr uint32_t kMaxDimension = 16384;

float ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<float, std::milli>(b - a).count();
}

uint64_t round_segment(uint64_t n) { return (n + kSegmentAlign - 1) / kSegmentAlign * kSegmentAlign; }

uint64_t initial_pairs(uint64_t splats) {
  return std::min<uint64_t>(std::max<uint64_t>(8 * splats, 1u << 20), 1u << 30);
}

// Relative device cost of a request: its splat slots (projection, depth sort,
// tile counts) plus its pixels (raster, pair sort). On batches of served
// requests both cost about 0.1 us per thousand, so they weigh alike.
double request_cost(uint32_t splats, const FrameParams &fp) {
  return (double)round_segment(splats) + (double)fp.width * fp.height;
}

}  // namespace

void *host_alloc(size_t bytes) {
  void *p = nullptr;
  GS_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  std::lock_guard<std::mutex> g(g_pinned_m);
  g_pinned[reinterpret_cast<uintptr_t>(p)] = PinnedBlock{std::max<size_t>(bytes, 1), 0};
  return p;
}

void host_free(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> g(g_pinned_m);
    auto it = g_pinned.find(reinterpret_cast<uintptr_t>(p));
    if (it == g_pinned.end()) throw Error(Code::InvalidArgument, "pointer was not returned by gsplat_host_alloc");
    if (it->second.users)
      throw Error(Code::InvalidArgument, "block is the target of " + std::to_string(it->second.users) +
                                             " queued or running request(s); wait for them first");
    g_pinned.erase(it);
  }
  (void)hipHostFree(p);
}

bool host_is_pinned(const void *p, size_t len) {
  std::lock_guard<std::mutex> g(g_pinned_m);
  return find_block(p, len) != g_pinned.end();
}

bool host_hold(const void *p, size_t len) {
  std::lock_guard<std::mutex> g(g_pinned_m);
  auto it = find_block(p, len);
  if (it == g_pinned.end()) return false;
  ++it->second.users;
  return true;
}

void host_release(const void *p) {
  std::lock_guard<std::mutex> g(g_pinned_m);
  auto it = find_block(p, 1);
  if (it != g_pinned.end() && it->second.users) --it->second.users;
}

Server::Server(const ServerConfig &cfg) : cfg_(cfg) {
  if (!device_present()) throw Error(Code::Unsupported, "gpu rasterizer unavailable");
  if (cfg_.max_batch_views == 0 || cfg_.max_batch_views > kMaxViews)
    throw Error(Code::InvalidArgument, "max_batch_views must be 1.." + std::to_string(kMaxViews));
  if (cfg_.max_queue == 0) throw Error(Code::InvalidArgument, "max_queue must be positive");
  if (cfg_.memory_budget == 0) {
    size_t free_b = 0, total_b = 0;
    GS_HIP(hipMemGetInfo(&free_b, &total_b));
    cfg_.memory_budget = free_b / 10 * 9;
  }
  if (cfg_.pipelines == 0 || cfg_.pipelines > 8) throw Error(Code::InvalidArgument, "pipelines must be 1..8");
  if (const char *e = std::getenv("GSPLAT_SERVER_BALANCE")) cfg_.balance = std::atoi(e) != 0;

  stats_.memory_budget = cfg_.memory_budget;
  for (uint32_t i = 0; i < cfg_.pipelines; ++i) {
    slots_.push_back(std::make_unique<Slot>());
    slots_.back()->pipe = std::make_unique<Pipeline>();
    slots_.back()->pipe->set_raster_variant(raster_variant_from_env());
    // Served batches mix many small requests: binning each at most 512
    // binning tiles keeps the pair sort one (9-bit) pass (a 24-request batch:
    // 3.26 ms, against 3.40 ms unbounded and 3.82 ms at 256).
    Tuning t = slots_.back()->pipe->tuning();
    if (!std::getenv("GSPLAT_BIN_MAX_CTILES")) t.bin_max_ctiles = 512;
    slots_.back()->pipe->set_tuning(t);
  }
  for (auto &s : slots_) s->completer = std::thread(&Server::complete_loop, this, s.get());
  submitter_ = std::thread(&Server::submit_loop, this);
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(m_);
    stop_ = true;
  }
  submit_cv_.notify_all();
  if (submitter_.joinable()) submitter_.join();
  for (auto &s : slots_) {
    s->cv.notify_all();
    if (s->completer.joinable()) s->completer.join();
  }
}

uint64_t Server::held_bytes() {
  uint64_t b = resident_bytes_;
  for (auto it = retired_.begin(); it != retired_.end();) {
    if (it->first.expired()) {
      it = retired_.erase(it);
    } else {
      b += it->second;
      ++it;
    }
  }
  return b;
}

uint64_t Server::slot_budget() {
  const uint64_t held = held_bytes();
  const uint64_t free_b = cfg_.memory_budget > held ? cfg_.memory_budget - held : 0;
  return free_b / cfg_.pipelines;
}

void Server::finish(Request &r) {
  if (r.out_pinned) host_release(r.out);
  r.out_pinned = false;
  r.done = true;
}

uint32_t Server::add_scene(std::shared_ptr<DeviceScene> scene) {
  if (!scene) throw Error(Code::InvalidArgument, "scene is null");
  if (scene->count > kLocalMask)
    throw Error(Code::Unsupported, "a served scene holds at most " + std::to_string(kLocalMask) + " splats");
  const uint64_t one_view = frame_working_bytes(round_segment(scene->count), 1920ull * 1080, 120ull * 68,
                                                initial_pairs(scene->count));
  std::lock_guard<std::mutex> g(m_);
  const uint64_t held = held_bytes();
  const uint64_t need = held + scene->bytes() + cfg_.pipelines * one_view;
  if (need > cfg_.memory_budget)
    throw Error(Code::Unsupported, "scene of " + std::to_string(scene->bytes()) + " bytes does not fit the memory budget (" +
                                       std::to_string(held) + " resident, " +
                                       std::to_string(cfg_.memory_budget) + " budget)");
  const uint32_t id = next_scene_++;
  resident_bytes_ += scene->bytes();
  scenes_[id] = std::move(scene);
  return id;
}

void Server::remove_scene(uint32_t id) {
  std::lock_guard<std::mutex> g(m_);
  auto it = scenes_.find(id);
  if (it == scenes_.end()) throw Error(Code::InvalidArgument, "unknown scene id " + std::to_string(id));
  // Queued and running requests (and the pipelines' last frames) keep their
  // own reference: the scene's HBM counts against the budget until the last
  // of them lets go.
  resident_bytes_ -= it->second->bytes();
  retired_.emplace_back(std::weak_ptr<DeviceScene>(it->second), it->second->bytes());
  scenes_.erase(it);
}

uint64_t Server::submit(uint32_t scene_id, const GsplatCamera &camera, uint32_t width, uint32_t height,
                        uint8_t *out, uint64_t out_len) {
  if (!camera_valid(camera)) throw Error(Code::InvalidArgument, "invalid camera");
  if (width == 0 || height == 0) throw Error(Code::InvalidArgument, "invalid renderer configuration");
  if (width > kMaxDimension || height > kMaxDimension)
    throw Error(Code::Unsupported, "render dimensions exceed the device 2D texture limit 16384");
  const uint64_t need = (uint64_t)width * height * 4;
  if (out && out_len < need) throw Error(Code::InvalidArgument, "output buffer smaller than width*height*4 bytes");
  auto r = std::make_shared<Request>();
  r->fp = make_frame_params(camera.position, camera.rotation_xyzw, camera.vertical_fov_radians, camera.near_plane,
                            camera.far_plane, width, height);
  r->out = out;
  std::lock_guard<std::mutex> g(m_);
  if (stop_) throw Error(Code::Internal, "server is shutting down");
  auto it = scenes_.find(scene_id);
  if (it == scenes_.end()) throw Error(Code::InvalidArgument, "unknown scene id " + std::to_string(scene_id));
  if (queue_.size() >= cfg_.max_queue) {
    ++stats_.rejected;
    throw Error(Code::Unsupported, "request queue full (" + std::to_string(cfg_.max_queue) + ")");
  }
  r->scene = it->second;
  r->cost = request_cost(r->scene->count, r->fp);
  // A pinned target is held (host_free refuses it) until the request ends.
  r->out_pinned = out && host_hold(out, need);
  r->ticket = next_ticket_++;
  r->arrival = std::chrono::steady_clock::now();
  tickets_[r->ticket] = r;
  queue_.push_back(r);
  queued_cost_ += r->cost;
  ++stats_.submitted;
  submit_cv_.notify_one();
  return r->ticket;
}

RequestResult Server::wait(uint64_t ticket) {
  std::unique_lock<std::mutex> lk(m_);
  auto it = tickets_.find(ticket);
  if (it == tickets_.end()) throw Error(Code::InvalidArgument, "unknown ticket " + std::to_string(ticket));
  std::shared_ptr<Request> r = it->second;
  done_cv_.wait(lk, [&] { return r->done; });
  tickets_.erase(ticket);
  return r->result;
}

ServerStats Server::stats() {
  std::lock_guard<std::mutex> g(m_);
  ServerStats s = stats_;
  s.resident_scene_bytes = held_bytes();
  s.memory_budget = cfg_.memory_budget;
  s.scenes = (uint32_t)scenes_.size();
  s.queued = (uint32_t)queue_.size();
  for (const auto &sl : slots_) {
    s.overflow_rerenders += sl->pipe->overflow_rerenders();
    s.buffer_grows += sl->pipe->buffer_grows();
  }
  return s;
}

void Server::submit_loop() {
  std::unique_lock<std::mutex> lk(m_);
  for (;;) {
    Slot *slot = nullptr;
    submit_cv_.wait(lk, [&] {
      if (stop_) return true;
      if (queue_.empty()) return false;
      for (auto &s : slots_)
        if (!s->busy) return true;
      return false;
    });
    if (stop_) break;
    for (auto &s : slots_)
      if (!s->busy) {
        slot = s.get();
        break;
      }
    // Fold queued requests, oldest first, while the batch's working set
    // fits this pipeline's share of the budget. Balanced (cfg_.balance): also
    // while the batch's cost stays within the pipeline's share of all work
    // outstanding (queued plus running on the other pipelines), so batches
    // of one size, in cost, run side by side: every request then waits
    // about one batch, where a batch of double cost beside a small one
    // would set the latency tail.
    double share = queued_cost_;
    for (auto &s : slots_)
      if (s->busy) share += s->cost;
    share /= (double)cfg_.pipelines;
    const uint64_t budget = slot_budget();
    const uint64_t pair_floor = slot->pipe->pair_capacity();
    std::vector<std::shared_ptr<Request>> batch;
    uint64_t slots = 0, pixels = 0, tiles = 0, splats = 0;
    double cost = 0;
    while (!queue_.empty() && batch.size() < cfg_.max_batch_views) {
      const std::shared_ptr<Request> r = queue_.front();
      if (cfg_.balance && !batch.empty() && cost + 0.5 * r->cost > share) break;
      const uint64_t s2 = slots + round_segment(r->scene->count);
      const uint64_t p2 = pixels + (uint64_t)r->fp.width * r->fp.height;
      const uint64_t t2 = tiles + (uint64_t)r->fp.tiles_x * r->fp.tiles_y;
      const uint64_t n2 = splats + r->scene->count;
      const uint64_t est = frame_working_bytes(s2, p2, t2, std::max<uint64_t>(initial_pairs(n2), pair_floor));
      if (est > budget || s2 > (1ull << 30)) {
        if (!batch.empty()) break;
        queue_.pop_front();
        queued_cost_ -= r->cost;
        r->result.status = (int32_t)Code::Unsupported;
        r->result.error = "request exceeds the memory budget";
        finish(*r);
        ++stats_.rejected;
        done_cv_.notify_all();
        continue;
      }
      queue_.pop_front();
      queued_cost_ -= r->cost;
      batch.push_back(r);
      cost += r->cost;
      slots = s2;
      pixels = p2;
      tiles = t2;
      splats = n2;
    }
    if (queue_.empty()) queued_cost_ = 0;  // no drift of the running sum
    if (batch.empty()) continue;
    const auto now = std::chrono::steady_clock::now();
    for (auto &r : batch) r->start = now;
    slot->batch = batch;
    slot->cost = cost;
    slot->busy = true;
    lk.unlock();
    std::vector<ViewSpec> specs;
    std::vector<uint8_t *> outs;
    specs.reserve(batch.size());
    outs.reserve(batch.size());
    for (auto &r : batch) {
      specs.push_back(ViewSpec{r->scene, r->fp});
      outs.push_back(r->out_pinned ? r->out : nullptr);
    }
    std::string error;
    int32_t code = 0;
    try {
      slot->pipe->enqueue(std::move(specs), std::move(outs));
    } catch (const Error &e) {
      code = (int32_t)e.code;
      error = e.what();
    } catch (const std::exception &e) {
      code = (int32_t)Code::Internal;
      error = e.what();
    }
    lk.lock();
    if (code != 0) {
      for (auto &r : batch) {
        r->result.status = code;
        r->result.error = error;
        finish(*r);
      }
      slot->batch.clear();
      slot->cost = 0;
      slot->busy = false;
      done_cv_.notify_all();
      continue;
    }
    ++stats_.batches;
    stats_.views_in_batches += batch.size();
    slot->ready = true;
    slot->cv.notify_all();
  }
  queued_cost_ = 0;
  for (auto &r : queue_) {
    r->result.status = (int32_t)Code::Internal;
    r->result.error = "server shut down before the request ran";
    finish(*r);
  }
  queue_.clear();
  done_cv_.notify_all();
}

void Server::complete_loop(Slot *slot) {
  std::unique_lock<std::mutex> lk(m_);
  for (;;) {
    slot->cv.wait(lk, [&] { return slot->ready || (stop_ && !slot->busy); });
    if (!slot->ready) break;
    const std::vector<std::shared_ptr<Request>> batch = slot->batch;
    lk.unlock();
    int32_t code = 0;
    std::string error;
    FrameTimes t;
    std::vector<uint32_t> vis(batch.size(), 0);
    try {
      slot->pipe->wait();
      t = slot->pipe->times();
      for (uint32_t i = 0; i < batch.size(); ++i) {
        vis[i] = slot->pipe->visible(i);
        const Request &r = *batch[i];
        if (r.out && !r.out_pinned) slot->pipe->readback_rgba8(i, r.out, (uint64_t)r.fp.width * r.fp.height * 4);
      }
    } catch (const Error &e) {
      code = (int32_t)e.code;
      error = e.what();
    } catch (const std::exception &e) {
      code = (int32_t)Code::Internal;
      error = e.what();
    }
    const auto now = std::chrono::steady_clock::now();
    lk.lock();
    for (uint32_t i = 0; i < batch.size(); ++i) {
      Request &r = *batch[i];
      r.result.status = code;
      r.result.error = error;
      r.result.visible = vis[i];
      r.result.batch_views = (uint32_t)batch.size();
      r.result.queue_ms = ms_between(r.arrival, r.start);
      r.result.device_ms = t.total_ms;
      r.result.total_ms = ms_between(r.arrival, now);
      finish(r);
    }
    stats_.completed += batch.size();
    slot->batch.clear();
    slot->cost = 0;
    slot->ready = false;
    slot->busy = false;
    done_cv_.notify_all();
    submit_cv_.notify_all();
  }
}

}  // namespace gs
