// serve.cpp - many clients, one GPU: request admission and continuous batching.
//
// A client records an ordinary command buffer (render passes, optionally preceded by copies)
// and hands it to the server instead of vkQueueSubmit. The server
//   * admits pending requests in arrival order while the device bytes they touch (targets,
//     textures, and the records and bins their triangles need) fit the HBM budget together with
//     the requests already on the GPU;
//   * merges the admitted requests' render passes into one setup / bin / raster sequence: every
//     target is a frame with its own tile range in the same launches, whatever its size;
//   * runs batches on several executor slots, each with its own stream and scratch, so a batch
//     formed while another is on the GPU starts at once and their work overlaps on the device;
//   * completes each request when its targets are back in host-visible memory, recording
//     submit -> present latency.
// A request is never merged with another that writes the same memory, and never runs while an
// earlier request writing that memory is in flight, so each result is the one the request gives
// alone (every frame's tiles see only that frame's triangles, in API order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <set>
#include <thread>

#include "internal.h"
#include "../../include/visor_icd.h"

namespace visor
{
struct Request
{
  uint64_t ticket = 0;
  Submission sub;
  bool mergeable = true;                       // copies first, then passes on distinct memories
  uint64_t bytes = 0;                          // admission cost
  std::vector<VkDeviceMemory_T *> writes;      // memories the request's passes and copies write
  double t_submit = 0.0, t_start = 0.0, t_done = 0.0;
  VkResult result = VK_SUCCESS;
  bool done = false;
};

static uint64_t ImageFootprint(const VkImage_T *img)
{
  uint64_t total = 0;
  uint32_t w = img->extent.width, h = img->extent.height;
  for(uint32_t m = 0; m < img->mipLevels; m++)
  {
    total += (uint64_t)w * h * img->bytesPerPixel;
    w = std::max(1u, w >> 1);
    h = std::max(1u, h >> 1);
  }
  return total * img->arrayLayers;
}

class Server
{
public:
  explicit Server(const VhServerConfig &c) : cfg(c)
  {
    if(cfg.slots < 1) cfg.slots = 1;
    if(cfg.slots > 4) cfg.slots = 4;
    if(cfg.max_batch < 1) cfg.max_batch = 1;
    if(cfg.hbm_budget_bytes == 0) cfg.hbm_budget_bytes = 16ull << 30;
    for(uint32_t i = 0; i < cfg.slots; i++)
    {
      auto q = std::make_unique<Queue>();
      q->mode = ExecMode::Batched;
      q->graphs = false;     // batch shapes change from batch to batch
      q->autotune = false;
      slots.push_back(std::move(q));
      slot_cv.push_back(std::make_unique<std::condition_variable>());
    }
    mailbox.resize(cfg.slots);
    slot_busy.assign(cfg.slots, false);
    for(uint32_t i = 0; i < cfg.slots; i++)
      threads.emplace_back([this, i] { SlotLoop(i); });
    threads.emplace_back([this] { DispatchLoop(); });
  }

  ~Server()
  {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    dispatch_cv.notify_all();
    for(auto &c : slot_cv)
      c->notify_all();
    for(auto &t : threads)
      t.join();
  }

  VkResult Submit(const VkCommandBuffer_T *cb, uint64_t *ticket)
  {
    auto r = std::make_shared<Request>();
    std::string err;
    if(!slots[0]->Decode(cb, r->sub, err))
    {
      LogError("vh_server_submit: %s", err.c_str());
      return VK_ERROR_FEATURE_NOT_PRESENT;
    }
    Classify(*r);
    r->t_submit = NowSeconds();
    std::lock_guard<std::mutex> lk(m);
    r->ticket = next_ticket++;
    *ticket = r->ticket;
    live[r->ticket] = r;
    pending.push_back(r);
    submitted++;
    dispatch_cv.notify_one();
    return VK_SUCCESS;
  }

  VkResult Wait(uint64_t ticket, uint64_t timeout_ns, double *latency_ms)
  {
    std::unique_lock<std::mutex> lk(m);
    auto it = live.find(ticket);
    if(it == live.end())
      return VK_ERROR_INITIALIZATION_FAILED;
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    if(timeout_ns == UINT64_MAX)
      done_cv.wait(lk, ready);
    else if(!done_cv.wait_for(lk, std::chrono::nanoseconds(timeout_ns), ready))
      return VK_TIMEOUT;
    if(latency_ms)
      *latency_ms = (r->t_done - r->t_submit) * 1e3;
    live.erase(ticket);
    return r->result;
  }

  void Stats(VhServerStats *out)
  {
    std::lock_guard<std::mutex> lk(m);
    memset(out, 0, sizeof(*out));
    out->submitted = submitted;
    out->completed = completed;
    out->batches = batches;
    out->fragments = fragments;
    out->shaded = shaded;
    out->busy_ms = busy_ms + (busy_slots > 0 ? (NowSeconds() - busy_since) * 1e3 : 0.0);
    out->mean_batch = batches ? (double)batched_requests / (double)batches : 0.0;
    out->max_inflight_bytes = max_inflight;
    if(!lat.empty())
    {
      std::vector<double> v = lat;
      std::sort(v.begin(), v.end());
      auto pct = [&](double p) {
        size_t i = (size_t)(p * (double)(v.size() - 1) + 0.5);
        return v[std::min(i, v.size() - 1)];
      };
      double sum = 0.0;
      for(double x : v)
        sum += x;
      out->p50_ms = pct(0.50);
      out->p99_ms = pct(0.99);
      out->mean_ms = sum / (double)v.size();
      out->max_ms = v.back();
    }
  }

  void ResetStats()
  {
    std::lock_guard<std::mutex> lk(m);
    lat.clear();
    submitted = completed = batches = fragments = shaded = batched_requests = 0;
    busy_ms = 0.0;
    busy_since = NowSeconds();
    max_inflight = inflight_bytes;
  }

private:
  // Admission cost and merge rules of a request.
  void Classify(Request &r)
  {
    std::set<VkDeviceMemory_T *> w;
    std::set<const void *> tex;
    bool seen_pass = false;
    uint64_t bytes = 0, tris = 0;
    for(Op &op : r.sub.ops)
    {
      if(op.kind == Op::Copy)
      {
        if(seen_pass)
          r.mergeable = false;  // a copy after a pass must see that pass's result
        if(op.copy.to_image && op.copy.dstImg && op.copy.dstImg->mem)
        {
          w.insert(op.copy.dstImg->mem);
          bytes += ImageFootprint(op.copy.dstImg);
        }
        else if(!op.copy.to_image && op.copy.dstBuf && op.copy.dstBuf->mem)
        {
          w.insert(op.copy.dstBuf->mem);
          bytes += op.copy.bb.size;
        }
        continue;
      }
      seen_pass = true;
      for(VkImage_T *img : {op.frame.color, op.frame.depth})
        if(img)
        {
          if(w.count(img->mem))
            r.mergeable = false;  // two passes on one memory must run in order
          w.insert(img->mem);
          bytes += ImageFootprint(img);
        }
      for(const DrawDesc &dd : op.draws)
      {
        tris += dd.d.ntris;
        for(const VaTex &t : dd.d.tex)
          if(t.base && tex.insert((const void *)(uintptr_t)t.base).second)
            bytes += t.slice_bytes * t.layers;
      }
    }
    // triangle records plus about two (tile, triangle) pairs per triangle, double-buffered
    bytes += tris * (128 + 2 * 16);
    r.bytes = bytes;
    r.writes.assign(w.begin(), w.end());
  }

  bool Conflicts(const Request &r, const std::set<VkDeviceMemory_T *> &w) const
  {
    for(VkDeviceMemory_T *x : r.writes)
      if(w.count(x))
        return true;
    return false;
  }

  void DispatchLoop()
  {
    std::unique_lock<std::mutex> lk(m);
    for(;;)
    {
      dispatch_cv.wait(lk, [&] {
        if(stop)
          return true;
        if(pending.empty())
          return false;
        for(bool b : slot_busy)
          if(!b)
            return true;
        return false;
      });
      if(stop)
        return;
      uint32_t slot = 0;
      while(slot_busy[slot])
        slot++;
      std::vector<std::shared_ptr<Request>> batch;
      std::set<VkDeviceMemory_T *> batch_writes;
      uint64_t bytes = 0;
      for(auto it = pending.begin(); it != pending.end() && batch.size() < cfg.max_batch;)
      {
        Request &r = **it;
        // memory written by a request on the GPU, or by one earlier in this batch, waits
        if(Conflicts(r, busy_writes) || Conflicts(r, batch_writes) ||
           (!batch.empty() && (!r.mergeable || !batch[0]->mergeable)))
        {
          // later requests keep their order behind a conflicting one on the same memory
          for(VkDeviceMemory_T *x : r.writes)
            batch_writes.insert(x);
          ++it;
          continue;
        }
        const bool alone = batch.empty() && inflight_bytes == 0;
        if(!alone && inflight_bytes + bytes + r.bytes > cfg.hbm_budget_bytes)
          break;  // admission in arrival order: wait for memory to free up
        bytes += r.bytes;
        for(VkDeviceMemory_T *x : r.writes)
          batch_writes.insert(x);
        batch.push_back(*it);
        it = pending.erase(it);
      }
      if(batch.empty())
      {
        // everything pending waits for a batch on the GPU: sleep until one completes
        uint64_t gen = completions;
        dispatch_cv.wait(lk, [&] { return stop || completions != gen; });
        if(stop)
          return;
        continue;
      }
      for(auto &r : batch)
        for(VkDeviceMemory_T *x : r->writes)
          busy_writes.insert(x);
      inflight_bytes += bytes;
      max_inflight = std::max(max_inflight, inflight_bytes);
      slot_busy[slot] = true;
      if(busy_slots++ == 0)
        busy_since = NowSeconds();
      mailbox[slot] = std::move(batch);
      slot_cv[slot]->notify_one();
    }
  }

  void SlotLoop(uint32_t i)
  {
    for(;;)
    {
      std::vector<std::shared_ptr<Request>> batch;
      {
        std::unique_lock<std::mutex> lk(m);
        slot_cv[i]->wait(lk, [&] { return stop || !mailbox[i].empty(); });
        if(stop)
          return;
        batch = std::move(mailbox[i]);
        mailbox[i].clear();
      }
      const double t0 = NowSeconds();
      std::vector<Submission *> subs;
      for(auto &r : batch)
      {
        r->t_start = t0;
        subs.push_back(&r->sub);
      }
      ExecStats st;
      VkResult res = slots[i]->ExecuteBatch(subs, st);
      const double t1 = NowSeconds();
      uint64_t bytes = 0;
      for(auto &r : batch)
        bytes += r->bytes;
      {
        std::lock_guard<std::mutex> lk(m);
        for(auto &r : batch)
        {
          r->t_done = t1;
          r->result = res;
          r->done = true;
          lat.push_back((t1 - r->t_submit) * 1e3);
          for(VkDeviceMemory_T *x : r->writes)
            busy_writes.erase(x);
        }
        completed += batch.size();
        batches++;
        batched_requests += batch.size();
        fragments += st.ctr.pixels_written;
        shaded += st.ctr.depth_passed;
        inflight_bytes -= bytes;
        slot_busy[i] = false;
        if(--busy_slots == 0)
          busy_ms += (t1 - busy_since) * 1e3;
        completions++;
      }
      done_cv.notify_all();
      dispatch_cv.notify_all();
    }
  }

  VhServerConfig cfg;
  std::vector<std::unique_ptr<Queue>> slots;
  std::vector<std::unique_ptr<std::condition_variable>> slot_cv;
  std::vector<std::vector<std::shared_ptr<Request>>> mailbox;
  std::vector<bool> slot_busy;
  std::vector<std::thread> threads;
  std::mutex m;
  std::condition_variable dispatch_cv, done_cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<uint64_t, std::shared_ptr<Request>> live;
  std::set<VkDeviceMemory_T *> busy_writes;
  uint64_t next_ticket = 1, inflight_bytes = 0, max_inflight = 0, completions = 0;
  bool stop = false;
  // statistics
  std::vector<double> lat;
  uint64_t submitted = 0, completed = 0, batches = 0, fragments = 0, shaded = 0, batched_requests = 0;
  double busy_ms = 0.0, busy_since = 0.0;
  int busy_slots = 0;
};

// Executes several admitted requests as one batch: uploads, their copies, every render pass in
// one batched launch sequence (one frame per pass), downloads, then waits for the stream.
VkResult Queue::ExecuteBatch(std::vector<Submission *> &subs, ExecStats &st)
{
  std::lock_guard<std::mutex> lk(submit_m);
  hipError_t e;
  if(!stream && (e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess)
    return HipToVk(e);
  std::set<VkDeviceMemory_T *> up, down;
  for(Submission *s : subs)
  {
    up.insert(s->uploads.begin(), s->uploads.end());
    down.insert(s->downloads.begin(), s->downloads.end());
  }
  for(VkDeviceMemory_T *mem : up)
    if((e = hipMemcpyAsync(mem->dev, mem->host, mem->size, hipMemcpyHostToDevice, stream)) != hipSuccess)
      return HipToVk(e);
  VkResult res = VK_SUCCESS;
  if(subs.size() == 1)
  {
    // one request: its operations in recorded order
    for(Op &op : subs[0]->ops)
    {
      if(res != VK_SUCCESS)
        break;
      res = op.kind == Op::Copy ? RunCopy(op.copy) : RunPass({&op}, true, st);
    }
  }
  else
  {
    // merged requests (each: copies, then passes on memories no other request writes)
    std::vector<Op *> passes;
    for(Submission *s : subs)
      for(Op &op : s->ops)
      {
        if(res != VK_SUCCESS)
          break;
        if(op.kind == Op::Copy)
          res = RunCopy(op.copy);
        else
          passes.push_back(&op);
      }
    if(res == VK_SUCCESS && !passes.empty())
      res = RunPass(passes, true, st);
  }
  for(VkDeviceMemory_T *mem : down)
    if((e = hipMemcpyAsync(mem->host, mem->dev, mem->size, hipMemcpyDeviceToHost, stream)) != hipSuccess)
      return HipToVk(e);
  if((e = hipStreamSynchronize(stream)) != hipSuccess && res == VK_SUCCESS)
    res = HipToVk(e);
  return res;
}
}  // namespace visor

using visor::Server;

struct VhServer_T
{
  Server *s;
};

extern "C" {

VkResult vh_server_create(VkDevice device, const VhServerConfig *cfg, VhServer *out)
{
  if(!device || !cfg || !out)
    return VK_ERROR_INITIALIZATION_FAILED;
  std::string err;
  if(!visor::Device::Get().Init(err))
    return VK_ERROR_INITIALIZATION_FAILED;
  *out = new VhServer_T{new Server(*cfg)};
  return VK_SUCCESS;
}

void vh_server_destroy(VhServer s)
{
  if(!s)
    return;
  delete s->s;
  delete s;
}

VkResult vh_server_submit(VhServer s, VkCommandBuffer cb, uint64_t *ticket)
{
  if(!s || !cb || !ticket)
    return VK_ERROR_INITIALIZATION_FAILED;
  return s->s->Submit(cb, ticket);
}

VkResult vh_server_wait(VhServer s, uint64_t ticket, uint64_t timeout_ns, double *latency_ms)
{
  if(!s)
    return VK_ERROR_INITIALIZATION_FAILED;
  return s->s->Wait(ticket, timeout_ns, latency_ms);
}

int vh_server_stats(VhServer s, VhServerStats *out)
{
  if(!s || !out)
    return -1;
  s->s->Stats(out);
  return 0;
}

int vh_server_reset_stats(VhServer s)
{
  if(!s)
    return -1;
  s->s->ResetStats();
  return 0;
}

}  // extern "C"
