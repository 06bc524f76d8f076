// Continuous-batching request server for list-ranking and tree-rooting jobs.
//
// Jobs are admitted against a device-memory budget and coalesced into mixed
// segmented batches (batch.hip).  Each of `num_slots` workers owns a device
// context (its own stream and scratch pool) and a host thread, so a batch can
// be formed and launched while other batches are still on the GPU: requests
// that arrive meanwhile are picked up by the next free slot.  Latency is
// measured from kc_submit to the moment the job's outputs are written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "common.hpp"

namespace kc {
bool is_device_ptr(const void* p);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Ticket {
  kc_job job;
  uint64_t id = 0;
  uint64_t bytes = 0;
  Clock::time_point submitted;
  Clock::time_point finished;
  bool device_ptrs = true;
  int state = 0;  // 0 queued, 1 running, 2 done
  kc_status status = KC_OK;
};

struct Server;

struct Slot {
  void* ctx = nullptr;
  std::thread thread;
};

struct Server {
  int device = 0;
  kc_server_cfg cfg{};
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work;
  std::condition_variable cv_done;
  std::deque<Ticket*> queue;
  // shared: a waiter keeps its ticket alive while another thread erases it
  std::unordered_map<uint64_t, std::shared_ptr<Ticket>> tickets;
  uint64_t next_id = 1;
  uint64_t bytes_in_use = 0;
  uint64_t peak_bytes = 0;
  bool stopping = false;
  // statistics since the last reset
  std::vector<double> latencies_ms;
  uint64_t elements_done = 0;
  uint64_t batches = 0;
  uint64_t batches_rerun = 0;
  uint64_t jobs_failed = 0;
  bool have_first = false;
  Clock::time_point first_submit;
  Clock::time_point last_finish;
};

uint64_t job_bytes(const kc_job& j) {
  // device footprint of a job inside a batch: records, outputs, unclaimed
  // lists and sublists; trees add the tour (two records per vertex) and sort
  // buffers
  const uint64_t per = j.kind == KC_JOB_LIST ? 96 : 208;
  return j.n * per + (1u << 20);
}

// One job as its own library call (the one-job-per-launch path).
kc_status run_alone(kc::Context* ctx, const kc_job& j, uint32_t tour_build) {
  if (j.kind == KC_JOB_LIST)
    return kc_rank_list(ctx, j.succ, j.weight, j.n, j.root_out, j.rank_out, nullptr, ctx->stream);
  kc_tree_cfg tc{};
  tc.tour_build = static_cast<int32_t>(tour_build);
  return kc_root_tree(ctx, j.succ, j.n, j.root_out, j.rank_out, &tc, ctx->stream);
}

void run_batch(Server* sv, void* vctx, std::vector<Ticket*>& batch) {
  auto* ctx = static_cast<kc::Context*>(vctx);
  kc_status status = KC_OK;
  std::vector<kc_status> per_job(batch.size(), KC_OK);
  try {
    KC_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    kc::Scratch scratch(ctx);
    if (sv->cfg.one_job_per_launch || batch.size() == 1) {
      for (size_t k = 0; k < batch.size(); k++) per_job[k] = run_alone(ctx, batch[k]->job, sv->cfg.tour_build);
    } else {
      std::vector<kc::BatchJob> jobs;
      struct Staging {
        Ticket* t;
        uint64_t* root;
        int64_t* rank;
      };
      std::vector<Staging> staged;
      for (Ticket* t : batch) {
        const kc_job& j = t->job;
        kc::BatchJob b{};
        b.n = j.n;
        b.kind = static_cast<uint32_t>(j.kind);
        if (t->device_ptrs) {
          b.succ = j.succ;
          b.w = j.weight;
          b.root_out = j.root_out;
          b.rank_out = j.rank_out;
        } else {
          uint64_t* s = scratch.get<uint64_t>(j.n);
          KC_HIP(hipMemcpyAsync(s, j.succ, j.n * sizeof(uint64_t), hipMemcpyHostToDevice, st));
          int64_t* w = nullptr;
          if (j.weight && j.kind == KC_JOB_LIST) {
            w = scratch.get<int64_t>(j.n);
            KC_HIP(hipMemcpyAsync(w, j.weight, j.n * sizeof(int64_t), hipMemcpyHostToDevice, st));
          }
          b.succ = s;
          b.w = w;
          b.root_out = scratch.get<uint64_t>(j.n);
          b.rank_out = scratch.get<int64_t>(j.n);
          staged.push_back(Staging{t, b.root_out, b.rank_out});
        }
        if (j.kind == KC_JOB_TREE) b.w = nullptr;
        jobs.push_back(b);
      }
      kc::rank_batch_device(ctx, jobs, nullptr, static_cast<int>(sv->cfg.tour_build), st);
      for (auto& s : staged) {
        KC_HIP(hipMemcpyAsync(s.t->job.root_out, s.root, s.t->job.n * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        KC_HIP(hipMemcpyAsync(s.t->job.rank_out, s.rank, s.t->job.n * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      }
      KC_HIP(hipStreamSynchronize(st));
    }
  } catch (const kc::KcError& e) {
    status = e.status;
    ctx->last_error = e.what;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipMemset(ctx->d_err, 0, sizeof(uint32_t));
  } catch (const std::bad_alloc&) {
    status = KC_ENOMEM;
  }
  const bool rerun = status != KC_OK;
  if (rerun) {
    // a bad job fails the coalesced pass: rerun the jobs one by one so that
    // only the offending job reports the error
    for (size_t k = 0; k < batch.size(); k++) per_job[k] = run_alone(ctx, batch[k]->job, sv->cfg.tour_build);
  }
  const Clock::time_point now = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  for (size_t k = 0; k < batch.size(); k++) {
    Ticket* t = batch[k];
    status = per_job[k];
    t->state = 2;
    t->status = status;
    t->finished = now;
    sv->bytes_in_use -= t->bytes;
    if (status == KC_OK) {
      sv->latencies_ms.push_back(std::chrono::duration<double, std::milli>(now - t->submitted).count());
      sv->elements_done += t->job.n;
    } else {
      sv->jobs_failed++;
    }
  }
  sv->batches++;
  if (rerun) sv->batches_rerun++;
  sv->last_finish = now;
  sv->cv_done.notify_all();
  sv->cv_work.notify_all();
}

void worker(Server* sv, Slot* slot) {
  const uint64_t max_jobs = sv->cfg.one_job_per_launch ? 1 : (sv->cfg.max_batch_jobs ? sv->cfg.max_batch_jobs : 4096);
  // 2^26 elements per batch: large enough for the ranking to run at its
  // large-list rate, small enough that several batches overlap
  const uint64_t max_elems = sv->cfg.max_batch_elements ? sv->cfg.max_batch_elements : (uint64_t{1} << 26);
  for (;;) {
    std::vector<Ticket*> batch;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->stopping && sv->queue.empty()) return;
        if (!sv->queue.empty()) {
          Ticket* head = sv->queue.front();
          if (sv->bytes_in_use + head->bytes <= sv->cfg.mem_budget_bytes || sv->bytes_in_use == 0) break;
        }
        sv->cv_work.wait(lk);
      }
      // FIFO admission: take jobs while they fit the batch and the budget
      uint64_t elems = 0;
      while (!sv->queue.empty() && batch.size() < max_jobs) {
        Ticket* t = sv->queue.front();
        if (!batch.empty() && elems + t->job.n > max_elems) break;
        if (sv->bytes_in_use + t->bytes > sv->cfg.mem_budget_bytes && !(batch.empty() && sv->bytes_in_use == 0))
          break;
        sv->queue.pop_front();
        t->state = 1;
        sv->bytes_in_use += t->bytes;
        sv->peak_bytes = std::max(sv->peak_bytes, sv->bytes_in_use);
        elems += t->job.n;
        batch.push_back(t);
      }
    }
    run_batch(sv, slot->ctx, batch);
  }
}

}  // namespace

extern "C" {

kc_status kc_server_create(void* vctx, const kc_server_cfg* cfg, void** server_out) {
  auto* ctx = static_cast<kc::Context*>(vctx);
  if (!ctx || !cfg || !server_out) return KC_EINVAL;
  if (cfg->tour_build > KC_TOUR_LINKED) return KC_EINVAL;
  *server_out = nullptr;
  auto* sv = new Server();
  sv->device = ctx->device;
  sv->cfg = *cfg;
  if (sv->cfg.mem_budget_bytes == 0) sv->cfg.mem_budget_bytes = uint64_t{200} << 30;
  uint32_t nslots = sv->cfg.num_slots ? sv->cfg.num_slots : (sv->cfg.one_job_per_launch ? 1 : 3);
  // Each coalescing slot's scratch is sized for a full batch up front.  A
  // slot whose scratch grows in the middle of serving consolidates it on its
  // next batch, and consolidation waits for the whole device, so it stalls
  // the batches of every other slot too (4 slots, batch cap 2^24, fresh
  // server: first pass 4344 jobs/s, second 6181 on MI355X).  About 48 bytes
  // per batch record (8-byte level-0 records, sublists, unclaimed lists);
  // a batch that needs more grows the scratch as before.
  const uint64_t cap_elems = sv->cfg.max_batch_elements ? sv->cfg.max_batch_elements : (uint64_t{1} << 26);
  const uint64_t reserve = sv->cfg.one_job_per_launch ? 0 : std::min<uint64_t>(cap_elems, uint64_t{1} << 28) * 48;
  for (uint32_t i = 0; i < nslots; i++) {
    auto slot = std::make_unique<Slot>();
    if (kc_context_create(ctx->device, 0, &slot->ctx) != KC_OK) {
      for (auto& s : sv->slots) kc_context_destroy(s->ctx);
      delete sv;
      return KC_EDEVICE;
    }
    static_cast<kc::Context*>(slot->ctx)->arena.reserve(reserve);
    sv->slots.push_back(std::move(slot));
  }
  for (auto& s : sv->slots) s->thread = std::thread(worker, sv, s.get());
  *server_out = sv;
  return KC_OK;
}

void kc_server_destroy(void* vsv) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stopping = true;
  }
  sv->cv_work.notify_all();
  for (auto& s : sv->slots) {
    if (s->thread.joinable()) s->thread.join();
    kc_context_destroy(s->ctx);
  }
  delete sv;
}

kc_status kc_submit(void* vsv, const kc_job* job, uint64_t* ticket_out) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv || !job || !ticket_out) return KC_EINVAL;
  if (job->kind != KC_JOB_LIST && job->kind != KC_JOB_TREE) return KC_EINVAL;
  if (job->n && (!job->succ || !job->root_out || !job->rank_out)) return KC_EINVAL;
  auto t = std::make_shared<Ticket>();
  t->job = *job;
  t->bytes = job_bytes(*job);
  if (t->bytes > sv->cfg.mem_budget_bytes) return KC_EBUDGET;
  (void)hipSetDevice(sv->device);
  t->device_ptrs = kc::is_device_ptr(job->succ) && kc::is_device_ptr(job->root_out) &&
                   kc::is_device_ptr(job->rank_out) && kc::is_device_ptr(job->weight);
  t->submitted = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  t->id = sv->next_id++;
  if (!sv->have_first) {
    sv->have_first = true;
    sv->first_submit = t->submitted;
  }
  *ticket_out = t->id;
  sv->queue.push_back(t.get());
  sv->tickets.emplace(t->id, std::move(t));
  sv->cv_work.notify_one();
  return KC_OK;
}

kc_status kc_poll(void* vsv, uint64_t ticket) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv) return KC_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->tickets.find(ticket);
  if (it == sv->tickets.end()) return KC_ENOTFOUND;
  if (it->second->state != 2) return KC_EAGAIN;
  kc_status s = it->second->status;
  sv->tickets.erase(it);
  return s;
}

kc_status kc_wait(void* vsv, uint64_t ticket) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv) return KC_EINVAL;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->tickets.find(ticket);
  if (it == sv->tickets.end()) return KC_ENOTFOUND;
  std::shared_ptr<Ticket> t = it->second;
  sv->cv_done.wait(lk, [&] { return t->state == 2; });
  // another waiter or kc_poll may have collected the ticket meanwhile
  sv->tickets.erase(ticket);
  return t->status;
}

kc_status kc_server_stats_get(void* vsv, kc_server_stats* out) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv || !out) return KC_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  kc_server_stats s{};
  s.jobs_completed = sv->latencies_ms.size();
  s.elements_completed = sv->elements_done;
  s.batches_launched = sv->batches;
  s.peak_device_bytes = sv->peak_bytes;
  s.batches_rerun = sv->batches_rerun;
  s.jobs_failed = sv->jobs_failed;
  for (auto& slot : sv->slots) s.scratch_bytes += kc_context_scratch_bytes(slot->ctx);
  if (!sv->latencies_ms.empty()) {
    std::vector<double> v = sv->latencies_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      size_t k = static_cast<size_t>(p * static_cast<double>(v.size() - 1) + 0.5);
      return v[std::min(k, v.size() - 1)];
    };
    s.p50_latency_ms = pct(0.50);
    s.p99_latency_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    s.mean_latency_ms = sum / static_cast<double>(v.size());
    const double span = std::chrono::duration<double>(sv->last_finish - sv->first_submit).count();
    if (span > 0) {
      s.jobs_per_s = static_cast<double>(v.size()) / span;
      s.elements_per_s = static_cast<double>(sv->elements_done) / span;
    }
  }
  *out = s;
  return KC_OK;
}

kc_status kc_server_reset_stats(void* vsv) {
  auto* sv = static_cast<Server*>(vsv);
  if (!sv) return KC_EINVAL;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->latencies_ms.clear();
  sv->elements_done = 0;
  sv->batches = 0;
  sv->batches_rerun = 0;
  sv->jobs_failed = 0;
  sv->have_first = false;
  sv->peak_bytes = sv->bytes_in_use;
  return KC_OK;
}

}  // extern "C"
