// ccd_serve.cpp -- continuous-batching fit server over the slot engine.
//
// No reference counterpart: Cyclops fits one model per call.  Requests are fits
// over host data (uploaded on admission) or over resident data.  Two host
// threads run the service:
//   admission -- FCFS.  The head request is admitted when one arena block can
//                hold its data and fit state and a slot is free (admission by
//                HBM footprint).  It is uploaded and set up (k_slot_init) on the
//                admission stream while the device keeps sweeping other fits.
//   scheduling -- one sweep launch over every fit in flight.  Each fit runs
//                whole CCD cycles until a time quantum has passed, stops at a
//                coordinate boundary and resumes there in the next launch, so a
//                large fit never holds small ones back.  Between launches,
//                finished fits leave (final log likelihood, coefficients copied
//                out, block released) and admitted fits join.
// A fit's arithmetic is the same alone or among others and wherever its sweeps
// pause: one workgroup owns it and its reductions are in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "ccd_host.h"
#include "ccd_slot.h"

using namespace ccdamd;
using Clock = std::chrono::steady_clock;

namespace {

double secondsBetween(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

struct Request {
  int64_t id = 0;
  ccdh_data_desc desc{};
  bool hasDesc = false;
  ccdh_data* data = nullptr;  // caller's resident data, or created on admission
  bool ownsData = false;
  std::vector<int32_t> ptype;
  std::vector<double> pvar, beta0, weights, heldout;
  std::vector<uint8_t> fixed;
  ccdh_fit_spec spec{};
  int model = 0, K = 0, J = 0;
  bool hasOffset = false;
  size_t dataBytes = 0, footprint = 0;
  SlotLayout layout{};
  char* block = nullptr;
  std::vector<unsigned char> slotHost;  // Slot<R> descriptor
  FitStatus st0{};
  int slot = -1;
  int boundCount = 1;
  bool oneGroup = false;  // restarted on one workgroup after its gang's wait gave up
  int64_t launches = 0;
  Clock::time_point tSubmit, tAdmit, tReady, tDone;
  bool done = false;
  int status = 0;
  std::string error;
  std::vector<double> beta;
  ccdh_fit_result result{};
};

}  // namespace

struct ccdh_server {
  virtual ~ccdh_server() = default;
  virtual int submit(Request* r) = 0;
  virtual int poll(int64_t id) = 0;
  virtual int wait(int64_t id, double timeoutMs, double* beta, ccdh_fit_result* res,
                   ccdh_request_info* info) = 0;
  virtual void stats(ccdh_server_stats* out) = 0;
  virtual void shutdown() = 0;
  int precision = 0;
};

namespace {

constexpr size_t kTail = 4096;  // per-request set-up descriptor, status and index

template <typename R>
class Server final : public ccdh_server {
 public:
  ~Server() override { shutdown(); }

  int init(const ccdh_server_options& o) {
    opt_ = o;
    precision = o.precision;
    maxSlots_ = o.max_slots > 0 ? o.max_slots : 1024;
    threads_ = o.threads > 0 ? o.threads : 256;
    // 512 threads (any size above 256) take 4 items per lane: the same pass
    // width as 256 x 8, two waves per SIMD
    itl_ = threads_ > 256 ? 4 : 8;
    maxGang_ = o.max_gang > 0 ? o.max_gang : kShards;
    capacity_ = slotGangCapacity(kMixed, threads_, itl_);
    if (capacity_ <= 0) return setError(CCDH_E_DEVICE, "sweep kernel occupancy query failed");
    (void)xcdCount();
    const double qms = o.quantum_ms > 0 ? o.quantum_ms : 20.0;
    quantumTicks_ = (long long)(qms * 1e5);  // device wall clock: 100 MHz
    size_t budget = (size_t)std::max<int64_t>(o.hbm_budget_bytes, 0);
    if (budget == 0) {
      size_t freeB = 0, totalB = 0;
      if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) return setError(CCDH_E_DEVICE, "hipMemGetInfo");
      budget = freeB / 2;
    }
    if (hipMalloc((void**)&arenaBase_, budget) != hipSuccess)
      return setError(CCDH_E_NOMEM, "server arena of " + std::to_string(budget) + " bytes");
    arena_.reset(arenaBase_, budget);
    if (hipMalloc((void**)&dSlots_, maxSlots_ * sizeof(Slot<R>)) != hipSuccess ||
        hipMalloc((void**)&dSt_, maxSlots_ * sizeof(FitStatus)) != hipSuccess ||
        hipMalloc((void**)&dMap_, std::max(maxSlots_, capacity_) * sizeof(GangEntry)) != hipSuccess ||
        hipMalloc((void**)&dXch_, maxSlots_ * sizeof(Xch)) != hipSuccess ||
        hipMalloc((void**)&dFinish_, maxSlots_ * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&dOne_, sizeof(int32_t)) != hipSuccess)
      return setError(CCDH_E_NOMEM, "server tables");
    if (hipHostMalloc((void**)&hSt_, maxSlots_ * sizeof(FitStatus), hipHostMallocDefault) != hipSuccess)
      return setError(CCDH_E_NOMEM, "server status buffer");
    if (hipStreamCreateWithFlags(&schedStream_, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&admitStream_, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ev0_) != hipSuccess || hipEventCreate(&ev1_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "server streams");
    hSlots_.resize(maxSlots_);
    slotReq_.assign(maxSlots_, nullptr);
    for (int i = maxSlots_ - 1; i >= 0; --i) freeSlots_.push_back(i);
    admitThread_ = std::thread([this] { admitLoop(); });
    schedThread_ = std::thread([this] { schedLoop(); });
    started_ = true;
    return 0;
  }

  int submit(Request* r) override {
    const size_t rs = sizeof(R);
    if (r->hasDesc) {
      const ccdh_data_desc& d = r->desc;
      int64_t maxLen = 1;
      for (int j = 0; j < d.J; ++j) maxLen = std::max<int64_t>(maxLen, d.col_ptr[j + 1] - d.col_ptr[j]);
      int nCap = d.K;
      if (d.model == CCDH_CONDITIONAL_LOGISTIC) {
        int ns = 1;
        if (d.strata)
          for (int k = 0; k < d.K; ++k) ns = std::max(ns, d.strata[k] + 1);
        nCap = ns;
      }
      r->dataBytes = (size_t)dataFootprint(&d);
      r->layout = slotLayout(d.model, d.K, d.J, nCap, maxLen, r->spec.heldout != nullptr, rs);
    } else {
      ccdh_data* d = r->data;
      r->dataBytes = 0;
      r->layout = slotLayout(d->model, d->K, d->J, slotGroups(d, r->spec.weights),
                             maxColumnLength(d), r->spec.heldout != nullptr, rs);
    }
    r->footprint = r->dataBytes + r->layout.bytes + kTail;
    std::lock_guard<std::mutex> g(mu_);
    if (stopping_) return setError(CCDH_E_STATE, "server is shutting down");
    if (broken_) return setError(CCDH_E_DEVICE, "server stopped after an error: " + brokenMsg_);
    r->id = ++nextId_;
    r->tSubmit = Clock::now();
    reqs_[r->id] = r;
    queue_.push_back(r);
    ++submitted_;
    cvQueue_.notify_one();
    return 0;
  }

  int poll(int64_t id) override {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return setError(CCDH_E_INVALID, "unknown request id");
    return it->second->done ? 1 : 0;
  }

  int wait(int64_t id, double timeoutMs, double* beta, ccdh_fit_result* res,
           ccdh_request_info* info) override {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) return setError(CCDH_E_INVALID, "unknown request id");
    Request* r = it->second;
    auto isDone = [r] { return r->done; };
    if (timeoutMs < 0) {
      cvDone_.wait(lk, isDone);
    } else if (!cvDone_.wait_for(lk, std::chrono::duration<double, std::milli>(timeoutMs), isDone)) {
      return setError(CCDH_E_STATE, "request not finished before the timeout");
    }
    reqs_.erase(it);
    lk.unlock();
    const int status = r->status;
    if (status == 0) {
      if (beta) std::memcpy(beta, r->beta.data(), r->beta.size() * sizeof(double));
      if (res) *res = r->result;
    }
    if (info) {
      info->latency = secondsBetween(r->tSubmit, r->tDone);
      info->queued = r->tAdmit.time_since_epoch().count() ? secondsBetween(r->tSubmit, r->tAdmit) : 0.0;
      info->prepare = r->tReady.time_since_epoch().count() ? secondsBetween(r->tAdmit, r->tReady) : 0.0;
      info->footprint = (int64_t)r->footprint;
      info->launches = r->launches;
    }
    const std::string err = r->error;
    delete r;
    if (status) return setError(status, err);
    return 0;
  }

  void stats(ccdh_server_stats* out) override {
    std::lock_guard<std::mutex> g(mu_);
    out->submitted = submitted_;
    out->completed = completed_;
    out->failed = failed_;
    out->in_flight = held_;
    out->queued = (int64_t)queue_.size();
    out->launches = launches_;
    out->kernel_ms = kernelMs_;
    out->arena_bytes = (int64_t)arena_.capacity();
    out->arena_used = (int64_t)arena_.used();
    out->arena_peak = (int64_t)arena_.peak();
    out->gang_restarts = gangRestarts_;
  }

  void shutdown() override {
    if (started_) {
      {
        std::lock_guard<std::mutex> g(mu_);
        stopping_ = true;
      }
      cvQueue_.notify_all();
      cvReady_.notify_all();
      if (admitThread_.joinable()) admitThread_.join();
      if (schedThread_.joinable()) schedThread_.join();
      started_ = false;
    }
    for (auto& kv : reqs_) {
      releaseData(kv.second);
      delete kv.second;
    }
    reqs_.clear();
    if (schedStream_) (void)hipStreamSynchronize(schedStream_);
    if (admitStream_) (void)hipStreamSynchronize(admitStream_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    ev0_ = ev1_ = nullptr;
    void* ptrs[] = {arenaBase_, dSlots_, dSt_, dMap_, dXch_, dFinish_, dOne_};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    arenaBase_ = nullptr;
    dSlots_ = nullptr;
    dSt_ = nullptr;
    dMap_ = nullptr;
    dXch_ = nullptr;
    dFinish_ = dOne_ = nullptr;
    if (hSt_) (void)hipHostFree(hSt_);
    hSt_ = nullptr;
    if (schedStream_) (void)hipStreamDestroy(schedStream_);
    if (admitStream_) (void)hipStreamDestroy(admitStream_);
    schedStream_ = admitStream_ = nullptr;
  }

 private:
  static constexpr bool kMixed = std::is_same<R, float>::value;

  // -------------------------------------------------------------- admission
  void admitLoop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      if (queue_.empty()) {
        if (stopping_ || broken_) {
          cvReady_.notify_all();
          return;
        }
        cvQueue_.wait(lk);
        continue;
      }
      if (broken_) {
        failAllQueuedLocked();
        cvReady_.notify_all();
        return;
      }
      Request* r = queue_.front();
      if (r->footprint > arena_.capacity()) {
        queue_.pop_front();
        finishLocked(r, CCDH_E_NOMEM,
                     "request needs " + std::to_string(r->footprint) +
                         " device bytes, more than the server's memory budget of " +
                         std::to_string(arena_.capacity()));
        continue;
      }
      if (held_ >= maxSlots_) {
        cvQueue_.wait(lk);
        continue;
      }
      void* blk = arena_.alloc(r->footprint);
      if (!blk) {
        cvQueue_.wait(lk);  // woken when a finished fit releases its block
        continue;
      }
      queue_.pop_front();
      ++held_;
      ++preparing_;
      r->block = (char*)blk;
      r->tAdmit = Clock::now();
      lk.unlock();
      const int rc = prepare(r);
      const std::string msg = rc ? std::string(ccdh_last_error()) : std::string();
      lk.lock();
      --preparing_;
      if (rc || broken_) {
        releaseData(r);
        --held_;
        finishLocked(r, rc ? rc : CCDH_E_DEVICE, rc ? msg : "server stopped after an error: " + brokenMsg_);
      } else {
        r->tReady = Clock::now();
        ready_.push_back(r);
      }
      cvReady_.notify_one();
    }
  }

  int prepare(Request* r) {
    int rc = 0;
    if (r->hasDesc) {
      Bump b;
      b.p = r->block;
      b.left = r->dataBytes;
      rc = dataCreate(&r->desc, &r->data, &b);
      if (rc) return rc;
      r->ownsData = true;
    }
    char* slotBlock = r->block + r->dataBytes;
    Slot<R> s;
    rc = slotPrepare<R>(r->data, &r->spec, 1, slotBlock, r->layout, &s, &r->st0, admitStream_);
    if (rc) return rc;
    s.ctl.spinLimit = opt_.spin_limit > 0 ? opt_.spin_limit : 0;
    // set-up kernel for this fit alone, on the admission stream
    char* tail = slotBlock + r->layout.bytes;
    auto* dS = (Slot<R>*)tail;
    auto* dF = (FitStatus*)(tail + 1024);
    auto* dA = (int32_t*)(tail + 2048);
    const int32_t zero = 0;
    static_assert(sizeof(Slot<R>) <= 1024 && sizeof(FitStatus) <= 1024, "set-up tail too small");
    if (hipMemcpyAsync(dS, &s, sizeof(s), hipMemcpyHostToDevice, admitStream_) != hipSuccess ||
        hipMemcpyAsync(dF, &r->st0, sizeof(FitStatus), hipMemcpyHostToDevice, admitStream_) != hipSuccess ||
        hipMemcpyAsync(dA, &zero, sizeof(zero), hipMemcpyHostToDevice, admitStream_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "request set-up copies");
    if (r->model == CCDH_COX) {  // risk-set groups under the request's weights
      rc = slotLaunchCoxGroups(kMixed, dS, dF, dA, 1, threads_, admitStream_);
      if (rc) return rc;
    }
    rc = slotLaunchInit(M_ANY, kMixed, dS, dF, dA, 1, threads_, admitStream_, true,
                        initRanks(r->data->nnz, r->K));
    if (rc) return rc;
    if (hipMemcpyAsync(&r->st0, dF, sizeof(FitStatus), hipMemcpyDeviceToHost, admitStream_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "request set-up status");
    const hipError_t e = hipStreamSynchronize(admitStream_);
    if (e != hipSuccess) return setError(CCDH_E_DEVICE, std::string("request set-up: ") + hipGetErrorString(e));
    s.N = r->st0.N;  // Cox: as the device built it
    r->slotHost.resize(sizeof(Slot<R>));
    std::memcpy(r->slotHost.data(), &s, sizeof(s));
    r->beta.assign(r->J, 0.0);
    return 0;
  }

  // ------------------------------------------------------------- scheduling
  void schedLoop() {
    std::vector<int32_t> active, finishing;
    for (;;) {
      std::vector<Request*> joining;
      {
        std::unique_lock<std::mutex> lk(mu_);
        while (ready_.empty() && active.empty()) {
          if (broken_ || (stopping_ && queue_.empty() && preparing_ == 0)) return;
          cvReady_.wait(lk);
        }
        joining.swap(ready_);
      }
      int rc = 0;
      for (Request* r : joining) {
        const int slot = freeSlots_.back();
        freeSlots_.pop_back();
        r->slot = slot;
        slotReq_[slot] = r;
        std::memcpy(&hSlots_[slot], r->slotHost.data(), sizeof(Slot<R>));
        hSlots_[slot].fit0 = slot;  // its status in the server's table
        if (!rc && (hipMemcpyAsync(dSlots_ + slot, &hSlots_[slot], sizeof(Slot<R>),
                                   hipMemcpyHostToDevice, schedStream_) != hipSuccess ||
                    hipMemcpyAsync(dSt_ + slot, &r->st0, sizeof(FitStatus), hipMemcpyHostToDevice,
                                   schedStream_) != hipSuccess))
          rc = setError(CCDH_E_DEVICE, "slot upload");
        active.push_back(slot);
      }
      if (!rc) rc = runLaunch(active);
      if (!rc) rc = retire(active, finishing);
      if (rc) {
        fail(active, ccdh_last_error());
        return;
      }
    }
  }

  // One launch over every fit in flight; each fit gets a gang sized by how many
  // fits share the device and by its own work (nonzeros; Cox adds its pid scans).
  int runLaunch(const std::vector<int32_t>& active) {
    std::vector<GangEntry> want;
    std::vector<double> prio;
    std::vector<int64_t> nnz;
    std::vector<int> models;
    for (int32_t slot : active) {
      const Request* r = slotReq_[slot];
      want.push_back(GangEntry{slot, 0, 1, 0});
      const int64_t w = slotWork(r->data, r->st0.N);  // Cox: nonzeros + a pid scan per coordinate
      nnz.push_back(w);
      prio.push_back((double)w);  // spare CUs to the largest fits
      models.push_back(r->model);
    }
    if (maxGang_ > 1) assignGangs(want, prio, nnz, models, capacity_, maxGang_, kGangMinNnz);
    for (GangEntry& g : want)
      if (slotReq_[g.slot]->oneGroup) g.G = 1;
    bool gangs = false;
    for (const GangEntry& g : want) gangs |= g.G > 1;
    if (gangs && !layoutGangs(want, capacity_, xcdCount(), map_)) gangs = false;
    if (!gangs) {
      map_.clear();
      for (int32_t slot : active) map_.push_back(GangEntry{slot, 0, 1, 0});
    }
    const int n = (int)map_.size();
    if (hipMemcpyAsync(dMap_, map_.data(), n * sizeof(GangEntry), hipMemcpyHostToDevice,
                       schedStream_) != hipSuccess ||
        hipEventRecord(ev0_, schedStream_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "launch set-up");
    int rc = slotLaunchSweeps(kMixed, dSlots_, dSt_, dMap_, n, dXch_, gangs, threads_, itl_, 1 << 30,
                              quantumTicks_, schedStream_);
    if (rc) return rc;
    if (hipEventRecord(ev1_, schedStream_) != hipSuccess ||
        hipMemcpyAsync(hSt_, dSt_, maxSlots_ * sizeof(FitStatus), hipMemcpyDeviceToHost,
                       schedStream_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "launch status copy");
    const hipError_t e = hipStreamSynchronize(schedStream_);
    if (e != hipSuccess) return setError(CCDH_E_DEVICE, std::string("sweep launch: ") + hipGetErrorString(e));
    for (int32_t slot : active) {
      if (hSt_[slot].fault == 1) {
        // A rank of the request's gang gave up waiting for its peers (other work
        // holding CUs can keep a rank from being scheduled): the fit starts over
        // from its starting coefficients with its current bound, on one
        // workgroup, where no rank waits for another; the other requests go on.
        Request* r = slotReq_[slot];
        r->oneGroup = true;
        hSt_[slot].fault = 0;
        if (hipMemcpyAsync(&dSt_[slot].fault, &hSt_[slot].fault, sizeof(int), hipMemcpyHostToDevice,
                           schedStream_) != hipSuccess ||
            hipMemcpyAsync(hSlots_[slot].beta, r->spec.beta0, r->J * sizeof(double),
                           hipMemcpyHostToDevice, schedStream_) != hipSuccess ||
            hipMemcpyAsync(dOne_, &slot, sizeof(int32_t), hipMemcpyHostToDevice, schedStream_) != hipSuccess)
          return setError(CCDH_E_DEVICE, "gang restart upload");
        int rc2 = slotLaunchInit(M_ANY, kMixed, dSlots_, dSt_, dOne_, 1, threads_, schedStream_, false,
                                 initRanks(r->data->nnz, r->K));
        if (rc2) return rc2;
        if (hipMemcpyAsync(&hSt_[slot], dSt_ + slot, sizeof(FitStatus), hipMemcpyDeviceToHost,
                           schedStream_) != hipSuccess ||
            hipStreamSynchronize(schedStream_) != hipSuccess)
          return setError(CCDH_E_DEVICE, "gang restart");
        std::lock_guard<std::mutex> g(mu_);
        ++gangRestarts_;
        continue;
      }
      if (hSt_[slot].fault == 2) {  // a gang spanned XCDs and backed off untouched
        maxGang_ = 1;
        hSt_[slot].fault = 0;
        if (hipMemcpyAsync(&dSt_[slot].fault, &hSt_[slot].fault, sizeof(int), hipMemcpyHostToDevice,
                           schedStream_) != hipSuccess ||
            hipStreamSynchronize(schedStream_) != hipSuccess)
          return setError(CCDH_E_DEVICE, "fault reset");
      }
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    std::lock_guard<std::mutex> g(mu_);
    kernelMs_ += ms;
    ++launches_;
    for (int32_t slot : active) slotReq_[slot]->launches++;
    return 0;
  }

  // Finished fits leave; ill-conditioned ones restart with a tenth of the bound
  // (CyclicCoordinateDescent::update, CyclicCoordinateDescent.cpp:586-615).
  int retire(std::vector<int32_t>& active, std::vector<int32_t>& finishing) {
    finishing.clear();
    for (int32_t slot : active) {
      const FitStatus& h = hSt_[slot];
      if (!h.done) continue;
      Request* r = slotReq_[slot];
      if (h.flag == CCDH_ILLCONDITIONED && r->boundCount < r->spec.max_bound_count) {
        r->boundCount++;
        Slot<R>& s = hSlots_[slot];
        s.initBound[0] /= 10.0;
        std::vector<double> hb(r->J, 0.0);
        if (r->hasOffset) hb[0] = r->spec.beta0[0];
        if (hipMemcpyAsync(s.beta, hb.data(), r->J * sizeof(double), hipMemcpyHostToDevice, schedStream_) != hipSuccess ||
            hipMemcpyAsync(dSlots_ + slot, &s, sizeof(s), hipMemcpyHostToDevice, schedStream_) != hipSuccess ||
            hipMemcpyAsync(dOne_, &slot, sizeof(int32_t), hipMemcpyHostToDevice, schedStream_) != hipSuccess)
          return setError(CCDH_E_DEVICE, "restart upload");
        int rc = slotLaunchInit(M_ANY, kMixed, dSlots_, dSt_, dOne_, 1, threads_, schedStream_, false,
                                initRanks(r->data->nnz, r->K));
        if (rc) return rc;
        if (hipStreamSynchronize(schedStream_) != hipSuccess) return setError(CCDH_E_DEVICE, "restart");
        continue;
      }
      finishing.push_back(slot);
    }
    if (finishing.empty()) return 0;
    const int n = (int)finishing.size();
    if (hipMemcpyAsync(dFinish_, finishing.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                       schedStream_) != hipSuccess)
      return setError(CCDH_E_DEVICE, "finish list");
    int rc = slotLaunchFinish(M_ANY, kMixed, dSlots_, dSt_, dFinish_, n, threads_, schedStream_);
    if (rc) return rc;
    for (int32_t slot : finishing) {
      Request* r = slotReq_[slot];
      if (hipMemcpyAsync(&hSt_[slot], dSt_ + slot, sizeof(FitStatus), hipMemcpyDeviceToHost, schedStream_) != hipSuccess ||
          hipMemcpyAsync(r->beta.data(), hSlots_[slot].beta, r->J * sizeof(double),
                         hipMemcpyDeviceToHost, schedStream_) != hipSuccess)
        return setError(CCDH_E_DEVICE, "result copy");
    }
    const hipError_t e = hipStreamSynchronize(schedStream_);
    if (e != hipSuccess) return setError(CCDH_E_DEVICE, std::string("finish: ") + hipGetErrorString(e));
    for (int32_t slot : finishing) {
      Request* r = slotReq_[slot];
      const FitStatus& h = hSt_[slot];
      ccdh_fit_result& res = r->result;
      res.log_likelihood = h.logLik;
      double lp = 0.0;
      for (int j = 0; j < r->J; ++j)
        lp += priorLogDensity(r->spec.prior_type[j], r->spec.prior_var[j], r->beta[j]);
      res.log_prior = lp;
      res.iterations = h.iteration;
      res.return_flag = h.flag;
      res.bound_count = r->boundCount;
      res.pred_log_likelihood = h.predLogLik;
      int st = 0;
      std::string msg;
      if (r->model == CCDH_COX && r->spec.heldout) {
        double v = 0.0;
        st = coxHeldOut(r->data, r->spec.weights, r->spec.heldout, r->beta.data(), &v);
        if (st) msg = ccdh_last_error();
        res.pred_log_likelihood = v;
      }
      slotReq_[slot] = nullptr;
      freeSlots_.push_back(slot);
      releaseData(r);
      std::lock_guard<std::mutex> g(mu_);
      --held_;
      r->tDone = Clock::now();
      res.seconds = secondsBetween(r->tSubmit, r->tDone);
      finishLocked(r, st, msg);
    }
    active.erase(std::remove_if(active.begin(), active.end(),
                                [&](int32_t s) { return slotReq_[s] == nullptr; }),
                 active.end());
    cvQueue_.notify_one();
    return 0;
  }

  // -------------------------------------------------------------- helpers
  void releaseData(Request* r) {
    if (r->ownsData && r->data) dataDestroy(r->data);  // placed: frees host copies only
    if (r->ownsData) r->data = nullptr;
    r->ownsData = false;
    if (r->block) arena_.release(r->block);
    r->block = nullptr;
  }

  void finishLocked(Request* r, int status, const std::string& msg) {
    r->status = status;
    r->error = msg;
    if (r->tDone.time_since_epoch().count() == 0) r->tDone = Clock::now();
    r->done = true;
    if (status)
      ++failed_;
    else
      ++completed_;
    cvDone_.notify_all();
    cvReady_.notify_all();  // the scheduler's drain condition may now hold
  }

  void failAllQueuedLocked() {
    while (!queue_.empty()) {
      Request* r = queue_.front();
      queue_.pop_front();
      finishLocked(r, CCDH_E_DEVICE, "server stopped after an error: " + brokenMsg_);
    }
  }

  void fail(std::vector<int32_t>& active, const std::string& msg) {
    std::lock_guard<std::mutex> g(mu_);
    broken_ = true;
    brokenMsg_ = msg;
    for (int32_t slot : active) {
      Request* r = slotReq_[slot];
      if (!r) continue;
      slotReq_[slot] = nullptr;
      --held_;
      releaseData(r);
      finishLocked(r, CCDH_E_DEVICE, msg);
    }
    active.clear();
    for (Request* r : ready_) {
      --held_;
      releaseData(r);
      finishLocked(r, CCDH_E_DEVICE, msg);
    }
    ready_.clear();
    failAllQueuedLocked();
    cvQueue_.notify_all();
  }

  static constexpr int64_t kGangMinNnz = 32768;
  ccdh_server_options opt_{};
  int64_t gangRestarts_ = 0;
  int maxSlots_ = 0, threads_ = 0, itl_ = 8, maxGang_ = 0, capacity_ = 0;
  std::vector<GangEntry> map_;
  GangEntry* dMap_ = nullptr;
  Xch* dXch_ = nullptr;
  long long quantumTicks_ = 0;
  Arena arena_;
  char* arenaBase_ = nullptr;
  Slot<R>* dSlots_ = nullptr;
  FitStatus* dSt_ = nullptr;
  FitStatus* hSt_ = nullptr;
  int32_t *dFinish_ = nullptr, *dOne_ = nullptr;
  hipStream_t schedStream_ = nullptr, admitStream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<Slot<R>> hSlots_;
  std::vector<Request*> slotReq_;
  std::vector<int> freeSlots_;

  std::mutex mu_;
  std::condition_variable cvQueue_, cvReady_, cvDone_;
  std::deque<Request*> queue_;
  std::vector<Request*> ready_;
  std::unordered_map<int64_t, Request*> reqs_;
  int64_t nextId_ = 0, submitted_ = 0, completed_ = 0, failed_ = 0, launches_ = 0;
  int held_ = 0, preparing_ = 0;
  double kernelMs_ = 0.0;
  bool stopping_ = false, broken_ = false, started_ = false;
  std::string brokenMsg_;
  std::thread admitThread_, schedThread_;
};

int copySpec(Request* r, const ccdh_fit_spec& sp, int K, int J) {
  r->ptype.assign(sp.prior_type, sp.prior_type + J);
  r->pvar.assign(sp.prior_var, sp.prior_var + J);
  r->fixed.assign(sp.fixed, sp.fixed + J);
  r->beta0.assign(sp.beta0, sp.beta0 + J);
  if (sp.weights) r->weights.assign(sp.weights, sp.weights + K);
  if (sp.heldout) r->heldout.assign(sp.heldout, sp.heldout + K);
  r->spec = sp;
  r->spec.prior_type = r->ptype.data();
  r->spec.prior_var = r->pvar.data();
  r->spec.fixed = r->fixed.data();
  r->spec.beta0 = r->beta0.data();
  r->spec.weights = sp.weights ? r->weights.data() : nullptr;
  r->spec.heldout = sp.heldout ? r->heldout.data() : nullptr;
  r->K = K;
  r->J = J;
  return 0;
}

int checkSpecArrays(const ccdh_fit_spec* sp, int J) {
  if (!sp || !sp->prior_type || !sp->prior_var || !sp->fixed || !sp->beta0)
    return setError(CCDH_E_INVALID, "fit spec arrays required");
  if (sp->convergence_type < 0 || sp->convergence_type > 3)
    return setError(CCDH_E_INVALID, "unsupported convergence type");
  for (int j = 0; j < J; ++j)
    if (sp->prior_type[j] < 0 || sp->prior_type[j] > 2)
      return setError(CCDH_E_INVALID, "unsupported prior type");
  return 0;
}

}  // namespace

extern "C" int ccdh_server_create(const ccdh_server_options* options, ccdh_server_t* out) {
  if (!out) return setError(CCDH_E_INVALID, "null argument");
  *out = nullptr;
  ccdh_server_options o{};
  if (options) o = *options;
  if (o.precision != CCDH_FP64 && o.precision != CCDH_MIXED)
    return setError(CCDH_E_INVALID, "unsupported precision");
  if (o.max_slots < 0 || o.threads < 0 || o.quantum_ms < 0 || o.hbm_budget_bytes < 0)
    return setError(CCDH_E_INVALID, "negative server option");
  try {
    ccdh_server* s = nullptr;
    int rc;
    if (o.precision == CCDH_MIXED) {
      auto* t = new Server<float>();
      s = t;
      rc = t->init(o);
    } else {
      auto* t = new Server<double>();
      s = t;
      rc = t->init(o);
    }
    if (rc) {
      delete s;
      return rc;
    }
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}

extern "C" int ccdh_server_submit(ccdh_server_t server, const ccdh_data_desc* desc,
                                  const ccdh_fit_spec* spec, int64_t* request_id) {
  if (!server || !desc || !spec || !request_id) return setError(CCDH_E_INVALID, "null argument");
  if (desc->precision != server->precision)
    return setError(CCDH_E_INVALID, "request precision differs from the server's");
  if (desc->K <= 0 || desc->J <= 0 || !desc->col_ptr || !desc->fmt || !desc->y)
    return setError(CCDH_E_INVALID, "empty problem");
  if (desc->model != CCDH_POISSON && desc->model != CCDH_LOGISTIC &&
      desc->model != CCDH_CONDITIONAL_LOGISTIC && desc->model != CCDH_COX)
    return setError(CCDH_E_INVALID, "unsupported model type " + std::to_string(desc->model));
  int rc = checkSpecArrays(spec, desc->J);
  if (rc) return rc;
  try {
    auto* r = new Request();
    r->desc = *desc;
    r->hasDesc = true;
    r->model = desc->model;
    r->hasOffset = desc->has_offset != 0;
    copySpec(r, *spec, desc->K, desc->J);
    rc = server->submit(r);
    if (rc) {
      delete r;
      return rc;
    }
    *request_id = r->id;
    return 0;
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}

extern "C" int ccdh_server_submit_resident(ccdh_server_t server, ccdh_data_t data,
                                           const ccdh_fit_spec* spec, int64_t* request_id) {
  if (!server || !data || !spec || !request_id) return setError(CCDH_E_INVALID, "null argument");
  if (data->precision != server->precision)
    return setError(CCDH_E_INVALID, "data precision differs from the server's");
  int rc = checkSpecArrays(spec, data->J);
  if (rc) return rc;
  try {
    auto* r = new Request();
    r->data = data;
    r->model = data->model;
    r->hasOffset = data->hasOffset != 0;
    copySpec(r, *spec, data->K, data->J);
    rc = server->submit(r);
    if (rc) {
      delete r;
      return rc;
    }
    *request_id = r->id;
    return 0;
  } catch (const std::exception& e) {
    return setError(CCDH_E_DEVICE, e.what());
  }
}

extern "C" int ccdh_server_poll(ccdh_server_t server, int64_t request_id) {
  if (!server) return setError(CCDH_E_INVALID, "null argument");
  return server->poll(request_id);
}

extern "C" int ccdh_server_wait(ccdh_server_t server, int64_t request_id, double timeout_ms,
                                double* beta_out, ccdh_fit_result* result, ccdh_request_info* info) {
  if (!server) return setError(CCDH_E_INVALID, "null argument");
  return server->wait(request_id, timeout_ms, beta_out, result, info);
}

extern "C" int ccdh_server_get_stats(ccdh_server_t server, ccdh_server_stats* out) {
  if (!server || !out) return setError(CCDH_E_INVALID, "null argument");
  server->stats(out);
  return 0;
}

extern "C" int ccdh_server_destroy(ccdh_server_t server) {
  if (!server) return setError(CCDH_E_INVALID, "null argument");
  delete server;
  return 0;
}
