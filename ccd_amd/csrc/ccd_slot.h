// ccd_slot.h -- fit slots of the device-resident engine.
//
// A slot is one fit in flight: a view of its (resident, read-only) data and
// pointers to its own per-row, per-group and per-coordinate state.  A sweep
// launch runs one workgroup per active slot, so fits over one data set
// (cross-validation folds x prior variances) and fits over different data
// sets (served requests of different N and P) share the same kernels.  All
// of a slot's state lives in one block of device memory, laid out by
// SlotLayout; the block is carved from a batch allocation or from the serving
// arena.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/ccdhip.h"
#include "ccd_device.h"

struct ccdh_data;

namespace ccdamd {

constexpr int M_ANY = -1;  // sweep kernel dispatching on each slot's model
constexpr int kMaxBundle = 8;  // fits of one slot moving in lock step

struct FitCtl {
  double tolerance;
  double initialBound;
  int maxIterations;
  int convergenceType;
  int useW;
  int solo;  // gangs: short columns on the lead rank alone (0) or never (-1)
  long long spinLimit;  // polls a gang rank waits for its peers before giving up (0: default)
};

// Status of one fit.  A slot's fits have consecutive statuses; the first also
// carries the slot's sweep position, fault and work counters.
struct FitStatus {
  double lastObj;
  double logLik;
  double predLogLik;
  int iteration;
  int flag;
  int done;
  int N;      // groups of this fit (Cox: tie groups under its weights)
  int jNext;  // coordinate the current sweep resumes at (a launch's time quantum ran out)
  int fault;  // a gang exchange gave up waiting (the launch's results are void)
  // work counters (thread 0), read by the host for the bandwidth figure
  long long gatherNnz, gatherRuns, updateNnz, updateRuns, scanPids, objRows;
  // wall-clock ticks (thread 0) spent per phase: single-pass columns, long columns, objective
  long long tSmall, tLarge, tObj, nLarge;
  // ticks per nonzero of the lead rank's last column of more than one pass,
  // carried from launch to launch (0 before the first)
  float longTpn;
};

// gather-pass records of long columns, read back by the update pass
template <typename R>
struct Records {
  int32_t* row;
  int32_t* key;
  uint8_t* flags;  // bit 0 own, bit 1 start
  R* o;
  R* xo;
  double* den;
};

// Per-row state: xBeta and exp(xBeta) side by side, so the random gather of
// one nonzero's row touches one cache line with one load (16 B in FP64).
template <typename R>
struct alignas(2 * sizeof(R)) RowSt {
  R x;  // xBeta
  R e;  // exp(xBeta) (offsExpXBeta)
};

// Per-group state (stratum / risk-set group): denominator and event weight nW.
struct alignas(16) GrpSt {
  double den;
  double nw;
};

// A slot holds F fits over one data set that move in lock step (a bundle;
// F = 1, 2, 4 or 8, and 1 for Cox): their per-row, per-group and per-coordinate
// arrays interleave the fits (entry i of fit f at i*F + f), so the F values one
// nonzero needs share a line.  The coordinates fixed at their start are common.
template <typename R>
struct Slot {
  DevX d;
  int model, K, J, N;
  int F, lF;  // fits, log2 F
  int fit0;   // status index of the first fit
  RowSt<R>* row;
  // logistic / Poisson: each (row, fit)'s denominator as a running sum of its
  // changes (the reference's per-row denomPid, ModelSpecifics.hpp:1567-1574);
  // interleaved like row; nullptr for the grouped models
  R* rden;
  double* kw;
  double* heldout;  // nullptr when the fits have no held-out weights
  GrpSt* grp;       // groups (grouped models only)
  double* numer;
  double* numer2;
  double* xjy;
  double* beta;
  double* bound;
  double* pvar;
  int8_t* ptype;
  uint8_t* fixed;
  int32_t* pid;
  int32_t* pidFirst;
  uint8_t* reset;
  int32_t* pidShard;  // Cox: first pid of each row shard (kShards + 1 entries, last = N)
  // Cox: 1 when the groups are risk-set pieces rather than tie groups (pids):
  // consecutive pids up to and including an event pid, split at strata and row
  // shards.  Only an event pid has nW != 0, so the gradient, Hessian and log
  // likelihood terms fall on pieces ending with one, whose accumulated sums
  // are the event pid's; the scans then walk pieces instead of pids.
  int coxPieces;
  Records<R> rec;
  int64_t recStride;  // records per buffer; the second buffer follows the first
  FitCtl ctl;
  double fixedLL[kMaxBundle];    // per fit: the Poisson -sum w log(y!) term
  double initBound[kMaxBundle];  // per fit: starting bound (/10 after each ill-conditioned restart)
};

// Cross-workgroup exchange of one slot's gang: an arrival counter (zeroed
// before each launch) and two alternating rounds of values per rank: kXchSum
// that are summed over the gang (gradient and Hessian of each fit), then
// kXchLead that only the lead rank publishes (coefficient and bound of each
// fit, end of quantum).
constexpr int kXchSum = 2 * kMaxBundle, kXchLead = 2 * kMaxBundle + 1;
struct Xch {
  unsigned cnt;
  unsigned pad[63];
  double v[2][kShards][kXchSum + kXchLead];
};

// Launch map entry: workgroup -> (slot, rank in the slot's gang, gang size).
struct GangEntry {
  int slot, rank, G, pad;
};

// Byte offsets of every array of one slot inside its block (256-byte aligned).
struct SlotLayout {
  size_t row, rden, kw, heldout, grp, numer, numer2, xjy, beta, bound, pvar, ptype, fixed,
      pid, pidFirst, reset, pidShard, recRow, recKey, recFlags, recO, recXo, recDen;
  int64_t recStride;
  size_t bytes;
};

// Layout for F fits of `model` over K rows, J coordinates, at most nCap groups
// and a longest column of maxLen nonzeros; rs = bytes of a per-row real.
SlotLayout slotLayout(int model, int K, int J, int nCap, int64_t maxLen, bool heldout, size_t rs,
                      int F = 1);

// Host-side preparation of one slot of F fits (specs sp[0..F), same fixed
// coordinates and controls; Cox: F = 1): groups (Cox pids under the fit's
// weights), the fixed log-likelihood terms, and the interleaved uploads of
// coefficients, priors, weights and groupings into the slot's block.  `out`
// receives the device pointers (fit0 = 0); st[0..F) the initial statuses.
// With sharedKw (F = 1) the slot reads those device weights, another slot's
// copy of the same host array, instead of uploading its own.
// Returns a ccdh status code.
template <typename R>
int slotPrepare(const ccdh_data* d, const ccdh_fit_spec* sp, int F, char* block,
                const SlotLayout& L, Slot<R>* out, FitStatus* st, hipStream_t s,
                const double* sharedKw = nullptr);

// Groups a fit of this data has under weights w (K for independent rows,
// strata for CLR, tie groups for Cox).
int slotGroups(const ccdh_data* d, const double* w);
int64_t maxColumnLength(const ccdh_data* d);

// Kernel launches over `n` slots listed in `active` (device array of slot
// indices into `slots`).  model is a CCDH model code or M_ANY.
// full = false: a restart from the starting coefficients that keeps the
// weight-dependent terms (nW, XjY) already in the slot.  ranks: workgroups per
// slot for the set-up (each takes a range of row shards); the starting
// objective follows in one workgroup per slot.  mask: the fits of each slot
// to set up (bit f; a restart sets up one fit of a bundle).
int slotLaunchInit(int model, bool mixed, const void* slots, FitStatus* st,
                   const int32_t* active, int n, int threads, hipStream_t s, bool full = true,
                   int ranks = 1, unsigned mask = ~0u);
// Set-up workgroups for a fit of nnz nonzeros over K rows (a power of two <= kShards).
int initRanks(int64_t nnz, int K);
// Cox slots listed in `active`: their risk-set groups under their weights
// (pid, first row and reset flag per group, rank boundaries, cleared
// numerators), the slot's N and the first fit's FitStatus::N, on the device.
// Runs before the set-up.
int slotLaunchCoxGroups(bool mixed, void* slots, FitStatus* st, const int32_t* active, int n,
                        int threads, hipStream_t s);
// One sweep launch over `n` workgroups described by `map` (device array).
// quantumTicks > 0: a slot stops at the first coordinate boundary after that
// many device clock ticks (100 MHz) and resumes there in the next launch.
// gangs: some slot runs on more than one workgroup (cooperative launch; the
// exchange counters of `xch`, indexed by slot, are reset first).
// itl: nonzeros per lane in one pass over a long column (8 or 4; 512-lane
// workgroups use 8).
int slotLaunchSweeps(bool mixed, const void* slots, FitStatus* st, const GangEntry* map, int n,
                     Xch* xch, bool gangs, int threads, int itl, int sweeps,
                     long long quantumTicks, hipStream_t s, int only = -1);
// Workgroups a cooperative sweep launch can hold at once.
int slotGangCapacity(bool mixed, int threads, int itl);
// XCDs of the device when workgroups are dealt to them round-robin by index
// (measured once with a probe kernel), else 0 (gangs are then not used).
int xcdCount();
// Map of one launch with every gang inside one XCD: gangs[i] = {slot, -, G};
// workgroup k*nX + b runs on XCD b.  False when they do not fit.
bool layoutGangs(const std::vector<GangEntry>& gangs, int capacity, int nX,
                 std::vector<GangEntry>& map);
// Gang size for one of n slots sharing `capacity` workgroups: the largest
// power of two <= min(kShards, maxGang) with n * G <= capacity and at least
// minNnzPerRank of the slot's work per rank (slotWork).
int gangSize(int n, int capacity, int64_t nnz, int model, int maxGang, int64_t minNnzPerRank);
// Gang sizes for slots sharing the device: the uniform size above, then spare
// CUs (one rank per CU) to the slots of highest priority, doubling one at a
// time; tiers = true: slots of equal priority double together, tier by tier.
void assignGangs(std::vector<GangEntry>& want, const std::vector<double>& priority,
                 const std::vector<int64_t>& nnz, const std::vector<int>& model, int capacity,
                 int maxGang, int64_t minNnzPerRank, bool tiers = false);
// Work of a fit in nonzero-equivalents for sizing its gang: its nonzeros, plus
// for Cox two per pid (the two scan passes of a coordinate).
int64_t slotWork(const ccdh_data* d, int N);
int slotLaunchFinish(int model, bool mixed, const void* slots, FitStatus* st,
                     const int32_t* active, int n, int threads, hipStream_t s);

}  // namespace ccdamd
