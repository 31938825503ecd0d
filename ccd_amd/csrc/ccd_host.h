// ccd_host.h -- host-side handle types and internal entry points of libccdhip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ccdhip.h"
#include "ccd_arena.h"
#include "ccd_device.h"

struct ccdh_data {
  int model = 0, precision = 0, K = 0, J = 0, nStrata = 0, hasOffset = 0;
  int64_t nnz = 0, nSeg = 0;
  int64_t devBytes = 0;
  // device
  int64_t* colPtr = nullptr;
  int32_t* rows = nullptr;
  double* vals = nullptr;
  float* valsF = nullptr;
  int8_t* fmt = nullptr;
  double* y = nullptr;
  double* time = nullptr;  // Cox only
  int32_t* strata = nullptr;
  int64_t* segPtr = nullptr;
  int32_t* segOff = nullptr;
  int32_t* segKey = nullptr;
  int32_t* stratStart = nullptr;
  int32_t* nzKey = nullptr;
  int32_t* shardRow = nullptr;
  int32_t* colShard = nullptr;
  int8_t* blk = nullptr;
  // row-major copy (ModelSpecifics::computeXBeta over the transpose) and the MM
  // step's column chunks; built on first use, never carved from `place`
  int64_t* tRowPtr = nullptr;   // [K+1]
  int32_t* tCol = nullptr;      // [entries] columns ascending within a row
  double* tVal = nullptr;       // [entries]; nullptr: every entry counts 1
  int32_t* mmChunk0 = nullptr;  // [J+1] first chunk of each column
  int32_t* mmChunkCol = nullptr;// [nChunks] column of each chunk
  int nChunks = 0;
  std::mutex lazyMu;  // clones on other threads may reach the first build together
  bool placed = false;  // device arrays carved from a caller's block (not freed here)
  // host copies of the small per-column / per-row vectors
  std::vector<int64_t> hColPtr, hSegPtr;
  std::vector<int8_t> hFmt;
  std::vector<double> hY, hOffs, hLogFact;
  std::vector<int32_t> hStrata;
  std::vector<int32_t> hShardRow;

  ccdamd::DevX view() const;
  int64_t hSegCount(int j) const { return hSegPtr.empty() ? 0 : hSegPtr[j + 1] - hSegPtr[j]; }
};

struct ccdh_fit {
  ccdh_data* d = nullptr;
  bool mixed = false;
  int N = 0;     // groups: K (independent), strata (CLR), tie groups (Cox)
  int nCap = 0;  // allocated group capacity (N+1 <= nCap)
  void* xb = nullptr;
  void* oeb = nullptr;
  // logistic / Poisson: each row's denominator (1 + exp(xBeta) / exp(xBeta)) as a
  // running sum of its changes, as the reference keeps denomPid per row
  // (ModelSpecifics.hpp:1567-1574, 1605-1609); nullptr for the grouped models
  void* rden = nullptr;
  double* kw = nullptr;
  double* w2 = nullptr;  // scratch weights (held-out scoring)
  double* xjy = nullptr;
  double* denom = nullptr;
  double* nw = nullptr;
  double* numer = nullptr;
  double* numer2 = nullptr;
  double* accTmp = nullptr;
  int32_t* pid = nullptr;
  int32_t* pidFirst = nullptr;
  uint8_t* resetFlag = nullptr;
  ccdamd::Seg<3>* tileAgg = nullptr;
  ccdamd::Seg<3>* tileCarry = nullptr;
  double* partials = nullptr;
  unsigned* ticket = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;  // pinned
  // MM step: per-row factor, per-chunk partial sums, results, coefficients, flags
  double* mmR = nullptr;
  double* mmPart = nullptr;
  double* mmGH = nullptr;
  double* mmBeta = nullptr;
  uint8_t* mmFixed = nullptr;
  hipEvent_t mmEv[4] = {nullptr, nullptr, nullptr, nullptr};  // chunk pass, xBeta: start/stop
  bool mmXbPending = false;  // an xBeta launch whose events are not yet read
  ccdh_mm_stats mmStats{};
  double fixedLL = 0.0;
  bool weightsSet = false, statsKnown = false;
  std::vector<double> hKw;
  hipStream_t stream = nullptr;
};

namespace ccdamd {

int setError(int code, const std::string& msg);

int dataCreate(const ccdh_data_desc* desc, ccdh_data** out, Bump* place = nullptr);
int64_t dataFootprint(const ccdh_data_desc* desc);
// the derived per-column tables of a data set whose arrays are resident (ccd_tables.hip)
int buildTablesDevice(ccdh_data* d, hipStream_t s);
void dataDestroy(ccdh_data* d);
int fitCreate(ccdh_data* d, ccdh_fit** out);
void fitDestroy(ccdh_fit* f);
int fitSetPids(ccdh_fit* f, const double* w);
int launchFill(ccdh_fit* f, double* p, int64_t n, double v);
void coxPids(const ccdh_data* d, const double* w, std::vector<int32_t>& pid, int* Nout,
             std::vector<uint8_t>& reset, std::vector<int32_t>& pidFirst);

int setWeights(ccdh_fit* f, const double* w, int useCV);
int fixedTermsLogLik(ccdh_fit* f, int useCV);
int fixedTermsGradHess(ccdh_fit* f, int useCV);
int zeroXBeta(ccdh_fit* f);
int axpyXBeta(ccdh_fit* f, double beta, int j);
int remainingStats(ccdh_fit* f, int useW);
int numerator(ccdh_fit* f, int j, int useW);
int gradHessLaunch(ccdh_fit* f, int j, int useW);
int gradHess(ccdh_fit* f, int j, int useW, double* g, double* h);
int updateXBeta(ccdh_fit* f, double delta, int j, int useW);
int logLik(ccdh_fit* f, int useCV, double* out);
int gradientObjective(ccdh_fit* f, int useCV, double* out);
int predLogLik(ccdh_fit* f, const double* w, double* out);
int getXBeta(ccdh_fit* f, double* out);
int computeXBetaRows(ccdh_fit* f, const double* beta);
int mmGradHess(ccdh_fit* f, const uint8_t* fixed, double* gh);
int mmStats(ccdh_fit* f, ccdh_mm_stats* out);

// batched / served engine helpers
int validateSpec(const ccdh_data* d, const ccdh_fit_spec& sp);
int coxHeldOut(ccdh_data* d, const double* weights, const double* heldout, const double* beta,
               double* out);

}  // namespace ccdamd
