// Host orchestration of the SPAI builders on the device.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "methods_c_api.h"
#include "spai_ext.h"

namespace spai {

struct EngineParams {
  int64_t max_iterations = 100;
  double tolerance = 1e-9;
  double max_density = 0.03;
  bool enable_dropping = true;
  int64_t num_threads = 1;  // number of contiguous column batches
  int64_t inner_iterations = 2;
  bool fixed_pattern = false;
  bool fp32 = false;
  int64_t top_k = 0;
  bool batch_ordered = false;
};

// Host array allocated with new[] and left uninitialised (it is filled by a
// download), so the C ABI can hand it to the caller as a Result array, which
// free_mem() releases with delete[].
template <class T>
class HostArray {
 public:
  void resize(size_t n) {
    p_.reset(n > 0 ? new T[n] : nullptr);
    n_ = n;
  }
  T* data() { return p_.get(); }
  const T* data() const { return p_.get(); }
  size_t size() const { return n_; }
  T* release() {
    n_ = 0;
    return p_.release();
  }

 private:
  std::unique_ptr<T[]> p_;
  size_t n_ = 0;
};

struct HostMatrix {
  int64_t rows = 0, cols = 0;
  HostArray<int64_t> ptr, idx;
  HostArray<double> val;
};

struct EngineOutput {
  HostMatrix M;
  int64_t iterations = 0;
  bool converged = false;
  std::vector<SingleIterationData> history;
  double device_seconds = 0.0;
};

// Validates like src/common/methods_c_api.cpp:16-29 (CSC input) and
// src/common/methods_common.cpp:12-36 (problem); throws std::invalid_argument
// with the reference's messages.
void ValidateMatrixData(const MatrixData& d);

// Runs one of the five methods (SpaiMethod) on the current HIP device.
// check_csc: validate the CSC arrays as the C ABI does
// (src/common/methods_c_api.cpp:16-29); the C++ interface passes false, its
// matrices being well formed by construction.
EngineOutput RunEngine(int64_t method, const MatrixData& A, const MatrixData* M0, const MatrixData* Pr,
                       const EngineParams& params, bool check_csc = true);

EngineParams ToEngineParams(const MethodParams& p, const SpaiExtParams* ext);

}  // namespace spai
