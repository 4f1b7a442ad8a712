// Fixed-pattern builders (M restricted to S = pattern(A) | pattern(M0) | I)
// in FP64 or FP32-bulk form.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>

#include "engine.hpp"
#include "ops.hpp"

namespace spai {

EngineOutput RunFixedEngine(int64_t method, const MatrixData& A, const MatrixData* M0, const MatrixData* Pr,
                            const EngineParams& params);

// Device-resident fixed-pattern problem: A, S and S2 = pattern(A S) | S | I
// set up once, builds replayable with every operand in HBM.
constexpr int kFixedInfoLen = 11;

class FixedProblem {
 public:
  virtual ~FixedProblem() = default;
  // Runs the build from M0; returns the iterations completed.
  virtual int64_t Build() = 0;
  // The same, calling enqueue_after() (which queues more work on the build's
  // stream) before the build's final read-back, so that work needs no
  // synchronisation of its own.
  virtual int64_t BuildThen(const std::function<void()>& enqueue_after) = 0;
  virtual void Output(EngineOutput& out) = 0;
  // Device copy of the built M as FP64 CSC on S (for the PCG).
  virtual DMat MatrixM() = 0;
  // M's values in the order perm gives: out_d[i] = M at position perm[i] of
  // S, and out_f (when non-null) the same rounded to FP32, for i < len.
  virtual void GatherM(const int* perm, int64_t len, double* out_d, float* out_f) = 0;
  virtual const DMat& MatrixA() const = 0;
  virtual Ctx& context() = 0;
  // HIP-event timing, on the build's stream, of the per-iteration pass of the
  // fused LOMR or MR build (off by default).
  virtual void SetKernelTiming(bool on) = 0;
  // Average microseconds of that pass over the iterations the last build ran
  // (launches receives their number); -1 when the last build did not time it.
  virtual double KernelMicros(int64_t* launches) = 0;
  // Launch shape of the per-iteration pass: out[0] = 1 when the build uses
  // the fused pass, 2 the diagonal pass, 0 neither; then lanes per column
  // (GL), columns per tile (64 / GL), tiles, blocks of the resident grid,
  // wavefronts per block, nnz of S and of S2, and the diagonals of A, S and
  // S2 that the diagonal pass reads (0 without it).
  virtual void FusedInfo(int64_t out[kFixedInfoLen]) const = 0;
};

std::unique_ptr<FixedProblem> MakeFixedProblem(int64_t method, const MatrixData& A, const MatrixData* M0,
                                               const MatrixData* Pr, const EngineParams& params);

}  // namespace spai
