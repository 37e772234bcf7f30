// pr_policy.h -- the PageRank pull policy (src/pr/omp_base.cc:25-40) for the
// generic iteration engine, shared by the single-graph and multi-graph plans.
#ifndef GD_PR_POLICY_H_
#define GD_PR_POLICY_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pull_engine.h"

namespace gd {

struct PrGather {
  const float* __restrict__ contrib;
  __device__ __forceinline__ float value(int /*e*/, int u) const { return contrib[u]; }
};

// NT: 0 plain, 1 streamed arrays non-temporal, 2 also the new contrib
template <int NT>
struct PrEpilogue {
  float* __restrict__ scores;
  float* __restrict__ contrib_next;
  const int32_t* __restrict__ outdeg;
  float base;
  float damp;
  int seed;
  struct Pre {
    float old;
    int deg;
  };
  __device__ __forceinline__ Pre pre(int v) const {
    return Pre{ld_stream<(NT > 0)>(scores + v), ld_stream<(NT > 0)>(outdeg + v)};
  }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  // scores[v] = base + kDamp * sum (two roundings), next contrib, |change|
  // (src/pr/omp_base.cc:38-40, :25-27)
  __device__ __forceinline__ double finish(int v, float sum, const Pre& p) const {
    const float b = (seed < 0 || v == seed) ? base : 0.0f;
    const float nv = b + damp * sum;
    st_stream<(NT > 0)>(scores + v, nv);
    st_stream<(NT > 1)>(contrib_next + v, nv / (float)p.deg);
    return (double)fabsf(nv - p.old);
  }
};

struct PrPolicy {
  // hub rows in the reference's order and rounding: its sequential FP32 sum of
  // a 10^5-edge row drifts by ~1e-4 relative, which a tree sum does not match
  static constexpr bool kExactHubs = true;
  // every row of a row tile summed by its owner thread in edge order (long
  // rows by a pipelined one-lane chain): with the exact hubs, every row of
  // the plan is the reference's, bit for bit
  static constexpr bool kSeqAll = true;
  // row tiles may run on 1024-thread workgroups (16K-edge tiles sorted by source)
  static constexpr bool kWideTiles = true;
  struct Args {
    float* scores;
    float* c0;
    float* c1;
    const int32_t* outdeg;
    float base;
    float damp;
    int seed;
  };
  template <int NT>
  __device__ static PrGather gather(const Args& a, int odd) {
    return PrGather{odd ? a.c1 : a.c0};
  }
  template <int NT>
  __device__ static PrEpilogue<NT> epilogue(const Args& a, int odd) {
    return PrEpilogue<NT>{a.scores, odd ? a.c0 : a.c1, a.outdeg, a.base, a.damp, a.seed};
  }
};

}  // namespace gd

#endif  // GD_PR_POLICY_H_
