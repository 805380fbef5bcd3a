// eg_pr.h -- the PageRank row update shared by the pull kernel (eg_pull.hip)
// and the binned two-phase kernels (eg_binned.hip).
//
// Reference: the PR iterator, pagerank_simple.c:110-133 (rank[v] =
// (1-d)/n + d * sum_{u in in(v)} prev[u]/deg(u)), with the dangling and
// convergence statistics of this library reduced in FP64.
#pragma once

#include "eg_internal.h"

struct PrArgs {
  const float *rank_old;  // for the L1 delta
  float *contrib_new;
  const uint32_t *out_deg;
  const double *dmass_in;  // dangling mass of the input iterate
  double *stats_out;       // [dangling of output, l1]
  float base;              // (1 - d) / n
  const float *base_vec;   // per-row term added to base (personalized PageRank), or null
  float damping;
  double damping_over_n;   // d / n, for the uniform dangling term
  int uniform;
  uint32_t hot;            // pull kernel: gathers of x[u] with u >= hot are non-temporal loads
};

__device__ __forceinline__ float pr_base(const PrArgs &a) {
  double b = (double)a.base;
  if (a.uniform) b += a.damping_over_n * (*a.dmass_in);
  return (float)b;
}

// Writes the row result; for PageRank also the new contribution and the FP64
// partials (dangling mass of the new iterate, |r' - r|).
template <bool PR>
__device__ __forceinline__ void finish_row(uint32_t r, float sum, float *__restrict__ y,
                                           const PrArgs &a, float basep, uint32_t deg,
                                           float rold, double &dang, double &l1) {
  if (PR) {
    float b = basep;
    if (a.base_vec) b += a.base_vec[r];
    const float v = b + a.damping * sum;
    y[r] = v;
    a.contrib_new[r] = deg ? v / (float)deg : 0.f;
    if (deg == 0) dang += (double)v;
    l1 += fabs((double)v - (double)rold);
  } else {
    y[r] = sum;
  }
}

// One binned PageRank iteration (eg_binned.hip) from contrib into y and
// a.contrib_new, with the statistics of the pull kernel; S is the fixed-point
// scale of eg_pb_scale.  The plan must have been built for the orientation.
// skip_static: leave out the work items of rows without entries.  Valid from
// the third iteration of a run with a constant base term (dangling mass
// dropped, no base vector): such a row's value is the base term from the
// first iteration on, so both iterate buffers hold it after two, and its
// item's statistics (dangling sum, zero change) repeat exactly.
int eg_pb_iterate(eg_graph *g, const float *contrib, float *y, const PrArgs &a, int S,
                  hipStream_t stream, bool skip_static = false);
