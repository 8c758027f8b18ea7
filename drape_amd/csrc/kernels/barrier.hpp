// The IPC log barrier of the vertex-collider and self-contact terms: b(d) = -kappa (d - dhat)^2 ln(d/dhat)
// for d < dhat (0 beyond), with its first and second derivatives in d (contact.rs:81-108), in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace drape {

struct BarrierTerms {
  double value, first, second;
};

__device__ inline BarrierTerms barrier_terms(double d, double dhat, double kappa) {
  BarrierTerms b{0.0, 0.0, 0.0};
  if (d >= dhat) return b;
  const double t = d - dhat;
  const double ln = log(d / dhat);
  b.value = d <= 0.0 ? INFINITY : -kappa * (t * t) * ln;
  b.first = -kappa * (2.0 * t * ln + t * t / d);
  b.second = -kappa * (2.0 * ln + 4.0 * t / d - t * t / (d * d));
  return b;
}

}  // namespace drape
