// Response functions: the state-tangent kernels and the FP64 instantiations
// of the adjoint kernels of rhg_rf.hpp (the FP32 form's are in
// rhg_rf32.hip, compiled with FMA contraction and the FP32 form's FP64
// reciprocal).
#include "rhg_rf.hpp"

namespace rhg {

// ---------------------------------------------------------------------------
// 1. state tangents

// out = atmos with row `row` shifted by delta (every column and depth)
__global__ void k_rf_perturb(const double *__restrict__ atmos, double *__restrict__ out,
                             int64_t n, int Nmax, int row, double delta) {
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = (int) ((i / Nmax) % 9);
  out[i] = atmos[i] + (r == row ? delta : 0.0);
}

// out = (a - b) * scale, elementwise, in the output type
template <typename O>
__global__ void k_rf_diff(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
                          double scale, O *__restrict__ out) {
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (O) ((a[i] - b[i]) * scale);
}
template __global__ void k_rf_diff<double>(const double *, const double *, int64_t, double, double *);
template __global__ void k_rf_diff<float>(const double *, const double *, int64_t, double, float *);

RHG_RF_INST(double)

}  // namespace rhg
