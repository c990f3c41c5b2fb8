// Response functions: the state-tangent kernels and the FP64 instantiations
// of the adjoint kernels of rhg_rf.hpp (the FP32 form's are in
// rhg_rf32.hip, compiled with FMA contraction and the FP32 form's FP64
// reciprocal).
#include "rhg_rf.hpp"

namespace rhg {

// ---------------------------------------------------------------------------
// 1. state tangents

// The perturbed copies of a batch's atmospheres [ncol][9][Nmax] (n values),
// side by side: copy s = blockIdx.y is the batch with row shift.row[s]
// moved by shift.h[s] at every column and depth.
__global__ void k_rf_perturb(const double *__restrict__ atmos, double *__restrict__ out,
                             int64_t n, int Nmax, RfShiftList shift) {
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = (int) blockIdx.y;
  const int r = (int) ((i / Nmax) % 9);
  out[(int64_t) s * n + i] = atmos[i] + (r == shift.row[s] ? shift.h[s] : 0.0);
}

// out[s * ncol + i] = ndep[i]: the depth counts of the copies side by side
__global__ void k_rf_rep_ndep(const int32_t *__restrict__ ndep, int ncol, int32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  out[(int64_t) blockIdx.y * ncol + i] = ndep[i];
}

// Central difference of two copies in a batch of row width W (a column
// state [row][W], a continuum [row][W]): out[r][c] = (S[r][off_p + c] -
// S[r][off_m + c]) * scale for nrows rows and ncol columns, in the output
// type.
template <typename O>
__global__ void k_rf_diff(const double *__restrict__ S, int64_t W, int64_t off_p, int64_t off_m,
                          int ncol, int64_t nrows, double scale, O *__restrict__ out) {
  const int64_t g = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nrows * ncol) return;
  const int64_t r = g / ncol, c = g - r * ncol;
  const double *row = S + r * W;
  out[g] = (O) ((row[off_p + c] - row[off_m + c]) * scale);
}
template __global__ void k_rf_diff<double>(const double *, int64_t, int64_t, int64_t, int, int64_t,
                                           double, double *);
template __global__ void k_rf_diff<float>(const double *, int64_t, int64_t, int64_t, int, int64_t,
                                          double, float *);

RHG_RF_INST(double)

}  // namespace rhg
