// The FP32 form of the stage-4 kernels (rhg_formal.hpp), in a translation
// unit of its own so that it can be compiled with FMA contraction while the
// FP64 form keeps the reference's uncontracted operation order.
#define RHG_FP32_FORM_TU 1
#include "rhg_formal.hpp"

namespace rhg {

#define RHG_KFORMAL(KNAME, R)                                                                 \
  template __global__ void KNAME<R>(const DevModel *, const int32_t *, const int32_t *, int, int, \
                                    const double *, const int32_t *, const int32_t *, int, int,  \
                                    const R *, const R *, double, double *, const int32_t *, \
                                    const int32_t *, const int32_t *);
#define RHG_KSTOKES(R, MO, LT)                                                                \
  template __global__ void k_formal_stokes<R, MO, LT>(const DevModel *, const int32_t *,            \
                                                      const int32_t *, int, int, const double *,    \
                                                      const int32_t *, const int32_t *, int, int,   \
                                                      const R *, const R *, double, double *,       \
                                                      const int32_t *, const int32_t *,             \
                                                      const int32_t *, int, int);
RHG_KSTOKES(float, false, false)
RHG_KSTOKES(float, true, false)
RHG_KSTOKES(float, false, true)
RHG_KSTOKES(float, true, true)
#undef RHG_KSTOKES
RHG_KFORMAL(k_formal_scalar, float)
#undef RHG_KFORMAL

}  // namespace rhg
