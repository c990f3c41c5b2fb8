// The FP32 form of the response-function kernels (rhg_rf.hpp): float Voigt
// functions and continuum, FP64 recursion, adjoint and points, in a
// translation unit compiled with FMA contraction and the FP32 form's FP64
// reciprocal (rhg_device.hpp).
#define RHG_FP32_FORM_TU 1
#include "rhg_rf.hpp"

namespace rhg {

RHG_RF_INST(float)

}  // namespace rhg
