// FP32 instantiation of every launcher (the FP32-mixed form: float fields
// and stencils, FP64 reductions and residuals). FMA contraction allowed.
#include "glx_launch_impl.hpp"

template struct glx::L<float>;
