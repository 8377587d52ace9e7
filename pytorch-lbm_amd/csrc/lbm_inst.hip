// Explicit instantiation of the launch sequences (and so of the step, init,
// observable and injection kernels) for one storage type, chosen by
// LBM_INST_PREC (0 fp32, 1 fp64, 2 bf16, 3 fp16): the build compiles this
// file once per precision, in parallel, and links the objects with
// lbm_api.hip.
#define LBM_INST_TU
#include "lbm_impl.hpp"

#if !defined(LBM_INST_PREC)
#error "LBM_INST_PREC must be 0 (fp32), 1 (fp64), 2 (bf16) or 3 (fp16)"
#elif LBM_INST_PREC == 0
LBM_SOLVER_TEMPLATES(, float, float, 2)
LBM_SOLVER_TEMPLATES(, float, float, 3)
LBM_BATCH_TEMPLATES(, float, float)
#elif LBM_INST_PREC == 1
LBM_SOLVER_TEMPLATES(, double, double, 2)
LBM_SOLVER_TEMPLATES(, double, double, 3)
#elif LBM_INST_PREC == 2
LBM_SOLVER_TEMPLATES(, float, __hip_bfloat16, 2)
LBM_SOLVER_TEMPLATES(, float, __hip_bfloat16, 3)
LBM_BATCH_TEMPLATES(, float, __hip_bfloat16)
#elif LBM_INST_PREC == 3
LBM_SOLVER_TEMPLATES(, float, __half, 2)
LBM_SOLVER_TEMPLATES(, float, __half, 3)
LBM_BATCH_TEMPLATES(, float, __half)
#else
#error "unknown LBM_INST_PREC"
#endif
