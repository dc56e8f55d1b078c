// aql_kernels.hip -- the two kernels aql_floor dispatches as AQL packets (built as a raw gfx950
// code object).  Neither reads the implicit kernel arguments, so the packets carry only the
// explicit ones.
#include <hip/hip_runtime.h>

// stores v into the flag after a system-scope fence
extern "C" __global__ void k_flag(int* flag, int v) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
// writes nothing when given a null pointer (it always is here)
extern "C" __global__ void k_nop(int* p) {
  if (p != nullptr && threadIdx.x == 0) p[0] = 1;
}
