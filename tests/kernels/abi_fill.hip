// A stand-alone kernel for the runtime-ABI tests: built as a gfx950 code
// object (not part of libbolt_hip.so), loaded with bolt_hip_module_load_data_ex
// and launched with the reference's kernel-parameter convention
// (src/exec/launch.rs:162-260: device pointers first, then 32-bit scalars).
#include <hip/hip_runtime.h>

extern "C" __global__ void abi_fill(unsigned long long* out, unsigned int n, unsigned int value) {
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (unsigned long long)value * i;
}
