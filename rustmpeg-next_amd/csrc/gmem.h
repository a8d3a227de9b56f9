// gmem.h - global-address-space views of device pointers (device code only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rmh {

// Picture planes, shadows and edge maps live in device memory, but reach the
// kernels as generic pointers inside descriptor structs, whose accesses the
// compiler emits as flat loads and stores. Flat accesses also count against
// the LDS wait counter, so every LDS wait then waits for the outstanding
// global loads as well. Casting to address space 1 gives global accesses.
#define RMH_GLOBAL __attribute__((address_space(1)))

// Plain vector types for wide copies through such pointers (HIP's vector
// classes have no assignment operators for address space 1).
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ const RMH_GLOBAL T* gsrc(const void* p) {
    return (const RMH_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ RMH_GLOBAL T* gdst(void* p) {
    return (RMH_GLOBAL T*)p;
}

}  // namespace rmh
