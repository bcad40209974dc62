// device_util.h -- addressing helpers shared by the step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hlb {

constexpr int kBlock = 256;  // threads (device slots) per workgroup tile
constexpr int kXcds = 8;     // accelerator complex dies on MI355X

// Loads and stores through a uniform plane base and a 32-bit element index:
// the byte offset stays 32-bit (planes are < 2^29 elements, checked at
// create time), so the address is an SGPR base plus one VGPR offset.  The
// casts name the global address space: pointers read from a table in memory
// would otherwise be generic and compile to flat accesses.
#define HLB_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ T ldp(const T* base, uint32_t i) {
  const char* p = reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T);
  return *(const HLB_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ T ldp_nt(const T* base, uint32_t i) {
  const char* p = reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T);
  return __builtin_nontemporal_load((const HLB_GLOBAL T*)p);
}
template <typename T>
__device__ __forceinline__ void stp(T* base, uint32_t i, T v) {
  char* p = reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T);
  *(HLB_GLOBAL T*)p = v;
}

template <typename T>
__device__ __forceinline__ T* plane(T* p, int64_t S, int j) {
  return p + (int64_t)j * S;
}

// XCD-contiguous workgroup order: the hardware hands workgroup w to XCD
// w mod 8, so logical tile (w mod 8) * (grid / 8) + w / 8 gives each XCD one
// contiguous eighth of the tiles (grid is a multiple of 8).
__device__ __forceinline__ uint32_t xcd_tile() {
  const uint32_t per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

__device__ __forceinline__ uint32_t site_index() { return xcd_tile() * kBlock + threadIdx.x; }

}  // namespace hlb
