// device_util.h -- addressing helpers shared by the step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "d3q19.h"
#include "lbm_kernels.h"

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

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// One lane of a slab: its slot, block-local (y, z) and boundary index.
struct Lane {
  uint32_t r;              // slab record (wave-uniform)
  uint32_t slot;
  int y, z;
  bool boundary;
  uint32_t b;              // compact boundary index (valid when boundary)
  uint64_t fm[kReach];     // reachable slabs' fluid masks (wave-uniform)
  int32_t base[kReach];    // and their first slots
};

// Number of set bits of the wave-uniform mask below this lane, plus add.
__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask, uint32_t add) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, add));
}

// Decodes slab record r for this lane; false for a position that is not
// fluid or a record past the end (the lane has nothing to do).  r must be
// wave-uniform; the record is read with scalar loads that do not wait on one
// another.
__device__ __forceinline__ bool slab_lane(const BrickTables& bt, uint32_t r, Lane& L) {
  if (r >= (uint64_t)bt.nslab) return false;
  const uint32_t lane = threadIdx.x & 63u;
  const HLB_GLOBAL SlabRec* g = (const HLB_GLOBAL SlabRec*)(bt.rec + r);
#pragma unroll
  for (int k = 0; k < kReach; ++k) {
    L.fm[k] = uniform64(g->fm[k]);
    L.base[k] = __builtin_amdgcn_readfirstlane(g->base[k]);
  }
  const uint64_t bm = uniform64(g->bmask);
  const int32_t bbase = __builtin_amdgcn_readfirstlane(g->bbase);
  const uint64_t fm = L.fm[reach_entry(0, 0, 0)];
  if (!((fm >> lane) & 1u)) return false;
  L.r = r;
  L.y = (int)(lane >> 3);
  L.z = (int)(lane & 7u);
  L.slot = mbcnt64(fm, (uint32_t)L.base[reach_entry(0, 0, 0)]);
  L.boundary = (bm >> lane) & 1u;
  L.b = L.boundary ? mbcnt64(bm, (uint32_t)bbase) : 0u;
  return true;
}

// Index among the lanes of the slab set in mask (uniform) below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  const uint32_t lane = threadIdx.x & 63u;
  return (uint32_t)__builtin_popcountll(mask & ((1ull << lane) - 1ull));
}

// Slot of the upstream site x - c_J of a lane (pull direction J), or the
// lane's own slot when that position is not fluid (the link is a wall or
// iolet link; its boundary delegate replaces the value).  The slab holding
// the position is one of at most four candidates from the record; the lane
// picks by whether its y / z step leaves the brick, then counts the fluid
// lanes below the position in that slab.
template <int J>
__device__ __forceinline__ uint32_t upstream_slot(const Lane& L) {
  constexpr int ox = -CX[J], oy = -CY[J], oz = -CZ[J];
  const int yy = L.y + oy, zz = L.z + oz;
  const bool yo = oy != 0 && (unsigned)yy > 7u;
  const bool zo = oz != 0 && (unsigned)zz > 7u;
  // constant entry indices only, so the masks and bases stay in scalar registers
  // (values are copied out before selecting: a select between array elements
  // would select their addresses and keep the record in scratch memory)
  constexpr int e00 = reach_entry(ox, 0, 0);
  uint64_t m = L.fm[e00];
  int32_t base = L.base[e00];
  if constexpr (oy != 0 && oz != 0) {
    static_assert(ox == 0, "D3Q19 links have at most two nonzero components");
    constexpr int e10 = reach_entry(0, oy, 0), e01 = reach_entry(0, 0, oz), e11 = reach_entry(0, oy, oz);
    const uint64_t m10 = L.fm[e10], m01 = L.fm[e01], m11 = L.fm[e11];
    const int32_t b10 = L.base[e10], b01 = L.base[e01], b11 = L.base[e11];
    m = yo ? (zo ? m11 : m10) : (zo ? m01 : m);
    base = yo ? (zo ? b11 : b10) : (zo ? b01 : base);
  } else if constexpr (oy != 0) {
    constexpr int e10 = reach_entry(ox, oy, 0);
    const uint64_t m10 = L.fm[e10];
    const int32_t b10 = L.base[e10];
    m = yo ? m10 : m;
    base = yo ? b10 : base;
  } else if constexpr (oz != 0) {
    constexpr int e01 = reach_entry(ox, 0, oz);
    const uint64_t m01 = L.fm[e01];
    const int32_t b01 = L.base[e01];
    m = zo ? m01 : m;
    base = zo ? b01 : base;
  }
  const uint32_t pos = ((uint32_t)(yy & 7) << 3) | (uint32_t)(zz & 7);
  const uint64_t bit = 1ull << pos;
  const uint32_t slot = (uint32_t)base + (uint32_t)__builtin_popcountll(m & (bit - 1ull));
  return (m & bit) ? slot : L.slot;
}

}  // namespace hlb
