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

// One lane of a slab: its slot, block-local coordinates and boundary index.
struct Lane {
  uint32_t r;        // slab record (wave-uniform)
  uint32_t slot;
  int x, y, z;       // x is wave-uniform
  bool boundary;
  uint32_t b;        // compact boundary index (valid when boundary)
  int32_t nb[18];    // the slab's reachable bricks (wave-uniform), as in SlabRec
};

// Decodes slab record r for this lane; false for a hole or a record past the
// end (the lane has nothing to do).  r must be wave-uniform; the record is
// read with scalar loads that do not wait on one another.
__device__ __forceinline__ bool slab_lane(const BrickTables& bt, uint32_t r, Lane& L) {
  if (r >= (uint64_t)bt.nslab) return false;
  const uint32_t lane = threadIdx.x & 63u;
  const HLB_GLOBAL SlabRec* g = (const HLB_GLOBAL SlabRec*)(bt.rec + r);
  const uint64_t fm = uniform64(g->fmask);
  const uint64_t bm = uniform64(g->bmask);
  const int32_t sid = __builtin_amdgcn_readfirstlane(g->sid);
  const int32_t bbase = __builtin_amdgcn_readfirstlane(g->bbase);
#pragma unroll
  for (int k = 0; k < 18; ++k) L.nb[k] = __builtin_amdgcn_readfirstlane(g->nb[k]);
  if (!((fm >> lane) & 1u)) return false;
  const uint32_t brick = (uint32_t)sid >> 3;
  L.r = r;
  L.x = sid & 7;
  L.y = (int)(lane >> 3);
  L.z = (int)(lane & 7u);
  L.slot = brick * (uint32_t)kBrickSites + (uint32_t)L.x * 64u + lane;
  L.boundary = (bm >> lane) & 1u;
  L.b = 0;
  if (L.boundary) {
    const uint64_t below = bm & ((1ull << lane) - 1ull);
    L.b = (uint32_t)bbase + (uint32_t)__builtin_popcountll(below);
  }
  return true;
}

// Index among the lanes of the slab set in mask (uniform) below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  const uint32_t lane = threadIdx.x & 63u;
  return (uint32_t)__builtin_popcountll(mask & ((1ull << lane) - 1ull));
}

// Slot of the upstream site x - c_J of a lane (pull direction J).  The brick
// holding it is one of at most four candidates from the slab's uniform
// neighbour list; the lane picks by whether its y / z step leaves the block.
// The slot is in bounds for every lane (a missing neighbour is the empty
// brick), so lanes whose link J is a wall or iolet read it too and the
// boundary delegates replace the value.
template <int J>
__device__ __forceinline__ uint32_t upstream_slot(const Lane& L) {
  constexpr int ox = -CX[J], oy = -CY[J], oz = -CZ[J];
  const bool xo = ox != 0 && (unsigned)(L.x + ox) > 7u;  // wave-uniform
  const bool yo = oy != 0 && (unsigned)(L.y + oy) > 7u;
  const bool zo = oz != 0 && (unsigned)(L.z + oz) > 7u;
  constexpr int iy = (oy + 1) * 3, iz = oz + 1;
  // Entries for (dy, dz) with each component either 0 or the step, from the
  // adjacent x layer of bricks when the x step leaves the brick (constant
  // indices only, so the list stays in scalar registers).
  auto pick = [&](int k) { return xo ? L.nb[9 + k] : L.nb[k]; };
  const int32_t c00 = pick(4);
  int32_t brick = c00;
  if constexpr (oy != 0 && oz != 0) {
    const int32_t c10 = pick(iy + 1);
    const int32_t c01 = pick(3 + iz);
    const int32_t c11 = pick(iy + iz);
    brick = yo ? (zo ? c11 : c10) : (zo ? c01 : c00);
  } else if constexpr (oy != 0) {
    const int32_t c10 = pick(iy + 1);
    brick = yo ? c10 : c00;
  } else if constexpr (oz != 0) {
    const int32_t c01 = pick(3 + iz);
    brick = zo ? c01 : c00;
  }
  const uint32_t local = ((uint32_t)((L.x + ox) & 7) * 8u + (uint32_t)((L.y + oy) & 7)) * 8u + (uint32_t)((L.z + oz) & 7);
  return (uint32_t)brick * (uint32_t)kBrickSites + local;
}

}  // namespace hlb
