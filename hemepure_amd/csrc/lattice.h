// lattice.h -- host-side LatticeData: site classification, the reference's
// class-major site order and the pull table the device kernels gather through.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/hemepure_amd.h"

namespace hlb {

// SiteType (src/geometry/SiteType.h:17-24).
enum : int32_t { kSolidType = 0, kFluidType = 1, kInletType = 2, kOutletType = 3 };

struct HostLattice {
  int32_t blocks[3] = {0, 0, 0};
  int32_t bs = 8;
  int64_t N = 0;
  int64_t off[7] = {0, 0, 0, 0, 0, 0, 0};  // class c occupies [off[c], off[c+1])
  std::vector<int32_t> coords;   // N*3
  std::vector<uint32_t> wall;    // N, bit d-1 set for a wall link d
  std::vector<uint32_t> iolet;   // N, bit d-1 set for an inlet/outlet link d
  std::vector<int32_t> ioletId;  // N
  std::vector<int32_t> siteType; // N
  std::vector<float> q;          // N*18 link distances as read
  std::vector<float> nrm;        // N*3 wall normal normalised in float (+inf if none)
  std::vector<uint32_t> pull;    // 18*N: source of direction j = 1..18, or the site itself
  // Site map: per block a slot (or -1), per slot 512 contiguous indices (or -1).
  std::vector<int64_t> blockSlot;
  std::vector<int32_t> slotSites;

  int64_t site_at(int64_t x, int64_t y, int64_t z) const;
  int collision_class(int64_t k) const;
};

// Builds the lattice; returns HLB_OK or an HLB_E* code with hlb_last_error set.
int build_lattice(const hlb_geometry& g, HostLattice& L);

// The reference's push table (n*19 int64, rubbish slot n*19), from the site map.
void push_table(const HostLattice& L, int64_t* out);

}  // namespace hlb
