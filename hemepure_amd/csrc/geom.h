// geom.h -- the owned geometry object behind hlb_geom (generators and the
// .gmy reader fill it; hlb_geom_view exposes it as an hlb_geometry).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/hemepure_amd.h"

struct hlb_geom {
  int32_t blocks[3] = {0, 0, 0};
  int32_t bs = 8;
  std::vector<int32_t> coords;
  std::vector<uint8_t> link_type;
  std::vector<int32_t> link_iolet;
  std::vector<float> link_dist;
  std::vector<uint8_t> normal_avail;
  std::vector<float> normal;
  std::vector<hlb_iolet_geom> iolets;
};
