// gmy_reader.cpp -- reads HemeLB geometry files (.gmy, version 4, and the
// .gmy+ variant with per-block weights) into an hlb_geom.
//
// Format (src/io/formats/geometry.h, src/io/formats/formats.h): XDR
// (big-endian) throughout.  A 32-byte preamble: HemeLB magic 0x686c6221,
// geometry magic 0x676d7904, version 4, blocks along x, y, z, sites per block
// side, one pad word.  Then one header record per block (x-major block
// order): fluid site count, [weight in .gmy+], compressed bytes, uncompressed
// bytes.  Then, for each block with data, a zlib stream of the block's site
// records in in-block order (z fastest): a fluid flag; for a fluid site, for
// each of the 26 neighbourhood displacements (geometry.cc order) a cut type,
// then a float distance for a wall cut or an iolet id and a float distance for
// an inlet/outlet cut; then a wall-normal flag and, if set, three floats.
// The parse follows GeometryReader::ReadPreamble / ReadHeader / ParseSite
// (src/geometry/GeometryReader.cc:382-446, 851-935); the 26 displacements
// are matched to D3Q19 directions, the other eight are read and dropped.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/hemepure_amd.h"
#include "d3q19.h"
#include "errors.h"
#include "geom.h"

namespace hlb {
namespace {

constexpr uint32_t kHemeLbMagic = 0x686c6221u;  // src/io/formats/formats.h:23
constexpr uint32_t kGmyMagic = 0x676d7904u;     // src/io/formats/geometry.h:46-49
constexpr uint32_t kGmyVersion = 4u;

// The 26 displacements of the file's neighbourhood, in file order
// (src/io/formats/geometry.cc:14-52): x slowest, then y, then z, skipping 0.
struct Disp {
  int x, y, z;
};
std::vector<Disp> neighbourhood() {
  std::vector<Disp> v;
  for (int x = -1; x <= 1; ++x)
    for (int y = -1; y <= 1; ++y)
      for (int z = -1; z <= 1; ++z)
        if (x || y || z) v.push_back({x, y, z});
  return v;
}

struct Xdr {
  const unsigned char* p;
  size_t n, at = 0;
  bool ok = true;
  uint32_t u32() {
    if (at + 4 > n) {
      ok = false;
      return 0;
    }
    const uint32_t v = ((uint32_t)p[at] << 24) | ((uint32_t)p[at + 1] << 16) | ((uint32_t)p[at + 2] << 8) | p[at + 3];
    at += 4;
    return v;
  }
  float f32() {
    const uint32_t b = u32();
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
  }
};

int inflate_block(const unsigned char* src, size_t n, std::vector<unsigned char>& out, size_t expect) {
  out.assign(expect, 0);
  z_stream zs;
  std::memset(&zs, 0, sizeof zs);
  if (inflateInit(&zs) != Z_OK) return -1;
  zs.next_in = const_cast<unsigned char*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = out.data();
  zs.avail_out = (uInt)expect;
  const int ret = inflate(&zs, Z_FINISH);
  const size_t got = expect - zs.avail_out;
  inflateEnd(&zs);
  if (ret != Z_STREAM_END || got != expect) return -1;
  return 0;
}

}  // namespace
}  // namespace hlb

using namespace hlb;

extern "C" int hlb_geom_read_gmy(const char* path, int gmy_plus, hlb_geom** out) {
  if (!path || !out) return fail(HLB_EINVAL, "hlb_geom_read_gmy: null argument");
  std::unique_ptr<FILE, int (*)(FILE*)> fh(std::fopen(path, "rb"), &std::fclose);
  if (!fh) return fail(HLB_EINVAL, std::string("hlb_geom_read_gmy: cannot open ") + path);
  std::vector<unsigned char> buf;
  try {
    std::fseek(fh.get(), 0, SEEK_END);
    const long size = std::ftell(fh.get());
    std::fseek(fh.get(), 0, SEEK_SET);
    if (size < 32) return fail(HLB_EINVAL, "hlb_geom_read_gmy: file shorter than the preamble");
    buf.resize((size_t)size);
    if (std::fread(buf.data(), 1, buf.size(), fh.get()) != buf.size())
      return fail(HLB_EINVAL, "hlb_geom_read_gmy: short read");
  } catch (const std::bad_alloc&) {
    return fail(HLB_EINVAL, "hlb_geom_read_gmy: out of host memory");
  }
  Xdr pre{buf.data(), 32};
  const uint32_t m1 = pre.u32(), m2 = pre.u32(), ver = pre.u32();
  if (m1 != kHemeLbMagic) return fail(HLB_EINVAL, "hlb_geom_read_gmy: not a HemeLB file (magic number)");
  if (m2 != kGmyMagic) return fail(HLB_EINVAL, "hlb_geom_read_gmy: not a geometry file (magic number)");
  if (ver != kGmyVersion)
    return fail(HLB_EINVAL, "hlb_geom_read_gmy: version " + std::to_string(ver) + ", only 4 is supported");
  const uint32_t bx = pre.u32(), by = pre.u32(), bz = pre.u32(), bs = pre.u32();
  if (bx == 0 || by == 0 || bz == 0 || bs == 0 || bs > 64)
    return fail(HLB_EINVAL, "hlb_geom_read_gmy: bad block counts in the preamble");
  const uint64_t nblocks = (uint64_t)bx * by * bz;
  const size_t rec = gmy_plus ? 16 : 12;
  if (32 + nblocks * rec > buf.size()) return fail(HLB_EINVAL, "hlb_geom_read_gmy: header runs past the end of the file");
  auto g = std::make_unique<hlb_geom>();
  g->blocks[0] = (int32_t)bx;
  g->blocks[1] = (int32_t)by;
  g->blocks[2] = (int32_t)bz;
  g->bs = (int32_t)bs;
  const std::vector<Disp> nbh = neighbourhood();
  int dirOf[26];
  for (int i = 0; i < 26; ++i) {
    dirOf[i] = -1;
    for (int d = 1; d < Q; ++d)
      if (CX[d] == nbh[i].x && CY[d] == nbh[i].y && CZ[d] == nbh[i].z) dirOf[i] = d;
  }
  Xdr hdr{buf.data() + 32, (size_t)(nblocks * rec)};
  size_t at = 32 + (size_t)(nblocks * rec);
  std::vector<unsigned char> block;
  const uint64_t spb = (uint64_t)bs * bs * bs;
  try {
    for (uint64_t b = 0; b < nblocks; ++b) {
      const uint32_t sites = hdr.u32();
      if (gmy_plus) (void)hdr.u32();  // weight
      const uint32_t bytes = hdr.u32(), raw = hdr.u32();
      if (bytes == 0) {
        if (sites != 0) return fail(HLB_EINVAL, "hlb_geom_read_gmy: block with sites but no data");
        continue;
      }
      if (at + bytes > buf.size()) return fail(HLB_EINVAL, "hlb_geom_read_gmy: block data runs past the end of the file");
      if (inflate_block(buf.data() + at, bytes, block, raw) != 0)
        return fail(HLB_EINVAL, "hlb_geom_read_gmy: decompression error for block " + std::to_string(b));
      at += bytes;
      const int64_t ox = (int64_t)(b / ((uint64_t)by * bz)) * bs, oy = (int64_t)((b / bz) % by) * bs,
                    oz = (int64_t)(b % bz) * bs;
      Xdr r{block.data(), block.size()};
      uint32_t fluid = 0;
      for (uint64_t l = 0; l < spb; ++l) {
        if (r.u32() == 0) continue;  // solid
        ++fluid;
        const int64_t x = ox + (int64_t)(l / (bs * bs)), y = oy + (int64_t)((l / bs) % bs), z = oz + (int64_t)(l % bs);
        g->coords.push_back((int32_t)x);
        g->coords.push_back((int32_t)y);
        g->coords.push_back((int32_t)z);
        uint8_t type[18];
        int32_t iolet[18];
        float dist[18];
        for (int d = 0; d < 18; ++d) {
          type[d] = HLB_LINK_NONE;
          iolet[d] = -1;
          dist[d] = -1.0f;
        }
        bool wallSite = false;
        for (int i = 0; i < 26; ++i) {
          const uint32_t t = r.u32();
          int32_t id = -1;
          float q = -1.0f;
          if (t == HLB_LINK_WALL) {
            wallSite = true;
            q = r.f32();
          } else if (t == HLB_LINK_INLET || t == HLB_LINK_OUTLET) {
            id = (int32_t)r.u32();
            q = r.f32();
          } else if (t != HLB_LINK_NONE) {
            return fail(HLB_EINVAL, "hlb_geom_read_gmy: unknown cut type " + std::to_string(t));
          }
          if (dirOf[i] > 0) {
            type[dirOf[i] - 1] = (uint8_t)t;
            iolet[dirOf[i] - 1] = id;
            dist[dirOf[i] - 1] = q;
          }
        }
        const uint32_t nav = r.u32();
        // GeometryReader.cc:916-926: a wall site must carry a normal and a bulk site must not.
        if ((nav == 1) != wallSite)
          return fail(HLB_EINVAL, std::string("hlb_geom_read_gmy: Malformed GMY file, ") +
                                      (wallSite ? "wall fluid site without" : "bulk fluid site with") +
                                      " a defined wall normal currently not allowed.");
        float n[3] = {0.0f, 0.0f, 0.0f};
        if (nav == 1)
          for (float& c : n) c = r.f32();
        g->link_type.insert(g->link_type.end(), type, type + 18);
        g->link_iolet.insert(g->link_iolet.end(), iolet, iolet + 18);
        g->link_dist.insert(g->link_dist.end(), dist, dist + 18);
        g->normal_avail.push_back(nav == 1 ? 1 : 0);
        g->normal.insert(g->normal.end(), n, n + 3);
      }
      if (!r.ok) return fail(HLB_EINVAL, "hlb_geom_read_gmy: block " + std::to_string(b) + " ends early");
      if (fluid != sites)
        return fail(HLB_EINVAL, "hlb_geom_read_gmy: block " + std::to_string(b) + " holds " + std::to_string(fluid) +
                                    " fluid sites, its header says " + std::to_string(sites));
    }
  } catch (const std::bad_alloc&) {
    return fail(HLB_EINVAL, "hlb_geom_read_gmy: out of host memory");
  }
  if (!hdr.ok) return fail(HLB_EINVAL, "hlb_geom_read_gmy: truncated header");
  *out = g.release();
  return HLB_OK;
}
