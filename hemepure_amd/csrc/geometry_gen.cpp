// geometry_gen.cpp -- procedural vessel geometries, emitted as the per-site
// records a .gmy file holds after GeometryReader::ParseSite
// (src/geometry/GeometryReader.cc:851-935): fluid flag, one link record per
// D3Q19 direction (type, iolet id, float distance along the link) and a float
// wall normal at sites with a wall link.
//
// A geometry is a union of clipped tubes.  Tube i has an axis p0 + s*a
// (|a| = 1), is clipped to s in [s_lo, s_hi] and has a radius profile r(s).
// Iolets are discs on the clip planes of tube ends.  A lattice point is fluid
// when it lies strictly inside some tube.  For a fluid site and direction d
// with a non-fluid neighbour, the first exit of the segment x -> x + c_d from
// the union is found by sampling plus bisection; it is an iolet link when the
// exit lies on an iolet disc (distance = exact plane crossing) and a wall
// link otherwise.  By construction every link with a fluid neighbour is a
// plain link and every link with a non-fluid neighbour is a wall or iolet
// link, which is the consistency the pull-scheme tables require.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hemepure_amd.h"
#include "d3q19.h"
#include "errors.h"
#include "geom.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Radius {
  double r0 = 1.0, r1 = 1.0;     // linear taper from s = 0 to s = len
  double len = 1.0;
  double stenA = 0.0;            // area reduction at s0
  double stenS0 = 0.0, stenW = 1.0;
  int nmodes = 0;                // jitter: r *= 1 + sum amp_k sin(k_k s + ph_k)
  double amp[4] = {0, 0, 0, 0}, wav[4] = {0, 0, 0, 0}, ph[4] = {0, 0, 0, 0};
  double operator()(double s) const {
    const double t = std::min(std::max(s / len, 0.0), 1.0);
    double r = r0 + (r1 - r0) * t;
    for (int k = 0; k < nmodes; ++k) r *= 1.0 + amp[k] * std::sin(wav[k] * s + ph[k]);
    if (stenA > 0.0) {
      const double g = std::exp(-0.5 * (s - stenS0) * (s - stenS0) / (stenW * stenW));
      r *= std::sqrt(1.0 - stenA * g);
    }
    return r;
  }
  double max_bound() const {
    double m = std::max(r0, r1);
    for (int k = 0; k < nmodes; ++k) m *= 1.0 + std::fabs(amp[k]);
    return m;
  }
};

struct Tube {
  double p0[3];
  double a[3];
  double s_lo, s_hi;
  Radius r;
};

struct Cap {
  double c[3];
  double n[3];  // into the fluid
  double r;     // disc radius (tube radius plus one lattice unit of slack)
  int is_inlet;
  int id;
};

struct Shape {
  std::vector<Tube> tubes;
  std::vector<Cap> caps;

  bool inside(const double p[3]) const {
    for (const Tube& t : tubes) {
      const double dx = p[0] - t.p0[0], dy = p[1] - t.p0[1], dz = p[2] - t.p0[2];
      const double s = dx * t.a[0] + dy * t.a[1] + dz * t.a[2];
      if (s < t.s_lo || s > t.s_hi) continue;
      const double rx = dx - s * t.a[0], ry = dy - s * t.a[1], rz = dz - s * t.a[2];
      const double rr = t.r(s);
      if (rx * rx + ry * ry + rz * rz < rr * rr) return true;
    }
    return false;
  }

  // Outward radial unit vector of the tube whose surface is nearest.
  void wall_normal(const double p[3], float out[3]) const {
    double best = 1e300;
    double nb[3] = {0, 0, 1};
    for (const Tube& t : tubes) {
      const double dx = p[0] - t.p0[0], dy = p[1] - t.p0[1], dz = p[2] - t.p0[2];
      double s = dx * t.a[0] + dy * t.a[1] + dz * t.a[2];
      s = std::min(std::max(s, t.s_lo), t.s_hi);
      const double rx = dx - s * t.a[0], ry = dy - s * t.a[1], rz = dz - s * t.a[2];
      const double rad = std::sqrt(rx * rx + ry * ry + rz * rz);
      const double gap = std::fabs(t.r(s) - rad);
      if (gap < best && rad > 1e-12) {
        best = gap;
        nb[0] = rx / rad;
        nb[1] = ry / rad;
        nb[2] = rz / rad;
      }
    }
    out[0] = (float)nb[0];
    out[1] = (float)nb[1];
    out[2] = (float)nb[2];
  }
};

// Integer grid with a bit per site.
struct Grid {
  int64_t nx, ny, nz;
  std::vector<uint8_t> fluid;
  bool at(int64_t x, int64_t y, int64_t z) const {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return false;
    return fluid[(x * ny + y) * nz + z] != 0;
  }
};

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

int32_t round_up8(double v) { return (int32_t)(std::ceil(v / 8.0) * 8.0); }

// Voxelise the shape into a box of size dims (multiples of 8) and emit records.
hlb_geom* voxelise(const Shape& sh, const int32_t dims[3]) {
  auto g = std::make_unique<hlb_geom>();
  g->bs = 8;
  for (int a = 0; a < 3; ++a) g->blocks[a] = dims[a] / 8;
  Grid grid{dims[0], dims[1], dims[2], {}};
  grid.fluid.assign((size_t)grid.nx * grid.ny * grid.nz, 0);
  // Bounding box of each tube limits the scan.
  for (const Tube& t : sh.tubes) {
    double lo[3], hi[3];
    const double rb = t.r.max_bound() + 1.0;
    for (int a = 0; a < 3; ++a) {
      const double e0 = t.p0[a] + t.s_lo * t.a[a], e1 = t.p0[a] + t.s_hi * t.a[a];
      lo[a] = std::min(e0, e1) - rb;
      hi[a] = std::max(e0, e1) + rb;
    }
    const int64_t x0 = std::max<int64_t>(0, (int64_t)std::floor(lo[0]));
    const int64_t x1 = std::min<int64_t>(grid.nx - 1, (int64_t)std::ceil(hi[0]));
    const int64_t y0 = std::max<int64_t>(0, (int64_t)std::floor(lo[1]));
    const int64_t y1 = std::min<int64_t>(grid.ny - 1, (int64_t)std::ceil(hi[1]));
    const int64_t z0 = std::max<int64_t>(0, (int64_t)std::floor(lo[2]));
    const int64_t z1 = std::min<int64_t>(grid.nz - 1, (int64_t)std::ceil(hi[2]));
    for (int64_t x = x0; x <= x1; ++x)
      for (int64_t y = y0; y <= y1; ++y)
        for (int64_t z = z0; z <= z1; ++z) {
          uint8_t& cell = grid.fluid[(x * grid.ny + y) * grid.nz + z];
          if (cell) continue;
          const double p[3] = {(double)x, (double)y, (double)z};
          if (sh.inside(p)) cell = 1;
        }
  }
  // Records.
  for (int64_t x = 0; x < grid.nx; ++x)
    for (int64_t y = 0; y < grid.ny; ++y)
      for (int64_t z = 0; z < grid.nz; ++z) {
        if (!grid.at(x, y, z)) continue;
        g->coords.push_back((int32_t)x);
        g->coords.push_back((int32_t)y);
        g->coords.push_back((int32_t)z);
        bool anyWall = false;
        const double px[3] = {(double)x, (double)y, (double)z};
        for (int d = 1; d < hlb::Q; ++d) {
          const int cx = hlb::CX[d], cy = hlb::CY[d], cz = hlb::CZ[d];
          uint8_t type = HLB_LINK_NONE;
          int32_t iolet = -1;
          float dist = -1.0f;
          if (!grid.at(x + cx, y + cy, z + cz)) {
            // First exit of the segment from the shape.
            double tin = 0.0, tout = 1.0;
            for (int k = 1; k <= 16; ++k) {
              const double t = k / 16.0;
              const double p[3] = {x + t * cx, y + t * cy, z + t * cz};
              if (!sh.inside(p)) {
                tout = t;
                break;
              }
              tin = t;
            }
            for (int it = 0; it < 60; ++it) {
              const double tm = 0.5 * (tin + tout);
              const double p[3] = {x + tm * cx, y + tm * cy, z + tm * cz};
              if (sh.inside(p)) tin = tm; else tout = tm;
            }
            const double texit = 0.5 * (tin + tout);
            type = HLB_LINK_WALL;
            dist = (float)texit;
            for (const Cap& c : sh.caps) {
              const double a0 = (px[0] - c.c[0]) * c.n[0] + (px[1] - c.c[1]) * c.n[1] +
                                (px[2] - c.c[2]) * c.n[2];
              const double dn = cx * c.n[0] + cy * c.n[1] + cz * c.n[2];
              if (!(a0 > 0.0) || !(dn < 0.0)) continue;
              const double tc = a0 / -dn;
              if (tc > 1.0 || std::fabs(tc - texit) > 1e-9) continue;
              const double q[3] = {px[0] + tc * cx - c.c[0], px[1] + tc * cy - c.c[1],
                                   px[2] + tc * cz - c.c[2]};
              if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] > c.r * c.r) continue;
              type = c.is_inlet ? HLB_LINK_INLET : HLB_LINK_OUTLET;
              iolet = c.id;
              dist = (float)tc;
              break;
            }
            if (type == HLB_LINK_WALL) anyWall = true;
          }
          g->link_type.push_back(type);
          g->link_iolet.push_back(iolet);
          g->link_dist.push_back(dist);
        }
        g->normal_avail.push_back(anyWall ? 1 : 0);
        float nv[3] = {0.0f, 0.0f, 0.0f};
        if (anyWall) sh.wall_normal(px, nv);
        g->normal.insert(g->normal.end(), nv, nv + 3);
      }
  for (const Cap& c : sh.caps) {
    hlb_iolet_geom io;
    io.is_inlet = c.is_inlet;
    io.id = c.id;
    for (int a = 0; a < 3; ++a) {
      io.position[a] = c.c[a];
      io.normal[a] = c.n[a];
    }
    io.radius = c.r - 1.0;
    g->iolets.push_back(io);
  }
  return g.release();
}

// Parallel voxeliser for large sparse shapes (the vessel tree): each tube's
// bounding box is scanned against that tube only, and each link search tests
// only the tubes listed for the site's 8^3 tile.  Worker threads own
// contiguous x ranges and their records are concatenated in x order, so the
// result is the serial voxeliser's.
hlb_geom* voxelise_sparse(const Shape& sh, const int32_t dims[3], int nthreads) {
  auto g = std::make_unique<hlb_geom>();
  g->bs = 8;
  for (int a = 0; a < 3; ++a) g->blocks[a] = dims[a] / 8;
  const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
  Grid grid{nx, ny, nz, {}};
  grid.fluid.assign((size_t)nx * ny * nz, 0);
  const int64_t tx = nx / 8, ty = ny / 8, tz = nz / 8;
  std::vector<std::vector<int32_t>> tileTubes((size_t)(tx * ty * tz));
  auto bbox = [&](const Tube& t, double pad, int64_t lo[3], int64_t hi[3]) {
    const double rb = t.r.max_bound() + pad;
    for (int a = 0; a < 3; ++a) {
      const double e0 = t.p0[a] + t.s_lo * t.a[a], e1 = t.p0[a] + t.s_hi * t.a[a];
      lo[a] = std::max<int64_t>(0, (int64_t)std::floor(std::min(e0, e1) - rb));
      hi[a] = std::min<int64_t>(dims[a] - 1, (int64_t)std::ceil(std::max(e0, e1) + rb));
    }
  };
  for (size_t ti = 0; ti < sh.tubes.size(); ++ti) {
    int64_t lo[3], hi[3];
    bbox(sh.tubes[ti], 3.0, lo, hi);
    for (int64_t x = lo[0] / 8; x <= hi[0] / 8; ++x)
      for (int64_t y = lo[1] / 8; y <= hi[1] / 8; ++y)
        for (int64_t z = lo[2] / 8; z <= hi[2] / 8; ++z) tileTubes[(size_t)((x * ty + y) * tz + z)].push_back((int32_t)ti);
  }
  auto insideNear = [&](const double p[3], int64_t tile) {
    for (int32_t ti : tileTubes[(size_t)tile]) {
      const Tube& t = sh.tubes[ti];
      const double dx = p[0] - t.p0[0], dy = p[1] - t.p0[1], dz = p[2] - t.p0[2];
      const double s = dx * t.a[0] + dy * t.a[1] + dz * t.a[2];
      if (s < t.s_lo || s > t.s_hi) continue;
      const double rx = dx - s * t.a[0], ry = dy - s * t.a[1], rz = dz - s * t.a[2];
      const double rr = t.r(s);
      if (rx * rx + ry * ry + rz * rz < rr * rr) return true;
    }
    return false;
  };
  nthreads = std::max(1, nthreads);
  auto parallel_x = [&](auto&& body) {
    std::vector<std::thread> pool;
    for (int w = 0; w < nthreads; ++w)
      pool.emplace_back([&, w] {
        const int64_t x0 = nx * w / nthreads, x1 = nx * (w + 1) / nthreads;
        body(w, x0, x1);
      });
    for (auto& th : pool) th.join();
  };
  // Fluid bits: a point is fluid when some tube holds it strictly inside.
  parallel_x([&](int, int64_t x0, int64_t x1) {
    for (const Tube& t : sh.tubes) {
      int64_t lo[3], hi[3];
      bbox(t, 1.0, lo, hi);
      for (int64_t x = std::max(lo[0], x0); x <= std::min(hi[0], x1 - 1); ++x)
        for (int64_t y = lo[1]; y <= hi[1]; ++y)
          for (int64_t z = lo[2]; z <= hi[2]; ++z) {
            uint8_t& cell = grid.fluid[(size_t)((x * ny + y) * nz + z)];
            if (cell) continue;
            const double dx = x - t.p0[0], dy = y - t.p0[1], dz = z - t.p0[2];
            const double s = dx * t.a[0] + dy * t.a[1] + dz * t.a[2];
            if (s < t.s_lo || s > t.s_hi) continue;
            const double rx = dx - s * t.a[0], ry = dy - s * t.a[1], rz = dz - s * t.a[2];
            const double rr = t.r(s);
            if (rx * rx + ry * ry + rz * rz < rr * rr) cell = 1;
          }
    }
  });
  // Records per worker, in (x, y, z) order.
  struct Part {
    std::vector<int32_t> coords, link_iolet;
    std::vector<uint8_t> link_type, normal_avail;
    std::vector<float> link_dist, normal;
  };
  std::vector<Part> parts((size_t)nthreads);
  parallel_x([&](int w, int64_t x0, int64_t x1) {
    Part& P = parts[(size_t)w];
    for (int64_t x = x0; x < x1; ++x)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t z = 0; z < nz; ++z) {
          if (!grid.at(x, y, z)) continue;
          const int64_t tile = ((x / 8) * ty + y / 8) * tz + z / 8;
          P.coords.push_back((int32_t)x);
          P.coords.push_back((int32_t)y);
          P.coords.push_back((int32_t)z);
          bool anyWall = false;
          const double px[3] = {(double)x, (double)y, (double)z};
          for (int d = 1; d < hlb::Q; ++d) {
            const int cx = hlb::CX[d], cy = hlb::CY[d], cz = hlb::CZ[d];
            uint8_t type = HLB_LINK_NONE;
            int32_t iolet = -1;
            float dist = -1.0f;
            if (!grid.at(x + cx, y + cy, z + cz)) {
              double tin = 0.0, tout = 1.0;
              for (int k = 1; k <= 16; ++k) {
                const double t = k / 16.0;
                const double p[3] = {x + t * cx, y + t * cy, z + t * cz};
                if (!insideNear(p, tile)) {
                  tout = t;
                  break;
                }
                tin = t;
              }
              for (int it = 0; it < 40; ++it) {
                const double tm = 0.5 * (tin + tout);
                const double p[3] = {x + tm * cx, y + tm * cy, z + tm * cz};
                if (insideNear(p, tile)) tin = tm; else tout = tm;
              }
              const double texit = 0.5 * (tin + tout);
              type = HLB_LINK_WALL;
              dist = (float)texit;
              for (const Cap& c : sh.caps) {
                const double a0 = (px[0] - c.c[0]) * c.n[0] + (px[1] - c.c[1]) * c.n[1] + (px[2] - c.c[2]) * c.n[2];
                const double dn = cx * c.n[0] + cy * c.n[1] + cz * c.n[2];
                if (!(a0 > 0.0) || !(dn < 0.0)) continue;
                const double tc = a0 / -dn;
                if (tc > 1.0 || std::fabs(tc - texit) > 1e-9) continue;
                const double q[3] = {px[0] + tc * cx - c.c[0], px[1] + tc * cy - c.c[1], px[2] + tc * cz - c.c[2]};
                if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] > c.r * c.r) continue;
                type = c.is_inlet ? HLB_LINK_INLET : HLB_LINK_OUTLET;
                iolet = c.id;
                dist = (float)tc;
                break;
              }
              if (type == HLB_LINK_WALL) anyWall = true;
            }
            P.link_type.push_back(type);
            P.link_iolet.push_back(iolet);
            P.link_dist.push_back(dist);
          }
          P.normal_avail.push_back(anyWall ? 1 : 0);
          float nv[3] = {0.0f, 0.0f, 0.0f};
          if (anyWall) sh.wall_normal(px, nv);
          P.normal.insert(P.normal.end(), nv, nv + 3);
        }
  });
  grid.fluid.clear();
  grid.fluid.shrink_to_fit();
  size_t n = 0;
  for (const Part& P : parts) n += P.normal_avail.size();
  g->coords.reserve(3 * n);
  g->link_type.reserve(18 * n);
  g->link_iolet.reserve(18 * n);
  g->link_dist.reserve(18 * n);
  g->normal_avail.reserve(n);
  g->normal.reserve(3 * n);
  for (Part& P : parts) {
    g->coords.insert(g->coords.end(), P.coords.begin(), P.coords.end());
    g->link_type.insert(g->link_type.end(), P.link_type.begin(), P.link_type.end());
    g->link_iolet.insert(g->link_iolet.end(), P.link_iolet.begin(), P.link_iolet.end());
    g->link_dist.insert(g->link_dist.end(), P.link_dist.begin(), P.link_dist.end());
    g->normal_avail.insert(g->normal_avail.end(), P.normal_avail.begin(), P.normal_avail.end());
    g->normal.insert(g->normal.end(), P.normal.begin(), P.normal.end());
    P = Part();
  }
  for (const Cap& c : sh.caps) {
    hlb_iolet_geom io;
    io.is_inlet = c.is_inlet;
    io.id = c.id;
    for (int a = 0; a < 3; ++a) {
      io.position[a] = c.c[a];
      io.normal[a] = c.n[a];
    }
    io.radius = c.r - 1.0;
    g->iolets.push_back(io);
  }
  return g.release();
}

}  // namespace

extern "C" int hlb_geom_cylinder(double radius, int32_t length, int32_t box_xy, hlb_geom** out) {
  try {
    if (!out || radius <= 0.0 || length <= 0 || box_xy <= 0 || box_xy % 8 != 0 || 2.0 * radius + 2.0 > box_xy)
      return hlb::fail(HLB_EINVAL, "hlb_geom_cylinder: bad arguments");
    const double c = 0.5 * (box_xy - 1);
    Shape sh;
    Tube t;
    t.p0[0] = c; t.p0[1] = c; t.p0[2] = -0.5;
    t.a[0] = 0; t.a[1] = 0; t.a[2] = 1;
    t.s_lo = 0.0;
    t.s_hi = (double)length;
    t.r.r0 = t.r.r1 = radius;
    t.r.len = length;
    sh.tubes.push_back(t);
    sh.caps.push_back(Cap{{c, c, -0.5}, {0, 0, 1}, radius + 1.0, 1, 0});
    sh.caps.push_back(Cap{{c, c, length - 0.5}, {0, 0, -1}, radius + 1.0, 0, 0});
    const int32_t dims[3] = {box_xy, box_xy, round_up8(length)};
    *out = voxelise(sh, dims);
    return HLB_OK;
  } catch (const std::bad_alloc&) {
    return hlb::fail(HLB_EINVAL, "hlb_geom_cylinder: out of host memory");
  }
}

extern "C" int hlb_geom_ybifurcation(double rp, double lp, double rd, double ld, double angle_deg,
                                     double jitter, uint64_t seed, hlb_geom** out) {
  try {
    if (!out || rp <= 0 || lp <= 0 || rd <= 0 || ld <= 0 || angle_deg <= 0 || angle_deg >= 90 ||
        jitter < 0 || jitter > 0.2)
      return hlb::fail(HLB_EINVAL, "hlb_geom_ybifurcation: bad arguments");
    const double th = angle_deg * kPi / 180.0;
    const double sn = std::sin(th), cs = std::cos(th);
    // Seeded smooth radius jitter, shared by parent and both daughters so the
    // Y stays mirror-symmetric.
    uint64_t st = seed;
    Radius base;
    base.nmodes = jitter > 0.0 ? 3 : 0;
    for (int k = 0; k < base.nmodes; ++k) {
      base.amp[k] = jitter * (0.5 + 0.5 * uniform01(st)) / 3.0;
      base.wav[k] = 2.0 * kPi * (1.0 + 2.0 * uniform01(st)) / 60.0;
      base.ph[k] = 2.0 * kPi * uniform01(st);
    }
    const double rdmax = rd * (1.0 + jitter), rpmax = rp * (1.0 + jitter);
    const double halfx = ld * sn + rdmax * cs + 3.0;
    const double halfy = std::max(rpmax, rdmax) + 3.0;
    const int32_t dims[3] = {round_up8(2.0 * halfx + 1.0), round_up8(2.0 * halfy + 1.0),
                             round_up8(lp + ld * cs + rdmax * sn + 3.0)};
    const double cx = 0.5 * (dims[0] - 1), cy = 0.5 * (dims[1] - 1);
    Shape sh;
    Tube parent;
    parent.p0[0] = cx; parent.p0[1] = cy; parent.p0[2] = -0.5;
    parent.a[0] = 0; parent.a[1] = 0; parent.a[2] = 1;
    parent.s_lo = 0.0;
    parent.s_hi = lp + 0.5 * rp;
    parent.r = base;
    parent.r.r0 = parent.r.r1 = rp;
    parent.r.len = lp;
    sh.tubes.push_back(parent);
    const double B[3] = {cx, cy, -0.5 + lp};
    for (int side = 0; side < 2; ++side) {
      const double sx = side == 0 ? sn : -sn;
      Tube dtube;
      for (int a = 0; a < 3; ++a) dtube.p0[a] = B[a];
      dtube.a[0] = sx; dtube.a[1] = 0; dtube.a[2] = cs;
      dtube.s_lo = -0.5 * rd;
      dtube.s_hi = ld;
      dtube.r = base;
      dtube.r.r0 = dtube.r.r1 = rd;
      dtube.r.len = ld;
      sh.tubes.push_back(dtube);
      const double e[3] = {B[0] + ld * sx, B[1], B[2] + ld * cs};
      sh.caps.push_back(Cap{{e[0], e[1], e[2]}, {-sx, 0, -cs}, rdmax + 1.0, 0, side});
    }
    sh.caps.push_back(Cap{{cx, cy, -0.5}, {0, 0, 1}, rpmax + 1.0, 1, 0});
    *out = voxelise(sh, dims);
    // The parabolic inlet profile uses the nominal parent radius.
    for (auto& io : (*out)->iolets)
      io.radius = io.is_inlet ? rpmax : rdmax;
    return HLB_OK;
  } catch (const std::bad_alloc&) {
    return hlb::fail(HLB_EINVAL, "hlb_geom_ybifurcation: out of host memory");
  }
}

extern "C" int hlb_geom_stenosed_pipe(double radius, int32_t length, double area_reduction, double width,
                                      hlb_geom** out) {
  try {
    if (!out || radius <= 0 || length <= 0 || area_reduction < 0 || area_reduction >= 1 || width <= 0)
      return hlb::fail(HLB_EINVAL, "hlb_geom_stenosed_pipe: bad arguments");
    const int32_t bxy = round_up8(2.0 * radius + 5.0);
    const double c = 0.5 * (bxy - 1);
    Shape sh;
    Tube t;
    t.p0[0] = c; t.p0[1] = c; t.p0[2] = -0.5;
    t.a[0] = 0; t.a[1] = 0; t.a[2] = 1;
    t.s_lo = 0.0;
    t.s_hi = (double)length;
    t.r.r0 = t.r.r1 = radius;
    t.r.len = length;
    t.r.stenA = area_reduction;
    t.r.stenS0 = 0.5 * length;
    t.r.stenW = width;
    sh.tubes.push_back(t);
    sh.caps.push_back(Cap{{c, c, -0.5}, {0, 0, 1}, radius + 1.0, 1, 0});
    sh.caps.push_back(Cap{{c, c, length - 0.5}, {0, 0, -1}, radius + 1.0, 0, 0});
    const int32_t dims[3] = {bxy, bxy, round_up8(length)};
    *out = voxelise(sh, dims);
    return HLB_OK;
  } catch (const std::bad_alloc&) {
    return hlb::fail(HLB_EINVAL, "hlb_geom_stenosed_pipe: out of host memory");
  }
}

// A binary tree of tapering tubes: the root enters through the z = -0.5
// face at the box's x-y centre along +z; every segment end splits into two
// children of radius taper * r, turned by independent angles U(angle_min,
// angle_max) about the parent axis to opposite sides of a random plane.  A
// segment is length_factor * r long (shortened to stay inside the box by
// two radii); children overlap their parent by half a radius.  The root
// start is inlet 0, the leaf ends are outlets 0..2^(generations-1)-1.
extern "C" int hlb_geom_vessel_tree(uint64_t seed, int32_t generations, double root_radius, double taper,
                                    double length_factor, double angle_min_deg, double angle_max_deg,
                                    const int32_t box[3], int32_t nthreads, hlb_geom** out) {
  try {
    if (!out || !box || generations < 1 || generations > 12 || root_radius <= 0 || taper <= 0 || taper > 1 ||
        length_factor <= 0 || angle_min_deg < 0 || angle_max_deg < angle_min_deg || angle_max_deg >= 90 ||
        box[0] <= 0 || box[1] <= 0 || box[2] <= 0 || box[0] % 8 || box[1] % 8 || box[2] % 8 ||
        2.0 * root_radius + 4.0 > std::min(box[0], box[1]))
      return hlb::fail(HLB_EINVAL, "hlb_geom_vessel_tree: bad arguments");
    uint64_t st = seed;
    Shape sh;
    struct Seg {
      double start[3], dir[3], r;
      int gen;
    };
    std::vector<Seg> work{{{0.5 * (box[0] - 1), 0.5 * (box[1] - 1), -0.5}, {0, 0, 1}, root_radius, 0}};
    int leaf = 0;
    for (size_t w = 0; w < work.size(); ++w) {
      const Seg sg = work[w];
      // longest length that keeps the axis end two radii inside the box
      double len = length_factor * sg.r;
      for (int a = 0; a < 3; ++a) {
        const double lo = 2.0 * sg.r, hi = box[a] - 1 - 2.0 * sg.r;
        if (sg.dir[a] > 1e-12) len = std::min(len, (hi - sg.start[a]) / sg.dir[a]);
        if (sg.dir[a] < -1e-12) len = std::min(len, (lo - sg.start[a]) / sg.dir[a]);
      }
      len = std::max(len, 0.5 * sg.r);
      Tube t;
      for (int a = 0; a < 3; ++a) {
        t.p0[a] = sg.start[a];
        t.a[a] = sg.dir[a];
      }
      t.s_lo = sg.gen == 0 ? 0.0 : -0.5 * sg.r;
      t.s_hi = len;
      t.r.r0 = t.r.r1 = sg.r;
      t.r.len = len;
      sh.tubes.push_back(t);
      const double end[3] = {sg.start[0] + len * sg.dir[0], sg.start[1] + len * sg.dir[1],
                             sg.start[2] + len * sg.dir[2]};
      if (sg.gen + 1 == generations) {
        sh.caps.push_back(Cap{{end[0], end[1], end[2]}, {-sg.dir[0], -sg.dir[1], -sg.dir[2]}, sg.r + 1.0, 0, leaf++});
        continue;
      }
      // an orthonormal pair (u, v) perpendicular to the axis
      const double* d = sg.dir;
      double u[3] = {0, 0, 0};
      const int k = std::fabs(d[0]) < 0.6 ? 0 : (std::fabs(d[1]) < 0.6 ? 1 : 2);
      u[k] = 1.0;
      double dot = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
      for (int a = 0; a < 3; ++a) u[a] -= dot * d[a];
      double m = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      for (int a = 0; a < 3; ++a) u[a] /= m;
      const double v[3] = {d[1] * u[2] - d[2] * u[1], d[2] * u[0] - d[0] * u[2], d[0] * u[1] - d[1] * u[0]};
      const double phi = 2.0 * kPi * uniform01(st);
      for (int side = 0; side < 2; ++side) {
        const double th = (angle_min_deg + (angle_max_deg - angle_min_deg) * uniform01(st)) * kPi / 180.0;
        const double ph = phi + side * kPi;
        Seg c;
        for (int a = 0; a < 3; ++a) {
          c.start[a] = end[a];
          c.dir[a] = std::cos(th) * d[a] + std::sin(th) * (std::cos(ph) * u[a] + std::sin(ph) * v[a]);
        }
        m = std::sqrt(c.dir[0] * c.dir[0] + c.dir[1] * c.dir[1] + c.dir[2] * c.dir[2]);
        for (int a = 0; a < 3; ++a) c.dir[a] /= m;
        c.r = sg.r * taper;
        c.gen = sg.gen + 1;
        work.push_back(c);
      }
    }
    sh.caps.push_back(Cap{{0.5 * (box[0] - 1), 0.5 * (box[1] - 1), -0.5}, {0, 0, 1}, root_radius + 1.0, 1, 0});
    const int32_t dims[3] = {box[0], box[1], box[2]};
    *out = voxelise_sparse(sh, dims, nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency()));
    return HLB_OK;
  } catch (const std::bad_alloc&) {
    return hlb::fail(HLB_EINVAL, "hlb_geom_vessel_tree: out of host memory");
  }
}

extern "C" int hlb_geom_view(const hlb_geom* g, hlb_geometry* v) {
  if (!g || !v) return hlb::fail(HLB_EINVAL, "hlb_geom_view: null argument");
  for (int a = 0; a < 3; ++a) v->blocks[a] = g->blocks[a];
  v->block_size = g->bs;
  v->n_sites = (int64_t)g->normal_avail.size();
  v->coords = g->coords.data();
  v->link_type = g->link_type.data();
  v->link_iolet = g->link_iolet.data();
  v->link_dist = g->link_dist.data();
  v->normal_avail = g->normal_avail.data();
  v->normal = g->normal.data();
  return HLB_OK;
}

extern "C" int hlb_geom_iolets(const hlb_geom* g, hlb_iolet_geom* out, int max_iolets) {
  if (!g) return hlb::fail(HLB_EINVAL, "hlb_geom_iolets: null geometry");
  const int n = (int)g->iolets.size();
  for (int i = 0; i < n && i < max_iolets && out; ++i) out[i] = g->iolets[i];
  return n;
}

extern "C" void hlb_geom_free(hlb_geom* g) { delete g; }
