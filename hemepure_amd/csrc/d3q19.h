// d3q19.h -- D3Q19 lattice constants and the double-precision site arithmetic
// of the collide step, written so that every rounding happens where the
// reference rounds it.
//
// The reference's default build (HEMELB_USE_SSE3=ON, src/CMakeLists.txt:17)
// evaluates the moments with two interleaved accumulation lanes
// (src/lb/lattices/Lattice.h:51-110) and the equilibrium with the SSE3
// grouping (Lattice.h:374-448).  Terms whose lattice coefficient is 0 only add
// a signed zero there, and coefficients of +-1 are exact, so the sums below
// keep the nonzero terms in the reference's lane order; the results are the
// same doubles.  Every translation unit including this header is compiled with
// -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HLB_HD __host__ __device__ __forceinline__
#else
#define HLB_HD inline
#endif

namespace hlb {

constexpr int Q = 19;

// src/lb/lattices/D3Q19.cc:15-33.
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int INV[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

constexpr double W0 = 1.0 / 3.0;
constexpr double W1 = 1.0 / 18.0;
constexpr double W2 = 1.0 / 36.0;
HLB_HD constexpr double Wt(int i) { return i == 0 ? W0 : (i < 7 ? W1 : W2); }

constexpr double CS2 = 1.0 / 3.0;  // src/constants.h:54

// Sum of the nonzero terms c_x*x, c_y*y, c_z*z in x, y, z order: the value of
// (CXD[i]*x + CYD[i]*y) + CZD[i]*z for a lattice vector (Lattice.h:394-401).
template <int I>
HLB_HD double cdot(double x, double y, double z) {
  constexpr int a = CX[I], b = CY[I], c = CZ[I];
  if constexpr (a == 0 && b == 0 && c == 0) {
    return 0.0;
  } else {
    double s = 0.0;
    bool first = true;
    if constexpr (a != 0) {
      s = (a > 0) ? x : -x;
      first = false;
    }
    if constexpr (b != 0) {
      const double t = (b > 0) ? y : -y;
      s = first ? t : s + t;
      first = false;
    }
    if constexpr (c != 0) {
      const double t = (c > 0) ? z : -z;
      s = first ? t : s + t;
    }
    return s;
  }
}

// Density and momentum in the SSE3 lane order (Lattice.h:51-110): lane 0
// accumulates the even directions 0..16, lane 1 starts from direction 18 and
// accumulates the odd directions 1..17; the horizontal add gives lane0 + lane1.
HLB_HD void moments(const double* f, double& rho, double& mx, double& my, double& mz) {
  double r0 = f[0];
  r0 += f[2]; r0 += f[4]; r0 += f[6]; r0 += f[8]; r0 += f[10]; r0 += f[12]; r0 += f[14]; r0 += f[16];
  double r1 = f[18];
  r1 += f[1]; r1 += f[3]; r1 += f[5]; r1 += f[7]; r1 += f[9]; r1 += f[11]; r1 += f[13]; r1 += f[15];
  r1 += f[17];
  rho = r0 + r1;
  // x: even lane 2(-) 8(-) 10(-) 12(-) 14(-); odd lane 1 7 9 11 13 (all +)
  double x0 = -f[2];
  x0 -= f[8]; x0 -= f[10]; x0 -= f[12]; x0 -= f[14];
  double x1 = f[1];
  x1 += f[7]; x1 += f[9]; x1 += f[11]; x1 += f[13];
  mx = x0 + x1;
  // y: even lane 4(-) 8(-) 10(+) 16(-); odd lane 18(-) 3(+) 7(+) 9(-) 15(+) 17(+)
  double y0 = -f[4];
  y0 -= f[8]; y0 += f[10]; y0 -= f[16];
  double y1 = -f[18];
  y1 += f[3]; y1 += f[7]; y1 -= f[9]; y1 += f[15]; y1 += f[17];
  my = y0 + y1;
  // z: even lane 6(-) 12(-) 14(+) 16(-); odd lane 18(+) 5(+) 11(+) 13(-) 15(+) 17(-)
  double z0 = -f[6];
  z0 -= f[12]; z0 += f[14]; z0 -= f[16];
  double z1 = f[18];
  z1 += f[5]; z1 += f[11]; z1 -= f[13]; z1 += f[15]; z1 -= f[17];
  mz = z0 + z1;
}

// Invariants of the SSE3 equilibrium (Lattice.h:383-392).
struct FeqPre {
  double tmp1;  // density - 1.5|j|^2 / density
  double nH;    // 4.5 / density
};

HLB_HD FeqPre feq_pre(double rho, double mx, double my, double mz) {
  const double tHMS = (3. / 2.) * (mx * mx + my * my + mz * mz);
  const double d1 = 1. / rho;
  FeqPre p;
  p.nH = (9. / 2.) * d1;
  p.tmp1 = rho - tHMS * d1;
  return p;
}

// f_eq[i] = w_i ((tmp1 + nH (c.j)^2) + 3 (c.j))   (Lattice.h:394-417)
template <int I>
HLB_HD double feq_dir(const FeqPre& p, double mx, double my, double mz) {
  const double mde = cdot<I>(mx, my, mz);
  return Wt(I) * ((p.tmp1 + p.nH * (mde * mde)) + 3. * mde);
}

// All 19 equilibria; unrolled so each direction's vector is a constant.
template <int I = 0>
HLB_HD void feq_all(const FeqPre& p, double mx, double my, double mz, double* feq) {
  if constexpr (I < Q) {
    feq[I] = feq_dir<I>(p, mx, my, mz);
    feq_all<I + 1>(p, mx, my, mz, feq);
  }
}

// Runtime-direction equilibrium for one direction (boundary ghosts).
HLB_HD double feq_one(int i, const FeqPre& p, double mx, double my, double mz) {
  switch (i) {
    case 0: return feq_dir<0>(p, mx, my, mz);
    case 1: return feq_dir<1>(p, mx, my, mz);
    case 2: return feq_dir<2>(p, mx, my, mz);
    case 3: return feq_dir<3>(p, mx, my, mz);
    case 4: return feq_dir<4>(p, mx, my, mz);
    case 5: return feq_dir<5>(p, mx, my, mz);
    case 6: return feq_dir<6>(p, mx, my, mz);
    case 7: return feq_dir<7>(p, mx, my, mz);
    case 8: return feq_dir<8>(p, mx, my, mz);
    case 9: return feq_dir<9>(p, mx, my, mz);
    case 10: return feq_dir<10>(p, mx, my, mz);
    case 11: return feq_dir<11>(p, mx, my, mz);
    case 12: return feq_dir<12>(p, mx, my, mz);
    case 13: return feq_dir<13>(p, mx, my, mz);
    case 14: return feq_dir<14>(p, mx, my, mz);
    case 15: return feq_dir<15>(p, mx, my, mz);
    case 16: return feq_dir<16>(p, mx, my, mz);
    case 17: return feq_dir<17>(p, mx, my, mz);
    default: return feq_dir<18>(p, mx, my, mz);
  }
}

// Second moments sum_v c_a c_b g_v of a distribution-like vector g in v order
// (the int products c_a*c_b are -1, 0 or +1): the qij of
// LBGKLESSpongeLayer.h:163-171 and the Pi sums of Lattice.h:1362-1369.
struct Sym3 {
  double xx, yx, yy, zx, zy, zz;
};

HLB_HD Sym3 second_moments(const double* g) {
  Sym3 s;
  // xx: directions with c_x != 0, all products +1.
  s.xx = g[1]; s.xx += g[2]; s.xx += g[7]; s.xx += g[8]; s.xx += g[9]; s.xx += g[10];
  s.xx += g[11]; s.xx += g[12]; s.xx += g[13]; s.xx += g[14];
  // yx: 7(+) 8(+) 9(-) 10(-)
  s.yx = g[7]; s.yx += g[8]; s.yx -= g[9]; s.yx -= g[10];
  // yy: 3 4 7 8 9 10 15 16 17 18
  s.yy = g[3]; s.yy += g[4]; s.yy += g[7]; s.yy += g[8]; s.yy += g[9]; s.yy += g[10];
  s.yy += g[15]; s.yy += g[16]; s.yy += g[17]; s.yy += g[18];
  // zx: 11(+) 12(+) 13(-) 14(-)
  s.zx = g[11]; s.zx += g[12]; s.zx -= g[13]; s.zx -= g[14];
  // zy: 15(+) 16(+) 17(-) 18(-)
  s.zy = g[15]; s.zy += g[16]; s.zy -= g[17]; s.zy -= g[18];
  // zz: 5 6 11 12 13 14 15 16 17 18
  s.zz = g[5]; s.zz += g[6]; s.zz += g[11]; s.zz += g[12]; s.zz += g[13]; s.zz += g[14];
  s.zz += g[15]; s.zz += g[16]; s.zz += g[17]; s.zz += g[18];
  return s;
}

}  // namespace hlb
