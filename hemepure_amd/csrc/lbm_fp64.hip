// lbm_fp64.hip -- the fused double-precision collide+stream pass for gfx950.
//
// One thread per site, 256 threads (four wave64s) per workgroup, sites in the
// reference's class-major order.  A step of the reference is
//   collide at every site, push fPost[d] to the neighbour (or a boundary
//   delegate writes the site's own slot), BFL post-step, swap
// (src/lb/lb.hpp:192-347).  Here the same step is one pull pass: site k first
// materialises the pre-collision f it would hold in the reference's fOld by
// gathering the previous pass's post-collision values (plain links: the
// neighbour's value; SimpleBounceBack / BFL: the site's own values,
// src/lb/streamers/SimpleBounceBackDelegate.h:36-44,
// BouzidiFirdaousLallemandDelegate.h:44-93; Nash / Ladd: the value its iolet
// delegate computed last pass), then collides and stores its own
// post-collision row.  Every store is coalesced; the 18 gathers are reads of
// neighbouring rows, which the block-structured order keeps in the same or
// adjacent 8^3 blocks.
//
// Sites k < off1 (collision class 0, about 90% of a vessel) take the bulk
// path: 18 index loads, 19 gathers, no boundary logic.  Sites of classes 1-5
// take the boundary path; its branches are uniform per direction for most
// wavefronts.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../../include/hemepure_amd.h"
#include "d3q19.h"
#include "lbm_kernels.h"

namespace hlb {

namespace {

constexpr int kBlock = 256;

template <int J = 1>
__device__ __forceinline__ void gather_bulk(const StepArgs& a, int64_t k, double* f) {
  if constexpr (J < Q) {
    const uint32_t src = __builtin_nontemporal_load(a.pull + (int64_t)(J - 1) * a.S + k);
    f[J] = a.Ain[(int64_t)J * a.S + src];
    gather_bulk<J + 1>(a, k, f);
  }
}

// Pre-collision values at a boundary site (classes 1-5): per direction j the
// writer is the site's link INV[j] -- iolet delegate (value stored in G),
// wall delegate (own post-collision values) or the neighbour (pull table).
template <int J = 1>
__device__ __forceinline__ void gather_boundary(const StepArgs& a, int64_t k, int64_t b, uint32_t wm,
                                                uint32_t im, double* f) {
  if constexpr (J < Q) {
    constexpr int IB = INV[J];
    constexpr uint32_t bit = 1u << (IB - 1);
    constexpr uint32_t bitJ = 1u << (J - 1);
    if (im & bit) {
      f[J] = a.G[(int64_t)J * a.SI + (k - a.off2)];
    } else if (wm & bit) {
      const double fpd = a.Ain[(int64_t)IB * a.S + k];
      if (a.wall_bc == HLB_WALL_SBB) {
        f[J] = fpd;
      } else {
        const double q = (double)a.bq[(int64_t)(IB - 1) * a.SB + b];
        if ((wm & bitJ) || q < 0.5) {
          f[J] = fpd;
        } else {
          const double fpi = a.Ain[(int64_t)J * a.S + k];
          f[J] = (fpd + (2.0 * q - 1) * fpi) / (2.0 * q);
        }
      }
    } else {
      const uint32_t src = a.pull[(int64_t)(J - 1) * a.S + k];
      f[J] = a.Ain[(int64_t)J * a.S + src];
    }
    gather_boundary<J + 1>(a, k, b, wm, im, f);
  }
}

// BFL post-step for the wall links d with q < 1/2 whose opposite link is not
// a wall (BouzidiFirdaousLallemandDelegate.h:72-93), in direction order.
template <int D = 1>
__device__ __forceinline__ void bfl_post(const StepArgs& a, int64_t b, uint32_t wm, double* f) {
  if constexpr (D < Q) {
    constexpr int J = INV[D];
    constexpr uint32_t bit = 1u << (D - 1);
    constexpr uint32_t bitJ = 1u << (J - 1);
    if (wm & bit) {
      const double q = (double)a.bq[(int64_t)(D - 1) * a.SB + b];
      if (!(wm & bitJ) && q < 0.5) f[J] = 2.0 * q * f[J] + (1.0 - 2.0 * q) * f[D];
    }
    bfl_post<D + 1>(a, b, wm, f);
  }
}

// Pre-collision gather for site k (shared by the step and the download pass).
__device__ __forceinline__ void gather(const StepArgs& a, int64_t k, double* f) {
  f[0] = a.Ain[k];
  if (k < a.off1) {
    gather_bulk(a, k, f);
  } else {
    const int64_t b = k - a.off1;
    const uint32_t wm = a.bwall[b], im = a.biolet[b];
    gather_boundary(a, k, b, wm, im, f);
    if (a.wall_bc == HLB_WALL_BFL && wm != 0) bfl_post(a, b, wm, f);
  }
}

// Iolet delegates of one site: the value each iolet link d hands to the next
// step in direction INV[d].  Nash: f_eq of the ghost state
// (NashZerothOrderPressureDelegate.h:36-76).  Ladd: bounce-back minus the
// moving-wall correction (LaddIoletDelegate.h:37-86).
template <int D = 1>
__device__ __forceinline__ void iolet_links(const StepArgs& a, int64_t i, uint32_t im, int gi, int bc,
                                            double rho, double vx, double vy, double vz,
                                            const double* fpost) {
  if constexpr (D < Q) {
    constexpr int U = INV[D];
    if (im & (1u << (D - 1))) {
      double val;
      if (bc == HLB_IOLET_NASH) {
        const double rg = a.io_dens[gi];
        const double nx = (double)a.io_nf[3 * gi], ny = (double)a.io_nf[3 * gi + 1],
                     nz = (double)a.io_nf[3 * gi + 2];
        const double component = vx * nx + vy * ny + vz * nz;
        const double gx = nx * component * rg, gy = ny * component * rg, gz = nz * component * rg;
        const FeqPre p = feq_pre(rg, gx, gy, gz);
        val = feq_dir<U>(p, gx, gy, gz);
      } else {
        const double m = a.io_vmax[gi] * a.ladd_s[(int64_t)(D - 1) * a.SI + i];
        const double wx = a.io_n[3 * gi] * m * rho, wy = a.io_n[3 * gi + 1] * m * rho,
                     wz = a.io_n[3 * gi + 2] * m * rho;
        const double correction =
            2. * Wt(D) * (wx * (double)CX[D] + wy * (double)CY[D] + wz * (double)CZ[D]) / CS2;
        val = fpost[D] - correction;
      }
      a.G[(int64_t)U * a.SI + i] = val;
    }
    iolet_links<D + 1>(a, i, im, gi, bc, rho, vx, vy, vz, fpost);
  }
}

template <int KERNEL, bool FIRST, bool PROPS>
__global__ __launch_bounds__(kBlock) void step_fp64(StepArgs a) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.N) return;
  double f[Q];
  if constexpr (FIRST) {
#pragma unroll
    for (int j = 0; j < Q; ++j) f[j] = a.Ain[(int64_t)j * a.S + k];
  } else {
    gather(a, k, f);
  }

  // Moments and equilibrium (Lattice.h:933-948), collision.
  double rho, mx, my, mz;
  moments(f, rho, mx, my, mz);
  const double ux = mx / rho, uy = my / rho, uz = mz / rho;
  const FeqPre pre = feq_pre(rho, mx, my, mz);
  double feq[Q];
  feq_all(pre, mx, my, mz, feq);
  double fneq[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) fneq[j] = f[j] - feq[j];
  double tau = a.tau;
  double fpost[Q];
  if constexpr (KERNEL == HLB_KERNEL_LBGK) {
    // LBGK.h:64-72
#pragma unroll
    for (int j = 0; j < Q; ++j) fpost[j] = f[j] + fneq[j] * a.omega;
  } else if constexpr (KERNEL == HLB_KERNEL_TRT) {
    // TRT.h:80-109; the (0,0) pair reduces to f0 + omega fneq0.
    fpost[0] = f[0] + a.omega * fneq[0];
#pragma unroll
    for (int i = 1; i < Q; i += 2) {
      const int ib = i + 1;
      const double sym = a.half_omega * (fneq[i] + fneq[ib]);
      const double asym = a.half_omega_minus * (fneq[i] - fneq[ib]);
      fpost[i] = f[i] + sym + asym;
      fpost[ib] = f[ib] + sym - asym;
    }
  } else {
    // LBGKLESSpongeLayer.h:150-182 with lifetime 0.
    const Sym3 s = second_moments(fneq);
    double q12 = s.xx * s.xx;
    q12 += s.yx * s.yx;
    q12 += s.zx * s.zx;
    q12 += s.yx * s.yx;
    q12 += s.yy * s.yy;
    q12 += s.zy * s.zy;
    q12 += s.zx * s.zx;
    q12 += s.zy * s.zy;
    q12 += s.zz * s.zz;
    q12 = __builtin_sqrt(q12);
    tau = 1. / 2. * (a.tau + __builtin_sqrt(a.tau * a.tau + a.les_k * q12));
    const double om = -1.0 / tau;
#pragma unroll
    for (int j = 0; j < Q; ++j) fpost[j] = f[j] + fneq[j] * om;
  }

#pragma unroll
  for (int j = 0; j < Q; ++j) a.Aout[(int64_t)j * a.S + k] = fpost[j];

  if (k >= a.off2) {
    const int64_t b = k - a.off1, i = k - a.off2;
    const uint32_t im = a.biolet[b];
    const int gi = a.bgi[b];
    const int bc = gi < a.n_in ? a.inlet_bc : a.outlet_bc;
    iolet_links(a, i, im, gi, bc, rho, ux, uy, uz, fpost);
  }

  if constexpr (PROPS) {
    // BaseStreamer::UpdateMinsAndMaxes (BaseStreamer.h:88-120) and the wall
    // shear stress of Lattice.h:1122-1156.
    a.rho[k] = rho;
    a.u[k] = ux;
    a.u[a.S + k] = uy;
    a.u[2 * a.S + k] = uz;
    double w = DBL_MAX;
    if (k >= a.off1) {
      const int64_t b = k - a.off1;
      if (a.bwall[b] != 0) {
        double g[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) g[j] = fpost[j] - f[j];
        const Sym3 s = second_moments(g);
        const double fac = -(tau - 0.5);
        const double p00 = s.xx * fac, p10 = s.yx * fac, p11 = s.yy * fac;
        const double p20 = s.zx * fac, p21 = s.zy * fac, p22 = s.zz * fac;
        const double n0 = (double)a.bnrm[b], n1 = (double)a.bnrm[a.SB + b], n2 = (double)a.bnrm[2 * a.SB + b];
        double sv0 = p00 * n0;
        sv0 += p10 * n1;
        sv0 += p20 * n2;
        double sv1 = p10 * n0;
        sv1 += p11 * n1;
        sv1 += p21 * n2;
        double sv2 = p20 * n0;
        sv2 += p21 * n1;
        sv2 += p22 * n2;
        double sq = sv0 * sv0;
        sq += sv1 * sv1;
        sq += sv2 * sv2;
        double nn = sv0 * n0;
        nn += sv1 * n1;
        nn += sv2 * n2;
        w = __builtin_sqrt(sq - nn * nn);
      }
    }
    a.wss[k] = w;
  }
}

// Materialises the pre-collision f (the reference's fOld entering the next
// step) into F, SoA.
__global__ __launch_bounds__(kBlock) void gather_fp64(StepArgs a, double* __restrict__ F) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.N) return;
  double f[Q];
  gather(a, k, f);
#pragma unroll
  for (int j = 0; j < Q; ++j) F[(int64_t)j * a.S + k] = f[j];
}

__global__ __launch_bounds__(kBlock) void fill_planes_fp64(double* __restrict__ F, int64_t N, int64_t S,
                                                           const double* __restrict__ vals) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= N) return;
#pragma unroll
  for (int j = 0; j < Q; ++j) F[(int64_t)j * S + k] = vals[j];
}

// StabilityTester: flag any !(f > 0) among the sites' distributions.
__global__ __launch_bounds__(kBlock) void unstable_fp64(const double* __restrict__ F, int64_t N, int64_t S,
                                                        int* __restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= N) return;
  bool bad = false;
#pragma unroll
  for (int j = 0; j < Q; ++j) bad |= !(F[(int64_t)j * S + k] > 0.0);
  if (bad) *flag = 1;
}

}  // namespace

unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

hipError_t launch_step_fp64(const StepArgs& a, int kernel, bool first, bool props, hipStream_t s) {
  const dim3 grid(grid_for(a.N)), block(kBlock);
#define HLB_LAUNCH(KN, FI, PR) hipLaunchKernelGGL((step_fp64<KN, FI, PR>), grid, block, 0, s, a)
#define HLB_KSEL(KN)                                   \
  if (first) {                                         \
    if (props) HLB_LAUNCH(KN, true, true);             \
    else HLB_LAUNCH(KN, true, false);                  \
  } else {                                             \
    if (props) HLB_LAUNCH(KN, false, true);            \
    else HLB_LAUNCH(KN, false, false);                 \
  }
  switch (kernel) {
    case HLB_KERNEL_LBGK: HLB_KSEL(HLB_KERNEL_LBGK); break;
    case HLB_KERNEL_TRT: HLB_KSEL(HLB_KERNEL_TRT); break;
    default: HLB_KSEL(HLB_KERNEL_LES); break;
  }
#undef HLB_KSEL
#undef HLB_LAUNCH
  return hipGetLastError();
}

hipError_t launch_gather_fp64(const StepArgs& a, double* F, hipStream_t s) {
  hipLaunchKernelGGL(gather_fp64, dim3(grid_for(a.N)), dim3(kBlock), 0, s, a, F);
  return hipGetLastError();
}

hipError_t launch_fill_fp64(double* F, int64_t N, int64_t S, const double* vals, hipStream_t s) {
  hipLaunchKernelGGL(fill_planes_fp64, dim3(grid_for(N)), dim3(kBlock), 0, s, F, N, S, vals);
  return hipGetLastError();
}

hipError_t launch_unstable_fp64(const double* F, int64_t N, int64_t S, int* flag, hipStream_t s) {
  hipLaunchKernelGGL(unstable_fp64, dim3(grid_for(N)), dim3(kBlock), 0, s, F, N, S, flag);
  return hipGetLastError();
}

}  // namespace hlb
