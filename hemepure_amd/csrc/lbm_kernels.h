// lbm_kernels.h -- device-side data layout shared by the kernels and the
// host launcher.
//
// HBM layout (one handle, N sites in the reference's contiguous order):
//   A[2]   post-collision distributions, 19 planes of S doubles (plane j holds
//          direction j for every site); S = N rounded up to 64.
//   F      pre-collision staging, same shape as A (initial state, uploads and
//          downloads; allocated on demand).
//   pull   18 planes of S uint32: pull[(j-1)*S + k] is the site whose
//          post-collision value of direction j arrives at k, or k itself
//          where a boundary condition supplies the value.
//   boundary arrays over the sites k >= off1 (index b = k - off1, stride SB):
//          wall/iolet link masks, global iolet index, 18 planes of float link
//          distances, 3 planes of float wall normals.
//   iolet arrays over the sites k >= off2 (index i = k - off2, stride SI):
//          G, 19 planes of the values the site's iolet delegates computed for
//          the next step, and 18 planes of Ladd profile factors.
#pragma once

#include <cstdint>

namespace hlb {

struct StepArgs {
  // Per-plane base pointers, passed as separate scalars so each access is one
  // SGPR base plus a 32-bit VGPR offset.
  const double* ain[19];
  double* aout[19];
  const uint32_t* pullp[18];
  const double* __restrict__ Ain;   // post-collision of the previous step (or F for the first)
  double* __restrict__ Aout;
  const uint32_t* __restrict__ pull;
  const uint32_t* __restrict__ bwall;
  const uint32_t* __restrict__ biolet;
  const int32_t* __restrict__ bgi;  // global iolet index: inlets 0..n_in-1, outlets follow
  const float* __restrict__ bq;
  const float* __restrict__ bnrm;
  double* __restrict__ G;
  const double* __restrict__ ladd_s;
  double* __restrict__ rho;
  double* __restrict__ u;
  double* __restrict__ wss;
  const float* __restrict__ io_nf;    // 3 per iolet: the normal narrowed to float (Nash)
  const double* __restrict__ io_n;    // 3 per iolet: the normal in double (Ladd)
  const double* __restrict__ io_dens; // this step's ghost density per iolet
  const double* __restrict__ io_vmax; // this step's peak speed per iolet
  int64_t N, S, off1, off2, SB, SI;
  int32_t n_in;
  int32_t wall_bc;
  int32_t inlet_bc;
  int32_t outlet_bc;
  double tau;         // LbmParameters tau
  double omega;       // -1/tau
  double half_omega;  // 0.5*omega (TRT)
  double half_omega_minus;
  double les_k;       // 18 sqrt(2) C_s^2 evaluated in the reference's order
};

}  // namespace hlb
