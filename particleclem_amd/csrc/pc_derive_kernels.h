// Plot variables of a level on the device: the state components and the
// reference's derived quantities (Source/Derive.cpp, Source/Derive.H), one
// thread per valid cell, written as one FAB per box (valid box, nplot comps,
// AMReX layout) into a staging buffer that the plotfile writer copies out.
#pragma once
#include "pc_level_kernels.h"

namespace pcd {

enum DeriveCode {
  DV_STATE = 0,    // state component `comp`
  DV_DIVRHO,       // comp / rho (x/y/z_velocity, massfrac): pc_derdividebyrho
  DV_MAGVEL,       // pc_dermagvel
  DV_KINENG,       // pc_derkineng
  DV_PRESSURE,     // pc_derpres
  DV_SOUNDSPEED,   // pc_dersoundspeed
  DV_MACH,         // pc_dermachnumber
  DV_MAGVORT,      // pc_dermagvort (regular cells, needs one ghost cell)
  DV_LOGDEN,       // pc_derlogden
};

constexpr int MAX_PLOT = 64;

struct DeriveArgs {
  int nplot;
  int code[MAX_PLOT];
  int comp[MAX_PLOT];
};

template <int NS>
PCD double vel_at(const V& U, int i, int j, int k, int c) {
  const double rhoInv = 1.0 / U(i, j, k, URHO);
  return U(i, j, k, UMX + c) * rhoInv;
}

// One cell of every requested plot variable; J.upd views the box's output
// FAB (valid box, nplot comps).
template <class EOS>
PCD void b_derive(const Sched& sc, const KArgs& ka, int lb_, const DeriveArgs& da) {
  constexpr int NS = EOS::NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_UPDATE, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const V& U = J.U;
  const double rho = U(i, j, k, URHO);
  const double rhoInv = 1.0 / rho;
  double Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) Y[s] = U(i, j, k, UFS + s) * rhoInv;
  const double T = U(i, j, k, UTEMP);
  const typename EOS::Thermo th(Y);
  for (int m = 0; m < da.nplot; ++m) {
    const int c = da.comp[m];
    double v = 0.0;
    switch (da.code[m]) {
      case DV_STATE: v = U(i, j, k, c); break;
      case DV_DIVRHO: v = U(i, j, k, c) / U(i, j, k, URHO); break;
      case DV_MAGVEL: {
        const double d1 = U(i, j, k, UMX) * rhoInv;
        const double d2 = U(i, j, k, UMY) * rhoInv;
        const double d3 = U(i, j, k, UMZ) * rhoInv;
        v = sqrt((d1 * d1) + (d2 * d2) + (d3 * d3));
        break;
      }
      case DV_KINENG: {
        const double x = U(i, j, k, UMX) * U(i, j, k, UMX);
        const double y = U(i, j, k, UMY) * U(i, j, k, UMY);
        const double z = U(i, j, k, UMZ) * U(i, j, k, UMZ);
        v = 0.5 / rho * (x + y + z);
        break;
      }
      case DV_PRESSURE: v = th.RTY2P(rho, T); break;
      case DV_SOUNDSPEED: v = th.RTY2Cs(rho, T); break;
      case DV_MACH: {
        const double cs = th.RTY2Cs(rho, T);
        const double x = U(i, j, k, UMX) * U(i, j, k, UMX);
        const double y = U(i, j, k, UMY) * U(i, j, k, UMY);
        const double z = U(i, j, k, UMZ) * U(i, j, k, UMZ);
        v = sqrt(x + y + z) / rho / cs;
        break;
      }
      case DV_MAGVORT: {
        const double vx = 0.5 * (vel_at<NS>(U, i + 1, j, k, 1) - vel_at<NS>(U, i - 1, j, k, 1)) / J.dx[0];
        const double uy = 0.5 * (vel_at<NS>(U, i, j + 1, k, 0) - vel_at<NS>(U, i, j - 1, k, 0)) / J.dx[1];
        const double v3 = vx - uy;
        const double wx = 0.5 * (vel_at<NS>(U, i + 1, j, k, 2) - vel_at<NS>(U, i - 1, j, k, 2)) / J.dx[0];
        const double wy = 0.5 * (vel_at<NS>(U, i, j + 1, k, 2) - vel_at<NS>(U, i, j - 1, k, 2)) / J.dx[1];
        const double uz = 0.5 * (vel_at<NS>(U, i, j, k + 1, 0) - vel_at<NS>(U, i, j, k - 1, 0)) / J.dx[2];
        const double vz = 0.5 * (vel_at<NS>(U, i, j, k + 1, 1) - vel_at<NS>(U, i, j, k - 1, 1)) / J.dx[2];
        const double v1 = wy - vz;
        const double v2 = uz - wx;
        v = sqrt(0. + v3 * v3 + v1 * v1 + v2 * v2);
        break;
      }
      case DV_LOGDEN: v = log10(rho); break;
      default: break;
    }
    J.upd(i, j, k, m) = v;
  }
}
template <class EOS>
__global__ void __launch_bounds__(BLOCK)
    k_derive(Sched sc, KArgs ka, DeriveArgs da) {
  b_derive<EOS>(sc, ka, first_block(sc), da);
}

}  // namespace pcd
