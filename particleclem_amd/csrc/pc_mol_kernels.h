// Method-of-lines hydro stage kernels for gfx950.
//
// The MOL step (do_mol_advance) is a two-stage Heun integration of
// dU/dt = -div(F)/V with fluxes from limited piecewise-linear face states
// and the same approximate Riemann solver as the Godunov path (or the
// Lax-Friedrichs flux).  Per stage the launches are
//   primitives on grow(bx,2) (ghost values read at their source cells) ->
//   one flux launch per direction (d-faces [L, H+1] x transverse [L, H]) ->
//   one update launch on [L, H].
// A flux thread recomputes the slopes of the two cells of its face from
// the primitives instead of reading a stored slope array: the slope of a
// cell costs ~60 flops on values that sit in L2, storing it would cost
// (4 + NS) words of HBM traffic per cell and direction.
#pragma once
#include "pc_level_kernels.h"
#include "pc_trace_lds.h"

namespace pcd {

// The monotonised-central limiter of mol_slope (Source/MOL.H:105-113).
template <class FP>
PCD FP mol_lim(FP l, FP r) {
  const FP dcen = FP(0.5) * (l + r);
  const FP dlim = l * r >= FP(0.0) ? FP(2.0) * mn(fabs(l), fabs(r)) : FP(0.0);
  return copysign(FP(1.0), dcen) * mn(dlim, fabs(dcen));
}

// Face state extrapolated half a cell from cell (i,j,k) in direction D:
// the slope of the cell (Source/MOL.H:31-120, regular cell; zero when
// `slopes` is off, as mol_iorder == 1 leaves dq at its setV value) and the
// state of Source/MOL.cpp:61-104.  LEFT: qtempl of face c+1/2 (+ half a
// slope); otherwise qtempr of face c-1/2 (- half a slope).  Outputs in the
// normal frame: rho, un, ut (first transverse), utt (second), p, Y.
template <int NS, int D, bool LEFT, class FP, class VW, class VA>
PCD void mol_face_state(const VW& q, const VA& qa, bool slopes, int i,
                        int j, int k, FP& rho, FP& un, FP& ut, FP& utt, FP& p,
                        FP* Y) {
  constexpr int IU = QU + D;
  constexpr int IV = (D == 0) ? QV : QU;
  constexpr int IV2 = (D == 2) ? QV : QW;
  int im = i, jm = j, km = k, ip = i, jp = j, kp = k;
  step<D>(im, jm, km, -1);
  step<D>(ip, jp, kp, 1);
  const FP rc = q(i, j, k, QRHO);
  const FP cc = qa(i, j, k, QC);
  const FP pc = q(i, j, k, QPRES);
  const FP uc = q(i, j, k, IU), vc = q(i, j, k, IV), wc = q(i, j, k, IV2);
  const FP rcc = FP(1.0) / cc;
  FP d0 = FP(0.0), d1 = FP(0.0), d2 = FP(0.0), d3 = FP(0.0);
  FP dY[NS];
#pragma unroll
  for (int n = 0; n < NS; ++n) dY[n] = FP(0.0);
  if (slopes) {
    const FP dpl = pc - q(im, jm, km, QPRES);
    const FP dpr = q(ip, jp, kp, QPRES) - pc;
    const FP dul = uc - q(im, jm, km, IU);
    const FP dur = q(ip, jp, kp, IU) - uc;
    const FP al = div_by(FP(0.5) * dpl, cc, rcc);
    const FP ar = div_by(FP(0.5) * dpr, cc, rcc);
    const FP hr = FP(0.5) * rc;
    d0 = mol_lim(al - hr * dul, ar - hr * dur);
    d1 = mol_lim(al + hr * dul, ar + hr * dur);
    d2 = mol_lim(vc - q(im, jm, km, IV), q(ip, jp, kp, IV) - vc);
    d3 = mol_lim(wc - q(im, jm, km, IV2), q(ip, jp, kp, IV2) - wc);
    const FP c2 = cc * cc;
    const FP rc2 = FP(1.0) / c2;
    const FP rm = q(im, jm, km, QRHO), rp = q(ip, jp, kp, QRHO);
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const FP yc = q(i, j, k, QFS + n);
      const FP l = rc * yc - rm * q(im, jm, km, QFS + n) - div_by(yc * dpl, c2, rc2);
      const FP r = rp * q(ip, jp, kp, QFS + n) - rc * yc - div_by(yc * dpr, c2, rc2);
      dY[n] = mol_lim(l, r);
    }
  }
  const FP dsum = d0 + d1;
  FP rY[NS];
  FP rs = FP(0.0);
  if constexpr (LEFT) {
    un = uc + FP(0.5) * ((d1 - d0) / rc);
    p = pc + FP(0.5) * dsum * cc;
    ut = vc + FP(0.5) * d2;
    utt = wc + FP(0.5) * d3;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const FP yc = q(i, j, k, QFS + n);
      rY[n] = yc * rc + FP(0.5) * (dY[n] + div_by(yc * dsum, cc, rcc));
      rs += rY[n];
    }
  } else {
    un = uc - FP(0.5) * ((d1 - d0) / rc);
    p = pc - FP(0.5) * dsum * cc;
    ut = vc - FP(0.5) * d2;
    utt = wc - FP(0.5) * d3;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const FP yc = q(i, j, k, QFS + n);
      rY[n] = yc * rc - FP(0.5) * (dY[n] + div_by(yc * dsum, cc, rcc));
      rs += rY[n];
    }
  }
  const FP rrs = FP(1.0) / rs;
#pragma unroll
  for (int n = 0; n < NS; ++n) Y[n] = div_by(rY[n], rs, rrs);
  rho = rs;
}

// pc_compute_hyp_mol_flux on d-face (i,j,k) between cells (i,j,k)-e_D and
// (i,j,k) (Source/MOL.cpp:6-190, regular cells, no rotating frame, no
// advected or linear scalars), the primitives read through the views q/qa;
// the extensive flux goes to flx[D] (the flux register starts from zero:
// 0 + F * area).
// FULL: store the zero UTEMP flux and the UEINT flux (the box entry point's
// caller reads them); the level and batch updates take the UTEMP divergence
// as zero and reset UEINT from UEDEN and the momenta, so they need neither.
template <class EOS, int D, bool LAXF, class FP, bool FULL, class VQ, class VA>
PCD void mol_face_flux(const JobT<FP>& J, const KArgs& ka, const VQ& q, const VA& qa,
                       int i, int j, int k) {
  constexpr int NS = EOS::NS;
  const bool slopes = ka.plm_iorder != 1;  // Source/Diffusion.cpp:371-373
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  FP rl, ul, vl, v2l, pl, rr, ur, vr, v2r, pr;
  FP Yl[NS], Yr[NS];
  mol_face_state<NS, D, true, FP>(q, qa, slopes, im, jm, km, rl, ul, vl, v2l,
                                  pl, Yl);
  mol_face_state<NS, D, false, FP>(q, qa, slopes, i, j, k, rr, ur, vr, v2r,
                                   pr, Yr);
  RiemannOut<EOS, FP> o;
  if constexpr (LAXF) {
    o = laxf<EOS, FP>(rl, ul, vl, v2l, pl, Yl, rr, ur, vr, v2r, pr, Yr);
  } else {
    const FP cavg = FP(0.5) * (qa(i, j, k, QC) + qa(im, jm, km, QC));
    o = riemann<EOS, FP>(rl, ul, vl, v2l, pl, Yl, rr, ur, vr, v2r, pr, Yr, cavg);
  }
  constexpr int MU = UMX + D;
  constexpr int MT = (D == 0) ? UMY : UMX;
  constexpr int MTT = (D == 2) ? UMY : UMZ;
  const double area = J.area[D].p ? J.area[D](i, j, k, 0) : J.area_u[D];
  if constexpr (sizeof(FP) == 8) {
    const auto& fo = J.flx[D];
    fo(i, j, k, URHO) = 0.0 + (double)o.f_rho * area;
    fo(i, j, k, MU) = 0.0 + (double)o.f_un * area;
    fo(i, j, k, MT) = 0.0 + (double)o.f_ut * area;
    fo(i, j, k, MTT) = 0.0 + (double)o.f_utt * area;
    fo(i, j, k, UEDEN) = 0.0 + (double)o.f_eden * area;
    if (FULL) {
      fo(i, j, k, UEINT) = 0.0 + (double)o.f_eint * area;
      fo(i, j, k, UTEMP) = 0.0;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) fo(i, j, k, UFS + s) = 0.0 + (double)o.f_Y[s] * area;
  } else {
    // mixed step (level and batch): the FP32 Riemann solve's face fluxes
    // stored in FP32 (the updates read them into FP64 arithmetic); neither
    // the zero UTEMP flux nor the UEINT flux is stored
    const VT<float> fo = as_fp32(J.flx[D]);
    fo(i, j, k, URHO) = (float)(0.0 + (double)o.f_rho * area);
    fo(i, j, k, MU) = (float)(0.0 + (double)o.f_un * area);
    fo(i, j, k, MT) = (float)(0.0 + (double)o.f_ut * area);
    fo(i, j, k, MTT) = (float)(0.0 + (double)o.f_utt * area);
    fo(i, j, k, UEDEN) = (float)(0.0 + (double)o.f_eden * area);
#pragma unroll
    for (int s = 0; s < NS; ++s) fo(i, j, k, UFS + s) = (float)(0.0 + (double)o.f_Y[s] * area);
  }
}

// The MOL fluxes of the d-faces of one box, one thread per face, the
// primitives read from the level arrays.
template <class EOS, int D, bool LAXF, class FP, bool FULL>
PCD void b_mol_flux(const Sched& sc, const KArgs& ka, int lb_) {
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_MOL_X + D, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const VB<FP> q = bufview(J.q), qa = bufview(J.qa);
  mol_face_flux<EOS, D, LAXF, FP, FULL>(J, ka, q, qa, i, j, k);
}
template <class EOS, int D, bool LAXF, class FP = double, bool PERSIST = false,
          bool FULL = true>
__global__ void __launch_bounds__(BLOCK)
    k_mol_flux(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_mol_flux<EOS, D, LAXF, FP, FULL>(sc, ka, lb_);
  } else {
    b_mol_flux<EOS, D, LAXF, FP, FULL>(sc, ka, first_block(sc));
  }
}

// ------------------------------------ MOL fluxes from tiles in LDS --
// A workgroup owns a tile of d-faces: for x, whole rows of x-faces (ta =
// n[0] faces) by tb rows along y; for y and z, ta <= 32 faces along x by tb
// along D (axis b = trace_axis_b(D) = D).  The primitives of the tile's
// cells, [f0 - 2, f0 + extent] along D (each face reads the three-cell slope
// stencils of its two cells), are loaded from HBM once, coalesced along x,
// into LDS (rho, the three velocities, p, the species and c: 6 + NS
// components); every face thread then reads its stencils there.  Arithmetic
// and its order are those of mol_face_flux with the level-array views, so
// the fluxes are the same bits.
constexpr int MOL_LDS_CELLS = BLOCK + 4 * 64;

__host__ __device__ inline TraceTile mol_tile(int D, const int n[3]) {
  TraceTile t;
  const int b = trace_axis_b(D), c = 3 - b;
  if (D == 0) {
    t.ta = n[0];
    t.tb = BLOCK / t.ta;
    const int fit = MOL_LDS_CELLS / (t.ta + 3);
    if (t.tb > fit) t.tb = fit;
    if (t.tb > n[b]) t.tb = n[b];
  } else {
    // rows of up to 32 cells along x (coalesced staging), as many faces
    // along D as the workgroup holds; a tile stages its faces' cells plus
    // three (two below, one above) along D
    t.ta = n[0] < 32 ? n[0] : 32;
    t.tb = BLOCK / t.ta;
    const int fit = MOL_LDS_CELLS / t.ta - 3;
    if (t.tb > fit) t.tb = fit;
    if (t.tb > n[b]) t.tb = n[b];
  }
  t.nta = (n[0] + t.ta - 1) / t.ta;
  t.ntb = (n[b] + t.tb - 1) / t.tb;
  t.nc = n[c];
  return t;
}

// Whether the x-face rows fit one workgroup (n: the valid box's extents;
// the x-face region is n[0] + 1 long).
__host__ __device__ inline bool mol_pencils_fit(const int n[3]) {
  return n[0] + 1 <= BLOCK;
}

// Window cells of the tile of direction D over the d-face region n.
__host__ __device__ inline int mol_window_cells(int D, const int n[3]) {
  const TraceTile t = mol_tile(D, n);
  return D == 0 ? (t.ta + 3) * t.tb : t.ta * (t.tb + 3);
}

// Component slot of a primitive in the window (QA: the qaux view, whose
// only component read is QC).
template <int NS, bool QA>
PCD int mol_slot(int comp) {
  if (QA) return 5 + NS;
  return comp == QRHO ? 0 : comp == QU ? 1 : comp == QV ? 2 : comp == QW ? 3
       : comp == QPRES ? 4 : 5 + (comp - QFS);
}

// A view of the window by absolute cell index: window cell (i - x0) +
// wa (c[AB] - b0), component slot m at lds[m * nwin + cell].
template <class FP, int NS, int AB, bool QA>
struct MolWin {
  const FP* lds;
  int nwin, wa, x0, b0;
  PCD FP operator()(int i, int j, int k, int comp) const {
    const int cb = AB == 1 ? j : k;
    return lds[mol_slot<NS, QA>(comp) * nwin + (i - x0) + wa * (cb - b0)];
  }
};

template <class EOS, int D, bool LAXF, class FP>
PCD void b_mol_flux_lds(const Sched& sc, const KArgs& ka, int lb_, FP* lds) {
  constexpr int NS = EOS::NS;
  constexpr int AB = (D == 2) ? 2 : 1;  // axis b
  constexpr int AC = 3 - AB;            // axis c
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3];
  stage_region(ST_MOL_X + D, J.lo, J.hi, lo, n);
  const TraceTile T = mol_tile(D, n);
  const int tile = local / BLOCK, t = local % BLOCK;
  if (tile >= T.nta * T.ntb * T.nc) return;  // whole workgroup: no barrier crossed
  const int ia = tile % T.nta, r = tile / T.nta;
  const int ib = r % T.ntb, ic = r / T.ntb;
  const int ta_ = t % T.ta, tb_ = t / T.ta;
  int ijk[3];
  ijk[0] = lo[0] + ia * T.ta + ta_;
  ijk[AB] = lo[AB] + ib * T.tb + tb_;
  ijk[AC] = lo[AC] + ic;
  const bool active = tb_ < T.tb && ijk[0] < lo[0] + n[0] && ijk[AB] < lo[AB] + n[AB];
  // window: the tile's cells, two more below and one more above along D
  constexpr int HA = (D == 0) ? 2 : 0, HB = (D == 0) ? 0 : 2;
  const int wa = T.ta + (D == 0 ? 3 : 0), wb = T.tb + (D == 0 ? 0 : 3);
  const int nwin = wa * wb;
  const int x0 = lo[0] + ia * T.ta - HA, b0 = lo[AB] + ib * T.tb - HB;
  // cells of the window inside the tile's part of the region (the last
  // tile of a row may be partial): along D [lo - 2, lo + n], across [lo, lo + n)
  const int xe = (D == 0) ? lo[0] + n[0] : lo[0] + n[0] - 1;
  const int be = (D == 0) ? lo[AB] + n[AB] - 1 : lo[AB] + n[AB];
  {
    const VB<FP> q = bufview(J.q), qa = bufview(J.qa);
    for (int w = threadIdx.x; w < nwin; w += BLOCK) {
      int g[3];
      g[0] = x0 + w % wa;
      g[AB] = b0 + w / wa;
      g[AC] = lo[AC] + ic;
      if (g[0] > xe || g[AB] > be) continue;
      lds[0 * nwin + w] = q(g[0], g[1], g[2], QRHO);
      lds[1 * nwin + w] = q(g[0], g[1], g[2], QU);
      lds[2 * nwin + w] = q(g[0], g[1], g[2], QV);
      lds[3 * nwin + w] = q(g[0], g[1], g[2], QW);
      lds[4 * nwin + w] = q(g[0], g[1], g[2], QPRES);
#pragma unroll
      for (int s = 0; s < NS; ++s) lds[(5 + s) * nwin + w] = q(g[0], g[1], g[2], QFS + s);
      lds[(5 + NS) * nwin + w] = qa(g[0], g[1], g[2], QC);
    }
  }
  __syncthreads();
  if (!active) return;
  const MolWin<FP, NS, AB, false> qv{lds, nwin, wa, x0, b0};
  const MolWin<FP, NS, AB, true> qav{lds, nwin, wa, x0, b0};
  mol_face_flux<EOS, D, LAXF, FP, false>(J, ka, qv, qav, ijk[0], ijk[1], ijk[2]);
}

// Dynamic LDS: (6 + NS) components of the largest window of the launch's
// boxes (mol_window_cells), sizeof(FP) each.
template <class EOS, int D, bool LAXF, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_mol_flux_lds(Sched sc, KArgs ka) {
  extern __shared__ __align__(16) unsigned char pc_mol_dyn[];
  FP* lds = (FP*)pc_mol_dyn;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x) {
      b_mol_flux_lds<EOS, D, LAXF, FP>(sc, ka, lb_, lds);
      __syncthreads();  // the next tile restages the buffer
    }
  } else {
    b_mol_flux_lds<EOS, D, LAXF, FP>(sc, ka, first_block(sc), lds);
  }
}

// pc_flux_div (Source/Diffterm.H:470-487): -div(F)/V of one cell.
// LEVEL: the level and batch updates, whose face fluxes store neither the
// zero UTEMP flux (its divergence -(0 - 0 + 0 - 0 + 0 - 0) / v = -0.0 is the
// value the sum gives) nor the UEINT flux (UEINT is reset after the update,
// D[UEINT] = 0 stands in).
template <int NS, bool LEVEL = false>
PCD void mol_divergence(const Job& J, int i, int j, int k, double* D) {
  constexpr int NVAR = 7 + NS;
  const double v = J.vol.p ? J.vol(i, j, k, 0) : J.vol_u;
  const V& f0 = J.flx[0];
  const V& f1 = J.flx[1];
  const V& f2 = J.flx[2];
#pragma unroll
  for (int m = 0; m < NVAR; ++m) {
    if (LEVEL && m == UTEMP) {
      D[m] = -0.0 / v;
      continue;
    }
    if (LEVEL && m == UEINT) {
      D[m] = 0.0;
      continue;
    }
    D[m] = -(f0(i + 1, j, k, m) - f0(i, j, k, m) + f1(i, j + 1, k, m) -
             f1(i, j, k, m) + f2(i, j, k + 1, m) - f2(i, j, k, m)) /
           v;
  }
}
// the same for the mixed step's FP32 face fluxes (FP64 arithmetic)
template <int NS>
PCD void mol_divergence_fp32(const Job& J, int i, int j, int k, double* D) {
  constexpr int NVAR = 7 + NS;
  const double v = J.vol.p ? J.vol(i, j, k, 0) : J.vol_u;
  const VB<float> f0 = bufview(as_fp32(J.flx[0]));
  const VB<float> f1 = bufview(as_fp32(J.flx[1]));
  const VB<float> f2 = bufview(as_fp32(J.flx[2]));
  auto F = [](const VB<float>& f, int a, int b, int c, int m) { return (double)f(a, b, c, m); };
#pragma unroll
  for (int m = 0; m < NVAR; ++m) {
    if (m == UTEMP) {
      D[m] = -0.0 / v;
      continue;
    }
    if (m == UEINT) {  // not stored (mol_face_flux); UEINT is reset
      D[m] = 0.0;
      continue;
    }
    D[m] = -(F(f0, i + 1, j, k, m) - F(f0, i, j, k, m) + F(f1, i, j + 1, k, m) -
             F(f1, i, j, k, m) + F(f2, i, j, k + 1, m) - F(f2, i, j, k, m)) /
           v;
  }
}

// The MOL source of one box into J.upd (getMOLSrcTerm's copy of Dterm to
// MOLSrcTerm, Source/Diffusion.cpp:654: overwritten, not accumulated).
template <int NS>
PCD void b_mol_div(const Sched& sc, const KArgs& ka, int lb_) {
  int local;
  const Job J = load_job(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_MOL_UPD, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  double D[7 + NS];
  mol_divergence<NS>(J, i, j, k, D);
#pragma unroll
  for (int m = 0; m < 7 + NS; ++m) J.upd(i, j, k, m) = D[m];
}
template <int NS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_mol_div(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_mol_div<NS>(sc, ka, lb_);
  } else {
    b_mol_div<NS>(sc, ka, first_block(sc));
  }
}

// One Heun stage of do_mol_advance on a cell (Source/Advance.cpp:110-176):
// STAGE 0: U* = 1.0 U^n + dt S(U^n), computeTemp, into the stage array
// J.sold (the state keeps U^n, so it is never copied; the second stage's
// primitives read U* there); STAGE 1: U^{n+1} = 0.5 U* + 0.5 U^n, += (0.5
// dt) S(U*), computeTemp and post_timestep's computeTemp, into the state;
// NaN check and the CFL partial minimum.
// F32: the face fluxes are the mixed step's FP32 ones (mol_divergence_fp32)
template <class EOS, int STAGE, bool F32>
PCD double b_mol_update(const Sched& sc, const KArgs& ka, int lb_,
                        DtState*& ds) {
  constexpr int NS = EOS::NS;
  constexpr int NVAR = 7 + NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  if (J.ds) ds = J.ds;
  int lo[3], n[3], i, j, k;
  stage_region(ST_MOL_UPD, J.lo, J.hi, lo, n);
  double mydt = 1.7976931348623157e308;
  if (region_ijk(sc, lo, n, local, i, j, k)) {
    const double dt = get_dt(J, ka);
    double D[NVAR];
    if constexpr (F32) mol_divergence_fp32<NS>(J, i, j, k, D);
    else mol_divergence<NS, true>(J, i, j, k, D);
    // UEINT is reset from UEDEN and the momenta below before it is read, so
    // neither stage reads it (nor does the first store it)
    double S[NVAR];
    if constexpr (STAGE == 0) {
#pragma unroll
      for (int m = 0; m < NVAR; ++m) S[m] = m == UEINT ? 0.0 : 1.0 * J.U(i, j, k, m) + dt * D[m];
    } else {
      const double hdt = 0.5 * dt;
#pragma unroll
      for (int m = 0; m < NVAR; ++m) {
        if (m == UEINT) {
          S[m] = 0.0;
          continue;
        }
        S[m] = 0.5 * J.sold(i, j, k, m) + 0.5 * J.U(i, j, k, m);
        S[m] += hdt * D[m];
      }
    }
    double Y[NS];
    state_Y<NS>(S, Y);
    const typename EOS::Thermo th(Y);
    reset_and_temp<EOS>(S, th);
    if constexpr (STAGE == 1) reset_and_temp<EOS>(S, th);
    bool bad = false;
    const V& out = STAGE == 0 ? J.sold : J.U;
#pragma unroll
    for (int m = 0; m < NVAR; ++m) {
      if (STAGE == 1 || m != UEINT) out(i, j, k, m) = S[m];
      bad |= !isfinite(S[m]);
    }
    if constexpr (STAGE == 1) {
      if (bad) ds->bad = 1;
      mydt = cell_estdt<EOS>(S, th, J.dx);
    }
  }
  return mydt;
}

template <class EOS, int STAGE, bool PERSIST = false, bool F32 = false>
__global__ void __launch_bounds__(BLOCK)
    k_mol_update(Sched sc, KArgs ka, DtState* ds) {
  double m = 1.7976931348623157e308;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      m = mn(m, b_mol_update<EOS, STAGE, F32>(sc, ka, lb_, ds));
  } else {
    m = b_mol_update<EOS, STAGE, F32>(sc, ka, first_block(sc), ds);
  }
  if constexpr (STAGE == 1) block_min_to(m, &ds->minbits);
}

}  // namespace pcd
