// PPM trace (Source/PPM.cpp:82-423) with the primitive stencil staged in
// LDS.
//
// A workgroup owns a tile of the trace region grow(bx,1): TA cells along x
// by TB along a second axis b (y for the x and y traces, z for the z trace),
// one index along the third.  The tile's primitives plus the 2-cell PPM
// halo along the traced direction are loaded from HBM once, coalesced along
// x, into LDS (the 5 + NS components the stencils read, one barrier); every
// thread then reads its cell and its 5-point stencils from LDS.  Arithmetic
// and operation order are the reference's (ppm_cell_states); outputs are the
// traces of the Compact layouts in pc_kernels.h.
#pragma once
#include "pc_kernels.h"

namespace pcd {

// Tile of the LDS trace for region extents n[3] (cells of grow(bx,1)):
// ta cells along x (the whole row when it is at most 64 wide, else an equal
// split into rows of at most 64), tb along axis b, tiles counted x fastest,
// then b, then the third axis c.
struct TraceTile {
  int ta, tb, nta, ntb, nc;
};

__host__ __device__ inline int trace_axis_b(int D) { return D == 2 ? 2 : 1; }

// Per-component staging footprints are at most (ta + 4) tb cells for the x
// trace and ta (tb + 4) for the others (ta tb <= BLOCK, ta <= 64).
constexpr int TRACE_LDS_CELLS = BLOCK + 4 * 64;

__host__ __device__ inline TraceTile trace_tile(int D, const int n[3]) {
  TraceTile t;
  const int b = trace_axis_b(D), c = 3 - b;
  const int parts = (n[0] + 63) / 64;
  t.ta = (n[0] + parts - 1) / parts;
  t.tb = BLOCK / t.ta;
  const int fit = (D == 0) ? TRACE_LDS_CELLS / (t.ta + 4) : TRACE_LDS_CELLS / t.ta - 4;
  if (t.tb > fit) t.tb = fit;
  if (t.tb > n[b]) t.tb = n[b];
  t.nta = (n[0] + t.ta - 1) / t.ta;
  t.ntb = (n[b] + t.tb - 1) / t.tb;
  t.nc = n[c];
  return t;
}

// Staging cells per component of the largest of the three traces' tiles.
__host__ __device__ inline int trace_lds_cells(const int n[3]) {
  int m = 0;
  for (int D = 0; D < 3; ++D) {
    const TraceTile t = trace_tile(D, n);
    const int c = (D == 0) ? (t.ta + 4) * t.tb : t.ta * (t.tb + 4);
    if (c > m) m = c;
  }
  return m;
}

// Logical blocks of one box for the three LDS traces (one launch plan serves
// all three directions: the largest of their tile counts; a workgroup whose
// tile lies beyond its direction's count has nothing to do).
__host__ __device__ inline long long trace_lds_blocks(const int n[3]) {
  long long m = 0;
  for (int D = 0; D < 3; ++D) {
    const TraceTile t = trace_tile(D, n);
    const long long c = (long long)t.nta * t.ntb * t.nc;
    if (c > m) m = c;
  }
  return m;
}

// PPM states of one cell from a staged window (component m of window cell w
// at lds[m * nwin + w]; rho, u_n, p, u_t, u_tt, then the species): the plus
// state P on the cell's low D-face and the minus state M on its high D-face
// (Source/PPM.cpp:82-423; the reference also traces rho e, a profile it
// never uses).  wc: the cell's window index, wstep: one cell along D.
template <class EOS, int D, class FP>
PCD void ppm_cell_states(const FP* lds, int nwin, int wc, int wstep, FP flatv,
                         FP dtdx, FaceState<EOS::NS, FP>& P,
                         FaceState<EOS::NS, FP>& M) {
  constexpr int NS = EOS::NS;
  constexpr int QUT = (D == 0) ? QV : ((D == 1) ? QW : QU);
  constexpr int QUTT = (D == 0) ? QW : ((D == 1) ? QU : QV);
  FP sm, sp, s6;
  FP Yc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) Yc[s] = lds[(5 + s) * nwin + wc];
  const FP cc = EOS::RPY2Cs(lds[wc], lds[2 * nwin + wc], Yc);
  const FP un = lds[1 * nwin + wc];
  const FP spd0 = un - cc, spd1 = un, spd2 = un + cc;
  auto rc = [&](int slot) {
    FP s[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) s[m] = lds[slot * nwin + wc + (m - 2) * wstep];
    ppm_reconstruct(s, flatv, sm, sp);
    s6 = ppm_s6(sm, sp, s[2]);
  };
  rc(0);
  const FP Im_r0 = ppm_im(sm, sp, s6, spd0, dtdx);
  const FP Im_r1 = ppm_im(sm, sp, s6, spd1, dtdx);
  const FP Ip_r2 = ppm_ip(sm, sp, s6, spd2, dtdx);
  const FP Ip_r1 = ppm_ip(sm, sp, s6, spd1, dtdx);
  rc(1);
  const FP Im_u0 = ppm_im(sm, sp, s6, spd0, dtdx);
  const FP Im_u2 = ppm_im(sm, sp, s6, spd2, dtdx);
  const FP Ip_u2 = ppm_ip(sm, sp, s6, spd2, dtdx);
  const FP Ip_u0 = ppm_ip(sm, sp, s6, spd0, dtdx);
  rc(2);
  const FP Im_p0 = ppm_im(sm, sp, s6, spd0, dtdx);
  const FP Im_p1 = ppm_im(sm, sp, s6, spd1, dtdx);
  const FP Im_p2 = ppm_im(sm, sp, s6, spd2, dtdx);
  const FP Ip_p0 = ppm_ip(sm, sp, s6, spd0, dtdx);
  const FP Ip_p1 = ppm_ip(sm, sp, s6, spd1, dtdx);
  const FP Ip_p2 = ppm_ip(sm, sp, s6, spd2, dtdx);
  rc(3);
  const FP Im_t1 = ppm_im(sm, sp, s6, spd1, dtdx);
  const FP Ip_t1 = ppm_ip(sm, sp, s6, spd1, dtdx);
  rc(4);
  const FP Im_tt1 = ppm_im(sm, sp, s6, spd1, dtdx);
  const FP Ip_tt1 = ppm_ip(sm, sp, s6, spd1, dtdx);
  FP YrefP[NS], YrefM[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rc(5 + s);
    YrefP[s] = ppm_im(sm, sp, s6, spd0, dtdx);
    P.Y[s] = ppm_im(sm, sp, s6, spd1, dtdx);
    YrefM[s] = ppm_ip(sm, sp, s6, spd2, dtdx);
    M.Y[s] = ppm_ip(sm, sp, s6, spd1, dtdx);
  }
  FP vel[3];
  // ---- plus state on face i (reference state: fastest wave to the left)
  {
    FP rho_ref = Im_r0;
    const FP un_ref = Im_u0;
    FP p_ref = Im_p0;
    rho_ref = mx(rho_ref, rmin<FP>());
    const FP rho_ref_inv = FP(1.0) / rho_ref;
    p_ref = mx(p_ref, rmin<FP>());
    const FP cc_ref = EOS::RPY2Cs(rho_ref, p_ref, YrefP);
    const FP csq_ref = cc_ref * cc_ref;
    const FP cc_ref_inv = FP(1.0) / cc_ref;
    const FP dum = un_ref - Im_u0;
    const FP dptotm = p_ref - Im_p0;
    const FP drho = rho_ref - Im_r1;
    const FP dptot = p_ref - Im_p1;
    const FP dup = un_ref - Im_u2;
    const FP dptotp = p_ref - Im_p2;
    FP alpham =
        FP(0.5) * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
    FP alphap =
        FP(0.5) * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
    FP alpha0r = drho - dptot / csq_ref;
    alpham = un - cc > FP(0.0) ? FP(0.0) : -alpham;
    alphap = un + cc > FP(0.0) ? FP(0.0) : -alphap;
    alpha0r = un > FP(0.0) ? FP(0.0) : -alpha0r;
    P.rho = mx(rmin<FP>(), rho_ref + alphap + alpham + alpha0r);
    vel[D] = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
    P.p = mx(rmin<FP>(), p_ref + (alphap + alpham) * csq_ref);
    vel[QUT - QU] = Im_t1;
    vel[QUTT - QU] = Im_tt1;
    P.u = vel[0];
    P.v = vel[1];
    P.w = vel[2];
    // the reference traces rho e = rho RYP2E(rho, Y, p) here
    // (Source/PPM.cpp:324-330); no flux reads it (store_trace)
    P.rhoe = FP(0.0);
  }
  // ---- minus state on face i+1 (fastest wave to the right)
  {
    FP rho_ref = Ip_r2;
    const FP un_ref = Ip_u2;
    FP p_ref = Ip_p2;
    rho_ref = mx(rho_ref, rmin<FP>());
    const FP rho_ref_inv = FP(1.0) / rho_ref;
    p_ref = mx(p_ref, rmin<FP>());
    const FP cc_ref = EOS::RPY2Cs(rho_ref, p_ref, YrefM);
    const FP csq_ref = cc_ref * cc_ref;
    const FP cc_ref_inv = FP(1.0) / cc_ref;
    const FP dum = un_ref - Ip_u0;
    const FP dptotm = p_ref - Ip_p0;
    const FP drho = rho_ref - Ip_r1;
    const FP dptot = p_ref - Ip_p1;
    const FP dup = un_ref - Ip_u2;
    const FP dptotp = p_ref - Ip_p2;
    FP alpham =
        FP(0.5) * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
    FP alphap =
        FP(0.5) * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
    FP alpha0r = drho - dptot / csq_ref;
    alpham = un - cc > FP(0.0) ? -alpham : FP(0.0);
    alphap = un + cc > FP(0.0) ? -alphap : FP(0.0);
    alpha0r = un > FP(0.0) ? -alpha0r : FP(0.0);
    M.rho = mx(rmin<FP>(), rho_ref + alphap + alpham + alpha0r);
    vel[D] = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
    M.p = mx(rmin<FP>(), p_ref + (alphap + alpham) * csq_ref);
    vel[QUT - QU] = Ip_t1;
    vel[QUTT - QU] = Ip_tt1;
    M.u = vel[0];
    M.v = vel[1];
    M.w = vel[2];
    M.rhoe = FP(0.0);  // Source/PPM.cpp:415-421; see the plus state
  }
}

// Stage the 5 + NS components the PPM stencils read for the window cells
// [0, nwin): window cell w is (a0 + w % wa, b0 + w / wa) on axes (x, AB) at
// index c on the third axis; cells outside the primitive box are skipped.
template <int NS, int D, int AB, class FP>
PCD void stage_ppm_window(const JobT<FP>& J, FP* lds, int nwin, int wa, int a0,
                          int b0, int c) {
  constexpr int AC = 3 - AB;
  constexpr int QUN = QU + D;
  constexpr int QUT = (D == 0) ? QV : ((D == 1) ? QW : QU);
  constexpr int QUTT = (D == 0) ? QW : ((D == 1) ? QU : QV);
  const VB<FP> q = bufview(J.q);
  const int qlo[3] = {J.q.b0, J.q.b1, J.q.b2};
  for (int w = threadIdx.x; w < nwin; w += BLOCK) {
    const int wa_ = w % wa, wb_ = w / wa;
    int g[3];
    g[0] = a0 + wa_;
    g[AB] = b0 + wb_;
    g[AC] = c;
    const bool in = g[0] >= qlo[0] && g[0] < J.qe[0] && g[1] >= qlo[1] &&
                    g[1] < J.qe[1] && g[2] >= qlo[2] && g[2] < J.qe[2];
    if (in) {
      lds[0 * nwin + w] = q(g[0], g[1], g[2], QRHO);
      lds[1 * nwin + w] = q(g[0], g[1], g[2], QUN);
      lds[2 * nwin + w] = q(g[0], g[1], g[2], QPRES);
      lds[3 * nwin + w] = q(g[0], g[1], g[2], QUT);
      lds[4 * nwin + w] = q(g[0], g[1], g[2], QUTT);
#pragma unroll
      for (int s = 0; s < NS; ++s) lds[(5 + s) * nwin + w] = q(g[0], g[1], g[2], QFS + s);
    }
  }
}

template <class EOS, int D, class FP>
PCD void b_trace_ppm_lds(const Sched& sc, const KArgs& ka, int lb_, FP* lds) {
  constexpr int NS = EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3];
  stage_region(ST_TRACE, J.lo, J.hi, lo, n);
  const TraceTile T = trace_tile(D, n);
  constexpr int AB = (D == 2) ? 2 : 1;  // axis b
  constexpr int AC = 3 - AB;            // axis c
  const int tile = local / BLOCK, t = local % BLOCK;
  if (tile >= T.nta * T.ntb * T.nc) return;  // whole workgroup: no barrier crossed
  const int ia = tile % T.nta, r = tile / T.nta;
  const int ib = r % T.ntb, ic = r / T.ntb;
  // cell of this thread
  const int ta_ = t % T.ta, tb_ = t / T.ta;
  int ijk[3];
  ijk[0] = lo[0] + ia * T.ta + ta_;
  ijk[AB] = lo[AB] + ib * T.tb + tb_;
  ijk[AC] = lo[AC] + ic;
  const bool active = tb_ < T.tb && ijk[0] < lo[0] + n[0] && ijk[AB] < lo[AB] + n[AB];
  const int i = ijk[0], j = ijk[1], k = ijk[2];
  // staging window: the tile plus 2 cells each side along D
  constexpr int HA = (D == 0) ? 2 : 0, HB = (D == 0) ? 0 : 2;
  const int wa = T.ta + 2 * HA, wb = T.tb + 2 * HB;
  const int nwin = wa * wb;
  const int a0 = lo[0] + ia * T.ta - HA, b0 = lo[AB] + ib * T.tb - HB;
  stage_ppm_window<NS, D, AB, FP>(J, lds, nwin, wa, a0, b0, lo[AC] + ic);
  __syncthreads();
  if (!active) return;
  const int wc = (ta_ + HA) + wa * (tb_ + HB);
  const int wstep = (D == 0) ? 1 : wa;
  const FP dtdx = get_dt(J, ka) / J.dx[D];
  const FP flatv = ka.use_flattening ? FP(J.flat(i, j, k, 0)) : FP(1.0);
  FaceState<NS, FP> P, M;
  ppm_cell_states<EOS, D, FP>(lds, nwin, wc, wstep, flatv, dtdx, P, M);
  const auto& tr = J.tr[D];
  store_trace<NS, FP>(tr, i, j, k, 1, P);
  int i1 = i, j1 = j, k1 = k;
  step<D>(i1, j1, k1, 1);
  store_trace<NS, FP>(tr, i1, j1, k1, 0, M);
}

// Dynamic LDS: (5 + NS) components of the largest window of the launch's
// boxes (trace_lds_cells), sizeof(FP) each.
template <class EOS, int D, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_trace_ppm_lds(Sched sc, KArgs ka) {
  extern __shared__ __align__(16) unsigned char pc_trace_dyn[];
  FP* lds = (FP*)pc_trace_dyn;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x) {
      b_trace_ppm_lds<EOS, D, FP>(sc, ka, lb_, lds);
      __syncthreads();  // the next tile restages the buffer
    }
  } else {
    b_trace_ppm_lds<EOS, D, FP>(sc, ka, first_block(sc), lds);
  }
}

// ------------------------------------------ fused trace + first fluxes --
// Pencil tiles: the whole trace-region extent along D in one workgroup (the
// extent must be at most BLOCK), with as many neighbouring pencils as fit
// (ta along x by tb along axis b, as trace_tile; along D the tile is the
// full extent).  Every D-face of the first-flux region then has both of its
// cells in the workgroup.
__host__ __device__ inline TraceTile pencil_tile(int D, const int n[3]) {
  TraceTile t;
  const int b = trace_axis_b(D), c = 3 - b;
  if (D == 0) {
    t.ta = n[0];
    t.tb = BLOCK / t.ta;
    const int fit = TRACE_LDS_CELLS / (t.ta + 4);
    if (t.tb > fit) t.tb = fit;
    if (t.tb > n[b]) t.tb = n[b];
  } else {
    t.tb = n[b];
    t.ta = BLOCK / t.tb;
    const int fit = TRACE_LDS_CELLS / (t.tb + 4);
    if (t.ta > fit) t.ta = fit;
    if (t.ta > n[0]) t.ta = n[0];
  }
  t.nta = (n[0] + t.ta - 1) / t.ta;
  t.ntb = (n[b] + t.tb - 1) / t.tb;
  t.nc = n[c];
  return t;
}

// The fused kernel applies when every extent of the trace region fits one
// workgroup's pencil.
__host__ __device__ inline bool pencil_fits(const int n[3]) {
  return n[0] <= BLOCK && n[1] <= BLOCK && n[2] <= BLOCK;
}

__host__ __device__ inline long long pencil_blocks(const int n[3]) {
  long long m = 0;
  for (int D = 0; D < 3; ++D) {
    const TraceTile t = pencil_tile(D, n);
    const long long c = (long long)t.nta * t.ntb * t.nc;
    if (c > m) m = c;
  }
  return m;
}

__host__ __device__ inline int pencil_cells(const int n[3]) {
  int m = 0;
  for (int D = 0; D < 3; ++D) {
    const TraceTile t = pencil_tile(D, n);
    const int c = (D == 0) ? (t.ta + 4) * t.tb : t.ta * (t.tb + 4);
    if (c > m) m = c;
  }
  return m;
}

// PPM trace of direction D and, from the traced states still in the
// workgroup, the first fluxes on the D-faces (Source/Godunov.cpp:182-242
// after the trace of :138-204): the plus state P of a cell stays in its
// thread, the minus state M goes through LDS to the thread of the next cell
// along D, whose thread solves the face between them.  Traces are stored for
// the transverse and final stages as by b_trace_ppm_lds; fluxes as b_flux1.
// LDS: the staging window, then reused for M (QT components per tile cell).
template <class EOS, int D, class FP>
PCD void b_trace_flux1_lds(const Sched& sc, const KArgs& ka, int lb_, FP* lds) {
  constexpr int NS = EOS::NS;
  constexpr int QT = 6 + NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3];
  stage_region(ST_TRACE, J.lo, J.hi, lo, n);
  const TraceTile T = pencil_tile(D, n);
  constexpr int AB = (D == 2) ? 2 : 1;  // axis b
  constexpr int AC = 3 - AB;            // axis c
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
  const int i = ijk[0], j = ijk[1], k = ijk[2];
  constexpr int HA = (D == 0) ? 2 : 0, HB = (D == 0) ? 0 : 2;
  const int wa = T.ta + 2 * HA, wb = T.tb + 2 * HB;
  const int nwin = wa * wb;
  stage_ppm_window<NS, D, AB, FP>(J, lds, nwin, wa, lo[0] + ia * T.ta - HA,
                                  lo[AB] + ib * T.tb - HB, lo[AC] + ic);
  __syncthreads();
  FaceState<NS, FP> P, M;
  if (active) {
    const int wc = (ta_ + HA) + wa * (tb_ + HB);
    const int wstep = (D == 0) ? 1 : wa;
    const FP dtdx = get_dt(J, ka) / J.dx[D];
    const FP flatv = ka.use_flattening ? FP(J.flat(i, j, k, 0)) : FP(1.0);
    ppm_cell_states<EOS, D, FP>(lds, nwin, wc, wstep, flatv, dtdx, P, M);
    const auto& tr = J.tr[D];
    store_trace<NS, FP>(tr, i, j, k, 1, P);
    int i1 = i, j1 = j, k1 = k;
    step<D>(i1, j1, k1, 1);
    store_trace<NS, FP>(tr, i1, j1, k1, 0, M);
  }
  __syncthreads();  // the window is read: reuse the buffer for M
  const int ncell = T.ta * T.tb;
  const int me = ta_ + T.ta * tb_;
  if (active) {
    lds[0 * ncell + me] = M.rho;
    lds[1 * ncell + me] = M.u;
    lds[2 * ncell + me] = M.v;
    lds[3 * ncell + me] = M.w;
    lds[4 * ncell + me] = M.p;
#pragma unroll
    for (int s = 0; s < NS; ++s) lds[(6 + s) * ncell + me] = M.Y[s];
  }
  __syncthreads();
  // this cell's low D-face; the region's first cell along D has none
  const int pd = (D == 0) ? ta_ : tb_;
  const int nd = (D == 0) ? n[0] : n[AB];
  const bool face = active && pd > 0;
  Flux<NS, FP> F;
  if (face) {
    const int prev = (D == 0) ? me - 1 : me - T.ta;
    FaceState<NS, FP> Lf;
    Lf.rho = lds[0 * ncell + prev];
    Lf.u = lds[1 * ncell + prev];
    Lf.v = lds[2 * ncell + prev];
    Lf.w = lds[3 * ncell + prev];
    Lf.p = lds[4 * ncell + prev];
    Lf.rhoe = FP(0.0);
#pragma unroll
    for (int s = 0; s < NS; ++s) Lf.Y[s] = lds[(6 + s) * ncell + prev];
    int im = i, jm = j, km = k;
    step<D>(im, jm, km, -1);
    const FP cav = FP(0.5) * (J.qa(i, j, k, QC) + J.qa(im, jm, km, QC));
    const int idx = (D == 0) ? i : ((D == 1) ? j : k);
    cmpflx<EOS, D>(Lf, P, cav, idx, J.dlo[D], J.dhi[D], J.bclo[D], J.bchi[D], F);
  }
  // the flux of the face above each cell comes from the next thread along D:
  // the cells [L, H] along D get the transverse correction their two faces
  // make (the transverse stage reads these per-cell values)
  __syncthreads();  // every M read: the buffer now takes the face fluxes
  if (face) {
#pragma unroll
    for (int m = 0; m < 4; ++m) lds[m * ncell + me] = F.f[m];  // no energy flux (store_flux)
#pragma unroll
    for (int s = 0; s < NS; ++s) lds[(5 + s) * ncell + me] = F.f[6 + s];
    lds[(5 + NS) * ncell + me] = F.pgd;
    lds[(6 + NS) * ncell + me] = F.ugd;
  }
  __syncthreads();
  if (!face || pd + 1 >= nd) return;
  const int next = (D == 0) ? me + 1 : me + T.ta;
  Flux<NS, FP> Fh;
#pragma unroll
  for (int m = 0; m < 4; ++m) Fh.f[m] = lds[m * ncell + next];
  Fh.f[4] = FP(0.0);
#pragma unroll
  for (int s = 0; s < NS; ++s) Fh.f[6 + s] = lds[(5 + s) * ncell + next];
  Fh.pgd = lds[(5 + NS) * ncell + next];
  Fh.ugd = lds[(6 + NS) * ncell + next];
  const FP cdtdx = FP(1.0) / FP(3.0) * get_dt(J, ka) / J.dx[D];
  TCorr<NS, FP> C;
  make_tcorr<NS, FP>(F, Fh, FP(J.qa(i, j, k, QGAMC)), cdtdx, C);
  store_tcorr<NS, FP>(J.dl[D], i, j, k, C);
  (void)QT;
}

// Dynamic LDS: max((5 + NS) * pencil_cells, (7 + NS) * BLOCK) elements of FP.
template <class EOS, int D, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_trace_flux1_lds(Sched sc, KArgs ka) {
  extern __shared__ __align__(16) unsigned char pc_trace_dyn[];
  FP* lds = (FP*)pc_trace_dyn;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x) {
      b_trace_flux1_lds<EOS, D, FP>(sc, ka, lb_, lds);
      __syncthreads();  // the next tile restages the buffer
    }
  } else {
    b_trace_flux1_lds<EOS, D, FP>(sc, ka, first_block(sc), lds);
  }
}

}  // namespace pcd
