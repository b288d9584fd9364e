// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Only tests/, the
// smoke check and the bench's cpu_baseline leg may load this code.
//
// Kernel-by-kernel CPU restatement of the reference Godunov (CTU/PPM) path
// behind PeleC::construct_hydro_source -> pc_umdrv -> pc_umeth_3D.  Every
// function names the reference file:line it follows; loop regions are the
// reference's own boxes (e.g. traces on grow(bx,2)), so the oracle does the
// reference's full amount of work, temporaries included.
//
// Compiled with -ffp-contract=off: products and sums round separately, the way
// the reference's expressions read.  amrex::min/max are std::min/std::max,
// restated as mn()/mx() below.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle_eos.hpp"

namespace orc {

// Source/IndexDefines.H:6-37
constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4, UEINT = 5,
              UTEMP = 6, UFS = 7;
constexpr int QRHO = 0, QU = 1, QV = 2, QW = 3, QGAME = 4, QPRES = 5,
              QREINT = 6, QTEMP = 7, QFS = 8;
constexpr int QGAMC = 0, QC = 1, QCSML = 2, QDPDR = 3, QDPDE = 4, QRSPEC = 5;
constexpr int GDRHO = 0, GDU = 1, GDV = 2, GDW = 3, GDPRES = 4, GDGAME = 5;
constexpr int NQAUX = 6, NGDNV = 6;

// Source/PeleC.H:48-58 (PCPhysBCType; amrex::PhysBCType uses the same codes)
constexpr int BC_INTERIOR = 0, BC_INFLOW = 1, BC_OUTFLOW = 2, BC_SYMMETRY = 3,
              BC_SLIPWALL = 4, BC_NOSLIPWALL = 5, BC_USER = 6;

// Source/Constants.H:18-27
constexpr double smallu = 1.0e-12;
constexpr double small_num = 1.0e-8;
constexpr double very_small_num =
    std::numeric_limits<double>::epsilon() * 1e-100;
constexpr double DMIN = std::numeric_limits<double>::min();

inline double mn(double a, double b) { return (b < a) ? b : a; }
inline double mx(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------- boxes/fabs
struct Box {
  int lo[3], hi[3];
  long npts() const {
    return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
  bool ok() const {
    return hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2];
  }
};
inline Box grow(Box b, int n) {
  for (int d = 0; d < 3; ++d) { b.lo[d] -= n; b.hi[d] += n; }
  return b;
}
inline Box grow(Box b, int d, int n) { b.lo[d] -= n; b.hi[d] += n; return b; }
inline Box growHi(Box b, int d, int n) { b.hi[d] += n; return b; }
// nodal in direction d: one more index on the high side
inline Box surroundingNodes(Box b, int d) { b.hi[d] += 1; return b; }

// Array4 restatement: element (i,j,k,n) of a FAB whose index box is
// [begin, end) with i fastest (AMReX layout).
struct A4 {
  double* p = nullptr;
  long js = 0, ks = 0, ns = 0;
  int b[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  int ncomp = 0;
  double& operator()(int i, int j, int k, int n = 0) const {
    return p[(i - b[0]) + (long)(j - b[1]) * js + (long)(k - b[2]) * ks +
             (long)n * ns];
  }
  bool contains(int i, int j, int k) const {
    return i >= b[0] && i < e[0] && j >= b[1] && j < e[1] && k >= b[2] &&
           k < e[2];
  }
  explicit operator bool() const { return p != nullptr; }
};

inline A4 make_a4(double* p, const Box& bx, int nc) {
  A4 a;
  a.p = p;
  for (int d = 0; d < 3; ++d) { a.b[d] = bx.lo[d]; a.e[d] = bx.hi[d] + 1; }
  a.js = bx.hi[0] - bx.lo[0] + 1;
  a.ks = a.js * (bx.hi[1] - bx.lo[1] + 1);
  a.ns = a.ks * (bx.hi[2] - bx.lo[2] + 1);
  a.ncomp = nc;
  return a;
}

// Temporary FAB; filled with NaN so a value read outside what was computed
// shows up in the outputs.
struct Fab {
  std::vector<double> v;
  A4 a;
  Fab(const Box& bx, int nc)
      : v((size_t)bx.npts() * nc, std::numeric_limits<double>::quiet_NaN()) {
    a = make_a4(v.data(), bx, nc);
  }
};

template <class F>
inline void ParallelFor(const Box& bx, F&& f) {
  for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
    for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
      for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) f(i, j, k);
}

struct IV {
  int v[3];
  int& operator[](int d) { return v[d]; }
  int operator[](int d) const { return v[d]; }
};
inline IV ivec(int i, int j, int k) { return IV{{i, j, k}}; }
inline IV shift(IV a, int d, int s) { a.v[d] += s; return a; }
inline double& at(const A4& a, IV iv, int n) { return a(iv[0], iv[1], iv[2], n); }

// =========================================================================
template <class EOS>
struct Hydro {
  static constexpr int NS = EOS::NS;
  static constexpr int NVAR = 7 + NS;  // Source/IndexDefines.H:69
  static constexpr int QVAR = 8 + NS;  // Source/IndexDefines.H:77

  // ------------------------------------------ Source/PeleCUtilities.H:242-320
  static void ctoprim(int i, int j, int k, const A4& u, const A4& q,
                      const A4& qa) {
    const double rho = u(i, j, k, URHO);
    const double rhoinv = 1.0 / rho;
    const double vx = u(i, j, k, UMX) * rhoinv;
    const double vy = u(i, j, k, UMY) * rhoinv;
    const double vz = u(i, j, k, UMZ) * rhoinv;
    const double kineng =
        0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * rho * 0.0 * 0.0 * 0.0 * 0.0;
    q(i, j, k, QRHO) = rho;
    q(i, j, k, QU) = vx;
    q(i, j, k, QV) = vy;
    q(i, j, k, QW) = vz;
    for (int n = 0; n < NS; n++) q(i, j, k, QFS + n) = u(i, j, k, UFS + n) / rho;
    const double e = (u(i, j, k, UEDEN) - kineng) * rhoinv;
    double T = u(i, j, k, UTEMP);
    double Y[NS];
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sp = 0; sp < NS; ++sp) {
      if ((-1e-4 * eps < q(i, j, k, sp + QFS)) &&
          (q(i, j, k, sp + QFS) < 1e-4 * eps)) {
        q(i, j, k, sp + QFS) = 0.0;
      }
      Y[sp] = q(i, j, k, sp + QFS);
    }
    double dpdr_e, dpde, gam1, cs, wbar, p;
    EOS::Y2WBAR(Y, wbar);
    EOS::REY2T(rho, e, Y, T);
    EOS::RTY2P(rho, T, Y, p);
    EOS::RTY2Cs(rho, T, Y, cs);
    EOS::RTY2G(rho, T, Y, gam1);
    EOS::RTY2dpde_dpdre(rho, T, Y, dpde, dpdr_e);
    q(i, j, k, QTEMP) = T;
    q(i, j, k, QREINT) = e * rho;
    q(i, j, k, QPRES) = p;
    q(i, j, k, QGAME) = p / (e * rho) + 1.0;
    qa(i, j, k, QDPDR) = dpdr_e;
    qa(i, j, k, QDPDE) = dpde;
    qa(i, j, k, QGAMC) = gam1;
    qa(i, j, k, QC) = cs;
    qa(i, j, k, QCSML) = mx(small_num, small_num * cs);
    qa(i, j, k, QRSPEC) = RU / wbar;
  }

  // ------------------------------------------------- Source/Hydro.H:14-68
  static void srctoprim(int i, int j, int k, const A4& q, const A4& qa,
                        const A4& src, const A4& srcq) {
    const double srcrho = src(i, j, k, URHO);
    const double rhoinv = 1.0 / q(i, j, k, QRHO);
    const double vx = q(i, j, k, QU);
    const double vy = q(i, j, k, QV);
    const double vz = q(i, j, k, QW);
    const double umx = src(i, j, k, UMX);
    const double umy = src(i, j, k, UMY);
    const double umz = src(i, j, k, UMZ);
    srcq(i, j, k, QRHO) = srcrho;
    srcq(i, j, k, QU) = (umx - vx * srcrho) * rhoinv;
    srcq(i, j, k, QV) = (umy - vy * srcrho) * rhoinv;
    srcq(i, j, k, QW) = (umz - vz * srcrho) * rhoinv;
    srcq(i, j, k, QREINT) = src(i, j, k, UEDEN) - vx * umx - vy * umy -
                            vz * umz + srcrho * (vx * vx + vy * vy + vz * vz);
    srcq(i, j, k, QPRES) =
        qa(i, j, k, QDPDE) *
            (srcq(i, j, k, QREINT) - q(i, j, k, QREINT) * srcrho * rhoinv) *
            rhoinv +
        qa(i, j, k, QDPDR) * srcrho;
    for (int n = 0; n < NS; n++)
      srcq(i, j, k, QFS + n) =
          (src(i, j, k, UFS + n) - q(i, j, k, QFS + n) * srcrho) * rhoinv;
  }

  // ------------------------------------------------- Source/Godunov.H:188-260
  static double flatten(int i, int j, int k, int dir, const A4& q) {
    const IV iv = ivec(i, j, k);
    const IV ivm3 = shift(iv, dir, -3), ivm2 = shift(iv, dir, -2),
             ivm = shift(iv, dir, -1), ivp = shift(iv, dir, 1),
             ivp2 = shift(iv, dir, 2), ivp3 = shift(iv, dir, 3);
    if (!q.contains(ivp3[0], ivp3[1], ivp3[2]) ||
        !q.contains(ivm3[0], ivm3[1], ivm3[2]))
      return 1.0;
    const int n = QPRES;
    const double small_pres = 1.e-200;
    const double shktst = 0.33e0;
    const double zcut1 = 0.75e0;
    const double zcut2 = 0.85e0;
    const double dzcut = 1.0 / (zcut2 - zcut1);
    const double qm = at(q, ivm, n);
    const double qp = at(q, ivp, n);
    const double qmun = at(q, ivm, QU + dir);
    const double qpun = at(q, ivp, QU + dir);
    const double qmm = at(q, ivm2, n);
    const double qpp = at(q, ivp2, n);
    double dp = qp - qm;
    const int ishft = (dp > 0.0) ? 1 : -1;
    double denom = mx(small_pres, std::abs(qpp - qmm));
    double zeta = std::abs(dp) / denom;
    const double z = mn(1.0, mx(0.0, dzcut * (zeta - zcut1)));
    double tst = (qmun - qpun >= 0.0) ? 1.0 : 0.0;
    double tmp = mn(qp, qm);
    const double chi = (std::abs(dp) / tmp) > shktst ? tst : 0.0;
    const double qpshft = at(q, shift(iv, dir, 1 - ishft), n);
    const double qmshft = at(q, shift(iv, dir, -(1 + ishft)), n);
    const double qpunshft = at(q, shift(iv, dir, 1 - ishft), QU + dir);
    const double qmunshft = at(q, shift(iv, dir, -(1 + ishft)), QU + dir);
    const double qppshft = at(q, shift(iv, dir, 2 - ishft), n);
    const double qmmshft = at(q, shift(iv, dir, -(2 + ishft)), n);
    dp = qpshft - qmshft;
    denom = mx(small_pres, std::abs(qppshft - qmmshft));
    zeta = std::abs(dp) / denom;
    const double z2 = mn(1.0, mx(0.0, dzcut * (zeta - zcut1)));
    tst = (qmunshft - qpunshft >= 0.0) ? 1.0 : 0.0;
    tmp = mn(qpshft, qmshft);
    const double chi2 = (std::abs(dp) / tmp) > shktst ? tst : 0.0;
    return 1.0 - mx(chi2 * z2, chi * z);
  }

  // ------------------------------------------------------ Source/PPM.H:28-115
  static void ppm_reconstruct(const double* s, double flatn, double& sm,
                              double& sp) {
    const int im2 = 0, im1 = 1, i0 = 2, ip1 = 3, ip2 = 4;
    double dsl = 2.0 * (s[im1] - s[im2]);
    double dsr = 2.0 * (s[i0] - s[im1]);
    double dsvl_l = 0.0;
    if (dsl * dsr > 0.0) {
      double dsc = 0.5 * (s[i0] - s[im2]);
      dsvl_l = std::copysign(1.0, dsc) *
               mn(std::abs(dsc), mn(std::abs(dsl), std::abs(dsr)));
    }
    dsl = 2.0 * (s[i0] - s[im1]);
    dsr = 2.0 * (s[ip1] - s[i0]);
    double dsvl_r = 0.0;
    if (dsl * dsr > 0.0) {
      double dsc = 0.5 * (s[ip1] - s[im1]);
      dsvl_r = std::copysign(1.0, dsc) *
               mn(std::abs(dsc), mn(std::abs(dsl), std::abs(dsr)));
    }
    sm = 0.5 * (s[i0] + s[im1]) - (1.0 / 6.0) * (dsvl_r - dsvl_l);
    sm = mx(sm, mn(s[i0], s[im1]));
    sm = mn(sm, mx(s[i0], s[im1]));
    dsl = 2.0 * (s[i0] - s[im1]);
    dsr = 2.0 * (s[ip1] - s[i0]);
    dsvl_l = 0.0;
    if (dsl * dsr > 0.0) {
      double dsc = 0.5 * (s[ip1] - s[im1]);
      dsvl_l = std::copysign(1.0, dsc) *
               mn(std::abs(dsc), mn(std::abs(dsl), std::abs(dsr)));
    }
    dsl = 2.0 * (s[ip1] - s[i0]);
    dsr = 2.0 * (s[ip2] - s[ip1]);
    dsvl_r = 0.0;
    if (dsl * dsr > 0.0) {
      double dsc = 0.5 * (s[ip2] - s[i0]);
      dsvl_r = std::copysign(1.0, dsc) *
               mn(std::abs(dsc), mn(std::abs(dsl), std::abs(dsr)));
    }
    sp = 0.5 * (s[ip1] + s[i0]) - (1.0 / 6.0) * (dsvl_r - dsvl_l);
    sp = mx(sp, mn(s[ip1], s[i0]));
    sp = mn(sp, mx(s[ip1], s[i0]));
    sm = flatn * sm + (1.0 - flatn) * s[i0];
    sp = flatn * sp + (1.0 - flatn) * s[i0];
    if ((sp - s[i0]) * (s[i0] - sm) <= 0.0) {
      sp = s[i0];
      sm = s[i0];
    } else if (std::abs(sp - s[i0]) >= 2.0 * std::abs(sm - s[i0])) {
      sp = 3.0 * s[i0] - 2.0 * sm;
    } else if (std::abs(sm - s[i0]) >= 2.0 * std::abs(sp - s[i0])) {
      sm = 3.0 * s[i0] - 2.0 * sp;
    }
  }

  // ----------------------------------------------------- Source/PPM.H:131-189
  static void ppm_int_profile(double sm, double sp, double sc, double u,
                              double c, double dtdx, double* Ip, double* Im) {
    const double s6 = 6.0 * sc - 3.0 * (sm + sp);
    const double spd[3] = {u - c, u, u + c};
    for (int w = 0; w < 3; ++w) {
      const double speed = spd[w];
      const double sigma = std::abs(speed) * dtdx;
      if (speed <= 0.0) {
        Ip[w] = sp;
        Im[w] = sm + 0.5 * sigma * (sp - sm + (1.0 - (2.0 / 3.0) * sigma) * s6);
      } else {
        Ip[w] = sp - 0.5 * sigma * (sp - sm - (1.0 - (2.0 / 3.0) * sigma) * s6);
        Im[w] = sm;
      }
    }
  }

  // ------------------------------------------------------ Source/PPM.cpp:6-424
  // (use_hybrid_weno = false: the WENO branches are out of scope)
  static void trace_ppm(const Box& bx, int idir, const A4& q, const A4& qm,
                        const A4& qp, const Box& vbx, double dt,
                        const double* dx, bool use_flattening) {
    const double dtdx = dt / dx[idir];
    const int* vlo = vbx.lo;
    const int* vhi = vbx.hi;
    int QUN = 0, QUT = 0, QUTT = 0;
    if (idir == 0) { QUN = QU; QUT = QV; QUTT = QW; }
    else if (idir == 1) { QUN = QV; QUT = QW; QUTT = QU; }
    else { QUN = QW; QUT = QU; QUTT = QV; }
    ParallelFor(bx, [&](int i, int j, int k) {
      const IV iv = ivec(i, j, k);
      const IV ivm2 = shift(iv, idir, -2), ivm1 = shift(iv, idir, -1),
               ivp1 = shift(iv, idir, 1), ivp2 = shift(iv, idir, 2);
      double Y[NS];
      for (int s = 0; s < NS; ++s) Y[s] = at(q, iv, s + QFS);
      double cc = 0;
      EOS::RPY2Cs(at(q, iv, QRHO), at(q, iv, QPRES), Y, cc);
      const double un = at(q, iv, QUN);
      double flat = 1.0;
      if (use_flattening)
        for (int df = 0; df < 3; df++) flat = mn(flat, flatten(i, j, k, df, q));
      double Ip[QVAR][3], Im[QVAR][3];
      for (int n = 0; n < QVAR; n++) {
        double s[5];
        s[0] = at(q, ivm2, n);
        s[1] = at(q, ivm1, n);
        s[2] = at(q, iv, n);
        s[3] = at(q, ivp1, n);
        s[4] = at(q, ivp2, n);
        double sm, sp;
        ppm_reconstruct(s, flat, sm, sp);
        ppm_int_profile(sm, sp, s[2], un, cc, dtdx, Ip[n], Im[n]);
      }
      const bool doP = (idir == 0 && i >= vlo[0]) || (idir == 1 && j >= vlo[1]) ||
                       (idir == 2 && k >= vlo[2]);
      const bool doM = (idir == 0 && i <= vhi[0]) || (idir == 1 && j <= vhi[1]) ||
                       (idir == 2 && k <= vhi[2]);
      for (int n = QFS; n < NS + QFS; n++) {
        if (doP) at(qp, iv, n) = Im[n][1];
        if (doM) at(qm, ivp1, n) = Ip[n][1];
      }
      if (doP) {
        double rho_ref = Im[QRHO][0];
        const double un_ref = Im[QUN][0];
        double p_ref = Im[QPRES][0];
        const double rhoe_g_ref = Im[QREINT][0];
        rho_ref = mx(rho_ref, DMIN);
        const double rho_ref_inv = 1.0 / rho_ref;
        p_ref = mx(p_ref, DMIN);
        double Yr[NS];
        for (int s = 0; s < NS; ++s) Yr[s] = Im[s + QFS][0];
        double cc_ref = 0;
        EOS::RPY2Cs(rho_ref, p_ref, Yr, cc_ref);
        const double csq_ref = cc_ref * cc_ref;
        const double cc_ref_inv = 1.0 / cc_ref;
        const double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv;
        const double dum = un_ref - Im[QUN][0];
        const double dptotm = p_ref - Im[QPRES][0];
        const double drho = rho_ref - Im[QRHO][1];
        const double dptot = p_ref - Im[QPRES][1];
        const double drhoe_g = rhoe_g_ref - Im[QREINT][1];
        const double dup = un_ref - Im[QUN][2];
        const double dptotp = p_ref - Im[QPRES][2];
        double alpham =
            0.5 * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
        double alphap =
            0.5 * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
        double alpha0r = drho - dptot / csq_ref;
        double alpha0e_g = drhoe_g - dptot * h_g_ref / csq_ref;
        alpham = un - cc > 0.0 ? 0.0 : -alpham;
        alphap = un + cc > 0.0 ? 0.0 : -alphap;
        alpha0r = un > 0.0 ? 0.0 : -alpha0r;
        alpha0e_g = un > 0.0 ? 0.0 : -alpha0e_g;
        (void)alpha0e_g;
        at(qp, iv, QRHO) = mx(DMIN, rho_ref + alphap + alpham + alpha0r);
        at(qp, iv, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
        at(qp, iv, QPRES) = mx(DMIN, p_ref + (alphap + alpham) * csq_ref);
        at(qp, iv, QUT) = Im[QUT][1];
        at(qp, iv, QUTT) = Im[QUTT][1];
        double eint = 0, Yp[NS];
        for (int s = 0; s < NS; ++s) Yp[s] = at(qp, iv, s + QFS);
        EOS::RYP2E(at(qp, iv, QRHO), Yp, at(qp, iv, QPRES), eint);
        at(qp, iv, QREINT) = at(qp, iv, QRHO) * eint;
      }
      if (doM) {
        double rho_ref = Ip[QRHO][2];
        const double un_ref = Ip[QUN][2];
        double p_ref = Ip[QPRES][2];
        const double rhoe_g_ref = Ip[QREINT][2];
        rho_ref = mx(rho_ref, DMIN);
        const double rho_ref_inv = 1.0 / rho_ref;
        p_ref = mx(p_ref, DMIN);
        double Yr[NS];
        for (int s = 0; s < NS; ++s) Yr[s] = Ip[s + QFS][2];
        double cc_ref = 0;
        EOS::RPY2Cs(rho_ref, p_ref, Yr, cc_ref);
        const double csq_ref = cc_ref * cc_ref;
        const double cc_ref_inv = 1.0 / cc_ref;
        const double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv;
        const double dum = un_ref - Ip[QUN][0];
        const double dptotm = p_ref - Ip[QPRES][0];
        const double drho = rho_ref - Ip[QRHO][1];
        const double dptot = p_ref - Ip[QPRES][1];
        const double drhoe_g = rhoe_g_ref - Ip[QREINT][1];
        const double dup = un_ref - Ip[QUN][2];
        const double dptotp = p_ref - Ip[QPRES][2];
        double alpham =
            0.5 * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
        double alphap =
            0.5 * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
        double alpha0r = drho - dptot / csq_ref;
        double alpha0e_g = drhoe_g - dptot * h_g_ref / csq_ref;
        alpham = un - cc > 0.0 ? -alpham : 0.0;
        alphap = un + cc > 0.0 ? -alphap : 0.0;
        alpha0r = un > 0.0 ? -alpha0r : 0.0;
        alpha0e_g = un > 0.0 ? -alpha0e_g : 0.0;
        (void)alpha0e_g;
        at(qm, ivp1, QRHO) = mx(DMIN, rho_ref + alphap + alpham + alpha0r);
        at(qm, ivp1, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
        at(qm, ivp1, QPRES) = mx(DMIN, p_ref + (alphap + alpham) * csq_ref);
        at(qm, ivp1, QUT) = Ip[QUT][1];
        at(qm, ivp1, QUTT) = Ip[QUTT][1];
        double eint = 0, Ym[NS];
        for (int s = 0; s < NS; ++s) Ym[s] = at(qm, ivp1, s + QFS);
        EOS::RYP2E(at(qm, ivp1, QRHO), Ym, at(qm, ivp1, QPRES), eint);
        at(qm, ivp1, QREINT) = at(qm, ivp1, QRHO) * eint;
      }
    });
  }

  // ------------------------------------------------------ Source/PLM.H:21-77
  static double plm_slope(int i, int j, int k, int n, int dir, const A4& q,
                          double flat, int order) {
    if (order == 1) return 0.0;
    const IV iv = ivec(i, j, k);
    double dlft, drgt, dcen, dfm = 0.0, dfp = 0.0, dlim, dsgn, dtemp;
    const double qc = at(q, iv, n);
    const double qm2 = at(q, shift(iv, dir, -2), n);
    const double qm = at(q, shift(iv, dir, -1), n);
    const double qp = at(q, shift(iv, dir, 1), n);
    const double qp2 = at(q, shift(iv, dir, 2), n);
    if (order == 4) {
      dlft = qm - qm2;
      drgt = qc - qm;
      dcen = 0.5 * (dlft + drgt);
      dsgn = std::copysign(1.0, dcen);
      dlim = (dlft * drgt >= 0.0) ? 2.0 * mn(std::abs(dlft), std::abs(drgt)) : 0.0;
      dfm = dsgn * mn(dlim, std::abs(dcen));
      dlft = qp - qc;
      drgt = qp2 - qp;
      dcen = 0.5 * (dlft + drgt);
      dsgn = std::copysign(1.0, dcen);
      dlim = (dlft * drgt >= 0.0) ? 2.0 * mn(std::abs(dlft), std::abs(drgt)) : 0.0;
      dfp = dsgn * mn(dlim, std::abs(dcen));
    }
    dlft = qc - qm;
    drgt = qp - qc;
    dcen = 0.5 * (dlft + drgt);
    dsgn = std::copysign(1.0, dcen);
    dlim = (dlft * drgt >= 0.0) ? 2.0 * mn(std::abs(dlft), std::abs(drgt)) : 0.0;
    dtemp = (order == 2) ? dcen : 4.0 / 3.0 * dcen - 1.0 / 6.0 * (dfp + dfm);
    return flat * dsgn * mn(dlim, std::abs(dtemp));
  }

  // ----------------------------------------------------- Source/PLM.H:167-309
  static void plm_d(int i, int j, int k, int dir, const A4& qm, const A4& qp,
                    const double* slope, const A4& q, double cc, double dx,
                    double dt) {
    const IV iv = ivec(i, j, k);
    const IV ivp = shift(iv, dir, 1);
    const int l_idx[3] = {dir == 0 ? 0 : (dir == 1 ? 1 : 2),
                          dir == 0 ? 1 : 0,
                          dir == 0 ? 2 : (dir == 1 ? 2 : 1)};
    const double dtdx = dt / dx;
    const double cs = cc * cc;
    const double rho = at(q, iv, QRHO);
    const double vel[3] = {at(q, iv, QU), at(q, iv, QV), at(q, iv, QW)};
    const double p = at(q, iv, QPRES);
    const double rhoe = at(q, iv, QREINT);
    const double enth = ((rhoe + p) / rho) / cs;
    const double drho = slope[QRHO];
    const double dvel[3] = {slope[QU], slope[QV], slope[QW]};
    const double dp = slope[QPRES];
    const double drhoe = slope[QREINT];
    const double alpham = 0.5 * (dp / (rho * cc) - dvel[dir]) * rho / cc;
    const double alphap = 0.5 * (dp / (rho * cc) + dvel[dir]) * rho / cc;
    const double alpha0r = drho - dp / cs;
    const double alpha0e = drhoe - dp * enth;
    const double alpha0v = dvel[l_idx[1]];
    const double alpha0w = dvel[l_idx[2]];
    const double wv[3] = {vel[dir] - cc, vel[dir], vel[dir] + cc};

    double rho_ref = rho - 0.5 * (1.0 + dtdx * mn(wv[0], 0.0)) * drho;
    double vel_ref[3];
    for (int dim = 0; dim < 3; dim++)
      vel_ref[dim] = vel[dim] - 0.5 * (1.0 + dtdx * mn(wv[0], 0.0)) * dvel[dim];
    double p_ref = p - 0.5 * (1.0 + dtdx * mn(wv[0], 0.0)) * dp;
    double rhoe_ref = rhoe - 0.5 * (1.0 + dtdx * mn(wv[0], 0.0)) * drhoe;
    const double apright =
        0.25 * dtdx * (wv[0] - wv[2]) * (1.0 - std::copysign(1.0, wv[2])) * alphap;
    const double amright = 0.0;
    const double azrright =
        0.25 * dtdx * (wv[0] - wv[1]) * (1.0 - std::copysign(1.0, wv[1])) * alpha0r;
    const double azeright =
        0.25 * dtdx * (wv[0] - wv[1]) * (1.0 - std::copysign(1.0, wv[1])) * alpha0e;
    const double azv1rght =
        0.25 * dtdx * (wv[0] - wv[1]) * (1.0 - std::copysign(1.0, wv[1])) * alpha0v;
    const double azw1rght =
        0.25 * dtdx * (wv[0] - wv[1]) * (1.0 - std::copysign(1.0, wv[1])) * alpha0w;
    at(qp, iv, QRHO) = rho_ref + apright + amright + azrright;
    at(qp, iv, QRHO) = mx(DMIN, at(qp, iv, QRHO));
    at(qp, iv, QU + l_idx[0]) = vel_ref[l_idx[0]] + (apright - amright) * cc / rho;
    at(qp, iv, QU + l_idx[1]) = vel_ref[l_idx[1]] + azv1rght;
    at(qp, iv, QU + l_idx[2]) = vel_ref[l_idx[2]] + azw1rght;
    at(qp, iv, QPRES) = p_ref + (apright + amright) * cs;
    at(qp, iv, QPRES) = mx(at(qp, iv, QPRES), DMIN);
    at(qp, iv, QREINT) = rhoe_ref + (apright + amright) * enth * cs + azeright;

    rho_ref = rho + 0.5 * (1.0 - dtdx * mx(wv[2], 0.0)) * drho;
    for (int dim = 0; dim < 3; dim++)
      vel_ref[dim] = vel[dim] + 0.5 * (1.0 - dtdx * mx(wv[2], 0.0)) * dvel[dim];
    p_ref = p + 0.5 * (1.0 - dtdx * mx(wv[2], 0.0)) * dp;
    rhoe_ref = rhoe + 0.5 * (1.0 - dtdx * mx(wv[2], 0.0)) * drhoe;
    const double apleft = 0.0;
    const double amleft =
        0.25 * dtdx * (wv[2] - wv[0]) * (1.0 + std::copysign(1.0, wv[0])) * alpham;
    const double azrleft =
        0.25 * dtdx * (wv[2] - wv[1]) * (1.0 + std::copysign(1.0, wv[1])) * alpha0r;
    const double azeleft =
        0.25 * dtdx * (wv[2] - wv[1]) * (1.0 + std::copysign(1.0, wv[1])) * alpha0e;
    const double azv1left =
        0.25 * dtdx * (wv[2] - wv[1]) * (1.0 + std::copysign(1.0, wv[1])) * alpha0v;
    const double azw1left =
        0.25 * dtdx * (wv[2] - wv[1]) * (1.0 + std::copysign(1.0, wv[1])) * alpha0w;
    at(qm, ivp, QRHO) = rho_ref + apleft + amleft + azrleft;
    at(qm, ivp, QRHO) = mx(at(qm, ivp, QRHO), DMIN);
    at(qm, ivp, QU + l_idx[0]) = vel_ref[l_idx[0]] + (apleft - amleft) * cc / rho;
    at(qm, ivp, QU + l_idx[1]) = vel_ref[l_idx[1]] + azv1left;
    at(qm, ivp, QU + l_idx[2]) = vel_ref[l_idx[2]] + azw1left;
    at(qm, ivp, QPRES) = p_ref + (apleft + amleft) * cs;
    at(qm, ivp, QPRES) = mx(at(qm, ivp, QPRES), DMIN);
    at(qm, ivp, QREINT) = rhoe_ref + (apleft + amleft) * enth * cs + azeleft;

    const double vel_ad = at(q, iv, QU + l_idx[0]);
    const double spzerom = (vel_ad > 0) ? -1.0 : vel_ad * dtdx;
    const double spzerop = (vel_ad >= 0) ? vel_ad * dtdx : 1.0;
    for (int s = 0; s < NS; s++) {
      const int n = QFS + s;  // Source/PLM.H:120-136
      const double acmprght = 0.5 * (-1.0 - spzerom) * slope[n];
      at(qp, iv, n) = at(q, iv, n) + acmprght;
      const double acmpleft = 0.5 * (1.0 - spzerop) * slope[n];
      at(qm, ivp, n) = at(q, iv, n) + acmpleft;
    }
  }

  // ---------------------------------------------------- Source/Riemann.H:9-200
  static void riemann(double rl, double ul, double vl, double v2l, double pl,
                      const double* spl, double rr, double ur, double vr,
                      double v2r, double pr, const double* spr, int bc_test_val,
                      double cav, double& ustar, double& uflx_rho,
                      double* uflx_rhoY, double& uflx_u, double& uflx_v,
                      double& uflx_w, double& uflx_eden, double& uflx_eint,
                      double& qint_iu, double& qint_iv1, double& qint_iv2,
                      double& qint_gdpres, double& qint_gdgame) {
    const double wsmall = DMIN;
    double gmf[NS];
    for (int n = 0; n < NS; n++) gmf[n] = spl[n];
    double cl = 0.0;
    EOS::RPY2Cs(rl, pl, gmf, cl);
    for (int n = 0; n < NS; n++) gmf[n] = spr[n];
    double cr = 0.0;
    EOS::RPY2Cs(rr, pr, gmf, cr);
    const double wl = mx(wsmall, cl * rl);
    const double wr = mx(wsmall, cr * rr);
    const double pstar =
        mx(DMIN, ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr));
    ustar = ((wl * ul + wr * ur) + (pl - pr)) / (wl + wr);
    bool mask = ustar > 0.0;
    double ro = 0.0;
    double rspo[NS];
    for (int n = 0; n < NS; n++) {
      rspo[n] = mask ? rl * spl[n] : rr * spr[n];
      ro += rspo[n];
    }
    double uo = mask ? ul : ur;
    double po = mask ? pl : pr;
    mask = std::abs(ustar) < smallu * 0.5 * (std::abs(ul) + std::abs(ur)) ||
           ustar == 0.0;
    ustar = mask ? 0.0 : ustar;
    ro = 0.0;
    for (int n = 0; n < NS; n++) {
      rspo[n] = mask ? 0.5 * (rl * spl[n] + rr * spr[n]) : rspo[n];
      ro += rspo[n];
    }
    uo = mask ? 0.5 * (ul + ur) : uo;
    po = mask ? 0.5 * (pl + pr) : po;
    double g_rho = ro, g_p = po;
    for (int n = 0; n < NS; n++) gmf[n] = rspo[n] / ro;
    double g_e;
    EOS::RYP2E(g_rho, gmf, g_p, g_e);
    double co;
    EOS::RPY2Cs(g_rho, g_p, gmf, co);
    const double drho = (pstar - po) / (co * co);
    double rstar = 0.0;
    double rspstar[NS];
    for (int n = 0; n < NS; n++) {
      const double spon = rspo[n] / ro;
      rspstar[n] = mx(0.0, rspo[n] + drho * spon);
      rstar += rspstar[n];
    }
    g_rho = rstar;
    g_p = pstar;
    for (int n = 0; n < NS; n++) gmf[n] = rspstar[n] / rstar;
    EOS::RYP2E(g_rho, gmf, g_p, g_e);
    double cstar;
    EOS::RPY2Cs(g_rho, g_p, gmf, cstar);
    const double sgnm = std::copysign(1.0, ustar);
    double spout = co - sgnm * uo;
    double spin = cstar - sgnm * ustar;
    const double ushock = 0.5 * (spin + spout);
    mask = pstar < po;
    spout = mask ? spout : ushock;
    spin = mask ? spin : ushock;
    const double scr = (std::abs(spout - spin) < very_small_num)
                           ? small_num * cav
                           : spout - spin;
    const double frac = mx(0.0, mn(1.0, (1.0 + (spout + spin) / scr) * 0.5));
    mask = ustar > 0.0;
    qint_iv1 = mask ? vl : vr;
    qint_iv2 = mask ? v2l : v2r;
    mask = (ustar == 0.0);
    qint_iv1 = mask ? 0.5 * (vl + vr) : qint_iv1;
    qint_iv2 = mask ? 0.5 * (v2l + v2r) : qint_iv2;
    double rgd = 0.0;
    double rspgd[NS];
    for (int n = 0; n < NS; n++) {
      rspgd[n] = frac * rspstar[n] + (1.0 - frac) * rspo[n];
      rgd += rspgd[n];
    }
    qint_iu = frac * ustar + (1.0 - frac) * uo;
    qint_gdpres = frac * pstar + (1.0 - frac) * po;
    g_rho = rgd;
    g_p = qint_gdpres;
    for (int n = 0; n < NS; n++) gmf[n] = rspgd[n] / rgd;
    EOS::RYP2E(g_rho, gmf, g_p, g_e);
    mask = (spout < 0.0);
    rgd = 0.0;
    for (int n = 0; n < NS; n++) {
      rspgd[n] = mask ? rspo[n] : rspgd[n];
      rgd += rspgd[n];
    }
    qint_iu = mask ? uo : qint_iu;
    qint_gdpres = mask ? po : qint_gdpres;
    mask = (spin >= 0.0);
    rgd = 0.0;
    for (int n = 0; n < NS; n++) {
      rspgd[n] = mask ? rspstar[n] : rspgd[n];
      rgd += rspgd[n];
    }
    qint_iu = mask ? ustar : qint_iu;
    qint_gdpres = mask ? pstar : qint_gdpres;
    g_rho = rgd;
    g_p = qint_gdpres;
    for (int n = 0; n < NS; n++) gmf[n] = rspgd[n] / rgd;
    EOS::RYP2E(g_rho, gmf, g_p, g_e);
    const double regd = g_rho * g_e;
    qint_gdgame = qint_gdpres / regd + 1.0;
    qint_iu = bc_test_val * qint_iu;
    uflx_rho = rgd * qint_iu;
    for (int n = 0; n < NS; n++) uflx_rhoY[n] = rspgd[n] * qint_iu;
    uflx_u = uflx_rho * qint_iu + qint_gdpres;
    uflx_v = uflx_rho * qint_iv1;
    uflx_w = uflx_rho * qint_iv2;
    const double rhoetot =
        regd + 0.5 * rgd * (qint_iu * qint_iu + qint_iv1 * qint_iv1 +
                            qint_iv2 * qint_iv2) -
        0.5 * rgd * 0.0 * 0.0 * 0.0 * 0.0;
    uflx_eden = qint_iu * (rhoetot + qint_gdpres);
    uflx_eint = qint_iu * regd;
  }

  // ------------------------------------------------- Source/Godunov.H:303-449
  static void cmpflx(int i, int j, int k, int bclo, int bchi, int domlo,
                     int domhi, const A4& ql, const A4& qr, const A4& flx,
                     const A4& q, const A4& qa, int dir) {
    double cav, ustar;
    double spl[NS], spr[NS];
    int IU, IV, IV2, GU, GV, GV2, f_idx[3];
    if (dir == 0) {
      IU = QU; IV = QV; IV2 = QW; GU = GDU; GV = GDV; GV2 = GDW;
      cav = 0.5 * (qa(i, j, k, QC) + qa(i - 1, j, k, QC));
      f_idx[0] = UMX; f_idx[1] = UMY; f_idx[2] = UMZ;
    } else if (dir == 1) {
      IU = QV; IV = QU; IV2 = QW; GU = GDV; GV = GDU; GV2 = GDW;
      cav = 0.5 * (qa(i, j, k, QC) + qa(i, j - 1, k, QC));
      f_idx[0] = UMY; f_idx[1] = UMX; f_idx[2] = UMZ;
    } else {
      IU = QW; IV = QU; IV2 = QV; GU = GDW; GV = GDU; GV2 = GDV;
      cav = 0.5 * (qa(i, j, k, QC) + qa(i, j, k - 1, QC));
      f_idx[0] = UMZ; f_idx[1] = UMX; f_idx[2] = UMY;
    }
    for (int sp = 0; sp < NS; ++sp) {
      spl[sp] = ql(i, j, k, QFS + sp);
      spr[sp] = qr(i, j, k, QFS + sp);
    }
    double ul = ql(i, j, k, IU), vl = ql(i, j, k, IV), v2l = ql(i, j, k, IV2);
    double pl = ql(i, j, k, QPRES), rhol = ql(i, j, k, QRHO);
    double ur = qr(i, j, k, IU), vr = qr(i, j, k, IV), v2r = qr(i, j, k, IV2);
    double pr = qr(i, j, k, QPRES), rhor = qr(i, j, k, QRHO);
    const int idx = (dir == 2) ? k : ((dir == 0) ? i : j);
    if (idx == domlo) {
      if (bclo == BC_NOSLIPWALL || bclo == BC_SLIPWALL || bclo == BC_SYMMETRY) {
        ul = -ur; vl = vr; v2l = v2r; pl = pr; rhol = rhor;
      } else if (bclo == BC_OUTFLOW) {
        ul = ur; vl = vr; v2l = v2r; pl = pr; rhol = rhor;
      }
    } else if (idx == domhi + 1) {
      if (bchi == BC_NOSLIPWALL || bchi == BC_SLIPWALL || bchi == BC_SYMMETRY) {
        ur = -ul; vr = vl; v2r = v2l; pr = pl; rhor = rhol;
      } else if (bchi == BC_OUTFLOW) {
        ur = ul; vr = vl; v2r = v2l; pr = pl; rhor = rhol;
      }
    }
    double dummy[NS];
    riemann(rhol, ul, vl, v2l, pl, spl, rhor, ur, vr, v2r, pr, spr, 1, cav,
            ustar, flx(i, j, k, URHO), dummy, flx(i, j, k, f_idx[0]),
            flx(i, j, k, f_idx[1]), flx(i, j, k, f_idx[2]), flx(i, j, k, UEDEN),
            flx(i, j, k, UEINT), q(i, j, k, GU), q(i, j, k, GV),
            q(i, j, k, GV2), q(i, j, k, GDPRES), q(i, j, k, GDGAME));
    const double flxrho = flx(i, j, k, URHO);
    for (int n = 0; n < NS; n++) {  // Source/Godunov.H:15-25
      const double qlv = ql(i, j, k, QFS + n), qrv = qr(i, j, k, QFS + n);
      flx(i, j, k, UFS + n) = (ustar > 0.0)   ? flxrho * qlv
                              : (ustar < 0.0) ? flxrho * qrv
                                              : flxrho * 0.5 * (qlv + qrv);
    }
  }

  // ------------------------------------------------- Source/Godunov.H:455-633
  static void transdo(int i, int j, int k, int dir, int other_dir, const A4& qm,
                      const A4& qp, const A4& qnormm, const A4& qnormp,
                      const A4& flxx, const A4& qa, const A4& qint,
                      double cdtdx) {
    const IV iv = ivec(i, j, k);
    const IV ivpn = shift(iv, dir, 1);
    const IV ivpt = shift(iv, other_dir, 1);
    const int qvidx = (other_dir == 0) ? GDU : ((other_dir == 1) ? GDV : GDW);
    const double flxrho = cdtdx * (at(flxx, ivpt, URHO) - at(flxx, iv, URHO));
    const double flxu = cdtdx * (at(flxx, ivpt, UMX) - at(flxx, iv, UMX));
    const double flxv = cdtdx * (at(flxx, ivpt, UMY) - at(flxx, iv, UMY));
    const double flxw = cdtdx * (at(flxx, ivpt, UMZ) - at(flxx, iv, UMZ));
    const double flxe = cdtdx * (at(flxx, ivpt, UEDEN) - at(flxx, iv, UEDEN));
    const double c = at(qa, iv, QGAMC);
    for (int n = 0; n < NS; n++) {  // Source/Godunov.H:30-73
      const int uc = UFS + n, qc = QFS + n;
      const double compn = cdtdx * (at(flxx, ivpt, uc) - at(flxx, iv, uc));
      double rr = at(qnormp, iv, QRHO);
      double rrnew = rr - flxrho;
      double compo = rr * at(qnormp, iv, qc) - compn;
      at(qp, iv, qc) = compo / rrnew;
      rr = at(qnormm, ivpn, QRHO);
      rrnew = rr - flxrho;
      compo = rr * at(qnormm, ivpn, qc) - compn;
      at(qm, ivpn, qc) = compo / rrnew;
    }
    const double pggp = at(qint, ivpt, GDPRES);
    const double pggm = at(qint, iv, GDPRES);
    const double ugp = at(qint, ivpt, qvidx);
    const double ugm = at(qint, iv, qvidx);
    const double dAup = pggp * ugp - pggm * ugm;
    const double pav = 0.5 * (pggp + pggm);
    const double dAu = ugp - ugm;
    {
      const double rrr = at(qnormp, iv, QRHO);
      double rur = at(qnormp, iv, QU);
      double rvr = at(qnormp, iv, QV);
      double rwr = at(qnormp, iv, QW);
      const double ekinr = 0.5 * rrr * (rur * rur + rvr * rvr + rwr * rwr);
      rur *= rrr;
      rvr *= rrr;
      rwr *= rrr;
      const double rer = at(qnormp, iv, QREINT) + ekinr;
      const double rrnewr = rrr - flxrho;
      const double runewr = rur - flxu;
      const double rvnewr = rvr - flxv;
      const double rwnewr = rwr - flxw;
      const double renewr = rer - flxe;
      at(qp, iv, QRHO) = rrnewr;
      at(qp, iv, QU) = runewr / rrnewr;
      at(qp, iv, QV) = rvnewr / rrnewr;
      at(qp, iv, QW) = rwnewr / rrnewr;
      const double rhoekinr =
          0.5 * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
      at(qp, iv, QREINT) = renewr - rhoekinr;
      const double pnewr =
          at(qnormp, iv, QPRES) - cdtdx * (dAup + pav * dAu * (c - 1.0));
      at(qp, iv, QPRES) = mx(pnewr, very_small_num);
    }
    {
      const double rrl = at(qnormm, ivpn, QRHO);
      double rul = at(qnormm, ivpn, QU);
      double rvl = at(qnormm, ivpn, QV);
      double rwl = at(qnormm, ivpn, QW);
      const double ekinl = 0.5 * rrl * (rul * rul + rvl * rvl + rwl * rwl);
      rul *= rrl;
      rvl *= rrl;
      rwl *= rrl;
      const double rel = at(qnormm, ivpn, QREINT) + ekinl;
      const double rrnewl = rrl - flxrho;
      const double runewl = rul - flxu;
      const double rvnewl = rvl - flxv;
      const double rwnewl = rwl - flxw;
      const double renewl = rel - flxe;
      at(qm, ivpn, QRHO) = rrnewl;
      at(qm, ivpn, QU) = runewl / rrnewl;
      at(qm, ivpn, QV) = rvnewl / rrnewl;
      at(qm, ivpn, QW) = rwnewl / rrnewl;
      const double rhoekinl =
          0.5 * (runewl * runewl + rvnewl * rvnewl + rwnewl * rwnewl) / rrnewl;
      at(qm, ivpn, QREINT) = renewl - rhoekinl;
      const double pnewl =
          at(qnormm, ivpn, QPRES) - cdtdx * (dAup + pav * dAu * (c - 1.0));
      at(qm, ivpn, QPRES) = mx(pnewl, very_small_num);
    }
  }

  // ------------------------------------------------- Source/Godunov.H:639-845
  static void transdd(int i, int j, int k, int dir, const A4& qm, const A4& qp,
                      const A4& qnormm, const A4& qnormp, const A4& flxx,
                      const A4& flxy, const A4& qx, const A4& qy, const A4& qa,
                      const A4& srcq, double hdt, double cdtdx0,
                      double cdtdx1) {
    const IV iv = ivec(i, j, k);
    const int qvidx0 = (dir == 0) ? GDV : GDU;
    const int qvidx1 = (dir == 2) ? GDV : GDW;
    const int l1 = (dir == 0) ? 1 : 0;
    const int l2 = (dir == 2) ? 1 : 2;
    const IV ivpt0 = shift(iv, l1, 1);
    const IV ivpt1 = shift(iv, l2, 1);
    const IV ivpn = shift(iv, dir, 1);
    auto dfl = [&](int n) {
      return cdtdx0 * (at(flxx, ivpt0, n) - at(flxx, iv, n)) +
             cdtdx1 * (at(flxy, ivpt1, n) - at(flxy, iv, n));
    };
    const double flxrho = dfl(URHO);
    const double flxu = dfl(UMX);
    const double flxv = dfl(UMY);
    const double flxw = dfl(UMZ);
    const double flxe = dfl(UEDEN);
    const double c = at(qa, iv, QGAMC);
    const double srcrho = at(srcq, iv, QRHO);
    const double srcu = at(srcq, iv, QU);
    const double srcv = at(srcq, iv, QV);
    const double srcw = at(srcq, iv, QW);
    const double srce = at(srcq, iv, QREINT);
    const double srcp = at(srcq, iv, QPRES);
    const double rrr = at(qnormp, iv, QRHO);
    const double rrl = at(qnormm, ivpn, QRHO);
    const double rrnewl = rrl - flxrho;
    const double rrnewr = rrr - flxrho;
    for (int n = 0; n < NS; n++) {  // Source/Godunov.H:78-132
      const int uc = UFS + n, qc = QFS + n;
      const double srcpass = at(srcq, iv, qc);
      const double compn = dfl(uc);
      double compo = rrr * at(qnormp, iv, qc) - compn;
      at(qp, iv, qc) = compo / rrnewr + hdt * srcpass;
      compo = rrl * at(qnormm, ivpn, qc) - compn;
      at(qm, ivpn, qc) = compo / rrnewl + hdt * srcpass;
    }
    const double pggpx = at(qx, ivpt0, GDPRES);
    const double pggmx = at(qx, iv, GDPRES);
    const double ugpx = at(qx, ivpt0, qvidx0);
    const double ugmx = at(qx, iv, qvidx0);
    const double dAupx = pggpx * ugpx - pggmx * ugmx;
    const double pavx = 0.5 * (pggpx + pggmx);
    const double dAux = ugpx - ugmx;
    const double pggpy = at(qy, ivpt1, GDPRES);
    const double pggmy = at(qy, iv, GDPRES);
    const double ugpy = at(qy, ivpt1, qvidx1);
    const double ugmy = at(qy, iv, qvidx1);
    const double dAupy = pggpy * ugpy - pggmy * ugmy;
    const double pavy = 0.5 * (pggpy + pggmy);
    const double dAuy = ugpy - ugmy;
    const double pxnew = cdtdx0 * (dAupx + pavx * dAux * (c - 1.0));
    const double pynew = cdtdx1 * (dAupy + pavy * dAuy * (c - 1.0));
    {
      const double rur = rrr * at(qnormp, iv, QU);
      const double rvr = rrr * at(qnormp, iv, QV);
      const double rwr = rrr * at(qnormp, iv, QW);
      const double ekinr = 0.5 * (rur * rur + rvr * rvr + rwr * rwr) / rrr;
      const double rer = at(qnormp, iv, QREINT) + ekinr;
      const double runewr = rur - flxu;
      const double rvnewr = rvr - flxv;
      const double rwnewr = rwr - flxw;
      const double renewr = rer - flxe;
      at(qp, iv, QRHO) = rrnewr + hdt * srcrho;
      at(qp, iv, QU) = runewr / rrnewr + hdt * srcu;
      at(qp, iv, QV) = rvnewr / rrnewr + hdt * srcv;
      at(qp, iv, QW) = rwnewr / rrnewr + hdt * srcw;
      const double rhoekinr =
          0.5 * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
      at(qp, iv, QREINT) = renewr - rhoekinr + hdt * srce;
      at(qp, iv, QPRES) = at(qnormp, iv, QPRES) - pxnew - pynew + hdt * srcp;
      at(qp, iv, QPRES) = mx(at(qp, iv, QPRES), DMIN);
    }
    {
      const double rul = rrl * at(qnormm, ivpn, QU);
      const double rvl = rrl * at(qnormm, ivpn, QV);
      const double rwl = rrl * at(qnormm, ivpn, QW);
      const double ekinl = 0.5 * (rul * rul + rvl * rvl + rwl * rwl) / rrl;
      const double rel = at(qnormm, ivpn, QREINT) + ekinl;
      const double runewl = rul - flxu;
      const double rvnewl = rvl - flxv;
      const double rwnewl = rwl - flxw;
      const double renewl = rel - flxe;
      at(qm, ivpn, QRHO) = rrnewl + hdt * srcrho;
      at(qm, ivpn, QU) = runewl / rrnewl + hdt * srcu;
      at(qm, ivpn, QV) = rvnewl / rrnewl + hdt * srcv;
      at(qm, ivpn, QW) = rwnewl / rrnewl + hdt * srcw;
      const double rhoekinl =
          0.5 * (runewl * runewl + rvnewl * rvnewl + rwnewl * rwnewl) / rrnewl;
      at(qm, ivpn, QREINT) = renewl - rhoekinl + hdt * srce;
      at(qm, ivpn, QPRES) = at(qnormm, ivpn, QPRES) - pxnew - pynew + hdt * srcp;
      at(qm, ivpn, QPRES) = mx(at(qm, ivpn, QPRES), DMIN);
    }
  }

  // ----------------------------------------------- Source/Godunov.H:1033-1060
  static void pdivu_cell(int i, int j, int k, const A4& pdivu, const A4& q1,
                         const A4& q2, const A4& q3, const A4& a1,
                         const A4& a2, const A4& a3, const A4& vol) {
    pdivu(i, j, k) =
        0.5 *
        ((q1(i + 1, j, k, GDPRES) + q1(i, j, k, GDPRES)) *
             (q1(i + 1, j, k, GDU) * a1(i + 1, j, k) - q1(i, j, k, GDU) * a1(i, j, k)) +
         (q2(i, j + 1, k, GDPRES) + q2(i, j, k, GDPRES)) *
             (q2(i, j + 1, k, GDV) * a2(i, j + 1, k) - q2(i, j, k, GDV) * a2(i, j, k)) +
         (q3(i, j, k + 1, GDPRES) + q3(i, j, k, GDPRES)) *
             (q3(i, j, k + 1, GDW) * a3(i, j, k + 1) - q3(i, j, k, GDW) * a3(i, j, k))) /
        vol(i, j, k);
  }

  // ------------------------------------------------ Source/Godunov.cpp:7-55
  static void low_order_boundary(const Box& bfbx, int bclo, int bchi,
                                 int domlo, int domhi, int plm_iorder,
                                 bool use_flattening, int idir, double dx,
                                 double dt, const A4& q, const A4& qa,
                                 const A4& flx, const A4& qdir) {
    Box bdbx = bfbx;  // enclosedCells(bfbx).grow(idir, 1)
    bdbx.hi[idir] -= 1;
    bdbx = grow(bdbx, idir, 1);
    const Box bdbx2 = growHi(bdbx, idir, 1);
    Fab qbm(bdbx2, QVAR), qbp(bdbx, QVAR);
    ParallelFor(bdbx, [&](int i, int j, int k) {
      double slope[QVAR];
      double flat = 1.0;
      if (use_flattening)
        for (int df = 0; df < 3; df++) flat = mn(flat, flatten(i, j, k, df, q));
      for (int n = 0; n < QVAR; ++n)
        slope[n] = plm_slope(i, j, k, n, idir, q, flat, plm_iorder);
      plm_d(i, j, k, idir, qbm.a, qbp.a, slope, q, qa(i, j, k, QC), dx, dt);
    });
    ParallelFor(bfbx, [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo, bchi, domlo, domhi, qbm.a, qbp.a, flx, qdir, qa,
             idir);
    });
  }

  // ------------------------------------------------ Source/Godunov.cpp:60-476
  static int umeth_3D(const Box& bx, const int* bclo, const int* bchi,
                      const int* domlo, const int* domhi, const A4& q,
                      const A4& qaux, const A4& srcQ, const A4* flx,
                      const A4* qec, const A4* a, const A4& pdivu,
                      const A4& vol, const double* del, double dt,
                      int ppm_type, int plm_iorder, bool use_flattening) {
    const double dx = del[0], dy = del[1], dz = del[2];
    const double hdtdx = 0.5 * dt / dx, hdtdy = 0.5 * dt / dy,
                 hdtdz = 0.5 * dt / dz;
    const double cdtdx = 1.0 / 3.0 * dt / dx, cdtdy = 1.0 / 3.0 * dt / dy,
                 cdtdz = 1.0 / 3.0 * dt / dz;
    const double hdt = 0.5 * dt;
    const Box bxg1 = grow(bx, 1), bxg2 = grow(bx, 2);
    Box mbx[3], flxbx[3];
    for (int d = 0; d < 3; ++d) {
      mbx[d] = growHi(bxg2, d, 1);
      flxbx[d] = surroundingNodes(grow(bxg2, d, -1), d);
    }
    Fab qxm(mbx[0], QVAR), qxp(bxg2, QVAR), qym(mbx[1], QVAR), qyp(bxg2, QVAR),
        qzm(mbx[2], QVAR), qzp(bxg2, QVAR);
    const A4* QM[3] = {&qxm.a, &qym.a, &qzm.a};
    const A4* QP[3] = {&qxp.a, &qyp.a, &qzp.a};
    if (ppm_type == 0) {
      ParallelFor(bxg2, [&](int i, int j, int k) {
        double slope[QVAR];
        double flat = 1.0;
        if (use_flattening)
          for (int df = 0; df < 3; df++) flat = mn(flat, flatten(i, j, k, df, q));
        for (int idir = 0; idir < 3; ++idir) {
          for (int n = 0; n < QVAR; ++n)
            slope[n] = plm_slope(i, j, k, n, idir, q, flat, plm_iorder);
          plm_d(i, j, k, idir, *QM[idir], *QP[idir], slope, q, qaux(i, j, k, QC),
                del[idir], dt);
        }
      });
    } else if (ppm_type == 1) {
      for (int idir = 0; idir < 3; ++idir)
        trace_ppm(bxg2, idir, q, *QM[idir], *QP[idir], bxg2, dt, del,
                  use_flattening);
    } else {
      return -3;  // "PeleC::ppm_type must be 0 (PLM) or 1 (PPM)"
    }
    // first fluxes
    Fab fx(flxbx[0], NVAR), qgdx(flxbx[0], NGDNV), fy(flxbx[1], NVAR),
        qgdy(flxbx[1], NGDNV), fz(flxbx[2], NVAR), qgdz(flxbx[2], NGDNV);
    ParallelFor(flxbx[0], [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[0], bchi[0], domlo[0], domhi[0], qxm.a, qxp.a, fx.a,
             qgdx.a, qaux, 0);
    });
    ParallelFor(flxbx[1], [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[1], bchi[1], domlo[1], domhi[1], qym.a, qyp.a, fy.a,
             qgdy.a, qaux, 1);
    });
    ParallelFor(flxbx[2], [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[2], bchi[2], domlo[2], domhi[2], qzm.a, qzp.a, fz.a,
             qgdz.a, qaux, 2);
    });
    // X interface corrections
    const Box txbx = grow(bxg1, 0, 1), txbxm = growHi(txbx, 0, 1);
    Fab qxym(txbxm, QVAR), qxyp(txbx, QVAR), qxzm(txbxm, QVAR), qxzp(txbx, QVAR);
    ParallelFor(txbx, [&](int i, int j, int k) {
      transdo(i, j, k, 0, 1, qxym.a, qxyp.a, qxm.a, qxp.a, fy.a, qaux, qgdy.a, cdtdy);
      transdo(i, j, k, 0, 2, qxzm.a, qxzp.a, qxm.a, qxp.a, fz.a, qaux, qgdz.a, cdtdz);
    });
    const Box txfxbx = surroundingNodes(bxg1, 0);
    Fab flxy(txfxbx, NVAR), flxz(txfxbx, NVAR), qxy(txfxbx, NGDNV), qxz(txfxbx, NGDNV);
    ParallelFor(txfxbx, [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[0], bchi[0], domlo[0], domhi[0], qxym.a, qxyp.a,
             flxy.a, qxy.a, qaux, 0);
      cmpflx(i, j, k, bclo[0], bchi[0], domlo[0], domhi[0], qxzm.a, qxzp.a,
             flxz.a, qxz.a, qaux, 0);
    });
    // Y interface corrections
    const Box tybx = grow(bxg1, 1, 1), tybxm = growHi(tybx, 1, 1);
    Fab qyxm(tybxm, QVAR), qyxp(tybx, QVAR), qyzm(tybxm, QVAR), qyzp(tybx, QVAR);
    ParallelFor(tybx, [&](int i, int j, int k) {
      transdo(i, j, k, 1, 0, qyxm.a, qyxp.a, qym.a, qyp.a, fx.a, qaux, qgdx.a, cdtdx);
      transdo(i, j, k, 1, 2, qyzm.a, qyzp.a, qym.a, qyp.a, fz.a, qaux, qgdz.a, cdtdz);
    });
    const Box tyfxbx = surroundingNodes(bxg1, 1);
    Fab flyx(tyfxbx, NVAR), flyz(tyfxbx, NVAR), qyx(tyfxbx, NGDNV), qyz(tyfxbx, NGDNV);
    ParallelFor(tyfxbx, [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[1], bchi[1], domlo[1], domhi[1], qyxm.a, qyxp.a,
             flyx.a, qyx.a, qaux, 1);
      cmpflx(i, j, k, bclo[1], bchi[1], domlo[1], domhi[1], qyzm.a, qyzp.a,
             flyz.a, qyz.a, qaux, 1);
    });
    // Z interface corrections
    const Box tzbx = grow(bxg1, 2, 1), tzbxm = growHi(tzbx, 2, 1);
    Fab qzxm(tzbxm, QVAR), qzxp(tzbx, QVAR), qzym(tzbxm, QVAR), qzyp(tzbx, QVAR);
    ParallelFor(tzbx, [&](int i, int j, int k) {
      transdo(i, j, k, 2, 0, qzxm.a, qzxp.a, qzm.a, qzp.a, fx.a, qaux, qgdx.a, cdtdx);
      transdo(i, j, k, 2, 1, qzym.a, qzyp.a, qzm.a, qzp.a, fy.a, qaux, qgdy.a, cdtdy);
    });
    const Box tzfxbx = surroundingNodes(bxg1, 2);
    Fab flzx(tzfxbx, NVAR), flzy(tzfxbx, NVAR), qzx(tzfxbx, NGDNV), qzy(tzfxbx, NGDNV);
    ParallelFor(tzfxbx, [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[2], bchi[2], domlo[2], domhi[2], qzxm.a, qzxp.a,
             flzx.a, qzx.a, qaux, 2);
      cmpflx(i, j, k, bclo[2], bchi[2], domlo[2], domhi[2], qzym.a, qzyp.a,
             flzy.a, qzy.a, qaux, 2);
    });
    // final fluxes
    Fab qmfab(bxg2, QVAR), qpfab(bxg1, QVAR);
    const A4& qmf = qmfab.a;
    const A4& qpf = qpfab.a;
    ParallelFor(grow(bx, 0, 1), [&](int i, int j, int k) {
      transdd(i, j, k, 0, qmf, qpf, qxm.a, qxp.a, flyz.a, flzy.a, qyz.a, qzy.a,
              qaux, srcQ, hdt, hdtdy, hdtdz);
    });
    ParallelFor(surroundingNodes(bx, 0), [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[0], bchi[0], domlo[0], domhi[0], qmf, qpf, flx[0],
             qec[0], qaux, 0);
    });
    ParallelFor(grow(bx, 1, 1), [&](int i, int j, int k) {
      transdd(i, j, k, 1, qmf, qpf, qym.a, qyp.a, flxz.a, flzx.a, qxz.a, qzx.a,
              qaux, srcQ, hdt, hdtdx, hdtdz);
    });
    ParallelFor(surroundingNodes(bx, 1), [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[1], bchi[1], domlo[1], domhi[1], qmf, qpf, flx[1],
             qec[1], qaux, 1);
    });
    ParallelFor(grow(bx, 2, 1), [&](int i, int j, int k) {
      transdd(i, j, k, 2, qmf, qpf, qzm.a, qzp.a, flxy.a, flyx.a, qxy.a, qyx.a,
              qaux, srcQ, hdt, hdtdx, hdtdy);
    });
    ParallelFor(surroundingNodes(bx, 2), [&](int i, int j, int k) {
      cmpflx(i, j, k, bclo[2], bchi[2], domlo[2], domhi[2], qmf, qpf, flx[2],
             qec[2], qaux, 2);
    });
    // Source/Godunov.cpp:443-468: Hard (inflow) and UserBC faces
    for (int idir = 0; idir < 3; idir++) {
      if (bclo[idir] == BC_INFLOW || bclo[idir] == BC_USER) {
        Box bfbx = surroundingNodes(bx, idir);
        bfbx.hi[idir] = domlo[idir];
        if (bfbx.ok())
          low_order_boundary(bfbx, bclo[idir], bchi[idir], domlo[idir],
                             domhi[idir], plm_iorder, use_flattening, idir,
                             del[idir], dt, q, qaux, flx[idir], qec[idir]);
      }
      if (bchi[idir] == BC_INFLOW || bchi[idir] == BC_USER) {
        Box bfbx = surroundingNodes(bx, idir);
        bfbx.lo[idir] = domhi[idir] + 1;
        if (bfbx.ok())
          low_order_boundary(bfbx, bclo[idir], bchi[idir], domlo[idir],
                             domhi[idir], plm_iorder, use_flattening, idir,
                             del[idir], dt, q, qaux, flx[idir], qec[idir]);
      }
    }
    ParallelFor(bx, [&](int i, int j, int k) {
      pdivu_cell(i, j, k, pdivu, qec[0], qec[1], qec[2], a[0], a[1], a[2], vol);
    });
    return 0;
  }

  // ---------------------------------------------------- Source/Hydro.H:476-517
  static void divu_node(int i, int j, int k, const A4& q, double dx, double dy,
                        double dz, const A4& divu) {
    const double ux =
        0.25 *
        (q(i, j, k, QU) - q(i - 1, j, k, QU) + q(i, j, k - 1, QU) -
         q(i - 1, j, k - 1, QU) + q(i, j - 1, k, QU) - q(i - 1, j - 1, k, QU) +
         q(i, j - 1, k - 1, QU) - q(i - 1, j - 1, k - 1, QU)) /
        dx;
    const double vy =
        0.25 *
        (q(i, j, k, QV) - q(i, j - 1, k, QV) + q(i, j, k - 1, QV) -
         q(i, j - 1, k - 1, QV) + q(i - 1, j, k, QV) - q(i - 1, j - 1, k, QV) +
         q(i - 1, j, k - 1, QV) - q(i - 1, j - 1, k - 1, QV)) /
        dy;
    const double wz =
        0.25 *
        (q(i, j, k, QW) - q(i, j, k - 1, QW) + q(i - 1, j, k, QW) -
         q(i - 1, j, k - 1, QW) + q(i, j - 1, k, QW) - q(i, j - 1, k - 1, QW) +
         q(i - 1, j - 1, k, QW) - q(i - 1, j - 1, k - 1, QW)) /
        dz;
    divu(i, j, k) = ux + vy + wz;
  }

  // ------------------------------------------------ Source/Godunov.H:1067-1138
  static void artif_visc(int i, int j, int k, const A4& flx, const A4& divu,
                         const A4& u, double dx, double difmag, int dir,
                         int domlo, int domhi, int bclo, int bchi) {
    const IV iv = ivec(i, j, k);
    const IV ivm = shift(iv, dir, -1);
    const int l1 = (dir == 0) ? 1 : 0;
    const int l2 = (dir == 2) ? 1 : 2;
    const IV ivpj = shift(iv, l1, 1);
    const IV ivpk = shift(iv, l2, 1);
    const IV ivpp = shift(ivpj, l2, 1);
    const double div =
        difmag * mn(0.0, 0.25 * (at(divu, iv, 0) + at(divu, ivpj, 0) +
                                 at(divu, ivpk, 0) + at(divu, ivpp, 0)));
    const int idx = iv[dir];
    const bool at_bndry = (idx == domlo && bclo == BC_NOSLIPWALL) ||
                          (idx == domhi + 1 && bchi == BC_NOSLIPWALL);
    for (int n = 0; n < NVAR; n++) {
      if (n != UTEMP) {
        const bool is_tang_vel = (dir == 0 && (n == UMY || n == UMZ)) ||
                                 (dir == 1 && (n == UMX || n == UMZ)) ||
                                 (dir == 2 && (n == UMX || n == UMY));
        if (!(at_bndry && is_tang_vel))
          at(flx, iv, n) += dx * div * (at(u, iv, n) - at(u, ivm, n));
      }
    }
    at(flx, iv, UTEMP) = 0.0;
  }

  // ------------------------------------------------- Source/Hydro.H:835-865
  static void norm_spec_flx(int i, int j, int k, const A4& flx) {
    double summ = 0.0;
    for (int n = UFS; n < UFS + NS; ++n) summ += flx(i, j, k, n);
    const double fac = (summ != 0.0) ? flx(i, j, k, URHO) / summ : 1.0;
    for (int n = UFS; n < UFS + NS; ++n) flx(i, j, k, n) *= fac;
  }
  static void ext_flx(int i, int j, int k, const A4& flx, const A4& a) {
    const double area = a(i, j, k);
    for (int n = 0; n < NVAR; ++n) flx(i, j, k, n) *= area;
  }

  // ------------------------------------------------ Source/Hydro.H:888-907
  static void update(int i, int j, int k, const A4& upd, const A4* flx,
                     const A4& vol, const A4& pdivu) {
    const double v = vol(i, j, k);
    for (int n = 0; n < NVAR; ++n)
      upd(i, j, k, n) += (flx[0](i, j, k, n) - flx[0](i + 1, j, k, n) +
                          flx[1](i, j, k, n) - flx[1](i, j + 1, k, n) +
                          flx[2](i, j, k, n) - flx[2](i, j, k + 1, n)) /
                         v;
    upd(i, j, k, UEINT) -= pdivu(i, j, k);
  }

  // ------------------------------------------------ Source/Hydro.cpp:304-413
  static int umdrv(const Box& bx, const int* domlo, const int* domhi,
                   const int* bclo, const int* bchi, const A4& uin,
                   const A4& uout, const A4& q, const A4& qaux, const A4& src_q,
                   const double* dx, double dt, int ppm_type, int plm_iorder,
                   bool use_flattening, double difmag, const A4* flx,
                   const A4* a, const A4& vol) {
    const Box bxg2 = grow(bx, 2);
    Fab qec0(surroundingNodes(bxg2, 0), NGDNV), qec1(surroundingNodes(bxg2, 1), NGDNV),
        qec2(surroundingNodes(bxg2, 2), NGDNV);
    const A4 qec[3] = {qec0.a, qec1.a, qec2.a};
    Fab divu(bxg2, 1), pdivu(bx, 1);
    int rc = umeth_3D(bx, bclo, bchi, domlo, domhi, q, qaux, src_q, flx, qec, a,
                      pdivu.a, vol, dx, dt, ppm_type, plm_iorder, use_flattening);
    if (rc) return rc;
    ParallelFor(bxg2, [&](int i, int j, int k) {
      divu_node(i, j, k, q, dx[0], dx[1], dx[2], divu.a);
    });
    for (int dir = 0; dir < 3; dir++) {  // pc_adjust_fluxes
      ParallelFor(surroundingNodes(bx, dir), [&](int i, int j, int k) {
        artif_visc(i, j, k, flx[dir], divu.a, uin, dx[dir], difmag, dir,
                   domlo[dir], domhi[dir], bclo[dir], bchi[dir]);
        norm_spec_flx(i, j, k, flx[dir]);
        ext_flx(i, j, k, flx[dir], a[dir]);
      });
    }
    ParallelFor(bx, [&](int i, int j, int k) {
      update(i, j, k, uout, flx, vol, pdivu.a);
    });
    return 0;
  }

  // ============================================================ MOL hydro
  // Normal / transverse velocity components and momentum components of a
  // direction (Source/MOL.cpp:37-48).
  static void mol_idx(int dir, int* q_idx, int* f_idx) {
    const int b0 = dir == 0, b1 = dir == 1, b2 = dir == 2;
    q_idx[0] = b0 * QU + b1 * QV + b2 * QW;
    q_idx[1] = b0 * QV + b1 * QU + b2 * QU;
    q_idx[2] = b0 * QW + b1 * QW + b2 * QV;
    f_idx[0] = b0 * UMX + b1 * UMY + b2 * UMZ;
    f_idx[1] = b0 * UMY + b1 * UMX + b2 * UMX;
    f_idx[2] = b0 * UMZ + b1 * UMZ + b2 * UMY;
  }

  // ------------------------------------------------- Source/MOL.H:31-120
  // (regular cells: both neighbours connected, not covered)
  static void mol_slope(int i, int j, int k, int dir, const int* q_idx,
                        const A4& q, const A4& qaux, const A4& dq) {
    const IV iv = ivec(i, j, k);
    const IV ivm = shift(iv, dir, -1), ivp = shift(iv, dir, 1);
    double dlft[QVAR], drgt[QVAR];
    for (int n = 0; n < QVAR; ++n) dlft[n] = drgt[n] = 0.0;
    dlft[0] = 0.5 * (at(q, iv, QPRES) - at(q, ivm, QPRES)) / at(qaux, iv, QC) -
              0.5 * at(q, iv, QRHO) * (at(q, iv, q_idx[0]) - at(q, ivm, q_idx[0]));
    dlft[1] = 0.5 * (at(q, iv, QPRES) - at(q, ivm, QPRES)) / at(qaux, iv, QC) +
              0.5 * at(q, iv, QRHO) * (at(q, iv, q_idx[0]) - at(q, ivm, q_idx[0]));
    dlft[2] = at(q, iv, q_idx[1]) - at(q, ivm, q_idx[1]);
    dlft[3] = at(q, iv, q_idx[2]) - at(q, ivm, q_idx[2]);
    drgt[0] = 0.5 * (at(q, ivp, QPRES) - at(q, iv, QPRES)) / at(qaux, iv, QC) -
              0.5 * at(q, iv, QRHO) * (at(q, ivp, q_idx[0]) - at(q, iv, q_idx[0]));
    drgt[1] = 0.5 * (at(q, ivp, QPRES) - at(q, iv, QPRES)) / at(qaux, iv, QC) +
              0.5 * at(q, iv, QRHO) * (at(q, ivp, q_idx[0]) - at(q, iv, q_idx[0]));
    drgt[2] = at(q, ivp, q_idx[1]) - at(q, iv, q_idx[1]);
    drgt[3] = at(q, ivp, q_idx[2]) - at(q, iv, q_idx[2]);
    for (int n = 0; n < NS; n++) {
      dlft[QFS + n] = at(q, iv, QRHO) * at(q, iv, QFS + n) -
                      at(q, ivm, QRHO) * at(q, ivm, QFS + n) -
                      at(q, iv, QFS + n) * (at(q, iv, QPRES) - at(q, ivm, QPRES)) /
                          (at(qaux, iv, QC) * at(qaux, iv, QC));
      drgt[QFS + n] = at(q, ivp, QRHO) * at(q, ivp, QFS + n) -
                      at(q, iv, QRHO) * at(q, iv, QFS + n) -
                      at(q, iv, QFS + n) * (at(q, ivp, QPRES) - at(q, iv, QPRES)) /
                          (at(qaux, iv, QC) * at(qaux, iv, QC));
    }
    for (int n = 0; n < QVAR; n++) {
      const double dcen = 0.5 * (dlft[n] + drgt[n]);
      const double dlim = dlft[n] * drgt[n] >= 0.0
                              ? 2.0 * mn(std::abs(dlft[n]), std::abs(drgt[n]))
                              : 0.0;
      at(dq, iv, n) = std::copysign(1.0, dcen) * mn(dlim, std::abs(dcen));
    }
  }

  // ------------------------------------------------- Source/Riemann.H:203-305
  // (no rotating frame: omega = rad = 0)
  static void laxfriedrich_flux(double rl, double ul, double vl, double v2l,
                                double pl, const double* spl, double rr,
                                double ur, double vr, double v2r, double pr,
                                const double* spr, double& ustar,
                                double& maxeigval, double& uflx_rho,
                                double* uflx_rhoY, double& uflx_u,
                                double& uflx_v, double& uflx_w,
                                double& uflx_eden, double& uflx_eint) {
    const double wsmall = DMIN;
    const double omegal = 0.0, radl = 0.0, omegar = 0.0, radr = 0.0;
    double cl = 0.0, cr = 0.0, el = 0.0, er = 0.0;
    double mfl[NS], mfr[NS];
    for (int n = 0; n < NS; n++) {
      mfl[n] = spl[n];
      mfr[n] = spr[n];
    }
    EOS::RPY2Cs(rl, pl, mfl, cl);
    EOS::RPY2Cs(rr, pr, mfr, cr);
    EOS::RYP2E(rl, mfl, pl, el);
    EOS::RYP2E(rr, mfr, pr, er);
    const double r_elt = rl * (el + 0.5 * (ul * ul + vl * vl + v2l * v2l) -
                               0.5 * omegal * omegal * radl * radl);
    const double r_ert = rr * (er + 0.5 * (ur * ur + vr * vr + v2r * v2r) -
                               0.5 * omegar * omegar * radr * radr);
    const double wl = mx(wsmall, cl * rl);
    const double wr = mx(wsmall, cr * rr);
    ustar = ((wl * ul + wr * ur) + (pl - pr)) / (wl + wr);
    const bool mask =
        std::abs(ustar) < smallu * 0.5 * (std::abs(ul) + std::abs(ur)) ||
        ustar == 0.0;
    ustar = mask ? 0.0 : ustar;
    const double chalf = 0.5 * (cr + cl);
    const double max_wavespd = std::abs(ustar) + chalf;
    maxeigval = max_wavespd;
    uflx_rho = 0.0;
    for (int n = 0; n < NS; n++) {
      uflx_rhoY[n] = 0.5 * ((rl * ul * spl[n] + rr * ur * spr[n]) -
                            max_wavespd * (rr * spr[n] - rl * spl[n]));
      uflx_rho += uflx_rhoY[n];
    }
    uflx_u = 0.5 * ((rl * ul * ul + pl + rr * ur * ur + pr) -
                    max_wavespd * (rr * ur - rl * ul));
    uflx_v = 0.5 * ((rl * vl * ul + rr * vr * ur) - max_wavespd * (rr * vr - rl * vl));
    uflx_w = 0.5 * ((rl * v2l * ul + rr * v2r * ur) -
                    max_wavespd * (rr * v2r - rl * v2l));
    uflx_eden = 0.5 * (((r_elt + pl) * ul + (r_ert + pr) * ur) -
                       max_wavespd * (r_ert - r_elt));
    uflx_eint = 0.5 * ((rl * el * ul + rr * er * ur) - max_wavespd * (rr * er - rl * el));
  }

  // ------------------------------------------------- Source/MOL.cpp:6-190
  // pc_compute_hyp_mol_flux (regular cells, no rotating frame, no advected or
  // linear scalars): flx[dir] += flux * area on the faces of grow(cbox,dir,-1)
  static void hyp_mol_flux(const Box& cbox, const A4& q, const A4& qaux,
                           const A4* flx, const A4* area, int mol_iorder,
                           bool use_laxf_flux) {
    const int bc_test_val = 1;
    for (int dir = 0; dir < 3; dir++) {
      Fab dq_fab(cbox, QVAR);
      const A4 dq = dq_fab.a;
      ParallelFor(cbox, [&](int i, int j, int k) {
        for (int n = 0; n < QVAR; ++n) dq(i, j, k, n) = 0.0;  // setV
      });
      int q_idx[3], f_idx[3];
      mol_idx(dir, q_idx, f_idx);
      if (mol_iorder != 1) {
        ParallelFor(cbox, [&](int i, int j, int k) {
          mol_slope(i, j, k, dir, q_idx, q, qaux, dq);
        });
      }
      Box tbox = cbox;
      tbox.lo[dir] += 1;
      tbox.hi[dir] -= 1;
      const Box ebox = surroundingNodes(tbox, dir);
      ParallelFor(ebox, [&](int i, int j, int k) {
        const IV iv = ivec(i, j, k);
        const IV ivm = shift(iv, dir, -1);
        // face states in the normal frame: rho, un, ut1, ut2, p, Y
        double l[5], r[5], spl[NS], spr[NS];
        l[1] = at(q, ivm, q_idx[0]) + 0.5 * ((at(dq, ivm, 1) - at(dq, ivm, 0)) / at(q, ivm, QRHO));
        l[4] = at(q, ivm, QPRES) + 0.5 * (at(dq, ivm, 0) + at(dq, ivm, 1)) * at(qaux, ivm, QC);
        l[2] = at(q, ivm, q_idx[1]) + 0.5 * at(dq, ivm, 2);
        l[3] = at(q, ivm, q_idx[2]) + 0.5 * at(dq, ivm, 3);
        l[0] = 0.0;
        for (int n = 0; n < NS; n++) {
          spl[n] = at(q, ivm, QFS + n) * at(q, ivm, QRHO) +
                   0.5 * (at(dq, ivm, QFS + n) +
                          at(q, ivm, QFS + n) * (at(dq, ivm, 0) + at(dq, ivm, 1)) /
                              at(qaux, ivm, QC));
          l[0] += spl[n];
        }
        for (int n = 0; n < NS; n++) spl[n] = spl[n] / l[0];
        r[1] = at(q, iv, q_idx[0]) - 0.5 * ((at(dq, iv, 1) - at(dq, iv, 0)) / at(q, iv, QRHO));
        r[4] = at(q, iv, QPRES) - 0.5 * (at(dq, iv, 0) + at(dq, iv, 1)) * at(qaux, iv, QC);
        r[2] = at(q, iv, q_idx[1]) - 0.5 * at(dq, iv, 2);
        r[3] = at(q, iv, q_idx[2]) - 0.5 * at(dq, iv, 3);
        r[0] = 0.0;
        for (int n = 0; n < NS; n++) {
          spr[n] = at(q, iv, QFS + n) * at(q, iv, QRHO) -
                   0.5 * (at(dq, iv, QFS + n) +
                          at(q, iv, QFS + n) * (at(dq, iv, 0) + at(dq, iv, 1)) /
                              at(qaux, iv, QC));
          r[0] += spr[n];
        }
        for (int n = 0; n < NS; n++) spr[n] = spr[n] / r[0];
        const double cavg = 0.5 * (at(qaux, iv, QC) + at(qaux, ivm, QC));
        double flux_tmp[NVAR];
        for (int n = 0; n < NVAR; ++n) flux_tmp[n] = 0.0;
        double ustar = 0.0;
        if (!use_laxf_flux) {
          double qint_iu = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
          riemann(l[0], l[1], l[2], l[3], l[4], spl, r[0], r[1], r[2], r[3],
                  r[4], spr, bc_test_val, cavg, ustar, flux_tmp[URHO],
                  &flux_tmp[UFS], flux_tmp[f_idx[0]], flux_tmp[f_idx[1]],
                  flux_tmp[f_idx[2]], flux_tmp[UEDEN], flux_tmp[UEINT],
                  qint_iu, t1, t2, t3, t4);
        } else {
          double maxeigval = 0.0;
          laxfriedrich_flux(l[0], l[1], l[2], l[3], l[4], spl, r[0], r[1],
                            r[2], r[3], r[4], spr, ustar, maxeigval,
                            flux_tmp[URHO], &flux_tmp[UFS], flux_tmp[f_idx[0]],
                            flux_tmp[f_idx[1]], flux_tmp[f_idx[2]],
                            flux_tmp[UEDEN], flux_tmp[UEINT]);
        }
        flux_tmp[UTEMP] = 0.0;
        for (int ivar = 0; ivar < NVAR; ivar++)
          at(flx[dir], iv, ivar) += flux_tmp[ivar] * area[dir](i, j, k);
      });
    }
  }

  // getMOLSrcTerm for one regular box with diffusion, filtering, EB and
  // refluxing off (Source/Diffusion.cpp:90-98,136-148,226-240,349-374,
  // 509-519,654): primitives on grow(vbox,4), face fluxes zeroed on
  // surroundingNodes(grow(vbox,3),d) (the diffusive ones are all cleared by
  // :332-346 when no diffuse_* switch is on), the MOL hydro fluxes added,
  // then pc_flux_div (Source/Diffterm.H:470-487) on vbox into src.
  static void mol_src(const Box& vbox, const A4& S, const A4& src,
                      const A4* area, const A4& vol, int mol_iorder,
                      bool use_laxf_flux, const A4* flx_out = nullptr) {
    const Box gbox = grow(vbox, 4), cbox = grow(vbox, 3);
    Fab q(gbox, QVAR), qa(gbox, NQAUX);
    ParallelFor(gbox, [&](int i, int j, int k) { ctoprim(i, j, k, S, q.a, qa.a); });
    Fab f0(surroundingNodes(cbox, 0), NVAR), f1(surroundingNodes(cbox, 1), NVAR),
        f2(surroundingNodes(cbox, 2), NVAR);
    for (Fab* f : {&f0, &f1, &f2}) std::fill(f->v.begin(), f->v.end(), 0.0);
    const A4 flx[3] = {f0.a, f1.a, f2.a};
    hyp_mol_flux(cbox, q.a, qa.a, flx, area, mol_iorder, use_laxf_flux);
    ParallelFor(vbox, [&](int i, int j, int k) {
      for (int n = 0; n < NVAR; ++n)
        src(i, j, k, n) =
            -(flx[0](i + 1, j, k, n) - flx[0](i, j, k, n) + flx[1](i, j + 1, k, n) -
              flx[1](i, j, k, n) + flx[2](i, j, k + 1, n) - flx[2](i, j, k, n)) /
            vol(i, j, k);
    });
    if (flx_out)
      for (int d = 0; d < 3; ++d)
        ParallelFor(surroundingNodes(vbox, d), [&](int i, int j, int k) {
          for (int n = 0; n < NVAR; ++n) flx_out[d](i, j, k, n) = flx[d](i, j, k, n);
        });
  }

  // ------------------------------------- Source/PeleCUtilities.H:70-84, 227-236
  static void rst_int_e_default(int i, int j, int k, const A4& S) {
    const double rho = S(i, j, k, URHO);
    const double rhoInv = 1.0 / rho;
    const double Up = S(i, j, k, UMX) * rhoInv;
    const double Vp = S(i, j, k, UMY) * rhoInv;
    const double Wp = S(i, j, k, UMZ) * rhoInv;
    const double ke = 0.5 * (Up * Up + Vp * Vp + Wp * Wp) - 0.5 * 0.0 * 0.0 * 0.0 * 0.0;
    S(i, j, k, UEINT) = S(i, j, k, UEDEN) - rho * ke;
  }
  static void cmpTemp(int i, int j, int k, const A4& S) {
    const double rhoInv = 1.0 / S(i, j, k, URHO);
    double T = S(i, j, k, UTEMP);
    const double e = S(i, j, k, UEINT) * rhoInv;
    double Y[NS];
    for (int n = 0; n < NS; ++n) Y[n] = S(i, j, k, UFS + n) * rhoInv;
    const double rho = S(i, j, k, URHO);
    EOS::REY2T(rho, e, Y, T);
    S(i, j, k, UTEMP) = T;
  }

  // ------------------------------------------------- Source/Timestep.H:51-88
  static double estdt_hydro(const Box& bx, const A4& u, double dx, double dy,
                            double dz) {
    double dt = std::numeric_limits<double>::max();
    ParallelFor(bx, [&](int i, int j, int k) {
      const double rho = u(i, j, k, URHO);
      const double rhoInv = 1.0 / rho;
      const double T = u(i, j, k, UTEMP);
      double Y[NS];
      for (int n = 0; n < NS; ++n) Y[n] = u(i, j, k, UFS + n) * rhoInv;
      double c;
      EOS::RTY2Cs(rho, T, Y, c);
      const double ux = u(i, j, k, UMX) * rhoInv;
      dt = mn(dt, dx / (c + std::abs(ux)));
      const double uy = u(i, j, k, UMY) * rhoInv;
      dt = mn(dt, dy / (c + std::abs(uy)));
      const double uz = u(i, j, k, UMZ) * rhoInv;
      dt = mn(dt, dz / (c + std::abs(uz)));
    });
    return dt;
  }
};

}  // namespace orc
