// Device physics of the Godunov step for gfx950: EOS, PPM/PLM reconstruction
// and tracing, the approximate Riemann solver and the CTU transverse
// corrections, as register-level functions of one cell or one face.
//
// Hydrodynamics arithmetic follows the reference expressions term by term
// (the same products, sums and roundings; no FMA contraction: built with
// -ffp-contract=off), so the gamma-law path reproduces a CPU restatement of
// the reference bit for bit.  The mixture EOS is the exception: its NASA
// polynomials are evaluated by Horner's rule in explicit FMAs from
// compile-time coefficients pre-scaled by R/W, and the e -> T Newton
// iteration runs on mass-weighted mixture coefficients formed once per call
// (same functions, different rounding: agreement to ~1e-15 relative).
// Data movement is this engine's own: states live in registers (FaceState),
// stencils are gathered by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>
#include <utility>

#define PCD __device__ __forceinline__

namespace pcd {

// Component maps (Source/IndexDefines.H:6-37)
constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4, UEINT = 5,
              UTEMP = 6, UFS = 7;
constexpr int QRHO = 0, QU = 1, QV = 2, QW = 3, QGAME = 4, QPRES = 5,
              QREINT = 6, QTEMP = 7, QFS = 8;
constexpr int QGAMC = 0, QC = 1, QCSML = 2, QDPDR = 3, QDPDE = 4, QRSPEC = 5;
constexpr int NQAUX = 6;

// PCPhysBCType (Source/PeleC.H:48-58)
constexpr int BC_INTERIOR = 0, BC_INFLOW = 1, BC_OUTFLOW = 2, BC_SYMMETRY = 3,
              BC_SLIPWALL = 4, BC_NOSLIPWALL = 5, BC_USER = 6;

// Source/Constants.H:18-27
constexpr double smallu = 1.0e-12;
constexpr double small_num = 1.0e-8;
constexpr double DEPS = std::numeric_limits<double>::epsilon();
constexpr double very_small_num = DEPS * 1e-100;
constexpr double DMIN = std::numeric_limits<double>::min();
constexpr double RU = 8.31446261815324e7;  // erg/(mol K)

// amrex::min / amrex::max (std::min / std::max semantics)
PCD double mn(double a, double b) { return (b < a) ? b : a; }
PCD double mx(double a, double b) { return (a < b) ? b : a; }
PCD float mn(float a, float b) { return (b < a) ? b : a; }
PCD float mx(float a, float b) { return (a < b) ? b : a; }

// ------------------------------------------------------------------- EOS --
// Calorically perfect gas, gamma = 1.4, one species of molecular weight
// 28.0134 (the T scale; nothing in the hydro depends on it).
struct EosGamma {
  static constexpr int NS = 1;
  static constexpr double gamma = 1.4;
  static constexpr double imw0 = 1.0 / 28.0134;

  template <class FP>
  PCD static FP wbar(const FP* Y) {
    FP s = FP(0.0);
    s += Y[0] * FP(imw0);
    return FP(1.0) / s;
  }
  // e(rho, p) and c(rho, p) in the precision of their arguments
  template <class FP>
  PCD static FP RYP2E(FP R, const FP* /*Y*/, FP P) {
    return P / (R * (FP(gamma) - FP(1.0)));
  }
  template <class FP>
  PCD static FP RPY2Cs(FP R, FP P, const FP* /*Y*/) {
    return sqrt(FP(gamma) * P / R);
  }
  PCD static double EY2T(double E, const double* Y, double /*Tguess*/) {
    const double w = wbar(Y);
    const double poverrho = (gamma - 1.0) * E;
    return poverrho * w / RU;
  }
  PCD static double RTY2P(double R, double T, const double* Y) {
    const double w = wbar(Y);
    return R * RU * T / w;
  }
  PCD static double RTY2Cs(double R, double T, const double* Y) {
    const double P = RTY2P(R, T, Y);
    return sqrt(gamma * P / R);
  }
  PCD static double RTY2G(double, double, const double*) { return gamma; }
  PCD static void RTY2dpde_dpdre(double R, double T, const double* Y,
                                 double& dpde, double& dpdr_e) {
    const double w = wbar(Y);
    const double E = RU * T / (w * (gamma - 1.0));
    dpde = (gamma - 1.0) * R;
    dpdr_e = (gamma - 1.0) * E;
  }
  PCD static double TY2E(double T, const double* Y) {
    double ei0 = RU * T * imw0 / (gamma - 1.0);
    double e = 0.0;
    e += Y[0] * ei0;
    return e;
  }
  // state functions of one composition (the interface of the mixture EOS)
  struct Thermo {
    double Y[1];
    PCD explicit Thermo(const double* Y_) { Y[0] = Y_[0]; }
    PCD double wbar() const { return EosGamma::wbar(Y); }
    PCD double EY2T(double E, double Tg) const { return EosGamma::EY2T(E, Y, Tg); }
    PCD double RTY2P(double R, double T) const { return EosGamma::RTY2P(R, T, Y); }
    PCD double RTY2G(double R, double T) const { return EosGamma::RTY2G(R, T, Y); }
    PCD double RTY2Cs(double R, double T) const { return EosGamma::RTY2Cs(R, T, Y); }
    PCD void RTY2dpde_dpdre(double R, double T, double& dpde, double& dpdr_e) const {
      EosGamma::RTY2dpde_dpdre(R, T, Y, dpde, dpdr_e);
    }
  };
};

// Ideal-gas mixture, 9 species (H2 O2 H2O H O OH HO2 H2O2 N2), NASA
// 7-coefficient polynomials with T_mid = 1000 K (Li-Dryer thermo data).
// Tables hold a0..a5 of each species for T < 1000 K (lo) and T >= 1000 K
// (hi):  cp/R = a0 + a1 T + a2 T^2 + a3 T^3 + a4 T^4,
//        h/RT = a0 + a1 T/2 + a2 T^2/3 + a3 T^3/4 + a4 T^4/5 + a5/T.
struct LiDryerTab {
  static constexpr double mw[9] = {2.01594,  31.9988,  18.01534,
                                   1.00797,  15.9994,  17.00737,
                                   33.00677, 34.01474, 28.0134};
  static constexpr double lo[9][6] = {
      {3.298124e+00, 8.249442e-04, -8.143015e-07, -9.475434e-11, 4.134872e-13,
       -1.012521e+03},
      {3.212936e+00, 1.127486e-03, -5.756150e-07, 1.313877e-09, -8.768554e-13,
       -1.005249e+03},
      {3.386842e+00, 3.474982e-03, -6.354696e-06, 6.968581e-09, -2.506588e-12,
       -3.020811e+04},
      {2.500000e+00, 0.0, 0.0, 0.0, 0.0, 2.547163e+04},
      {2.946429e+00, -1.638166e-03, 2.421032e-06, -1.602843e-09, 3.890696e-13,
       2.914764e+04},
      {4.12530561e+00, -3.22544939e-03, 6.52764691e-06, -5.79853643e-09,
       2.06237379e-12, 3.34630913e+03},
      {4.30179801e+00, -4.74912051e-03, 2.11582891e-05, -2.42763894e-08,
       9.29225124e-12, 2.94808040e+02},
      {3.388754e+00, 6.569226e-03, -1.485013e-07, -4.625806e-09, 2.471515e-12,
       -1.766315e+04},
      {3.298677e+00, 1.408240e-03, -3.963222e-06, 5.641515e-09, -2.444855e-12,
       -1.020900e+03}};
  static constexpr double hi[9][6] = {
      {2.991423e+00, 7.000644e-04, -5.633829e-08, -9.231578e-12, 1.582752e-15,
       -8.350340e+02},
      {3.697578e+00, 6.135197e-04, -1.258842e-07, 1.775281e-11, -1.136435e-15,
       -1.233930e+03},
      {2.672146e+00, 3.056293e-03, -8.730260e-07, 1.200996e-10, -6.391618e-15,
       -2.989921e+04},
      {2.500000e+00, 0.0, 0.0, 0.0, 0.0, 2.547163e+04},
      {2.542060e+00, -2.755062e-05, -3.102803e-09, 4.551067e-12, -4.368052e-16,
       2.923080e+04},
      {2.86472886e+00, 1.05650448e-03, -2.59082758e-07, 3.05218674e-11,
       -1.33195876e-15, 3.68362875e+03},
      {4.01721090e+00, 2.23982013e-03, -6.33658150e-07, 1.14246370e-10,
       -1.07908535e-14, 1.11856713e+02},
      {4.573167e+00, 4.336136e-03, -1.474689e-06, 2.348904e-10, -1.431654e-14,
       -1.800696e+04},
      {2.926640e+00, 1.487977e-03, -5.684761e-07, 1.009704e-10, -6.753351e-15,
       -9.227977e+02}};
};

// A compile-time FP64 constant that is the addend of an FMA, held in a scalar
// register pair.  gfx950's three-operand (VOP3) encoding takes no 64-bit
// literal, and left alone the compiler puts each such addend in a vector
// register pair (two v_mov_b32, vector issue slots) to use the two-address
// v_fmac form; from an SGPR pair (two scalar moves, issued beside other
// waves' vector work) the addend feeds v_fma_f64 directly.  The empty asm
// only pins the register class: the value and the arithmetic are unchanged.
template <class FP>
PCD FP kc(FP c) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(FP) == 8) asm("" : "+s"(c));
#endif
  return c;
}

// Species K's polynomials in one NASA range with every coefficient
// pre-multiplied by RU / W_K (compile-time literals):
//   cv_K(T) = v0 + v1 T + v2 T^2 + v3 T^3 + v4 T^4                erg/(g K)
//   e_K(T)  = u0 + u1 T + u2 T^2 + u3 T^3 + u4 T^4 + u5 T^5       erg/g
// v = RU/W (a0 - 1, a1, a2, a3, a4), u = RU/W (a5, a0 - 1, a1/2, a2/3,
// a3/4, a4/5): e = RT (a0 - 1 + a1 T/2 + ... + a4 T^4/5 + a5/T) multiplied
// through by T, so no division.  Horner's rule in fused multiply-adds.
template <int K, bool LO>
struct Nasa {
  static constexpr const double* a = LO ? LiDryerTab::lo[K] : LiDryerTab::hi[K];
  static constexpr double s = RU / LiDryerTab::mw[K];
  static constexpr double imw = 1.0 / LiDryerTab::mw[K];
  static constexpr bool flat =
      a[1] == 0.0 && a[2] == 0.0 && a[3] == 0.0 && a[4] == 0.0;
  static constexpr double v0 = (a[0] - 1.0) * s, v1 = a[1] * s, v2 = a[2] * s,
                          v3 = a[3] * s, v4 = a[4] * s;
  static constexpr double u0 = a[5] * s, u1 = v0, u2 = a[1] / 2.0 * s,
                          u3 = a[2] / 3.0 * s, u4 = a[3] / 4.0 * s,
                          u5 = a[4] / 5.0 * s;
  template <class FP>
  PCD static FP cv(FP T) {
    if constexpr (flat) return FP(v0);
    return fma(T, fma(T, fma(T, fma(T, FP(v4), kc(FP(v3))), kc(FP(v2))), kc(FP(v1))),
               kc(FP(v0)));
  }
  template <class FP>
  PCD static FP e(FP T) {
    if constexpr (flat) return fma(T, FP(u1), kc(FP(u0)));
    return fma(T, fma(T, fma(T, fma(T, fma(T, FP(u5), kc(FP(u4))), kc(FP(u3))),
                             kc(FP(u2))),
                      kc(FP(u1))),
               kc(FP(u0)));
  }
};

// e(T) = sum_m U[m] T^m and its derivative cv(T), Horner / FMA
PCD double poly_e(const double* U, double T) {
  return fma(T, fma(T, fma(T, fma(T, fma(T, U[5], U[4]), U[3]), U[2]), U[1]),
             U[0]);
}
PCD double poly_cv(const double* U, double T) {
  return fma(T, fma(T, fma(T, fma(T, 5.0 * U[5], 4.0 * U[4]), 3.0 * U[3]),
                    2.0 * U[2]),
             U[1]);
}

struct EosLiDryer {
  static constexpr int NS = 9;
  static constexpr double TMID = 1000.0;
  using Seq = std::make_integer_sequence<int, 9>;

  // sum_k Y_k e_k(T) (erg/g) and sum_k Y_k cv_k(T), one range, in the
  // precision FP of the arguments
  template <bool LO, class FP, int... K>
  PCD static FP e_seq(FP T, const FP* Y, std::integer_sequence<int, K...>) {
    FP E = FP(0.0);
    ((E = fma(Y[K], Nasa<K, LO>::e(T), E)), ...);
    return E;
  }
  template <bool LO, class FP, int... K>
  PCD static FP cv_seq(FP T, const FP* Y, std::integer_sequence<int, K...>) {
    FP cv = FP(0.0);
    ((cv = fma(Y[K], Nasa<K, LO>::cv(T), cv)), ...);
    return cv;
  }
  template <class FP>
  PCD static FP ubms(FP T, const FP* Y) {
    return (T < FP(TMID)) ? e_seq<true>(T, Y, Seq{}) : e_seq<false>(T, Y, Seq{});
  }
  template <class FP>
  PCD static FP cvbs(FP T, const FP* Y) {
    return (T < FP(TMID)) ? cv_seq<true>(T, Y, Seq{}) : cv_seq<false>(T, Y, Seq{});
  }

  template <class FP, int... K>
  PCD static FP sum_y_imw(const FP* Y, std::integer_sequence<int, K...>) {
    FP s = FP(0.0);
    ((s = fma(Y[K], FP(Nasa<K, true>::imw), s)), ...);
    return s;
  }
  template <class FP>
  PCD static FP wbar(const FP* Y) { return FP(1.0) / sum_y_imw(Y, Seq{}); }
  template <class FP>
  PCD static FP RYP2E(FP R, const FP* Y, FP P) {
    const FP w = wbar(Y);
    const FP T = P * w / (R * FP(RU));
    return ubms(T, Y);
  }
  template <class FP>
  PCD static FP RPY2Cs(FP R, FP P, const FP* Y) {
    const FP w = wbar(Y);
    const FP T = P * w / (R * FP(RU));
    const FP cv = cvbs(T, Y);
    const FP G = (w * cv + FP(RU)) / (w * cv);
    return sqrt(G * P / R);
  }
  PCD static double TY2E(double T, const double* Y) { return ubms(T, Y); }

  // mixture coefficients of one composition in one range
  template <bool LO, int... K>
  PCD static void mix_seq(const double* Y, double* U,
                          std::integer_sequence<int, K...>) {
#pragma unroll
    for (int m = 0; m < 6; ++m) U[m] = 0.0;
    ((U[0] = fma(Y[K], Nasa<K, LO>::u0, U[0]),
      U[1] = fma(Y[K], Nasa<K, LO>::u1, U[1]),
      U[2] = fma(Y[K], Nasa<K, LO>::u2, U[2]),
      U[3] = fma(Y[K], Nasa<K, LO>::u3, U[3]),
      U[4] = fma(Y[K], Nasa<K, LO>::u4, U[4]),
      U[5] = fma(Y[K], Nasa<K, LO>::u5, U[5])),
     ...);
  }

  // Thermodynamics of one composition: the mass-weighted polynomial
  // coefficients of both ranges are formed once, after which e(T) and cv(T)
  // cost one Horner evaluation each (the e -> T Newton loop, and the state
  // functions at the solved T).
  struct Thermo {
    double lo[6], hi[6];
    double sy;  // sum_k Y_k / W_k = 1 / Wbar
    PCD explicit Thermo(const double* Y) {
      mix_seq<true>(Y, lo, Seq{});
      mix_seq<false>(Y, hi, Seq{});
      sy = sum_y_imw(Y, Seq{});
    }
    PCD double e(double T) const {
      return (T < TMID) ? poly_e(lo, T) : poly_e(hi, T);
    }
    PCD double cv(double T) const {
      return (T < TMID) ? poly_cv(lo, T) : poly_cv(hi, T);
    }
    PCD double wbar() const { return 1.0 / sy; }
    PCD double RTY2P(double R, double T) const { return R * RU * T * sy; }
    PCD double RTY2G(double, double T) const {
      const double w = wbar();
      const double c = cv(T);
      return (w * c + RU) / (w * c);
    }
    PCD double RTY2Cs(double R, double T) const {
      const double P = RTY2P(R, T);
      const double G = RTY2G(R, T);
      return sqrt(G * P / R);
    }
    PCD void RTY2dpde_dpdre(double R, double T, double& dpde,
                            double& dpdr_e) const {
      const double w = wbar();
      dpde = R * RU / (w * cv(T));
      dpdr_e = RU * T / w;
    }
    // e -> T Newton (Fuego stopping rule: |dT| < 1e-12 or T + dT == T,
    // |dT| clipped to 100 K, bracket [90, 4000] K, 200 iterations; linear
    // extrapolation outside the bracket)
    PCD double EY2T(double E, double Tguess) const {
      const double tmin = 90.0, tmax = 4000.0;
      const double emin = poly_e(lo, tmin);
      const double emax = poly_e(hi, tmax);
      if (E < emin) return tmin - (emin - E) / poly_cv(lo, tmin);
      if (E > emax) return tmax - (emax - E) / poly_cv(hi, tmax);
      double t1 = Tguess;
      if (t1 < tmin || t1 > tmax)
        t1 = tmin + (tmax - tmin) / (emax - emin) * (E - emin);
      for (int it = 0; it < 200; ++it) {
        const bool l = t1 < TMID;
        double c[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) c[m] = l ? lo[m] : hi[m];
        double dt = (E - poly_e(c, t1)) / poly_cv(c, t1);
        if (dt > 100.0) {
          dt = 100.0;
        } else if (dt < -100.0) {
          dt = -100.0;
        } else if (fabs(dt) < 1.0e-12) {
          break;
        } else if (t1 + dt == t1) {
          break;
        }
        t1 += dt;
      }
      return t1;
    }
  };
};

// Quotient a / b given rb = 1 / b (both correctly rounded): Markstein's
// correction of a * rb, q0 = a*rb, q = q0 + rb*(a - b*q0) with the remainder
// exact by FMA.  It equals the correctly rounded a / b whenever neither
// overflows nor underflows, so one division serves every quotient by the
// same divisor (the species fractions of a state) at three operations each.
template <class FP>
PCD FP div_by(FP a, FP b, FP rb) {
  const FP q0 = a * rb;
  return fma(fma(-q0, b, a), rb, q0);
}

// Floors of the reference in precision FP: numeric_limits<FP>::min(), and
// very_small_num (epsilon * 1e-100), which underflows in FP32 and becomes
// FLT_MIN there.
template <class FP>
PCD constexpr FP rmin() { return std::numeric_limits<FP>::min(); }
template <class FP>
PCD constexpr FP vsmall() {
  if constexpr (std::is_same<FP, double>::value) return very_small_num;
  return std::numeric_limits<FP>::min();
}

// ---------------------------------------------------------------- states --
// One primitive state on a face: lab-frame velocity, pressure, rho*e and
// mass fractions.
template <int NS, class FP = double>
struct FaceState {
  FP rho, u, v, w, p, rhoe;
  FP Y[NS];
};

// ----------------------------------------------------- flattening (A6) --
// Source/Godunov.H:188-FP(260.)  p(o) and un(o) give the pressure and the
// velocity normal to `dir` at offset o in [-3, 3] along dir; `inside` is
// whether both +-3 neighbours lie in the primitive box.
template <class FP = double, class PF, class UF>
PCD FP flatten_line(bool inside, PF&& pq, UF&& unq) {
  if (!inside) return FP(1.0);
  const FP small_pres = FP(1.e-200);
  const FP shktst = FP(0.33e0);
  const FP zcut1 = FP(0.75e0);
  const FP zcut2 = FP(0.85e0);
  const FP dzcut = FP(1.0) / (zcut2 - zcut1);
  const FP qm = pq(-1);
  const FP qp = pq(1);
  const FP qmun = unq(-1);
  const FP qpun = unq(1);
  const FP qmm = pq(-2);
  const FP qpp = pq(2);
  FP dp = qp - qm;
  const int ishft = (dp > FP(0.0)) ? 1 : -1;
  FP denom = mx(small_pres, fabs(qpp - qmm));
  FP zeta = fabs(dp) / denom;
  const FP z = mn(FP(1.0), mx(FP(0.0), dzcut * (zeta - zcut1)));
  FP tst = (qmun - qpun >= FP(0.0)) ? FP(1.0) : FP(0.0);
  FP tmp = mn(qp, qm);
  const FP chi = (fabs(dp) / tmp) > shktst ? tst : FP(0.0);
  const FP qpshft = pq(1 - ishft);
  const FP qmshft = pq(-(1 + ishft));
  const FP qpunshft = unq(1 - ishft);
  const FP qmunshft = unq(-(1 + ishft));
  const FP qppshft = pq(2 - ishft);
  const FP qmmshft = pq(-(2 + ishft));
  dp = qpshft - qmshft;
  denom = mx(small_pres, fabs(qppshft - qmmshft));
  zeta = fabs(dp) / denom;
  const FP z2 = mn(FP(1.0), mx(FP(0.0), dzcut * (zeta - zcut1)));
  tst = (qmunshft - qpunshft >= FP(0.0)) ? FP(1.0) : FP(0.0);
  tmp = mn(qpshft, qmshft);
  const FP chi2 = (fabs(dp) / tmp) > shktst ? tst : FP(0.0);
  return FP(1.0) - mx(chi2 * z2, chi * z);
}

// --------------------------------------------------------------- PPM (A7) --
// Source/PPM.H:28-115: CW84 edge values, van Leer slopes, flattening blend,
// Colella-Sekora parabola limiter.  The reference's sign(dsc) * min(|.|)
// is written copysign(min(|.|), dsc): the minimum is never negative, so the
// product by +-1 is that value with dsc's sign, bit for bit (zeros included).
template <class FP>
PCD void ppm_reconstruct(const FP s[5], FP flatn, FP& sm,
                         FP& sp) {
  FP dsl = FP(2.0) * (s[1] - s[0]);
  FP dsr = FP(2.0) * (s[2] - s[1]);
  FP dsvl_l = FP(0.0);
  if (dsl * dsr > FP(0.0)) {
    const FP dsc = FP(0.5) * (s[2] - s[0]);
    dsvl_l = copysign(mn(fabs(dsc), mn(fabs(dsl), fabs(dsr))), dsc);
  }
  dsl = FP(2.0) * (s[2] - s[1]);
  dsr = FP(2.0) * (s[3] - s[2]);
  FP dsvl_r = FP(0.0);
  if (dsl * dsr > FP(0.0)) {
    const FP dsc = FP(0.5) * (s[3] - s[1]);
    dsvl_r = copysign(mn(fabs(dsc), mn(fabs(dsl), fabs(dsr))), dsc);
  }
  sm = FP(0.5) * (s[2] + s[1]) - (FP(1.0) / FP(6.0)) * (dsvl_r - dsvl_l);
  sm = mx(sm, mn(s[2], s[1]));
  sm = mn(sm, mx(s[2], s[1]));
  // the left slope of the right edge equals the right slope of the left edge
  dsvl_l = dsvl_r;
  dsl = FP(2.0) * (s[3] - s[2]);
  dsr = FP(2.0) * (s[4] - s[3]);
  dsvl_r = FP(0.0);
  if (dsl * dsr > FP(0.0)) {
    const FP dsc = FP(0.5) * (s[4] - s[2]);
    dsvl_r = copysign(mn(fabs(dsc), mn(fabs(dsl), fabs(dsr))), dsc);
  }
  sp = FP(0.5) * (s[3] + s[2]) - (FP(1.0) / FP(6.0)) * (dsvl_r - dsvl_l);
  sp = mx(sp, mn(s[3], s[2]));
  sp = mn(sp, mx(s[3], s[2]));
  sm = flatn * sm + (FP(1.0) - flatn) * s[2];
  sp = flatn * sp + (FP(1.0) - flatn) * s[2];
  if ((sp - s[2]) * (s[2] - sm) <= FP(0.0)) {
    sp = s[2];
    sm = s[2];
  } else if (fabs(sp - s[2]) >= FP(2.0) * fabs(sm - s[2])) {
    sp = FP(3.0) * s[2] - FP(2.0) * sm;
  } else if (fabs(sm - s[2]) >= FP(2.0) * fabs(sp - s[2])) {
    sm = FP(3.0) * s[2] - FP(2.0) * sp;
  }
}

// One wave of Source/PPM.H:131-189: integral of the parabola over the
// domain of dependence of a wave of speed `speed`, from the left (Im) or the
// right (Ip) edge of the zone.
template <class FP>
PCD FP ppm_im(FP sm, FP sp, FP s6, FP speed, FP dtdx) {
  const FP sigma = fabs(speed) * dtdx;
  if (speed <= FP(0.0))
    return sm + FP(0.5) * sigma * (sp - sm + (FP(1.0) - (FP(2.0) / FP(3.0)) * sigma) * s6);
  return sm;
}
template <class FP>
PCD FP ppm_ip(FP sm, FP sp, FP s6, FP speed, FP dtdx) {
  const FP sigma = fabs(speed) * dtdx;
  if (speed <= FP(0.0)) return sp;
  return sp - FP(0.5) * sigma * (sp - sm - (FP(1.0) - (FP(2.0) / FP(3.0)) * sigma) * s6);
}
template <class FP>
PCD FP ppm_s6(FP sm, FP sp, FP sc) {
  return FP(6.0) * sc - FP(3.0) * (sm + sp);
}

// ---------------------------------------------------------------- PLM (A8) --
// Source/PLM.H:21-77 on a 5-point line q[0..4] centred at q[2].
template <class FP>
PCD FP plm_slope(const FP q[5], FP flat, int order) {
  if (order == 1) return FP(0.0);
  FP dlft, drgt, dcen, dfm = FP(0.0), dfp = FP(0.0), dlim, dsgn, dtemp;
  const FP qc = q[2], qm2 = q[0], qm = q[1], qp = q[3], qp2 = q[4];
  if (order == 4) {
    dlft = qm - qm2;
    drgt = qc - qm;
    dcen = FP(0.5) * (dlft + drgt);
    dsgn = copysign(FP(1.0), dcen);
    dlim = (dlft * drgt >= FP(0.0)) ? FP(2.0) * mn(fabs(dlft), fabs(drgt)) : FP(0.0);
    dfm = dsgn * mn(dlim, fabs(dcen));
    dlft = qp - qc;
    drgt = qp2 - qp;
    dcen = FP(0.5) * (dlft + drgt);
    dsgn = copysign(FP(1.0), dcen);
    dlim = (dlft * drgt >= FP(0.0)) ? FP(2.0) * mn(fabs(dlft), fabs(drgt)) : FP(0.0);
    dfp = dsgn * mn(dlim, fabs(dcen));
  }
  dlft = qc - qm;
  drgt = qp - qc;
  dcen = FP(0.5) * (dlft + drgt);
  dsgn = copysign(FP(1.0), dcen);
  dlim = (dlft * drgt >= FP(0.0)) ? FP(2.0) * mn(fabs(dlft), fabs(drgt)) : FP(0.0);
  dtemp = (order == 2) ? dcen : FP(4.0) / FP(3.0) * dcen - FP(1.0) / FP(6.0) * (dfp + dfm);
  return flat * dsgn * mn(dlim, fabs(dtemp));
}

// Source/PLM.H:167-FP(309.)  Cell state c (rho, vel[3], p, rhoe, Y), slopes sl
// of the same layout, sound speed cc; writes the right state at the low face
// (qp) and the left state at the high face (qm) of the cell in direction dir.
template <int NS, class FP>
PCD void plm_trace(int dir, const FaceState<NS, FP>& c, const FaceState<NS, FP>& sl,
                   FP cc, FP dtdx, FaceState<NS, FP>& qp,
                   FaceState<NS, FP>& qm) {
  const int l1 = (dir == 0) ? 1 : 0;
  const int l2 = (dir == 2) ? 1 : 2;
  const FP cs = cc * cc;
  const FP rho = c.rho;
  const FP vel[3] = {c.u, c.v, c.w};
  const FP p = c.p;
  const FP rhoe = c.rhoe;
  const FP enth = ((rhoe + p) / rho) / cs;
  const FP drho = sl.rho;
  const FP dvel[3] = {sl.u, sl.v, sl.w};
  const FP dp = sl.p;
  const FP drhoe = sl.rhoe;
  const FP alpham = FP(0.5) * (dp / (rho * cc) - dvel[dir]) * rho / cc;
  const FP alphap = FP(0.5) * (dp / (rho * cc) + dvel[dir]) * rho / cc;
  const FP alpha0r = drho - dp / cs;
  const FP alpha0e = drhoe - dp * enth;
  const FP alpha0v = dvel[l1];
  const FP alpha0w = dvel[l2];
  const FP wv[3] = {vel[dir] - cc, vel[dir], vel[dir] + cc};
  FP out_vel[3];
  {
    const FP rho_ref = rho - FP(0.5) * (FP(1.0) + dtdx * mn(wv[0], FP(0.0))) * drho;
    FP vel_ref[3];
#pragma unroll
    for (int dim = 0; dim < 3; dim++)
      vel_ref[dim] = vel[dim] - FP(0.5) * (FP(1.0) + dtdx * mn(wv[0], FP(0.0))) * dvel[dim];
    const FP p_ref = p - FP(0.5) * (FP(1.0) + dtdx * mn(wv[0], FP(0.0))) * dp;
    const FP rhoe_ref = rhoe - FP(0.5) * (FP(1.0) + dtdx * mn(wv[0], FP(0.0))) * drhoe;
    const FP apright =
        FP(0.25) * dtdx * (wv[0] - wv[2]) * (FP(1.0) - copysign(FP(1.0), wv[2])) * alphap;
    const FP amright = FP(0.0);
    const FP azrright =
        FP(0.25) * dtdx * (wv[0] - wv[1]) * (FP(1.0) - copysign(FP(1.0), wv[1])) * alpha0r;
    const FP azeright =
        FP(0.25) * dtdx * (wv[0] - wv[1]) * (FP(1.0) - copysign(FP(1.0), wv[1])) * alpha0e;
    const FP azv1rght =
        FP(0.25) * dtdx * (wv[0] - wv[1]) * (FP(1.0) - copysign(FP(1.0), wv[1])) * alpha0v;
    const FP azw1rght =
        FP(0.25) * dtdx * (wv[0] - wv[1]) * (FP(1.0) - copysign(FP(1.0), wv[1])) * alpha0w;
    qp.rho = mx(rmin<FP>(), rho_ref + apright + amright + azrright);
    out_vel[dir] = vel_ref[dir] + (apright - amright) * cc / rho;
    out_vel[l1] = vel_ref[l1] + azv1rght;
    out_vel[l2] = vel_ref[l2] + azw1rght;
    qp.u = out_vel[0];
    qp.v = out_vel[1];
    qp.w = out_vel[2];
    qp.p = mx(p_ref + (apright + amright) * cs, rmin<FP>());
    qp.rhoe = rhoe_ref + (apright + amright) * enth * cs + azeright;
  }
  {
    const FP rho_ref = rho + FP(0.5) * (FP(1.0) - dtdx * mx(wv[2], FP(0.0))) * drho;
    FP vel_ref[3];
#pragma unroll
    for (int dim = 0; dim < 3; dim++)
      vel_ref[dim] = vel[dim] + FP(0.5) * (FP(1.0) - dtdx * mx(wv[2], FP(0.0))) * dvel[dim];
    const FP p_ref = p + FP(0.5) * (FP(1.0) - dtdx * mx(wv[2], FP(0.0))) * dp;
    const FP rhoe_ref = rhoe + FP(0.5) * (FP(1.0) - dtdx * mx(wv[2], FP(0.0))) * drhoe;
    const FP apleft = FP(0.0);
    const FP amleft =
        FP(0.25) * dtdx * (wv[2] - wv[0]) * (FP(1.0) + copysign(FP(1.0), wv[0])) * alpham;
    const FP azrleft =
        FP(0.25) * dtdx * (wv[2] - wv[1]) * (FP(1.0) + copysign(FP(1.0), wv[1])) * alpha0r;
    const FP azeleft =
        FP(0.25) * dtdx * (wv[2] - wv[1]) * (FP(1.0) + copysign(FP(1.0), wv[1])) * alpha0e;
    const FP azv1left =
        FP(0.25) * dtdx * (wv[2] - wv[1]) * (FP(1.0) + copysign(FP(1.0), wv[1])) * alpha0v;
    const FP azw1left =
        FP(0.25) * dtdx * (wv[2] - wv[1]) * (FP(1.0) + copysign(FP(1.0), wv[1])) * alpha0w;
    qm.rho = mx(rho_ref + apleft + amleft + azrleft, rmin<FP>());
    out_vel[dir] = vel_ref[dir] + (apleft - amleft) * cc / rho;
    out_vel[l1] = vel_ref[l1] + azv1left;
    out_vel[l2] = vel_ref[l2] + azw1left;
    qm.u = out_vel[0];
    qm.v = out_vel[1];
    qm.w = out_vel[2];
    qm.p = mx(p_ref + (apleft + amleft) * cs, rmin<FP>());
    qm.rhoe = rhoe_ref + (apleft + amleft) * enth * cs + azeleft;
  }
  const FP vel_ad = vel[dir];
  const FP spzerom = (vel_ad > 0) ? -FP(1.0) : vel_ad * dtdx;
  const FP spzerop = (vel_ad >= 0) ? vel_ad * dtdx : FP(1.0);
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const FP acmprght = FP(0.5) * (-FP(1.0) - spzerom) * sl.Y[s];
    qp.Y[s] = c.Y[s] + acmprght;
    const FP acmpleft = FP(0.5) * (FP(1.0) - spzerop) * sl.Y[s];
    qm.Y[s] = c.Y[s] + acmpleft;
  }
}

// --------------------------------------------------------- Riemann (A10) --
// Source/Riemann.H:9-200 for one face; ql/qr in the face-normal frame:
// (rho, un, ut, utt, p, Y).  Outputs: ustar, the fluxes of rho, normal and
// transverse momenta, rho E, rho e, rho Y_k (optional), and the Godunov
// state (u_gd, ut_gd, utt_gd, p_gd, game_gd).
template <class EOS, class FP = double>
struct RiemannOut {
  FP ustar, f_rho, f_un, f_ut, f_utt, f_eden, f_eint;
  FP u_gd, ut_gd, utt_gd, p_gd, game_gd;
  FP f_Y[EOS::NS];  // uflx_rhoY (read by the MOL flux; dead code otherwise)
};

template <class EOS, class FP>
PCD RiemannOut<EOS, FP> riemann(FP rl, FP ul, FP vl, FP v2l,
                            FP pl, const FP* spl, FP rr, FP ur,
                            FP vr, FP v2r, FP pr, const FP* spr,
                            FP cav) {
  constexpr int NS = EOS::NS;
  RiemannOut<EOS, FP> o;
  const FP wsmall = rmin<FP>();
  const FP cl = EOS::RPY2Cs(rl, pl, spl);
  const FP cr = EOS::RPY2Cs(rr, pr, spr);
  const FP wl = mx(wsmall, cl * rl);
  const FP wr = mx(wsmall, cr * rr);
  const FP pstar =
      mx(rmin<FP>(), ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr));
  FP ustar = ((wl * ul + wr * ur) + (pl - pr)) / (wl + wr);
  const bool up = ustar > FP(0.0);
  FP uo = up ? ul : ur;
  FP po = up ? pl : pr;
  const bool tiny =
      fabs(ustar) < FP(smallu) * FP(0.5) * (fabs(ul) + fabs(ur)) || ustar == FP(0.0);
  ustar = tiny ? FP(0.0) : ustar;
  FP rspo[NS];
  FP ro = FP(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP a = up ? rl * spl[n] : rr * spr[n];
    rspo[n] = tiny ? FP(0.5) * (rl * spl[n] + rr * spr[n]) : a;
    ro += rspo[n];
  }
  uo = tiny ? FP(0.5) * (ul + ur) : uo;
  po = tiny ? FP(0.5) * (pl + pr) : po;
  FP gmf[NS];
  const FP rro = FP(1.0) / ro;
#pragma unroll
  for (int n = 0; n < NS; n++) gmf[n] = div_by(rspo[n], ro, rro);
  const FP co = EOS::RPY2Cs(ro, po, gmf);
  const FP drho = (pstar - po) / (co * co);
  FP rspstar[NS];
  FP rstar = FP(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP spon = gmf[n];  // rspo[n] / ro
    rspstar[n] = mx(FP(0.0), rspo[n] + drho * spon);
    rstar += rspstar[n];
  }
  const FP rrstar = FP(1.0) / rstar;
#pragma unroll
  for (int n = 0; n < NS; n++) gmf[n] = div_by(rspstar[n], rstar, rrstar);
  const FP cstar = EOS::RPY2Cs(rstar, pstar, gmf);
  const FP sgnm = copysign(FP(1.0), ustar);
  FP spout = co - sgnm * uo;
  FP spin = cstar - sgnm * ustar;
  const FP ushock = FP(0.5) * (spin + spout);
  const bool rare = pstar < po;
  spout = rare ? spout : ushock;
  spin = rare ? spin : ushock;
  const FP scr =
      (fabs(spout - spin) < vsmall<FP>()) ? FP(small_num) * cav : spout - spin;
  const FP frac = mx(FP(0.0), mn(FP(1.0), (FP(1.0) + (spout + spin) / scr) * FP(0.5)));
  const bool upz = ustar > FP(0.0);
  FP qv1 = upz ? vl : vr;
  FP qv2 = upz ? v2l : v2r;
  const bool zero = (ustar == FP(0.0));
  qv1 = zero ? FP(0.5) * (vl + vr) : qv1;
  qv2 = zero ? FP(0.5) * (v2l + v2r) : qv2;
  FP rspgd[NS];
  FP qiu = frac * ustar + (FP(1.0) - frac) * uo;
  FP qp_ = frac * pstar + (FP(1.0) - frac) * po;
  const bool outside = (spout < FP(0.0));
  const bool inside = (spin >= FP(0.0));
  FP rgd = FP(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    FP v = frac * rspstar[n] + (FP(1.0) - frac) * rspo[n];
    v = outside ? rspo[n] : v;
    v = inside ? rspstar[n] : v;
    rspgd[n] = v;
    rgd += v;
  }
  qiu = outside ? uo : qiu;
  qp_ = outside ? po : qp_;
  qiu = inside ? ustar : qiu;
  qp_ = inside ? pstar : qp_;
  const FP rrgd = FP(1.0) / rgd;
#pragma unroll
  for (int n = 0; n < NS; n++) gmf[n] = div_by(rspgd[n], rgd, rrgd);
  const FP ge = EOS::RYP2E(rgd, gmf, qp_);
  const FP regd = rgd * ge;
  o.game_gd = qp_ / regd + FP(1.0);
  qiu = 1 * qiu;
  o.f_rho = rgd * qiu;
  o.f_un = o.f_rho * qiu + qp_;
  o.f_ut = o.f_rho * qv1;
  o.f_utt = o.f_rho * qv2;
  const FP rhoetot =
      regd + FP(0.5) * rgd * (qiu * qiu + qv1 * qv1 + qv2 * qv2) -
      FP(0.5) * rgd * FP(0.0) * FP(0.0) * FP(0.0) * FP(0.0);
  o.f_eden = qiu * (rhoetot + qp_);
  o.f_eint = qiu * regd;
#pragma unroll
  for (int n = 0; n < NS; n++) o.f_Y[n] = rspgd[n] * qiu;
  o.ustar = ustar;
  o.u_gd = qiu;
  o.ut_gd = qv1;
  o.utt_gd = qv2;
  o.p_gd = qp_;
  return o;
}

// laxfriedrich_flux (Source/Riemann.H:205-307) without a rotating frame:
// the Rusanov flux with wave speed |u*| + (cl + cr)/2.  Same output record
// as riemann(); only ustar and the fluxes are set.
template <class EOS, class FP>
PCD RiemannOut<EOS, FP> laxf(FP rl, FP ul, FP vl, FP v2l, FP pl, const FP* spl,
                             FP rr, FP ur, FP vr, FP v2r, FP pr, const FP* spr) {
  constexpr int NS = EOS::NS;
  RiemannOut<EOS, FP> o;
  const FP wsmall = rmin<FP>();
  const FP cl = EOS::RPY2Cs(rl, pl, spl);
  const FP cr = EOS::RPY2Cs(rr, pr, spr);
  const FP el = EOS::RYP2E(rl, spl, pl);
  const FP er = EOS::RYP2E(rr, spr, pr);
  const FP r_elt = rl * (el + FP(0.5) * (ul * ul + vl * vl + v2l * v2l) -
                         FP(0.5) * FP(0.0) * FP(0.0) * FP(0.0) * FP(0.0));
  const FP r_ert = rr * (er + FP(0.5) * (ur * ur + vr * vr + v2r * v2r) -
                         FP(0.5) * FP(0.0) * FP(0.0) * FP(0.0) * FP(0.0));
  const FP wl = mx(wsmall, cl * rl);
  const FP wr = mx(wsmall, cr * rr);
  FP ustar = ((wl * ul + wr * ur) + (pl - pr)) / (wl + wr);
  const bool tiny =
      fabs(ustar) < FP(smallu) * FP(0.5) * (fabs(ul) + fabs(ur)) || ustar == FP(0.0);
  ustar = tiny ? FP(0.0) : ustar;
  const FP chalf = FP(0.5) * (cr + cl);
  const FP a = fabs(ustar) + chalf;
  o.ustar = ustar;
  FP frho = FP(0.0);
#pragma unroll
  for (int n = 0; n < NS; n++) {
    o.f_Y[n] = FP(0.5) * ((rl * ul * spl[n] + rr * ur * spr[n]) -
                          a * (rr * spr[n] - rl * spl[n]));
    frho += o.f_Y[n];
  }
  o.f_rho = frho;
  o.f_un = FP(0.5) * ((rl * ul * ul + pl + rr * ur * ur + pr) - a * (rr * ur - rl * ul));
  o.f_ut = FP(0.5) * ((rl * vl * ul + rr * vr * ur) - a * (rr * vr - rl * vl));
  o.f_utt = FP(0.5) * ((rl * v2l * ul + rr * v2r * ur) - a * (rr * v2r - rl * v2l));
  o.f_eden = FP(0.5) * (((r_elt + pl) * ul + (r_ert + pr) * ur) - a * (r_ert - r_elt));
  o.f_eint = FP(0.5) * ((rl * el * ul + rr * er * ur) - a * (rr * er - rl * el));
  return o;
}

// ------------------------------------------------------ cmpflx BCs (A9) --
// Source/Godunov.H:378-418: wall/symmetry mirror or outflow copy at domain
// faces, on the normal-frame states (rho, un, ut, utt, p) of both sides.
// The empty asm statement ends the left-face branch differently from the
// right-face one, so the compiler does not merge the two symmetric bodies
// into one block that addresses the states through pointers (that would put
// all ten values in scratch memory).
template <class FP>
PCD void face_bc(int idx, int domlo, int domhi, int bclo, int bchi,
                 FP& rl, FP& ul, FP& vl, FP& v2l, FP& pl,
                 FP& rr, FP& ur, FP& vr, FP& v2r, FP& pr) {
  auto wall = [](int b) {
    return b == BC_NOSLIPWALL || b == BC_SLIPWALL || b == BC_SYMMETRY;
  };
  if (idx == domlo) {
    if (wall(bclo) || bclo == BC_OUTFLOW) {
      ul = wall(bclo) ? -ur : ur;
      vl = vr;
      v2l = v2r;
      pl = pr;
      rl = rr;
      asm volatile("");
    }
  } else if (idx == domhi + 1) {
    if (wall(bchi) || bchi == BC_OUTFLOW) {
      ur = wall(bchi) ? -ul : ul;
      vr = vl;
      v2r = v2l;
      pr = pl;
      rr = rl;
    }
  }
}

// Flux of one face in lab-frame layout: [rho, mx, my, mz, rhoE, rhoe, Y...]
// plus the Godunov pressure and normal velocity.
template <int NS, class FP = double>
struct Flux {
  FP f[6 + NS];  // URHO, UMX, UMY, UMZ, UEDEN, UEINT, species
  FP pgd, ugd;
};

// pc_cmpflx (Source/Godunov.H:303-449): BC-aware Riemann solve and passive
// upwinding by the mass flux.
template <class EOS, int DIR, class FP>
PCD void cmpflx(const FaceState<EOS::NS, FP>& ql, const FaceState<EOS::NS, FP>& qr,
                FP cav, int idx, int domlo, int domhi, int bclo, int bchi,
                Flux<EOS::NS, FP>& F) {
  constexpr int NS = EOS::NS;
  constexpr int IU = DIR, IV = (DIR == 0) ? 1 : 0, IV2 = (DIR == 2) ? 1 : 2;
  auto comp = [](const FaceState<NS, FP>& q, int c) {
    return c == 0 ? q.u : (c == 1 ? q.v : q.w);
  };
  FP rl = ql.rho, ul = comp(ql, IU), vl = comp(ql, IV), v2l = comp(ql, IV2),
         pl = ql.p;
  FP rr = qr.rho, ur = comp(qr, IU), vr = comp(qr, IV), v2r = comp(qr, IV2),
         pr = qr.p;
  face_bc(idx, domlo, domhi, bclo, bchi, rl, ul, vl, v2l, pl, rr, ur, vr, v2r,
          pr);
  const RiemannOut<EOS, FP> o = riemann<EOS>(rl, ul, vl, v2l, pl, ql.Y, rr, ur, vr,
                                         v2r, pr, qr.Y, cav);
  F.f[0] = o.f_rho;
  F.f[1 + IU] = o.f_un;
  F.f[1 + IV] = o.f_ut;
  F.f[1 + IV2] = o.f_utt;
  F.f[4] = o.f_eden;
  F.f[5] = o.f_eint;
  F.pgd = o.p_gd;
  F.ugd = o.u_gd;
  const FP flxrho = o.f_rho;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP qlv = ql.Y[n], qrv = qr.Y[n];
    F.f[6 + n] = (o.ustar > FP(0.0))   ? flxrho * qlv
                 : (o.ustar < FP(0.0)) ? flxrho * qrv
                                   : flxrho * FP(0.5) * (qlv + qrv);
  }
}

// ---------------------------------------------- transverse corrections --
// Flux differences across one cell in a transverse direction:
// d[0..4] = c*(F_hi - F_lo) of rho, mx, my, mz, rhoE; ds[k] of rho Y_k.
template <int NS, class FP = double>
struct DFlux {
  FP d[5];
  FP ds[NS];
  FP pgp, pgm, ugp, ugm;  // Godunov p and normal u on the two faces
};

// pc_transdo, one side (Source/Godunov.H:455-633 with the passive part
// :30-73): the dir-state `s` corrected by the transverse flux difference of
// its cell; gamc = QGAMC of that cell.
template <int NS, class FP>
PCD void transdo_side(const FaceState<NS, FP>& s, const DFlux<NS, FP>& D, FP gamc,
                      FP cdtdx, FaceState<NS, FP>& o) {
  const FP flxrho = D.d[0], flxu = D.d[1], flxv = D.d[2], flxw = D.d[3],
               flxe = D.d[4];
  const FP rrnew = s.rho - flxrho;
  const FP rrinv = FP(1.0) / rrnew;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP rr = s.rho;
    const FP compo = rr * s.Y[n] - D.ds[n];
    o.Y[n] = div_by(compo, rrnew, rrinv);
  }
  const FP dAup = D.pgp * D.ugp - D.pgm * D.ugm;
  const FP pav = FP(0.5) * (D.pgp + D.pgm);
  const FP dAu = D.ugp - D.ugm;
  const FP rrr = s.rho;
  FP rur = s.u, rvr = s.v, rwr = s.w;
  const FP ekinr = FP(0.5) * rrr * (rur * rur + rvr * rvr + rwr * rwr);
  rur *= rrr;
  rvr *= rrr;
  rwr *= rrr;
  const FP rer = s.rhoe + ekinr;
  const FP rrnewr = rrr - flxrho;
  const FP runewr = rur - flxu;
  const FP rvnewr = rvr - flxv;
  const FP rwnewr = rwr - flxw;
  const FP renewr = rer - flxe;
  o.rho = rrnewr;
  o.u = div_by(runewr, rrnewr, rrinv);
  o.v = div_by(rvnewr, rrnewr, rrinv);
  o.w = div_by(rwnewr, rrnewr, rrinv);
  const FP rhoekinr = div_by(
      FP(0.5) * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr), rrnewr, rrinv);
  o.rhoe = renewr - rhoekinr;
  const FP pnewr = s.p - cdtdx * (dAup + pav * dAu * (gamc - FP(1.0)));
  o.p = mx(pnewr, vsmall<FP>());
}

// The same correction with the Godunov pressure term already formed per
// cell: pt = cdtdx (dAup + pav dAu (Gamma1 - 1)) of that cell's two faces.
template <int NS, class FP = double>
struct TCorr {
  FP d[5];
  FP ds[NS];
  FP pt;
};

template <int NS, class FP>
PCD void transdo_side(const FaceState<NS, FP>& s, const TCorr<NS, FP>& D,
                      FaceState<NS, FP>& o) {
  const FP flxrho = D.d[0], flxu = D.d[1], flxv = D.d[2], flxw = D.d[3],
               flxe = D.d[4];
  const FP rrnew = s.rho - flxrho;
  const FP rrinv = FP(1.0) / rrnew;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP rr = s.rho;
    const FP compo = rr * s.Y[n] - D.ds[n];
    o.Y[n] = div_by(compo, rrnew, rrinv);
  }
  const FP rrr = s.rho;
  FP rur = s.u, rvr = s.v, rwr = s.w;
  const FP ekinr = FP(0.5) * rrr * (rur * rur + rvr * rvr + rwr * rwr);
  rur *= rrr;
  rvr *= rrr;
  rwr *= rrr;
  const FP rer = s.rhoe + ekinr;
  const FP rrnewr = rrr - flxrho;
  const FP runewr = rur - flxu;
  const FP rvnewr = rvr - flxv;
  const FP rwnewr = rwr - flxw;
  const FP renewr = rer - flxe;
  o.rho = rrnewr;
  o.u = div_by(runewr, rrnewr, rrinv);
  o.v = div_by(rvnewr, rrnewr, rrinv);
  o.w = div_by(rwnewr, rrnewr, rrinv);
  const FP rhoekinr = div_by(
      FP(0.5) * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr), rrnewr, rrinv);
  o.rhoe = renewr - rhoekinr;
  const FP pnewr = s.p - D.pt;
  o.p = mx(pnewr, vsmall<FP>());
}

// pc_transdd, one side (Source/Godunov.H:639-845, passive :78-132): the
// dir-state corrected by both transverse flux differences (D0 scaled by
// cdtdx0, D1 by cdtdx1 when they were formed) plus hdt * srcQ.
template <int NS, class FP>
PCD void transdd_side(const FaceState<NS, FP>& s, const FP* dd,
                      const FP* dds, FP pxnew, FP pynew,
                      const FP* src /* rho,u,v,w,rhoe,p,Y */, FP hdt,
                      FaceState<NS, FP>& o) {
  const FP flxrho = dd[0], flxu = dd[1], flxv = dd[2], flxw = dd[3],
               flxe = dd[4];
  const FP rrr = s.rho;
  const FP rrnewr = rrr - flxrho;
  const FP rrinv = FP(1.0) / rrnewr;
#pragma unroll
  for (int n = 0; n < NS; n++) {
    const FP compo = rrr * s.Y[n] - dds[n];
    o.Y[n] = div_by(compo, rrnewr, rrinv) + hdt * src[6 + n];
  }
  const FP rur = rrr * s.u;
  const FP rvr = rrr * s.v;
  const FP rwr = rrr * s.w;
  const FP ekinr = FP(0.5) * (rur * rur + rvr * rvr + rwr * rwr) / rrr;
  const FP rer = s.rhoe + ekinr;
  const FP runewr = rur - flxu;
  const FP rvnewr = rvr - flxv;
  const FP rwnewr = rwr - flxw;
  const FP renewr = rer - flxe;
  o.rho = rrnewr + hdt * src[0];
  o.u = div_by(runewr, rrnewr, rrinv) + hdt * src[1];
  o.v = div_by(rvnewr, rrnewr, rrinv) + hdt * src[2];
  o.w = div_by(rwnewr, rrnewr, rrinv) + hdt * src[3];
  const FP rhoekinr = div_by(
      FP(0.5) * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr), rrnewr, rrinv);
  o.rhoe = renewr - rhoekinr + hdt * src[4];
  FP p = s.p - pxnew - pynew + hdt * src[5];
  o.p = mx(p, rmin<FP>());
}

}  // namespace pcd
