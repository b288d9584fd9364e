// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Only tests/, the
// smoke check and the bench's cpu_baseline leg may load this code.  It is the
// checker, never the thing measured or shipped.
//
// Equation-of-state restatements used by the oracle.  The reference calls the
// PelePhysics EOS (absent from the reference tree; PelePhysics
// v1.0.0-28-g6e8026f per Exec/RegTests/CrossJet/run_small/run.log:13) through
// the names kept here: Y2WBAR, REY2T, RTY2P, RTY2Cs, RTY2G, RTY2dpde_dpdre,
// RPY2Cs, RYP2E, EY2T, T2Ei (call sites: Source/PeleCUtilities.H:300-305,82;
// Source/PPM.cpp:101,259,329,359,420; Source/Riemann.H:51-178;
// Source/Timestep.H:73; Exec/RegTests/Sod/prob.cpp:215-241).
//
// GammaLaw: closed forms of a calorically perfect gas (gamma = 1.4).  The only
// free constant is the molecular weight of the single species (28.0134, N2),
// which scales T and nothing else.  Parity against the reference's T scale is
// unpinned; every hydro quantity is independent of it.
//
// LiDryer: ideal-gas mixture of 9 species (H2, O2, H2O, H, O, OH, HO2, H2O2,
// N2) with NASA 7-coefficient polynomials (T_mid = 1000 K), the classic
// CHEMKIN thermodynamic data the Li-Dryer H2/O2 mechanism ships with, and a
// Newton e->T inversion with the Fuego-generated stopping rule (T bracket
// [90, 4000] K, |dT| < 1e-12 or T+dT == T, |dT| clipped to 100 K, 200
// iterations).  Evaluation scheme (this restatement's own; the reference's
// EOS source is absent): polynomials by Horner's rule in std::fma with
// coefficients scaled by RU/W_k; species sums as fma chains in species
// order; functions of a known composition (RTY2G, RTY2Cs, RTY2dpde_dpdre,
// EY2T/REY2T) on mass-weighted mixture coefficients formed once per call.
// The same operations on the GPU give the same bits.  Pinned by ideal-gas identities, NASA continuity at T_mid, the
// PMF flame file's molecular weights and the adiabatic enthalpy balances of
// Exec/RegTests/PMF and Exec/RegTests/TGReact (tests/test_oracle_eos.py).
// The exact coefficient digits and Newton details of the reference build are
// parity unpinned (no reference EOS source is available).
#pragma once
#include <cmath>

namespace orc {

constexpr double RU = 8.31446261815324e7;  // erg / (mol K)
constexpr double PATM = 1013250.0;

struct GammaLaw {
  static constexpr int NS = 1;
  static constexpr double gamma = 1.4;
  static constexpr double mw[1] = {28.0134};
  static constexpr double imw[1] = {1.0 / 28.0134};

  static void Y2WBAR(const double* Y, double& wbar) {
    double s = 0.0;
    for (int k = 0; k < NS; ++k) s += Y[k] * imw[k];
    wbar = 1.0 / s;
  }
  static void RYP2E(double R, const double* /*Y*/, double P, double& E) {
    E = P / (R * (gamma - 1.0));
  }
  static void RPY2Cs(double R, double P, const double* /*Y*/, double& Cs) {
    Cs = std::sqrt(gamma * P / R);
  }
  static void EY2T(double E, const double* Y, double& T) {
    double wbar;
    Y2WBAR(Y, wbar);
    const double poverrho = (gamma - 1.0) * E;
    T = poverrho * wbar / RU;
  }
  static void REY2T(double /*R*/, double E, const double* Y, double& T) {
    EY2T(E, Y, T);
  }
  static void RTY2P(double R, double T, const double* Y, double& P) {
    double wbar;
    Y2WBAR(Y, wbar);
    P = R * RU * T / wbar;
  }
  static void RTY2Cs(double R, double T, const double* Y, double& Cs) {
    double P;
    RTY2P(R, T, Y, P);
    Cs = std::sqrt(gamma * P / R);
  }
  static void RTY2G(double, double, const double*, double& G) { G = gamma; }
  static void RTY2dpde_dpdre(double R, double T, const double* Y, double& dpde,
                             double& dpdr_e) {
    double wbar;
    Y2WBAR(Y, wbar);
    const double E = RU * T / (wbar * (gamma - 1.0));
    dpde = (gamma - 1.0) * R;
    dpdr_e = (gamma - 1.0) * E;
  }
  static void T2Ei(double T, double* Ei) {
    for (int k = 0; k < NS; ++k) Ei[k] = RU * T * imw[k] / (gamma - 1.0);
  }
  static void TY2E(double T, const double* Y, double& E) {
    double ei[NS];
    T2Ei(T, ei);
    E = 0.0;
    for (int k = 0; k < NS; ++k) E += Y[k] * ei[k];
  }
};

// NASA-7 data, species order H2 O2 H2O H O OH HO2 H2O2 N2.
// Row: a0..a6 (a5 = enthalpy constant, a6 = entropy constant).
struct LiDryerData {
  static constexpr double mw[9] = {2.01594,  31.9988,  18.01534,
                                   1.00797,  15.9994,  17.00737,
                                   33.00677, 34.01474, 28.0134};
  static constexpr double lo[9][7] = {
      {3.298124e+00, 8.249442e-04, -8.143015e-07, -9.475434e-11, 4.134872e-13,
       -1.012521e+03, -3.294094e+00},
      {3.212936e+00, 1.127486e-03, -5.756150e-07, 1.313877e-09, -8.768554e-13,
       -1.005249e+03, 6.034738e+00},
      {3.386842e+00, 3.474982e-03, -6.354696e-06, 6.968581e-09, -2.506588e-12,
       -3.020811e+04, 2.590233e+00},
      {2.500000e+00, 0.0, 0.0, 0.0, 0.0, 2.547163e+04, -4.601176e-01},
      {2.946429e+00, -1.638166e-03, 2.421032e-06, -1.602843e-09, 3.890696e-13,
       2.914764e+04, 2.963995e+00},
      {4.12530561e+00, -3.22544939e-03, 6.52764691e-06, -5.79853643e-09,
       2.06237379e-12, 3.34630913e+03, -6.90432960e-01},
      {4.30179801e+00, -4.74912051e-03, 2.11582891e-05, -2.42763894e-08,
       9.29225124e-12, 2.94808040e+02, 3.71666245e+00},
      {3.388754e+00, 6.569226e-03, -1.485013e-07, -4.625806e-09, 2.471515e-12,
       -1.766315e+04, 6.785363e+00},
      {3.298677e+00, 1.408240e-03, -3.963222e-06, 5.641515e-09, -2.444855e-12,
       -1.020900e+03, 3.950372e+00}};
  static constexpr double hi[9][7] = {
      {2.991423e+00, 7.000644e-04, -5.633829e-08, -9.231578e-12, 1.582752e-15,
       -8.350340e+02, -1.355110e+00},
      {3.697578e+00, 6.135197e-04, -1.258842e-07, 1.775281e-11, -1.136435e-15,
       -1.233930e+03, 3.189166e+00},
      {2.672146e+00, 3.056293e-03, -8.730260e-07, 1.200996e-10, -6.391618e-15,
       -2.989921e+04, 6.862817e+00},
      {2.500000e+00, 0.0, 0.0, 0.0, 0.0, 2.547163e+04, -4.601176e-01},
      {2.542060e+00, -2.755062e-05, -3.102803e-09, 4.551067e-12, -4.368052e-16,
       2.923080e+04, 4.920308e+00},
      {2.86472886e+00, 1.05650448e-03, -2.59082758e-07, 3.05218674e-11,
       -1.33195876e-15, 3.68362875e+03, 5.70164073e+00},
      {4.01721090e+00, 2.23982013e-03, -6.33658150e-07, 1.14246370e-10,
       -1.07908535e-14, 1.11856713e+02, 3.78510215e+00},
      {4.573167e+00, 4.336136e-03, -1.474689e-06, 2.348904e-10, -1.431654e-14,
       -1.800696e+04, 5.011370e-01},
      {2.926640e+00, 1.487977e-03, -5.684761e-07, 1.009704e-10, -6.753351e-15,
       -9.227977e+02, 5.980528e+00}};
};

struct LiDryer {
  static constexpr int NS = 9;
  using D = LiDryerData;
  static constexpr double TMID = 1000.0;

  static double imw(int k) { return 1.0 / D::mw[k]; }
  static const double* coef(int k, double T) {
    return (T < TMID) ? D::lo[k] : D::hi[k];
  }
  // Species k's polynomials with the coefficients scaled by RU / W_k:
  //   cv_k(T) = v0 + v1 T + ... + v4 T^4,  v = RU/W (a0 - 1, a1, a2, a3, a4)
  //   e_k(T)  = u0 + u1 T + ... + u5 T^5,  u = RU/W (a5, a0 - 1, a1/2, a2/3,
  //             a3/4, a4/5)
  // i.e. e_k = RT (a0 - 1 + a1 T/2 + ... + a4 T^4/5 + a5/T) multiplied
  // through by T; both by Horner's rule in fused multiply-adds.
  static void ucoef(const double* a, int k, double* u) {
    const double s = RU / D::mw[k];
    u[0] = a[5] * s;
    u[1] = (a[0] - 1.0) * s;
    u[2] = a[1] / 2.0 * s;
    u[3] = a[2] / 3.0 * s;
    u[4] = a[3] / 4.0 * s;
    u[5] = a[4] / 5.0 * s;
  }
  static double e_k(int k, double T) {
    double u[6];
    ucoef(coef(k, T), k, u);
    return std::fma(T, std::fma(T, std::fma(T, std::fma(T, std::fma(T, u[5], u[4]), u[3]), u[2]), u[1]), u[0]);
  }
  static double cv_k(int k, double T) {
    const double* a = coef(k, T);
    const double s = RU / D::mw[k];
    const double v0 = (a[0] - 1.0) * s, v1 = a[1] * s, v2 = a[2] * s,
                 v3 = a[3] * s, v4 = a[4] * s;
    return std::fma(T, std::fma(T, std::fma(T, std::fma(T, v4, v3), v2), v1), v0);
  }
  // h_k / (R T)  (for tests: enthalpy balances)
  static void hRT(double T, double* h) {
    const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, invT = 1.0 / T;
    for (int k = 0; k < NS; ++k) {
      const double* a = coef(k, T);
      h[k] = a[0] + (a[1] / 2.0) * T + (a[2] / 3.0) * T2 + (a[3] / 4.0) * T3 +
             (a[4] / 5.0) * T4 + a[5] * invT;
    }
  }
  // species internal energies, erg/g
  static void CKUMS(double T, double* ei) {
    for (int k = 0; k < NS; ++k) ei[k] = e_k(k, T);
  }
  static void CKUBMS(double T, const double* Y, double& E) {
    E = 0.0;
    for (int k = 0; k < NS; ++k) E = std::fma(Y[k], e_k(k, T), E);
  }
  static void CKCVBS(double T, const double* Y, double& cv) {
    cv = 0.0;
    for (int k = 0; k < NS; ++k) cv = std::fma(Y[k], cv_k(k, T), cv);
  }
  static double sum_y_imw(const double* Y) {
    double s = 0.0;
    for (int k = 0; k < NS; ++k) s = std::fma(Y[k], imw(k), s);
    return s;
  }

  // Mass-weighted mixture polynomial e(T) = sum_m U[m] T^m of one
  // composition in each NASA range (cv = de/dT); the state functions of a
  // known composition (RTY2*, EY2T) evaluate these.
  struct Mix {
    double lo[6], hi[6], sy;
    explicit Mix(const double* Y) {
      for (int m = 0; m < 6; ++m) lo[m] = hi[m] = 0.0;
      for (int k = 0; k < NS; ++k) {
        double ul[6], uh[6];
        ucoef(D::lo[k], k, ul);
        ucoef(D::hi[k], k, uh);
        for (int m = 0; m < 6; ++m) {
          lo[m] = std::fma(Y[k], ul[m], lo[m]);
          hi[m] = std::fma(Y[k], uh[m], hi[m]);
        }
      }
      sy = sum_y_imw(Y);
    }
    static double pe(const double* U, double T) {
      return std::fma(T, std::fma(T, std::fma(T, std::fma(T, std::fma(T, U[5], U[4]), U[3]), U[2]), U[1]), U[0]);
    }
    static double pcv(const double* U, double T) {
      return std::fma(T, std::fma(T, std::fma(T, std::fma(T, 5.0 * U[5], 4.0 * U[4]), 3.0 * U[3]), 2.0 * U[2]), U[1]);
    }
    double e(double T) const { return (T < TMID) ? pe(lo, T) : pe(hi, T); }
    double cv(double T) const { return (T < TMID) ? pcv(lo, T) : pcv(hi, T); }
  };

  static void Y2WBAR(const double* Y, double& wbar) { wbar = 1.0 / sum_y_imw(Y); }
  static void RYP2E(double R, const double* Y, double P, double& E) {
    double wbar;
    Y2WBAR(Y, wbar);
    const double T = P * wbar / (R * RU);
    CKUBMS(T, Y, E);
  }
  static void RPY2Cs(double R, double P, const double* Y, double& Cs) {
    double wbar;
    Y2WBAR(Y, wbar);
    const double T = P * wbar / (R * RU);
    double cv;
    CKCVBS(T, Y, cv);
    const double G = (wbar * cv + RU) / (wbar * cv);
    Cs = std::sqrt(G * P / R);
  }
  static void RTY2P(double R, double T, const double* Y, double& P) {
    P = R * RU * T * sum_y_imw(Y);
  }
  static void RTY2G(double /*R*/, double T, const double* Y, double& G) {
    const Mix m(Y);
    const double wbar = 1.0 / m.sy;
    const double cv = m.cv(T);
    G = (wbar * cv + RU) / (wbar * cv);
  }
  static void RTY2Cs(double R, double T, const double* Y, double& Cs) {
    double P, G;
    RTY2P(R, T, Y, P);
    RTY2G(R, T, Y, G);
    Cs = std::sqrt(G * P / R);
  }
  static void RTY2dpde_dpdre(double R, double T, const double* Y, double& dpde,
                             double& dpdr_e) {
    const Mix m(Y);
    const double wbar = 1.0 / m.sy;
    dpde = R * RU / (wbar * m.cv(T));
    dpdr_e = RU * T / wbar;
  }
  static void T2Ei(double T, double* Ei) { CKUMS(T, Ei); }
  static void TY2E(double T, const double* Y, double& E) { CKUBMS(T, Y, E); }
  static void EY2T(double E, const double* Y, double& T) {
    const int maxiter = 200;
    const double tol = 1.0e-12;
    const double tmin = 90.0, tmax = 4000.0;
    const Mix m(Y);
    const double emin = Mix::pe(m.lo, tmin);
    const double emax = Mix::pe(m.hi, tmax);
    if (E < emin) {
      T = tmin - (emin - E) / Mix::pcv(m.lo, tmin);
      return;
    }
    if (E > emax) {
      T = tmax - (emax - E) / Mix::pcv(m.hi, tmax);
      return;
    }
    double t1 = T;
    if (t1 < tmin || t1 > tmax)
      t1 = tmin + (tmax - tmin) / (emax - emin) * (E - emin);
    for (int it = 0; it < maxiter; ++it) {
      double dt = (E - m.e(t1)) / m.cv(t1);
      if (dt > 100.0) {
        dt = 100.0;
      } else if (dt < -100.0) {
        dt = -100.0;
      } else if (std::fabs(dt) < tol) {
        break;
      } else if (t1 + dt == t1) {
        break;
      }
      t1 += dt;
    }
    T = t1;
  }
  static void REY2T(double /*R*/, double E, const double* Y, double& T) {
    EY2T(E, Y, T);
  }
};

}  // namespace orc
