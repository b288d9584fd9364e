// TEST INFRASTRUCTURE ONLY -- C entry points of the CPU oracle (see
// pc_oracle.h).  Used by tests/, the smoke check and the bench's cpu_baseline
// leg; never linked into the product.
#include "pc_oracle.h"

#include <cstdio>
#include <cstdlib>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_hydro.hpp"

using namespace orc;

namespace {

A4 to_a4(const pc_array4_t& v) {
  A4 a;
  a.p = v.p;
  a.js = v.jstride;
  a.ks = v.kstride;
  a.ns = v.nstride;
  for (int d = 0; d < 3; ++d) { a.b[d] = v.begin[d]; a.e[d] = v.end[d]; }
  a.ncomp = v.ncomp;
  return a;
}
Box to_box(const pc_box_t& b) {
  Box r;
  for (int d = 0; d < 3; ++d) { r.lo[d] = b.lo[d]; r.hi[d] = b.hi[d]; }
  return r;
}
bool covers(const pc_array4_t& v, const Box& b) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < v.begin[d] || b.hi[d] >= v.end[d]) return false;
  return true;
}

template <class EOS>
int hydro_source_box_t(const Box& bx, const int* domlo, const int* domhi,
                       const int* bclo, const int* bchi, const A4& sarr,
                       const A4* src, const A4& hyd, const double* dx,
                       double dt, const pc_params_t& prm, const A4* flx,
                       const A4* a, const A4& vol) {
  using H = Hydro<EOS>;
  const Box qbx = grow(bx, 4);
  Fab q(qbx, H::QVAR), qaux(qbx, NQAUX), srcq(qbx, H::QVAR);
  ParallelFor(qbx, [&](int i, int j, int k) { H::ctoprim(i, j, k, sarr, q.a, qaux.a); });
  Fab zero(qbx, H::NVAR);
  std::fill(zero.v.begin(), zero.v.end(), 0.0);
  const A4& s = src ? *src : zero.a;
  ParallelFor(qbx, [&](int i, int j, int k) { H::srctoprim(i, j, k, q.a, qaux.a, s, srcq.a); });
  for (int d = 0; d < 3; ++d) {  // flux[dir].setVal(0.0) (Source/Hydro.cpp:88)
    ParallelFor(surroundingNodes(bx, d), [&](int i, int j, int k) {
      for (int n = 0; n < H::NVAR; ++n) flx[d](i, j, k, n) = 0.0;
    });
  }
  return H::umdrv(bx, domlo, domhi, bclo, bchi, sarr, hyd, q.a, qaux.a, srcq.a,
                  dx, dt, prm.ppm_type, prm.plm_iorder, prm.use_flattening != 0,
                  prm.difmag, flx, a, vol);
}

// ------------------------------------------------------------- level (host)
struct HostLevel {
  int nvar;
  std::vector<Box> boxes;
  std::vector<long long> off;
  Box dom;
  int bclo[3], bchi[3];
  double dx[3];
  bool self;
  std::vector<int> cellmap;  // domain cell -> box id (domain mode)
  long long total = 0;
  A4 fab(double* arena, int b) const {
    return make_a4(arena + off[b], grow(boxes[b], 4), nvar);
  }
};

HostLevel make_level(int nvar, const pc_level_desc_t* d) {
  HostLevel L;
  L.nvar = nvar;
  L.self = d->periodic_self != 0;
  for (int dd = 0; dd < 3; ++dd) {
    L.dom.lo[dd] = d->domlo[dd];
    L.dom.hi[dd] = d->domhi[dd];
    L.bclo[dd] = d->bclo[dd];
    L.bchi[dd] = d->bchi[dd];
    L.dx[dd] = d->dx[dd];
  }
  long long o = 0;
  for (int b = 0; b < d->nboxes; ++b) {
    L.boxes.push_back(to_box(d->boxes[b]));
    L.off.push_back(o);
    o += grow(L.boxes.back(), 4).npts() * nvar;
  }
  L.total = o;
  return L;
}

void build_cellmap(HostLevel& L) {
  if (L.self || !L.cellmap.empty()) return;
  const Box& D = L.dom;
  const long nx = D.hi[0] - D.lo[0] + 1, ny = D.hi[1] - D.lo[1] + 1;
  L.cellmap.assign(D.npts(), -1);
  for (int b = 0; b < (int)L.boxes.size(); ++b)
    ParallelFor(L.boxes[b], [&](int i, int j, int k) {
      L.cellmap[(i - D.lo[0]) + nx * ((j - D.lo[1]) + ny * (long)(k - D.lo[2]))] = b;
    });
}

// Map index x in direction d of a ghost cell to its source index; sets odd
// when the reflecting side flips normal (or, for no-slip, all) momenta.
int map_index(int x, int lo, int hi, int bcl, int bch, int* refl) {
  *refl = 0;
  if (x >= lo && x <= hi) return x;
  const int L = hi - lo + 1;
  const int bc = (x < lo) ? bcl : bch;
  if (bc == BC_INTERIOR) return lo + (((x - lo) % L) + L) % L;
  if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) {
    *refl = (bc == BC_NOSLIPWALL) ? 2 : 1;
    return (x < lo) ? 2 * lo - 1 - x : 2 * hi + 1 - x;
  }
  return (x < lo) ? lo : hi;  // outflow (FOExtrap); inflow/user as FOExtrap
}

void fill_ghosts(HostLevel& L, double* arena) {
  build_cellmap(L);
  const int nb = (int)L.boxes.size();
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nb; ++b) {
    const Box& vb = L.boxes[b];
    const A4 dst = L.fab(arena, b);
    const Box& D = L.self ? vb : L.dom;
    int bl[3], bh[3];
    for (int d = 0; d < 3; ++d) {
      bl[d] = L.self ? BC_INTERIOR : L.bclo[d];
      bh[d] = L.self ? BC_INTERIOR : L.bchi[d];
    }
    const long nx = D.hi[0] - D.lo[0] + 1, ny = D.hi[1] - D.lo[1] + 1;
    ParallelFor(grow(vb, 4), [&](int i, int j, int k) {
      if (i >= vb.lo[0] && i <= vb.hi[0] && j >= vb.lo[1] && j <= vb.hi[1] &&
          k >= vb.lo[2] && k <= vb.hi[2])
        return;
      int r[3];
      const int s[3] = {map_index(i, D.lo[0], D.hi[0], bl[0], bh[0], &r[0]),
                        map_index(j, D.lo[1], D.hi[1], bl[1], bh[1], &r[1]),
                        map_index(k, D.lo[2], D.hi[2], bl[2], bh[2], &r[2])};
      int sb = b;
      if (!L.self)
        sb = L.cellmap[(s[0] - D.lo[0]) + nx * ((s[1] - D.lo[1]) + ny * (long)(s[2] - D.lo[2]))];
      const A4 src = L.fab(arena, sb);
      for (int n = 0; n < L.nvar; ++n) {
        double sg = 1.0;
        for (int d = 0; d < 3; ++d) {
          if (r[d] == 0) continue;
          const bool mom = (n == UMX || n == UMY || n == UMZ);
          if ((r[d] == 2 && mom) || (r[d] == 1 && n == UMX + d)) sg = -sg;
        }
        dst(i, j, k, n) = sg * src(s[0], s[1], s[2], n);
      }
    });
  }
}

template <class EOS>
double estdt_level(const HostLevel& L, const pc_params_t& prm,
                   const double* arena, const double* dx) {
  // Source/PeleC.cpp:829-961
  if (prm.fixed_dt > 0.0) return prm.fixed_dt;
  double estdt = prm.max_dt;
  double estdt_hydro = prm.max_dt / prm.cfl;
  double dtc = std::numeric_limits<double>::max();
  for (int b = 0; b < (int)L.boxes.size(); ++b) {
    const A4 u = L.fab(const_cast<double*>(arena), b);
    dtc = mn(dtc, Hydro<EOS>::estdt_hydro(L.boxes[b], u, dx[0], dx[1], dx[2]));
  }
  estdt_hydro = mn(estdt_hydro, dtc);
  estdt_hydro *= prm.cfl;
  if (estdt_hydro < estdt) estdt = estdt_hydro;
  return estdt;
}

// The stop_time limit of computeNewDt / computeInitialDt (Source/PeleC.cpp:
// 1018-1027, 1061-1068): a step that would end within 0.001 dt of stop_time
// (or past it) is shortened to land on it.  stop_time < 0: no limit.
inline double clamp_stop_time(double dt0, double cur_time, double stop_time) {
  const double dt_eps = 0.001 * dt0;
  if (stop_time >= 0.0) {
    if ((cur_time + dt0) > (stop_time - dt_eps)) dt0 = stop_time - cur_time;
  }
  return dt0;
}

// computeNewDt (Source/PeleC.cpp:964-1034) for a single level after a step
// of dt that ended at cur_time.
template <class EOS>
double new_dt(const HostLevel& L, const pc_params_t& prm, const double* arena,
              const double* dx, double dt, double cur_time) {
  double dtn = estdt_level<EOS>(L, prm, arena, dx);
  if (prm.fixed_dt <= 0.0) dtn = mn(dtn, prm.change_max * dt);
  return clamp_stop_time(dtn, cur_time, prm.stop_time);
}

// Uniform-geometry face areas and cell volumes of a box (Cartesian), the
// areas over the faces of grow(bx,3) that the MOL flux loop visits (the
// reference's area MultiFab carries ghost cells).
struct BoxGeom {
  Fab a0, a1, a2, vol;
  BoxGeom(const Box& bx, const double* dx)
      : a0(surroundingNodes(grow(bx, 3), 0), 1),
        a1(surroundingNodes(grow(bx, 3), 1), 1),
        a2(surroundingNodes(grow(bx, 3), 2), 1), vol(bx, 1) {
    std::fill(a0.v.begin(), a0.v.end(), dx[1] * dx[2]);
    std::fill(a1.v.begin(), a1.v.end(), dx[2] * dx[0]);
    std::fill(a2.v.begin(), a2.v.end(), dx[0] * dx[1]);
    std::fill(vol.v.begin(), vol.v.end(), dx[0] * dx[1] * dx[2]);
  }
};

// do_mol_advance (Source/Advance.cpp:49-227) for a hydro-only single level
// (no reactions or other sources, mol_iters = 1): MOL source at t^n,
// U* = 1.0 U^n + dt S^n, computeTemp; MOL source of U* (ghosts filled again),
// U^{n+1} = 0.5 U* + 0.5 U^n, += (0.5 dt) S*, computeTemp; then
// post_timestep's computeTemp (Source/PeleC.cpp:1108-1110) and computeNewDt.
template <class EOS>
int mol_level_step_t(HostLevel& L, const pc_params_t& prm, double* arena,
                     int nsteps, double* dtp, double* timep) {
  using H = Hydro<EOS>;
  const int nb = (int)L.boxes.size();
  const int iorder = prm.plm_iorder;  // Source/Diffusion.cpp:371-373
  const bool laxf = prm.use_laxf_flux != 0;
  double dt = *dtp;
  std::vector<std::vector<double>> sold(nb);
  for (int s = 0; s < nsteps; ++s) {
    for (int stage = 0; stage < 2; ++stage) {
      fill_ghosts(L, arena);
#pragma omp parallel for schedule(dynamic)
      for (int b = 0; b < nb; ++b) {
        const Box bx = L.boxes[b];
        const A4 S = L.fab(arena, b);
        BoxGeom g(bx, L.dx);
        const A4 a[3] = {g.a0.a, g.a1.a, g.a2.a};
        Fab src(bx, H::NVAR);
        H::mol_src(bx, S, src.a, a, g.vol.a, iorder, laxf);
        if (stage == 0) {
          sold[b].resize((size_t)bx.npts() * H::NVAR);
          const A4 O = make_a4(sold[b].data(), bx, H::NVAR);
          ParallelFor(bx, [&](int i, int j, int k) {
            for (int n = 0; n < H::NVAR; ++n) {
              O(i, j, k, n) = S(i, j, k, n);
              S(i, j, k, n) = 1.0 * S(i, j, k, n) + dt * src.a(i, j, k, n);
            }
            H::rst_int_e_default(i, j, k, S);
            H::cmpTemp(i, j, k, S);
          });
        } else {
          const A4 O = make_a4(sold[b].data(), bx, H::NVAR);
          ParallelFor(bx, [&](int i, int j, int k) {
            for (int n = 0; n < H::NVAR; ++n) {
              S(i, j, k, n) = 0.5 * S(i, j, k, n) + 0.5 * O(i, j, k, n);
              S(i, j, k, n) += (0.5 * dt) * src.a(i, j, k, n);
            }
            H::rst_int_e_default(i, j, k, S);
            H::cmpTemp(i, j, k, S);
            H::rst_int_e_default(i, j, k, S);
            H::cmpTemp(i, j, k, S);
          });
        }
      }
    }
    *timep += dt;
    dt = new_dt<EOS>(L, prm, arena, L.dx, dt, *timep);
  }
  *dtp = dt;
  return 0;
}

template <class EOS>
int level_step_t(const pc_level_desc_t* desc, const pc_params_t& prm,
                 double* arena, int nsteps, double* dtp, double* timep,
                 int nthreads) {
  using H = Hydro<EOS>;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  HostLevel L = make_level(H::NVAR, desc);
  const double* dx = desc->dx;
  const double vol_v = dx[0] * dx[1] * dx[2];
  const double area_v[3] = {dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1]};
  int dlo[3], dhi[3], blo[3], bhi[3];
  for (int d = 0; d < 3; ++d) {
    dlo[d] = desc->domlo[d]; dhi[d] = desc->domhi[d];
    blo[d] = desc->bclo[d]; bhi[d] = desc->bchi[d];
  }
  double dt = *dtp;
  if (prm.do_mol) return mol_level_step_t<EOS>(L, prm, arena, nsteps, dtp, timep);
  // The reference's CPU path works on tiles (MFIter with TilingIfNotGPU,
  // Source/Hydro.cpp:60-73; AMReX's default CPU tile is the whole x extent
  // by 8 by 8): the hydro source of every tile is formed from the old state
  // (each tile's primitives on grow(tile,4) of its box's FAB), then every
  // box is updated.  Each cell's source depends only on its stencil, so the
  // tiling changes no value; it spreads a box over threads.
  const int nb = (int)L.boxes.size();
  std::vector<std::pair<int, Box>> tiles;
  for (int b = 0; b < nb; ++b) {
    const Box& bx = L.boxes[b];
    for (int k0 = bx.lo[2]; k0 <= bx.hi[2]; k0 += 8)
      for (int j0 = bx.lo[1]; j0 <= bx.hi[1]; j0 += 8) {
        Box t = bx;
        t.lo[1] = j0;
        t.hi[1] = std::min(j0 + 7, bx.hi[1]);
        t.lo[2] = k0;
        t.hi[2] = std::min(k0 + 7, bx.hi[2]);
        tiles.push_back({b, t});
      }
  }
  std::vector<Fab> hsrcs;
  hsrcs.reserve(nb);
  for (int b = 0; b < nb; ++b) hsrcs.emplace_back(L.boxes[b], H::NVAR);
  for (int s = 0; s < nsteps; ++s) {
    fill_ghosts(L, arena);
    int err = 0;
    const int nt = (int)tiles.size();
#pragma omp parallel for schedule(dynamic) reduction(min : err)
    for (int t = 0; t < nt; ++t) {
      const int b = tiles[t].first;
      const Box bx = tiles[t].second;
      const Box& box = L.boxes[b];
      int sdlo[3], sdhi[3], sblo[3], sbhi[3];
      for (int d = 0; d < 3; ++d) {
        sdlo[d] = L.self ? box.lo[d] : dlo[d];
        sdhi[d] = L.self ? box.hi[d] : dhi[d];
        sblo[d] = L.self ? BC_INTERIOR : blo[d];
        sbhi[d] = L.self ? BC_INTERIOR : bhi[d];
      }
      const A4 S = L.fab(arena, b);
      const A4& hs = hsrcs[b].a;
      ParallelFor(bx, [&](int i, int j, int k) {  // hydro_source.setVal(0)
        for (int n = 0; n < H::NVAR; ++n) hs(i, j, k, n) = 0.0;
      });
      Fab f0(surroundingNodes(bx, 0), H::NVAR), f1(surroundingNodes(bx, 1), H::NVAR),
          f2(surroundingNodes(bx, 2), H::NVAR);
      const A4 flx[3] = {f0.a, f1.a, f2.a};
      Fab a0(surroundingNodes(bx, 0), 1), a1(surroundingNodes(bx, 1), 1),
          a2(surroundingNodes(bx, 2), 1), vol(bx, 1);
      std::fill(a0.v.begin(), a0.v.end(), area_v[0]);
      std::fill(a1.v.begin(), a1.v.end(), area_v[1]);
      std::fill(a2.v.begin(), a2.v.end(), area_v[2]);
      std::fill(vol.v.begin(), vol.v.end(), vol_v);
      const A4 a[3] = {a0.a, a1.a, a2.a};
      int rc = hydro_source_box_t<EOS>(bx, sdlo, sdhi, sblo, sbhi, S, nullptr,
                                       hs, dx, dt, prm, flx, a, vol.a);
      if (rc) err = rc;
    }
    if (err) return err;
    // construct_Snew (Source/Advance.cpp:397-416) + computeTemp twice
    // (Source/Advance.cpp:352,388) + post_timestep computeTemp
    // (Source/PeleC.cpp:1108-1110).  The second construct_Snew rebuilds
    // the same S_new from S_old, so the sequence reduces to one update,
    // reset + e->T from the old T, then reset + e->T from that T.
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < nt; ++t) {
      const int b = tiles[t].first;
      const A4 S = L.fab(arena, b);
      const A4& hs = hsrcs[b].a;
      ParallelFor(tiles[t].second, [&](int i, int j, int k) {
        for (int n = 0; n < H::NVAR; ++n) S(i, j, k, n) += dt * hs(i, j, k, n);
        H::rst_int_e_default(i, j, k, S);
        H::cmpTemp(i, j, k, S);
        H::rst_int_e_default(i, j, k, S);
        H::cmpTemp(i, j, k, S);
      });
    }
    // computeNewDt (Source/PeleC.cpp:964-1034), single level, at the time
    // the step ended
    *timep += dt;
    dt = new_dt<EOS>(L, prm, arena, dx, dt, *timep);
  }
  *dtp = dt;
  return 0;
}

// pc_initdata restatement (Exec/RegTests/Sod/prob.H:25-117 with
// Exec/RegTests/Sod/prob.cpp:215-241 for mode 0; the PMF pattern
// PYT2R + RTY2E for mode 1).  prim comps: mode 0: rho, p, u, v, w, Y;
// mode 1: p, T, u, v, w, Y.
template <class EOS>
int prim_to_cons_t(const Box& bx, const A4& P, int mode, const A4& U) {
  constexpr int NS = EOS::NS;
  ParallelFor(bx, [&](int i, int j, int k) {
    double Y[NS];
    for (int n = 0; n < NS; ++n) Y[n] = P(i, j, k, 5 + n);
    const double u = P(i, j, k, 2), v = P(i, j, k, 3), w = P(i, j, k, 4);
    double rho, e, T;
    if (mode == 0) {
      rho = P(i, j, k, 0);
      const double p = P(i, j, k, 1);
      EOS::RYP2E(rho, Y, p, e);
      T = 1.0;
      EOS::EY2T(e, Y, T);
    } else {
      const double p = P(i, j, k, 0);
      T = P(i, j, k, 1);
      double wbar;
      EOS::Y2WBAR(Y, wbar);
      rho = p * wbar / (RU * T);
      EOS::TY2E(T, Y, e);
    }
    const double rhoe = rho * e;
    U(i, j, k, URHO) = rho;
    U(i, j, k, UMX) = rho * u;
    U(i, j, k, UMY) = rho * v;
    U(i, j, k, UMZ) = rho * w;
    U(i, j, k, UEINT) = rhoe;
    U(i, j, k, UEDEN) = rhoe + 0.5 * rho * (u * u + v * v + w * w);
    U(i, j, k, UTEMP) = T;
    for (int n = 0; n < NS; ++n) U(i, j, k, UFS + n) = rho * Y[n];
  });
  return 0;
}

template <class EOS>
int eos_point(int which, double a, double b, double c, const double* Y,
              double* out) {
  constexpr int NS = EOS::NS;
  double y[NS];
  for (int k = 0; k < NS; ++k) y[k] = Y ? Y[k] : (k == 0 ? 1.0 : 0.0);
  switch (which) {
    case 0: EOS::RYP2E(a, y, b, out[0]); return 0;
    case 1: EOS::RPY2Cs(a, b, y, out[0]); return 0;
    case 2: { double T = c; EOS::REY2T(a, b, y, T); out[0] = T; return 0; }
    case 3: EOS::RTY2P(a, b, y, out[0]); return 0;
    case 4: EOS::RTY2Cs(a, b, y, out[0]); return 0;
    case 5: EOS::RTY2G(a, b, y, out[0]); return 0;
    case 6: EOS::Y2WBAR(y, out[0]); return 0;
    case 7: EOS::T2Ei(a, out); return 0;
    case 8:
      if constexpr (NS == 9) { LiDryer::CKCVBS(a, y, out[0]); return 0; }
      else { out[0] = RU * GammaLaw::imw[0] / (GammaLaw::gamma - 1.0); return 0; }
    case 9:
      if constexpr (NS == 9) { LiDryer::hRT(a, out); return 0; }
      else { out[0] = GammaLaw::gamma / (GammaLaw::gamma - 1.0); return 0; }
    case 10: EOS::RTY2dpde_dpdre(a, b, y, out[0], out[1]); return 0;
    case 11:  // cv of the composition from mixture coefficients
      if constexpr (NS == 9) { out[0] = LiDryer::Mix(y).cv(a); return 0; }
      else { out[0] = RU * GammaLaw::imw[0] / (GammaLaw::gamma - 1.0); return 0; }
  }
  return PC_ERR_ARG;
}

template <class EOS>
int riemann_point(const double* L5, const double* spl, const double* R5,
                  const double* spr, double cav, double* o) {
  double fy[EOS::NS];
  Hydro<EOS>::riemann(L5[0], L5[1], L5[2], L5[3], L5[4], spl, R5[0], R5[1],
                      R5[2], R5[3], R5[4], spr, 1, cav, o[0], o[1], fy, o[2],
                      o[3], o[4], o[5], o[6], o[7], o[8], o[9], o[10], o[11]);
  return 0;
}

// pc_transdo (Source/Godunov.H:455-633) on the cell (0,0,0) of a 2^3 patch:
// qplus = qnormp(cell), qminus = qnormm(cell + e_dir); the transverse-t
// fluxes and Godunov states on the cell's low (fl_lo, qi_lo) and high t-faces.
template <class EOS>
int transdo_point(int dir, int t, const double* qplus, const double* qminus,
                  const double* fl_lo, const double* fl_hi, const double* qi_lo,
                  const double* qi_hi, double gamc, double cdtdx,
                  double* out_p, double* out_m) {
  using H = Hydro<EOS>;
  const Box b2{{0, 0, 0}, {1, 1, 1}};
  Fab qm(b2, H::QVAR), qp(b2, H::QVAR), qnm(b2, H::QVAR), qnp(b2, H::QVAR),
      fl(b2, H::NVAR), qa(b2, NQAUX), qi(b2, NGDNV);
  const IV c = ivec(0, 0, 0), cn = shift(c, dir, 1), ct = shift(c, t, 1);
  for (int n = 0; n < H::QVAR; ++n) {
    at(qnp.a, c, n) = qplus[n];
    at(qnm.a, cn, n) = qminus[n];
  }
  for (int n = 0; n < H::NVAR; ++n) {
    at(fl.a, c, n) = fl_lo[n];
    at(fl.a, ct, n) = fl_hi[n];
  }
  for (int n = 0; n < NGDNV; ++n) {
    at(qi.a, c, n) = qi_lo[n];
    at(qi.a, ct, n) = qi_hi[n];
  }
  at(qa.a, c, QGAMC) = gamc;
  H::transdo(0, 0, 0, dir, t, qm.a, qp.a, qnm.a, qnp.a, fl.a, qa.a, qi.a, cdtdx);
  for (int n = 0; n < H::QVAR; ++n) {
    out_p[n] = at(qp.a, c, n);
    out_m[n] = at(qm.a, cn, n);
  }
  return 0;
}

// pc_transdd (Source/Godunov.H:639-845) on the cell (0,0,0): fl0/qi0 on the
// low and high faces of the first transverse direction (y for dir x, x
// otherwise), fl1/qi1 of the second (z for dir x or y, y for dir z).
template <class EOS>
int transdd_point(int dir, const double* qplus, const double* qminus,
                  const double* fl0_lo, const double* fl0_hi,
                  const double* fl1_lo, const double* fl1_hi,
                  const double* qi0_lo, const double* qi0_hi,
                  const double* qi1_lo, const double* qi1_hi, double gamc,
                  const double* srcq, double hdt, double c0, double c1,
                  double* out_p, double* out_m) {
  using H = Hydro<EOS>;
  const Box b2{{0, 0, 0}, {1, 1, 1}};
  Fab qm(b2, H::QVAR), qp(b2, H::QVAR), qnm(b2, H::QVAR), qnp(b2, H::QVAR),
      f0(b2, H::NVAR), f1(b2, H::NVAR), qa(b2, NQAUX), q0(b2, NGDNV),
      q1(b2, NGDNV), sq(b2, H::QVAR);
  const int l1 = (dir == 0) ? 1 : 0, l2 = (dir == 2) ? 1 : 2;
  const IV c = ivec(0, 0, 0), cn = shift(c, dir, 1), c0p = shift(c, l1, 1),
           c1p = shift(c, l2, 1);
  for (int n = 0; n < H::QVAR; ++n) {
    at(qnp.a, c, n) = qplus[n];
    at(qnm.a, cn, n) = qminus[n];
    at(sq.a, c, n) = srcq ? srcq[n] : 0.0;
  }
  for (int n = 0; n < H::NVAR; ++n) {
    at(f0.a, c, n) = fl0_lo[n];
    at(f0.a, c0p, n) = fl0_hi[n];
    at(f1.a, c, n) = fl1_lo[n];
    at(f1.a, c1p, n) = fl1_hi[n];
  }
  for (int n = 0; n < NGDNV; ++n) {
    at(q0.a, c, n) = qi0_lo[n];
    at(q0.a, c0p, n) = qi0_hi[n];
    at(q1.a, c, n) = qi1_lo[n];
    at(q1.a, c1p, n) = qi1_hi[n];
  }
  at(qa.a, c, QGAMC) = gamc;
  H::transdd(0, 0, 0, dir, qm.a, qp.a, qnm.a, qnp.a, f0.a, f1.a, q0.a, q1.a,
             qa.a, sq.a, hdt, c0, c1);
  for (int n = 0; n < H::QVAR; ++n) {
    out_p[n] = at(qp.a, c, n);
    out_m[n] = at(qm.a, cn, n);
  }
  return 0;
}

template <class EOS>
int laxf_point(const double* L5, const double* spl, const double* R5,
               const double* spr, double* o) {
  Hydro<EOS>::laxfriedrich_flux(L5[0], L5[1], L5[2], L5[3], L5[4], spl, R5[0],
                                R5[1], R5[2], R5[3], R5[4], spr, o[0], o[1],
                                o[2], o + 8, o[3], o[4], o[5], o[6], o[7]);
  return 0;
}

template <class EOS>
int mol_src_box_t(const Box& bx, const A4& S, const A4& src,
                  const pc_params_t& prm, const A4* flx, const A4* a,
                  const A4& vol) {
  // The flux loop visits the faces of grow(bx,3); only those of bx enter
  // the result.  The caller's areas cover surroundingNodes(bx,d): extend
  // them with NaN, so a flux outside bx could not reach the output unseen.
  Fab ext[3] = {Fab(surroundingNodes(grow(bx, 3), 0), 1),
                Fab(surroundingNodes(grow(bx, 3), 1), 1),
                Fab(surroundingNodes(grow(bx, 3), 2), 1)};
  A4 ae[3];
  for (int d = 0; d < 3; ++d) {
    ae[d] = ext[d].a;
    ParallelFor(surroundingNodes(bx, d),
                [&](int i, int j, int k) { ae[d](i, j, k) = a[d](i, j, k); });
  }
  Hydro<EOS>::mol_src(bx, S, src, ae, vol, prm.plm_iorder,
                      prm.use_laxf_flux != 0, flx);
  return 0;
}

// Plot variables of Source/Derive.cpp on the valid cells of one box (U with
// at least one filled ghost cell for magvort).  codes: 0 state comp,
// 1 comp / rho (pc_derdividebyrho), 2 magvel (:10-31), 3 kineng (:109-128),
// 4 pressure (:531-555), 5 soundspeed (:446-469), 6 MachNumber (:498-529),
// 7 magvort (:201-271, regular cells), 8 logden (:181-198).
template <class EOS>
int derive_t(const Box& bx, const A4& U, const int* codes, const int* comps,
             int nplot, const A4& out, const double* dx) {
  constexpr int NS = EOS::NS;
  auto vel = [&](int i, int j, int k, int c) {
    const double rhoInv = 1.0 / U(i, j, k, URHO);
    return U(i, j, k, UMX + c) * rhoInv;
  };
  ParallelFor(bx, [&](int i, int j, int k) {
    const double rho = U(i, j, k, URHO);
    const double rhoInv = 1.0 / rho;
    const double T = U(i, j, k, UTEMP);
    double Y[NS];
    for (int n = 0; n < NS; ++n) Y[n] = U(i, j, k, UFS + n) * rhoInv;
    for (int m = 0; m < nplot; ++m) {
      const int c = comps[m];
      double v = 0.0;
      switch (codes[m]) {
        case 0: v = U(i, j, k, c); break;
        case 1: v = U(i, j, k, c) / U(i, j, k, URHO); break;
        case 2: {
          const double d1 = U(i, j, k, UMX) * rhoInv, d2 = U(i, j, k, UMY) * rhoInv,
                       d3 = U(i, j, k, UMZ) * rhoInv;
          v = std::sqrt((d1 * d1) + (d2 * d2) + (d3 * d3));
          break;
        }
        case 3: {
          const double x = U(i, j, k, UMX) * U(i, j, k, UMX),
                       y = U(i, j, k, UMY) * U(i, j, k, UMY),
                       z = U(i, j, k, UMZ) * U(i, j, k, UMZ);
          v = 0.5 / U(i, j, k, URHO) * (x + y + z);
          break;
        }
        case 4: EOS::RTY2P(rho, T, Y, v); break;
        case 5: EOS::RTY2Cs(rho, T, Y, v); break;
        case 6: {
          double cs;
          EOS::RTY2Cs(rho, T, Y, cs);
          const double x = U(i, j, k, UMX) * U(i, j, k, UMX),
                       y = U(i, j, k, UMY) * U(i, j, k, UMY),
                       z = U(i, j, k, UMZ) * U(i, j, k, UMZ);
          v = std::sqrt(x + y + z) / U(i, j, k, URHO) / cs;
          break;
        }
        case 7: {
          const double wi = 0.5, wj = 0.5, wk = 0.5;  // get_weight(i-1, i+1)
          const double vx = wi * (vel(i + 1, j, k, 1) - vel(i - 1, j, k, 1)) / dx[0];
          const double uy = wj * (vel(i, j + 1, k, 0) - vel(i, j - 1, k, 0)) / dx[1];
          const double v3 = vx - uy;
          const double wx = wi * (vel(i + 1, j, k, 2) - vel(i - 1, j, k, 2)) / dx[0];
          const double wy = wj * (vel(i, j + 1, k, 2) - vel(i, j - 1, k, 2)) / dx[1];
          const double uz = wk * (vel(i, j, k + 1, 0) - vel(i, j, k - 1, 0)) / dx[2];
          const double vz = wk * (vel(i, j, k + 1, 1) - vel(i, j, k - 1, 1)) / dx[2];
          const double v1 = wy - vz, v2 = uz - wx;
          v = std::sqrt(0. + v3 * v3 + v1 * v1 + v2 * v2);
          break;
        }
        case 8: v = std::log10(U(i, j, k, URHO)); break;
        default: return;
      }
      out(i, j, k, m) = v;
    }
  });
  return 0;
}

}  // namespace

#define DISPATCH(eos, call)                                   \
  do {                                                        \
    if ((eos) == PC_EOS_GAMMALAW) { using E = GammaLaw; call; } \
    else if ((eos) == PC_EOS_LIDRYER) { using E = LiDryer; call; } \
    else return PC_ERR_ARG;                                   \
  } while (0)

extern "C" {

void pc_ref_params_default(pc_params_t* p) {
  p->ppm_type = 0;
  p->plm_iorder = 4;
  p->use_flattening = 1;
  p->use_hybrid_weno = 0;
  p->weno_scheme = 1;
  p->difmag = 0.1;
  p->cfl = 0.8;
  p->init_shrink = 1.0;
  p->change_max = 1.1;
  p->max_dt = 1.e200;
  p->fixed_dt = -1.0;
  p->dt_cutoff = 0.0;
  p->stop_time = -1.0;
  p->allow_small_energy = 1;
  p->allow_negative_energy = 1;
  p->dual_energy_update_E_from_e = 1;
  p->dual_energy_eta2 = 1.0e-4;
  p->do_mol = 0;
  p->mol_iorder = 2;
  p->use_laxf_flux = 0;
}

int pc_ref_ctoprim(int eos, pc_box_t qbx, pc_array4_t u, pc_array4_t q,
                   pc_array4_t qaux) {
  const Box b = to_box(qbx);
  if (!covers(u, b) || !covers(q, b) || !covers(qaux, b)) return PC_ERR_ARG;
  const A4 U = to_a4(u), Q = to_a4(q), QA = to_a4(qaux);
  DISPATCH(eos, ParallelFor(b, [&](int i, int j, int k) { Hydro<E>::ctoprim(i, j, k, U, Q, QA); }));
  return 0;
}

int pc_ref_srctoprim(int eos, pc_box_t qbx, pc_array4_t q, pc_array4_t qaux,
                     pc_array4_t src, pc_array4_t srcq) {
  const Box b = to_box(qbx);
  const A4 Q = to_a4(q), QA = to_a4(qaux), S = to_a4(src), SQ = to_a4(srcq);
  DISPATCH(eos, ParallelFor(b, [&](int i, int j, int k) { Hydro<E>::srctoprim(i, j, k, Q, QA, S, SQ); }));
  return 0;
}

int pc_ref_umdrv(int eos, double /*time*/, pc_box_t bx, const int* domlo,
                 const int* domhi, const int* bclo, const int* bchi,
                 pc_array4_t uin, pc_array4_t uout, pc_array4_t q,
                 pc_array4_t qaux, pc_array4_t src_q, const double* dx,
                 double dt, int ppm_type, int plm_iorder, int use_flattening,
                 int use_hybrid_weno, int /*weno_scheme*/, double difmag,
                 const pc_array4_t* flx, const pc_array4_t* a, pc_array4_t vol,
                 double* /*cflLoc*/) {
  if (use_hybrid_weno) return PC_ERR_UNSUPPORTED;
  const Box b = to_box(bx);
  if (!covers(q, grow(b, 4)) || !covers(uin, grow(b, 4))) return PC_ERR_GROW;
  const A4 F[3] = {to_a4(flx[0]), to_a4(flx[1]), to_a4(flx[2])};
  const A4 AA[3] = {to_a4(a[0]), to_a4(a[1]), to_a4(a[2])};
  int rc = 0;
  std::vector<double> zero;
  A4 SQ = to_a4(src_q);
  if (!SQ.p) {
    const int qv = q.ncomp;
    zero.assign((size_t)grow(b, 4).npts() * qv, 0.0);
    SQ = make_a4(zero.data(), grow(b, 4), qv);
  }
  DISPATCH(eos, rc = Hydro<E>::umdrv(b, domlo, domhi, bclo, bchi, to_a4(uin), to_a4(uout),
                                     to_a4(q), to_a4(qaux), SQ, dx, dt, ppm_type,
                                     plm_iorder, use_flattening != 0, difmag, F, AA,
                                     to_a4(vol)));
  return rc;
}

int pc_ref_hydro_source_box(int eos, pc_box_t bx, const int* domlo,
                            const int* domhi, const int* bclo, const int* bchi,
                            pc_array4_t sarr, pc_array4_t src,
                            pc_array4_t hyd_src, const double* dx, double dt,
                            const pc_params_t* prm, const pc_array4_t* flx,
                            const pc_array4_t* a, pc_array4_t vol) {
  const Box b = to_box(bx);
  if (!covers(sarr, grow(b, 4))) return PC_ERR_GROW;
  const A4 F[3] = {to_a4(flx[0]), to_a4(flx[1]), to_a4(flx[2])};
  const A4 AA[3] = {to_a4(a[0]), to_a4(a[1]), to_a4(a[2])};
  const A4 S = to_a4(src);
  int rc = 0;
  DISPATCH(eos, rc = hydro_source_box_t<E>(b, domlo, domhi, bclo, bchi, to_a4(sarr),
                                           S.p ? &S : nullptr, to_a4(hyd_src), dx, dt,
                                           *prm, F, AA, to_a4(vol)));
  return rc;
}

long long pc_ref_level_arena_size(int eos, const pc_level_desc_t* desc,
                                  long long* offsets) {
  HostLevel L = make_level(7 + eos, desc);
  if (offsets)
    for (size_t b = 0; b < L.off.size(); ++b) offsets[b] = L.off[b];
  return L.total;
}

int pc_ref_level_fill_ghosts(int eos, const pc_level_desc_t* desc,
                             double* arena) {
  if (eos != 1 && eos != 9) return PC_ERR_ARG;
  HostLevel L = make_level(7 + eos, desc);
  fill_ghosts(L, arena);
  return 0;
}

int pc_ref_level_estdt(int eos, const pc_level_desc_t* desc,
                       const pc_params_t* prm, const double* arena,
                       double* dt_out) {
  HostLevel L = make_level(7 + eos, desc);
  DISPATCH(eos, *dt_out = estdt_level<E>(L, *prm, arena, desc->dx));
  return 0;
}

int pc_ref_level_init_dt(int eos, const pc_level_desc_t* desc,
                         const pc_params_t* prm, const double* arena,
                         double time, double* dt_out) {
  // computeInitialDt (Source/PeleC.cpp:1041-1075) -> initialTimeStep
  // (:811-826) for one level: init_shrink * estTimeStep, stop_time limit
  HostLevel L = make_level(7 + eos, desc);
  double est = 0.0;
  DISPATCH(eos, est = estdt_level<E>(L, *prm, arena, desc->dx));
  *dt_out = clamp_stop_time(prm->init_shrink * est, time, prm->stop_time);
  return 0;
}

int pc_ref_level_step(int eos, const pc_level_desc_t* desc,
                      const pc_params_t* prm, double* arena, int nsteps,
                      double* dt, double* time, int nthreads) {
  double t_local = 0.0;
  if (!time) time = &t_local;
  if (prm->use_hybrid_weno || !prm->allow_small_energy ||
      !prm->allow_negative_energy)
    return PC_ERR_UNSUPPORTED;
  // Source/PeleC.cpp:388-392
  if (prm->do_mol && prm->mol_iorder != 1 && prm->mol_iorder != 2)
    return PC_ERR_MOL_IORDER;
  int rc = 0;
  DISPATCH(eos, rc = level_step_t<E>(desc, *prm, arena, nsteps, dt, time, nthreads));
  return rc;
}

int pc_ref_prim_to_cons(int eos, pc_box_t bx, pc_array4_t prim, int mode,
                        pc_array4_t u) {
  if (mode != 0 && mode != 1) return PC_ERR_ARG;
  int rc = 0;
  DISPATCH(eos, rc = prim_to_cons_t<E>(to_box(bx), to_a4(prim), mode, to_a4(u)));
  return rc;
}

int pc_ref_eos(int eos, int which, double a, double b, double c,
               const double* Y, double* out) {
  int rc = 0;
  DISPATCH(eos, rc = eos_point<E>(which, a, b, c, Y, out));
  return rc;
}

int pc_ref_riemann(int eos, const double* left5, const double* spl,
                   const double* right5, const double* spr, double cav,
                   double* out12) {
  int rc = 0;
  DISPATCH(eos, rc = riemann_point<E>(left5, spl, right5, spr, cav, out12));
  return rc;
}

int pc_ref_ppm_reconstruct(const double* s5, double flatn, double* smsp) {
  Hydro<GammaLaw>::ppm_reconstruct(s5, flatn, smsp[0], smsp[1]);
  return 0;
}

int pc_ref_ppm_int_profile(double sm, double sp, double sc, double u, double c,
                           double dtdx, double* Ip3, double* Im3) {
  Hydro<GammaLaw>::ppm_int_profile(sm, sp, sc, u, c, dtdx, Ip3, Im3);
  return 0;
}

int pc_ref_mol_src_box(int eos, pc_box_t bx, pc_array4_t sarr,
                       pc_array4_t molsrc, const pc_params_t* prm,
                       const pc_array4_t* flx, const pc_array4_t* a,
                       pc_array4_t vol) {
  const Box b = to_box(bx);
  if (!covers(sarr, grow(b, 4))) return PC_ERR_GROW;
  if (!covers(molsrc, b) || !covers(vol, b)) return PC_ERR_ARG;
  A4 F[3];
  for (int d = 0; d < 3; ++d) {
    if (!covers(a[d], surroundingNodes(b, d))) return PC_ERR_ARG;
    if (flx) {
      if (!covers(flx[d], surroundingNodes(b, d))) return PC_ERR_ARG;
      F[d] = to_a4(flx[d]);
    }
  }
  const A4 AA[3] = {to_a4(a[0]), to_a4(a[1]), to_a4(a[2])};
  int rc = 0;
  DISPATCH(eos, rc = mol_src_box_t<E>(b, to_a4(sarr), to_a4(molsrc), *prm,
                                      flx ? F : nullptr, AA, to_a4(vol)));
  return rc;
}

int pc_ref_derive(int eos, pc_box_t bx, pc_array4_t u, const int* codes,
                  const int* comps, int nplot, pc_array4_t out,
                  const double* dx) {
  const Box b = to_box(bx);
  if (!covers(u, grow(b, 1)) || !covers(out, b) || nplot <= 0) return PC_ERR_ARG;
  for (int m = 0; m < nplot; ++m)
    if (codes[m] < 0 || codes[m] > 8) return PC_ERR_ARG;
  int rc = 0;
  DISPATCH(eos, rc = derive_t<E>(b, to_a4(u), codes, comps, nplot, to_a4(out), dx));
  return rc;
}

int pc_ref_transdo_point(int eos, int dir, int t, const double* qplus,
                         const double* qminus, const double* fl_lo,
                         const double* fl_hi, const double* qi_lo,
                         const double* qi_hi, double gamc, double cdtdx,
                         double* out_p, double* out_m) {
  if (dir < 0 || dir > 2 || t < 0 || t > 2 || t == dir) return PC_ERR_ARG;
  int rc = 0;
  DISPATCH(eos, rc = transdo_point<E>(dir, t, qplus, qminus, fl_lo, fl_hi, qi_lo,
                                      qi_hi, gamc, cdtdx, out_p, out_m));
  return rc;
}

int pc_ref_transdd_point(int eos, int dir, const double* qplus,
                         const double* qminus, const double* fl0_lo,
                         const double* fl0_hi, const double* fl1_lo,
                         const double* fl1_hi, const double* qi0_lo,
                         const double* qi0_hi, const double* qi1_lo,
                         const double* qi1_hi, double gamc, const double* srcq,
                         double hdt, double c0, double c1, double* out_p,
                         double* out_m) {
  if (dir < 0 || dir > 2) return PC_ERR_ARG;
  int rc = 0;
  DISPATCH(eos, rc = transdd_point<E>(dir, qplus, qminus, fl0_lo, fl0_hi, fl1_lo,
                                      fl1_hi, qi0_lo, qi0_hi, qi1_lo, qi1_hi,
                                      gamc, srcq, hdt, c0, c1, out_p, out_m));
  return rc;
}

double pc_ref_pdivu_point(const double* pg, const double* un,
                          const double* area, double vol) {
  // pc_pdivu (Source/Godunov.H:1033-1060) of the cell (0,0,0); per direction
  // d the low face holds pg[2d], un[2d], area[2d] and the high face the
  // entries 2d+1.
  const Box b2{{0, 0, 0}, {1, 1, 1}};
  Fab qe[3] = {Fab(b2, NGDNV), Fab(b2, NGDNV), Fab(b2, NGDNV)};
  Fab ar[3] = {Fab(b2, 1), Fab(b2, 1), Fab(b2, 1)};
  Fab v(b2, 1), out(b2, 1);
  const int gdv[3] = {GDU, GDV, GDW};
  for (int d = 0; d < 3; ++d) {
    const IV c = ivec(0, 0, 0), ch = shift(c, d, 1);
    at(qe[d].a, c, GDPRES) = pg[2 * d];
    at(qe[d].a, ch, GDPRES) = pg[2 * d + 1];
    at(qe[d].a, c, gdv[d]) = un[2 * d];
    at(qe[d].a, ch, gdv[d]) = un[2 * d + 1];
    at(ar[d].a, c, 0) = area[2 * d];
    at(ar[d].a, ch, 0) = area[2 * d + 1];
  }
  v.a(0, 0, 0) = vol;
  Hydro<GammaLaw>::pdivu_cell(0, 0, 0, out.a, qe[0].a, qe[1].a, qe[2].a,
                              ar[0].a, ar[1].a, ar[2].a, v.a);
  return out.a(0, 0, 0);
}

int pc_ref_laxf(int eos, const double* left5, const double* spl,
                const double* right5, const double* spr, double* out) {
  int rc = 0;
  DISPATCH(eos, rc = laxf_point<E>(left5, spl, right5, spr, out));
  return rc;
}

double pc_ref_plm_slope_line(const double* q5, double flat, int order) {
  // a 5-cell line in x, centre at i = 2
  double buf[5];
  for (int i = 0; i < 5; ++i) buf[i] = q5[i];
  Box b{{0, 0, 0}, {4, 0, 0}};
  A4 a = make_a4(buf, b, 1);
  return Hydro<GammaLaw>::plm_slope(2, 0, 0, 0, 0, a, flat, order);
}

}  // extern "C"

// ------------------------------------------------------------ flux register
// Dense restatement of the coarse/fine flux register (see include/pc_hydro.h,
// pc_fluxreg_*): one register value per coarse face of the domain and
// component, a face marked when exactly one of its two cells is refined and
// the other exists (in the domain or as a periodic image).
namespace {
struct RefFluxReg {
  int nc, ratio;
  Box dom;
  int periodic[3];
  std::vector<Box> cboxes, fc;  // coarse boxes, coarsened fine boxes
  // per direction: faces [domlo, domhi+1] x cells transverse; side: 0 none,
  // +1 outside cell above the face, -1 below
  std::vector<signed char> side[3];
  std::vector<double> reg[3];
  long nface(int d) const {
    long n = 1;
    for (int e = 0; e < 3; ++e) n *= (dom.hi[e] - dom.lo[e] + 1) + (e == d ? 1 : 0);
    return n;
  }
  long fid(int d, const int* f) const {
    long id = 0, m = 1;
    for (int e = 0; e < 3; ++e) {
      id += (f[e] - dom.lo[e]) * m;
      m *= (dom.hi[e] - dom.lo[e] + 1) + (e == d ? 1 : 0);
    }
    return id;
  }
  bool covered(const int* x) const {
    for (const Box& b : fc)
      if (x[0] >= b.lo[0] && x[0] <= b.hi[0] && x[1] >= b.lo[1] && x[1] <= b.hi[1] &&
          x[2] >= b.lo[2] && x[2] <= b.hi[2])
        return true;
    return false;
  }
  // the cell x moved into the domain through a periodic direction d, or
  // false when it lies outside a non-periodic one
  bool wrap(int d, int* x) const {
    if (x[d] >= dom.lo[d] && x[d] <= dom.hi[d]) return true;
    if (!periodic[d]) return false;
    const int L = dom.hi[d] - dom.lo[d] + 1;
    x[d] += x[d] < dom.lo[d] ? L : -L;
    return true;
  }
};

bool in_box(const Box& b, const int* x) {
  return x[0] >= b.lo[0] && x[0] <= b.hi[0] && x[1] >= b.lo[1] && x[1] <= b.hi[1] &&
         x[2] >= b.lo[2] && x[2] <= b.hi[2];
}
}  // namespace

extern "C" {

void* pc_ref_fluxreg_create(int nc, const pc_box_t* crse_boxes, int ncrse,
                            const pc_box_t* fine_boxes, int nfine,
                            const int* domlo, const int* domhi,
                            const int* periodic, int ratio) {
  RefFluxReg* R = new RefFluxReg();
  R->nc = nc;
  R->ratio = ratio;
  for (int d = 0; d < 3; ++d) {
    R->dom.lo[d] = domlo[d];
    R->dom.hi[d] = domhi[d];
    R->periodic[d] = periodic[d];
  }
  for (int b = 0; b < ncrse; ++b) R->cboxes.push_back(to_box(crse_boxes[b]));
  for (int b = 0; b < nfine; ++b) {
    Box c;
    for (int d = 0; d < 3; ++d) {
      c.lo[d] = fine_boxes[b].lo[d] / ratio;
      c.hi[d] = (fine_boxes[b].hi[d] + 1) / ratio - 1;
    }
    R->fc.push_back(c);
  }
  for (int d = 0; d < 3; ++d) {
    R->side[d].assign(R->nface(d), 0);
    R->reg[d].assign(R->nface(d) * nc, 0.0);
    Box fb = R->dom;
    fb.hi[d] += 1;
    // in a periodic direction the face domhi+1 is the face domlo
    if (R->periodic[d]) fb.hi[d] -= 1;
    ParallelFor(fb, [&](int i, int j, int k) {
      int a[3] = {i, j, k}, b[3] = {i, j, k};
      b[d] -= 1;  // the cell below the face
      const bool ia = R->wrap(d, a), ib = R->wrap(d, b);
      if (!ia || !ib) return;
      const bool ca = R->covered(a), cb = R->covered(b);
      const int f[3] = {i, j, k};
      if (ca && !cb) R->side[d][R->fid(d, f)] = -1;
      if (cb && !ca) R->side[d][R->fid(d, f)] = +1;
    });
  }
  return R;
}

void pc_ref_fluxreg_destroy(void* r) { delete (RefFluxReg*)r; }

int pc_ref_fluxreg_nfaces(void* r) {
  const RefFluxReg* R = (const RefFluxReg*)r;
  int n = 0;
  for (int d = 0; d < 3; ++d)
    for (signed char s : R->side[d]) n += s != 0;
  return n;
}

// reg += scale * flux of coarse box bx on the register faces whose outside
// cell is in bx
int pc_ref_fluxreg_crse_add(void* r, pc_box_t bx, const pc_array4_t* flx,
                            double scale) {
  RefFluxReg* R = (RefFluxReg*)r;
  const Box b = to_box(bx);
  for (int d = 0; d < 3; ++d) {
    const A4 F = to_a4(flx[d]);
    Box fb = R->dom;
    fb.hi[d] += R->periodic[d] ? 0 : 1;
    ParallelFor(fb, [&](int i, int j, int k) {
      const int f[3] = {i, j, k};
      const long id = R->fid(d, f);
      const int s = R->side[d][id];
      if (!s) return;
      int out[3] = {i, j, k};
      if (s < 0) out[d] -= 1;
      R->wrap(d, out);
      if (!in_box(b, out)) return;
      int face[3] = {out[0], out[1], out[2]};  // in the box's index space
      if (s < 0) face[d] += 1;
      for (int n = 0; n < R->nc; ++n)
        R->reg[d][id * R->nc + n] += scale * F(face[0], face[1], face[2], n);
    });
  }
  return 0;
}

// reg += scale * (sum of the ratio^2 fine fluxes of fine box bx)
int pc_ref_fluxreg_fine_add(void* r, pc_box_t bx, const pc_array4_t* flx,
                            double scale) {
  RefFluxReg* R = (RefFluxReg*)r;
  Box c;  // bx coarsened
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = bx.lo[d] / R->ratio;
    c.hi[d] = (bx.hi[d] + 1) / R->ratio - 1;
  }
  for (int d = 0; d < 3; ++d) {
    const A4 F = to_a4(flx[d]);
    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    Box fb = R->dom;
    fb.hi[d] += R->periodic[d] ? 0 : 1;
    ParallelFor(fb, [&](int i, int j, int k) {
      const int f[3] = {i, j, k};
      const long id = R->fid(d, f);
      const int s = R->side[d][id];
      if (!s) return;
      // the refined cell next to the face, in bx's unwrapped frame
      for (int img = -1; img <= 1; ++img) {
        if (img != 0 && !R->periodic[d]) continue;
        int in[3] = {i, j, k};
        in[d] += img * (R->dom.hi[d] - R->dom.lo[d] + 1);
        const int face_d = in[d];
        if (s > 0) in[d] -= 1;  // outside above: refined cell below
        if (!in_box(c, in)) continue;
        for (int n = 0; n < R->nc; ++n) {
          double sum = 0.0;
          for (int bb = 0; bb < R->ratio; ++bb)
            for (int aa = 0; aa < R->ratio; ++aa) {
              int x[3];
              x[d] = face_d * R->ratio;
              x[t0] = in[t0] * R->ratio + aa;
              x[t1] = in[t1] * R->ratio + bb;
              sum += F(x[0], x[1], x[2], n);
            }
          R->reg[d][id * R->nc + n] += scale * sum;
        }
      }
    });
  }
  return 0;
}

// U += +reg/V (cell above its register face) or -reg/V (below) on the
// outside cells in bx, per cell in direction order, low face first
int pc_ref_fluxreg_reflux(void* r, pc_box_t bx, pc_array4_t state, double vol) {
  RefFluxReg* R = (RefFluxReg*)r;
  const Box b = to_box(bx);
  const A4 U = to_a4(state);
  ParallelFor(b, [&](int i, int j, int k) {
    for (int n = 0; n < R->nc; ++n) {
      double u = U(i, j, k, n);
      bool touched = false;
      for (int d = 0; d < 3; ++d)
        for (int hi = 0; hi <= 1; ++hi) {
          int f[3] = {i, j, k};
          f[d] += hi;  // the cell's low (hi = 0) or high face
          if (R->periodic[d] && f[d] == R->dom.hi[d] + 1) f[d] = R->dom.lo[d];
          if (f[d] < R->dom.lo[d] || f[d] > R->dom.hi[d] + 1) continue;
          const long id = R->fid(d, f);
          const int s = R->side[d][id];
          if (hi == 0 && s == +1) {
            u += R->reg[d][id * R->nc + n] / vol;
            touched = true;
          }
          if (hi == 1 && s == -1) {
            u += -R->reg[d][id * R->nc + n] / vol;
            touched = true;
          }
        }
      if (touched) U(i, j, k, n) = u;
    }
  });
  return 0;
}

int pc_ref_average_down(int nc, pc_box_t crse_region, pc_array4_t crse,
                        pc_array4_t fine, int ratio) {
  const A4 C = to_a4(crse), F = to_a4(fine);
  const double volfrac = 1.0 / (double)(ratio * ratio * ratio);
  ParallelFor(to_box(crse_region), [&](int i, int j, int k) {
    for (int n = 0; n < nc; ++n) {
      double c = 0.0;
      for (int kk = 0; kk < ratio; ++kk)
        for (int jj = 0; jj < ratio; ++jj)
          for (int ii = 0; ii < ratio; ++ii)
            c += F(ratio * i + ii, ratio * j + jj, ratio * k + kk, n);
      C(i, j, k, n) = volfrac * c;
    }
  });
  return 0;
}

}  // extern "C"
