// Level-path kernels: ghost fill over a box table, the fused cell update
// (flux divergence + construct_Snew + reset_internal_energy + computeTemp +
// estTimeStep partial minimum), the device-side dt finalisation and the
// primitive -> conservative initialiser.
#pragma once
#include "pc_kernels.h"

namespace pcd {


struct LevelArgs {
  int periodic_self;
  int dlo[3], dhi[3], bclo[3], bchi[3];
  const int* cellmap;  // domain cell -> box (domain mode)
  double cfl, max_dt, fixed_dt, change_max, init_shrink, dt_cutoff;
  double stop_time;  // < 0: no stop_time limit
};

// Source index of ghost coordinate x (Source/Setup.cpp:42-121 BC map:
// Interior -> periodic, Outflow/Inflow/UserBC -> first-order extrapolation,
// Symmetry/SlipWall -> reflect (normal momentum odd), NoSlipWall -> reflect
// (all momenta odd)).  refl: 0 none, 1 normal-odd, 2 all-momenta-odd.
PCD int ghost_src(int x, int lo, int hi, int bcl, int bch, int& refl) {
  refl = 0;
  if (x >= lo && x <= hi) return x;
  const int L = hi - lo + 1;
  const int bc = (x < lo) ? bcl : bch;
  if (bc == BC_INTERIOR) {
    // a ghost cell is at most 4 cells out: one period, unless the domain is
    // shorter than that
    if (x >= lo - L && x <= hi + L) return x < lo ? x + L : x - L;
    return lo + (((x - lo) % L) + L) % L;
  }
  if (bc == BC_SYMMETRY || bc == BC_SLIPWALL || bc == BC_NOSLIPWALL) {
    refl = (bc == BC_NOSLIPWALL) ? 2 : 1;
    return (x < lo) ? 2 * lo - 1 - x : 2 * hi + 1 - x;
  }
  return (x < lo) ? lo : hi;
}

// Where ghost cell (i,j,k) of job J takes its values from: the source cell
// (s[3]), the state it lies in (the job's own, or in a level whose boxes tile
// the domain the box that owns the source cell) and the reflection kind per
// direction (r[3], ghost_src).
template <class J_>
PCD void ghost_source(const Sched& sc, const J_& J, const LevelArgs& la, int i,
                      int j, int k, int s[3], int r[3], V& src, bool sold = false) {
  // periodic_self: 0 = boxes tile the level domain, 1 = every box is its own
  // periodic domain, 2 = every box is its own domain with its own BCs
  int D0[3], D1[3], bl[3], bh[3];
  for (int d = 0; d < 3; ++d) {
    D0[d] = la.periodic_self ? J.lo[d] : la.dlo[d];
    D1[d] = la.periodic_self ? J.hi[d] : la.dhi[d];
    bl[d] = la.periodic_self == 1 ? BC_INTERIOR
                                  : (la.periodic_self == 2 ? J.bclo[d] : la.bclo[d]);
    bh[d] = la.periodic_self == 1 ? BC_INTERIOR
                                  : (la.periodic_self == 2 ? J.bchi[d] : la.bchi[d]);
  }
  s[0] = ghost_src(i, D0[0], D1[0], bl[0], bh[0], r[0]);
  s[1] = ghost_src(j, D0[1], D1[1], bl[1], bh[1], r[1]);
  s[2] = ghost_src(k, D0[2], D1[2], bl[2], bh[2], r[2]);
  src = sold ? J.sold : J.U;
  if (!la.periodic_self) {
    const long nx = D1[0] - D0[0] + 1, ny = D1[1] - D0[1] + 1;
    const int sb =
        la.cellmap[(s[0] - D0[0]) + nx * ((s[1] - D0[1]) + ny * (long)(s[2] - D0[2]))];
    src = sold ? sc.jobs[sb].sold : sc.jobs[sb].U;
  }
}

// Sign of component m of a ghost value: momenta flip across a reflecting
// face (the normal one for Symmetry/SlipWall, all three for NoSlipWall).
PCD double ghost_sign(const int r[3], int m) {
  double sg = 1.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool mom = (m == UMX || m == UMY || m == UMZ);
    if ((r[d] == 2 && mom) || (r[d] == 1 && m == UMX + d)) sg = -sg;
  }
  return sg;
}

template <int NS>
PCD void b_ghost(const Sched& sc, const KArgs& ka, int lb_, LevelArgs la) {
  constexpr int NVAR = 7 + NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  int i, j, k;
  if (!ghost_shell_cell(J.lo, J.hi, local, i, j, k)) return;
  int s[3], r[3];
  V src;
  ghost_source(sc, J, la, i, j, k, s, r, src);
#pragma unroll
  for (int m = 0; m < NVAR; ++m) J.U(i, j, k, m) = ghost_sign(r, m) * src(s[0], s[1], s[2], m);
}
template <int NS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_ghost(Sched sc, KArgs ka, LevelArgs la) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_ghost<NS>(sc, ka, lb_, la);
  } else {
    b_ghost<NS>(sc, ka, first_block(sc), la);
  }
}

// The primitives with the ghost fill folded in (the level and batch steps):
// a valid cell reads its own state; a ghost cell reads its source cell's
// state with the reflection signs, the values the ghost fill would have
// written, so the primitives are the same bits.  REG: the Godunov step's
// grow(bx,4) (ST_CTOPRIM), of whose ghost cells the first layer (grow(bx,1),
// which the final fluxes' artificial viscosity reads) is written back to the
// state, or the method-of-lines step's grow(bx,2) (ST_MOL_PRIM), which writes
// none back (its kernels read the state on the valid cells only).  The
// other ghost cells are not stored (pc_level_fill_ghosts_v1 fills all four
// layers).  SOLD: the second MOL stage's primitives, of the first stage's
// result U*, which the first update leaves in each box's stage array sold
// (the valid cells and every ghost source read there).
template <class EOS, class FP, int REG, bool SOLD = false>
PCD void b_ctoprim_ghost(const Sched& sc, const KArgs& ka, int lb_,
                         const LevelArgs& la) {
  constexpr int NVAR = 7 + EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(REG, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  double Uc[NVAR];
  // distance outside the valid box (0 inside)
  auto out = [](int x, int l, int h) { return x < l ? l - x : (x > h ? x - h : 0); };
  const int d0 = out(i, J.lo[0], J.hi[0]), d1 = out(j, J.lo[1], J.hi[1]),
            d2 = out(k, J.lo[2], J.hi[2]);
  const int dist = d0 > d1 ? (d0 > d2 ? d0 : d2) : (d1 > d2 ? d1 : d2);
  if (dist <= 0) {
    const V& u = SOLD ? J.sold : J.U;
#pragma unroll
    for (int m = 0; m < NVAR; ++m)
      if (m != UEINT) Uc[m] = u(i, j, k, m);
  } else {
    int s[3], r[3];
    V src;
    ghost_source(sc, J, la, i, j, k, s, r, src, SOLD);
    const bool store = REG == ST_CTOPRIM && dist == 1;
#pragma unroll
    for (int m = 0; m < NVAR; ++m) {
      if (m == UEINT && !store) continue;
      Uc[m] = ghost_sign(r, m) * src(s[0], s[1], s[2], m);
      if (store) J.U(i, j, k, m) = Uc[m];
    }
  }
  ctoprim_cell<EOS, FP, REG == ST_MOL_PRIM>(J, ka, i, j, k, Uc);
}
template <class EOS, class FP = double, bool PERSIST = false, int REG = ST_CTOPRIM,
          bool SOLD = false>
__global__ void __launch_bounds__(BLOCK)
    k_ctoprim_ghost(Sched sc, KArgs ka, LevelArgs la) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_ctoprim_ghost<EOS, FP, REG, SOLD>(sc, ka, lb_, la);
  } else {
    b_ctoprim_ghost<EOS, FP, REG, SOLD>(sc, ka, first_block(sc), la);
  }
}

PCD double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = mn(v, __shfl_xor(v, off, 64));
  return v;
}

// block-wide minimum of positive doubles -> atomic minimum on the bits
PCD void block_min_to(double v, unsigned long long* dst) {
  __shared__ double red[BLOCK / 64];
  v = wave_min(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < BLOCK / 64; ++w) m = mn(m, red[w]);
    if (m < 1.0e300) atomicMin(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// Mass fractions of a conserved state
template <int NS>
PCD void state_Y(const double* S, double* Y) {
  const double rhoInv = 1.0 / S[URHO];
#pragma unroll
  for (int s = 0; s < NS; ++s) Y[s] = S[UFS + s] * rhoInv;
}

// pc_estdt_hydro for one cell (Source/Timestep.H:62-84) from its conserved
// state and the thermodynamics of its composition
template <class EOS>
PCD double cell_estdt(const double* S, const typename EOS::Thermo& th,
                      const double* dx) {
  const double rho = S[URHO];
  const double rhoInv = 1.0 / rho;
  const double T = S[UTEMP];
  const double c = th.RTY2Cs(rho, T);
  double dt = 1.7976931348623157e308;  // numeric_limits<double>::max()
  const double ux = S[UMX] * rhoInv;
  dt = mn(dt, dx[0] / (c + fabs(ux)));
  const double uy = S[UMY] * rhoInv;
  dt = mn(dt, dx[1] / (c + fabs(uy)));
  const double uz = S[UMZ] * rhoInv;
  dt = mn(dt, dx[2] / (c + fabs(uz)));
  return dt;
}

// reset_internal_energy with allow_small_energy = allow_negative_energy = 1
// (Source/PeleCUtilities.H:227-236) and pc_cmpTemp (:70-84).  Neither
// changes rho or rho Y, so th (the composition's thermodynamics) holds.
template <class EOS>
PCD void reset_and_temp(double* S, const typename EOS::Thermo& th) {
  {
    const double rho = S[URHO];
    const double rhoInv = 1.0 / rho;
    const double Up = S[UMX] * rhoInv;
    const double Vp = S[UMY] * rhoInv;
    const double Wp = S[UMZ] * rhoInv;
    const double ke =
        0.5 * (Up * Up + Vp * Vp + Wp * Wp) - 0.5 * 0.0 * 0.0 * 0.0 * 0.0;
    S[UEINT] = S[UEDEN] - rho * ke;
  }
  const double rhoInv = 1.0 / S[URHO];
  const double e = S[UEINT] * rhoInv;
  S[UTEMP] = th.EY2T(e, S[UTEMP]);
}

// Fused per-cell step end: dU/dt from the fluxes, S_new = S_old + dt dU/dt
// (construct_Snew, Source/Advance.cpp:397-416), reset + e->T twice
// (computeTemp in do_sdc_iteration and in post_timestep), NaN check and the
// CFL partial minimum on the new state.
// F32: the final fluxes are the mixed step's FP32 ones (cell_divergence_fp32)
template <class EOS, bool F32>
PCD double b_update(const Sched& sc, const KArgs& ka, int lb_, DtState*& ds) {
  constexpr int NS = EOS::NS;
  constexpr int NVAR = 7 + NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  if (J.ds) ds = J.ds;  // this job's dt state (the same for the whole block)
  int lo[3], n[3], i, j, k;
  stage_region(ST_UPDATE, J.lo, J.hi, lo, n);
  double mydt = 1.7976931348623157e308;
  if (region_ijk(sc, lo, n, local, i, j, k)) {
    const double dt = get_dt(J, ka);
    double h[NVAR];
#pragma unroll
    for (int m = 0; m < NVAR; ++m) h[m] = 0.0;
    if constexpr (F32) cell_divergence_fp32<NS>(J, ka, i, j, k, h);
    else cell_divergence<NS, true>(J, ka, i, j, k, h);
    // UEINT is reset from UEDEN and the momenta below before it is read
    double S[NVAR];
#pragma unroll
    for (int m = 0; m < NVAR; ++m) S[m] = m == UEINT ? 0.0 : J.U(i, j, k, m) + dt * h[m];
    double Y[NS];
    state_Y<NS>(S, Y);
    const typename EOS::Thermo th(Y);
    reset_and_temp<EOS>(S, th);
    reset_and_temp<EOS>(S, th);
    bool bad = false;
#pragma unroll
    for (int m = 0; m < NVAR; ++m) {
      J.U(i, j, k, m) = S[m];
      bad |= !isfinite(S[m]);
    }
    if (bad) ds->bad = 1;
    mydt = cell_estdt<EOS>(S, th, J.dx);
  }
  return mydt;
}

template <class EOS, bool PERSIST = false, bool F32 = false>
__global__ void __launch_bounds__(BLOCK)
    k_update(Sched sc, KArgs ka, DtState* ds) {
  double m = 1.7976931348623157e308;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      m = mn(m, b_update<EOS, F32>(sc, ka, lb_, ds));
  } else {
    m = b_update<EOS, F32>(sc, ka, first_block(sc), ds);
  }
  block_min_to(m, &ds->minbits);
}

// estTimeStep partial minimum over the valid cells (no update)
template <class EOS>
PCD double b_estdt(const Sched& sc, const KArgs& ka, int lb_, DtState*& ds) {
  constexpr int NS = EOS::NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  if (J.ds) ds = J.ds;
  int lo[3], n[3], i, j, k;
  stage_region(ST_UPDATE, J.lo, J.hi, lo, n);
  double mydt = 1.7976931348623157e308;
  if (region_ijk(sc, lo, n, local, i, j, k)) {
    double S[7 + NS];
#pragma unroll
    for (int m = 0; m < 7 + NS; ++m) S[m] = J.U(i, j, k, m);
    double Y[NS];
    state_Y<NS>(S, Y);
    const typename EOS::Thermo th(Y);
    mydt = cell_estdt<EOS>(S, th, J.dx);
  }
  return mydt;
}

template <class EOS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_estdt(Sched sc, KArgs ka, DtState* ds) {
  double m = 1.7976931348623157e308;
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      m = mn(m, b_estdt<EOS>(sc, ka, lb_, ds));
  } else {
    m = b_estdt<EOS>(sc, ka, first_block(sc), ds);
  }
  block_min_to(m, &ds->minbits);
}

// What dt_finalize computes from the partial minimum.
enum DtMode {
  DT_NEW = 0,   // computeNewDt after a step of ds->dt: time += dt first
  DT_INIT = 1,  // computeInitialDt at ds->time
  DT_EST = 2    // estTimeStep alone (no init_shrink, no stop_time)
};

// estTimeStep (Source/PeleC.cpp:829-961) from the partial minimum, then
// computeNewDt (:964-1034, single level) or computeInitialDt (:1041-1075 ->
// initialTimeStep :811-826), both limited by stop_time (:1018-1027,
// 1061-1068) at the time the state has reached.
PCD void dt_finalize(DtState* ds, const LevelArgs& la, int mode) {
  double est;
  if (la.fixed_dt > 0.0) {
    est = la.fixed_dt;
  } else {
    est = la.max_dt;
    double eh = la.max_dt / la.cfl;
    const double dtc = __longlong_as_double((long long)ds->minbits);
    eh = mn(eh, dtc);
    eh *= la.cfl;
    if (eh < est) est = eh;
  }
  double dtn;
  if (mode == DT_EST) {
    dtn = est;
  } else {
    if (mode == DT_INIT) {
      dtn = la.init_shrink * est;
    } else {
      ds->time += ds->dt;
      dtn = est;
      if (la.fixed_dt <= 0.0) dtn = mn(dtn, la.change_max * ds->dt);
    }
    const double dt_eps = 0.001 * dtn;
    if (la.stop_time >= 0.0) {
      if ((ds->time + dtn) > (la.stop_time - dt_eps)) dtn = la.stop_time - ds->time;
    }
  }
  ds->dt = dtn;
  if (dtn < la.dt_cutoff) ds->small = 1;
  ds->minbits = 0x7FEFFFFFFFFFFFFFull;
}

__global__ void k_dt_finalize(DtState* ds, LevelArgs la, int mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  dt_finalize(ds, la, mode);
}

// the same for the requests ids[0..n) of a batch, one thread each
__global__ void k_dt_finalize_list(DtState* base, const int* ids, int n,
                                   LevelArgs la, int mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dt_finalize(base + ids[t], la, mode);
}

// primitive -> conservative (pc_initdata pattern; see workloads.py)
template <class EOS>
PCD void b_prim_to_cons(const Sched& sc, const KArgs& ka, int lb_, V prim, int mode) {
  constexpr int NS = EOS::NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_UPDATE, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  double Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) Y[s] = prim(i, j, k, 5 + s);
  const double u = prim(i, j, k, 2), v = prim(i, j, k, 3), w = prim(i, j, k, 4);
  double rho, e, T;
  if (mode == 0) {
    rho = prim(i, j, k, 0);
    const double p = prim(i, j, k, 1);
    e = EOS::RYP2E(rho, Y, p);
    const typename EOS::Thermo th(Y);
    T = th.EY2T(e, 1.0);
  } else {
    const double p = prim(i, j, k, 0);
    T = prim(i, j, k, 1);
    const double wb = EOS::wbar(Y);
    rho = p * wb / (RU * T);
    e = EOS::TY2E(T, Y);
  }
  const double rhoe = rho * e;
  const V& U = J.U;
  U(i, j, k, URHO) = rho;
  U(i, j, k, UMX) = rho * u;
  U(i, j, k, UMY) = rho * v;
  U(i, j, k, UMZ) = rho * w;
  U(i, j, k, UEINT) = rhoe;
  U(i, j, k, UEDEN) = rhoe + 0.5 * rho * (u * u + v * v + w * w);
  U(i, j, k, UTEMP) = T;
#pragma unroll
  for (int s = 0; s < NS; ++s) U(i, j, k, UFS + s) = rho * Y[s];
}
template <class EOS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_prim_to_cons(Sched sc, KArgs ka, V prim, int mode) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_prim_to_cons<EOS>(sc, ka, lb_, prim, mode);
  } else {
    b_prim_to_cons<EOS>(sc, ka, first_block(sc), prim, mode);
  }
}

}  // namespace pcd
