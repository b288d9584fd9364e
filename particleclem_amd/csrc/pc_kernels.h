// Stage kernels of the Godunov step for gfx950.
//
// One launch per stage covers every box of a box table (one box for the
// drop-in pc_umdrv_v1, a whole level for the level stepper).  Blocks of 256
// threads (4 wave64) map to boxes through a block->box table; inside a box
// the thread index runs i-fastest over the stage's region, so every global
// access is a unit-stride (coalesced) walk along i of an SoA component.
//
// Stage regions (bx = [L, H] the valid box) are the dependency cone of the
// valid outputs, smaller than the reference's grow-2 temporaries:
//   ctoprim      cells grow(bx,4)
//   divu         [L-1, H+1]^3: divu on the nodes [L, H+1]^3, the flattening
//                coefficient on the cells grow(bx,1)
//   trace_d      cells grow(bx,1); writes qm on face c+1 and qp on face c
//   flux1_d      d-faces [L, H+1] x transverse [L-1, H+1]
//   trans_d      d-faces [L, H+1] x [L-1, H+1]^2: (d|t) where t is in
//                [L, H], for both transverse t (under the stage of the
//                first t; the stage of the second is empty)
//   final_d      d-faces [L, H+1] x transverse [L, H]
//   update       cells [L, H]
// and of the method-of-lines step:
//   mol_prim     cells grow(bx,2)
//   mol_d        d-faces [L, H+1] x transverse [L, H]
//   mol_update   cells [L, H]
#pragma once
#include "pc_phys.h"
#include "pc_shell.h"

namespace pcd {

constexpr int BLOCK = 256;

// SoA view of one FAB (Array4 with 32-bit strides) of element type T.
// Element references are in the global address space (address space 1), so
// loads and stores are global_load/global_store rather than flat accesses.
template <class T>
struct VT {
  typedef __attribute__((address_space(1))) T gT;
  T* p;
  int js, ks, ns;
  int b0, b1, b2;
  PCD gT& operator()(int i, int j, int k, int n) const {
    return ((gT*)p)[(i - b0) + (j - b1) * js + (k - b2) * ks + n * ns];
  }
  PCD bool contains(int i, int j, int k, int e0, int e1, int e2) const {
    return i >= b0 && i < e0 && j >= b1 && j < e1 && k >= b2 && k < e2;
  }
};
using V = VT<double>;

// The same view with the component offset added to the (wave-uniform) base
// and the cell offset a 32-bit byte offset (never negative: every access lies
// in the view's box): loads take the SGPR-base + VGPR-offset form and all
// components of a cell share one offset register instead of a 64-bit
// address each.  Each component then holds its own SGPR base, so this pays
// only in kernels with few arrays (the transverse kernels: 171 -> 168
// VGPRs, 2 -> 3 waves per SIMD); with many arrays the bases spill.
template <class T>
struct VS : VT<T> {
  PCD typename VT<T>::gT& operator()(int i, int j, int k, int n) const {
    typedef __attribute__((address_space(1))) char gC;
    const unsigned off = (unsigned)((i - this->b0) + (j - this->b1) * this->js +
                                    (k - this->b2) * this->ks) *
                         (unsigned)sizeof(T);
    return *(typename VT<T>::gT*)((gC*)((typename VT<T>::gT*)this->p +
                                        (long)n * this->ns) + off);
  }
};
template <class T>
PCD VS<T> split(const VT<T>& v) {
  return VS<T>{v};
}

// A read-only view whose loads are buffer loads: the array's base in a
// buffer resource (4 SGPRs for the whole array), the component offset in the
// instruction's scalar offset and the cell offset a 32-bit VGPR byte offset
// (never negative: every access lies in the view's box), so a load needs no
// per-component VGPR address arithmetic and no per-component base register.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t BufRsrc;
#else
struct BufRsrc {  // the host pass parses device code but never runs it
  int w[4];
};
#endif
template <class T>
struct VB {
  BufRsrc r;
  int js, ks, ns, b0, b1, b2;
  PCD T operator()(int i, int j, int k, int n) const {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned off =
        (unsigned)((i - b0) + (j - b1) * js + (k - b2) * ks) * (unsigned)sizeof(T);
    const int soff = n * ns * (int)sizeof(T);
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
#else
    return T(0);
#endif
  }
};
// The same array seen as FP32 elements with the same element strides (the
// data occupies the first half of each component's bytes): the mixed step
// stores its final fluxes and qec so.
PCD VT<float> as_fp32(const VT<double>& v) {
  VT<float> f;
  f.p = (float*)v.p;
  f.js = v.js;
  f.ks = v.ks;
  f.ns = v.ns;
  f.b0 = v.b0;
  f.b1 = v.b1;
  f.b2 = v.b2;
  return f;
}

template <class T>
PCD VB<T> bufview(const VT<T>& v) {
  VB<T> b;
#if defined(__HIP_DEVICE_COMPILE__)
  // raw buffer (stride 0), 2 GiB range, the data format word of gfx9 parts
  b.r = __builtin_amdgcn_make_buffer_rsrc((void*)v.p, (short)0, 0x7fffffff, 0x00020000);
#endif
  b.js = v.js;
  b.ks = v.ks;
  b.ns = v.ns;
  b.b0 = v.b0;
  b.b1 = v.b1;
  b.b2 = v.b2;
  return b;
}

// Compact layouts of the stage intermediates.
//   traces: comps [0, QT) = minus state (left of face), [QT, 2QT) = plus
//           state; per state rho, u, v, w, p, rhoe (slot not stored), Y[NS]
//           (QT = 6 + NS)
//   fluxes: rho, mx, my, mz, rhoE, p_gd, un_gd, rhoY[NS]   (FT = 7 + NS)
// FP is the element type of the primitives and of the stage intermediates
// (double, or float in the mixed-precision step); the state, the final
// fluxes, qec and divu are always double.  Both instantiations have the same
// layout (views differ only in the pointee type), so one job table serves
// either.
// Device-resident time-step state of a level or of one request: dt of the
// next step, the running minimum of dx/(c+|u|) (bits of a positive double),
// the NaN and dt_cutoff flags and the time the state has reached.
struct DtState {
  double dt;
  unsigned long long minbits;
  int bad;
  int small;
  double time;
};

template <class FP>
struct JobT {
  int lo[3], hi[3];                 // valid box
  int dlo[3], dhi[3], bclo[3], bchi[3];
  int qe[3];                        // exclusive end of the q box (contains)
  double dx[3];                     // cell size
  double area_u[3], vol_u;          // uniform face areas / volume (no area/vol)
  const double* dtp;                // device address of this job's dt
  double dt_val;                    // the dt of a box-driver call (dtp -> here)
  DtState* ds;                      // dt state the CFL minimum goes to (or 0)
  VT<double> U;                     // state with >= 4 ghosts (uin)
  VT<FP> q, qa, srcq;               // primitives on grow(bx,4); srcq.p may be 0
  VT<FP> tr[3];                     // traces per direction
  VT<FP> f1[3];                     // first fluxes
  VT<FP> dl[3];                     // their per-cell transverse corrections
  VT<FP> ft[6];                     // transverse fluxes, [2*d + slot(d,t)]
  VT<double> flx[3], qec[3];        // final fluxes (NVAR) and p_gd, un_gd
  VT<double> area[3], vol;          // p == 0: uniform geometry
  VT<double> upd;                   // uout (dU/dt accumulator) or p == 0
  VT<double> divu;                  // nodes [L, H+1]^3
  VT<double> sold;                  // MOL: the first stage's U* on the valid box (or p == 0)
  VT<FP> flat;                      // flattening coefficient on grow(bx,1)
};
using Job = JobT<double>;
static_assert(sizeof(JobT<float>) == sizeof(JobT<double>),
              "job layouts of both precisions must agree");

struct KArgs {
  double difmag;
  int plm_iorder;
  int use_flattening;
  int nblocks;
  int full_prims;  // 1: write every q/qaux comp; 0: only what the step reads
  int skip_reint;  // 1: q's rho e is not stored (the level and batch steps)
};

template <class J_>
PCD double get_dt(const J_& J, const KArgs&) {
  return *(const __attribute__((address_space(1))) double*)J.dtp;
}

PCD int tslot(int d, int t) { return (t < d) ? t : t - 1; }

// Logical block -> (job, cells).  A launch covers `nlog` logical blocks of
// BLOCK threads (one workgroup per logical block, or fewer workgroups that
// loop over them: a persistent launch).  A logical block is either a run of
// BLOCK consecutive cells of the stage region in i-fastest order (lx == -1)
// or a tile of 2^lx x 2^ly x 2^lz cells (lx + ly + lz = log2 BLOCK).
struct Sched {
  const Job* jobs;
  const int* blkmap;    // logical block -> job (null: search blkstart)
  const int* blkstart;  // job -> its first logical block (null: one job)
  int nlog;             // logical blocks (0: one per workgroup)
  int lx, ly, lz;       // tile shape (log2), lx = -1 for runs
  int njobs;            // jobs in blkstart (used without blkmap)
  int rev = 0;          // 1: each XCD walks its range of logical blocks backwards
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i % 8).  Renumber them so that each XCD walks one contiguous range of the
// logical block order: stencil neighbours (the next rows and planes of a
// box) then sit in the L2 of the XCD that just read them.
PCD int xcd_swizzle(int i, int n) {
  const int x = i & 7, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + (i >> 3);
}

// The job table is read-only during a launch: read it through the constant
// address space so the (block-uniform) descriptor fields are scalar loads.
typedef const __attribute__((address_space(4))) int CInt;

PCD int first_block() { return xcd_swizzle(blockIdx.x, gridDim.x); }
// One workgroup per logical block: the XCD's range in order, or backwards
// (s.rev), so that a launch can start where the previous one ended.
PCD int first_block(const Sched& s) {
  const int i = blockIdx.x, n = gridDim.x;
  if (!s.rev) return xcd_swizzle(i, n);
  const int x = i & 7, q = n >> 3, r = n & 7;
  const int cnt = q + (x < r ? 1 : 0);
  return x * q + (x < r ? x : r) + (cnt - 1 - (i >> 3));
}
PCD int nlogical(const Sched& s) { return s.nlog ? s.nlog : (int)gridDim.x; }

template <class FP = double>
PCD JobT<FP> load_job(const Sched& s, int lb, int& local) {
  typedef const __attribute__((address_space(4))) JobT<FP> CJob;
#if defined(__HIP_DEVICE_COMPILE__)
  int b = 0, start = 0;
  if (s.blkmap) {
    b = ((CInt*)s.blkmap)[lb];
    start = ((CInt*)s.blkstart)[b];
  } else if (s.blkstart) {
    // last job whose first block is <= lb (uniform: scalar loads)
    int lo = 0, hi = s.njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (((CInt*)s.blkstart)[mid] <= lb) lo = mid;
      else hi = mid - 1;
    }
    b = lo;
    start = ((CInt*)s.blkstart)[b];
  }
  local = (lb - start) * BLOCK + threadIdx.x;
  return ((CJob*)s.jobs)[b];
#else
  // the host compilation pass parses device code but never runs it
  local = lb;
  return *(const JobT<FP>*)s.jobs;
#endif
}

// Cell (i,j,k) of thread `local` of a job over the region [lo, lo + n).
PCD bool region_ijk(const Sched& s, const int lo[3], const int n[3], int local,
                    int& i, int& j, int& k) {
  if (s.lx < 0) {
    const int nxy = n[0] * n[1];
    if (local >= nxy * n[2]) return false;
    k = local / nxy;
    const int r = local - k * nxy;
    j = r / n[0];
    i = r - j * n[0];
  } else {
    const int bib = local / BLOCK, t = local % BLOCK;
    const int tx = 1 << s.lx, ty = 1 << s.ly, tz = 1 << s.lz;
    const int nt0 = (n[0] + tx - 1) >> s.lx, nt1 = (n[1] + ty - 1) >> s.ly;
    const int b0 = bib % nt0, r = bib / nt0;
    const int b1 = r % nt1, b2 = r / nt1;
    i = (b0 << s.lx) + (t & (tx - 1));
    j = (b1 << s.ly) + ((t >> s.lx) & (ty - 1));
    k = (b2 << s.lz) + (t >> (s.lx + s.ly));
    (void)tz;
    if (i >= n[0] || j >= n[1] || k >= n[2]) return false;
  }
  i += lo[0];
  j += lo[1];
  k += lo[2];
  return true;
}

// Stage regions (see file comment).
enum Stage {
  ST_CTOPRIM = 0,
  ST_DIVU,
  ST_TRACE,
  ST_FLUX1_X, ST_FLUX1_Y, ST_FLUX1_Z,
  ST_TRANS_XY, ST_TRANS_XZ, ST_TRANS_YX, ST_TRANS_YZ, ST_TRANS_ZX, ST_TRANS_ZY,
  ST_FINAL_X, ST_FINAL_Y, ST_FINAL_Z,
  ST_UPDATE,
  ST_GHOST,
  ST_MOL_PRIM,
  ST_MOL_X, ST_MOL_Y, ST_MOL_Z,
  ST_MOL_UPD,
  ST_COUNT
};

__host__ __device__ inline void stage_region(int st, const int L[3],
                                             const int H[3], int lo[3],
                                             int n[3]) {
  int a[3], b[3];
  for (int d = 0; d < 3; ++d) { a[d] = L[d]; b[d] = H[d]; }
  auto set = [&](int d, int x, int y) { a[d] = x; b[d] = y; };
  switch (st) {
    case ST_CTOPRIM:
    case ST_GHOST:
      for (int d = 0; d < 3; ++d) set(d, L[d] - 4, H[d] + 4);
      break;
    case ST_DIVU:
      for (int d = 0; d < 3; ++d) set(d, L[d] - 1, H[d] + 1);
      break;
    case ST_TRACE:
      for (int d = 0; d < 3; ++d) set(d, L[d] - 1, H[d] + 1);
      break;
    case ST_FLUX1_X: case ST_FLUX1_Y: case ST_FLUX1_Z: {
      const int D = st - ST_FLUX1_X;
      for (int d = 0; d < 3; ++d) set(d, L[d] - 1, H[d] + 1);
      set(D, L[D], H[D] + 1);
      break;
    }
    // both transverse fluxes of direction D in one launch under the first
    // slot, (D|t0) where t0 is in [L, H] and (D|t1) where t1 is; the second
    // slot is empty
    case ST_TRANS_XY: case ST_TRANS_YX: case ST_TRANS_ZX: {
      const int D = (st - ST_TRANS_XY) / 2;
      for (int d = 0; d < 3; ++d) set(d, L[d] - 1, H[d] + 1);
      set(D, L[D], H[D] + 1);
      break;
    }
    case ST_TRANS_XZ: case ST_TRANS_YZ: case ST_TRANS_ZY:
      for (int d = 0; d < 3; ++d) set(d, L[d], L[d] - 1);
      break;
    case ST_FINAL_X: case ST_FINAL_Y: case ST_FINAL_Z: {
      const int D = st - ST_FINAL_X;
      set(D, L[D], H[D] + 1);
      break;
    }
    case ST_MOL_PRIM:
      for (int d = 0; d < 3; ++d) set(d, L[d] - 2, H[d] + 2);
      break;
    case ST_MOL_X: case ST_MOL_Y: case ST_MOL_Z: {
      const int D = st - ST_MOL_X;
      set(D, L[D], H[D] + 1);
      break;
    }
    case ST_UPDATE:
    default:
      break;
  }
  for (int d = 0; d < 3; ++d) { lo[d] = a[d]; n[d] = b[d] - a[d] + 1; }
}

// ----------------------------------------------------------------- loads --
template <int NS, class FP, class VW>
PCD void load_trace(const VW& tr, int i, int j, int k, int side,
                    FaceState<NS, FP>& s) {
  constexpr int QT = 6 + NS;
  const int o = side * QT;
  s.rho = tr(i, j, k, o + 0);
  s.u = tr(i, j, k, o + 1);
  s.v = tr(i, j, k, o + 2);
  s.w = tr(i, j, k, o + 3);
  s.p = tr(i, j, k, o + 4);
  s.rhoe = FP(0.0);  // not stored (store_trace)
#pragma unroll
  for (int n = 0; n < NS; ++n) s.Y[n] = tr(i, j, k, o + 6 + n);
}
// A traced state's rho e (slot 5) is not stored: the transverse and final
// corrections carry it into their corrected states, but the Riemann solver
// and cmpflx read rho, the velocities, p and the species only
// (Source/Godunov.H:303-449, Source/Riemann.H:9-200), so no flux depends on
// it.
template <int NS, class FP>
PCD void store_trace(const VT<FP>& tr, int i, int j, int k, int side,
                     const FaceState<NS, FP>& s) {
  constexpr int QT = 6 + NS;
  const int o = side * QT;
  tr(i, j, k, o + 0) = s.rho;
  tr(i, j, k, o + 1) = s.u;
  tr(i, j, k, o + 2) = s.v;
  tr(i, j, k, o + 3) = s.w;
  tr(i, j, k, o + 4) = s.p;
#pragma unroll
  for (int n = 0; n < NS; ++n) tr(i, j, k, o + 6 + n) = s.Y[n];
}
template <int NS, class FP, class VW>
PCD void store_flux(const VW& f, int i, int j, int k, const Flux<NS, FP>& F) {
  f(i, j, k, 0) = F.f[0];
  f(i, j, k, 1) = F.f[1];
  f(i, j, k, 2) = F.f[2];
  f(i, j, k, 3) = F.f[3];
  // slot 4, the energy flux, is not stored: in the transverse corrections it
  // only updates the corrected states' rho e, which no flux reads
  // (store_trace); without it the Riemann solves of the first and transverse
  // fluxes need no EOS call at the Godunov state
  f(i, j, k, 5) = F.pgd;
  f(i, j, k, 6) = F.ugd;
#pragma unroll
  for (int n = 0; n < NS; ++n) f(i, j, k, 7 + n) = F.f[6 + n];
}

template <int D>
PCD void step(int& i, int& j, int& k, int s) {
  if (D == 0) i += s;
  else if (D == 1) j += s;
  else k += s;
}

// ------------------------------------------------------------ K1 ctoprim --
// Source/PeleCUtilities.H:242-320: the primitives and auxiliaries of cell
// (i,j,k) from its conserved state Uc (every component but UEINT is read).
// MOLQ: for the MOL fluxes, which read rho, the velocities, p, the species
// and c only (QREINT and QGAMC are not stored).
template <class EOS, class FP, bool MOLQ = false>
PCD void ctoprim_cell(const JobT<FP>& J, const KArgs& ka, int i, int j, int k,
                      const double* Uc) {
  constexpr int NS = EOS::NS;
  const auto& q = J.q;
  const auto& qa = J.qa;
  const double rho = Uc[URHO];
  const double rhoinv = 1.0 / rho;
  const double vx = Uc[UMX] * rhoinv;
  const double vy = Uc[UMY] * rhoinv;
  const double vz = Uc[UMZ] * rhoinv;
  const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz) -
                        0.5 * rho * 0.0 * 0.0 * 0.0 * 0.0;
  double Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double y = div_by(Uc[UFS + s], rho, rhoinv);
    if ((-1e-4 * DEPS < y) && (y < 1e-4 * DEPS)) y = 0.0;
    Y[s] = y;
  }
  const double e = (Uc[UEDEN] - kineng) * rhoinv;
  const typename EOS::Thermo th(Y);
  const double T = th.EY2T(e, Uc[UTEMP]);
  const double p = th.RTY2P(rho, T);
  const double cs = th.RTY2Cs(rho, T);
  const double gam1 = th.RTY2G(rho, T);
  q(i, j, k, QRHO) = rho;
  q(i, j, k, QU) = vx;
  q(i, j, k, QV) = vy;
  q(i, j, k, QW) = vz;
#pragma unroll
  for (int s = 0; s < NS; ++s) q(i, j, k, QFS + s) = Y[s];
  if (!MOLQ && !ka.skip_reint) q(i, j, k, QREINT) = e * rho;
  q(i, j, k, QPRES) = p;
  if (!MOLQ) qa(i, j, k, QGAMC) = gam1;
  qa(i, j, k, QC) = cs;
  if (ka.full_prims) {
    const double w = th.wbar();
    double dpde, dpdr_e;
    th.RTY2dpde_dpdre(rho, T, dpde, dpdr_e);
    q(i, j, k, QTEMP) = T;
    q(i, j, k, QGAME) = p / (e * rho) + 1.0;
    qa(i, j, k, QDPDR) = dpdr_e;
    qa(i, j, k, QDPDE) = dpde;
    qa(i, j, k, QCSML) = mx(small_num, small_num * cs);
    qa(i, j, k, QRSPEC) = RU / w;
  }
}

// (REG: the region's stage, ST_CTOPRIM or the MOL step's ST_MOL_PRIM)
template <class EOS, class FP, int REG>
PCD void b_ctoprim(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NVAR = 7 + EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(REG, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  double Uc[NVAR];
#pragma unroll
  for (int m = 0; m < NVAR; ++m)
    if (m != UEINT) Uc[m] = J.U(i, j, k, m);
  ctoprim_cell<EOS, FP, REG == ST_MOL_PRIM>(J, ka, i, j, k, Uc);
}
template <class EOS, class FP = double, bool PERSIST = false,
          int REG = ST_CTOPRIM>
__global__ void __launch_bounds__(BLOCK)
    k_ctoprim(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_ctoprim<EOS, FP, REG>(sc, ka, lb_);
  } else {
    b_ctoprim<EOS, FP, REG>(sc, ka, first_block(sc));
  }
}

// ---------------------------------------------------------- K2 srctoprim --
// Source/Hydro.H:14-68 (src: NVAR comps over the q box)
template <int NS>
PCD void b_srctoprim(const Sched& sc, const KArgs& ka, int lb_, V src) {
  int local;
  const Job J = load_job(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_CTOPRIM, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const V& q = J.q;
  const V& qa = J.qa;
  const V& srcq = J.srcq;
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
  const double sre = src(i, j, k, UEDEN) - vx * umx - vy * umy - vz * umz +
                     srcrho * (vx * vx + vy * vy + vz * vz);
  srcq(i, j, k, QREINT) = sre;
  srcq(i, j, k, QPRES) =
      qa(i, j, k, QDPDE) * (sre - q(i, j, k, QREINT) * srcrho * rhoinv) *
          rhoinv +
      qa(i, j, k, QDPDR) * srcrho;
#pragma unroll
  for (int s = 0; s < NS; s++)
    srcq(i, j, k, QFS + s) =
        (src(i, j, k, UFS + s) - q(i, j, k, QFS + s) * srcrho) * rhoinv;
}
template <int NS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_srctoprim(Sched sc, KArgs ka, V src) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_srctoprim<NS>(sc, ka, lb_, src);
  } else {
    b_srctoprim<NS>(sc, ka, first_block(sc), src);
  }
}

// --------------------------------------------------------------- K24 divu --
// Source/Hydro.H:476-517 on the nodes of the valid box
template <class FP>
PCD FP flat_cell(const JobT<FP>& J, int i, int j, int k);

// Also the flattening coefficient of the cell (i,j,k) for the traces of all
// three directions (the reference recomputes it in each direction's trace).
template <class FP>
PCD void b_divu(const Sched& sc, const KArgs& ka, int lb_) {
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_DIVU, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  if (ka.use_flattening) J.flat(i, j, k, 0) = flat_cell(J, i, j, k);
  if (i < J.lo[0] || j < J.lo[1] || k < J.lo[2]) return;
  const auto& q = J.q;
  const double ux =
      0.25 *
      (q(i, j, k, QU) - q(i - 1, j, k, QU) + q(i, j, k - 1, QU) -
       q(i - 1, j, k - 1, QU) + q(i, j - 1, k, QU) - q(i - 1, j - 1, k, QU) +
       q(i, j - 1, k - 1, QU) - q(i - 1, j - 1, k - 1, QU)) /
      J.dx[0];
  const double vy =
      0.25 *
      (q(i, j, k, QV) - q(i, j - 1, k, QV) + q(i, j, k - 1, QV) -
       q(i, j - 1, k - 1, QV) + q(i - 1, j, k, QV) - q(i - 1, j - 1, k, QV) +
       q(i - 1, j, k - 1, QV) - q(i - 1, j - 1, k - 1, QV)) /
      J.dx[1];
  const double wz =
      0.25 *
      (q(i, j, k, QW) - q(i, j, k - 1, QW) + q(i - 1, j, k, QW) -
       q(i - 1, j, k - 1, QW) + q(i, j - 1, k, QW) - q(i, j - 1, k - 1, QW) +
       q(i - 1, j - 1, k, QW) - q(i - 1, j - 1, k - 1, QW)) /
      J.dx[2];
  J.divu(i, j, k, 0) = ux + vy + wz;
}
template <class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_divu(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_divu<FP>(sc, ka, lb_);
  } else {
    b_divu<FP>(sc, ka, first_block(sc));
  }
}

// flattening coefficient of cell (i,j,k): min over the three directions
template <class FP>
PCD FP flat_cell(const JobT<FP>& J, int i, int j, int k) {
  const auto& q = J.q;
  FP flat = FP(1.0);
#pragma unroll
  for (int df = 0; df < 3; ++df) {
    int oi = (df == 0), oj = (df == 1), ok = (df == 2);
    const bool inside =
        q.contains(i + 3 * oi, j + 3 * oj, k + 3 * ok, J.qe[0], J.qe[1],
                   J.qe[2]) &&
        q.contains(i - 3 * oi, j - 3 * oj, k - 3 * ok, J.qe[0], J.qe[1],
                   J.qe[2]);
    auto pq = [&](int m) { return q(i + m * oi, j + m * oj, k + m * ok, QPRES); };
    auto uq = [&](int m) {
      return q(i + m * oi, j + m * oj, k + m * ok, QU + df);
    };
    flat = mn(flat, flatten_line<FP>(inside, pq, uq));
  }
  return flat;
}

// ------------------------------------------------------ K3-5 trace (PPM) --
// The PPM trace (Source/PPM.cpp:82-423) runs with its stencils staged in
// LDS: ppm_cell_states and the trace kernels of pc_trace_lds.h.

// --------------------------------------------------- K3' trace (PLM) ----
// Source/Godunov.cpp:139-180 (one direction per launch)
template <class EOS, int D, class FP>
PCD void b_trace_plm(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NS = EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_TRACE, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const auto& q = J.q;
  const FP dtdx = get_dt(J, ka) / J.dx[D];
  const FP flat = ka.use_flattening ? FP(J.flat(i, j, k, 0)) : FP(1.0);
  auto slope = [&](int comp) {
    FP s[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      int ii = i, jj = j, kk = k;
      step<D>(ii, jj, kk, m - 2);
      s[m] = q(ii, jj, kk, comp);
    }
    return plm_slope(s, flat, ka.plm_iorder);
  };
  FaceState<NS, FP> c, sl, P, M;
  c.rho = q(i, j, k, QRHO);
  c.u = q(i, j, k, QU);
  c.v = q(i, j, k, QV);
  c.w = q(i, j, k, QW);
  c.p = q(i, j, k, QPRES);
  c.rhoe = q(i, j, k, QREINT);
  sl.rho = slope(QRHO);
  sl.u = slope(QU);
  sl.v = slope(QV);
  sl.w = slope(QW);
  sl.p = slope(QPRES);
  sl.rhoe = slope(QREINT);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    c.Y[s] = q(i, j, k, QFS + s);
    sl.Y[s] = slope(QFS + s);
  }
  plm_trace<NS>(D, c, sl, J.qa(i, j, k, QC), dtdx, P, M);
  const auto& tr = J.tr[D];
  store_trace<NS, FP>(tr, i, j, k, 1, P);
  int i1 = i, j1 = j, k1 = k;
  step<D>(i1, j1, k1, 1);
  store_trace<NS, FP>(tr, i1, j1, k1, 0, M);
}
template <class EOS, int D, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_trace_plm(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_trace_plm<EOS, D, FP>(sc, ka, lb_);
  } else {
    b_trace_plm<EOS, D, FP>(sc, ka, first_block(sc));
  }
}

// -------------------------------------------------- K6-8 first fluxes ----
template <class EOS, int D, class FP>
PCD void b_flux1(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NS = EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_FLUX1_X + D, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  FaceState<NS, FP> L, R;
  load_trace<NS, FP>(J.tr[D], i, j, k, 0, L);
  load_trace<NS, FP>(J.tr[D], i, j, k, 1, R);
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  const FP cav = FP(0.5) * (J.qa(i, j, k, QC) + J.qa(im, jm, km, QC));
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  Flux<NS, FP> F;
  cmpflx<EOS, D>(L, R, cav, idx, J.dlo[D], J.dhi[D], J.bclo[D], J.bchi[D], F);
  store_flux<NS, FP>(J.f1[D], i, j, k, F);
}
template <class EOS, int D, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_flux1(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_flux1<EOS, D, FP>(sc, ka, lb_);
  } else {
    b_flux1<EOS, D, FP>(sc, ka, first_block(sc));
  }
}

// The transverse correction a cell's first fluxes in direction T make to the
// traces of the other two directions (pc_transdo's flux differences and its
// Godunov pressure term, Source/Godunov.H:455-633): cdtdx (F_hi - F_lo) of
// rho, mx, my, mz, rhoE and rho Y, and cdtdx (dAup + pav dAu (Gamma1 - 1)),
// in the order pc_transdo forms them.  Stored per cell in dl[T] with the
// components d[0..4], ds[0..NS), pt.
template <int NS, class FP>
PCD void make_tcorr(const Flux<NS, FP>& lo, const Flux<NS, FP>& hi, FP gamc,
                    FP c, TCorr<NS, FP>& D) {
#pragma unroll
  for (int m = 0; m < 4; ++m) D.d[m] = c * (hi.f[m] - lo.f[m]);
  D.d[4] = FP(0.0);  // the energy flux difference feeds only rho e (store_flux)
#pragma unroll
  for (int s = 0; s < NS; ++s) D.ds[s] = c * (hi.f[6 + s] - lo.f[6 + s]);
  const FP dAup = hi.pgd * hi.ugd - lo.pgd * lo.ugd;
  const FP pav = FP(0.5) * (hi.pgd + lo.pgd);
  const FP dAu = hi.ugd - lo.ugd;
  D.pt = c * (dAup + pav * dAu * (gamc - FP(1.0)));
}
template <int NS, class FP, class VW>
PCD void store_tcorr(const VW& v, int i, int j, int k, const TCorr<NS, FP>& D) {
#pragma unroll
  for (int m = 0; m < 4; ++m) v(i, j, k, m) = D.d[m];  // slot 4 not stored (make_tcorr)
#pragma unroll
  for (int s = 0; s < NS; ++s) v(i, j, k, 5 + s) = D.ds[s];
  v(i, j, k, 5 + NS) = D.pt;
}
template <int NS, class FP, class VW>
PCD void load_tcorr(const VW& v, int i, int j, int k, TCorr<NS, FP>& D) {
#pragma unroll
  for (int m = 0; m < 4; ++m) D.d[m] = v(i, j, k, m);
  D.d[4] = FP(0.0);
#pragma unroll
  for (int s = 0; s < NS; ++s) D.ds[s] = v(i, j, k, 5 + s);
  D.pt = v(i, j, k, 5 + NS);
}
// a face's flux from a compact flux array (the components make_tcorr reads)
template <int NS, class FP, class VW>
PCD void load_flux(const VW& f, int i, int j, int k, Flux<NS, FP>& F) {
#pragma unroll
  for (int m = 0; m < 4; ++m) F.f[m] = f(i, j, k, m);
  F.f[4] = FP(0.0);  // not stored (store_flux)
  F.pgd = f(i, j, k, 5);
  F.ugd = f(i, j, k, 6);
#pragma unroll
  for (int s = 0; s < NS; ++s) F.f[6 + s] = f(i, j, k, 7 + s);
}

// ------------------------------------------ K8' per-cell corrections ----
// The first fluxes' transverse corrections of the cells [L, H] along D
// (x [L-1, H+1] across) from the face fluxes f1[D] (the unfused path; the
// fused trace kernel forms them from the faces in its workgroup).
template <int D, class FP, int NS>
PCD void b_tcorr(const Sched& sc, const KArgs& ka, int lb_) {
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_FLUX1_X + D, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  if (idx > J.hi[D]) return;  // the last face row has no cell above it
  int ip = i, jp = j, kp = k;
  step<D>(ip, jp, kp, 1);
  Flux<NS, FP> Flo, Fhi;
  load_flux<NS, FP>(J.f1[D], i, j, k, Flo);
  load_flux<NS, FP>(J.f1[D], ip, jp, kp, Fhi);
  const FP cdtdx = FP(1.0) / FP(3.0) * get_dt(J, ka) / J.dx[D];
  TCorr<NS, FP> C;
  make_tcorr<NS, FP>(Flo, Fhi, FP(J.qa(i, j, k, QGAMC)), cdtdx, C);
  store_tcorr<NS, FP>(J.dl[D], i, j, k, C);
}
template <int D, int NS, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_tcorr(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_tcorr<D, FP, NS>(sc, ka, lb_);
  } else {
    b_tcorr<D, FP, NS>(sc, ka, first_block(sc));
  }
}

// flux differences across cell c in direction T from a compact flux array
template <int NS, int T, class FP, class VW>
PCD void dflux(const VW& f, int i, int j, int k, FP c, DFlux<NS, FP>& D) {
  int ip = i, jp = j, kp = k;
  step<T>(ip, jp, kp, 1);
#pragma unroll
  for (int m = 0; m < 5; ++m) D.d[m] = c * (f(ip, jp, kp, m) - f(i, j, k, m));
#pragma unroll
  for (int s = 0; s < NS; ++s)
    D.ds[s] = c * (f(ip, jp, kp, 7 + s) - f(i, j, k, 7 + s));
  D.pgp = f(ip, jp, kp, 5);
  D.pgm = f(i, j, k, 5);
  D.ugp = f(ip, jp, kp, 6);
  D.ugm = f(i, j, k, 6);
}

// ------------------------------------- K9-16 transverse (d|t) fluxes ----
// One of two views of the job by a per-lane condition.  Each field is first
// pinned in a register (an empty asm), so the selection is made on the
// values: a select between the two fields' addresses would keep the job's
// copy in private (scratch) memory.
template <class T>
PCD VT<T> lane_select(bool c, const VT<T>& a, const VT<T>& b) {
  unsigned long long pa = (unsigned long long)a.p, pb = (unsigned long long)b.p;
  int fa[6] = {a.js, a.ks, a.ns, a.b0, a.b1, a.b2};
  int fb[6] = {b.js, b.ks, b.ns, b.b0, b.b1, b.b2};
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(pa), "+v"(pb));
#pragma unroll
  for (int m = 0; m < 6; ++m) asm("" : "+v"(fa[m]), "+v"(fb[m]));
#endif
  VT<T> r;
  r.p = (T*)(c ? pa : pb);
  r.js = c ? fa[0] : fb[0];
  r.ks = c ? fa[1] : fb[1];
  r.ns = c ? fa[2] : fb[2];
  r.b0 = c ? fa[3] : fb[3];
  r.b1 = c ? fa[4] : fb[4];
  r.b2 = c ? fa[5] : fb[5];
  return r;
}

// pc_transdo on the two cells of d-face (i,j,k) with the corrections of
// direction T (view dl), then pc_cmpflx (Source/Godunov.cpp:246-384).
template <class EOS, int D, class FP, class TV, class DV>
PCD void trans_face(const JobT<FP>& J, const TV& tr, const DV& dl, int i, int j,
                    int k, Flux<EOS::NS, FP>& F) {
  constexpr int NS = EOS::NS;
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  FaceState<NS, FP> s, L, R;
  TCorr<NS, FP> df;
  const VS<FP> qa = split(J.qa);
  // minus side: cell f - e_D, state on its high face
  load_trace<NS, FP>(tr, i, j, k, 0, s);
  load_tcorr<NS, FP>(dl, im, jm, km, df);
  transdo_side<NS>(s, df, L);
  // plus side: cell f, state on its low face
  load_trace<NS, FP>(tr, i, j, k, 1, s);
  load_tcorr<NS, FP>(dl, i, j, k, df);
  transdo_side<NS>(s, df, R);
  const FP cav = FP(0.5) * (qa(i, j, k, QC) + qa(im, jm, km, QC));
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  cmpflx<EOS, D>(L, R, cav, idx, J.dlo[D], J.dhi[D], J.bclo[D], J.bchi[D], F);
}

// Threads per face of the transverse stage: two in FP64 (below), one in FP32.
template <class FP>
__host__ __device__ constexpr int trans_lanes() {
  return sizeof(FP) == 8 ? 2 : 1;
}

// Both transverse fluxes of a d-face, (D|t0) and (D|t1), each where its t
// lies in [L, H].  FP64: a pair of neighbouring lanes per face, the even
// lane solving (D|t0), the odd lane (D|t1); the two lanes load the face's
// traces by the same instruction, so each trace is fetched from HBM once for
// both fluxes (one thread solving both read them twice, the second time
// mostly from HBM again), and each lane holds one pair of states; the face
// index of a thread is its local index halved.  FP32: one thread solves
// both, reading the face's traces again for the second flux (half the bytes:
// the second read finds them in L2).
template <class EOS, int D, class FP>
PCD void b_trans(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NS = EOS::NS;
  constexpr int T0 = (D == 0) ? 1 : 0, T1 = (D == 2) ? 1 : 2;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_TRANS_XY + 2 * D, J.lo, J.hi, lo, n);
  (void)ka;
  const VS<FP> tr = split(J.tr[D]);
  if constexpr (trans_lanes<FP>() == 2) {
    const bool second = (local & 1) != 0;
    if (!region_ijk(sc, lo, n, local >> 1, i, j, k)) return;
    const int c[3] = {i, j, k};
    const int ct = second ? c[T1] : c[T0];
    const int tl = second ? J.lo[T1] : J.lo[T0], th = second ? J.hi[T1] : J.hi[T0];
    if (ct < tl || ct > th) return;
    const VT<FP> dl = lane_select(second, J.dl[T1], J.dl[T0]);
    Flux<NS, FP> F;
    trans_face<EOS, D, FP>(J, tr, dl, i, j, k, F);
    const VT<FP> ft =
        lane_select(second, J.ft[2 * D + tslot(D, T1)], J.ft[2 * D + tslot(D, T0)]);
    store_flux<NS, FP>(ft, i, j, k, F);
  } else {
    if (!region_ijk(sc, lo, n, local, i, j, k)) return;
    const int c[3] = {i, j, k};
    if (c[T0] >= J.lo[T0] && c[T0] <= J.hi[T0]) {
      Flux<NS, FP> F;
      trans_face<EOS, D, FP>(J, tr, split(J.dl[T0]), i, j, k, F);
      store_flux<NS, FP>(split(J.ft[2 * D + tslot(D, T0)]), i, j, k, F);
    }
    if (c[T1] >= J.lo[T1] && c[T1] <= J.hi[T1]) {
      Flux<NS, FP> F;
      trans_face<EOS, D, FP>(J, tr, split(J.dl[T1]), i, j, k, F);
      store_flux<NS, FP>(split(J.ft[2 * D + tslot(D, T1)]), i, j, k, F);
    }
  }
}
template <class EOS, int D, class FP = double, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_trans(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_trans<EOS, D, FP>(sc, ka, lb_);
  } else {
    b_trans<EOS, D, FP>(sc, ka, first_block(sc));
  }
}

// transdd flux differences across cell (i,j,k) for final direction D
template <int NS, int D, bool SRC, class FP>
PCD void final_corr(const JobT<FP>& J, int i, int j, int k, FP hdt,
                    FP c0, FP c1, const FaceState<NS, FP>& s,
                    FaceState<NS, FP>& o) {
  constexpr int T0 = (D == 0) ? 1 : 0;
  constexpr int T1 = (D == 2) ? 1 : 2;
  const auto& fx = J.ft[2 * T0 + tslot(T0, T1)];  // F_{T0|T1} on T0-faces
  const auto& fy = J.ft[2 * T1 + tslot(T1, T0)];  // F_{T1|T0} on T1-faces
  int i0 = i, j0 = j, k0 = k;
  step<T0>(i0, j0, k0, 1);
  int i1 = i, j1 = j, k1 = k;
  step<T1>(i1, j1, k1, 1);
  FP dd[5], dds[NS];
#pragma unroll
  for (int m = 0; m < 4; ++m)
    dd[m] = c0 * (fx(i0, j0, k0, m) - fx(i, j, k, m)) +
            c1 * (fy(i1, j1, k1, m) - fy(i, j, k, m));
  dd[4] = FP(0.0);  // the energy fluxes are not stored (store_flux)
#pragma unroll
  for (int s2 = 0; s2 < NS; ++s2)
    dds[s2] = c0 * (fx(i0, j0, k0, 7 + s2) - fx(i, j, k, 7 + s2)) +
              c1 * (fy(i1, j1, k1, 7 + s2) - fy(i, j, k, 7 + s2));
  const FP gamc = J.qa(i, j, k, QGAMC);
  const FP pggpx = fx(i0, j0, k0, 5), pggmx = fx(i, j, k, 5);
  const FP ugpx = fx(i0, j0, k0, 6), ugmx = fx(i, j, k, 6);
  const FP dAupx = pggpx * ugpx - pggmx * ugmx;
  const FP pavx = FP(0.5) * (pggpx + pggmx);
  const FP dAux = ugpx - ugmx;
  const FP pggpy = fy(i1, j1, k1, 5), pggmy = fy(i, j, k, 5);
  const FP ugpy = fy(i1, j1, k1, 6), ugmy = fy(i, j, k, 6);
  const FP dAupy = pggpy * ugpy - pggmy * ugmy;
  const FP pavy = FP(0.5) * (pggpy + pggmy);
  const FP dAuy = ugpy - ugmy;
  const FP pxnew = c0 * (dAupx + pavx * dAux * (gamc - FP(1.0)));
  const FP pynew = c1 * (dAupy + pavy * dAuy * (gamc - FP(1.0)));
  FP src[6 + NS];
  if (SRC) {
    src[0] = J.srcq(i, j, k, QRHO);
    src[1] = J.srcq(i, j, k, QU);
    src[2] = J.srcq(i, j, k, QV);
    src[3] = J.srcq(i, j, k, QW);
    src[4] = J.srcq(i, j, k, QREINT);
    src[5] = J.srcq(i, j, k, QPRES);
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) src[6 + s2] = J.srcq(i, j, k, QFS + s2);
  } else {
#pragma unroll
    for (int m = 0; m < 6 + NS; ++m) src[m] = FP(0.0);
  }
  transdd_side<NS>(s, dd, dds, pxnew, pynew, src, hdt, o);
}

// PLM state of cell (i,j,k) in direction D (pc_low_order_boundary)
template <class EOS, int D, class FP>
PCD void plm_cell(const JobT<FP>& J, const KArgs& ka, FP dtdx, int i, int j,
                  int k, FaceState<EOS::NS, FP>& P, FaceState<EOS::NS, FP>& M) {
  constexpr int NS = EOS::NS;
  const auto& q = J.q;
  const FP flat = ka.use_flattening ? FP(J.flat(i, j, k, 0)) : FP(1.0);
  auto slope = [&](int comp) {
    FP s[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      int ii = i, jj = j, kk = k;
      step<D>(ii, jj, kk, m - 2);
      s[m] = q(ii, jj, kk, comp);
    }
    return plm_slope(s, flat, ka.plm_iorder);
  };
  FaceState<NS, FP> c, sl;
  c.rho = q(i, j, k, QRHO);
  c.u = q(i, j, k, QU);
  c.v = q(i, j, k, QV);
  c.w = q(i, j, k, QW);
  c.p = q(i, j, k, QPRES);
  c.rhoe = q(i, j, k, QREINT);
  sl.rho = slope(QRHO);
  sl.u = slope(QU);
  sl.v = slope(QV);
  sl.w = slope(QW);
  sl.p = slope(QPRES);
  sl.rhoe = slope(QREINT);
  for (int s = 0; s < NS; ++s) {
    c.Y[s] = q(i, j, k, QFS + s);
    sl.Y[s] = slope(QFS + s);
  }
  plm_trace<NS>(D, c, sl, J.qa(i, j, k, QC), dtdx, P, M);
}

// ----------------------------- K17-22 + K25-27 final fluxes + adjust ----
// pc_transdd + pc_cmpflx (Source/Godunov.cpp:392-438), then pc_artif_visc,
// pc_norm_spec_flx and pc_ext_flx (Source/Hydro.cpp:380-413) on the face.
// Faces of Hard/UserBC domain boundaries are left to k_final_loworder.
template <class FP>
PCD bool loworder_face(const JobT<FP>& J, int D, int idx) {
  return (idx == J.dlo[D] && (J.bclo[D] == BC_INFLOW || J.bclo[D] == BC_USER)) ||
         (idx == J.dhi[D] + 1 &&
          (J.bchi[D] == BC_INFLOW || J.bchi[D] == BC_USER));
}

// FULL: store every face quantity the box entry points' callers read.  The
// level and batch steps store neither the zero UTEMP flux nor the UEINT flux
// and qec (p_gd, u_gd): their update resets UEINT to UEDEN - rho |u|^2 / 2
// (reset_internal_energy with allow_small_energy = allow_negative_energy =
// 1), so the UEINT flux divergence and pdivu never reach the new state.
template <class EOS, int D, class FP, bool FULL = true>
PCD void final_tail(const JobT<FP>& J, const KArgs& ka, int i, int j, int k,
                    const FaceState<EOS::NS, FP>& L, const FaceState<EOS::NS, FP>& R) {
  constexpr int NS = EOS::NS;
  constexpr int NVAR = 7 + NS;
  constexpr int T0 = (D == 0) ? 1 : 0;
  constexpr int T1 = (D == 2) ? 1 : 2;
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  const FP cav = FP(0.5) * (J.qa(i, j, k, QC) + J.qa(im, jm, km, QC));
  Flux<NS, FP> F;
  cmpflx<EOS, D>(L, R, cav, idx, J.dlo[D], J.dhi[D], J.bclo[D], J.bchi[D], F);
  // lab-frame NVAR layout (UTEMP flux starts at zero)
  double fl[NVAR];
  fl[URHO] = F.f[0];
  fl[UMX] = F.f[1];
  fl[UMY] = F.f[2];
  fl[UMZ] = F.f[3];
  fl[UEDEN] = F.f[4];
  fl[UEINT] = F.f[5];
  fl[UTEMP] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) fl[UFS + s] = F.f[6 + s];
  // pc_artif_visc (Source/Godunov.H:1067-1138)
  {
    const auto& dv = J.divu;
    int a0 = i, b0 = j, e0 = k;
    step<T0>(a0, b0, e0, 1);
    int a1 = i, b1 = j, e1 = k;
    step<T1>(a1, b1, e1, 1);
    int a2 = a0, b2 = b0, e2 = e0;
    step<T1>(a2, b2, e2, 1);
    const double div =
        ka.difmag * mn(0.0, 0.25 * (dv(i, j, k, 0) + dv(a0, b0, e0, 0) +
                                    dv(a1, b1, e1, 0) + dv(a2, b2, e2, 0)));
    const bool at_bndry = (idx == J.dlo[D] && J.bclo[D] == BC_NOSLIPWALL) ||
                          (idx == J.dhi[D] + 1 && J.bchi[D] == BC_NOSLIPWALL);
    const double dxd = J.dx[D];
#pragma unroll
    for (int m = 0; m < NVAR; ++m) {
      if (m == UTEMP) continue;
      const bool tang = (m == UMX + T0) || (m == UMX + T1);
      if (!(at_bndry && tang))
        fl[m] += dxd * div * (J.U(i, j, k, m) - J.U(im, jm, km, m));
    }
    fl[UTEMP] = 0.0;
  }
  // pc_norm_spec_flx (Source/Hydro.H:835-849)
  {
    double summ = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) summ += fl[UFS + s];
    const double fac = (summ != 0.0) ? fl[URHO] / summ : 1.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) fl[UFS + s] *= fac;
  }
  // pc_ext_flx (Source/Hydro.H:854-865)
  const double area = J.area[D].p ? J.area[D](i, j, k, 0) : J.area_u[D];
  if constexpr (sizeof(FP) == 8) {
    const auto& fo = J.flx[D];
#pragma unroll
    for (int m = 0; m < NVAR; ++m)
      if (FULL || (m != UTEMP && m != UEINT)) fo(i, j, k, m) = fl[m] * area;
    if (FULL) {
      J.qec[D](i, j, k, 0) = F.pgd;
      J.qec[D](i, j, k, 1) = F.ugd;
    }
  } else {
    // mixed step (level and batch): the face fluxes in FP32, half the bytes
    // written here and read by the update; the density flux is not stored
    // (the update takes the sum of the species fluxes, so sum_k rho Y_k =
    // rho holds to FP64 rounding), nor are the zero UTEMP flux, the UEINT
    // flux and qec (above)
    const VT<float> fo = as_fp32(J.flx[D]);
#pragma unroll
    for (int m = 0; m < NVAR; ++m)
      if (m != URHO && m != UTEMP && m != UEINT) fo(i, j, k, m) = (float)(fl[m] * area);
  }
}

template <class EOS, int D, bool SRC, class FP, bool FULL>
PCD void b_final(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NS = EOS::NS;
  constexpr int T0 = (D == 0) ? 1 : 0;
  constexpr int T1 = (D == 2) ? 1 : 2;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_FINAL_X + D, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  if (loworder_face(J, D, idx)) return;
  const FP dt = get_dt(J, ka);
  const FP hdt = FP(0.5) * dt;
  const FP c0 = FP(0.5) * dt / J.dx[T0];
  const FP c1 = FP(0.5) * dt / J.dx[T1];
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  FaceState<NS, FP> L, R, s;
  load_trace<NS, FP>(J.tr[D], i, j, k, 0, s);
  final_corr<NS, D, SRC, FP>(J, im, jm, km, hdt, c0, c1, s, L);
  load_trace<NS, FP>(J.tr[D], i, j, k, 1, s);
  final_corr<NS, D, SRC, FP>(J, i, j, k, hdt, c0, c1, s, R);
  final_tail<EOS, D, FP, FULL>(J, ka, i, j, k, L, R);
}
template <class EOS, int D, bool SRC, class FP = double, bool PERSIST = false,
          bool FULL = true>
__global__ void __launch_bounds__(BLOCK)
    k_final(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_final<EOS, D, SRC, FP, FULL>(sc, ka, lb_);
  } else {
    b_final<EOS, D, SRC, FP, FULL>(sc, ka, first_block(sc));
  }
}

// pc_low_order_boundary (Source/Godunov.cpp:7-55,440-468): at Hard/UserBC
// domain faces the flux comes from first-order-in-time PLM states without
// transverse terms.  Launched only when such a boundary exists.
template <class EOS, int D, class FP, bool FULL>
PCD void b_final_loworder(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NS = EOS::NS;
  int local;
  const JobT<FP> J = load_job<FP>(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_FINAL_X + D, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  const int idx = (D == 0) ? i : ((D == 1) ? j : k);
  if (!loworder_face(J, D, idx)) return;
  const FP dtdx = get_dt(J, ka) / J.dx[D];
  int im = i, jm = j, km = k;
  step<D>(im, jm, km, -1);
  FaceState<NS, FP> L, R, unused;
  plm_cell<EOS, D, FP>(J, ka, dtdx, im, jm, km, unused, L);
  plm_cell<EOS, D, FP>(J, ka, dtdx, i, j, k, R, unused);
  final_tail<EOS, D, FP, FULL>(J, ka, i, j, k, L, R);
}
template <class EOS, int D, class FP = double, bool PERSIST = false, bool FULL = true>
__global__ void __launch_bounds__(BLOCK)
    k_final_loworder(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_final_loworder<EOS, D, FP, FULL>(sc, ka, lb_);
  } else {
    b_final_loworder<EOS, D, FP, FULL>(sc, ka, first_block(sc));
  }
}

// pdivu (Source/Godunov.H:1033-1060) and the flux divergence of one cell
// (Source/Hydro.H:888-907); h receives dU/dt.
// LEVEL: the level and batch updates, whose final fluxes store neither the
// zero UTEMP flux (adding its (0 - 0 + 0 - 0 + 0 - 0) / v = +0.0 changes no
// bit) nor the UEINT flux and qec (the update resets UEINT; final_tail):
// h[UTEMP] and h[UEINT] are left as they are.
template <int NS, bool LEVEL = false>
PCD void cell_divergence(const Job& J, const KArgs& ka, int i, int j, int k,
                         double* h) {
  constexpr int NVAR = 7 + NS;
  auto A = [&](int d, int a, int b, int c) {
    return J.area[d].p ? J.area[d](a, b, c, 0) : J.area_u[d];
  };
  const double v = J.vol.p ? J.vol(i, j, k, 0) : J.vol_u;
  const VB<double> f0 = bufview(J.flx[0]);
  const VB<double> f1 = bufview(J.flx[1]);
  const VB<double> f2 = bufview(J.flx[2]);
#pragma unroll
  for (int m = 0; m < NVAR; ++m) {
    if (LEVEL && (m == UTEMP || m == UEINT)) continue;
    h[m] += (f0(i, j, k, m) - f0(i + 1, j, k, m) + f1(i, j, k, m) -
             f1(i, j + 1, k, m) + f2(i, j, k, m) - f2(i, j, k + 1, m)) /
            v;
  }
  if constexpr (!LEVEL) {
    const V& q1 = J.qec[0];
    const V& q2 = J.qec[1];
    const V& q3 = J.qec[2];
    const double pdivu =
        0.5 *
        ((q1(i + 1, j, k, 0) + q1(i, j, k, 0)) *
             (q1(i + 1, j, k, 1) * A(0, i + 1, j, k) - q1(i, j, k, 1) * A(0, i, j, k)) +
         (q2(i, j + 1, k, 0) + q2(i, j, k, 0)) *
             (q2(i, j + 1, k, 1) * A(1, i, j + 1, k) - q2(i, j, k, 1) * A(1, i, j, k)) +
         (q3(i, j, k + 1, 0) + q3(i, j, k, 0)) *
             (q3(i, j, k + 1, 1) * A(2, i, j, k + 1) - q3(i, j, k, 1) * A(2, i, j, k))) /
        v;
    h[UEINT] -= pdivu;
  }
}

// The same, as the level update with LEVEL, for the mixed step's FP32 face
// fluxes (final_tail), in FP64 arithmetic: the density flux of a face is the
// sum of its species fluxes (species first to last).
template <int NS>
PCD void cell_divergence_fp32(const Job& J, const KArgs& ka, int i, int j, int k,
                              double* h) {
  constexpr int NVAR = 7 + NS;
  (void)ka;
  const double v = J.vol.p ? J.vol(i, j, k, 0) : J.vol_u;
  const VB<float> f[3] = {bufview(as_fp32(J.flx[0])), bufview(as_fp32(J.flx[1])),
                          bufview(as_fp32(J.flx[2]))};
  double drho = 0.0;
#pragma unroll
  for (int m = 0; m < NVAR; ++m) {
    if (m == URHO || m == UTEMP || m == UEINT) continue;
    double dm = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int a = i + (d == 0), b = j + (d == 1), c = k + (d == 2);
      dm += (double)f[d](i, j, k, m) - (double)f[d](a, b, c, m);
    }
    h[m] += dm / v;
    if (m >= UFS) drho += dm;
  }
  h[URHO] += drho / v;
}

// ----------------------------------------------------------- K28 consup --
template <int NS>
PCD void b_consup(const Sched& sc, const KArgs& ka, int lb_) {
  constexpr int NVAR = 7 + NS;
  int local;
  const Job J = load_job(sc, lb_, local);
  int lo[3], n[3], i, j, k;
  stage_region(ST_UPDATE, J.lo, J.hi, lo, n);
  if (!region_ijk(sc, lo, n, local, i, j, k)) return;
  double h[NVAR];
#pragma unroll
  for (int m = 0; m < NVAR; ++m) h[m] = J.upd(i, j, k, m);
  cell_divergence<NS>(J, ka, i, j, k, h);
#pragma unroll
  for (int m = 0; m < NVAR; ++m) J.upd(i, j, k, m) = h[m];
}
template <int NS, bool PERSIST = false>
__global__ void __launch_bounds__(BLOCK)
    k_consup(Sched sc, KArgs ka) {
  if constexpr (PERSIST) {
    for (int lb_ = first_block(); lb_ < nlogical(sc); lb_ += gridDim.x)
      b_consup<NS>(sc, ka, lb_);
  } else {
    b_consup<NS>(sc, ka, first_block(sc));
  }
}

}  // namespace pcd
