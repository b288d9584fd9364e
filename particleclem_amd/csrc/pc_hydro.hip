// Host side of the gfx950 Godunov hydro engine: contexts, stage workspaces,
// launch plans over box tables, the level stepper and the C ABI declared in
// include/pc_hydro.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pc_hydro.h"
#include "pc_derive_kernels.h"
#include "pc_fluxreg.h"
#include "pc_mol_kernels.h"
#include "pc_plotfile.h"
#include "pc_trace_lds.h"

using namespace pcd;

// ----------------------------------------------------------------- errors --
struct pc_ctx {
  int eos = 0;
  int ns = 0;
  int device = 0;
  std::string err;
  // drop-in scratch: one workspace per stream the box drivers are called on
  // (reused in stream order).  A caller may spread its boxes over several
  // streams, as AMReX's MFIter does on GPUs; calls on different streams then
  // run concurrently and must not share a workspace.
  struct Scratch {
    hipStream_t st = nullptr;
    double* p = nullptr;
    size_t bytes = 0;
  };
  static constexpr int MAXWS = 16;
  std::vector<Scratch> wsl;
  int ws_evict = 0;
  // job descriptors of the box drivers: a ring of device slots filled from
  // pinned host slots; a slot is reused once the event of its last use fired
  static constexpr int RING = 64;
  Job* djobs = nullptr;
  Job* hjobs = nullptr;
  hipEvent_t jev[RING];
  bool jev_used[RING] = {};
  int jnext = 0;
};

static int fail(pc_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(ctx, call)                                                   \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(ctx, PC_ERR_HIP, "%s failed: %s", #call,                  \
                  hipGetErrorString(e_));                                   \
  } while (0)

// ------------------------------------------------------------------ boxes --
struct IBox {
  int lo[3], hi[3];
  long long npts() const {
    return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) *
           (hi[2] - lo[2] + 1);
  }
};
static IBox mkbox(const int* lo, const int* hi) {
  IBox b;
  for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
  return b;
}
static IBox growb(IBox b, int n) {
  for (int d = 0; d < 3; ++d) { b.lo[d] -= n; b.hi[d] += n; }
  return b;
}

// view of a contiguous FAB over box b with nc comps at p
static V fab_view(double* p, const IBox& b) {
  V v;
  v.p = p;
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1,
            nz = b.hi[2] - b.lo[2] + 1;
  v.js = nx;
  v.ks = nx * ny;
  v.ns = nx * ny * nz;
  v.b0 = b.lo[0];
  v.b1 = b.lo[1];
  v.b2 = b.lo[2];
  return v;
}
static V null_view() {
  V v;
  memset(&v, 0, sizeof(v));
  return v;
}

// Largest FAB the kernels address: views index elements with 32-bit ints,
// and the buffer-load views (VB) give the hardware the whole array as one
// buffer whose byte offsets (component offset + cell offset) must stay below
// 2^31.  Every array a kernel reads or writes is held to this many bytes.
constexpr long long MAX_FAB_BYTES = 2147483647LL;

// Array4 -> V; fails when the FAB spans more than MAX_FAB_BYTES (its byte
// offsets would not fit the kernels' 32-bit offsets) or the view does not
// cover `need` with ncomp >= nc.
static int a4_view(pc_ctx* c, const pc_array4_t& a, const IBox& need, int nc,
                   const char* name, V& v) {
  if (!a.p) return fail(c, PC_ERR_ARG, "%s: null pointer", name);
  for (int d = 0; d < 3; ++d)
    if (need.lo[d] < a.begin[d] || need.hi[d] >= a.end[d])
      return fail(c, (nc > 0 && strcmp(name, "flx") && strcmp(name, "a") &&
                      strcmp(name, "vol") && strcmp(name, "uout") &&
                      strcmp(name, "molsrc"))
                         ? PC_ERR_GROW
                         : PC_ERR_ARG,
                  "%s does not cover the required box in dir %d", name, d);
  if (a.ncomp < nc)
    return fail(c, PC_ERR_ARG, "%s: ncomp %d < %d", name, a.ncomp, nc);
  const long long ext = (long long)(a.end[0] - a.begin[0]) +
                        a.jstride * (a.end[1] - a.begin[1]) +
                        a.kstride * (a.end[2] - a.begin[2]) +
                        a.nstride * (long long)(a.ncomp > 0 ? a.ncomp : 1);
  if (a.jstride < 0 || a.kstride < 0 || a.nstride < 0 ||
      ext > MAX_FAB_BYTES / (long long)sizeof(double))
    return fail(c, PC_ERR_UNSUPPORTED,
                "%s: FAB of %lld doubles exceeds the %lld bytes the kernels "
                "address with 32-bit offsets", name, ext, MAX_FAB_BYTES);
  v.p = a.p;
  v.js = (int)a.jstride;
  v.ks = (int)a.kstride;
  v.ns = (int)a.nstride;
  v.b0 = a.begin[0];
  v.b1 = a.begin[1];
  v.b2 = a.begin[2];
  return PC_OK;
}

// ------------------------------------------------------ stage workspace --
// Sizes (doubles) of one box's intermediates; `full` adds q/qaux/srcq (the
// box driver computes them) and the final fluxes (level path).
struct WsLayout {
  IBox q, tr[3], f1[3], dl[3], ft[6], fl[3], node, g1;
  size_t off_q = 0, off_qa = 0, off_srcq = 0, off_tr[3], off_f1[3], off_dl[3], off_ft[6],
         off_fl[3], off_qec[3], off_divu = 0, off_sold = 0, off_flat = 0,
         total = 0;
  bool mol = false;
};

// mol: the method-of-lines step's workspace instead of the CTU one:
// primitives, the three face-flux arrays and U^n of the valid box.
static WsLayout ws_layout(const IBox& bx, int ns, bool with_q, bool with_src,
                          bool with_flx, bool mol = false) {
  WsLayout L;
  const int QVAR = 8 + ns, NVAR = 7 + ns, QT = 6 + ns, FT = 7 + ns;
  size_t o = 0;
  auto take = [&](size_t n) {
    size_t r = o;
    o += (n + 31) & ~size_t(31);  // 256-byte aligned sub-arrays
    return r;
  };
  L.q = growb(bx, 4);
  if (with_q) {
    L.off_q = take(L.q.npts() * QVAR);
    L.off_qa = take(L.q.npts() * NQAUX);
    if (with_src) L.off_srcq = take(L.q.npts() * QVAR);
  }
  if (mol) {
    L.mol = true;
    for (int d = 0; d < 3; ++d) {
      IBox b = bx;
      b.hi[d] += 1;
      L.fl[d] = b;
      if (with_flx) L.off_fl[d] = take(b.npts() * NVAR);
    }
    L.off_sold = take(bx.npts() * NVAR);
    L.node = bx;
    L.total = o;
    return L;
  }
  for (int d = 0; d < 3; ++d) {
    IBox b = growb(bx, 1);
    b.hi[d] += 1;  // faces [L-1, H+2]
    L.tr[d] = b;
    L.off_tr[d] = take(b.npts() * 2 * QT);
  }
  for (int d = 0; d < 3; ++d) {
    IBox b = growb(bx, 1);
    b.lo[d] = bx.lo[d];
    b.hi[d] = bx.hi[d] + 1;
    L.f1[d] = b;
    L.off_f1[d] = take(b.npts() * FT);
  }
  for (int d = 0; d < 3; ++d) {  // per-cell corrections: cells [L, H] along d
    IBox b = growb(bx, 1);
    b.lo[d] = bx.lo[d];
    b.hi[d] = bx.hi[d];
    L.dl[d] = b;
    L.off_dl[d] = take(b.npts() * (6 + ns));
  }
  for (int s = 0; s < 6; ++s) {
    const int D = s / 2;
    const int T = (s % 2 == 0) ? (D == 0 ? 1 : 0) : (D == 2 ? 1 : 2);
    const int Tp = 3 - D - T;
    IBox b = bx;
    b.hi[D] += 1;
    b.lo[Tp] -= 1;
    b.hi[Tp] += 1;
    L.ft[s] = b;
    L.off_ft[s] = take(b.npts() * FT);
  }
  for (int d = 0; d < 3; ++d) {
    IBox b = bx;
    b.hi[d] += 1;
    L.fl[d] = b;
    if (with_flx) L.off_fl[d] = take(b.npts() * NVAR);
    L.off_qec[d] = take(b.npts() * 2);
  }
  L.node = bx;
  for (int d = 0; d < 3; ++d) L.node.hi[d] += 1;
  L.off_divu = take(L.node.npts());
  L.g1 = growb(bx, 1);
  L.off_flat = take(L.g1.npts());
  L.total = o;
  return L;
}

// The largest array of a box's state and workspace in bytes: the state FAB
// (grow(bx,4) x NVAR) and every stage array of the layout (q x QVAR, traces x
// 2 QT, fluxes, ...), each of which one view addresses.
static long long box_max_array_bytes(const IBox& bx, int ns, const WsLayout& W) {
  const int QVAR = 8 + ns, NVAR = 7 + ns, QT = 6 + ns, FT = 7 + ns;
  long long m = growb(bx, 4).npts() * (long long)NVAR;
  auto mx_ = [&](long long v) { m = v > m ? v : m; };
  mx_(W.q.npts() * (long long)QVAR);
  for (int d = 0; d < 3; ++d) {
    mx_(W.fl[d].npts() * (long long)NVAR);
    if (!W.mol) {
      mx_(W.tr[d].npts() * 2LL * QT);
      mx_(W.f1[d].npts() * (long long)FT);
      mx_(W.dl[d].npts() * (long long)(6 + ns));
    }
  }
  if (!W.mol)
    for (int t = 0; t < 6; ++t) mx_(W.ft[t].npts() * (long long)FT);
  return m * (long long)sizeof(double);
}

static int ws_guard(pc_ctx* c, const IBox& bx, const WsLayout& W) {
  const long long big = box_max_array_bytes(bx, c->ns, W);
  if (big > MAX_FAB_BYTES)
    return fail(c, PC_ERR_UNSUPPORTED,
                "box: a stage array of %lld bytes exceeds the %lld the kernels "
                "address with 32-bit offsets", big, MAX_FAB_BYTES);
  return PC_OK;
}

static void ws_fill_job(Job& J, double* base, const WsLayout& L, bool with_q,
                        bool with_src, bool with_flx) {
  if (with_q) {
    J.q = fab_view(base + L.off_q, L.q);
    J.qa = fab_view(base + L.off_qa, L.q);
    J.srcq = with_src ? fab_view(base + L.off_srcq, L.q) : null_view();
  }
  if (L.mol) {
    for (int d = 0; d < 3; ++d) {
      if (with_flx) J.flx[d] = fab_view(base + L.off_fl[d], L.fl[d]);
      J.tr[d] = J.f1[d] = J.dl[d] = J.qec[d] = null_view();
    }
    for (int s = 0; s < 6; ++s) J.ft[s] = null_view();
    J.divu = J.flat = null_view();
    J.sold = fab_view(base + L.off_sold, L.node);
    return;
  }
  J.sold = null_view();
  J.flat = fab_view(base + L.off_flat, L.g1);
  for (int d = 0; d < 3; ++d) {
    J.tr[d] = fab_view(base + L.off_tr[d], L.tr[d]);
    J.f1[d] = fab_view(base + L.off_f1[d], L.f1[d]);
    J.dl[d] = fab_view(base + L.off_dl[d], L.dl[d]);
    if (with_flx) J.flx[d] = fab_view(base + L.off_fl[d], L.fl[d]);
    J.qec[d] = fab_view(base + L.off_qec[d], L.fl[d]);
  }
  for (int s = 0; s < 6; ++s) J.ft[s] = fab_view(base + L.off_ft[s], L.ft[s]);
  J.divu = fab_view(base + L.off_divu, L.node);
}

// Schedule of a launch over one job: one workgroup per block of cells.
static Sched one_job_sched(const Job* j) {
  return Sched{j, nullptr, nullptr, 0, -1, 0, 0, 1};
}

// Stage job J for the kernels of one box-driver call on stream st.
static int ctx_job(pc_ctx* c, const Job& J, hipStream_t st, Sched* sc,
                   int* slot) {
  if (!c->djobs) {
    if (hipMalloc(&c->djobs, sizeof(Job) * pc_ctx::RING) != hipSuccess ||
        hipHostMalloc(&c->hjobs, sizeof(Job) * pc_ctx::RING) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, PC_ERR_NOMEM, "job ring allocation failed");
    }
    for (int e = 0; e < pc_ctx::RING; ++e)
      HIPCHK(c, hipEventCreateWithFlags(&c->jev[e], hipEventDisableTiming));
  }
  const int s = c->jnext;
  c->jnext = (c->jnext + 1) % pc_ctx::RING;
  if (c->jev_used[s]) HIPCHK(c, hipEventSynchronize(c->jev[s]));
  c->hjobs[s] = J;
  c->hjobs[s].dtp = &c->djobs[s].dt_val;  // the call's dt travels in its slot
  HIPCHK(c, hipMemcpyAsync(c->djobs + s, c->hjobs + s, sizeof(Job),
                           hipMemcpyHostToDevice, st));
  *sc = one_job_sched(c->djobs + s);
  *slot = s;
  return PC_OK;
}

static int ctx_job_done(pc_ctx* c, int slot, hipStream_t st) {
  HIPCHK(c, hipEventRecord(c->jev[slot], st));
  c->jev_used[slot] = true;
  return PC_OK;
}

// The workspace of stream st, at least `bytes` large, in *base.
static int ctx_scratch(pc_ctx* c, size_t bytes, hipStream_t st, double** base) {
  int k = -1;
  for (int i = 0; i < (int)c->wsl.size(); ++i)
    if (c->wsl[i].st == st) k = i;
  if (k < 0) {
    if ((int)c->wsl.size() < pc_ctx::MAXWS) {
      c->wsl.push_back(pc_ctx::Scratch{});
      k = (int)c->wsl.size() - 1;
    } else {
      // more streams than workspaces: take one over once the device is idle
      // (its stream may no longer exist)
      k = c->ws_evict;
      c->ws_evict = (c->ws_evict + 1) % pc_ctx::MAXWS;
      HIPCHK(c, hipDeviceSynchronize());
    }
  }
  pc_ctx::Scratch& s = c->wsl[k];
  s.st = st;
  if (bytes > s.bytes) {
    if (s.p) {
      // the previous calls on this stream may still run
      HIPCHK(c, hipDeviceSynchronize());
      HIPCHK(c, hipFree(s.p));
      s.p = nullptr;
      s.bytes = 0;
    }
    const size_t nb = std::max(bytes, (size_t)64 << 20);
    if (hipMalloc(&s.p, nb) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, PC_ERR_NOMEM, "scratch allocation of %zu bytes failed", nb);
    }
    s.bytes = nb;
  }
  *base = s.p;
  return PC_OK;
}

// ----------------------------------------------------------- launching --
static long long region_npts(int st, const Job& J) {
  int lo[3], n[3];
  stage_region(st, J.lo, J.hi, lo, n);
  return (long long)n[0] * n[1] * n[2];
}

// launch plan of one stage over a job table
struct StagePlan {
  int nblocks = 0;          // logical blocks
  int grid = 0;             // workgroups launched (< nblocks: persistent)
  int lx = -1, ly = 0, lz = 0;  // tile shape (log2), lx = -1: runs
  int* blkmap = nullptr;    // device
  int* blkstart = nullptr;  // device
};

// logical blocks of one job for a stage and block shape
// lx = -2: the LDS tiles of the PPM trace, -3: the pencil tiles of the fused
// PPM trace + first fluxes (pc_trace_lds.h), -4: the face pencils of the MOL
// fluxes (pc_mol_kernels.h)
// fp32: the launch is the mixed-precision step's (the transverse stage's
// threads per face, trans_lanes, differ by precision)
static long long stage_blocks(int st, const Job& J, int lx, int ly, int lz,
                              bool fp32 = false) {
  const int tl = fp32 ? trans_lanes<float>() : trans_lanes<double>();
  if (lx == -2 || lx == -3) {
    int lo[3], n[3];
    stage_region(st, J.lo, J.hi, lo, n);
    return lx == -2 ? trace_lds_blocks(n) : pencil_blocks(n);
  }
  if (lx == -4) {  // the face pencils of the MOL fluxes in LDS
    int lo[3], n[3];
    stage_region(st, J.lo, J.hi, lo, n);
    const TraceTile t = mol_tile(st - ST_MOL_X, n);
    return (long long)t.nta * t.ntb * t.nc;
  }
  // the ghost fill enumerates its cells itself (launch shapes apply only
  // through persistence)
  if (st == ST_GHOST) return (ghost_shell_count(J.lo, J.hi) + BLOCK - 1) / BLOCK;
  int lo[3], n[3];
  stage_region(st, J.lo, J.hi, lo, n);
  if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return 0;
  const int per = (st == ST_TRANS_XY || st == ST_TRANS_YX || st == ST_TRANS_ZX) ? tl : 1;
  if (lx < 0) return (per * (long long)n[0] * n[1] * n[2] + BLOCK - 1) / BLOCK;
  auto nt = [](int m, int l) { return (long long)((m + (1 << l) - 1) >> l); };
  return per * nt(n[0], lx) * nt(n[1], ly) * nt(n[2], lz);
}

static int grid1(int st, const Job& J) {
  return (int)stage_blocks(st, J, -1, 0, 0);
}


// While a level step is captured into a hipGraph, every launch appends its
// stage tag (stage * 8 + direction, 7 for none) to t_tags, so the captured
// chain of kernel nodes can be rewired into the branch-stream dependencies.
static thread_local std::vector<int>* t_tags = nullptr;
static thread_local int t_tag = -1;
constexpr int TAG_DT = ST_COUNT * 8 + 7;
static inline void tag_launch() {
  if (t_tags) t_tags->push_back(t_tag);
}

#define LAUNCH(kern, grid, stream, ...)                                      \
  do {                                                                       \
    if ((grid) > 0) {                                                        \
      hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, stream,           \
                         __VA_ARGS__);                                       \
      tag_launch();                                                          \
    }                                                                        \
  } while (0)
#define LAUNCH_SH(kern, grid, shmem, stream, ...)                            \
  do {                                                                       \
    if ((grid) > 0) {                                                        \
      hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), shmem, stream,       \
                         __VA_ARGS__);                                       \
      tag_launch();                                                          \
    }                                                                        \
  } while (0)

// Largest LDS trace window (cells per component) over a set of jobs, for
// the trace tiles or (pencil) the fused trace + first-flux tiles.
static int trace_cells_of(const Job* jobs, int n, bool pencil = false) {
  int m = 0;
  for (int b = 0; b < n; ++b) {
    int lo[3], e[3];
    stage_region(ST_TRACE, jobs[b].lo, jobs[b].hi, lo, e);
    m = std::max(m, pencil ? pencil_cells(e) : trace_lds_cells(e));
  }
  return m;
}

// Whether every job's trace region fits the pencils of the fused PPM trace +
// first-flux kernel (else the trace and the first fluxes run separately).
static bool pencils_fit(const Job* jobs, int n) {
  for (int b = 0; b < n; ++b) {
    int lo[3], e[3];
    stage_region(ST_TRACE, jobs[b].lo, jobs[b].hi, lo, e);
    if (!pencil_fits(e)) return false;
  }
  return true;
}

// Whether every job's MOL face pencils fit one workgroup, and the largest
// MOL flux window (cells per component) over a set of jobs.
static bool mol_fit(const Job* jobs, int n) {
  for (int b = 0; b < n; ++b) {
    const int e[3] = {jobs[b].hi[0] - jobs[b].lo[0] + 1, jobs[b].hi[1] - jobs[b].lo[1] + 1,
                      jobs[b].hi[2] - jobs[b].lo[2] + 1};
    if (!mol_pencils_fit(e)) return false;
  }
  return true;
}
static int mol_cells_of(const Job* jobs, int n) {
  int m = 0;
  for (int b = 0; b < n; ++b)
    for (int D = 0; D < 3; ++D) {
      int lo[3], e[3];
      stage_region(ST_MOL_X + D, jobs[b].lo, jobs[b].hi, lo, e);
      m = std::max(m, mol_window_cells(D, e));
    }
  return m;
}

// Dynamic LDS bytes of the fused kernel: the staging window, then the minus
// states and the face fluxes of a tile's cells.
template <int NS, class FP>
static size_t pencil_lds_bytes(int cells) {
  return (size_t)std::max(cells * (5 + NS), BLOCK * (7 + NS)) * sizeof(FP);
}

// The CTU stages from primitives to final fluxes for one EOS.  With a job
// table (sc.jobs) the grids come from `plans`; otherwise from job j0.
template <class EOS>
static void run_ctu(const Job& j0, const Sched& s1, const KArgs& ka,
                    int ppm_type, hipStream_t st) {
  auto G = [&](int stage) { return grid1(stage, j0); };
  auto S = [&](int) { return s1; };
  LAUNCH(k_divu<>, G(ST_DIVU), st, S(ST_DIVU), ka);
  const bool fused = ppm_type == 1 && pencils_fit(&j0, 1);
  if (fused) {
    const int gt = (int)stage_blocks(ST_TRACE, j0, -3, 0, 0);
    const size_t sh = pencil_lds_bytes<EOS::NS, double>(trace_cells_of(&j0, 1, true));
    LAUNCH_SH((k_trace_flux1_lds<EOS, 0>), gt, sh, st, S(ST_TRACE), ka);
    LAUNCH_SH((k_trace_flux1_lds<EOS, 1>), gt, sh, st, S(ST_TRACE), ka);
    LAUNCH_SH((k_trace_flux1_lds<EOS, 2>), gt, sh, st, S(ST_TRACE), ka);
  } else if (ppm_type == 1) {
    const int gt = (int)stage_blocks(ST_TRACE, j0, -2, 0, 0);
    const size_t sh = (size_t)trace_cells_of(&j0, 1) * (5 + EOS::NS) * sizeof(double);
    LAUNCH_SH((k_trace_ppm_lds<EOS, 0>), gt, sh, st, S(ST_TRACE), ka);
    LAUNCH_SH((k_trace_ppm_lds<EOS, 1>), gt, sh, st, S(ST_TRACE), ka);
    LAUNCH_SH((k_trace_ppm_lds<EOS, 2>), gt, sh, st, S(ST_TRACE), ka);
  } else {
    LAUNCH((k_trace_plm<EOS, 0>), G(ST_TRACE), st, S(ST_TRACE), ka);
    LAUNCH((k_trace_plm<EOS, 1>), G(ST_TRACE), st, S(ST_TRACE), ka);
    LAUNCH((k_trace_plm<EOS, 2>), G(ST_TRACE), st, S(ST_TRACE), ka);
  }
  if (!fused) {
    LAUNCH((k_flux1<EOS, 0>), G(ST_FLUX1_X), st, S(ST_FLUX1_X), ka);
    LAUNCH((k_flux1<EOS, 1>), G(ST_FLUX1_Y), st, S(ST_FLUX1_Y), ka);
    LAUNCH((k_flux1<EOS, 2>), G(ST_FLUX1_Z), st, S(ST_FLUX1_Z), ka);
    LAUNCH((k_tcorr<0, EOS::NS>), G(ST_FLUX1_X), st, S(ST_FLUX1_X), ka);
    LAUNCH((k_tcorr<1, EOS::NS>), G(ST_FLUX1_Y), st, S(ST_FLUX1_Y), ka);
    LAUNCH((k_tcorr<2, EOS::NS>), G(ST_FLUX1_Z), st, S(ST_FLUX1_Z), ka);
  }
  LAUNCH((k_trans<EOS, 0>), G(ST_TRANS_XY), st, S(ST_TRANS_XY), ka);
  LAUNCH((k_trans<EOS, 1>), G(ST_TRANS_YX), st, S(ST_TRANS_YX), ka);
  LAUNCH((k_trans<EOS, 2>), G(ST_TRANS_ZX), st, S(ST_TRANS_ZX), ka);
  if (j0.srcq.p) {
    LAUNCH((k_final<EOS, 0, true>), G(ST_FINAL_X), st, S(ST_FINAL_X), ka);
    LAUNCH((k_final<EOS, 1, true>), G(ST_FINAL_Y), st, S(ST_FINAL_Y), ka);
    LAUNCH((k_final<EOS, 2, true>), G(ST_FINAL_Z), st, S(ST_FINAL_Z), ka);
  } else {
    LAUNCH((k_final<EOS, 0, false>), G(ST_FINAL_X), st, S(ST_FINAL_X), ka);
    LAUNCH((k_final<EOS, 1, false>), G(ST_FINAL_Y), st, S(ST_FINAL_Y), ka);
    LAUNCH((k_final<EOS, 2, false>), G(ST_FINAL_Z), st, S(ST_FINAL_Z), ka);
  }
  auto lowo = [&](int d) {
    auto is = [](int b) { return b == PC_BC_INFLOW || b == PC_BC_USERBC; };
    return is(j0.bclo[d]) || is(j0.bchi[d]);
  };
  if (lowo(0)) LAUNCH((k_final_loworder<EOS, 0>), G(ST_FINAL_X), st, S(ST_FINAL_X), ka);
  if (lowo(1)) LAUNCH((k_final_loworder<EOS, 1>), G(ST_FINAL_Y), st, S(ST_FINAL_Y), ka);
  if (lowo(2)) LAUNCH((k_final_loworder<EOS, 2>), G(ST_FINAL_Z), st, S(ST_FINAL_Z), ka);
}

// ------------------------------------------------------------------ C ABI --
// Workspace views plus the caller's flx/area/vol/uout views of a box-driver
// job; ws points at the CTU part of the context scratch.
static int finish_umdrv_job(pc_ctx* c, Job& J, const IBox& bx, double* ws,
                            const pc_array4_t* flx, const pc_array4_t* a,
                            const pc_array4_t& vol, const pc_array4_t& uout) {
  int rc;
  const WsLayout L = ws_layout(bx, c->ns, false, false, false);
  ws_fill_job(J, ws, L, false, false, false);
  for (int d = 0; d < 3; ++d) {
    IBox fb = bx;
    fb.hi[d] += 1;
    if ((rc = a4_view(c, flx[d], fb, 7 + c->ns, "flx", J.flx[d]))) return rc;
    if ((rc = a4_view(c, a[d], fb, 1, "a", J.area[d]))) return rc;
  }
  if ((rc = a4_view(c, vol, bx, 1, "vol", J.vol))) return rc;
  if ((rc = a4_view(c, uout, bx, 7 + c->ns, "uout", J.upd))) return rc;
  return PC_OK;
}

// uniform Cartesian geometry of a job (amrex::Geometry on a Cartesian grid:
// face areas dx1*dx2 etc., volume dx0*dx1*dx2)
static void set_geom(Job& J, const double* dx) {
  for (int d = 0; d < 3; ++d) J.dx[d] = dx[d];
  J.area_u[0] = dx[1] * dx[2];
  J.area_u[1] = dx[2] * dx[0];
  J.area_u[2] = dx[0] * dx[1];
  J.vol_u = dx[0] * dx[1] * dx[2];
}

static KArgs umdrv_args(const double* dx, double dt, int plm_iorder,
                        int use_flattening, double difmag) {
  KArgs ka;
  memset(&ka, 0, sizeof(ka));
  (void)dx;
  (void)dt;
  ka.difmag = difmag;
  ka.plm_iorder = plm_iorder;
  ka.use_flattening = use_flattening;
  return ka;
}

template <class EOS>
static void launch_ctu_consup(const Job& J, const Sched& sc, const KArgs& ka,
                              int ppm_type, hipStream_t st) {
  run_ctu<EOS>(J, sc, ka, ppm_type, st);
  LAUNCH(k_consup<EOS::NS>, grid1(ST_UPDATE, J), st, sc, ka);
}

extern "C" {

void pc_params_default(pc_params_t* p) {
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
  p->allow_small_energy = 1;
  p->allow_negative_energy = 1;
  p->dual_energy_update_E_from_e = 1;
  p->dual_energy_eta2 = 1.0e-4;
  p->do_mol = 0;
  p->mol_iorder = 2;
  p->use_laxf_flux = 0;
  p->stop_time = -1.0;
}

int pc_ctx_create(int eos_kind, pc_ctx_t** out) {
  if (!out) return PC_ERR_ARG;
  *out = nullptr;
  if (eos_kind != PC_EOS_GAMMALAW && eos_kind != PC_EOS_LIDRYER)
    return PC_ERR_ARG;
  pc_ctx* c = new pc_ctx();
  c->eos = eos_kind;
  c->ns = eos_kind;
  // no device is needed to create a context or validate arguments; the
  // first launch reports a missing device as PC_ERR_HIP
  if (hipGetDevice(&c->device) != hipSuccess) {
    (void)hipGetLastError();
    c->device = -1;
  }
  *out = c;
  return PC_OK;
}

int pc_ctx_destroy(pc_ctx_t* c) {
  if (!c) return PC_ERR_ARG;
  if (!c->wsl.empty() || c->djobs) (void)hipDeviceSynchronize();
  for (auto& s : c->wsl)
    if (s.p) (void)hipFree(s.p);
  if (c->djobs) {
    (void)hipFree(c->djobs);
    (void)hipHostFree(c->hjobs);
    for (int e = 0; e < pc_ctx::RING; ++e) (void)hipEventDestroy(c->jev[e]);
  }
  delete c;
  return PC_OK;
}

int pc_ctx_num_species(const pc_ctx_t* c) { return c ? c->ns : PC_ERR_ARG; }
int pc_ctx_nvar(const pc_ctx_t* c) { return c ? 7 + c->ns : PC_ERR_ARG; }
int pc_ctx_qvar(const pc_ctx_t* c) { return c ? 8 + c->ns : PC_ERR_ARG; }
const char* pc_last_error(const pc_ctx_t* c) {
  return c ? c->err.c_str() : "null context";
}

static bool box_ok(const pc_box_t& b) {
  return b.hi[0] >= b.lo[0] && b.hi[1] >= b.lo[1] && b.hi[2] >= b.lo[2];
}

int pc_ctoprim_v1(pc_ctx_t* c, pc_box_t qbx, pc_array4_t u, pc_array4_t q,
                  pc_array4_t qaux, void* stream) {
  if (!c || !box_ok(qbx)) return fail(c, PC_ERR_ARG, "bad context or box");
  const IBox b = mkbox(qbx.lo, qbx.hi);
  Job J;
  memset(&J, 0, sizeof(J));
  // the ctoprim region is grow(bx,4): express qbx as bx = shrink(qbx,4)
  for (int d = 0; d < 3; ++d) { J.lo[d] = b.lo[d] + 4; J.hi[d] = b.hi[d] - 4; }
  int rc;
  if ((rc = a4_view(c, u, b, 7 + c->ns, "u", J.U))) return rc;
  if ((rc = a4_view(c, q, b, 8 + c->ns, "q", J.q))) return rc;
  if ((rc = a4_view(c, qaux, b, NQAUX, "qaux", J.qa))) return rc;
  KArgs ka;
  memset(&ka, 0, sizeof(ka));
  ka.full_prims = 1;
  hipStream_t st = (hipStream_t)stream;
  const int g = (int)((b.npts() + BLOCK - 1) / BLOCK);
  Sched sc;
  int slot;
  if ((rc = ctx_job(c, J, st, &sc, &slot))) return rc;
  if (c->eos == PC_EOS_GAMMALAW)
    LAUNCH(k_ctoprim<EosGamma>, g, st, sc, ka);
  else
    LAUNCH(k_ctoprim<EosLiDryer>, g, st, sc, ka);
  HIPCHK(c, hipGetLastError());
  return ctx_job_done(c, slot, st);
}

int pc_srctoprim_v1(pc_ctx_t* c, pc_box_t qbx, pc_array4_t q, pc_array4_t qaux,
                    pc_array4_t src, pc_array4_t srcq, void* stream) {
  if (!c || !box_ok(qbx)) return fail(c, PC_ERR_ARG, "bad context or box");
  const IBox b = mkbox(qbx.lo, qbx.hi);
  Job J;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) { J.lo[d] = b.lo[d] + 4; J.hi[d] = b.hi[d] - 4; }
  int rc;
  V vs;
  if ((rc = a4_view(c, q, b, 8 + c->ns, "q", J.q))) return rc;
  if ((rc = a4_view(c, qaux, b, NQAUX, "qaux", J.qa))) return rc;
  if ((rc = a4_view(c, src, b, 7 + c->ns, "src", vs))) return rc;
  if ((rc = a4_view(c, srcq, b, 8 + c->ns, "srcq", J.srcq))) return rc;
  KArgs ka;
  memset(&ka, 0, sizeof(ka));
  hipStream_t st = (hipStream_t)stream;
  const int g = (int)((b.npts() + BLOCK - 1) / BLOCK);
  Sched sc;
  int slot;
  if ((rc = ctx_job(c, J, st, &sc, &slot))) return rc;
  if (c->eos == PC_EOS_GAMMALAW)
    LAUNCH(k_srctoprim<1>, g, st, sc, ka, vs);
  else
    LAUNCH(k_srctoprim<9>, g, st, sc, ka, vs);
  HIPCHK(c, hipGetLastError());
  return ctx_job_done(c, slot, st);
}

static int check_opts(pc_ctx* c, int ppm_type, int plm_iorder,
                      int use_hybrid_weno) {
  if (ppm_type != 0 && ppm_type != 1)
    return fail(c, PC_ERR_PPM_TYPE, "PeleC::ppm_type must be 0 (PLM) or 1 (PPM)");
  if (plm_iorder != 1 && plm_iorder != 2 && plm_iorder != 4)
    return fail(c, PC_ERR_PLM_IORDER, "plm_iorder must be 1, 2 or 4");
  if (use_hybrid_weno)
    return fail(c, PC_ERR_UNSUPPORTED, "use_hybrid_weno is not implemented");
  return PC_OK;
}

// Source/PeleC.cpp:384-392 with do_mol = 1: only mol_iorder is checked
// (ppm_type and plm_iorder are not; plm_iorder != 1 switches the slopes on)
static int check_mol_opts(pc_ctx* c, const pc_params_t* prm) {
  if (prm->mol_iorder != 1 && prm->mol_iorder != 2)
    return fail(c, PC_ERR_MOL_IORDER, "PeleC::mol_iorder must be 1, or 2.");
  if (prm->use_hybrid_weno)
    return fail(c, PC_ERR_UNSUPPORTED, "use_hybrid_weno is not implemented");
  return PC_OK;
}

// shared tail of pc_umdrv_v1 / pc_hydro_source_box_v1: J.q/qa/srcq/U are
// set; fills the stage workspace views and launches the CTU + consup.
static void set_domain(Job& J, const int* domlo, const int* domhi,
                       const int* bclo, const int* bchi) {
  for (int d = 0; d < 3; ++d) {
    J.dlo[d] = domlo[d];
    J.dhi[d] = domhi[d];
    J.bclo[d] = bclo[d];
    J.bchi[d] = bchi[d];
  }
}

int pc_umdrv_v1(pc_ctx_t* c, double /*time*/, pc_box_t bx, const int* domlo,
                const int* domhi, const int* bclo, const int* bchi,
                pc_array4_t uin, pc_array4_t uout, pc_array4_t q,
                pc_array4_t qaux, pc_array4_t src_q, const double* dx,
                double dt, int ppm_type, int plm_iorder, int use_flattening,
                int use_hybrid_weno, int /*weno_scheme*/, double difmag,
                const pc_array4_t* flx, const pc_array4_t* a, pc_array4_t vol,
                double* /*cflLoc*/, void* stream) {
  if (!c || !box_ok(bx) || !domlo || !domhi || !bclo || !bchi || !dx || !flx ||
      !a)
    return fail(c, PC_ERR_ARG, "null argument or empty box");
  int rc;
  if ((rc = check_opts(c, ppm_type, plm_iorder, use_hybrid_weno))) return rc;
  const IBox b = mkbox(bx.lo, bx.hi);
  const IBox g4 = growb(b, 4);
  Job J;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) { J.lo[d] = b.lo[d]; J.hi[d] = b.hi[d]; }
  set_domain(J, domlo, domhi, bclo, bchi);
  if ((rc = a4_view(c, uin, g4, 7 + c->ns, "uin", J.U))) return rc;
  if ((rc = a4_view(c, q, g4, 8 + c->ns, "q", J.q))) return rc;
  if ((rc = a4_view(c, qaux, g4, NQAUX, "qaux", J.qa))) return rc;
  for (int d = 0; d < 3; ++d) J.qe[d] = q.end[d];
  if (src_q.p) {
    if ((rc = a4_view(c, src_q, g4, 8 + c->ns, "src_q", J.srcq))) return rc;
  } else {
    J.srcq = null_view();
  }
  hipStream_t st = (hipStream_t)stream;
  const WsLayout L = ws_layout(b, c->ns, false, false, false);
  if ((rc = ws_guard(c, b, L))) return rc;
  double* ws;
  if ((rc = ctx_scratch(c, L.total * sizeof(double), st, &ws))) return rc;
  if ((rc = finish_umdrv_job(c, J, b, ws, flx, a, vol, uout))) return rc;
  set_geom(J, dx);
  J.dt_val = dt;
  const KArgs ka = umdrv_args(dx, dt, plm_iorder, use_flattening, difmag);
  Sched sc;
  int slot;
  if ((rc = ctx_job(c, J, st, &sc, &slot))) return rc;
  if (c->eos == PC_EOS_GAMMALAW)
    launch_ctu_consup<EosGamma>(J, sc, ka, ppm_type, st);
  else
    launch_ctu_consup<EosLiDryer>(J, sc, ka, ppm_type, st);
  HIPCHK(c, hipGetLastError());
  return ctx_job_done(c, slot, st);
}

}  // extern "C"

// One construct_hydro_source box (Source/Hydro.cpp:73-252): argument checks
// and the views of the caller's arrays; the workspace comes after.
struct HsbCall {
  IBox b;
  Job J;
  V vsrc;
  KArgs ka;
  WsLayout Lq, L;  // [q, qaux, srcq] part and CTU part of the workspace
  size_t qpart = 0, total = 0;  // doubles
};

static int hsb_setup(pc_ctx* c, pc_box_t bx, const int* domlo,
                     const int* domhi, const int* bclo, const int* bchi,
                     const pc_array4_t& sarr, const pc_array4_t& src,
                     const double* dx, double dt, const pc_params_t* prm,
                     const pc_array4_t* flx, const pc_array4_t* a, HsbCall& H) {
  if (!c || !prm || !box_ok(bx) || !domlo || !domhi || !bclo || !bchi || !dx ||
      !flx || !a)
    return fail(c, PC_ERR_ARG, "null argument or empty box");
  int rc;
  if ((rc = check_opts(c, prm->ppm_type, prm->plm_iorder, prm->use_hybrid_weno)))
    return rc;
  if (!prm->allow_small_energy || !prm->allow_negative_energy)
    return fail(c, PC_ERR_UNSUPPORTED,
                "allow_small_energy = allow_negative_energy = 1 required");
  H.b = mkbox(bx.lo, bx.hi);
  const IBox g4 = growb(H.b, 4);
  Job& J = H.J;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) { J.lo[d] = H.b.lo[d]; J.hi[d] = H.b.hi[d]; }
  set_domain(J, domlo, domhi, bclo, bchi);
  if ((rc = a4_view(c, sarr, g4, 7 + c->ns, "sarr", J.U))) return rc;
  H.vsrc = null_view();
  if (src.p && (rc = a4_view(c, src, g4, 7 + c->ns, "src", H.vsrc))) return rc;
  H.Lq = ws_layout(H.b, c->ns, true, src.p != nullptr, false);
  H.L = ws_layout(H.b, c->ns, false, false, false);
  H.qpart = H.Lq.off_tr[0];
  H.total = H.qpart + H.L.total;
  if ((rc = ws_guard(c, H.b, H.L))) return rc;
  set_geom(J, dx);
  J.dt_val = dt;
  H.ka = umdrv_args(dx, dt, prm->plm_iorder, prm->use_flattening, prm->difmag);
  // srctoprim reads qaux's dp/de and dp/drho, which ctoprim only writes
  // when asked for the full primitive set
  if (src.p) H.ka.full_prims = 1;
  return PC_OK;
}

// Workspace views (ws: H.total doubles) and the caller's output views.
static int hsb_views(pc_ctx* c, HsbCall& H, double* ws, const pc_array4_t* flx,
                     const pc_array4_t* a, const pc_array4_t& vol,
                     const pc_array4_t& hyd_src) {
  Job& J = H.J;
  J.q = fab_view(ws + H.Lq.off_q, H.Lq.q);
  J.qa = fab_view(ws + H.Lq.off_qa, H.Lq.q);
  J.srcq = H.vsrc.p ? fab_view(ws + H.Lq.off_srcq, H.Lq.q) : null_view();
  for (int d = 0; d < 3; ++d) J.qe[d] = H.Lq.q.hi[d] + 1;
  return finish_umdrv_job(c, J, H.b, ws + H.qpart, flx, a, vol, hyd_src);
}

// K1, K2 and pc_umdrv's launches for the job the kernels read through sc.
static int hsb_launch(pc_ctx* c, const HsbCall& H, const Sched& sc, int ppm_type,
                      hipStream_t st) {
  const int gq = grid1(ST_CTOPRIM, H.J);
  const bool src = H.vsrc.p != nullptr;
  if (c->eos == PC_EOS_GAMMALAW) {
    LAUNCH(k_ctoprim<EosGamma>, gq, st, sc, H.ka);
    if (src) LAUNCH(k_srctoprim<1>, gq, st, sc, H.ka, H.vsrc);
    launch_ctu_consup<EosGamma>(H.J, sc, H.ka, ppm_type, st);
  } else {
    LAUNCH(k_ctoprim<EosLiDryer>, gq, st, sc, H.ka);
    if (src) LAUNCH(k_srctoprim<9>, gq, st, sc, H.ka, H.vsrc);
    launch_ctu_consup<EosLiDryer>(H.J, sc, H.ka, ppm_type, st);
  }
  HIPCHK(c, hipGetLastError());
  return PC_OK;
}

// ------------------------------------------- allocation-free box driver --
// The workspace of pc_hydro_source_box_noalloc_v1: a job descriptor, then
// q/qaux/srcq and the CTU intermediates of one box.  Sized for any box of
// up to 32^3 cells with 9 species and sources (102.4 MiB); every element a
// call reads it has written earlier in the same call, so nothing needs
// clearing.  One copy per GPU, shared by every call on it.
constexpr size_t NOALLOC_JOB_DOUBLES = (sizeof(Job) + 255) / 256 * 32;
constexpr size_t NOALLOC_WS_DOUBLES = (size_t)112 << 17;  // 112 MiB of doubles
static __device__ __align__(256) double scratch[NOALLOC_JOB_DOUBLES + NOALLOC_WS_DOUBLES];

// Device address of `scratch` per device, looked up on the first call there
// (the array's name is a host-side stand-in, never a device address).
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64];

static int noalloc_scratch(pc_ctx* c, double** out) {
  int dev = 0;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(c, PC_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(c, PC_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                  hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return PC_OK;
}

// The job descriptor travels as a kernel argument into the scratch array.
__global__ void k_put_job(Job j, Job* dst) {
  const int* s = (const int*)&j;
  int* d = (int*)dst;
  for (int w = threadIdx.x; w < (int)(sizeof(Job) / sizeof(int)); w += blockDim.x)
    d[w] = s[w];
}
static_assert(sizeof(Job) % sizeof(int) == 0, "job descriptor in whole words");

extern "C" {

int pc_hydro_source_box_v1(pc_ctx_t* c, pc_box_t bx, const int* domlo,
                           const int* domhi, const int* bclo, const int* bchi,
                           pc_array4_t sarr, pc_array4_t src,
                           pc_array4_t hyd_src, const double* dx, double dt,
                           const pc_params_t* prm, const pc_array4_t* flx,
                           const pc_array4_t* a, pc_array4_t vol,
                           void* stream) {
  HsbCall H;
  int rc;
  if ((rc = hsb_setup(c, bx, domlo, domhi, bclo, bchi, sarr, src, dx, dt, prm,
                      flx, a, H)))
    return rc;
  hipStream_t st = (hipStream_t)stream;
  double* ws;
  if ((rc = ctx_scratch(c, H.total * sizeof(double), st, &ws))) return rc;
  if ((rc = hsb_views(c, H, ws, flx, a, vol, hyd_src))) return rc;
  Sched sc;
  int slot;
  if ((rc = ctx_job(c, H.J, st, &sc, &slot))) return rc;
  if ((rc = hsb_launch(c, H, sc, prm->ppm_type, st))) return rc;
  return ctx_job_done(c, slot, st);
}

int pc_hydro_source_box_noalloc_v1(pc_ctx_t* c, pc_box_t bx, const int* domlo,
                                   const int* domhi, const int* bclo,
                                   const int* bchi, pc_array4_t sarr,
                                   pc_array4_t src, pc_array4_t hyd_src,
                                   const double* dx, double dt,
                                   const pc_params_t* prm,
                                   const pc_array4_t* flx, const pc_array4_t* a,
                                   pc_array4_t vol, void* stream) {
  HsbCall H;
  int rc;
  if ((rc = hsb_setup(c, bx, domlo, domhi, bclo, bchi, sarr, src, dx, dt, prm,
                      flx, a, H)))
    return rc;
  if (H.total > NOALLOC_WS_DOUBLES)
    return fail(c, PC_ERR_UNSUPPORTED,
                "box needs %zu bytes of workspace, more than the %zu of the "
                "allocation-free scratch (boxes up to 32^3 cells fit)",
                H.total * sizeof(double), NOALLOC_WS_DOUBLES * sizeof(double));
  double* base;
  if ((rc = noalloc_scratch(c, &base))) return rc;
  Job* djob = (Job*)base;
  if ((rc = hsb_views(c, H, base + NOALLOC_JOB_DOUBLES, flx, a, vol, hyd_src)))
    return rc;
  H.J.dtp = &djob->dt_val;  // the call's dt travels in the job descriptor
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, st, H.J, djob);
  HIPCHK(c, hipGetLastError());
  return hsb_launch(c, H, one_job_sched(djob), prm->ppm_type, st);
}

}  // extern "C"

// primitives, MOL fluxes and the divergence of one box-driver job
template <class EOS>
static void launch_mol_src(const Job& J, const Sched& sc, const KArgs& ka,
                           bool laxf, hipStream_t st) {
  LAUNCH((k_ctoprim<EOS, double, false, ST_MOL_PRIM>), grid1(ST_MOL_PRIM, J), st,
         sc, ka);
  if (laxf) {
    LAUNCH((k_mol_flux<EOS, 0, true>), grid1(ST_MOL_X, J), st, sc, ka);
    LAUNCH((k_mol_flux<EOS, 1, true>), grid1(ST_MOL_Y, J), st, sc, ka);
    LAUNCH((k_mol_flux<EOS, 2, true>), grid1(ST_MOL_Z, J), st, sc, ka);
  } else {
    LAUNCH((k_mol_flux<EOS, 0, false>), grid1(ST_MOL_X, J), st, sc, ka);
    LAUNCH((k_mol_flux<EOS, 1, false>), grid1(ST_MOL_Y, J), st, sc, ka);
    LAUNCH((k_mol_flux<EOS, 2, false>), grid1(ST_MOL_Z, J), st, sc, ka);
  }
  LAUNCH(k_mol_div<EOS::NS>, grid1(ST_MOL_UPD, J), st, sc, ka);
}

extern "C" {

int pc_mol_src_box_v1(pc_ctx_t* c, pc_box_t bx, pc_array4_t sarr,
                      pc_array4_t molsrc, const pc_params_t* prm,
                      const pc_array4_t* flx, const pc_array4_t* a,
                      pc_array4_t vol, void* stream) {
  if (!c || !prm || !box_ok(bx) || !a)
    return fail(c, PC_ERR_ARG, "null argument or empty box");
  int rc;
  if ((rc = check_mol_opts(c, prm))) return rc;
  const IBox b = mkbox(bx.lo, bx.hi);
  const IBox g4 = growb(b, 4);
  const int NVAR = 7 + c->ns;
  Job J;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) { J.lo[d] = b.lo[d]; J.hi[d] = b.hi[d]; }
  if ((rc = a4_view(c, sarr, g4, NVAR, "sarr", J.U))) return rc;
  if ((rc = a4_view(c, molsrc, b, NVAR, "molsrc", J.upd))) return rc;
  if ((rc = a4_view(c, vol, b, 1, "vol", J.vol))) return rc;
  for (int d = 0; d < 3; ++d) {
    IBox fb = b;
    fb.hi[d] += 1;
    if ((rc = a4_view(c, a[d], fb, 1, "a", J.area[d]))) return rc;
    if (flx && (rc = a4_view(c, flx[d], fb, NVAR, "flx", J.flx[d]))) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  const WsLayout W = ws_layout(b, c->ns, true, false, flx == nullptr, true);
  if ((rc = ws_guard(c, b, W))) return rc;
  double* ws;
  if ((rc = ctx_scratch(c, W.total * sizeof(double), st, &ws))) return rc;
  J.q = fab_view(ws + W.off_q, W.q);
  J.qa = fab_view(ws + W.off_qa, W.q);
  for (int d = 0; d < 3; ++d) {
    J.qe[d] = W.q.hi[d] + 1;
    if (!flx) J.flx[d] = fab_view(ws + W.off_fl[d], W.fl[d]);
  }
  KArgs ka;
  memset(&ka, 0, sizeof(ka));
  ka.plm_iorder = prm->plm_iorder;
  Sched sc;
  int slot;
  if ((rc = ctx_job(c, J, st, &sc, &slot))) return rc;
  if (c->eos == PC_EOS_GAMMALAW)
    launch_mol_src<EosGamma>(J, sc, ka, prm->use_laxf_flux != 0, st);
  else
    launch_mol_src<EosLiDryer>(J, sc, ka, prm->use_laxf_flux != 0, st);
  HIPCHK(c, hipGetLastError());
  return ctx_job_done(c, slot, st);
}

}  // extern "C"

// ==================================================================== level
// Three streams for the three directions of a step, forked from and joined
// back to the stream the step is issued on.  Event ev[k] marks the end of a
// launch (or of the fork point on the step's stream); a launch on one branch
// waits only for the launches of other branches whose output it reads.
struct Branches {
  static constexpr int NEV = 16;
  bool on = false;
  bool made = false;
  hipStream_t br[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev[NEV];
  hipError_t make() {
    if (made) return hipSuccess;
    hipError_t e = hipSuccess;
    int nb = 0, ne = 0;  // streams and events created so far
    while (nb < 3 &&
           (e = hipStreamCreateWithFlags(&br[nb], hipStreamNonBlocking)) == hipSuccess)
      ++nb;
    while (e == hipSuccess && ne < NEV &&
           (e = hipEventCreateWithFlags(&ev[ne], hipEventDisableTiming)) == hipSuccess)
      ++ne;
    if (e != hipSuccess) {  // undo the part that was created
      for (int b = 0; b < nb; ++b) (void)hipStreamDestroy(br[b]);
      for (int x = 0; x < ne; ++x) (void)hipEventDestroy(ev[x]);
      return e;
    }
    made = true;
    return hipSuccess;
  }
  void release() {
    if (!made) return;
    for (int b = 0; b < 3; ++b) (void)hipStreamDestroy(br[b]);
    for (int x = 0; x < NEV; ++x) (void)hipEventDestroy(ev[x]);
    made = false;
  }
  // event slots: 0 fork point, 1-3 first fluxes of d, 4-12 transverse flux
  // (d|t) at 4 + 3d + t, 13-15 last launch of branch d
  static int F1(int d) { return 1 + d; }
  static int TR(int d, int t) { return 4 + 3 * d + t; }
  static int END(int d) { return 13 + d; }
  void fork(hipStream_t st) {  // the branches wait for the step's stream
    (void)hipEventRecord(ev[0], st);
    for (int b = 0; b < 3; ++b) (void)hipStreamWaitEvent(br[b], ev[0], 0);
  }
  void done(int k, int b) { (void)hipEventRecord(ev[k], br[b]); }
  void wait(int b, int k) { (void)hipStreamWaitEvent(br[b], ev[k], 0); }
  void join(hipStream_t st) {  // the step's stream waits for every branch
    for (int b = 0; b < 3; ++b) (void)hipStreamWaitEvent(st, ev[END(b)], 0);
  }
};

// Stages whose workgroups walk their logical blocks backwards (Sched::rev):
// such a launch starts on the boxes its producer wrote last, whose arrays
// are still in the last-level cache.  The divergence after the primitives,
// the transverse fluxes after the traces, the update after the final fluxes
// (the next primitives then run forwards after it), and the MOL primitives
// after the MOL update.  On a level of 64 boxes of 32^3 this takes ~1% off
// the Godunov step and 2-4% off the MOL step; reversing every stage is
// slower than reversing none.
constexpr unsigned kReversedStages =
    (1u << ST_DIVU) | (1u << ST_TRANS_XY) | (1u << ST_TRANS_YX) |
    (1u << ST_TRANS_ZX) | (1u << ST_UPDATE) | (1u << ST_MOL_PRIM);

struct pc_level {
  unsigned rev_mask = kReversedStages;
  pc_ctx* ctx = nullptr;
  pc_params_t prm;
  std::vector<IBox> boxes;
  std::vector<long long> off;  // state arena offsets (doubles)
  long long state_total = 0;
  double* state = nullptr;     // device arena
  double* ws = nullptr;        // device stage workspace
  size_t ws_total = 0;
  Job* jobs = nullptr;         // device job table
  int* cellmap = nullptr;      // device domain cell -> box
  StagePlan plans[ST_COUNT];
  Sched scs[ST_COUNT];
  DtState* ds = nullptr;       // device
  LevelArgs la;
  KArgs ka;
  int lowo[3] = {0, 0, 0};  // a Hard/UserBC face exists in direction d
  int timing = 0;
  hipEvent_t ev[PC_NKERN + 1];
  bool ev_made = false;
  double ktime[PC_NKERN];
  int nsteps_timed = 0;
  std::vector<Job> hjobs;      // host copy of the job table (plans)
  int mixed = 0;               // 1: FP32 reconstruction/Riemann/intermediates
  bool mol = false;            // method-of-lines step (do_mol)
  int trace_cells = 0;         // largest LDS trace window of the boxes
  bool fuse = false;           // PPM trace and first fluxes in one kernel
  int mol_cells = 0;           // largest MOL flux window of the boxes
  bool mol_lds = false;        // MOL fluxes from face pencils in LDS
  bool gbranched = false;      // the captured graph carries the branch DAG
  int gnodes = 0;              // kernel nodes of the captured graph
  bool laxf = false;           // MOL with the Lax-Friedrichs flux
  double tuned_ms[ST_COUNT] = {};
  // hipGraph of gsteps consecutive steps, replayed by pc_level_step_v1
  int gsteps = 0;
  bool gvalid = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t cap = nullptr;
  Branches brs;  // per-direction streams (pc_level_set_branches_v1)
};

static void plan_free(StagePlan& P) {
  if (P.blkmap) (void)hipFree(P.blkmap);
  if (P.blkstart) (void)hipFree(P.blkstart);
  P.blkmap = P.blkstart = nullptr;
}

// Launch plan of stage s over the level's boxes with block shape (lx,ly,lz)
// and at most `cap` workgroups (0: one per logical block).
static bool plan_make(pc_level* L, int s, int lx, int ly, int lz, int cap,
                      StagePlan& P) {
  // the PPM trace always runs on its LDS tiles (only persistence varies)
  if (s == ST_TRACE && L->prm.ppm_type == 1 && !L->mol) lx = L->fuse ? -3 : -2;
  // so do the MOL fluxes on their face tiles (mixed: the x fluxes only)
  if (s >= ST_MOL_X && s <= ST_MOL_Z)
    lx = (L->mol_lds && (!L->mixed || s == ST_MOL_X)) ? -4 : (lx == -4 ? -1 : lx);
  const int nb = (int)L->hjobs.size();
  std::vector<int> start(nb), map;
  for (int b = 0; b < nb; ++b) {
    start[b] = (int)map.size();
    const long long n = stage_blocks(s, L->hjobs[b], lx, ly, lz, L->mixed != 0);
    for (long long k = 0; k < n; ++k) map.push_back(b);
  }
  P.nblocks = (int)map.size();
  P.grid = (cap > 0 && cap < P.nblocks) ? cap : P.nblocks;
  P.lx = lx;
  P.ly = ly;
  P.lz = lz;
  P.blkmap = P.blkstart = nullptr;
  if (hipMalloc(&P.blkmap, sizeof(int) * std::max<size_t>(map.size(), 1)) !=
          hipSuccess ||
      hipMalloc(&P.blkstart, sizeof(int) * start.size()) != hipSuccess) {
    (void)hipGetLastError();
    plan_free(P);
    return false;
  }
  if (hipMemcpy(P.blkmap, map.data(), sizeof(int) * map.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(P.blkstart, start.data(), sizeof(int) * start.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    plan_free(P);
    return false;
  }
  return true;
}

static Sched plan_sched(const pc_level* L, const StagePlan& P, int stage) {
  Sched s{L->jobs, P.blkmap, P.blkstart, P.nblocks, P.lx, P.ly, P.lz,
          (int)L->hjobs.size()};
  s.rev = P.grid >= P.nblocks && ((L->rev_mask >> stage) & 1u);
  return s;
}

static void graph_drop(pc_level* L) {
  if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
  if (L->graph) (void)hipGraphDestroy(L->graph);
  L->gexec = nullptr;
  L->graph = nullptr;
  L->gvalid = false;
}

static void level_free(pc_level* L) {
  graph_drop(L);
  if (L->cap) (void)hipStreamDestroy(L->cap);
  if (L->state) (void)hipFree(L->state);
  if (L->ws) (void)hipFree(L->ws);
  if (L->jobs) (void)hipFree(L->jobs);
  if (L->cellmap) (void)hipFree(L->cellmap);
  if (L->ds) (void)hipFree(L->ds);
  for (int s = 0; s < ST_COUNT; ++s) plan_free(L->plans[s]);
  if (L->ev_made)
    for (int e = 0; e <= PC_NKERN; ++e) (void)hipEventDestroy(L->ev[e]);
  L->brs.release();
  delete L;
}

// The MOL fluxes of direction D: from face tiles staged in LDS when every
// box's x-face rows fit one workgroup (in the mixed step only the x fluxes:
// staged, the FP32 y and z kernels ran faster alone but the step with its
// three directions on branch streams slower), else one thread per face from
// the level arrays.
template <class EOS, int D, class FP, bool P, class H>
static void launch_mol_flux(H* L, int grid, hipStream_t st, const Sched& sc) {
  const KArgs& ka = L->ka;
  if (L->mol_lds && (sizeof(FP) == 8 || D == 0)) {
    const size_t sh = (size_t)L->mol_cells * (6 + EOS::NS) * sizeof(FP);
    if (L->laxf) LAUNCH_SH((k_mol_flux_lds<EOS, D, true, FP, P>), grid, sh, st, sc, ka);
    else LAUNCH_SH((k_mol_flux_lds<EOS, D, false, FP, P>), grid, sh, st, sc, ka);
  } else {
    if (L->laxf) LAUNCH((k_mol_flux<EOS, D, true, FP, P, false>), grid, st, sc, ka);
    else LAUNCH((k_mol_flux<EOS, D, false, FP, P, false>), grid, st, sc, ka);
  }
}

// Launch every kernel of stage s once with schedule sc on `grid` workgroups
// (sub >= 0: only that direction of the trace stage).  P: the persistent
// instantiations, which loop over the logical blocks.
template <class EOS, class FP, bool P, class H>
static void launch_stage_t(H* L, int s, const Sched& sc, int grid,
                           hipStream_t st, int sub) {
  const KArgs& ka = L->ka;
  auto tr = [&](int d) { return sub < 0 || sub == d; };
  switch (s) {
    case ST_GHOST: LAUNCH((k_ghost<EOS::NS, P>), grid, st, sc, ka, L->la); break;
    // the Godunov step's primitives read ghost values from their source
    // cells (no ghost fill before them)
    case ST_CTOPRIM: LAUNCH((k_ctoprim_ghost<EOS, FP, P>), grid, st, sc, ka, L->la); break;
    case ST_DIVU: LAUNCH((k_divu<FP, P>), grid, st, sc, ka); break;
    case ST_TRACE:
      if (L->prm.ppm_type == 1 && L->fuse) {
        const size_t sh = pencil_lds_bytes<EOS::NS, FP>(L->trace_cells);
        if (tr(0)) LAUNCH_SH((k_trace_flux1_lds<EOS, 0, FP, P>), grid, sh, st, sc, ka);
        if (tr(1)) LAUNCH_SH((k_trace_flux1_lds<EOS, 1, FP, P>), grid, sh, st, sc, ka);
        if (tr(2)) LAUNCH_SH((k_trace_flux1_lds<EOS, 2, FP, P>), grid, sh, st, sc, ka);
      } else if (L->prm.ppm_type == 1) {
        const size_t sh = (size_t)L->trace_cells * (5 + EOS::NS) * sizeof(FP);
        if (tr(0)) LAUNCH_SH((k_trace_ppm_lds<EOS, 0, FP, P>), grid, sh, st, sc, ka);
        if (tr(1)) LAUNCH_SH((k_trace_ppm_lds<EOS, 1, FP, P>), grid, sh, st, sc, ka);
        if (tr(2)) LAUNCH_SH((k_trace_ppm_lds<EOS, 2, FP, P>), grid, sh, st, sc, ka);
      } else {
        if (tr(0)) LAUNCH((k_trace_plm<EOS, 0, FP, P>), grid, st, sc, ka);
        if (tr(1)) LAUNCH((k_trace_plm<EOS, 1, FP, P>), grid, st, sc, ka);
        if (tr(2)) LAUNCH((k_trace_plm<EOS, 2, FP, P>), grid, st, sc, ka);
      }
      break;
    // with the fused PPM trace the first fluxes come from the trace stage
    case ST_FLUX1_X:
      if (!L->fuse) {
        LAUNCH((k_flux1<EOS, 0, FP, P>), grid, st, sc, ka);
        LAUNCH((k_tcorr<0, EOS::NS, FP, P>), grid, st, sc, ka);
      }
      break;
    case ST_FLUX1_Y:
      if (!L->fuse) {
        LAUNCH((k_flux1<EOS, 1, FP, P>), grid, st, sc, ka);
        LAUNCH((k_tcorr<1, EOS::NS, FP, P>), grid, st, sc, ka);
      }
      break;
    case ST_FLUX1_Z:
      if (!L->fuse) {
        LAUNCH((k_flux1<EOS, 2, FP, P>), grid, st, sc, ka);
        LAUNCH((k_tcorr<2, EOS::NS, FP, P>), grid, st, sc, ka);
      }
      break;
    // both transverse fluxes of a direction under its first stage; the
    // second stage's region is empty
    case ST_TRANS_XY: LAUNCH((k_trans<EOS, 0, FP, P>), grid, st, sc, ka); break;
    case ST_TRANS_YX: LAUNCH((k_trans<EOS, 1, FP, P>), grid, st, sc, ka); break;
    case ST_TRANS_ZX: LAUNCH((k_trans<EOS, 2, FP, P>), grid, st, sc, ka); break;
    case ST_TRANS_XZ: case ST_TRANS_YZ: case ST_TRANS_ZY: break;
    case ST_FINAL_X:
      LAUNCH((k_final<EOS, 0, false, FP, P, false>), grid, st, sc, ka);
      if (L->lowo[0]) LAUNCH((k_final_loworder<EOS, 0, FP, P, false>), grid, st, sc, ka);
      break;
    case ST_FINAL_Y:
      LAUNCH((k_final<EOS, 1, false, FP, P, false>), grid, st, sc, ka);
      if (L->lowo[1]) LAUNCH((k_final_loworder<EOS, 1, FP, P, false>), grid, st, sc, ka);
      break;
    case ST_FINAL_Z:
      LAUNCH((k_final<EOS, 2, false, FP, P, false>), grid, st, sc, ka);
      if (L->lowo[2]) LAUNCH((k_final_loworder<EOS, 2, FP, P, false>), grid, st, sc, ka);
      break;
    case ST_UPDATE:
      LAUNCH((k_update<EOS, P, sizeof(FP) == 4>), grid, st, sc, ka, L->ds);
      break;
    case ST_MOL_PRIM:  // sub: the Heun stage (1: of U* in the stage arrays)
      if (sub == 1)
        LAUNCH((k_ctoprim_ghost<EOS, FP, P, ST_MOL_PRIM, true>), grid, st, sc, ka, L->la);
      else
        LAUNCH((k_ctoprim_ghost<EOS, FP, P, ST_MOL_PRIM>), grid, st, sc, ka, L->la);
      break;
    case ST_MOL_X: launch_mol_flux<EOS, 0, FP, P>(L, grid, st, sc); break;
    case ST_MOL_Y: launch_mol_flux<EOS, 1, FP, P>(L, grid, st, sc); break;
    case ST_MOL_Z: launch_mol_flux<EOS, 2, FP, P>(L, grid, st, sc); break;
    case ST_MOL_UPD:  // sub: the Heun stage (0 or 1)
      if (sub == 0) LAUNCH((k_mol_update<EOS, 0, P, sizeof(FP) == 4>), grid, st, sc, ka, L->ds);
      else LAUNCH((k_mol_update<EOS, 1, P, sizeof(FP) == 4>), grid, st, sc, ka, L->ds);
      break;
    default: break;
  }
}

// H: the level or the batch (ka, la, lowo, ds, prm.ppm_type, mixed)
template <class EOS, class H>
static void launch_stage(H* L, int s, const Sched& sc,
                         const StagePlan& pl, hipStream_t st, int sub = -1) {
  const bool persist = pl.grid < pl.nblocks;
  if (L->mixed) {
    if (persist)
      launch_stage_t<EOS, float, true>(L, s, sc, pl.grid, st, sub);
    else
      launch_stage_t<EOS, float, false>(L, s, sc, pl.grid, st, sub);
  } else {
    if (persist)
      launch_stage_t<EOS, double, true>(L, s, sc, pl.grid, st, sub);
    else
      launch_stage_t<EOS, double, false>(L, s, sc, pl.grid, st, sub);
  }
}

extern "C" {

int pc_level_create_v1(pc_ctx_t* c, const pc_level_desc_t* d,
                       const pc_params_t* prm, pc_level_t** out) {
  if (!c || !d || !prm || !out || d->nboxes <= 0 || !d->boxes)
    return fail(c, PC_ERR_ARG, "bad level description");
  int rc;
  if (prm->do_mol) {
    if ((rc = check_mol_opts(c, prm))) return rc;
  } else if ((rc = check_opts(c, prm->ppm_type, prm->plm_iorder,
                              prm->use_hybrid_weno))) {
    return rc;
  }
  if (!prm->allow_small_energy || !prm->allow_negative_energy)
    return fail(c, PC_ERR_UNSUPPORTED,
                "allow_small_energy = allow_negative_energy = 1 required");
  *out = nullptr;
  pc_level* L = new pc_level();
  L->mol = prm->do_mol != 0;
  L->laxf = prm->use_laxf_flux != 0;
  L->ctx = c;
  L->prm = *prm;
  const int ns = c->ns, NVAR = 7 + ns;
  long long o = 0;
  size_t wso = 0;
  std::vector<size_t> wsoff;
  for (int b = 0; b < d->nboxes; ++b) {
    const IBox bx = mkbox(d->boxes[b].lo, d->boxes[b].hi);
    if (bx.hi[0] < bx.lo[0] || bx.hi[1] < bx.lo[1] || bx.hi[2] < bx.lo[2]) {
      level_free(L);
      return fail(c, PC_ERR_ARG, "empty box %d", b);
    }
    if (!d->periodic_self)
      for (int dd = 0; dd < 3; ++dd)
        if (bx.lo[dd] < d->domlo[dd] || bx.hi[dd] > d->domhi[dd]) {
          level_free(L);
          return fail(c, PC_ERR_ARG, "box %d outside the domain", b);
        }
    const WsLayout W = ws_layout(bx, ns, true, false, true, L->mol);
    const long long big = box_max_array_bytes(bx, ns, W);
    if (big > MAX_FAB_BYTES) {
      level_free(L);
      return fail(c, PC_ERR_UNSUPPORTED,
                  "box %d: an array of %lld bytes exceeds the %lld the kernels "
                  "address with 32-bit offsets", b, big, MAX_FAB_BYTES);
    }
    L->boxes.push_back(bx);
    L->off.push_back(o);
    o += growb(bx, 4).npts() * NVAR;
    wsoff.push_back(wso);
    wso += W.total;
  }
  L->state_total = o;
  L->ws_total = wso;
  auto dalloc = [&](void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return true;
  };
  if (!dalloc((void**)&L->state, o * sizeof(double)) ||
      !dalloc((void**)&L->ws, wso * sizeof(double)) ||
      !dalloc((void**)&L->ds, sizeof(DtState))) {
    level_free(L);
    return fail(c, PC_ERR_NOMEM, "level allocation failed");
  }
  // the Godunov step stores only the first ghost layer: the outer layers
  // start at zero rather than as whatever the allocation held
  if (hipMemset(L->state, 0, o * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    level_free(L);
    return fail(c, PC_ERR_HIP, "level state initialisation failed");
  }
  // jobs
  std::vector<Job> jobs(d->nboxes);
  for (int b = 0; b < d->nboxes; ++b) {
    Job& J = jobs[b];
    memset(&J, 0, sizeof(J));
    const IBox& bx = L->boxes[b];
    for (int dd = 0; dd < 3; ++dd) {
      J.lo[dd] = bx.lo[dd];
      J.hi[dd] = bx.hi[dd];
      J.dlo[dd] = d->periodic_self ? bx.lo[dd] : d->domlo[dd];
      J.dhi[dd] = d->periodic_self ? bx.hi[dd] : d->domhi[dd];
      J.bclo[dd] = d->periodic_self ? PC_BC_INTERIOR : d->bclo[dd];
      J.bchi[dd] = d->periodic_self ? PC_BC_INTERIOR : d->bchi[dd];
      J.qe[dd] = bx.hi[dd] + 5;
    }
    set_geom(J, d->dx);
    J.dtp = &L->ds->dt;
    J.U = fab_view(L->state + L->off[b], growb(bx, 4));
    const WsLayout W = ws_layout(bx, ns, true, false, true, L->mol);
    ws_fill_job(J, L->ws + wsoff[b], W, true, false, true);
    J.srcq = null_view();
    for (int dd = 0; dd < 3; ++dd) J.area[dd] = null_view();
    J.vol = null_view();
    J.upd = null_view();
  }
  if (!dalloc((void**)&L->jobs, sizeof(Job) * jobs.size())) {
    level_free(L);
    return fail(c, PC_ERR_NOMEM, "level allocation failed");
  }
  if (hipMemcpy(L->jobs, jobs.data(), sizeof(Job) * jobs.size(),
                hipMemcpyHostToDevice) != hipSuccess) {
    level_free(L);
    return fail(c, PC_ERR_HIP, "job upload failed");
  }
  // per-stage launch plans: runs of BLOCK cells, one workgroup each
  L->hjobs = jobs;
  L->fuse = !L->mol && prm->ppm_type == 1 && pencils_fit(jobs.data(), (int)jobs.size());
  L->trace_cells = trace_cells_of(jobs.data(), (int)jobs.size(), L->fuse);
  L->mol_lds = L->mol && mol_fit(jobs.data(), (int)jobs.size());
  L->mol_cells = L->mol_lds ? mol_cells_of(jobs.data(), (int)jobs.size()) : 0;
  for (int s = 0; s < ST_COUNT; ++s) {
    if (!plan_make(L, s, -1, 0, 0, 0, L->plans[s])) {
      level_free(L);
      return fail(c, PC_ERR_NOMEM, "plan allocation failed");
    }
    L->scs[s] = plan_sched(L, L->plans[s], s);
  }
  // ghost-fill cell map over the domain
  memset(&L->la, 0, sizeof(L->la));
  L->la.periodic_self = d->periodic_self;
  for (int dd = 0; dd < 3; ++dd) {
    L->la.dlo[dd] = d->domlo[dd];
    L->la.dhi[dd] = d->domhi[dd];
    L->la.bclo[dd] = d->bclo[dd];
    L->la.bchi[dd] = d->bchi[dd];
  }
  if (!d->periodic_self) {
    const long long nx = d->domhi[0] - d->domlo[0] + 1,
                    ny = d->domhi[1] - d->domlo[1] + 1,
                    nz = d->domhi[2] - d->domlo[2] + 1;
    std::vector<int> cm((size_t)(nx * ny * nz), -1);
    for (int b = 0; b < d->nboxes; ++b) {
      const IBox& bx = L->boxes[b];
      for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
        for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
          for (int i = bx.lo[0]; i <= bx.hi[0]; ++i)
            cm[(i - d->domlo[0]) + nx * ((j - d->domlo[1]) + ny * (long long)(k - d->domlo[2]))] = b;
    }
    for (size_t x = 0; x < cm.size(); ++x)
      if (cm[x] < 0) {
        level_free(L);
        return fail(c, PC_ERR_UNSUPPORTED,
                    "the boxes must tile the domain (or use periodic_self)");
      }
    if (!dalloc((void**)&L->cellmap, sizeof(int) * cm.size())) {
      level_free(L);
      return fail(c, PC_ERR_NOMEM, "cellmap allocation failed");
    }
    (void)hipMemcpy(L->cellmap, cm.data(), sizeof(int) * cm.size(),
                    hipMemcpyHostToDevice);
  }
  L->la.cellmap = L->cellmap;
  for (int dd = 0; dd < 3; ++dd) {
    auto is = [](int b) { return b == PC_BC_INFLOW || b == PC_BC_USERBC; };
    L->lowo[dd] = !d->periodic_self && (is(d->bclo[dd]) || is(d->bchi[dd]));
  }
  L->la.cfl = prm->cfl;
  L->la.max_dt = prm->max_dt;
  L->la.fixed_dt = prm->fixed_dt;
  L->la.change_max = prm->change_max;
  L->la.init_shrink = prm->init_shrink;
  L->la.dt_cutoff = prm->dt_cutoff;
  L->la.stop_time = prm->stop_time;
  memset(&L->ka, 0, sizeof(L->ka));
  L->ka.difmag = prm->difmag;
  L->ka.plm_iorder = prm->plm_iorder;
  L->ka.use_flattening = prm->use_flattening;
  // q's rho e feeds only the traced states' rho e, which no flux reads
  // (store_trace)
  L->ka.skip_reint = 1;
  DtState z;
  z.dt = 0.0;
  z.minbits = 0x7FEFFFFFFFFFFFFFull;
  z.bad = 0;
  z.small = 0;
  z.time = 0.0;
  (void)hipMemcpy(L->ds, &z, sizeof(z), hipMemcpyHostToDevice);
  if (hipDeviceSynchronize() != hipSuccess) {
    level_free(L);
    return fail(c, PC_ERR_HIP, "level setup failed");
  }
  *out = L;
  return PC_OK;
}

int pc_level_destroy_v1(pc_level_t* L) {
  if (!L) return PC_ERR_ARG;
  (void)hipDeviceSynchronize();
  level_free(L);
  return PC_OK;
}

long long pc_level_box_offset_v1(const pc_level_t* L, int b) {
  if (!L || b < 0 || b >= (int)L->boxes.size()) return -1;
  return L->off[b];
}
long long pc_level_box_size_v1(const pc_level_t* L, int b) {
  if (!L || b < 0 || b >= (int)L->boxes.size()) return -1;
  return growb(L->boxes[b], 4).npts() * (7 + L->ctx->ns);
}
double* pc_level_state_ptr_v1(pc_level_t* L) { return L ? L->state : nullptr; }
long long pc_level_state_size_v1(const pc_level_t* L) {
  return L ? L->state_total : -1;
}

int pc_level_write_box_v1(pc_level_t* L, int b, const double* buf) {
  if (!L || !buf || b < 0 || b >= (int)L->boxes.size()) return PC_ERR_ARG;
  const size_t n = (size_t)growb(L->boxes[b], 4).npts() * (7 + L->ctx->ns);
  HIPCHK(L->ctx, hipDeviceSynchronize());
  HIPCHK(L->ctx, hipMemcpy(L->state + L->off[b], buf, n * sizeof(double),
                           hipMemcpyDefault));
  return PC_OK;
}

int pc_level_read_box_v1(pc_level_t* L, int b, double* buf) {
  if (!L || !buf || b < 0 || b >= (int)L->boxes.size()) return PC_ERR_ARG;
  const size_t n = (size_t)growb(L->boxes[b], 4).npts() * (7 + L->ctx->ns);
  HIPCHK(L->ctx, hipDeviceSynchronize());
  HIPCHK(L->ctx, hipMemcpy(buf, L->state + L->off[b], n * sizeof(double),
                           hipMemcpyDefault));
  return PC_OK;
}

int pc_level_set_dt_v1(pc_level_t* L, double dt) {
  if (!L) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipMemcpy(&L->ds->dt, &dt, sizeof(double), hipMemcpyHostToDevice));
  return PC_OK;
}
int pc_level_get_dt_v1(pc_level_t* L, double* dt) {
  if (!L || !dt) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipDeviceSynchronize());
  HIPCHK(L->ctx, hipMemcpy(dt, &L->ds->dt, sizeof(double), hipMemcpyDeviceToHost));
  return PC_OK;
}
int pc_level_set_time_v1(pc_level_t* L, double time) {
  if (!L) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipDeviceSynchronize());
  HIPCHK(L->ctx, hipMemcpy(&L->ds->time, &time, sizeof(double), hipMemcpyHostToDevice));
  return PC_OK;
}
int pc_level_get_time_v1(pc_level_t* L, double* time) {
  if (!L || !time) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipDeviceSynchronize());
  HIPCHK(L->ctx, hipMemcpy(time, &L->ds->time, sizeof(double), hipMemcpyDeviceToHost));
  return PC_OK;
}

static int level_fill(pc_level* L, hipStream_t st) {
  if (L->ctx->ns == 1)
    launch_stage<EosGamma>(L, ST_GHOST, L->scs[ST_GHOST], L->plans[ST_GHOST], st);
  else
    launch_stage<EosLiDryer>(L, ST_GHOST, L->scs[ST_GHOST], L->plans[ST_GHOST], st);
  HIPCHK(L->ctx, hipGetLastError());
  return PC_OK;
}

int pc_level_fill_ghosts_v1(pc_level_t* L, void* stream) {
  if (!L) return PC_ERR_ARG;
  return level_fill(L, (hipStream_t)stream);
}

}  // extern "C"

template <class EOS>
static void level_estdt_launch(pc_level* L, hipStream_t st) {
  const StagePlan& P = L->plans[ST_UPDATE];
  if (P.grid < P.nblocks)
    LAUNCH((k_estdt<EOS, true>), P.grid, st, L->scs[ST_UPDATE], L->ka, L->ds);
  else
    LAUNCH((k_estdt<EOS, false>), P.grid, st, L->scs[ST_UPDATE], L->ka, L->ds);
}

extern "C" {

int pc_level_init_dt_v1(pc_level_t* L, void* stream) {
  if (!L) return PC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (L->ctx->ns == 1) level_estdt_launch<EosGamma>(L, st);
  else level_estdt_launch<EosLiDryer>(L, st);
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, L->ds, L->la, DT_INIT);
  HIPCHK(L->ctx, hipGetLastError());
  return PC_OK;
}

int pc_level_estdt_v1(pc_level_t* L, double* dt_out, void* stream) {
  if (!L || !dt_out) return PC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  DtState save;
  HIPCHK(L->ctx, hipStreamSynchronize(st));
  HIPCHK(L->ctx, hipMemcpy(&save, L->ds, sizeof(save), hipMemcpyDeviceToHost));
  if (L->ctx->ns == 1) level_estdt_launch<EosGamma>(L, st);
  else level_estdt_launch<EosLiDryer>(L, st);
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, L->ds, L->la, DT_EST);
  HIPCHK(L->ctx, hipStreamSynchronize(st));
  DtState now;
  HIPCHK(L->ctx, hipMemcpy(&now, L->ds, sizeof(now), hipMemcpyDeviceToHost));
  *dt_out = now.dt;
  HIPCHK(L->ctx, hipMemcpy(L->ds, &save, sizeof(save), hipMemcpyHostToDevice));
  return PC_OK;
}

int pc_level_enable_timing_v1(pc_level_t* L, int on) {
  if (!L) return PC_ERR_ARG;
  if (on && !L->ev_made) {
    for (int e = 0; e <= PC_NKERN; ++e) HIPCHK(L->ctx, hipEventCreate(&L->ev[e]));
    L->ev_made = true;
  }
  L->timing = on;
  for (int s = 0; s < PC_NKERN; ++s) L->ktime[s] = 0.0;
  L->nsteps_timed = 0;
  return PC_OK;
}

}  // extern "C"

// One step of the whole level, each launch covering every box: the
// Godunov step (14 to 23 launches) or the MOL step (two Heun stages of ghost fill,
// primitives, three flux launches and an update; 13 launches).  With timing
// on, an event follows every launch and the step ends with a host wait, so
// per-kernel device times accumulate in ktime[] under the kernel's id.
// The directional stages of a CTU step over the three branch streams:
// branch d runs trace_d and flux1_d, then the two transverse fluxes (d|t)
// in one launch (after the first fluxes of both t, computed on branches t),
// then final_d (after the transverse fluxes (t0|t1) and (t1|t0) of the other
// two directions, which the other two branches compute).  Every read of another branch's output
// waits for the event recorded after the launch that writes it.
template <class EOS, class H>
static void branch_ctu_fluxes(H* L, const Sched* S, const StagePlan* P,
                              hipStream_t st, Branches& B) {
  B.fork(st);
  for (int d = 0; d < 3; ++d) {
    launch_stage<EOS>(L, ST_TRACE, S[ST_TRACE], P[ST_TRACE], B.br[d], d);
    launch_stage<EOS>(L, ST_FLUX1_X + d, S[ST_FLUX1_X + d], P[ST_FLUX1_X + d], B.br[d]);
    B.done(Branches::F1(d), d);
  }
  for (int d = 0; d < 3; ++d) {  // (d|t0) and (d|t1) in one launch
    const int x = ST_TRANS_XY + 2 * d;
    for (int t = 0; t < 3; ++t)
      if (t != d) B.wait(d, Branches::F1(t));
    launch_stage<EOS>(L, x, S[x], P[x], B.br[d]);
    for (int t = 0; t < 3; ++t)
      if (t != d) B.done(Branches::TR(d, t), d);
  }
  for (int d = 0; d < 3; ++d) {
    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    B.wait(d, Branches::TR(t0, t1));
    B.wait(d, Branches::TR(t1, t0));
    launch_stage<EOS>(L, ST_FINAL_X + d, S[ST_FINAL_X + d], P[ST_FINAL_X + d], B.br[d]);
    B.done(Branches::END(d), d);
  }
  B.join(st);
}

// The three MOL flux directions of one Heun stage over the branch streams.
template <class EOS, class H>
static void branch_mol_fluxes(H* L, const Sched* S, const StagePlan* P,
                              hipStream_t st, Branches& B) {
  B.fork(st);
  for (int d = 0; d < 3; ++d) {
    launch_stage<EOS>(L, ST_MOL_X + d, S[ST_MOL_X + d], P[ST_MOL_X + d], B.br[d]);
    B.done(Branches::END(d), d);
  }
  B.join(st);
}

template <class EOS>
static int level_step_eos(pc_level* L, int nsteps, hipStream_t st,
                          bool capturing = false) {
  const Sched* S = L->scs;
  const StagePlan* P = L->plans;
  for (int s = 0; s < nsteps; ++s) {
    const bool tm = L->timing;
    int ne = 0;
    int kid[PC_NKERN + 1];
    auto mark = [&](int id) {
      if (tm) {
        kid[ne] = id;
        (void)hipEventRecord(L->ev[ne++], st);
      }
    };
    auto run = [&](int stage, int id, int sub = -1) {
      t_tag = stage * 8 + (sub < 0 ? 7 : sub);
      launch_stage<EOS>(L, stage, S[stage], P[stage], st, sub);
      mark(id);
    };
    mark(-1);
    // Branch mode (not while timing kernels one by one): the stages of the
    // three directions that do not depend on each other run on three
    // streams.  A graph capture stays on one stream; the captured chain is
    // rewired into the same dependencies afterwards (rewire_branches).
    const bool brm = L->brs.on && !tm && !capturing;
    if (L->mol) {
      for (int h = 0; h < 2; ++h) {
        run(ST_MOL_PRIM, PC_K_CTOPRIM, h);
        if (brm) {
          branch_mol_fluxes<EOS>(L, S, P, st, L->brs);
        } else {
          for (int d = 0; d < 3; ++d) run(ST_MOL_X + d, PC_K_MOL_X + d);
        }
        run(ST_MOL_UPD, h == 0 ? PC_K_MOL_UPD1 : PC_K_MOL_UPD2, h);
      }
    } else if (brm) {
      run(ST_CTOPRIM, PC_K_CTOPRIM);
      run(ST_DIVU, PC_K_DIVU);
      branch_ctu_fluxes<EOS>(L, S, P, st, L->brs);
      run(ST_UPDATE, PC_K_UPDATE);
    } else {
      run(ST_CTOPRIM, PC_K_CTOPRIM);
      run(ST_DIVU, PC_K_DIVU);
      for (int d = 0; d < 3; ++d) run(ST_TRACE, PC_K_TRACE_X + d, d);
      for (int x = ST_FLUX1_X; x <= ST_FLUX1_Z; ++x)
        run(x, PC_K_FLUX1_X + (x - ST_FLUX1_X));
      for (int x = ST_TRANS_XY; x <= ST_TRANS_ZY; ++x)
        run(x, PC_K_TRANS_XY + (x - ST_TRANS_XY));
      for (int x = ST_FINAL_X; x <= ST_FINAL_Z; ++x)
        run(x, PC_K_FINAL_X + (x - ST_FINAL_X));
      run(ST_UPDATE, PC_K_UPDATE);
    }
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, L->ds, L->la, DT_NEW);
    t_tag = TAG_DT;
    tag_launch();
    mark(PC_K_DT);
    if (tm) {
      (void)hipEventSynchronize(L->ev[ne - 1]);
      for (int e = 0; e + 1 < ne; ++e) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, L->ev[e], L->ev[e + 1]);
        L->ktime[kid[e + 1]] += ms;
      }
      L->nsteps_timed++;
    }
  }
  HIPCHK(L->ctx, hipGetLastError());
  return PC_OK;
}

// Candidate block shapes of the autotuner: {lx, ly, lz, workgroups per CU}
// (lx = -1: runs of BLOCK cells in i-fastest order; a non-zero last entry
// makes the launch persistent with that many workgroups per compute unit).
static const int kTuneCand[][4] = {
    {-1, 0, 0, 0}, {-1, 0, 0, 2}, {-1, 0, 0, 4}, {6, 2, 0, 0}, {5, 3, 0, 0},
    {5, 2, 1, 0},  {4, 4, 0, 0},  {4, 2, 2, 0},  {3, 3, 2, 0}};
static constexpr int kNTune = sizeof(kTuneCand) / sizeof(kTuneCand[0]);

// Time each candidate plan of every stage on the level as it stands (the
// intermediates of one step are computed first; the update stage runs on a
// saved copy of the state, restored before each launch) and keep the
// fastest.  The state and dt are left as they were.
template <class EOS>
static int level_autotune(pc_level* L, int reps, hipStream_t st) {
  pc_ctx* c = L->ctx;
  graph_drop(L);
  int ncu = 0;
  HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                  c->device < 0 ? 0 : c->device));
  const size_t sbytes = sizeof(double) * (size_t)L->state_total;
  double* saved = nullptr;
  if (hipMalloc(&saved, sbytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, PC_ERR_NOMEM, "autotune: no room for a state copy");
  }
  DtState ds0;
  HIPCHK(c, hipMemcpy(&ds0, L->ds, sizeof(DtState), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpyAsync(saved, L->state, sbytes, hipMemcpyDeviceToDevice, st));
  // intermediates of one step (the MOL first-stage update also leaves U* in
  // the stage arrays, which the second-stage update timed below reads; the
  // state is restored before each of its runs)
  std::vector<int> tune;
  if (L->mol) {
    for (int s : {ST_GHOST, ST_MOL_PRIM, ST_MOL_X, ST_MOL_Y, ST_MOL_Z})
      launch_stage<EOS>(L, s, L->scs[s], L->plans[s], st);
    launch_stage<EOS>(L, ST_MOL_UPD, L->scs[ST_MOL_UPD], L->plans[ST_MOL_UPD], st, 0);
    tune = {ST_GHOST, ST_MOL_PRIM, ST_MOL_X, ST_MOL_Y, ST_MOL_Z, ST_MOL_UPD};
  } else {
    for (int s : {ST_GHOST, ST_CTOPRIM, ST_DIVU, ST_TRACE, ST_FLUX1_X, ST_FLUX1_Y,
                  ST_FLUX1_Z, ST_TRANS_XY, ST_TRANS_XZ, ST_TRANS_YX, ST_TRANS_YZ,
                  ST_TRANS_ZX, ST_TRANS_ZY, ST_FINAL_X, ST_FINAL_Y, ST_FINAL_Z})
      launch_stage<EOS>(L, s, L->scs[s], L->plans[s], st);
    for (int s = 0; s <= ST_GHOST; ++s) tune.push_back(s);
  }
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int rc = PC_OK;
  for (int s : tune) {
    if (rc != PC_OK) break;
    double best = 1e300;
    int bestc = -1;
    StagePlan bestP;
    for (int k = 0; k < kNTune && rc == PC_OK; ++k) {
      // the ghost fill enumerates its cells itself: only the persistence of
      // the linear shapes differs
      if (s == ST_GHOST && kTuneCand[k][0] >= 0) continue;
      StagePlan P;
      const int cap = kTuneCand[k][3] * ncu;
      if (!plan_make(L, s, kTuneCand[k][0], kTuneCand[k][1], kTuneCand[k][2],
                     cap, P)) {
        rc = fail(c, PC_ERR_NOMEM, "autotune: plan allocation failed");
        break;
      }
      const Sched sc = plan_sched(L, P, s);
      double tot = 0.0;
      for (int r = 0; r <= reps; ++r) {  // r = 0 warms up
        if (s == ST_UPDATE || s == ST_MOL_UPD)
          (void)hipMemcpyAsync(L->state, saved, sbytes, hipMemcpyDeviceToDevice, st);
        (void)hipEventRecord(e0, st);
        launch_stage<EOS>(L, s, sc, P, st);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = fail(c, PC_ERR_HIP, "autotune: launch failed");
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tot += ms;
      }
      const double avg = tot / reps;
      if (rc == PC_OK && avg < best) {
        if (bestc >= 0) plan_free(bestP);
        best = avg;
        bestc = k;
        bestP = P;
      } else {
        plan_free(P);
      }
    }
    if (bestc >= 0) {
      plan_free(L->plans[s]);
      L->plans[s] = bestP;
      L->scs[s] = plan_sched(L, bestP, s);
      L->tuned_ms[s] = best;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipMemcpyAsync(L->state, saved, sbytes, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(L->ds, &ds0, sizeof(DtState), hipMemcpyHostToDevice, st);
  (void)hipStreamSynchronize(st);
  (void)hipFree(saved);
  HIPCHK(c, hipGetLastError());
  return rc;
}

// Rewire a graph captured on one stream (a chain of kernel nodes, tags[i] the
// stage tag of node i) into the dependencies of the branch streams
// (branch_ctu_fluxes, branch_mol_fluxes): per direction d, trace_d (and the
// first fluxes) after divu; the transverse fluxes of d (one node for both t)
// after the first fluxes of all three directions; final_d after the transverse fluxes (t0|t1) and (t1|t0)
// and the first fluxes of d; the update after the three finals; the MOL
// fluxes of the three directions after the primitives, the update after all
// three; every other node after its predecessor.  Returns false, with the
// graph unchanged, when the chain does not have that shape.
static bool rewire_branches(hipGraph_t g, const std::vector<int>& tags) {
  const size_t n = tags.size();
  std::vector<hipGraphNode_t> chain;
  size_t nr = 0;
  if (hipGraphGetRootNodes(g, nullptr, &nr) != hipSuccess || nr != 1) return false;
  hipGraphNode_t cur;
  if (hipGraphGetRootNodes(g, &cur, &nr) != hipSuccess) return false;
  while (true) {
    chain.push_back(cur);
    size_t nd = 0;
    if (hipGraphNodeGetDependentNodes(cur, nullptr, &nd) != hipSuccess || nd > 1)
      return false;
    if (nd == 0) break;
    if (hipGraphNodeGetDependentNodes(cur, &cur, &nd) != hipSuccess) return false;
    if (chain.size() > n) return false;
  }
  if (chain.size() != n) return false;
  std::vector<std::vector<int>> dep(n);
  int F1[3] = {-1, -1, -1}, FIN[3] = {-1, -1, -1}, MOL[3] = {-1, -1, -1};
  int TR[3][3];
  for (auto& r : TR) r[0] = r[1] = r[2] = -1;
  int join = -1;
  for (size_t i = 0; i < n; ++i) {
    const int st = tags[i] / 8, sub = tags[i] % 8;
    const int k = (int)i;
    if (st == ST_TRACE && sub < 3) {
      dep[i] = {join};
      F1[sub] = k;
    } else if (st >= ST_FLUX1_X && st <= ST_FLUX1_Z) {
      const int d = st - ST_FLUX1_X;
      dep[i] = {F1[d]};
      F1[d] = k;
    } else if (st >= ST_TRANS_XY && st <= ST_TRANS_ZY) {
      // both transverse fluxes of d (the second stage launches nothing)
      const int d = (st - ST_TRANS_XY) / 2;
      const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
      dep[i] = {F1[d], F1[t0], F1[t1]};
      TR[d][t0] = TR[d][t1] = k;
    } else if (st >= ST_FINAL_X && st <= ST_FINAL_Z) {
      const int d = st - ST_FINAL_X;
      const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
      if (FIN[d] < 0 || (i > 0 && tags[i - 1] != tags[i]))
        dep[i] = {TR[t0][t1], TR[t1][t0], F1[d]};
      else
        dep[i] = {FIN[d]};  // the low-order faces of the same stage
      FIN[d] = k;
    } else if (st == ST_UPDATE) {
      dep[i] = {FIN[0], FIN[1], FIN[2]};
      join = k;
      for (int d = 0; d < 3; ++d) F1[d] = FIN[d] = -1;
      for (auto& r : TR) r[0] = r[1] = r[2] = -1;
    } else if (st >= ST_MOL_X && st <= ST_MOL_Z) {
      dep[i] = {join};
      MOL[st - ST_MOL_X] = k;
    } else if (st == ST_MOL_UPD) {
      dep[i] = {MOL[0], MOL[1], MOL[2]};
      join = k;
      MOL[0] = MOL[1] = MOL[2] = -1;
    } else {  // ghost, primitives, divu, dt: after the previous node
      if (i > 0) dep[i] = {k - 1};
      join = k;
    }
    for (int x : dep[i])
      if (x < 0) return false;
  }
  for (size_t i = 0; i + 1 < n; ++i)
    if (hipGraphRemoveDependencies(g, &chain[i], &chain[i + 1], 1) != hipSuccess)
      return false;
  for (size_t i = 0; i < n; ++i)
    for (int x : dep[i])
      if (hipGraphAddDependencies(g, &chain[x], &chain[i], 1) != hipSuccess)
        return false;
  return true;
}

template <class EOS>
static int level_step_graph(pc_level* L, int nsteps, hipStream_t st) {
  pc_ctx* c = L->ctx;
  if (!L->gvalid) {
    if (!L->cap) HIPCHK(c, hipStreamCreateWithFlags(&L->cap, hipStreamNonBlocking));
    HIPCHK(c, hipStreamSynchronize(st));
    HIPCHK(c, hipStreamBeginCapture(L->cap, hipStreamCaptureModeThreadLocal));
    std::vector<int> tags;
    t_tags = &tags;
    const int rc = level_step_eos<EOS>(L, L->gsteps, L->cap, true);
    t_tags = nullptr;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(L->cap, &g);
    if (rc != PC_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(c, PC_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    // with branch streams on, the replayed graph runs the directions'
    // stages concurrently, as the direct launches do
    L->gbranched = L->brs.on && rewire_branches(g, tags);
    L->gnodes = (int)tags.size();
    L->graph = g;
    HIPCHK(c, hipGraphInstantiate(&L->gexec, L->graph, nullptr, nullptr, 0));
    L->gvalid = true;
  }
  for (int r = 0; r < nsteps / L->gsteps; ++r) HIPCHK(c, hipGraphLaunch(L->gexec, st));
  return level_step_eos<EOS>(L, nsteps % L->gsteps, st);
}

extern "C" {

int pc_level_step_v1(pc_level_t* L, int nsteps, void* stream) {
  if (!L || nsteps < 0) return PC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const bool g = L->gsteps > 0 && !L->timing;
  if (L->ctx->ns == 1)
    return g ? level_step_graph<EosGamma>(L, nsteps, st)
             : level_step_eos<EosGamma>(L, nsteps, st);
  return g ? level_step_graph<EosLiDryer>(L, nsteps, st)
           : level_step_eos<EosLiDryer>(L, nsteps, st);
}

int pc_level_autotune_v1(pc_level_t* L, int reps, void* stream) {
  if (!L || reps < 1) return PC_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (L->ctx->ns == 1) return level_autotune<EosGamma>(L, reps, st);
  return level_autotune<EosLiDryer>(L, reps, st);
}

int pc_level_plan_v1(const pc_level_t* L, int stage, int* shape, double* ms) {
  if (!L || stage < 0 || stage >= ST_COUNT || !shape) return PC_ERR_ARG;
  const StagePlan& P = L->plans[stage];
  shape[0] = P.lx;
  shape[1] = P.ly;
  shape[2] = P.lz;
  shape[3] = P.grid;
  shape[4] = P.nblocks;
  if (ms) *ms = L->tuned_ms[stage];
  return PC_OK;
}

int pc_level_set_plan_v1(pc_level_t* L, int stage, int lx, int ly, int lz,
                         int wg_per_cu) {
  if (!L || stage < 0 || stage >= ST_COUNT || wg_per_cu < 0)
    return fail(L ? L->ctx : nullptr, PC_ERR_ARG, "bad plan request");
  if (lx >= 0 && (ly < 0 || lz < 0 || (1 << (lx + ly + lz)) != BLOCK))
    return fail(L->ctx, PC_ERR_ARG, "tile 2^%d x 2^%d x 2^%d is not %d cells",
                lx, ly, lz, BLOCK);
  int ncu = 0;
  HIPCHK(L->ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                       L->ctx->device < 0 ? 0 : L->ctx->device));
  HIPCHK(L->ctx, hipDeviceSynchronize());
  StagePlan P;
  if (!plan_make(L, stage, lx < 0 ? -1 : lx, ly, lz, wg_per_cu * ncu, P))
    return fail(L->ctx, PC_ERR_NOMEM, "plan allocation failed");
  graph_drop(L);
  plan_free(L->plans[stage]);
  L->plans[stage] = P;
  L->scs[stage] = plan_sched(L, P, stage);
  L->tuned_ms[stage] = 0.0;
  return PC_OK;
}

int pc_level_set_precision_v1(pc_level_t* L, int mixed) {
  if (!L || (mixed != 0 && mixed != 1)) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipDeviceSynchronize());
  graph_drop(L);
  if (L->mixed == mixed) return PC_OK;
  L->mixed = mixed;
  // the transverse stage's threads per face and the MOL flux tiles follow
  // the precision: re-plan those launches, keeping their shapes
  const int st[6] = {ST_TRANS_XY, ST_TRANS_YX, ST_TRANS_ZX, ST_MOL_X, ST_MOL_Y, ST_MOL_Z};
  StagePlan P[6];
  for (int x = 0; x < 6; ++x) {
    const StagePlan& O = L->plans[st[x]];
    const int cap = O.grid < O.nblocks ? O.grid : 0;
    if (!plan_make(L, st[x], O.lx, O.ly, O.lz, cap, P[x])) {
      for (int y = 0; y < x; ++y) plan_free(P[y]);
      L->mixed = !mixed;
      return fail(L->ctx, PC_ERR_NOMEM, "set_precision: plan allocation failed");
    }
  }
  for (int x = 0; x < 6; ++x) {
    plan_free(L->plans[st[x]]);
    L->plans[st[x]] = P[x];
    L->scs[st[x]] = plan_sched(L, P[x], st[x]);
  }
  return PC_OK;
}

int pc_level_set_graph_v1(pc_level_t* L, int steps_per_graph) {
  if (!L || steps_per_graph < 0) return PC_ERR_ARG;
  graph_drop(L);
  L->gsteps = steps_per_graph;
  return PC_OK;
}

int pc_level_graph_info_v1(const pc_level_t* L, int* nodes, int* branched) {
  if (!L || !nodes || !branched) return PC_ERR_ARG;
  *nodes = L->gvalid ? L->gnodes : 0;
  *branched = L->gvalid && L->gbranched ? 1 : 0;
  return PC_OK;
}

int pc_level_set_branches_v1(pc_level_t* L, int on) {
  if (!L || (on != 0 && on != 1)) return PC_ERR_ARG;
  if (on) HIPCHK(L->ctx, L->brs.make());
  graph_drop(L);
  L->brs.on = on != 0;
  return PC_OK;
}

int pc_level_kernel_times_v1(pc_level_t* L, double* ms, int n) {
  if (!L || !ms || n <= 0) return PC_ERR_ARG;
  for (int s = 0; s < n && s < PC_NKERN; ++s)
    ms[s] = L->nsteps_timed ? L->ktime[s] / L->nsteps_timed : 0.0;
  return PC_OK;
}

}  // extern "C"

// ------------------------------------------------------------- plotfiles --
// Plot variables this writer derives, in the order the reference's plotfile
// lists them: state components in component order (Source/Setup.cpp:292-360),
// then derived quantities in derive-list registration order (:440-560).
struct PlotVar {
  std::string name;
  int code, comp;
};

static const char* species_name(int ns, int s) {
  // LiDryer mechanism order; the single gamma-law species carries N2's
  // molecular weight
  static const char* ld[9] = {"H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "N2"};
  return ns == 9 ? ld[s] : "N2";
}

static std::vector<PlotVar> plot_catalog(int ns) {
  std::vector<PlotVar> v;
  const char* st[7] = {"density", "xmom", "ymom", "zmom", "rho_E", "rho_e", "Temp"};
  for (int c = 0; c < 7; ++c) v.push_back({st[c], DV_STATE, c});
  for (int s = 0; s < ns; ++s)
    v.push_back({std::string("rho_") + species_name(ns, s), DV_STATE, UFS + s});
  v.push_back({"pressure", DV_PRESSURE, 0});
  v.push_back({"kineng", DV_KINENG, 0});
  v.push_back({"soundspeed", DV_SOUNDSPEED, 0});
  v.push_back({"MachNumber", DV_MACH, 0});
  v.push_back({"magvort", DV_MAGVORT, 0});
  v.push_back({"logden", DV_LOGDEN, 0});
  for (int s = 0; s < ns; ++s)
    v.push_back({std::string("Y(") + species_name(ns, s) + ")", DV_DIVRHO, UFS + s});
  v.push_back({"x_velocity", DV_DIVRHO, UMX});
  v.push_back({"y_velocity", DV_DIVRHO, UMY});
  v.push_back({"z_velocity", DV_DIVRHO, UMZ});
  v.push_back({"magvel", DV_MAGVEL, 0});
  return v;
}

extern "C" {

int pc_level_write_plotfile_v1(pc_level_t* L, const char* dir, const char* vars,
                               double time, int step, const double* prob_lo,
                               void* stream) {
  if (!L || !dir || !*dir) return PC_ERR_ARG;
  pc_ctx* c = L->ctx;
  const int ns = c->ns;
  const std::vector<PlotVar> cat = plot_catalog(ns);
  std::vector<char> sel(cat.size(), 0);
  if (!vars || !*vars) {
    for (int k = 0; k < 7 + ns; ++k) sel[k] = 1;  // every state component
  } else {
    std::istringstream in(vars);
    std::string tok;
    while (in >> tok) {
      bool hit = false;
      for (size_t k = 0; k < cat.size(); ++k) {
        const bool spec = cat[k].name.compare(0, 2, "Y(") == 0;
        if (tok == "ALL" || tok == cat[k].name || (tok == "massfrac" && spec)) {
          sel[k] = 1;
          hit = true;
        }
      }
      if (!hit) return fail(c, PC_ERR_ARG, "unknown plot variable '%s'", tok.c_str());
    }
  }
  DeriveArgs da;
  memset(&da, 0, sizeof(da));
  pcplot::PlotMeta meta;
  bool vort = false;
  for (size_t k = 0; k < cat.size(); ++k) {
    if (!sel[k]) continue;
    if (da.nplot == MAX_PLOT) return fail(c, PC_ERR_ARG, "more than %d plot variables", MAX_PLOT);
    da.code[da.nplot] = cat[k].code;
    da.comp[da.nplot] = cat[k].comp;
    da.nplot++;
    meta.names.push_back(cat[k].name);
    vort |= cat[k].code == DV_MAGVORT;
  }
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if (vort && (rc = level_fill(L, st))) return rc;  // magvort reads one ghost cell
  const int nb = (int)L->boxes.size();
  std::vector<long long> off(nb + 1, 0);
  for (int b = 0; b < nb; ++b) off[b + 1] = off[b] + L->boxes[b].npts() * da.nplot;
  double* dout = nullptr;
  Job* djobs = nullptr;
  StagePlan P;
  auto cleanup = [&]() {
    if (dout) (void)hipFree(dout);
    if (djobs) (void)hipFree(djobs);
    plan_free(P);
  };
  if (hipMalloc(&dout, sizeof(double) * off[nb]) != hipSuccess ||
      hipMalloc(&djobs, sizeof(Job) * nb) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return fail(c, PC_ERR_NOMEM, "plotfile staging allocation failed");
  }
  std::vector<Job> jobs = L->hjobs;
  for (int b = 0; b < nb; ++b) jobs[b].upd = fab_view(dout + off[b], L->boxes[b]);
  if (hipMemcpyAsync(djobs, jobs.data(), sizeof(Job) * nb, hipMemcpyHostToDevice,
                     st) != hipSuccess ||
      !plan_make(L, ST_UPDATE, -1, 0, 0, 0, P)) {
    cleanup();
    return fail(c, PC_ERR_HIP, "plotfile job table upload failed");
  }
  const Sched sc{djobs, P.blkmap, P.blkstart, P.nblocks, -1, 0, 0, nb};
  if (ns == 1) LAUNCH(k_derive<EosGamma>, P.nblocks, st, sc, L->ka, da);
  else LAUNCH(k_derive<EosLiDryer>, P.nblocks, st, sc, L->ka, da);
  std::vector<double> host((size_t)off[nb]);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess ||
      hipMemcpy(host.data(), dout, sizeof(double) * off[nb], hipMemcpyDeviceToHost) !=
          hipSuccess) {
    cleanup();
    return fail(c, PC_ERR_HIP, "plot variable derivation failed");
  }
  cleanup();
  // the level's domain (or the bounding box of self-periodic boxes)
  for (int d = 0; d < 3; ++d) {
    meta.domlo[d] = L->la.dlo[d];
    meta.domhi[d] = L->la.dhi[d];
    if (L->la.periodic_self) {
      meta.domlo[d] = L->boxes[0].lo[d];
      meta.domhi[d] = L->boxes[0].hi[d];
      for (const IBox& b : L->boxes) {
        meta.domlo[d] = std::min(meta.domlo[d], b.lo[d]);
        meta.domhi[d] = std::max(meta.domhi[d], b.hi[d]);
      }
    }
    meta.prob_lo[d] = prob_lo ? prob_lo[d] : 0.0;
    meta.dx[d] = L->hjobs[0].dx[d];
  }
  meta.time = time;
  meta.step = step;
  std::vector<pcplot::PlotBox> pb(nb);
  for (int b = 0; b < nb; ++b) {
    for (int d = 0; d < 3; ++d) {
      pb[b].lo[d] = L->boxes[b].lo[d];
      pb[b].hi[d] = L->boxes[b].hi[d];
    }
    pb[b].data = host.data() + off[b];
  }
  std::string msg;
  if (pcplot::write_plotfile(dir, meta, pb, msg) != 0)
    return fail(c, PC_ERR_IO, "%s", msg.c_str());
  return PC_OK;
}

}  // extern "C"

// ---------------------------------------------------------- flux register --
struct pc_fluxreg {
  pc_ctx* ctx = nullptr;
  int ratio = 2, nc = 0;
  int ncrse = 0, nfine = 0;
  int nfaces = 0;
  RegFace* faces = nullptr;  // device
  double* reg = nullptr;     // device, nfaces * nc
  // per coarse box: its faces (outside cell in the box) and reflux cells;
  // per fine box: its faces
  std::vector<int> c_first, c_count, cell_first, cell_count, f_first, f_count;
  int* cids = nullptr;        // device face ids grouped by coarse box
  int* fids = nullptr;        // device face ids grouped by fine box
  RegCell* cells = nullptr;   // device reflux cells grouped by coarse box
  int* cell_faces = nullptr;  // device
};

static void fluxreg_free(pc_fluxreg* R) {
  for (void* p : {(void*)R->faces, (void*)R->reg, (void*)R->cids, (void*)R->fids,
                  (void*)R->cells, (void*)R->cell_faces})
    if (p) (void)hipFree(p);
  delete R;
}

template <class T>
static bool upload(T** dst, const std::vector<T>& v) {
  *dst = nullptr;
  if (hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return v.empty() ||
         hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) ==
             hipSuccess;
}

extern "C" {

int pc_fluxreg_create_v1(pc_ctx_t* c, const pc_box_t* crse_boxes, int ncrse,
                         const pc_box_t* fine_boxes, int nfine,
                         const int* crse_domlo, const int* crse_domhi,
                         const int* periodic, int ratio, pc_fluxreg_t** out) {
  if (!c || !crse_boxes || ncrse <= 0 || !fine_boxes || nfine <= 0 ||
      !crse_domlo || !crse_domhi || !periodic || ratio < 2 || !out)
    return fail(c, PC_ERR_ARG, "bad flux register arguments");
  *out = nullptr;
  std::vector<IBox> cb(ncrse), fc(nfine);  // coarse boxes, coarsened fine boxes
  for (int b = 0; b < ncrse; ++b) cb[b] = mkbox(crse_boxes[b].lo, crse_boxes[b].hi);
  for (int b = 0; b < nfine; ++b)
    for (int d = 0; d < 3; ++d) {
      const int lo = fine_boxes[b].lo[d], hi = fine_boxes[b].hi[d];
      if (hi < lo || lo % ratio != 0 || (hi + 1) % ratio != 0 || lo < 0)
        return fail(c, PC_ERR_ARG, "fine box %d is not aligned to the ratio %d", b, ratio);
      fc[b].lo[d] = lo / ratio;
      fc[b].hi[d] = (hi + 1) / ratio - 1;
    }
  auto in = [](const IBox& b, const int* x) {
    return x[0] >= b.lo[0] && x[0] <= b.hi[0] && x[1] >= b.lo[1] && x[1] <= b.hi[1] &&
           x[2] >= b.lo[2] && x[2] <= b.hi[2];
  };
  auto covered = [&](const int* x) {
    for (const IBox& b : fc)
      if (in(b, x)) return true;
    return false;
  };
  auto owner = [&](const int* x) {
    for (int b = 0; b < ncrse; ++b)
      if (in(cb[b], x)) return b;
    return -1;
  };
  std::vector<RegFace> faces;
  std::vector<int> face_crse, face_fine;
  for (int b = 0; b < nfine; ++b)
    for (int d = 0; d < 3; ++d)
      for (int side = -1; side <= 1; side += 2) {
        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
        const int I = side < 0 ? fc[b].lo[d] : fc[b].hi[d] + 1;  // coarse face
        for (int K = fc[b].lo[t1]; K <= fc[b].hi[t1]; ++K)
          for (int J = fc[b].lo[t0]; J <= fc[b].hi[t0]; ++J) {
            int out_cell[3];
            out_cell[d] = side < 0 ? I - 1 : I;
            out_cell[t0] = J;
            out_cell[t1] = K;
            if (covered(out_cell)) continue;  // a face between fine boxes
            int shift = 0;                    // periodic image of the cell
            if (out_cell[d] < crse_domlo[d] || out_cell[d] > crse_domhi[d]) {
              if (!periodic[d]) continue;     // a domain face: no coarse cell
              const int L = crse_domhi[d] - crse_domlo[d] + 1;
              shift = out_cell[d] < crse_domlo[d] ? L : -L;
              out_cell[d] += shift;
            }
            const int ob = owner(out_cell);
            if (ob < 0)
              return fail(c, PC_ERR_ARG, "coarse cell next to fine box %d is on no coarse box", b);
            RegFace F;
            F.d = d;
            int fx[3];
            fx[d] = I + shift;
            fx[t0] = J;
            fx[t1] = K;
            F.i = fx[0];
            F.j = fx[1];
            F.k = fx[2];
            F.side = side < 0 ? -1 : 1;
            int ff[3];
            ff[d] = I * ratio;
            ff[t0] = J * ratio;
            ff[t1] = K * ratio;
            F.fi = ff[0];
            F.fj = ff[1];
            F.fk = ff[2];
            faces.push_back(F);
            face_crse.push_back(ob);
            face_fine.push_back(b);
          }
      }
  pc_fluxreg* R = new pc_fluxreg();
  R->ctx = c;
  R->ratio = ratio;
  R->nc = 7 + c->ns;
  R->ncrse = ncrse;
  R->nfine = nfine;
  R->nfaces = (int)faces.size();
  std::vector<int> cids, fids, cell_faces;
  std::vector<RegCell> cells;
  R->c_first.assign(ncrse, 0);
  R->c_count.assign(ncrse, 0);
  R->cell_first.assign(ncrse, 0);
  R->cell_count.assign(ncrse, 0);
  for (int b = 0; b < ncrse; ++b) {
    R->c_first[b] = (int)cids.size();
    R->cell_first[b] = (int)cells.size();
    std::map<std::array<int, 3>, std::vector<int>> by_cell;
    for (int f = 0; f < R->nfaces; ++f) {
      if (face_crse[f] != b) continue;
      cids.push_back(f);
      const RegFace& F = faces[f];
      std::array<int, 3> x = {F.i, F.j, F.k};
      if (F.side < 0) x[F.d] -= 1;
      by_cell[{x[2], x[1], x[0]}].push_back(f);
    }
    R->c_count[b] = (int)cids.size() - R->c_first[b];
    for (auto& kv : by_cell) {
      // a cell's faces in direction order, its low face before its high face
      std::stable_sort(kv.second.begin(), kv.second.end(), [&](int a, int b2) {
        const RegFace &A = faces[a], &B = faces[b2];
        return A.d != B.d ? A.d < B.d : A.side > B.side;
      });
      RegCell C;
      C.k = kv.first[0];
      C.j = kv.first[1];
      C.i = kv.first[2];
      C.first = (int)cell_faces.size();
      C.count = (int)kv.second.size();
      for (int f : kv.second) cell_faces.push_back(f);
      cells.push_back(C);
    }
    R->cell_count[b] = (int)cells.size() - R->cell_first[b];
  }
  R->f_first.assign(nfine, 0);
  R->f_count.assign(nfine, 0);
  for (int b = 0; b < nfine; ++b) {
    R->f_first[b] = (int)fids.size();
    for (int f = 0; f < R->nfaces; ++f)
      if (face_fine[f] == b) fids.push_back(f);
    R->f_count[b] = (int)fids.size() - R->f_first[b];
  }
  if (!upload(&R->faces, faces) || !upload(&R->cids, cids) || !upload(&R->fids, fids) ||
      !upload(&R->cells, cells) || !upload(&R->cell_faces, cell_faces) ||
      hipMalloc(&R->reg, sizeof(double) * std::max<size_t>((size_t)R->nfaces * R->nc, 1)) !=
          hipSuccess ||
      hipMemset(R->reg, 0, sizeof(double) * std::max<size_t>((size_t)R->nfaces * R->nc, 1)) !=
          hipSuccess) {
    (void)hipGetLastError();
    fluxreg_free(R);
    return fail(c, PC_ERR_NOMEM, "flux register allocation failed");
  }
  *out = R;
  return PC_OK;
}

int pc_fluxreg_destroy_v1(pc_fluxreg_t* R) {
  if (!R) return PC_ERR_ARG;
  (void)hipDeviceSynchronize();
  fluxreg_free(R);
  return PC_OK;
}

int pc_fluxreg_nfaces_v1(const pc_fluxreg_t* R) { return R ? R->nfaces : -1; }

int pc_fluxreg_reset_v1(pc_fluxreg_t* R, void* stream) {
  if (!R) return PC_ERR_ARG;
  // a fill kernel rather than a memset node, so the call can be captured
  const size_t n = std::max<size_t>((size_t)R->nfaces * R->nc, 1);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, R->reg, (long long)n, 0.0);
  HIPCHK(R->ctx, hipGetLastError());
  return PC_OK;
}

static int fluxreg_views(pc_fluxreg* R, const IBox& b, const pc_array4_t* flx,
                         V* f) {
  int rc;
  for (int d = 0; d < 3; ++d) {
    IBox fb = b;
    fb.hi[d] += 1;
    if ((rc = a4_view(R->ctx, flx[d], fb, R->nc, "flx", f[d]))) return rc;
  }
  return PC_OK;
}

int pc_fluxreg_crse_add_v1(pc_fluxreg_t* R, int crse_box, pc_box_t bx,
                           const pc_array4_t* flx, double scale, void* stream) {
  if (!R || !flx || crse_box < 0 || crse_box >= R->ncrse || !box_ok(bx))
    return fail(R ? R->ctx : nullptr, PC_ERR_ARG, "bad crse_add arguments");
  V f[3];
  int rc;
  if ((rc = fluxreg_views(R, mkbox(bx.lo, bx.hi), flx, f))) return rc;
  const int n = R->c_count[crse_box];
  if (n > 0)
    hipLaunchKernelGGL(k_fluxreg_crse, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, R->faces, R->cids + R->c_first[crse_box],
                       n, R->reg, R->nc, f[0], f[1], f[2], scale);
  HIPCHK(R->ctx, hipGetLastError());
  return PC_OK;
}

int pc_fluxreg_fine_add_v1(pc_fluxreg_t* R, int fine_box, pc_box_t bx,
                           const pc_array4_t* flx, double scale, void* stream) {
  if (!R || !flx || fine_box < 0 || fine_box >= R->nfine || !box_ok(bx))
    return fail(R ? R->ctx : nullptr, PC_ERR_ARG, "bad fine_add arguments");
  V f[3];
  int rc;
  if ((rc = fluxreg_views(R, mkbox(bx.lo, bx.hi), flx, f))) return rc;
  const int n = R->f_count[fine_box];
  if (n > 0)
    hipLaunchKernelGGL(k_fluxreg_fine, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, R->faces, R->fids + R->f_first[fine_box],
                       n, R->reg, R->nc, f[0], f[1], f[2], R->ratio, scale);
  HIPCHK(R->ctx, hipGetLastError());
  return PC_OK;
}

int pc_fluxreg_reflux_v1(pc_fluxreg_t* R, int crse_box, pc_box_t bx,
                         pc_array4_t state, pc_array4_t vol, double vol_u,
                         void* stream) {
  if (!R || crse_box < 0 || crse_box >= R->ncrse || !box_ok(bx))
    return fail(R ? R->ctx : nullptr, PC_ERR_ARG, "bad reflux arguments");
  const IBox b = mkbox(bx.lo, bx.hi);
  V U, vv = null_view();
  int rc;
  if ((rc = a4_view(R->ctx, state, b, R->nc, "uout", U))) return rc;
  if (vol.p && (rc = a4_view(R->ctx, vol, b, 1, "vol", vv))) return rc;
  const int n = R->cell_count[crse_box];
  if (n > 0)
    hipLaunchKernelGGL(k_fluxreg_reflux, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, R->faces, R->cells + R->cell_first[crse_box],
                       R->cell_faces, n, R->reg, R->nc, U, vv, vol_u);
  HIPCHK(R->ctx, hipGetLastError());
  return PC_OK;
}

int pc_average_down_v1(pc_ctx_t* c, pc_box_t crse_region, pc_array4_t crse,
                       pc_array4_t fine, int ratio, void* stream) {
  if (!c || !box_ok(crse_region) || ratio < 1)
    return fail(c, PC_ERR_ARG, "bad average_down arguments");
  const IBox cbx = mkbox(crse_region.lo, crse_region.hi);
  IBox fbx;
  for (int d = 0; d < 3; ++d) {
    fbx.lo[d] = cbx.lo[d] * ratio;
    fbx.hi[d] = (cbx.hi[d] + 1) * ratio - 1;
  }
  V C, F;
  int rc;
  const int nc = 7 + c->ns;
  if ((rc = a4_view(c, crse, cbx, nc, "uout", C))) return rc;
  if ((rc = a4_view(c, fine, fbx, nc, "uout", F))) return rc;
  const long long n = cbx.npts();
  hipLaunchKernelGGL(k_average_down, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, C, F, cbx.lo[0], cbx.lo[1], cbx.lo[2],
                     cbx.hi[0] - cbx.lo[0] + 1, cbx.hi[1] - cbx.lo[1] + 1,
                     cbx.hi[2] - cbx.lo[2] + 1, ratio, nc);
  HIPCHK(c, hipGetLastError());
  return PC_OK;
}

int pc_write_plotfile_v1(const char* dir, int nvars, const char* const* names,
                         int nboxes, const pc_box_t* boxes,
                         const double* const* data, const int* domlo,
                         const int* domhi, const double* prob_lo,
                         const double* dx, double time, int step) {
  if (!dir || !*dir || nvars <= 0 || !names || nboxes <= 0 || !boxes || !data ||
      !domlo || !domhi || !dx)
    return PC_ERR_ARG;
  pcplot::PlotMeta meta;
  for (int v = 0; v < nvars; ++v) {
    if (!names[v] || !*names[v]) return PC_ERR_ARG;
    meta.names.push_back(names[v]);
  }
  for (int d = 0; d < 3; ++d) {
    meta.domlo[d] = domlo[d];
    meta.domhi[d] = domhi[d];
    meta.prob_lo[d] = prob_lo ? prob_lo[d] : 0.0;
    meta.dx[d] = dx[d];
  }
  meta.time = time;
  meta.step = step;
  std::vector<pcplot::PlotBox> pb(nboxes);
  for (int b = 0; b < nboxes; ++b) {
    if (!box_ok(boxes[b]) || !data[b]) return PC_ERR_ARG;
    for (int d = 0; d < 3; ++d) {
      pb[b].lo[d] = boxes[b].lo[d];
      pb[b].hi[d] = boxes[b].hi[d];
    }
    pb[b].data = data[b];
  }
  std::string msg;
  return pcplot::write_plotfile(dir, meta, pb, msg) == 0 ? PC_OK : PC_ERR_IO;
}

int pc_level_status_v1(pc_level_t* L, int* bad, int* small_dt) {
  if (!L) return PC_ERR_ARG;
  HIPCHK(L->ctx, hipDeviceSynchronize());
  DtState s;
  HIPCHK(L->ctx, hipMemcpy(&s, L->ds, sizeof(s), hipMemcpyDeviceToHost));
  if (bad) *bad = s.bad;
  if (small_dt) *small_dt = s.small;
  return s.bad ? PC_ERR_NAN : (s.small ? PC_ERR_DT : PC_OK);
}

int pc_level_prim_to_cons_v1(pc_level_t* L, int b, pc_array4_t prim, int mode,
                             void* stream) {
  if (!L || b < 0 || b >= (int)L->boxes.size() || (mode != 0 && mode != 1))
    return PC_ERR_ARG;
  const IBox& bx = L->boxes[b];
  Job J;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) { J.lo[d] = bx.lo[d]; J.hi[d] = bx.hi[d]; }
  J.U = fab_view(L->state + L->off[b], growb(bx, 4));
  V pv;
  int rc;
  if ((rc = a4_view(L->ctx, prim, bx, 5 + L->ctx->ns, "prim", pv))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int g = grid1(ST_UPDATE, J);
  const Sched s0 = one_job_sched(L->jobs + b);
  if (L->ctx->ns == 1)
    LAUNCH(k_prim_to_cons<EosGamma>, g, st, s0, L->ka, pv, mode);
  else
    LAUNCH(k_prim_to_cons<EosLiDryer>, g, st, s0, L->ka, pv, mode);
  HIPCHK(L->ctx, hipGetLastError());
  return PC_OK;
}

}  // extern "C"

// ==================================================================== batch
// Independent requests (one box each, its own domain, BCs, cell size and
// dt) resident in one device arena.  A step advances any subset of them with
// one launch per stage over all their boxes; the job table of the subset is
// rebuilt only when the subset changes.

// First-fit allocator over one device allocation (host-side bookkeeping).
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  int refs = 0;  // batches drawing on this arena
  std::map<size_t, size_t> fr;  // free ranges: offset -> bytes
  static size_t round(size_t b) { return (b + 255) & ~size_t(255); }
  long long alloc(size_t bytes) {
    bytes = round(bytes);
    for (auto it = fr.begin(); it != fr.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      fr.erase(it);
      if (rest) fr[off + bytes] = rest;
      used += bytes;
      return (long long)off;
    }
    return -1;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
    used -= bytes;
    auto it = fr.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != fr.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      fr.erase(nx);
    }
    if (it != fr.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        fr.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (const auto& f : fr) m = std::max(m, f.second);
    return m;
  }
};

struct BatchSlot {
  bool live = false;
  IBox box;
  size_t off = 0, bytes = 0;
  long long nstate = 0;  // doubles of the state FAB
  Job job;
};

struct pc_batch {
  pc_ctx* ctx = nullptr;
  pc_params_t prm;
  int mixed = 0;
  bool mol = false;       // method-of-lines steps (do_mol)
  int trace_cells = 0;    // largest LDS trace window of the stepped set
  bool fuse = false;      // PPM trace and first fluxes in one kernel
  int mol_cells = 0;      // largest MOL flux window of the stepped set
  bool mol_lds = false;   // MOL fluxes from face pencils in LDS
  bool laxf = false;      // MOL with the Lax-Friedrichs flux
  KArgs ka;
  LevelArgs la;
  int lowo[3] = {0, 0, 0};
  DtState* ds = nullptr;  // device, one per slot
  Arena* arena = nullptr;  // possibly shared with other batches
  std::vector<BatchSlot> slots;
  int max_slots = 0;
  // The tables of the stepped subset in NTAB rotating sets, so a new set is
  // written (pinned host -> device, on the step's stream) while steps that
  // read an older one may still be queued: set k is rewritten only after
  // tev[k], recorded behind the last step that read it, has completed.
  static constexpr int NTAB = 4;
  Job* djobs = nullptr;   // device job table of the stepped subset (set tab)
  int* dstart = nullptr;  // device [ST_COUNT][max_slots] first blocks
  int* dids = nullptr;    // device slot ids of the stepped subset
  Job* tjobs[NTAB] = {};  // device sets
  int* tstart[NTAB] = {};
  int* tids[NTAB] = {};
  Job* hjobs[NTAB] = {};  // pinned host staging of each set
  int* hstart[NTAB] = {};
  int* hids[NTAB] = {};
  hipEvent_t tev[NTAB] = {};
  bool tev_made = false;
  int tab = 0;
  Job* dslot = nullptr;     // device job of each slot (admission kernels)
  Job* hslot = nullptr;     // pinned staging of dslot
  DtState* hds0 = nullptr;  // pinned initial dt state of each slot
  DtState* hsnap = nullptr; // pinned snapshots of each slot's dt state
  std::vector<int> cur;   // the stepped subset the tables describe
  StagePlan plans[ST_COUNT];
  Sched scs[ST_COUNT];
  Branches brs;  // per-direction streams (pc_batch_set_branches_v1)
};

static size_t batch_slot_bytes(const pc_batch* B, const IBox& bx) {
  const int ns = B->ctx->ns;
  const size_t st = (size_t)growb(bx, 4).npts() * (7 + ns);
  const size_t ws = ws_layout(bx, ns, true, false, true, B->mol).total;
  return Arena::round(st * sizeof(double)) + ws * sizeof(double);
}

static void batch_free(pc_batch* B) {
  if (B->arena && --B->arena->refs == 0) {
    if (B->arena->base) (void)hipFree(B->arena->base);
    delete B->arena;
  }
  if (B->ds) (void)hipFree(B->ds);
  for (int k = 0; k < pc_batch::NTAB; ++k) {
    if (B->tjobs[k]) (void)hipFree(B->tjobs[k]);
    if (B->tstart[k]) (void)hipFree(B->tstart[k]);
    if (B->tids[k]) (void)hipFree(B->tids[k]);
    if (B->hjobs[k]) (void)hipHostFree(B->hjobs[k]);
    if (B->hstart[k]) (void)hipHostFree(B->hstart[k]);
    if (B->hids[k]) (void)hipHostFree(B->hids[k]);
    if (B->tev_made) (void)hipEventDestroy(B->tev[k]);
  }
  if (B->dslot) (void)hipFree(B->dslot);
  if (B->hslot) (void)hipHostFree(B->hslot);
  if (B->hds0) (void)hipHostFree(B->hds0);
  if (B->hsnap) (void)hipHostFree(B->hsnap);
  B->brs.release();
  delete B;
}

// Job table, per-stage first blocks and slot ids of `ids` on the device, in
// the next of the rotating table sets (copied on stream st, no host wait
// unless that set's last reader is still queued).
static int batch_tables(pc_batch* B, const std::vector<int>& ids,
                        hipStream_t st) {
  if (ids == B->cur) return PC_OK;
  pc_ctx* c = B->ctx;
  const int n = (int)ids.size();
  const int t = (B->tab + 1) % pc_batch::NTAB;
  HIPCHK(c, hipEventSynchronize(B->tev[t]));
  Job* jobs = B->hjobs[t];
  int* start = B->hstart[t];
  memset(start, 0, sizeof(int) * ST_COUNT * B->max_slots);
  for (int k = 0; k < n; ++k) jobs[k] = B->slots[ids[k]].job;
  B->fuse = !B->mol && B->prm.ppm_type == 1 && pencils_fit(jobs, n);
  B->trace_cells = trace_cells_of(jobs, n, B->fuse);
  B->mol_lds = B->mol && mol_fit(jobs, n);
  B->mol_cells = B->mol_lds ? mol_cells_of(jobs, n) : 0;
  B->tab = t;
  B->djobs = B->tjobs[t];
  B->dstart = B->tstart[t];
  B->dids = B->tids[t];
  for (int s = 0; s < ST_COUNT; ++s) {
    int tot = 0;
    const int lx = (s == ST_TRACE && B->prm.ppm_type == 1 && !B->mol)
                       ? (B->fuse ? -3 : -2)
                       : (s >= ST_MOL_X && s <= ST_MOL_Z && B->mol_lds && (!B->mixed || s == ST_MOL_X)) ? -4
                                                                                   : -1;
    for (int k = 0; k < n; ++k) {
      start[(size_t)s * B->max_slots + k] = tot;
      tot += (int)stage_blocks(s, jobs[k], lx, 0, 0, B->mixed != 0);
    }
    B->plans[s].nblocks = B->plans[s].grid = tot;
    B->plans[s].lx = -1;
    B->scs[s] = Sched{B->djobs, nullptr, B->dstart + (size_t)s * B->max_slots,
                      tot, -1, 0, 0, n};
  }
  memcpy(B->hids[t], ids.data(), sizeof(int) * n);
  HIPCHK(c, hipMemcpyAsync(B->djobs, jobs, sizeof(Job) * n, hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B->dstart, start, sizeof(int) * ST_COUNT * B->max_slots,
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B->dids, B->hids[t], sizeof(int) * n, hipMemcpyHostToDevice, st));
  B->cur = ids;
  return PC_OK;
}

template <class EOS>
static void batch_step_eos(pc_batch* B, int n, hipStream_t st) {
  const Sched* S = B->scs;
  const StagePlan* P = B->plans;
  if (B->mol) {
    for (int h = 0; h < 2; ++h) {
      launch_stage<EOS>(B, ST_MOL_PRIM, S[ST_MOL_PRIM], P[ST_MOL_PRIM], st, h);
      if (B->brs.on) {
        branch_mol_fluxes<EOS>(B, S, P, st, B->brs);
      } else {
        for (int x : {ST_MOL_X, ST_MOL_Y, ST_MOL_Z})
          launch_stage<EOS>(B, x, S[x], P[x], st);
      }
      launch_stage<EOS>(B, ST_MOL_UPD, S[ST_MOL_UPD], P[ST_MOL_UPD], st, h);
    }
    hipLaunchKernelGGL(k_dt_finalize_list, dim3((n + 63) / 64), dim3(64), 0, st,
                       B->ds, (const int*)B->dids, n, B->la, DT_NEW);
    return;
  }
  launch_stage<EOS>(B, ST_CTOPRIM, S[ST_CTOPRIM], P[ST_CTOPRIM], st);
  launch_stage<EOS>(B, ST_DIVU, S[ST_DIVU], P[ST_DIVU], st);
  if (B->brs.on) {
    branch_ctu_fluxes<EOS>(B, S, P, st, B->brs);
  } else {
    launch_stage<EOS>(B, ST_TRACE, S[ST_TRACE], P[ST_TRACE], st);
    for (int x = ST_FLUX1_X; x <= ST_FINAL_Z; ++x)
      launch_stage<EOS>(B, x, S[x], P[x], st);
  }
  launch_stage<EOS>(B, ST_UPDATE, S[ST_UPDATE], P[ST_UPDATE], st);
  hipLaunchKernelGGL(k_dt_finalize_list, dim3((n + 63) / 64), dim3(64), 0, st,
                     B->ds, (const int*)B->dids, n, B->la, DT_NEW);
}

extern "C" {

int pc_batch_create_v1(pc_ctx_t* c, const pc_params_t* prm, int mixed,
                       long long budget_bytes, pc_batch_t* share,
                       int max_slots, pc_batch_t** out) {
  if (!c || !prm || !out || (budget_bytes <= 0 && !share) || max_slots <= 0 ||
      (mixed != 0 && mixed != 1))
    return fail(c, PC_ERR_ARG, "bad batch arguments");
  int rc;
  if (prm->do_mol) {
    if ((rc = check_mol_opts(c, prm))) return rc;
  } else if ((rc = check_opts(c, prm->ppm_type, prm->plm_iorder,
                              prm->use_hybrid_weno))) {
    return rc;
  }
  if (!prm->allow_small_energy || !prm->allow_negative_energy)
    return fail(c, PC_ERR_UNSUPPORTED,
                "allow_small_energy = allow_negative_energy = 1 required");
  pc_batch* B = new pc_batch;
  B->ctx = c;
  B->prm = *prm;
  B->mixed = mixed;
  B->mol = prm->do_mol != 0;
  B->laxf = prm->use_laxf_flux != 0;
  B->max_slots = max_slots;
  B->slots.resize(max_slots);
  memset(&B->ka, 0, sizeof(B->ka));
  B->ka.difmag = prm->difmag;
  B->ka.plm_iorder = prm->plm_iorder;
  B->ka.use_flattening = prm->use_flattening;
  B->ka.skip_reint = 1;  // as the level's
  memset(&B->la, 0, sizeof(B->la));
  B->la.periodic_self = 2;
  B->la.cfl = prm->cfl;
  B->la.max_dt = prm->max_dt;
  B->la.fixed_dt = prm->fixed_dt;
  B->la.change_max = prm->change_max;
  B->la.init_shrink = prm->init_shrink;
  B->la.dt_cutoff = prm->dt_cutoff;
  B->la.stop_time = prm->stop_time;
  auto dalloc = [&](void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return true;
  };
  if (share) {
    B->arena = share->arena;
    B->arena->refs++;
  } else {
    B->arena = new Arena;
    B->arena->refs = 1;
    B->arena->size = (size_t)budget_bytes;
    if (!dalloc((void**)&B->arena->base, B->arena->size)) {
      batch_free(B);
      return fail(c, PC_ERR_NOMEM, "batch arena of %lld bytes failed",
                  budget_bytes);
    }
    B->arena->fr[0] = B->arena->size;
  }
  auto halloc = [&](void** p, size_t bytes) {
    if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return false;
    }
    return true;
  };
  bool ok = dalloc((void**)&B->ds, sizeof(DtState) * max_slots) &&
            dalloc((void**)&B->dslot, sizeof(Job) * max_slots) &&
            halloc((void**)&B->hslot, sizeof(Job) * max_slots) &&
            halloc((void**)&B->hds0, sizeof(DtState) * max_slots) &&
            halloc((void**)&B->hsnap, sizeof(DtState) * max_slots);
  for (int k = 0; ok && k < pc_batch::NTAB; ++k)
    ok = dalloc((void**)&B->tjobs[k], sizeof(Job) * max_slots) &&
         dalloc((void**)&B->tstart[k], sizeof(int) * ST_COUNT * max_slots) &&
         dalloc((void**)&B->tids[k], sizeof(int) * max_slots) &&
         halloc((void**)&B->hjobs[k], sizeof(Job) * max_slots) &&
         halloc((void**)&B->hstart[k], sizeof(int) * ST_COUNT * max_slots) &&
         halloc((void**)&B->hids[k], sizeof(int) * max_slots);
  if (ok) {
    int made = 0;
    while (made < pc_batch::NTAB &&
           hipEventCreateWithFlags(&B->tev[made], hipEventDisableTiming) == hipSuccess)
      ++made;
    if (made < pc_batch::NTAB) {
      for (int k = 0; k < made; ++k) (void)hipEventDestroy(B->tev[k]);
      (void)hipGetLastError();
      ok = false;
    } else {
      B->tev_made = true;
    }
  }
  if (!ok) {
    batch_free(B);
    return fail(c, PC_ERR_NOMEM, "batch tables allocation failed");
  }
  B->tab = 0;
  B->djobs = B->tjobs[0];
  B->dstart = B->tstart[0];
  B->dids = B->tids[0];
  *out = B;
  return PC_OK;
}

int pc_batch_destroy_v1(pc_batch_t* B) {
  if (!B) return PC_ERR_ARG;
  (void)hipDeviceSynchronize();
  batch_free(B);
  return PC_OK;
}

long long pc_batch_slot_bytes_v1(const pc_batch_t* B, pc_box_t bx) {
  if (!B || !box_ok(bx)) return -1;
  return (long long)batch_slot_bytes(B, mkbox(bx.lo, bx.hi));
}

int pc_batch_memory_v1(const pc_batch_t* B, long long* total, long long* used,
                       long long* largest_free) {
  if (!B) return PC_ERR_ARG;
  if (total) *total = (long long)B->arena->size;
  if (used) *used = (long long)B->arena->used;
  if (largest_free) *largest_free = (long long)B->arena->largest();
  return PC_OK;
}

int pc_batch_free_ranges_v1(const pc_batch_t* B, long long* offsets,
                            long long* sizes, int cap, int* count) {
  if (!B || !count || cap < 0 || (cap > 0 && (!offsets || !sizes)))
    return PC_ERR_ARG;
  int n = 0;
  for (const auto& f : B->arena->fr) {
    if (n < cap) {
      offsets[n] = (long long)f.first;
      sizes[n] = (long long)f.second;
    }
    ++n;
  }
  *count = n;
  return PC_OK;
}

int pc_batch_slot_extent_v1(const pc_batch_t* B, int slot, long long* offset,
                            long long* bytes) {
  if (!B || slot < 0 || slot >= B->max_slots || !B->slots[slot].live ||
      !offset || !bytes)
    return PC_ERR_ARG;
  *offset = (long long)B->slots[slot].off;
  *bytes = (long long)Arena::round(B->slots[slot].bytes);
  return PC_OK;
}

int pc_batch_add_v1(pc_batch_t* B, pc_box_t bx, const int* bclo,
                    const int* bchi, const double* dx, pc_array4_t prim,
                    int prim_mode, void* stream, int* slot_out) {
  if (!B || !box_ok(bx) || !bclo || !bchi || !dx || !slot_out ||
      (prim_mode != 0 && prim_mode != 1))
    return fail(B ? B->ctx : nullptr, PC_ERR_ARG, "bad request");
  pc_ctx* c = B->ctx;
  hipStream_t st = (hipStream_t)stream;
  int id = -1;
  for (int k = 0; k < B->max_slots; ++k)
    if (!B->slots[k].live) { id = k; break; }
  if (id < 0) return fail(c, PC_ERR_NOMEM, "all %d request slots in use", B->max_slots);
  const IBox box = mkbox(bx.lo, bx.hi);
  int rc;
  V pv;
  if ((rc = a4_view(c, prim, box, 5 + c->ns, "prim", pv))) return rc;
  {
    const WsLayout W = ws_layout(box, c->ns, true, false, true, B->mol);
    const long long big = box_max_array_bytes(box, c->ns, W);
    if (big > MAX_FAB_BYTES)
      return fail(c, PC_ERR_UNSUPPORTED,
                  "request box: an array of %lld bytes exceeds the %lld the "
                  "kernels address with 32-bit offsets", big, MAX_FAB_BYTES);
  }
  const size_t bytes = batch_slot_bytes(B, box);
  const long long off = B->arena->alloc(bytes);
  if (off < 0)
    return fail(c, PC_ERR_NOMEM, "request of %zu bytes does not fit the budget",
                bytes);
  BatchSlot& S = B->slots[id];
  S.box = box;
  S.off = (size_t)off;
  S.bytes = bytes;
  S.nstate = growb(box, 4).npts() * (7 + c->ns);
  double* base = (double*)(B->arena->base + off);
  Job& J = S.job;
  memset(&J, 0, sizeof(J));
  for (int d = 0; d < 3; ++d) {
    J.lo[d] = J.dlo[d] = box.lo[d];
    J.hi[d] = J.dhi[d] = box.hi[d];
    J.bclo[d] = bclo[d];
    J.bchi[d] = bchi[d];
    J.qe[d] = box.hi[d] + 5;
  }
  set_geom(J, dx);
  J.U = fab_view(base, growb(box, 4));
  const WsLayout W = ws_layout(box, c->ns, true, false, true, B->mol);
  ws_fill_job(J, base + (Arena::round(S.nstate * sizeof(double)) / sizeof(double)),
              W, true, false, true);
  J.srcq = null_view();
  for (int d = 0; d < 3; ++d) J.area[d] = null_view();
  J.vol = null_view();
  J.upd = null_view();
  J.ds = B->ds + id;
  J.dtp = &B->ds[id].dt;
  for (int d = 0; d < 3; ++d) {
    auto lo_order = [](int b) { return b == PC_BC_INFLOW || b == PC_BC_USERBC; };
    if (lo_order(bclo[d]) || lo_order(bchi[d])) B->lowo[d] = 1;
  }
  // conservative state, then dt = init_shrink * estTimeStep of this request
  DtState z;
  z.dt = 0.0;
  z.minbits = 0x7FEFFFFFFFFFFFFFull;
  z.bad = 0;
  z.small = 0;
  z.time = 0.0;
  // the slot's staging is free: a slot is reused only after its previous
  // request's work has completed (pc_batch_remove_v1)
  B->hslot[id] = J;
  B->hds0[id] = z;
  HIPCHK(c, hipMemcpyAsync(B->dslot + id, B->hslot + id, sizeof(Job),
                           hipMemcpyHostToDevice, st));
  HIPCHK(c, hipMemcpyAsync(B->ds + id, B->hds0 + id, sizeof(z), hipMemcpyHostToDevice, st));
  const Sched s0 = one_job_sched(B->dslot + id);
  const int g = grid1(ST_UPDATE, J);
  // the Godunov step stores only the first ghost layer: the outer layers
  // start at zero rather than as the arena's previous contents
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((S.nstate + 255) / 256)), dim3(256), 0, st,
                     base, (long long)S.nstate, 0.0);
  if (c->ns == 1) {
    LAUNCH(k_prim_to_cons<EosGamma>, g, st, s0, B->ka, pv, prim_mode);
    LAUNCH((k_estdt<EosGamma, false>), g, st, s0, B->ka, B->ds + id);
  } else {
    LAUNCH(k_prim_to_cons<EosLiDryer>, g, st, s0, B->ka, pv, prim_mode);
    LAUNCH((k_estdt<EosLiDryer, false>), g, st, s0, B->ka, B->ds + id);
  }
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, B->ds + id, B->la, DT_INIT);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  S.live = true;
  *slot_out = id;
  return PC_OK;
}

int pc_batch_set_branches_v1(pc_batch_t* B, int on) {
  if (!B || (on != 0 && on != 1)) return PC_ERR_ARG;
  if (on) HIPCHK(B->ctx, B->brs.make());
  B->brs.on = on != 0;
  return PC_OK;
}

int pc_batch_step_v1(pc_batch_t* B, const int* slots, int n, void* stream) {
  if (!B || n < 0 || (n > 0 && !slots)) return PC_ERR_ARG;
  if (n == 0) return PC_OK;
  pc_ctx* c = B->ctx;
  std::vector<int> ids(slots, slots + n);
  std::vector<char> seen(B->max_slots, 0);
  for (int k = 0; k < n; ++k) {
    const int id = ids[k];
    if (id < 0 || id >= B->max_slots || !B->slots[id].live || seen[id])
      return fail(c, PC_ERR_ARG, "slot %d is not a live, distinct request", id);
    seen[id] = 1;
  }
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = batch_tables(B, ids, st))) return rc;
  if (c->ns == 1)
    batch_step_eos<EosGamma>(B, n, st);
  else
    batch_step_eos<EosLiDryer>(B, n, st);
  HIPCHK(c, hipGetLastError());
  // the current table set is in use until this step has run
  HIPCHK(c, hipEventRecord(B->tev[B->tab], st));
  return PC_OK;
}

int pc_batch_snapshot_v1(pc_batch_t* B, int slot, void* stream) {
  if (!B || slot < 0 || slot >= B->max_slots || !B->slots[slot].live)
    return PC_ERR_ARG;
  HIPCHK(B->ctx, hipMemcpyAsync(B->hsnap + slot, B->ds + slot, sizeof(DtState),
                                hipMemcpyDeviceToHost, (hipStream_t)stream));
  return PC_OK;
}

int pc_batch_snapshot_status_v1(const pc_batch_t* B, int slot, double* dt,
                                int* bad, int* small_dt) {
  if (!B || slot < 0 || slot >= B->max_slots || !B->slots[slot].live)
    return PC_ERR_ARG;
  const DtState s = B->hsnap[slot];
  if (dt) *dt = s.dt;
  if (bad) *bad = s.bad;
  if (small_dt) *small_dt = s.small;
  return s.bad ? PC_ERR_NAN : (s.small ? PC_ERR_DT : PC_OK);
}

int pc_batch_read_v1(pc_batch_t* B, int slot, double* buf) {
  if (!B || !buf || slot < 0 || slot >= B->max_slots || !B->slots[slot].live)
    return PC_ERR_ARG;
  const BatchSlot& S = B->slots[slot];
  HIPCHK(B->ctx, hipDeviceSynchronize());
  HIPCHK(B->ctx, hipMemcpy(buf, B->arena->base + S.off, S.nstate * sizeof(double),
                           hipMemcpyDefault));
  return PC_OK;
}

int pc_batch_status_v1(pc_batch_t* B, int slot, double* dt, int* bad,
                       int* small_dt) {
  if (!B || slot < 0 || slot >= B->max_slots || !B->slots[slot].live)
    return PC_ERR_ARG;
  DtState s;
  HIPCHK(B->ctx, hipDeviceSynchronize());
  HIPCHK(B->ctx, hipMemcpy(&s, B->ds + slot, sizeof(s), hipMemcpyDeviceToHost));
  if (dt) *dt = s.dt;
  if (bad) *bad = s.bad;
  if (small_dt) *small_dt = s.small;
  return s.bad ? PC_ERR_NAN : (s.small ? PC_ERR_DT : PC_OK);
}

int pc_batch_remove_v1(pc_batch_t* B, int slot, void* stream) {
  if (!B || slot < 0 || slot >= B->max_slots || !B->slots[slot].live)
    return PC_ERR_ARG;
  if (stream) HIPCHK(B->ctx, hipStreamSynchronize((hipStream_t)stream));
  BatchSlot& S = B->slots[slot];
  B->arena->release(S.off, S.bytes);
  S.live = false;
  for (int id : B->cur)
    if (id == slot) {
      B->cur.clear();  // tables must be rebuilt
      break;
    }
  return PC_OK;
}

}  // extern "C"
