// Flux register and average-down for a coarse level and a level refined by
// an integer ratio r (the operations PeleC::update_flux_registers, reflux
// and avgDown perform through AMReX, Source/PeleC.cpp:1343-1437,
// Source/Hydro.cpp:280-288, for regular Cartesian boxes).
//
// A register face is a coarse face on the boundary of the refined region:
// one side is covered by a fine box, the other ("outside") coarse cell is
// not.  Each face accumulates
//     reg = sum_fine_steps s_f * (sum of the r^2 extensive fine fluxes on it)
//         + s_c * (extensive coarse flux on it)
// (the caller passes s_c = -dt_coarse and s_f = +dt_fine), and reflux adds
// -reg / V to an outside cell below the face (the face is its high face) and
// +reg / V to one above it: the coarse update is redone with the fine
// time-integrated flux on that face.
#pragma once
#include "pc_kernels.h"

namespace pcd {

// One register face: direction, coarse face index (i,j,k: the face between
// cells index-1 and index in direction d, in the index space of the coarse
// box that holds the outside cell), the outside cell's side (+1: above the
// face, -1: below) and the fine index of the first of its r^2 fine faces.
struct RegFace {
  int d;
  int i, j, k;      // coarse face index
  int side;         // +1: outside cell (i,j,k); -1: outside cell (i,j,k)-e_d
  int fi, fj, fk;   // fine face index of the first of the r^2 fine faces
};

__global__ void __launch_bounds__(BLOCK) k_fill(double* p, long long n, double v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// reg[f*nc + n] += scale * flux_c(face f, n) for the faces of one coarse box
__global__ void __launch_bounds__(BLOCK)
    k_fluxreg_crse(const RegFace* faces, const int* ids, int nfaces, double* reg,
                   int nc, V f0, V f1, V f2, double scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfaces) return;
  const int f = ids[t];
  const RegFace F = faces[f];
  const V& fl = F.d == 0 ? f0 : (F.d == 1 ? f1 : f2);
  for (int n = 0; n < nc; ++n)
    reg[(size_t)f * nc + n] += scale * fl(F.i, F.j, F.k, n);
}

// reg[f*nc + n] += scale * (sum of the r^2 fine fluxes), the sum taken over
// the two transverse fine offsets (second transverse direction outermost)
__global__ void __launch_bounds__(BLOCK)
    k_fluxreg_fine(const RegFace* faces, const int* ids, int nfaces, double* reg,
                   int nc, V f0, V f1, V f2, int ratio, double scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfaces) return;
  const int f = ids[t];
  const RegFace F = faces[f];
  const V& fl = F.d == 0 ? f0 : (F.d == 1 ? f1 : f2);
  const int t0 = F.d == 0 ? 1 : 0, t1 = F.d == 2 ? 1 : 2;
  for (int n = 0; n < nc; ++n) {
    double sum = 0.0;
    for (int b = 0; b < ratio; ++b)
      for (int a = 0; a < ratio; ++a) {
        int x[3] = {F.fi, F.fj, F.fk};
        x[t0] += a;
        x[t1] += b;
        sum += fl(x[0], x[1], x[2], n);
      }
    reg[(size_t)f * nc + n] += scale * sum;
  }
}

// Reflux of the outside cells of one coarse box: each listed cell adds, for
// each of its register faces (in direction order, the cell's low face before
// its high face), +reg / V (the cell is above the face) or -reg / V (below).
// cells[c] = (i, j, k, first, count) into cell_faces.
struct RegCell {
  int i, j, k;
  int first, count;
};

__global__ void __launch_bounds__(BLOCK)
    k_fluxreg_reflux(const RegFace* faces, const RegCell* cells,
                     const int* cell_faces, int ncells, const double* reg,
                     int nc, V U, V vol, double vol_u) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells) return;
  const RegCell C = cells[t];
  const double v = vol.p ? vol(C.i, C.j, C.k, 0) : vol_u;
  for (int n = 0; n < nc; ++n) {
    double u = U(C.i, C.j, C.k, n);
    for (int m = 0; m < C.count; ++m) {
      const int f = cell_faces[C.first + m];
      const double r = reg[(size_t)f * nc + n];
      u += faces[f].side > 0 ? r / v : -r / v;
    }
    U(C.i, C.j, C.k, n) = u;
  }
}

// Average-down of conserved quantities over one coarse box region [lo, hi]
// (uniform Cartesian volumes): crse = (1/r^3) * sum of the r^3 fine cells,
// the sum over the fine offsets with i fastest.
__global__ void __launch_bounds__(BLOCK)
    k_average_down(V crse, V fine, int lo0, int lo1, int lo2, int n0, int n1,
                   int n2, int ratio, int nc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n0 * n1 * n2) return;
  const int i = lo0 + (int)(t % n0);
  const int j = lo1 + (int)((t / n0) % n1);
  const int k = lo2 + (int)(t / ((long long)n0 * n1));
  const double volfrac = 1.0 / (double)(ratio * ratio * ratio);
  for (int n = 0; n < nc; ++n) {
    double c = 0.0;
    for (int kk = 0; kk < ratio; ++kk)
      for (int jj = 0; jj < ratio; ++jj)
        for (int ii = 0; ii < ratio; ++ii)
          c += fine(ratio * i + ii, ratio * j + jj, ratio * k + kk, n);
    crse(i, j, k, n) = volfrac * c;
  }
}

}  // namespace pcd
