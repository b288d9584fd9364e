// Enumeration of the ghost cells of a box (host and device; no HIP
// dependency, so the CPU tests compile it on its own).
#pragma once
#ifndef PC_HD
#if defined(__HIPCC__)
#define PC_HD __host__ __device__
#else
#define PC_HD
#endif
#endif

// The ghost cells of grow(bx,4) in a compact order: the 4 planes below and
// above the box (whole rows of grow(bx,4)), then per plane of the box the 4
// rows below and above it (whole rows) and the 4 + 4 cells at the two ends of
// each of its rows.  Threads of a launch map to cells in this order, so every
// lane of a wave has a ghost cell and rows are written in unit stride.
PC_HD inline long long ghost_shell_count(const int L[3], const int H[3]) {
  const long long n0 = H[0] - L[0] + 1, n1 = H[1] - L[1] + 1, n2 = H[2] - L[2] + 1;
  return (n0 + 8) * (n1 + 8) * (n2 + 8) - n0 * n1 * n2;
}
PC_HD inline bool ghost_shell_cell(const int L[3], const int H[3],
                                                 long long t, int& i, int& j, int& k) {
  const int n1 = H[1] - L[1] + 1, n2 = H[2] - L[2] + 1;
  const int N0 = H[0] - L[0] + 9, N1 = n1 + 8;
  const long long plane = (long long)N0 * N1, cap = 4 * plane;
  if (t < 2 * cap) {  // the 4 planes below, then the 4 above
    const bool above = t >= cap;
    const long long r = above ? t - cap : t;
    const int kk = (int)(r / plane), rr = (int)(r - kk * plane);
    k = above ? H[2] + 1 + kk : L[2] - 4 + kk;
    j = L[1] - 4 + rr / N0;
    i = L[0] - 4 + rr % N0;
    return true;
  }
  const long long per = 8LL * N0 + 8LL * n1;  // ghost cells of one box plane
  const long long u = t - 2 * cap;
  if (u >= per * n2) return false;
  const int kk = (int)(u / per);
  int r = (int)(u - kk * per);
  k = L[2] + kk;
  if (r < 8 * N0) {  // the 4 rows below, then the 4 above
    const bool above = r >= 4 * N0;
    if (above) r -= 4 * N0;
    j = above ? H[1] + 1 + r / N0 : L[1] - 4 + r / N0;
    i = L[0] - 4 + r % N0;
    return true;
  }
  r -= 8 * N0;  // 8 end cells per row of the box: 4 below L[0], 4 above H[0]
  j = L[1] + r / 8;
  const int x = r % 8;
  i = x < 4 ? L[0] - 4 + x : H[0] + 1 + (x - 4);
  return true;
}

