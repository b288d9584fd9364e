// AMReX plotfile writer (host side of pc_level_write_plotfile_v1).
//
// Layout of a single-level plotfile as amrex::WriteMultiLevelPlotfile writes
// it (AMReX is an external dependency of the reference; this follows its
// published plotfile format, "HyperCLaw-V1.1" with a VisMF version-1 data
// header):
//   Header              variable names, dimension, time, geometry, per-level
//                       box list in physical coordinates, "Level_0/Cell"
//   Level_0/Cell_H      VisMF header: version, NFiles layout, ncomp, ngrow,
//                       the BoxArray, one "FabOnDisk: <file> <offset>" per
//                       box, per-box per-component minima and maxima
//   Level_0/Cell_D_00000  the FABs back to back, each a one-line text header
//                       (IEEE double descriptor, little-endian byte order,
//                       box, ncomp) followed by the raw components
#include "pc_plotfile.h"

#include <sys/stat.h>

#include <cerrno>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace pcplot {

namespace {

std::string ivec(const int* v) {
  std::ostringstream s;
  s << '(' << v[0] << ',' << v[1] << ',' << v[2] << ')';
  return s.str();
}

// amrex operator<<(Box): ((lo) (hi) (index type)), cell-centred
std::string box_str(const int* lo, const int* hi) {
  static const int cell[3] = {0, 0, 0};
  return "(" + ivec(lo) + " " + ivec(hi) + " " + ivec(cell) + ")";
}

int make_dir(const std::string& d, std::string& msg) {
  if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) {
    msg = "cannot create " + d + ": " + strerror(errno);
    return errno ? errno : -1;
  }
  return 0;
}

long long npts(const PlotBox& b) {
  return (long long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) *
         (b.hi[2] - b.lo[2] + 1);
}

}  // namespace

int write_plotfile(const std::string& dir, const PlotMeta& m,
                   const std::vector<PlotBox>& boxes, std::string& msg) {
  int rc;
  if ((rc = make_dir(dir, msg))) return rc;
  if ((rc = make_dir(dir + "/Level_0", msg))) return rc;
  const int nc = (int)m.names.size();

  // ---- data file: FABs back to back
  std::vector<long long> offsets;
  {
    std::ofstream d(dir + "/Level_0/Cell_D_00000", std::ios::binary);
    if (!d) {
      msg = "cannot write " + dir + "/Level_0/Cell_D_00000";
      return EIO;
    }
    for (const PlotBox& b : boxes) {
      offsets.push_back((long long)d.tellp());
      d << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
        << box_str(b.lo, b.hi) << ' ' << nc << '\n';
      d.write(reinterpret_cast<const char*>(b.data),
              (std::streamsize)(sizeof(double) * npts(b) * nc));
    }
    if (!d) {
      msg = "write error in " + dir + "/Level_0/Cell_D_00000";
      return EIO;
    }
  }

  // ---- VisMF header
  {
    std::ofstream h(dir + "/Level_0/Cell_H");
    if (!h) {
      msg = "cannot write " + dir + "/Level_0/Cell_H";
      return EIO;
    }
    h.setf(std::ios::scientific, std::ios::floatfield);
    h.precision(15);
    h << 1 << '\n' << 1 << '\n' << nc << '\n' << 0 << '\n';
    h << '(' << boxes.size() << ' ' << 0 << '\n';
    for (const PlotBox& b : boxes) h << box_str(b.lo, b.hi) << '\n';
    h << ')' << '\n';
    h << boxes.size() << '\n';
    for (size_t i = 0; i < boxes.size(); ++i)
      h << "FabOnDisk: Cell_D_00000 " << offsets[i] << '\n';
    h << '\n';
    for (int pass = 0; pass < 2; ++pass) {
      h << boxes.size() << ',' << nc << '\n';
      for (const PlotBox& b : boxes) {
        const long long n = npts(b);
        for (int c = 0; c < nc; ++c) {
          const double* p = b.data + (long long)c * n;
          double v = p[0];
          for (long long x = 1; x < n; ++x)
            v = pass == 0 ? (p[x] < v ? p[x] : v) : (p[x] > v ? p[x] : v);
          h << v << ',';
        }
        h << '\n';
      }
      h << '\n';
    }
    if (!h) {
      msg = "write error in " + dir + "/Level_0/Cell_H";
      return EIO;
    }
  }

  // ---- plotfile header
  {
    std::ofstream h(dir + "/Header");
    if (!h) {
      msg = "cannot write " + dir + "/Header";
      return EIO;
    }
    h.precision(17);
    h << "HyperCLaw-V1.1" << '\n' << nc << '\n';
    for (const std::string& s : m.names) h << s << '\n';
    h << 3 << '\n' << m.time << '\n' << 0 << '\n';
    double hi[3];
    for (int d = 0; d < 3; ++d)
      hi[d] = m.prob_lo[d] + m.dx[d] * (m.domhi[d] - m.domlo[d] + 1);
    for (int d = 0; d < 3; ++d) h << m.prob_lo[d] << ' ';
    h << '\n';
    for (int d = 0; d < 3; ++d) h << hi[d] << ' ';
    h << '\n';
    h << '\n';  // no refinement ratios on a single level
    h << box_str(m.domlo, m.domhi) << ' ' << '\n';
    h << m.step << ' ' << '\n';
    for (int d = 0; d < 3; ++d) h << m.dx[d] << ' ';
    h << '\n';
    h << 0 << '\n' << 0 << '\n';  // Cartesian, no boundary width
    h << 0 << ' ' << boxes.size() << ' ' << m.time << '\n';
    h << m.step << '\n';
    for (const PlotBox& b : boxes)
      for (int d = 0; d < 3; ++d) {
        const double lo = m.prob_lo[d] + m.dx[d] * (b.lo[d] - m.domlo[d]);
        const double up = m.prob_lo[d] + m.dx[d] * (b.hi[d] - m.domlo[d] + 1);
        h << lo << ' ' << up << '\n';
      }
    h << "Level_0/Cell" << '\n';
    if (!h) {
      msg = "write error in " + dir + "/Header";
      return EIO;
    }
  }
  return 0;
}

}  // namespace pcplot
