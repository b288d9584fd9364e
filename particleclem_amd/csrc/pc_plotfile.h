// AMReX plotfile writer (host): the on-disk format of the reference's
// plotfiles (PeleCAmr::writePlotFileDoit, Source/PeleCAmr.cpp:225-267, which
// calls amrex::WriteMultiLevelPlotfile), single level, cell-centred data.
#pragma once
#include <string>
#include <vector>

namespace pcplot {

struct PlotBox {
  int lo[3], hi[3];
  const double* data;  // nvars components of the box, AMReX FAB layout
};

struct PlotMeta {
  std::vector<std::string> names;
  double time = 0.0;
  int step = 0;
  int domlo[3], domhi[3];
  double prob_lo[3], dx[3];
};

// Writes dir/Header, dir/Level_0/Cell_H and dir/Level_0/Cell_D_00000.
// Returns 0 or an errno-style code; msg receives a description on failure.
int write_plotfile(const std::string& dir, const PlotMeta& meta,
                   const std::vector<PlotBox>& boxes, std::string& msg);

}  // namespace pcplot
