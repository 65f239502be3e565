// Pixel grid build and preplace on the device: internal interface between
// the host import (lg_host.cpp, compiled by the host compiler) and the
// device code (lg_grid.hip).  Plain C++ types only.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/skyline_lg.h"

// Inputs of lg_grid_preplace_device, all host arrays.  Grid coordinates
// count sites (x) and rows (y) from (lx, ly); device pixels are row-major.
struct LgGridBuild {
  int xs = 0, ys = 0, sw = 1, rh = 1, lx = 0, ly = 0;
  int core_lx = 0, core_ly = 0, core_ux = 0, core_uy = 0;
  const int32_t* vdd = nullptr;         // [ys] 1: VDD on top of row y
  const int32_t* span_start = nullptr;  // [ys + 1] CSR of the rows' x-spans per grid row
  const int32_t* span = nullptr;        // [2 * spans] first, last x (inclusive)
  int num_paint = 0;                    // blockage and fence regions in region order
  const int32_t* paint = nullptr;       // [5 * num_paint] x1 y1 x2 y2 (inclusive), pixel code
  int num_fences = 0;                   // fences (disjoint boxes): edge tests
  const int32_t* fence = nullptr;       // [12 * num_fences] expanded x1 y1 x2 y2, shrunk x1 y1 x2 y2, dbu lx ly ux uy
  int num_cells = 0;
  const int32_t* cell = nullptr;        // [8 * num_cells] lx ly w h ilx ily group gnd (dbu)
  int num_regions = 0;
  const int32_t* region = nullptr;      // [4 * num_regions] lx ly ux uy (dbu)
  int group_min = 0;                    // group lists are indexed by group - group_min
  int num_group_lists = 0;
  const int32_t* glist_start = nullptr;  // [num_group_lists + 1]
  const int32_t* glist = nullptr;        // region ids per group, in region order
  int num_dirs = 0;
  const int32_t* dirs = nullptr;        // [2 * num_dirs] checkHopeless's moves (x, y)
};

// Builds the pixel codes (valid ? group + 2 : 0, row-major, int16) in a new
// device buffer *d_pcode (the caller frees it with hipFree, or hands it to
// lg_create_device_grid) and preplaces the cells: out_lx / out_ly receive
// the preplaced lower-left corners (dbu), *fails the cells no region could
// take.  Synchronous on `stream` (NULL: the null stream).
int lg_grid_preplace_device(const LgGridBuild& in, void** d_pcode, int32_t* out_lx, int32_t* out_ly, int* fails,
                            void* stream, std::string& err);

// lg_create with the pixel codes already on the device (a buffer from
// lg_grid_preplace_device, owned by the handle from then on); the grid
// descriptor's pixel arrays are not read.
int lg_create_device_grid(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, void* d_pcode,
                          LgHandle** out);

// Host copy of device pixel codes (row-major int16) and their release.
int lg_grid_pcode_to_host(const void* d_pcode, size_t num_pixels, int16_t* out, std::string& err);
void lg_grid_free(void* d_pcode);
