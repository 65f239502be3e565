// ORACLE -- test infrastructure only.  CPU restatement of the SkyLine ADMM
// negotiation legalizer path, used by tests/, __graft_entry__.smoke() and the
// bench's cpu_baseline leg as the checker.  It is never linked into, loaded
// by, or called from the product library.
//
// Every function names the reference file:line range it restates (paths are
// relative to src/skyline/legalizer/ unless stated otherwise).
//
// Floating-point contract (shared with the HIP kernels): IEEE binary32 with
// round-to-nearest, no FMA contraction (-ffp-contract=off), IEEE division.
// The reference was compiled with --use_fast_math (CMakeLists.txt:12), whose
// exact bits cannot be reproduced; bit-parity is claimed against this contract.
// Reductions the reference does with order-dependent FP32 thrust reduces are
// made order-independent here (documented at each site).
//
// Parity pins (tests/test_oracle_*.py, test_tech_cpu.py, test_lefdef_cpu.py):
// brute-force enumeration of every candidate of a primal, refine and
// adaptive move (ICCAD17 weight, pin-access and edge-spacing terms
// included), closed forms for the dual step and rho / lambda0, the colouring
// and sort keys recounted, invariants after every iteration, the bitonic
// network against a reference sort, Thomas against a dense solve, the
// cut-row QP run to convergence against the bounded isotonic-regression
// closed form, the site-alignment DP against exhaustive enumeration, the
// checkPlace technology counts and the technology import against numpy
// recounts.  Each brute-force pin is shown to catch planted mistakes.
#include "../include/skyline_lg.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cctype>
#include <deque>
#include <fstream>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ref {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr int kIntMax = std::numeric_limits<int>::max();
constexpr int kIntMin = std::numeric_limits<int>::min();
constexpr int kDefaultGroup = -1;   // Legalizer.h:33
constexpr int kBlockageGroup = -2;  // Legalizer.h:34

static thread_local std::string g_err;

// Planted mistakes for the pin tests (0: none).  A pin test sets one with
// ref_set_mutation() and must then see the oracle disagree with the
// independent check it is pinned by; the solve paths never set it.
enum Mutation {
  kMutNone = 0,
  kMutPenaltySwapped = 1,     // standard and negotiation penalty exchanged
  kMutNoExcessTerm = 2,       // displacement beyond the threshold not charged
  kMutThresholdInSites = 3,   // threshold max_disp rows taken as sites
  kMutDualNoHalf = 4,         // dual step rho * g^2 instead of rho/2 * g^2
  kMutEvenNoDoubleStep = 5,   // even heights move by one row per dy
  kMutTieKeepsUpper = 6,      // tree-min replaces on >= (ties keep the upper slot)
  kMutNoFoldOffset = 7,       // partition colouring without the fold offset
  kMutKeyOverflowWeight1 = 8, // sort key overflow weight 1 instead of 10
  kMutNoStraddleCheck = 9,    // a cell takes its lower-left pixel's partition
  kMutNoProjection = 10,      // adaptive candidates never jump past a region
  kMutAdaptiveDispFull = 11,  // adaptive cost without the displacement divisor
  kMutIccadWeightInverted = 12,  // ICCAD17 weight count(h) / N
  kMutPinNoRotation = 13,     // pin boxes taken unrotated on a rotated row
  kMutEdgeWindowWide = 14,    // edge-spacing windows one site too wide
  kMutPinLayerSame = 15,      // access violation counted as a short (no layer + 1)
};
static int g_mut = kMutNone;

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int pow2_ceil(int n) {  // GlobalUtil.h:17-23
  int r = 1;
  while (r < n) r <<= 1;
  return r;
}

// ---------------------------------------------------------------------------
// Regions (objects/LGRegion.h) and the pixel grid (Grid.h / Grid.cpp)
// ---------------------------------------------------------------------------
enum RegionType { R_DEFAULT = 0, R_FENCE = 1, R_BLOCKAGE = 2 };

struct Region {
  int id, lx, ly, ux, uy, group, type;
  int width() const { return ux - lx; }
};

struct Grid {
  int lx, ly, w, h;  // core box (Grid.cpp:36-39)
  int sw, rh, x_size, y_size;
  std::vector<int> vdd_up;
  std::vector<uint8_t> valid;
  std::vector<int> h_ov, v_ov, group, region;

  size_t id(int x, int y) const { return (size_t)x * y_size + y; }
  int startX(int d) const { return (d - lx) / sw; }  // Grid.cpp:69-73
  int startY(int d) const { return (d - ly) / rh; }  // Grid.cpp:75-79
  int endX(int d) const {                            // Grid.cpp:81-88
    int quo = (d - lx) / sw, rem = (d - lx) % sw;
    int e = rem == 0 ? quo - 1 : quo;
    return std::max(std::min(x_size - 1, e), 0);
  }
  int endY(int d) const {  // Grid.cpp:90-97
    int quo = (d - ly) / rh, rem = (d - ly) % rh;
    int e = rem == 0 ? quo - 1 : quo;
    return std::max(std::min(y_size - 1, e), 0);
  }
  int dbuX(int gx) const { return gx * sw + lx; }  // Grid.cpp:125-129
  int dbuY(int gy) const { return gy * rh + ly; }  // Grid.cpp:131-135
  // Grid.cpp:99-107
  void expanded(int rlx, int rly, int rux, int ruy, int& x1, int& y1, int& x2, int& y2) const {
    x1 = startX(rlx); y1 = startY(rly); x2 = endX(rux); y2 = endY(ruy);
  }
  // Grid.cpp:109-123
  void shrunk(int rlx, int rly, int rux, int ruy, int& x1, int& y1, int& x2, int& y2) const {
    x1 = endX(rlx + sw); y1 = endY(rly + rh); x2 = startX(rux - sw); y2 = startY(ruy - rh);
  }
};

struct Cell {
  int lx, ly, w, h;  // dbu
  int init_lx, init_ly;
  int group, gnd, inst;
};

// The host side of one legalization: Legalizer's members (Legalizer.h:159-217)
struct Host {
  LgParams p;
  int core_lx, core_ly, core_ux, core_uy;
  int sw, rh;
  int dbu;
  int num_insts;
  std::vector<Cell> cells;
  std::vector<std::array<int, 4>> macro_rects, fixed_std_rects;  // lx ly ux uy
  std::vector<Region> regions;
  std::map<int, std::vector<int>> index2regions;  // group -> region ids
  std::map<int, int> height_distribution;         // height dbu -> count
  Grid grid;
  std::vector<int> dir_default, dir_wide;
  int num_preplace_fail = 0;
};

// Legalizer::makeDirectionVectors (Legalizer.cpp:249-354), default and wide sets
static void make_directions(Host& H) {
  H.dir_default.clear();
  for (int y = -6; y <= 6; y++) {
    static const int xr[7] = {40, 20, 10, 5, 2, 1, 0};
    int r = xr[std::abs(y)];
    for (int x = -r; x <= r; x++) { H.dir_default.push_back(x); H.dir_default.push_back(y); }
  }
  H.dir_wide.clear();
  for (int y = -12; y <= 12; y++) {
    static const int xr[13] = {280, 240, 200, 160, 120, 80, 40, 20, 10, 4, 2, 1, 0};
    int r = xr[std::abs(y)];
    for (int x = -r; x <= r; x++) { H.dir_wide.push_back(x); H.dir_wide.push_back(y); }
  }
}

// Grid::Grid / initNonGroupGrid / markDbRow (Grid.cpp:19-67, 316-342)
static void create_grid(Host& H, const LgDesign* d) {
  Grid& g = H.grid;
  g.lx = d->core_lx; g.ly = d->core_ly;
  g.w = d->core_ux - d->core_lx; g.h = d->core_uy - d->core_ly;
  int min_rh = INT_MAX;
  for (int r = 0; r < d->num_rows; r++) min_rh = std::min(min_rh, d->row_dy[r]);
  g.rh = min_rh;
  g.sw = d->row_site_width[d->num_rows - 1];
  g.x_size = g.w / g.sw;
  g.y_size = g.h / g.rh;
  size_t P = (size_t)g.x_size * g.y_size;
  g.vdd_up.assign(g.y_size, 0);
  g.valid.assign(P, 0);
  g.h_ov.assign(P, 0);
  g.v_ov.assign(P, 0);
  g.group.assign(P, kDefaultGroup);
  g.region.assign(P, -1);
  for (int r = 0; r < d->num_rows; r++) {
    int x1, y1, x2, y2;
    g.shrunk(d->row_lx[r], d->row_ly[r], d->row_lx[r] + d->row_dx[r], d->row_ly[r] + d->row_dy[r],
             x1, y1, x2, y2);
    if (y1 > y2 || x1 > x2) continue;
    for (int i = x1; i <= x2; i++)
      for (int j = y1; j <= y2; j++) g.valid[g.id(i, j)] = 1;
    int o = d->row_orient[r];
    g.vdd_up[y1] = (o == LG_ORIENT_N || o == LG_ORIENT_FN) ? 1 : 0;
  }
}

static void add_region(Host& H, int lx, int ly, int ux, int uy, int group, int type) {
  Region r{(int)H.regions.size(), lx, ly, ux, uy, group, type};
  H.regions.push_back(r);  // Legalizer::addNewRegion (DatabaseIO.cpp:653-659)
  H.index2regions[group].push_back(r.id);
}

// Legalizer::importCells (DatabaseIO.cpp:79-169)
// tokenize (Util.cpp:22-47): delimiters and white space separate tokens
static std::vector<std::string> tokenize(const std::string& line, const std::string& dels) {
  std::vector<std::string> out;
  std::string tok;
  for (char ch : line) {
    if (dels.find(ch) != std::string::npos || std::isspace((unsigned char)ch)) {
      if (!tok.empty()) { out.push_back(tok); tok.clear(); }
    } else {
      tok.push_back(ch);
    }
  }
  if (!tok.empty()) out.push_back(tok);
  return out;
}

// importCells (DatabaseIO.cpp:79-169) with parseSizeFile (Util.cpp:220-234)
static int import_cells(Host& H, const LgDesign* d) {
  int row_h = d->row_dy[d->num_rows - 1];
  int site_w = d->row_site_width[d->num_rows - 1];
  std::map<std::string, std::pair<int, int>> sizes;
  const bool modified = d->size_file != nullptr && d->size_file[0] != 0;
  if (modified) {
    std::ifstream in(d->size_file);
    if (!in) { g_err = "cannot open size file"; return LG_ERR_ARG; }
    std::string line;
    while (std::getline(in, line)) {
      auto t = tokenize(line, "=");
      if (t.size() >= 3) sizes[t[0]] = {std::stoi(t[1]), std::stoi(t[2])};
    }
  }
  for (int i = 0; i < d->num_insts; i++) {
    if (d->inst_dx[i] == 0 || d->inst_dy[i] < row_h) continue;
    int lx = d->inst_lx[i], ly = d->inst_ly[i];
    int ux = lx + d->inst_dx[i], uy = ly + d->inst_dy[i];
    if (d->inst_macro[i]) {
      if (!d->inst_fixed[i]) { g_err = "Cannot handle movable macro yet"; return LG_ERR_DESIGN; }
      H.macro_rects.push_back({lx, ly, ux, uy});
    } else if (!d->inst_fixed[i]) {
      Cell c;
      c.lx = c.init_lx = lx; c.ly = c.init_ly = ly;
      c.w = d->inst_dx[i]; c.h = d->inst_dy[i];
      if (modified && d->inst_name && d->inst_name[i]) {
        auto it = sizes.find(d->inst_name[i]);
        if (it != sizes.end()) { c.w = it->second.first * site_w; c.h = it->second.second * row_h; }
      }
      c.group = d->inst_group ? d->inst_group[i] : -1;  // assignGroupMembers (DatabaseIO.cpp:171-185)
      c.gnd = d->inst_ground_at_bottom[i] ? 1 : 0;
      c.inst = i;
      H.cells.push_back(c);
      H.height_distribution[c.h] += 1;
    } else {
      H.fixed_std_rects.push_back({lx, ly, ux, uy});
    }
  }
  return LG_OK;
}

// importBlockageRegions / importGroupRegions / importNonGroupRegions
// (DatabaseIO.cpp:210-379)
static void import_regions(Host& H, const LgDesign* d) {
  auto clip_add = [&](const std::array<int, 4>& r, int group, int type) {
    add_region(H, std::max(H.core_lx, r[0]), std::max(H.core_ly, r[1]), std::min(H.core_ux, r[2]),
               std::min(H.core_uy, r[3]), group, type);
  };
  for (auto& r : H.macro_rects) clip_add(r, kBlockageGroup, R_BLOCKAGE);
  for (auto& r : H.fixed_std_rects) clip_add(r, kBlockageGroup, R_BLOCKAGE);
  for (int b = 0; b < d->num_region_boxes; b++)
    clip_add({d->box_lx[b], d->box_ly[b], d->box_ux[b], d->box_uy[b]}, d->box_group[b], R_FENCE);

  // line sweep over fence + blockage rectangles (DatabaseIO.cpp:272-379)
  struct Event { int rect_id, y; bool is_enter; };
  std::vector<int> rects;
  for (auto& r : H.regions)
    if (r.type == R_FENCE || r.type == R_BLOCKAGE) rects.push_back(r.id);
  std::vector<Event> events;
  for (int k = 0; k < (int)rects.size(); k++) {
    events.push_back({k, H.regions[rects[k]].ly, true});
    events.push_back({k, H.regions[rects[k]].uy, false});
  }
  std::sort(events.begin(), events.end(), [](const Event& a, const Event& b) {
    if (a.y == b.y) return !a.is_enter && b.is_enter;
    return a.y < b.y;
  });
  std::set<int> active;
  std::unordered_map<int, std::vector<int>> lx_to_regions;
  auto create = [&](int lx, int ly, int ux, int uy) {
    if ((ux - lx) < H.sw || (uy - ly) < H.rh) return;
    int width = ux - lx;
    auto& same = lx_to_regions[lx];
    for (int rid : same) {
      Region& r = H.regions[rid];
      if (r.width() == width && r.uy == ly) { r.uy = uy; return; }  // merge (setHeight)
    }
    int id = (int)H.regions.size();
    add_region(H, lx, ly, ux, uy, kDefaultGroup, R_DEFAULT);
    same.push_back(id);
  };
  auto partial = [&](int ly, int uy) {
    std::vector<std::pair<int, int>> iv;
    for (int k : active) iv.push_back({H.regions[rects[k]].lx, H.regions[rects[k]].ux});
    std::sort(iv.begin(), iv.end());
    int lx = H.core_lx;
    for (auto& [a, b] : iv) { create(lx, ly, a, uy); lx = b; }
    create(lx, ly, H.core_ux, uy);
  };
  int ly = H.core_ly;
  for (auto& e : events) {
    partial(ly, e.y);
    if (e.is_enter) active.insert(e.rect_id); else active.erase(e.rect_id);
    ly = e.y;
  }
  create(H.core_lx, ly, H.core_ux, H.core_uy);
}

// markRegion + Grid::mark{Blockage,Group,NonGroup}Region + handleRegionSegments
// (DatabaseIO.cpp:196-208, 661-681; Grid.cpp:184-314)
static void mark_regions(Host& H) {
  Grid& g = H.grid;
  for (const Region& r : H.regions) {
    if (r.type == R_BLOCKAGE) {
      int x1, y1, x2, y2;
      g.expanded(r.lx, r.ly, r.ux, r.uy, x1, y1, x2, y2);
      x1 = std::max(0, x1); y1 = std::max(0, y1);
      x2 = std::min(g.x_size - 1, x2); y2 = std::min(g.y_size - 1, y2);
      for (int i = x1; i <= x2; i++)
        for (int j = y1; j <= y2; j++) {
          size_t p = g.id(i, j);
          g.valid[p] = 0; g.group[p] = r.group; g.region[p] = r.id;
        }
    } else if (r.type == R_FENCE) {
      int x1e, y1e, x2e, y2e, x1, y1, x2, y2;
      g.expanded(r.lx, r.ly, r.ux, r.uy, x1e, y1e, x2e, y2e);
      g.shrunk(r.lx, r.ly, r.ux, r.uy, x1, y1, x2, y2);
      for (int i = x1e; i <= x2e; i++)
        for (int j = y1e; j <= y2e; j++) {
          size_t p = g.id(i, j);
          g.group[p] = r.group; g.h_ov[p] += g.sw; g.v_ov[p] += g.rh;
          g.region[p] = r.id; g.valid[p] = 1;
        }
      if (x1e == x1 - 1) {
        int ov = x1 * g.sw - r.lx;
        for (int y = y1e; y <= y2e; y++) g.h_ov[g.id(x1e, y)] += ov - g.sw;
      }
      if (x2e == x2 + 1) {
        int ov = r.ux - x2e * g.sw;
        for (int y = y1e; y <= y2e; y++) g.h_ov[g.id(x2e, y)] += ov - g.sw;
      }
      if (y1e == y1 - 1) {
        int ov = y1 * g.rh - r.ly;
        for (int x = x1e; x <= x2e; x++) g.v_ov[g.id(x, y1e)] += ov - g.rh;
      }
      if (y2e == y2 + 1) {
        int ov = r.uy - y2e * g.rh;
        for (int x = x1e; x <= x2e; x++) g.v_ov[g.id(x, y2e)] += ov - g.rh;
      }
    } else {
      int x1, y1, x2, y2;
      g.shrunk(r.lx, r.ly, r.ux, r.uy, x1, y1, x2, y2);
      int a = std::min(g.x_size - 1, std::max(0, x1)), b = std::min(g.y_size - 1, std::max(0, y1));
      int c = std::min(g.x_size - 1, std::max(0, x2)), e = std::min(g.y_size - 1, std::max(0, y2));
      for (int i = a; i <= c; i++)
        for (int j = b; j <= e; j++) g.region[g.id(i, j)] = r.id;
    }
  }
  for (size_t p = 0; p < g.valid.size(); p++)
    if (g.h_ov[p] > 0 || g.v_ov[p] > 0)
      if (g.h_ov[p] < g.sw || g.v_ov[p] < g.rh) g.valid[p] = 0;
}

// Legalizer::importDB (DatabaseIO.cpp:31-77) + adjustYHint (Legalizer.cpp:230-247)
static int import_db(Host& H, const LgDesign* d, const LgParams* p) {
  if (!d || d->num_rows <= 0) { g_err = "design has no rows"; return LG_ERR_ARG; }
  H.p = *p;
  H.core_lx = d->core_lx; H.core_ly = d->core_ly; H.core_ux = d->core_ux; H.core_uy = d->core_uy;
  H.dbu = d->dbu_per_micron;
  H.num_insts = d->num_insts;
  int st = import_cells(H, d);
  if (st) return st;
  create_grid(H, d);
  H.sw = H.grid.sw; H.rh = H.grid.rh;
  import_regions(H, d);
  mark_regions(H);
  int max_h = 0;
  for (auto& [hh, n] : H.height_distribution) max_h = std::max(max_h, hh);
  int max_h_grid = max_h / H.rh;
  if (H.p.y_hint < 3 * max_h_grid) H.p.y_hint = 3 * max_h_grid;
  make_directions(H);
  return LG_OK;
}

// ---------------------------------------------------------------------------
// Preplace (Preplace.cpp)
// ---------------------------------------------------------------------------
enum PixStatus { PS_VALID = 0, PS_BLOCKAGE = 1, PS_FENCE = 2 };

// Grid::checkPixelStatus (Grid.cpp:365-383)
static int check_pixel_status(const Host& H, const Cell& c) {
  const Grid& g = H.grid;
  int x1, y1, x2, y2;
  g.expanded(c.lx, c.ly, c.lx + c.w, c.ly + c.h, x1, y1, x2, y2);
  for (int i = x1; i <= x2; i++)
    for (int j = y1; j <= y2; j++) {
      size_t p = g.id(i, j);
      if (!g.valid[p]) return PS_BLOCKAGE;
      if (g.group[p] != c.group) return PS_FENCE;
    }
  return PS_VALID;
}

// Legalizer::isOutOfRegion (Preplace.cpp:244-265)
static bool out_of_region(const Region& r, const Cell& c, int lx, int ly) {
  return !(lx >= r.lx && ly >= r.ly && lx + c.w <= r.ux && ly + c.h <= r.uy);
}

struct Proj { int lx, ly, disp; };

// Legalizer::findProjectionPoint (Preplace.cpp:267-353)
static Proj find_projection(const Host& H, const Cell& c, const Region& r) {
  const Grid& g = H.grid;
  int w_g = c.w / H.sw, h_g = c.h / H.rh;
  int gx1, gy1, gx2, gy2;
  g.shrunk(r.lx, r.ly, r.ux, r.uy, gx1, gy1, gx2, gy2);
  if (gy1 > gy2 || gx1 > gx2) return {-1, -1, kIntMax};
  int rlx = g.dbuX(gx1), rly = g.dbuY(gy1), rux = g.dbuX(gx2), ruy = g.dbuY(gy2);
  int nlx = c.lx, nly = c.ly;
  if (c.lx + c.w > rux) nlx = g.dbuX(gx2 - w_g + 1);
  else if (c.lx < rlx) nlx = g.dbuX(gx1);
  if (c.ly + c.h > ruy) nly = g.dbuY(gy2 - h_g + 1);
  else if (c.ly < rly) nly = g.dbuY(gy1);
  if (h_g % 2 == 0) {
    int ny = g.startY(nly);
    if (g.vdd_up[ny] != c.gnd) {
      int cand[2] = {nly - H.rh, nly + H.rh};
      int best = kIntMax, best_ly = -1;
      for (int yc : cand) {
        int dd = std::abs(nlx - c.init_lx) + std::abs(yc - c.init_ly);
        if (!out_of_region(r, c, nlx, yc) && dd < best) { best = dd; best_ly = yc; }
      }
      nly = best_ly;
    }
  }
  bool invalid = false;
  bool oor = out_of_region(r, c, nlx, nly);
  if (h_g % 2 == 0 && !oor) {
    int nx = g.startX(nlx), ny = g.startY(nly);
    for (int x = nx; x < nx + w_g; x++)
      for (int y = ny; y < ny + h_g; y++)
        if (!g.valid[g.id(x, y)]) invalid = true;
  }
  int disp = std::abs(c.init_lx - nlx) + std::abs(c.init_ly - nly);
  if (!oor && !invalid) return {nlx, nly, disp};
  return {-1, -1, kIntMax};
}

static bool project_to_group(const Host& H, Cell& c) {
  auto it = H.index2regions.find(c.group);
  if (it == H.index2regions.end() || it->second.empty()) return false;
  Proj best{-1, -1, kIntMax};
  bool first = true;
  for (int rid : it->second) {  // std::min_element keeps the first minimum
    Proj pr = find_projection(H, c, H.regions[rid]);
    if (first || pr.disp < best.disp) { best = pr; first = false; }
  }
  if (best.disp < kIntMax) { c.lx = best.lx; c.ly = best.ly; return true; }
  return false;
}

// Legalizer::roundToNearsetPoint (Preplace.cpp:111-157)
static void round_to_nearest(const Host& H, Cell& c) {
  const Grid& g = H.grid;
  int lx = c.lx, ly = c.ly;
  int w_g = c.w / H.sw, h_g = c.h / H.rh;
  int gx = std::min(g.x_size - w_g, std::max(0, g.startX(lx)));
  int x1 = g.dbuX(gx), x2 = g.dbuX(std::min(g.x_size - w_g, gx + 1));
  int nlx = (std::abs(lx - x1) < std::abs(lx - x2)) ? x1 : x2;
  int gy = std::min(g.y_size - h_g, std::max(0, g.startY(ly)));
  int y1 = g.dbuY(gy), y2 = g.dbuY(std::min(g.y_size - h_g, gy + 1));
  if (y1 == y2) y2 = g.dbuY(std::max(0, gy - 1));
  int nly = (std::abs(ly - y1) < std::abs(ly - y2)) ? y1 : y2;
  if (h_g % 2 == 0) {
    int ny = g.startY(nly);
    if (g.vdd_up[ny] != c.gnd) {
      if (nly == y1) nly = (y2 + c.h <= H.core_uy) ? y2 : y1 - H.rh;
      else if (nly == y2) nly = (y1 >= H.core_ly) ? y1 : y2 + H.rh;
    }
  }
  c.lx = nlx; c.ly = nly;
}

// Legalizer::checkHopeless (Preplace.cpp:159-207).  The reference loops over
// direction_default_.size() (twice the number of directions) and reads past the
// vector for j >= 245; only the 245 defined directions are tried here.
static bool check_hopeless(const Host& H, Cell& c) {
  const Grid& g = H.grid;
  int nd = (int)H.dir_default.size() / 2;
  int cx = g.startX(c.lx), cy = g.startY(c.ly);
  int save_x = c.lx, save_y = c.ly;
  bool hopeless = true;
  for (int j = 0; j < nd; j++) {
    int nx = cx + H.dir_default[2 * j], ny = cy + H.dir_default[2 * j + 1];
    c.lx = g.dbuX(nx); c.ly = g.dbuY(ny);
    if (c.lx + c.w > H.core_ux || c.ly + c.h > H.core_uy || c.lx < H.core_lx || c.ly < H.core_ly)
      continue;  // isOutOfCore (Preplace.cpp:229-242)
    if (g.vdd_up[ny] == c.gnd && check_pixel_status(H, c) == PS_VALID) { hopeless = false; break; }
  }
  c.lx = save_x; c.ly = save_y;
  return hopeless;
}

// Legalizer::initialLegalize (Preplace.cpp:18-72)
static void preplace(Host& H) {
  int n = (int)H.cells.size();
  int threads = std::max(1, H.p.num_cpu_threads);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int i = 0; i < n; i++) {
    Cell& c = H.cells[i];
    round_to_nearest(H, c);                          // preplaceCell (Preplace.cpp:74-83)
    if (check_pixel_status(H, c) != PS_VALID) project_to_group(H, c);
  }
  int fails = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : fails)
  for (int i = 0; i < n; i++) {
    Cell& c = H.cells[i];
    if (check_hopeless(H, c))
      if (!project_to_group(H, c)) fails++;
  }
  H.num_preplace_fail = fails;
}

// ---------------------------------------------------------------------------
// Device-layer restatement: CudaDatabase + CudaAdmmLegalizer + CudaPostProcess
// ---------------------------------------------------------------------------
struct Dev {
  LgParams p;
  int x_size, y_size, sw, rh, core_lx, core_ly;
  std::vector<int> valid, pgroup, pregion, vdd_up;
  std::vector<int> rxmin, rymin, rxmax, rymax;
  int N;
  std::vector<int> ox, oy, lx, ly, w, h, cgroup, gnd, need_rot;
  std::vector<int> usage, color;
  std::vector<float> dual;
  std::vector<int> k_height_to_num;  // [max_h]
  std::vector<int> dir_default, dir_wide;
  std::vector<int> random_priority;
  // partitioning
  int ncol, nrow, K;
  std::vector<int> part_of, grouped, offset, key;
  int max_in_part;
  // ADMM scalars
  float rho, dual_step, initial_dual;
  int min_ovf_so_far, last_slow;
  std::deque<int> ovf_deque;
  std::vector<int> trace;
  int iterations = 0, converged = 0, adaptive_calls = 0;
  int64_t final_ovf = 0;
  // ICCAD 2017 technology (CudaDatabase.h:182-213): power metal per pixel,
  // edge types and pin boxes per cell, the rule table, and the four pixel
  // edge maps (cells whose left / right edge of type 1 / 2 lies on the
  // pixel), live and as they stood when the current primal or refine step
  // began (`snap`, what a partition reads outside its own pixels)
  bool tech = false;
  std::vector<int> pmetal, ledge, redge, cmacro, mpin, pbgn, pend, play, rules;
  std::vector<int> eL1, eL2, eR1, eR2, sL1, sL2, sR1, sR2;
  size_t pid(int x, int y) const { return (size_t)x * y_size + y; }
};

static bool correct_row(int h, int gnd, int vdd) {  // DeviceUtil.hpp:11-22
  return !((h & 1) == 0 && gnd != vdd);
}

// willBeRotated (AdmmKernels.hpp:532-538): odd heights take the row's rail
static bool will_rotate(int h, int vdd, int gnd) { return (h % 2 == 0) ? false : (vdd != gnd); }

// changePixelEdgeType (AdmmKernels.hpp:428-465): the left column of the
// footprint carries the left edge type, the right column the right one
// (both, for a one-site cell)
static void edge_change(Dev& D, int x0, int y0, int w, int h, int lt, int rt, int delta) {
  for (int y = y0; y < y0 + h; y++) {
    if (lt == 1) D.eL1[D.pid(x0, y)] += delta;
    else if (lt == 2) D.eL2[D.pid(x0, y)] += delta;
    if (rt == 1) D.eR1[D.pid(x0 + w - 1, y)] += delta;
    else if (rt == 2) D.eR2[D.pid(x0 + w - 1, y)] += delta;
  }
}

// rip-up / commit of a cell's edges at (x, y) (doCommitOrRipup,
// AdmmKernels.hpp:467-506), its orientation from willBeRotated
static void edge_move(Dev& D, int c, int x, int y, int delta) {
  if (!D.tech) return;
  const bool rot = will_rotate(D.h[c], D.vdd_up[y], D.gnd[c]);
  edge_change(D, x, y, D.w[c], D.h[c], rot ? D.redge[c] : D.ledge[c], rot ? D.ledge[c] : D.redge[c], delta);
}

// updatePixelEdgeMapInGrid + computePixelEdgeMapKernel (CudaDatabase.cu:465-493,
// DatabaseKernels.hpp:105-164): the map of the current positions, the
// orientation taken as rotated whenever the row's rail differs from the
// cell's (for every height, as that kernel does)
static void edge_init(Dev& D) {
  if (!D.tech) return;
  const size_t P = (size_t)D.x_size * D.y_size;
  for (auto* v : {&D.eL1, &D.eL2, &D.eR1, &D.eR2}) v->assign(P, 0);
  for (int c = 0; c < D.N; c++) {
    const bool rot = D.vdd_up[D.ly[c]] != D.gnd[c];
    edge_change(D, D.lx[c], D.ly[c], D.w[c], D.h[c], rot ? D.redge[c] : D.ledge[c],
                rot ? D.ledge[c] : D.redge[c], +1);
  }
}

static void edge_snapshot(Dev& D) {
  if (!D.tech) return;
  D.sL1 = D.eL1; D.sL2 = D.eL2; D.sR1 = D.eR1; D.sR2 = D.eR2;
}

// computeNumPinVio (AdmmKernels.hpp:193-264): a pin box over a pixel whose
// power metal is the pin's layer is a short, one layer up an access
// violation; each counts once per cell
static void pin_vio_split(const Dev& D, int c, bool rot, int nx, int ny, int& shrt, int& acc) {
  shrt = acc = 0;
  if (D.cmacro.empty()) return;
  const int w = D.w[c], h = D.h[c], m = D.cmacro[c];
  for (int i = D.mpin[m]; i < D.mpin[m + 1]; i++) {
    int bx = D.pbgn[i] / h, by = D.pbgn[i] % h, ex = D.pend[i] / h, ey = D.pend[i] % h;
    if (rot && g_mut != kMutPinNoRotation) {
      const int ox = bx, oy = by;
      bx = w - 1 - ex; by = h - 1 - ey;
      ex = w - 1 - ox; ey = h - 1 - oy;
    }
    const int x1 = std::min(bx, ex) + nx, y1 = std::min(by, ey) + ny;
    const int x2 = std::max(bx, ex) + nx, y2 = std::max(by, ey) + ny;
    for (int x = x1; x <= x2; x++)
      for (int y = y1; y <= y2; y++) {
        const int l = D.pmetal[D.pid(x, y)];
        if (l > 0) {
          if (l == D.play[i]) shrt = 1;
          else if (l == D.play[i] + (g_mut == kMutPinLayerSame ? 0 : 1)) acc = 1;
        }
      }
  }
}

static int pin_vio(const Dev& D, int c, bool rot, int nx, int ny) {
  int s, a;
  pin_vio_split(D, c, rot, nx, ny, s, a);
  return s + a;
}

// the edge counts a cell sees beside a footprint at (nx, ny): reads of a
// pixel of partition `part` come from the live map, others from the snapshot
// (part < 0: everything live, as after the solve)
static int edge_count(const Dev& D, int part, const std::vector<int>& live, const std::vector<int>& snap, int x0,
                      int x1, int ny, int h) {
  int n = 0;
  if (g_mut == kMutEdgeWindowWide) x1++;
  for (int x = x0; x < x1; x++)
    for (int y = ny; y < ny + h; y++) {
      if (x < 0 || x >= D.x_size) continue;
      const size_t p = D.pid(x, y);
      n += (part < 0 || D.color[p] == part) ? live[p] : snap[p];
    }
  return n;
}

// computeEdgeSpacingCost (AdmmKernels.hpp:266-331) as counts: neighbours'
// right edges within the spacing left of the cell, left edges right of it
static void edge_sides(const Dev& D, int part, int c, bool rot, int nx, int ny, int& left, int& right) {
  const int lt = rot ? D.redge[c] : D.ledge[c], rt = rot ? D.ledge[c] : D.redge[c];
  const int w = D.w[c], h = D.h[c];
  left = right = 0;
  if (lt > 0) {
    const int s1 = D.rules[lt + 1], s2 = D.rules[lt + 2];
    left = edge_count(D, part, D.eR1, D.sR1, nx - s1, nx, ny, h) + edge_count(D, part, D.eR2, D.sR2, nx - s2, nx, ny, h);
  }
  if (rt > 0) {
    const int s1 = D.rules[rt + 1], s2 = D.rules[rt + 2];
    right = edge_count(D, part, D.eL1, D.sL1, nx + w, nx + w + s1, ny, h) +
            edge_count(D, part, D.eL2, D.sL2, nx + w, nx + w + s2, ny, h);
  }
}

static float edge_cost(const Dev& D, int part, int c, bool rot, int nx, int ny) {
  int l, r;
  edge_sides(D, part, c, rot, nx, ny, l, r);
  const float cl = float(l), cr = float(r);
  return D.p.edge_spacing_penalty * (cl + cr);
}

static void init_dev(Dev& D, const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p) {
  D.p = *p;
  D.x_size = g->x_size; D.y_size = g->y_size; D.sw = g->site_width; D.rh = g->row_height;
  D.core_lx = g->core_lx; D.core_ly = g->core_ly;
  size_t P = (size_t)D.x_size * D.y_size;
  D.valid.assign(g->pixel_valid, g->pixel_valid + P);
  D.pgroup.assign(g->pixel_group, g->pixel_group + P);
  if (g->pixel_region) D.pregion.assign(g->pixel_region, g->pixel_region + P);
  else D.pregion.assign(P, -1);
  D.vdd_up.assign(g->row_vdd_up, g->row_vdd_up + D.y_size);
  D.rxmin.assign(g->region_x_min, g->region_x_min + g->num_regions);
  D.rymin.assign(g->region_y_min, g->region_y_min + g->num_regions);
  D.rxmax.assign(g->region_x_max, g->region_x_max + g->num_regions);
  D.rymax.assign(g->region_y_max, g->region_y_max + g->num_regions);
  D.N = c->num_cells;
  D.ox.assign(c->orig_lx_dbu, c->orig_lx_dbu + D.N);
  D.oy.assign(c->orig_ly_dbu, c->orig_ly_dbu + D.N);
  D.w.assign(c->w_grid, c->w_grid + D.N);
  D.h.assign(c->h_grid, c->h_grid + D.N);
  D.cgroup.assign(c->group, c->group + D.N);
  D.gnd.assign(c->ground_at_bottom, c->ground_at_bottom + D.N);
  D.lx.resize(D.N); D.ly.resize(D.N);
  for (int i = 0; i < D.N; i++) {  // dbu_to_grid_functor (DatabaseKernels.hpp:7-20)
    D.lx[i] = (c->lx_dbu[i] - D.core_lx) / D.sw;
    D.ly[i] = (c->ly_dbu[i] - D.core_ly) / D.rh;
  }
  D.need_rot.assign(D.N, 0);
  int max_h = 0;
  for (int i = 0; i < D.N; i++) max_h = std::max(max_h, D.h[i]);
  D.k_height_to_num.assign(std::max(max_h, 1), 0);
  for (int i = 0; i < D.N; i++) D.k_height_to_num[D.h[i] - 1]++;
  Host tmp;
  make_directions(tmp);
  D.dir_default = tmp.dir_default;
  D.dir_wide = tmp.dir_wide;
  // CudaAdmmLegalizer ctor (CudaAdmmLegalizer.cu:86-130)
  D.ncol = D.x_size % p->x_hint == 0 ? D.x_size / p->x_hint : D.x_size / p->x_hint + 1;
  D.nrow = D.y_size % p->y_hint == 0 ? D.y_size / p->y_hint : D.y_size / p->y_hint + 1;
  D.K = D.ncol * D.nrow;
  D.random_priority.assign(D.N, 0);
  if (p->sorting_policy == LG_SORT_POLICY_RANDOM) {
    std::iota(D.random_priority.begin(), D.random_priority.end(), 0);
    std::mt19937 gen(p->random_sorting_seed);
    std::shuffle(D.random_priority.begin(), D.random_priority.end(), gen);
  }
  D.usage.assign(P, 0);
  D.dual.assign(P, 0.0f);
  D.color.assign(P, 0);
  // technology: absent parts are zero (no power metal, edge type 0, no pins)
  D.tech = g->pixel_power_layer || g->num_edge_rules > 0 || c->ledge_type || c->redge_type || c->macro_id;
  if (D.tech) {
    if (g->pixel_power_layer) D.pmetal.assign(g->pixel_power_layer, g->pixel_power_layer + P);
    else D.pmetal.assign(P, -1);
    // rule table read at indices up to 4 (types 1 and 2); entries beyond the
    // given table are 0
    D.rules.assign(5, 0);
    for (int i = 0; i < std::min(g->num_edge_rules, 5); i++) D.rules[i] = g->edge_spacing_rules[i];
    D.ledge.assign(D.N, 0);
    D.redge.assign(D.N, 0);
    if (c->ledge_type) D.ledge.assign(c->ledge_type, c->ledge_type + D.N);
    if (c->redge_type) D.redge.assign(c->redge_type, c->redge_type + D.N);
    if (c->macro_id && c->num_macros > 0) {
      D.cmacro.assign(c->macro_id, c->macro_id + D.N);
      D.mpin.assign(c->macro_pin_start, c->macro_pin_start + c->num_macros + 1);
      const int np = D.mpin[c->num_macros];
      D.pbgn.assign(c->pin_bgn, c->pin_bgn + np);
      D.pend.assign(c->pin_end, c->pin_end + np);
      D.play.assign(c->pin_layer, c->pin_layer + np);
    }
    for (auto* v : {&D.eL1, &D.eL2, &D.eR1, &D.eR2}) v->assign(P, 0);
  }
}

// computeUsageGridKernel (DatabaseKernels.hpp:69-103)
static void update_usage(Dev& D) {
  std::fill(D.usage.begin(), D.usage.end(), 0);
  for (int i = 0; i < D.N; i++) {
    int x0 = std::max(0, D.lx[i]), y0 = std::max(0, D.ly[i]);
    int x1 = std::min(D.x_size, D.lx[i] + D.w[i]), y1 = std::min(D.y_size, D.ly[i] + D.h[i]);
    for (int x = x0; x < x1; x++)
      for (int y = y0; y < y1; y++) D.usage[D.pid(x, y)]++;
  }
}

static int64_t compute_overflow(const Dev& D) {  // CudaDatabase.cu:562-572
  int64_t s = 0;
  const int64_t P = (int64_t)D.usage.size();
  const int threads = std::max(1, D.p.num_cpu_threads);
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : s)
  for (int64_t p = 0; p < P; p++) s += std::max(D.usage[p] - 1, 0);
  return s;
}

// computeDisplace (CudaDatabase.cu:580-603); exact int64 instead of an FP32 reduce
static int64_t compute_displace(const Dev& D) {
  int64_t s = 0;
  for (int i = 0; i < D.N; i++) {
    int dx = D.lx[i] * D.sw + D.core_lx, dy = D.ly[i] * D.rh + D.core_ly;
    s += std::abs(dx - D.ox[i]) + std::abs(dy - D.oy[i]);
  }
  return s;
}

// initializeRho + computeDemandKernel (CudaAdmmLegalizer.cu:321-350;
// AdmmKernels.hpp:44-81).  The demand values are integers; their sum is taken
// exactly (int64) and rounded once to float.
static void initialize_rho(Dev& D) {
  int64_t sum = 0;
  for (int i = 0; i < D.N; i++) {
    int64_t s = 0;
    bool off = false;
    for (int x = D.lx[i]; x < D.lx[i] + D.w[i] && !off; x++) {
      if (x < 0 || x >= D.x_size) { off = true; break; }
      for (int y = D.ly[i]; y < D.ly[i] + D.h[i]; y++) {
        if (y < 0 || y >= D.y_size) { off = true; break; }
        s += D.usage[D.pid(x, y)];
      }
    }
    if (!off) sum += s;  // early return leaves the zero-initialised slot
  }
  float demand_sum = (float)sum;
  D.rho = D.p.coeff_to_init_rho * demand_sum / float(D.N);
}

// Per-cell inverse augmented demand (AdmmKernels.hpp:83-120), FP32 in the
// reference's operation order.
static float inv_aug_demand(const Dev& D, int i, bool& off) {
  float expected = float(D.rh) / float(D.sw);
  float aug = 0.0f;
  off = false;
  for (int x = D.lx[i]; x < D.lx[i] + D.w[i]; x++) {
    if (x < 0 || x >= D.x_size) { off = true; return 0.0f; }
    for (int y = D.ly[i]; y < D.ly[i] + D.h[i]; y++) {
      if (y < 0 || y >= D.y_size) { off = true; return 0.0f; }
      float u = float(D.usage[D.pid(x, y)]);
      float t = D.rho * 0.5f;
      t = t * u;
      t = t * u;
      float a = 1.0f + u;
      aug += a + t;
    }
  }
  return expected / aug;
}

// initializeDual (CudaAdmmLegalizer.cu:352-395).  The per-cell values are
// summed in fixed point (round toward zero at 2^-32) so the sum does not depend
// on the reduction order; the result is rounded once to float.
static void initialize_dual(Dev& D) {
  int64_t fx = 0;
  for (int i = 0; i < D.N; i++) {
    bool off;
    float v = inv_aug_demand(D, i, off);
    if (off) continue;
    fx += (int64_t)(v * 4294967296.0f);
  }
  float sum = (float)((double)fx * (1.0 / 4294967296.0));
  float init = D.p.coeff_to_init_dual * sum / float(D.N);
  std::fill(D.dual.begin(), D.dual.end(), init);
  D.initial_dual = init;
  D.dual_step = init;
}

// updatePartition (CudaAdmmLegalizer.cu:166-228) with paintPixelsKernel and
// assignPartitionKernel (AdmmKernels.hpp:122-191)
static void update_partition(Dev& D, int iter) {
  int tf = (iter / D.p.iter_update_partition) % 3;
  int ox = D.p.x_hint / 3 * tf, oy = D.p.y_hint / 3 * tf;
  if (g_mut == kMutNoFoldOffset) ox = oy = 0;
  const int threads = std::max(1, D.p.num_cpu_threads);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int x = 0; x < D.x_size; x++)
    for (int y = 0; y < D.y_size; y++) {
      int col = ((x + ox) / D.p.x_hint) % D.ncol;
      int row = ((y + oy) / D.p.y_hint) % D.nrow;
      D.color[D.pid(x, y)] = col * D.nrow + row;
    }
  D.part_of.resize(D.N);
  for (int i = 0; i < D.N; i++) {
    int lxg = D.lx[i] + ox, lyg = D.ly[i] + oy;
    int uxg = lxg + D.w[i] - 1, uyg = lyg + D.h[i] - 1;
    int c1 = (lxg / D.p.x_hint) % D.ncol, c2 = (uxg / D.p.x_hint) % D.ncol;
    int r1 = (lyg / D.p.y_hint) % D.nrow, r2 = (uyg / D.p.y_hint) % D.nrow;
    D.part_of[i] = (c1 == c2 && r1 == r2) || g_mut == kMutNoStraddleCheck ? c1 * D.nrow + r1 : D.K;
  }
  // stable grouping by partition id (thrust::stable_sort_by_key on a sequence)
  std::vector<int> cnt(D.K + 1, 0);
  for (int i = 0; i < D.N; i++) cnt[D.part_of[i]]++;
  D.offset.assign(D.K + 2, D.N);
  int acc = 0;
  for (int k = 0; k <= D.K; k++) { D.offset[k] = acc; acc += cnt[k]; }
  D.offset[D.K + 1] = D.N;
  D.grouped.resize(D.N);
  std::vector<int> fill(D.K + 1, 0);
  for (int i = 0; i < D.N; i++) { int k = D.part_of[i]; D.grouped[D.offset[k] + fill[k]++] = i; }
  D.max_in_part = 0;
  for (int k = 0; k < D.K; k++) D.max_in_part = std::max(D.max_in_part, cnt[k]);
}

// bitonicSortDescending (AdmmKernels.hpp:899-945): the compare-exchange
// network on n = 2^m slots.  Pairs of one (k, j) step are disjoint, so a
// sequential sweep equals the parallel one.
void bitonic_desc(int n, int* key, int* val) {
  for (int k = 2; k <= n; k *= 2)
    for (int j = k / 2; j > 0; j /= 2)
      for (int i = 0; i < n; i++) {
        int ij = i ^ j;
        if (ij > i) {
          bool swap = ((i & k) == 0) ? (key[i] < key[ij]) : (key[i] > key[ij]);
          if (swap) { std::swap(key[i], key[ij]); std::swap(val[i], val[ij]); }
        }
      }
}

// mergeSortDescending + mergeSubroutine (SortingKernels.hpp:143-219)
static void merge_desc(int n, std::vector<int>& key, std::vector<int>& val) {
  std::vector<int> bk(key), bv(val);  // buffer initialised as the input
  for (int w = 1; w < n; w *= 2) {
    for (int j = 0; j < n / 2; j++) {
      int left = j * 2 * w;
      int mid = std::min(left + w, n), right = std::min(left + 2 * w, n);
      int a = left, b = mid, k = left;
      while (a < mid && b < right) {
        if (key[a] > key[b]) { bk[k] = key[a]; bv[k] = val[a]; a++; }
        else { bk[k] = key[b]; bv[k] = val[b]; b++; }
        k++;
      }
      while (a < mid) { bk[k] = key[a]; bv[k] = val[a]; k++; a++; }
      while (b < right) { bk[k] = key[b]; bv[k] = val[b]; k++; b++; }
    }
    key = bk; val = bv;
  }
}

static int cell_overflow(const Dev& D, int x0, int y0, int w, int h) {  // AdmmKernels.hpp:393-406
  int s = 0;
  for (int x = x0; x < x0 + w; x++)
    for (int y = y0; y < y0 + h; y++) s += std::max(D.usage[D.pid(x, y)] - 1, 0);
  return s;
}

// executeSorting (CudaAdmmLegalizer.cu:397-556, SortingKernels.hpp:9-55, 221-331)
static void execute_sorting(Dev& D) {
  int cr = 0, co = g_mut == kMutKeyOverflowWeight1 ? 1 : 10, ca = 1;
  if (D.p.sorting_policy == LG_SORT_POLICY_RANDOM) { cr = 1; co = 0; ca = 0; }
  D.key.resize(D.N);
  const int threads = std::max(1, D.p.num_cpu_threads);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int t = 0; t < D.N; t++) {
    int c = D.grouped[t];
    int ovf = cell_overflow(D, D.lx[c], D.ly[c], D.w[c], D.h[c]);
    D.key[t] = cr * D.random_priority[c] + co * ovf + ca * D.w[c] * D.h[c];
  }
  int n = pow2_ceil(D.max_in_part);
  int algo = D.p.sorting_algorithm;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 8)
  for (int k = 0; k < D.K; k++) {
    int off = D.offset[k], m = D.offset[k + 1] - off;
    if (algo == LG_SORT_CUB) {  // segmented radix sort: stable, descending
      std::vector<std::pair<int, int>> kv(m);
      for (int i = 0; i < m; i++) kv[i] = {D.key[off + i], D.grouped[off + i]};
      std::stable_sort(kv.begin(), kv.end(), [](auto& a, auto& b) { return a.first > b.first; });
      for (int i = 0; i < m; i++) { D.key[off + i] = kv[i].first; D.grouped[off + i] = kv[i].second; }
      continue;
    }
    std::vector<int> kk(n, kIntMin), vv(n, -1);
    for (int i = 0; i < m; i++) { kk[i] = D.key[off + i]; vv[i] = D.grouped[off + i]; }
    if (algo == LG_SORT_S_MERGE || algo == LG_SORT_M_MERGE) merge_desc(n, kk, vv);
    else bitonic_desc(n, kk.data(), vv.data());
    for (int i = 0; i < m; i++) { D.key[off + i] = kk[i]; D.grouped[off + i] = vv[i]; }
  }
}

// computeDispCost (AdmmKernels.hpp:508-530)
static inline float disp_cost(const Dev& D, int nx, int ny, int c) {
  float coeff = D.p.flag_iccad17 == 0 ? 1.0f
                : float(D.N) / float(D.k_height_to_num[D.h[c] - 1]);
  if (g_mut == kMutIccadWeightInverted && D.p.flag_iccad17)
    coeff = float(D.k_height_to_num[D.h[c] - 1]) / float(D.N);
  int dx = nx * D.sw + D.core_lx, dy = ny * D.rh + D.core_ly;
  int disp = std::abs(dx - D.ox[c]) + std::abs(dy - D.oy[c]);
  float ds = float(disp) / float(D.sw);
  float thr = float(D.p.max_disp_in_row_height) * float(g_mut == kMutThresholdInSites ? D.sw : D.rh);
  thr = thr / float(D.sw);
  float ex = ds - thr;
  float a = coeff * ds;
  float b = D.p.max_disp_coeff * std::max(ex, 0.0f);
  if (g_mut == kMutNoExcessTerm) b = 0.0f;
  return a + b;
}

// computePenaltySum (AdmmKernels.hpp:333-391)
static inline float penalty_sum(const Dev& D, bool standard, bool exclusive, bool final_refine,
                                int part, int nx, int ny, int c) {
  int w = D.w[c], h = D.h[c];
  if (nx + w > D.x_size || ny + h > D.y_size || nx < 0 || ny < 0) return kInf;
  if (!correct_row(h, D.gnd[c], D.vdd_up[ny])) return kInf;
  float s = 0.0f;
  for (int x = nx; x < nx + w; x++)
    for (int y = ny; y < ny + h; y++) {
      size_t p = D.pid(x, y);
      if (D.valid[p] == 0 || D.pgroup[p] != D.cgroup[c]) return kInf;
      if (part != D.color[p] && exclusive) return kInf;
      float u = float(D.usage[p]);
      float l = D.dual[p];
      if (final_refine && u > 0.0f) return kInf;
      float t;
      if (standard != (g_mut == kMutPenaltySwapped)) {
        float a = u - 1.0f;
        float b = l / D.rho;
        t = D.rho * std::max(a + b, 0.0f);
      } else {
        float a = u - 1.0f;
        float m = l * a;
        float b = l / D.rho;
        t = D.rho * std::max(m + b, 0.0f);
      }
      s += t;
    }
  return final_refine ? 0.0f : s;
}

// getNewCellPosition (AdmmKernels.hpp:540-560)
static inline void new_position(const Dev& D, const std::vector<int>& dir, int d, int c, int cx, int cy,
                                int& nx, int& ny) {
  int eps = (D.h[c] % 2 == 0 && g_mut != kMutEvenNoDoubleStep) ? 2 : 1;
  nx = cx + dir[2 * d];
  ny = cy + eps * dir[2 * d + 1];
  nx = std::max(0, std::min(D.x_size - D.w[c], nx));
  ny = std::max(0, std::min(D.y_size - D.h[c], ny));
}

// getNewCellPositionWithProjection (AdmmKernels.hpp:562-635)
static inline void new_position_projected(const Dev& D, int d, int c, int cx, int cy, int& nx, int& ny) {
  int dx = D.dir_wide[2 * d], dy = D.dir_wide[2 * d + 1];
  int eps = (D.h[c] % 2 == 0) ? 2 : 1;
  nx = cx + dx;
  ny = cy + eps * dy;
  int rid = -1;
  bool inval = false;
  for (int x = nx; x < nx + D.w[c] && !inval; x++)
    for (int y = ny; y < ny + D.h[c]; y++) {
      if (x >= D.x_size || y >= D.y_size || x < 0 || y < 0) { inval = true; break; }
      if (D.valid[D.pid(x, y)] == 0) { inval = true; rid = D.pregion[D.pid(x, y)]; break; }
    }
  if (rid != -1 && g_mut != kMutNoProjection) {
    if (std::abs(dx) >= std::abs(dy)) nx = dx > 0 ? D.rxmax[rid] + 1 : D.rxmin[rid] - 1;
    else ny = dy > 0 ? D.rymax[rid] + 1 : D.rymin[rid] - 1;
  }
  nx = std::max(0, std::min(D.x_size - D.w[c], nx));
  ny = std::max(0, std::min(D.y_size - D.h[c], ny));
}

// Tree-min of doPrimalUpdateKernel (AdmmKernels.hpp:837-862): halving stride,
// a slot takes its partner only if strictly smaller.
int tree_argmin(float* cost, int n_padded) {
  std::vector<int> idx(n_padded);
  std::iota(idx.begin(), idx.end(), 0);
  for (int s = n_padded / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; t++)
      if (cost[t] > cost[t + s] || (g_mut == kMutTieKeepsUpper && cost[t] == cost[t + s])) {
        cost[t] = cost[t + s];
        idx[t] = idx[t + s];
      }
  return idx[0];
}

static void change_usage(Dev& D, int x0, int y0, int w, int h, int delta) {
  for (int i = 0; i < w * h; i++) D.usage[D.pid(x0 + i / h, y0 + i % h)] += delta;
}

// One cell of doPrimalUpdateKernel / doFinalRefineKernel (AdmmKernels.hpp:710-896,
// 1246-1415).  Returns after committing the cell (possibly in place).
static void move_cell(Dev& D, int part, int c, bool final_refine, std::vector<float>& buf,
                      float* costs_out = nullptr) {
  const int nd = (int)D.dir_default.size() / 2;
  const int np = pow2_ceil(nd);
  int cx = D.lx[c], cy = D.ly[c];
  change_usage(D, cx, cy, D.w[c], D.h[c], -1);  // rip-up
  edge_move(D, c, cx, cy, -1);
  bool standard = final_refine ? false : (D.p.standard_admm != 0);
  // tech_penalty * row_height_in_grid (AdmmKernels.hpp:814): float * int
  const float pin_unit = D.p.tech_penalty * float(D.rh / D.sw);
  for (int d = 0; d < np; d++) {
    if (d >= nd) { buf[d] = kInf; continue; }
    int nx, ny;
    new_position(D, D.dir_default, d, c, cx, cy, nx, ny);
    float dc = disp_cost(D, nx, ny, c);
    float pc = penalty_sum(D, standard, true, final_refine, part, nx, ny, c);
    // pin-access and edge-spacing terms (AdmmKernels.hpp:803-830); zero
    // without technology
    float pin = 0.0f;
    float edge = (pc == kInf) ? kInf : 0.0f;
    if (D.tech && pc != kInf) {
      const bool rot = will_rotate(D.h[c], D.vdd_up[ny], D.gnd[c]);
      pin = pin_unit * float(pin_vio(D, c, rot, nx, ny));
      edge = edge_cost(D, part, c, rot, nx, ny);
    }
    buf[d] = dc + pc + pin + edge;
  }
  if (costs_out) std::copy(buf.begin(), buf.begin() + np, costs_out);
  int best = tree_argmin(buf.data(), np);
  int nx = cx, ny = cy;
  if (buf[0] != kInf) new_position(D, D.dir_default, best, c, cx, cy, nx, ny);
  change_usage(D, nx, ny, D.w[c], D.h[c], +1);  // commit
  edge_move(D, c, nx, ny, +1);
  D.lx[c] = nx;
  D.ly[c] = ny;
}

// doPrimalUpdateKernel (AdmmKernels.hpp:637-897), one partition
static void primal_partition(Dev& D, int k, bool perturbation, std::vector<float>& buf) {
  for (int t = D.offset[k]; t < D.offset[k + 1]; t++) {
    int c = D.grouped[t];
    int cx = D.lx[c], cy = D.ly[c];
    bool ok_row = correct_row(D.h[c], D.gnd[c], D.vdd_up[cy]);
    int ovf = cell_overflow(D, cx, cy, D.w[c], D.h[c]);
    if (!perturbation && ovf == 0 && ok_row) continue;
    move_cell(D, k, c, false, buf);
  }
}

// doFinalRefineKernel (AdmmKernels.hpp:1176-1416), one partition
static void final_refine_partition(Dev& D, int k, std::vector<float>& buf) {
  for (int t = D.offset[k]; t < D.offset[k + 1]; t++) {
    int c = D.grouped[t];
    if (D.h[c] <= 1) continue;
    move_cell(D, k, c, true, buf);
  }
}

// dual_update_functor_with_step_size (AdmmKernels.hpp:26-42): the literal 0.5
// promotes the quadratic term to double before the store rounds it to float.
static void dual_update(Dev& D) {
  const float rho = D.rho, step = D.dual_step;
  const float neg = -1.0f / rho;
  const int64_t P = (int64_t)D.dual.size();
  const int threads = std::max(1, D.p.num_cpu_threads);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t p = 0; p < P; p++) {
    float u = float(D.usage[p]);
    float g = std::max(u - 1.0f, neg);
    double q = (double)rho * (g_mut == kMutDualNoHalf ? 1.0 : 0.5);
    q = q * (double)g;
    q = q * (double)g;
    float delta = (float)((double)g + q);
    float m = step * delta;
    float l = D.dual[p] + m;
    D.dual[p] = std::max(l, 0.0f);
  }
}

// One cell of projectionAwareSearchKernel (AdmmKernels.hpp:1029-1172): skip
// without overflow, rip up, 4093 wide candidates with projection, cost =
// displacement / coeff_to_neglect_displace + negotiation penalty (no
// partition restriction), tree-min, commit.
static void adaptive_move(Dev& D, int c, std::vector<float>& buf, float* costs_out = nullptr) {
  const int nd = (int)D.dir_wide.size() / 2;
  const int np = pow2_ceil(nd);
  int cx = D.lx[c], cy = D.ly[c];
  if (cell_overflow(D, cx, cy, D.w[c], D.h[c]) == 0) return;
  change_usage(D, cx, cy, D.w[c], D.h[c], -1);
  edge_move(D, c, cx, cy, -1);  // edges follow the move; no pin / edge cost here
  for (int d = 0; d < np; d++) {
    if (d >= nd) { buf[d] = kInf; continue; }
    int nx, ny;
    new_position_projected(D, d, c, cx, cy, nx, ny);
    float dc = disp_cost(D, nx, ny, c);
    float pc = penalty_sum(D, false, false, false, -1, nx, ny, c);
    float q = g_mut == kMutAdaptiveDispFull ? dc : dc / D.p.coeff_to_neglect_displace;
    buf[d] = q + pc;
  }
  if (costs_out) std::copy(buf.begin(), buf.begin() + np, costs_out);
  int best = tree_argmin(buf.data(), np);
  int nx = cx, ny = cy;
  if (buf[0] != kInf) new_position_projected(D, best, c, cx, cy, nx, ny);
  change_usage(D, nx, ny, D.w[c], D.h[c], +1);
  edge_move(D, c, nx, ny, +1);
  D.lx[c] = nx;
  D.ly[c] = ny;
}

// projectionAwareSearchKernel (AdmmKernels.hpp:974-1174) after
// extractCongestedCellsKernel + streamCompaction (CudaAdmmLegalizer.cu:878-993):
// the overflowing cells in id order, one after another
static void adaptive_search(Dev& D) {
  std::vector<int> ovf_cells;
  for (int c = 0; c < D.N; c++)
    if (cell_overflow(D, D.lx[c], D.ly[c], D.w[c], D.h[c]) != 0) ovf_cells.push_back(c);
  std::vector<float> buf(pow2_ceil((int)D.dir_wide.size() / 2));
  for (int c : ovf_cells) adaptive_move(D, c, buf);
}

// detectSlowConvergence (CudaAdmmLegalizer.cu:799-840)
static bool detect_slow(Dev& D, int iter, int ovf) {
  if (iter < D.p.min_iter_to_call_adaptive_search) return false;
  if (D.min_ovf_so_far == 0) return false;
  if (D.last_slow > 0 && iter < D.last_slow + D.p.cooltime_to_call_adaptive_search) return false;
  bool by_ratio = float(ovf) / float(D.min_ovf_so_far) >= D.p.ovf_ratio_to_call_adaptive_search;
  bool by_stag = true;
  if ((int)D.ovf_deque.size() == D.p.stagnation_window) {
    int f = D.ovf_deque.front();
    for (int v : D.ovf_deque)
      if (v != f) { by_stag = false; break; }
  }
  bool det = by_ratio || by_stag;
  if (det) D.last_slow = iter;
  return det;
}

// CudaAdmmLegalizer::solve + applyAdmm (CudaAdmmLegalizer.cu:250-319)
static void solve(Dev& D) {
  D.min_ovf_so_far = kIntMax;
  D.last_slow = -1;
  std::fill(D.dual.begin(), D.dual.end(), 0.0f);
  update_usage(D);
  edge_init(D);  // CudaAdmmLegalizer.cu:270
  initialize_rho(D);
  initialize_dual(D);
  D.trace.clear();
  D.ovf_deque.clear();
  bool exceptional = false;
  int threads = std::max(1, D.p.num_cpu_threads);
  for (int iter = 0; iter < D.p.max_admm_iter; iter++) {
    if (iter % D.p.iter_update_partition == 0) update_partition(D, iter);
    if (iter % D.p.iter_update_dual_step == 0 && iter > D.p.iter_threshold_dual_update)
      D.dual_step += D.p.coeff_to_update_dual_step * D.initial_dual;
    if (!exceptional || !D.p.use_adaptive_search) {
      execute_sorting(D);
      // The reference's partitions read their neighbours' edge maps while
      // those move cells (an unsynchronised race between thread blocks);
      // here a partition reads the pixels outside its own as they stood when
      // the step began, one of the outcomes that race allows, and the same
      // for every schedule.
      edge_snapshot(D);
      bool pert = iter < D.p.iter_perturbation_on;
#pragma omp parallel num_threads(threads)
      {
        std::vector<float> buf(256);
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < D.K; k++) primal_partition(D, k, pert, buf);
      }
    } else {
      adaptive_search(D);
      D.adaptive_calls++;
    }
    dual_update(D);
    int64_t ovf = compute_overflow(D);
    D.min_ovf_so_far = (int)std::min<int64_t>(ovf, D.min_ovf_so_far);
    D.trace.push_back((int)ovf);
    D.ovf_deque.push_back((int)ovf);
    if ((int)D.ovf_deque.size() > D.p.stagnation_window) D.ovf_deque.pop_front();
    exceptional = detect_slow(D, iter, (int)ovf);
    D.iterations = iter + 1;
    D.final_ovf = ovf;
    bool conv = ovf <= D.p.admm_terminate_thr;
    if (D.p.verbose && (conv || iter % D.p.log_freq == 0)) {
      int64_t disp = compute_displace(D);
      printf("Iter: %5d Ovf: %7lld Disp: %10lld (%5.2f sites) MinOvf: %3d\n", iter, (long long)ovf,
             (long long)disp, double(disp) / D.N / D.sw, D.min_ovf_so_far);
    }
    if (conv) {
      D.converged = 1;
      edge_snapshot(D);
#pragma omp parallel num_threads(threads)
      {
        std::vector<float> buf(256);
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < D.K; k++) final_refine_partition(D, k, buf);
      }
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// Post-process (post_process/): rotation and the QP refine, QP in FP64
// ---------------------------------------------------------------------------
static void rotate_cells(Dev& D) {  // rotateKernel (PostProcessKernel.hpp:6-26)
  for (int c = 0; c < D.N; c++)
    if ((D.h[c] & 1) == 1) D.need_rot[c] = (D.gnd[c] == D.vdd_up[D.ly[c]]) ? 0 : 1;
}

// bitonicSortAscending (PostProcessKernel.hpp:215-261)
void bitonic_asc(int n, int* key, int* val) {
  for (int k = 2; k <= n; k *= 2)
    for (int j = k / 2; j > 0; j /= 2)
      for (int i = 0; i < n; i++) {
        int ij = i ^ j;
        if (ij < i) {
          bool swap = ((i & k) == 0) ? (key[i] < key[ij]) : (key[i] > key[ij]);
          if (swap) { std::swap(key[i], key[ij]); std::swap(val[i], val[ij]); }
        }
      }
}

// Thomas algorithm (QpSolverKernels.hpp:68-94) in FP64
static void thomas(int n, double a, double b, const double* d, const double* cp, double* dp, double* x) {
  for (int i = 0; i < n; i++) dp[i] = (i == 0) ? d[i] / b : (d[i] - a * dp[i - 1]) / (b - a * cp[i - 1]);
  for (int i = n - 1; i >= 0; i--) x[i] = (i == n - 1) ? dp[i] : dp[i] - cp[i] * x[i + 1];
}

struct CutRows {
  int num = 0;
  std::vector<int> y, xb, xe, offset, ncell, cells;
};

// stampCells / makeCutRows / insertCellsToCutRows / sortCellsEachCutRow
// (CudaPostProcess.cu:94-301; PostProcessKernel.hpp:28-312)
static void make_cut_rows(const Dev& D, CutRows& R) {
  size_t P = (size_t)D.x_size * D.y_size;
  std::vector<int> stamp(P, -1);
  for (int c = 0; c < D.N; c++)
    for (int x = D.lx[c]; x < D.lx[c] + D.w[c]; x++)
      for (int y = D.ly[c]; y < D.ly[c] + D.h[c]; y++) stamp[D.pid(x, y)] = c;
  auto pix = [&](int x, int y, int& g, int& v, int& hh) {
    size_t p = D.pid(x, y);
    g = D.pgroup[p]; v = D.valid[p];
    hh = stamp[p] >= 0 ? D.h[stamp[p]] : 0;
  };
  std::vector<int> count(D.y_size, 0);
  for (int y = 0; y < D.y_size; y++) {  // countCutRowsKernel
    int pg = -999, pv = -999, ps = -999;
    for (int x = 0; x < D.x_size; x++) {
      int g, v, hh;
      pix(x, y, g, v, hh);
      int single = hh <= 1 ? 1 : 0;
      bool ends = (pg != g) || (pv != v) || (hh > 1);
      if (ends && pv == 1 && ps == 1) count[y]++;
      if (x == D.x_size - 1 && !ends && v == 1 && hh <= 1) count[y]++;
      pg = g; pv = v; ps = single;
    }
  }
  std::vector<int> yoff(D.y_size + 1, 0);
  for (int y = 0; y < D.y_size; y++) yoff[y + 1] = yoff[y] + count[y];
  R.num = yoff[D.y_size];
  R.y.assign(R.num, 0); R.xb.assign(R.num, 0); R.xe.assign(R.num, 0);
  for (int y = 0; y < D.y_size; y++) {  // generateCutRowsKernel
    int made = 0, pg = -999, pv = -999, ps = -999, xb = 0;
    auto create = [&](int b, int e) {
      int id = yoff[y] + made++;
      if (id < R.num) { R.y[id] = y; R.xb[id] = b; R.xe[id] = e; }
    };
    for (int x = 0; x < D.x_size; x++) {
      int g, v, hh;
      pix(x, y, g, v, hh);
      int single = hh <= 1 ? 1 : 0;
      bool ends = (pg != g) || (pv != v) || (ps != single);
      if (ends) {
        if (pv == 1 && ps == 1) create(xb, x - 1);
        xb = x;
      }
      if (x == D.x_size - 1 && !ends && v == 1 && hh <= 1) create(xb, D.x_size - 1);
      pg = g; pv = v; ps = single;
    }
  }
  std::vector<int> pix_row(P, -1);  // markCutRowToGridKernel; higher id wins on overlap
  for (int r = 0; r < R.num; r++)
    for (int x = R.xb[r]; x <= R.xe[r]; x++) pix_row[D.pid(x, R.y[r])] = r;
  std::vector<int> row_of(D.N);
  for (int c = 0; c < D.N; c++) {  // markCutRowToCellKernel
    int r = pix_row[D.pid(D.lx[c], D.ly[c])];
    row_of[c] = (D.h[c] == 1 && r >= 0) ? r : R.num;
  }
  std::vector<int> cnt(R.num + 1, 0);
  for (int c = 0; c < D.N; c++) cnt[row_of[c]]++;
  R.offset.assign(R.num + 2, 0);
  for (int r = 0; r <= R.num; r++) R.offset[r + 1] = R.offset[r] + cnt[r];
  R.ncell.resize(R.num);
  for (int r = 0; r < R.num; r++) R.ncell[r] = cnt[r];
  R.cells.resize(D.N);
  std::vector<int> fill(R.num + 1, 0);
  for (int c = 0; c < D.N; c++) R.cells[R.offset[row_of[c]] + fill[row_of[c]]++] = c;
  int mx = 0;
  for (int r = 0; r < R.num; r++) mx = std::max(mx, R.ncell[r]);
  int n = pow2_ceil(mx);
  for (int r = 0; r < R.num; r++) {  // sortCellsEachCutRowKernel
    std::vector<int> k(n, kIntMax), v(n, -1);
    for (int i = 0; i < R.ncell[r]; i++) { int c = R.cells[R.offset[r] + i]; k[i] = D.lx[c]; v[i] = c; }
    bitonic_asc(n, k.data(), v.data());
    for (int i = 0; i < R.ncell[r]; i++) R.cells[R.offset[r] + i] = v[i];
  }
}

// generateConstraintsKernel (QpSolverKernels.hpp): one constraint per cell
// boundary of every cut row (n + 1 per row, constraint id off + r + i), reading
// x[pos] - x[neg] >= bound.  wd[l] is cell l's width in DBU; rlx/rux the row's
// DBU extent.
static void qp_constraints(int nr, const int* ncell, const int* offset, const int* wd, const int* rlx,
                           const int* rux, int* pos_c, int* neg_c, int* c_pos, int* c_neg, double* bound) {
  for (int r = 0; r < nr; r++) {
    int n = ncell[r], off = offset[r];
    for (int i = 0; i < n + 1; i++) {
      int cid = off + r + i;
      int lid = (i == n) ? -1 : i + off;
      int plid = (i == 0) ? -1 : i + off - 1;
      int pw = (i == 0) ? 0 : wd[plid];
      if (lid != -1) pos_c[lid] = cid;
      if (plid != -1) neg_c[plid] = cid;
      c_pos[cid] = lid;
      c_neg[cid] = plid;
      bound[cid] = (n == 0) ? 0.0 : (i == 0) ? double(rlx[r]) : (i == n) ? double(pw - rux[r]) : double(pw);
    }
  }
}

// CudaQpSolver::solve (CudaQpSolver.cu; QpSolverKernels.hpp:96-300): ADMM on
// min sum (x - x0)^2 subject to the cut-row constraints; the x-update is one
// tridiagonal (Thomas) solve per row.  Stops once no constraint is violated by
// more than 1 DBU and min_iter iterations have run.  Returns the iteration count.
static int qp_admm(int nr, const int* ncell, const int* offset, int nm, const double* x0, const int* pos_c,
                   const int* neg_c, const int* c_pos, const int* c_neg, const double* bound, double rho,
                   int min_iter, int max_iter, double* x) {
  const int nc = nm + nr;
  std::vector<double> pp(nc, 0.0), slack(nc, 0.0), dualv(nc, 0.0);
  std::vector<double> cp(nm), din(nm), dp(nm);
  for (int r = 0; r < nr; r++)  // precomputeCPrimeKernel
    for (int i = 0; i < ncell[r]; i++) {
      int l = offset[r] + i;
      cp[l] = (i == 0) ? -rho / (1 + 2 * rho) : -rho / (1 + 2 * rho + rho * cp[l - 1]);
    }
  int it_done = 0;
  for (int it = 0; it < max_iter; it++) {
    for (int l = 0; l < nm; l++) {  // updatePrimalXKernel
      double v = x0[l];
      int a = pos_c[l], b = neg_c[l];
      v -= rho * (slack[a] - bound[a] + dualv[a] / rho);
      v += rho * (slack[b] - bound[b] + dualv[b] / rho);
      din[l] = v;
    }
    for (int r = 0; r < nr; r++) {  // runThomasKernel
      int off = offset[r];
      thomas(ncell[r], -rho, 1 + 2 * rho, din.data() + off, cp.data() + off, dp.data() + off, x + off);
    }
    for (int k = 0; k < nc; k++) {  // computeConstraintValueKernel
      int a = c_pos[k], b = c_neg[k];
      pp[k] = (a == -1 && b == -1) ? 0.0 : (a == -1) ? -x[b] : (b == -1) ? x[a] : x[a] - x[b];
    }
    for (int k = 0; k < nc; k++) slack[k] = std::min(0.0, -pp[k] + bound[k] - dualv[k] / rho);
    for (int k = 0; k < nc; k++) dualv[k] = dualv[k] + rho * (pp[k] + slack[k] - bound[k]);
    double mn = std::numeric_limits<double>::infinity();
    for (int k = 0; k < nc; k++) mn = std::min(mn, pp[k] - bound[k]);
    it_done = it + 1;
    if (!(mn < -1 || it < min_iter)) break;
  }
  return it_done;
}

// siteAlignDPKernel (QpSolverKernels.hpp:300-475) for one cut row: every cell
// goes to the site at or just right of its QP position (clamped to the row's
// left end); a two-state DP over "left site" / "right site" minimises the summed
// |site - x0| with no overlap and the right site inside the row; ties go left.
// site[i] receives the chosen site index of cell i.
static void site_align_dp_row(int n, const double* x, const double* x0, const int* wdv, int rlx, int rux, int sw,
                              int core_lx, int* bl, int* br, int* site) {
  const double INF = std::numeric_limits<double>::infinity();
  double cl = 0, cr = 0;
  int uxl = rlx, uxr = rlx;
  for (int i = 0; i < n; i++) {
    int wd = wdv[i];
    int orig = (int)x0[i];
    int clamp = std::max(rlx, (int)x[i]);
    int dist = clamp - core_lx;
    int nl = (dist / sw) * sw + core_lx, nrt = (dist / sw + 1) * sw + core_lx;
    double dl = std::abs(nl - orig), dr = std::abs(nrt - orig);
    double ml, mr;
    if (i == 0) {
      ml = dl + ((nl < rlx) ? INF : 0.0);
      mr = dr + ((nrt < rlx || nrt + wd > rux) ? INF : 0.0);
      bl[i] = -1; br[i] = -1;
    } else {
      double a1 = cl + dl + ((nl < uxl) ? INF : 0.0);
      double a2 = cr + dl + ((nl < uxr) ? INF : 0.0);
      if (a1 <= a2) { ml = a1; bl[i] = 0; } else { ml = a2; bl[i] = 1; }
      double bp = (nrt + wd > rux) ? INF : 0.0;
      double b1 = cl + dr + (((nrt < uxl) ? INF : 0.0) + bp);
      double b2 = cr + dr + (((nrt < uxr) ? INF : 0.0) + bp);
      if (b1 <= b2) { mr = b1; br[i] = 0; } else { mr = b2; br[i] = 1; }
    }
    cl = ml; cr = mr;
    uxl = nl + wd; uxr = nrt + wd;
  }
  int st = (cl <= cr) ? 0 : 1;
  for (int i = n - 1; i >= 0; i--) {
    int clamp = std::max(rlx, (int)x[i]);
    int dist = clamp - core_lx;
    site[i] = (st == 0) ? dist / sw : dist / sw + 1;
    st = (st == 0) ? bl[i] : br[i];
  }
}

// CudaQpSolver::initializeSolver / solve / siteAlignDP (CudaQpSolver.cu:22-353;
// QpSolverKernels.hpp:96-475), FP64 throughout
static int qp_refine(Dev& D) {
  CutRows R;
  make_cut_rows(D, R);
  const int nr = R.num;
  int nm = 0;
  for (int r = 0; r < nr; r++) nm += R.ncell[r];
  const int nc = nm + nr;
  const double rho = D.p.qp_admm_rho;
  std::vector<int> l2g(nm), pos_c(nm), neg_c(nm), c_pos(nc), c_neg(nc), wd(nm), rlx(nr), rux(nr);
  std::vector<double> x(nm), x0(nm), bound(nc);
  for (int l = 0; l < nm; l++) {  // assignLocalIndexKernel
    int g = R.cells[l];
    l2g[l] = g;
    x[l] = double(D.core_lx + D.sw * D.lx[g]);
    x0[l] = double(D.ox[g]);
    wd[l] = D.sw * D.w[g];
  }
  for (int r = 0; r < nr; r++) {
    rlx[r] = D.sw * R.xb[r] + D.core_lx;
    rux[r] = D.sw * (R.xe[r] + 1) + D.core_lx;
  }
  qp_constraints(nr, R.ncell.data(), R.offset.data(), wd.data(), rlx.data(), rux.data(), pos_c.data(),
                 neg_c.data(), c_pos.data(), c_neg.data(), bound.data());
  const int it_done = qp_admm(nr, R.ncell.data(), R.offset.data(), nm, x0.data(), pos_c.data(), neg_c.data(), c_pos.data(),
          c_neg.data(), bound.data(), rho, D.p.qp_admm_min_iter, D.p.qp_admm_max_iter, x.data());
  std::vector<int> lx_new(D.lx);
  if (D.p.qp_admm_dp_site_align) {
    std::vector<int> bl(nm), br(nm), site(nm);
    for (int r = 0; r < nr; r++) {  // siteAlignDPKernel
      int n = R.ncell[r];
      if (n == 0) continue;
      int rlx = D.sw * R.xb[r] + D.core_lx, rux = D.sw * (R.xe[r] + 1) + D.core_lx;
      int off = R.offset[r];
      site_align_dp_row(n, x.data() + off, x0.data() + off, wd.data() + off, rlx, rux, D.sw, D.core_lx,
                        bl.data() + off, br.data() + off, site.data() + off);
      for (int i = n - 1; i >= 0; i--) lx_new[l2g[i + off]] = site[i + off];
    }
  } else {
    for (int r = 0; r < nr; r++) {  // siteAlignNaiveKernel
      int n = R.ncell[r], off = R.offset[r];
      int rlx = D.sw * R.xb[r] + D.core_lx, rux = D.sw * (R.xe[r] + 1) + D.core_lx;
      int prev_ux = rlx;
      for (int i = 0; i < n; i++) {
        int l = i + off, g = l2g[l];
        int wd = D.w[g] * D.sw;
        int orig = (int)x0[l];
        int clamp = std::max(rlx, (int)x[l]);
        int dist = clamp - D.core_lx;
        int nl = (dist / D.sw) * D.sw + D.core_lx, nrt = (dist / D.sw + 1) * D.sw + D.core_lx;
        double dl = std::abs(nl - orig), dr = std::abs(nrt - orig);
        int nx = nl;
        if (!(dl <= dr)) {
          if (i < n - 1) {
            int nd = (int)x[l + 1] - D.core_lx;
            int nll = (nd / D.sw) * D.sw + D.core_lx;
            if (nll >= nrt + wd) nx = nrt;
          }
          if (i == n - 1 && nrt + wd <= rux) nx = nrt;
        }
        if (prev_ux > nx) nx = nrt;
        prev_ux = nx + wd;
        lx_new[g] = (nx - D.core_lx) / D.sw;
        x[l] = double(nx);
      }
    }
  }
  D.lx = lx_new;
  return it_done;
}

// countRowOrientVioKernel (DatabaseKernels.hpp:468-504) + overflow recount
static LgVioStat check_place(Dev& D) {
  update_usage(D);
  LgVioStat v{};
  v.total_ovf = (int)compute_overflow(D);
  for (int c = 0; c < D.N; c++) {
    int vdd = (D.ly[c] >= 0 && D.ly[c] < D.y_size) ? D.vdd_up[D.ly[c]] : -1;
    int viol = 0;
    if ((D.h[c] & 1) == 1) {
      if (D.need_rot[c] == 0 && D.gnd[c] != vdd) viol = 1;
      else if (D.need_rot[c] == 1 && D.gnd[c] == vdd) viol = 1;
    } else if (D.gnd[c] != vdd) viol = 1;
    v.num_row_orient_vio += viol;
  }
  // strict counts: every footprint pixel inside the grid, a valid site, of
  // the cell's own group; even heights on their own rail
  for (int c = 0; c < D.N; c++) {
    bool site = false, fence = false, core = false;
    for (int x = D.lx[c]; x < D.lx[c] + D.w[c]; x++)
      for (int y = D.ly[c]; y < D.ly[c] + D.h[c]; y++) {
        if (x < 0 || y < 0 || x >= D.x_size || y >= D.y_size) { core = true; continue; }
        const size_t p = D.pid(x, y);
        if (!D.valid[p]) site = true;
        else if (D.pgroup[p] != D.cgroup[c]) fence = true;
      }
    v.num_site_vio += site;
    v.num_fence_vio += fence;
    v.num_core_vio += core;
    const bool row_in = D.ly[c] >= 0 && D.ly[c] < D.y_size;
    if (D.h[c] % 2 == 0 && (!row_in || D.gnd[c] != D.vdd_up[D.ly[c]])) v.num_rail_vio++;
  }
  if (D.tech) {
    // countPinVioKernel / countEdgeSpacingVioKernel (DatabaseKernels.hpp:
    // 286-504), each cell's orientation from its row's rail for every
    // height.  The edge maps are rebuilt from the final positions first;
    // the reference counts on the maps the solve left, which the QP refine
    // (it moves cells in x) does not update.
    edge_init(D);
    int64_t edge_sum = 0;
    for (int c = 0; c < D.N; c++) {
      const bool rot = D.vdd_up[D.ly[c]] != D.gnd[c];
      int s, a;
      pin_vio_split(D, c, rot, D.lx[c], D.ly[c], s, a);
      v.num_pin_short_vio += s;
      v.num_pin_access_vio += a;
      int l, r;
      edge_sides(D, -1, c, rot, D.lx[c], D.ly[c], l, r);
      edge_sum += l + r;
    }
    v.num_edge_spacing_vio = (int)(edge_sum / 2);  // pairs, counted from both sides
  }
  return v;
}

// computeHpwlFromDB (Legalizer.cpp:402-437): per net, the bounding box of
// its pins' average positions (dbITerm::getAvgXY, dbITerm.cpp:94-151, for
// orientation N and, on flipped cells, FS: y measured down from the top).
static int64_t hpwl_from_db(const LgDesign* d, const int32_t* lx, const int32_t* ly, const uint8_t* fs) {
  int64_t sum = 0;
  for (int k = 0; k < d->num_nets; k++) {
    int lo_x = kIntMax, lo_y = kIntMax, hi_x = kIntMin, hi_y = kIntMin;
    bool any = false;
    for (int q = d->net_start[k]; q < d->net_start[k + 1]; q++) {
      const int i = d->pin_inst[q];
      int ax = d->pin_dx[q], ay = d->pin_dy[q];
      if (i >= 0) {
        ax = lx[i] + d->pin_dx[q];
        ay = (fs && fs[i]) ? (ly[i] + d->inst_dy[i]) - d->pin_dy[q] : ly[i] + d->pin_dy[q];
      }
      lo_x = std::min(lo_x, ax); lo_y = std::min(lo_y, ay);
      hi_x = std::max(hi_x, ax); hi_y = std::max(hi_y, ay);
      any = true;
    }
    if (any) sum += int64_t(hi_x - lo_x) + int64_t(hi_y - lo_y);
  }
  return sum;
}

// parseICCAD17Constraint (Util.cpp:190-217)
static void parse_constraints(const char* path, int rh, float& util, float& disp) {
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    auto t = tokenize(line, "=");
    if (t.empty()) continue;
    if (t.front() == "maximum_utilization") util = std::stof(t.back().substr(0, t.back().size() - 1));
    if (t.front() == "maximum_movement") disp = std::stof(t.back().substr(0, t.back().size() - 4)) * float(rh);
  }
}

// Export of the host grid into the device-level descriptors
// (CudaDatabase::copyGridInfo / copyFromHostCells, CudaDatabase.cu:43-230)
struct Export {
  std::vector<int> valid, group, region, vdd;
  std::vector<int> rx0, ry0, rx1, ry1;
  std::vector<int> ox, oy, lx, ly, w, h, grp, gnd;
  // technology (importTechInfo)
  std::vector<int> pmetal, rules, ledge, redge, macro, mstart, pbgn, pend, play;
  LgGridDesc g{};
  LgCellsSoA c{};
};

// Legalizer::importTechInfo (DatabaseIO.cpp:381-572) into grid units:
// importEdgeSpcingRules (416-441), importCellMacro (443-526: a tech info per
// non-BLOCK master, pins as bounding boxes in whole sites / rows, upper
// corner clipped to the master, flattened x * h + y), importPowerMetals
// (528-572) with Grid::markPowerMetalLayer (Grid.cpp:345-362).
static int import_tech(const Host& H, const LgDesign* d, Export& E) {
  const Grid& g = H.grid;
  const int M = d->num_masters;
  int max_type = 0;
  for (int k = 0; k < d->num_edge_spacings; k++) {
    max_type = std::max(max_type, d->es_type1[k]);
    max_type = std::max(max_type, d->es_type2[k]);
  }
  E.rules.assign(max_type * 2 + 1, 0);
  for (int k = 0; k < d->num_edge_spacings; k++) {
    int sp_dbu = static_cast<int>(d->es_spacing_um[k] * static_cast<float>(d->dbu_per_micron));
    E.rules[d->es_type1[k] + d->es_type2[k]] = sp_dbu / g.sw;
  }
  std::vector<int> id_of(M, -1), size_from(M, -1);
  for (const Cell& c : H.cells)
    if (size_from[d->inst_master[c.inst]] < 0) size_from[d->inst_master[c.inst]] = c.inst;
  E.mstart.assign(1, 0);
  int next_id = 0;
  for (int m = 0; m < M; m++) {
    if (d->master_is_block[m]) continue;  // DatabaseIO.cpp:454-456
    id_of[m] = next_id++;
    int wg = 0, hg = 0;
    if (size_from[m] >= 0) {
      wg = d->inst_dx[size_from[m]] / g.sw;
      hg = d->inst_dy[size_from[m]] / g.rh;
    }
    for (int i = d->master_pin_start[m]; wg > 0 && i < d->master_pin_start[m + 1]; i++) {
      int bx0 = d->mpin_lx[i] / g.sw, by0 = d->mpin_ly[i] / g.rh;
      int bx1 = std::min(wg - 1, d->mpin_ux[i] / g.sw), by1 = std::min(hg - 1, d->mpin_uy[i] / g.rh);
      if (bx0 < 0 || by0 < 0 || bx0 > bx1 || by0 > by1) return LG_ERR_DESIGN;
      E.pbgn.push_back(bx0 * hg + by0);
      E.pend.push_back(bx1 * hg + by1);
      E.play.push_back(d->mpin_layer[i]);
    }
    E.mstart.push_back((int)E.pbgn.size());
  }
  for (const Cell& c : H.cells) {
    int m = d->inst_master[c.inst];
    if (m < 0 || id_of[m] < 0) return LG_ERR_DESIGN;
    E.ledge.push_back(d->master_ledge[m]);
    E.redge.push_back(d->master_redge[m]);
    E.macro.push_back(id_of[m]);
  }
  E.pmetal.assign((size_t)g.x_size * g.y_size, -1);
  for (int k = 0; k < d->num_pwires; k++) {
    int layer = d->pwire_layer[k];
    if (layer < 0) continue;               // not a routing layer
    if (layer == 0 || layer > 2) continue;  // M1 and above M3 ignored
    int x1 = d->pwire_x1[k], x2 = d->pwire_x2[k], y1 = d->pwire_y1[k], y2 = d->pwire_y2[k];
    if (x1 > x2) std::swap(x1, x2);
    if (y1 > y2) std::swap(y1, y2);
    if (x1 == x2 && y1 == y2) continue;
    int width = d->pwire_width[k];
    int slx = std::max(H.core_lx, x1 - width / 2), sux = std::min(H.core_ux, x2 + width / 2);
    int sly = std::max(H.core_ly, y1 - width / 2), suy = std::min(H.core_uy, y2 + width / 2);
    int ax, ay, bx, by;
    g.expanded(slx, sly, sux, suy, ax, ay, bx, by);
    for (int x = ax; x <= bx; x++)
      for (int y = ay; y <= by; y++) {
        if (x < 0 || y < 0 || x >= g.x_size || y >= g.y_size) continue;
        int& v = E.pmetal[g.id(x, y)];
        v = v == -1 ? layer : std::min(layer, v);
      }
  }
  E.g.pixel_power_layer = E.pmetal.data();
  E.g.num_edge_rules = (int)E.rules.size();
  E.g.edge_spacing_rules = E.rules.data();
  E.c.ledge_type = E.ledge.data();
  E.c.redge_type = E.redge.data();
  if (!H.cells.empty() && E.mstart.size() > 1) {
    E.c.macro_id = E.macro.data();
    E.c.num_macros = (int)E.mstart.size() - 1;
    E.c.macro_pin_start = E.mstart.data();
    E.c.pin_bgn = E.pbgn.data();
    E.c.pin_end = E.pend.data();
    E.c.pin_layer = E.play.data();
  }
  return LG_OK;
}

static void export_host(const Host& H, Export& E) {
  const Grid& g = H.grid;
  E.valid.assign(g.valid.begin(), g.valid.end());
  E.group = g.group;
  E.region = g.region;
  E.vdd = g.vdd_up;
  for (auto& r : H.regions) {
    int x1, y1, x2, y2;
    g.shrunk(r.lx, r.ly, r.ux, r.uy, x1, y1, x2, y2);
    E.rx0.push_back(x1); E.ry0.push_back(y1); E.rx1.push_back(x2); E.ry1.push_back(y2);
  }
  for (auto& c : H.cells) {
    E.ox.push_back(c.init_lx); E.oy.push_back(c.init_ly);
    E.lx.push_back(c.lx); E.ly.push_back(c.ly);
    E.w.push_back(c.w / H.sw); E.h.push_back(c.h / H.rh);
    E.grp.push_back(c.group); E.gnd.push_back(c.gnd);
  }
  E.g = LgGridDesc{g.x_size, g.y_size, g.sw, g.rh, g.lx, g.ly, E.valid.data(), E.group.data(),
                   E.region.data(), E.vdd.data(), (int)H.regions.size(), E.rx0.data(), E.ry0.data(),
                   E.rx1.data(), E.ry1.data()};
  E.c = LgCellsSoA{(int)H.cells.size(), E.ox.data(), E.oy.data(), E.lx.data(), E.ly.data(),
                   E.w.data(), E.h.data(), E.grp.data(), E.gnd.data()};
}

}  // namespace ref

// ---------------------------------------------------------------------------
// C entry points of the oracle library (oracle/_build/libref_lg.so)
// ---------------------------------------------------------------------------
extern "C" {

const char* ref_last_error(void) { return ref::g_err.c_str(); }

// Grid size of the design (x_size, y_size) and number of movable cells.
int ref_grid_shape(const LgDesign* d, const LgParams* p, int32_t* x_size, int32_t* y_size,
                   int32_t* num_cells, int32_t* y_hint) {
  ref::Host H;
  int st = ref::import_db(H, d, p);
  if (st) return st;
  *x_size = H.grid.x_size; *y_size = H.grid.y_size;
  *num_cells = (int)H.cells.size();
  if (y_hint) *y_hint = H.p.y_hint;
  return LG_OK;
}

// Pixel maps after importDB (valid, group, region, vdd_up), shrunk region
// bounds, and the preplaced positions of the movable cells (dbu), in
// movable-cell order.  shape[4] = x_size, y_size, num_cells, num_regions.
int ref_import_and_preplace(const LgDesign* d, const LgParams* p, int32_t* shape, int32_t* valid,
                            int32_t* group, int32_t* region, int32_t* vdd_up, int32_t* region_bounds,
                            int32_t* cell_lx, int32_t* cell_ly, int32_t* num_fail) {
  ref::Host H;
  int st = ref::import_db(H, d, p);
  if (st) return st;
  const auto& g = H.grid;
  shape[0] = g.x_size; shape[1] = g.y_size; shape[2] = (int)H.cells.size(); shape[3] = (int)H.regions.size();
  if (!valid) return LG_OK;
  for (size_t i = 0; i < g.valid.size(); i++) {
    valid[i] = g.valid[i]; group[i] = g.group[i]; region[i] = g.region[i];
  }
  for (int y = 0; y < g.y_size; y++) vdd_up[y] = g.vdd_up[y];
  for (size_t k = 0; k < H.regions.size(); k++) {
    const auto& r = H.regions[k];
    int x1, y1, x2, y2;
    g.shrunk(r.lx, r.ly, r.ux, r.uy, x1, y1, x2, y2);
    region_bounds[4 * k] = x1; region_bounds[4 * k + 1] = y1;
    region_bounds[4 * k + 2] = x2; region_bounds[4 * k + 3] = y2;
  }
  ref::preplace(H);
  for (size_t i = 0; i < H.cells.size(); i++) { cell_lx[i] = H.cells[i].lx; cell_ly[i] = H.cells[i].ly; }
  if (num_fail) *num_fail = H.num_preplace_fail;
  return LG_OK;
}

// The grid-level technology of a design (importTechInfo): the power-metal
// map (x-major, x_size * y_size), the rule table (up to 16 entries; count in
// *n_rules) and the pins of the tech masters (up to cap; count in *n_pins),
// with the per-cell edge types and tech master ids (movable-cell order).
int ref_import_tech(const LgDesign* d, const LgParams* p, int32_t* pmetal, int32_t* rules, int32_t* n_rules,
                    int32_t* mstart, int32_t* pbgn, int32_t* pend, int32_t* play, int32_t cap, int32_t* n_pins,
                    int32_t* ledge, int32_t* redge, int32_t* macro) {
  ref::Host H;
  int st = ref::import_db(H, d, p);
  if (st) return st;
  ref::Export E;
  ref::export_host(H, E);
  if ((st = ref::import_tech(H, d, E))) return st;
  std::copy(E.pmetal.begin(), E.pmetal.end(), pmetal);
  *n_rules = (int)E.rules.size();
  for (int i = 0; i < std::min(*n_rules, 16); i++) rules[i] = E.rules[i];
  *n_pins = (int)E.pbgn.size();
  if (*n_pins > cap) return LG_ERR_ARG;
  std::copy(E.mstart.begin(), E.mstart.end(), mstart);
  std::copy(E.pbgn.begin(), E.pbgn.end(), pbgn);
  std::copy(E.pend.begin(), E.pend.end(), pend);
  std::copy(E.play.begin(), E.play.end(), play);
  std::copy(E.ledge.begin(), E.ledge.end(), ledge);
  std::copy(E.redge.begin(), E.redge.end(), redge);
  std::copy(E.macro.begin(), E.macro.end(), macro);
  return LG_OK;
}

// Device-layer run from explicit grid/cell descriptors: ADMM (+ optional
// post-process).  Outputs grid positions, rotation flags, pixel state, stats.
int ref_solve(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, int32_t do_post,
              int32_t* lx_grid, int32_t* ly_grid, int32_t* need_rot, int32_t* usage, float* dual,
              LgStats* stats, int32_t* trace, int32_t trace_cap) {
  ref::Dev D;
  ref::init_dev(D, g, c, p);
  ref::solve(D);
  if (usage) std::copy(D.usage.begin(), D.usage.end(), usage);
  if (dual) std::copy(D.dual.begin(), D.dual.end(), dual);
  if (do_post) {
    ref::rotate_cells(D);
    if (D.p.qp_admm_on) ref::qp_refine(D);
  }
  LgVioStat v = ref::check_place(D);
  for (int i = 0; i < D.N; i++) {
    if (lx_grid) lx_grid[i] = D.lx[i];
    if (ly_grid) ly_grid[i] = D.ly[i];
    if (need_rot) need_rot[i] = D.need_rot[i];
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = D.iterations;
    stats->converged = D.converged;
    stats->final_overflow = D.final_ovf;
    stats->rho = D.rho;
    stats->initial_dual = D.initial_dual;
    stats->num_cells = D.N;
    stats->num_adaptive_calls = D.adaptive_calls;
    stats->vio = v;
    stats->total_disp_dbu = ref::compute_displace(D);
  }
  if (trace)
    for (int i = 0; i < std::min<int>(trace_cap, (int)D.trace.size()); i++) trace[i] = D.trace[i];
  return LG_OK;
}

// Whole Legalizer::run (Legalizer.cpp:197-228) on the CPU.
int ref_run_design(const LgDesign* d, const LgParams* p, LgResult* r) {
  double t0 = ref::now_s();
  ref::Host H;
  int st = ref::import_db(H, d, p);
  if (st) return st;
  double t1 = ref::now_s();
  ref::preplace(H);
  double t2 = ref::now_s();
  ref::Export E;
  ref::export_host(H, E);
  if (d->num_masters > 0 && (st = ref::import_tech(H, d, E))) return st;
  LgParams pp = H.p;  // y_hint possibly adjusted
  ref::Dev D;
  ref::init_dev(D, &E.g, &E.c, &pp);
  ref::solve(D);
  double t3 = ref::now_s();
  ref::rotate_cells(D);
  if (pp.qp_admm_on) ref::qp_refine(D);
  double t4 = ref::now_s();
  LgVioStat v = ref::check_place(D);
  for (int i = 0; i < d->num_insts; i++) {
    r->out_lx[i] = d->inst_lx[i];
    r->out_ly[i] = d->inst_ly[i];
    r->out_flipped[i] = 0;
  }
  int64_t tot = 0, mx = 0;
  for (int k = 0; k < D.N; k++) {
    int inst = H.cells[k].inst;
    int lx = D.lx[k] * D.sw + D.core_lx, ly = D.ly[k] * D.rh + D.core_ly;
    r->out_lx[inst] = lx;
    r->out_ly[inst] = ly;
    r->out_flipped[inst] = (uint8_t)D.need_rot[k];
    int64_t dd = std::abs(H.cells[k].init_lx - lx) + std::abs(H.cells[k].init_ly - ly);
    tot += dd;
    mx = std::max(mx, dd);
  }
  LgStats& s = r->stats;
  std::memset(&s, 0, sizeof(s));
  s.iterations = D.iterations;
  s.converged = D.converged;
  s.final_overflow = D.final_ovf;
  s.total_disp_dbu = tot;
  s.max_disp_dbu = mx;
  double dbu = double(H.dbu);
  s.total_disp_um = double(tot) / dbu;
  s.max_disp_um = double(mx) / dbu;
  s.avg_disp_um = s.total_disp_um / double(d->num_insts);
  s.rho = D.rho;
  s.initial_dual = D.initial_dual;
  s.num_cells = D.N;
  s.num_adaptive_calls = D.adaptive_calls;
  s.vio = v;
  s.t_import_s = t1 - t0;
  s.t_preplace_s = t2 - t1;
  s.t_admm_s = t3 - t2;
  s.t_post_s = t4 - t3;
  // computeStat (Legalizer.cpp:439-476): HPWL before and after, in microns
  const int64_t h0 = ref::hpwl_from_db(d, d->inst_lx, d->inst_ly, nullptr);
  const int64_t h1 = ref::hpwl_from_db(d, r->out_lx, r->out_ly, r->out_flipped);
  s.hpwl_before_um = double(h0) / dbu;
  s.hpwl_after_um = double(h1) / dbu;
  s.delta_hpwl_pct = h0 > 0 ? s.hpwl_after_um / s.hpwl_before_um * 100.0 - 100.0 : 0.0;
  // printICCAD17Metric (Util.cpp:141-188)
  if (d->constraint_file && d->constraint_file[0]) {
    float util = 0.f, disp = 0.f;
    ref::parse_constraints(d->constraint_file, d->row_dy[d->num_rows - 1], util, disp);
    s.max_util_constraint = util;
    s.max_disp_constraint = disp;
  }
  {
    const double swd = H.sw, rhd = d->row_dy[d->num_rows - 1];
    std::map<int, double> per_h;
    int max_dbu = 0;
    for (int k = 0; k < D.N; k++) {
      const int inst = H.cells[k].inst;
      const int dd = std::abs(r->out_lx[inst] - H.cells[k].init_lx) + std::abs(r->out_ly[inst] - H.cells[k].init_ly);
      max_dbu = std::max(max_dbu, dd);
      per_h[H.cells[k].h] += double(dd) / swd;
    }
    double total_sw = 0.0;
    for (auto& kv : per_h) total_sw += kv.second / double(std::max(1, H.height_distribution.at(kv.first)));
    const double large_k = double(H.height_distribution.size());
    s.iccad17_s_am = large_k > 0 ? total_sw / large_k / (rhd / swd) : 0.0;
    s.iccad17_m_max = max_dbu / rhd;
    const int np = v.num_pin_short_vio + v.num_pin_access_vio, ne = v.num_edge_spacing_vio;
    const double pen = double(np + ne) / double(std::max<size_t>(H.cells.size(), 1));
    s.iccad17_score = (1.0 + s.delta_hpwl_pct / 100.0 + pen) * (1.0 + s.iccad17_m_max / 100.0) * s.iccad17_s_am;
  }
  s.t_total_s = ref::now_s() - t0;
  if (r->ovf_trace)
    for (int i = 0; i < std::min<int>(r->ovf_trace_cap, (int)D.trace.size()); i++) r->ovf_trace[i] = D.trace[i];
  return LG_OK;
}

// ---- single steps for the parity pins (tests/test_oracle_pins_cpu.py) ----
void ref_set_mutation(int m) { ref::g_mut = m; }

// One move of one cell on an explicit state (x-major usage, dual and
// partition colour; positions from c->lx_dbu/ly_dbu).  mode 0: primal step of
// partition `part` (doPrimalUpdateKernel's move), 1: final refine, 2: adaptive
// search (projected wide candidates).  out_xy receives the new grid position,
// costs the padded candidate costs before the tree-min (256 or 4096), usage
// the state after the commit.
int ref_move_one(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, int32_t* usage,
                 const float* dual, const int32_t* color, float rho, int32_t cell, int32_t part, int32_t mode,
                 int32_t* out_xy, float* costs) {
  ref::Dev D;
  ref::init_dev(D, g, c, p);
  const size_t P = (size_t)D.x_size * D.y_size;
  std::copy(usage, usage + P, D.usage.begin());
  std::copy(dual, dual + P, D.dual.begin());
  std::copy(color, color + P, D.color.begin());
  D.rho = rho;
  if (cell < 0 || cell >= D.N) return LG_ERR_ARG;
  // edge maps of the given positions, live and snapshot alike
  ref::edge_init(D);
  ref::edge_snapshot(D);
  std::vector<float> buf(mode == 2 ? 4096 : 256);
  if (mode == 2) ref::adaptive_move(D, cell, buf, costs);
  else ref::move_cell(D, part, cell, mode == 1, buf, costs);
  out_xy[0] = D.lx[cell];
  out_xy[1] = D.ly[cell];
  std::copy(D.usage.begin(), D.usage.end(), usage);
  return LG_OK;
}

// dual_update_functor_with_step_size over n pixels, in place.
void ref_dual_update(int64_t n, const int32_t* usage, float* dual, float rho, float step) {
  ref::Dev D;
  D.p.num_cpu_threads = 1;
  D.usage.assign(usage, usage + n);
  D.dual.assign(dual, dual + n);
  D.rho = rho;
  D.dual_step = step;
  ref::dual_update(D);
  std::copy(D.dual.begin(), D.dual.end(), dual);
}

// updatePartition of iteration `iter` (colour map x-major, partition of every
// cell, K meaning excluded) and, with usage given, executeSorting (processing
// order per partition: grouped[offset[k] .. offset[k+1]) and the sort keys).
int ref_partition_order(const LgGridDesc* g, const LgCellsSoA* c, const LgParams* p, int32_t iter,
                        const int32_t* usage, int32_t* color, int32_t* part_of, int32_t* grouped,
                        int32_t* offset, int32_t* key) {
  ref::Dev D;
  ref::init_dev(D, g, c, p);
  ref::update_partition(D, iter);
  std::copy(D.color.begin(), D.color.end(), color);
  std::copy(D.part_of.begin(), D.part_of.end(), part_of);
  if (usage) {
    std::copy(usage, usage + D.usage.size(), D.usage.begin());
    ref::execute_sorting(D);
    std::copy(D.grouped.begin(), D.grouped.end(), grouped);
    std::copy(D.offset.begin(), D.offset.end(), offset);
    std::copy(D.key.begin(), D.key.end(), key);
  }
  return LG_OK;
}

// Small exported pieces for the parity pins
void ref_bitonic_desc(int n, int* key, int* val) { ref::bitonic_desc(n, key, val); }
void ref_bitonic_asc(int n, int* key, int* val) { ref::bitonic_asc(n, key, val); }
int ref_tree_argmin(float* cost, int n) { return ref::tree_argmin(cost, n); }
void ref_direction_tables(int32_t* dflt /*[2*245]*/, int32_t* wide /*[2*4093]*/) {
  ref::Host H;
  ref::make_directions(H);
  std::copy(H.dir_default.begin(), H.dir_default.end(), dflt);
  std::copy(H.dir_wide.begin(), H.dir_wide.end(), wide);
}
void ref_thomas(int n, double a, double b, const double* d, const double* cp, double* dp, double* x) {
  ref::thomas(n, a, b, d, cp, dp, x);
}
// One cut row's site alignment alone; site receives site indices.
void ref_site_align_dp_row(int n, const double* x, const double* x0, const int* wd, int rlx, int rux, int sw,
                           int core_lx, int* site) {
  std::vector<int> bl(n), br(n);
  ref::site_align_dp_row(n, x, x0, wd, rlx, rux, sw, core_lx, bl.data(), br.data(), site);
}
// The cut-row QP alone (before site alignment): rows given by ncell; x0, wd in
// DBU per cell, row extents rlx/rux.  x receives the ADMM solution; returns the
// iteration count.
int ref_qp_admm(int nr, const int* ncell, const double* x0, const int* wd, const int* rlx, const int* rux,
                double rho, int min_iter, int max_iter, double* x) {
  std::vector<int> offset(nr);
  int nm = 0;
  for (int r = 0; r < nr; r++) { offset[r] = nm; nm += ncell[r]; }
  const int nc = nm + nr;
  std::vector<int> pos_c(nm), neg_c(nm), c_pos(nc), c_neg(nc);
  std::vector<double> bound(nc);
  ref::qp_constraints(nr, ncell, offset.data(), wd, rlx, rux, pos_c.data(), neg_c.data(), c_pos.data(),
                      c_neg.data(), bound.data());
  for (int l = 0; l < nm; l++) x[l] = x0[l];
  return ref::qp_admm(nr, ncell, offset.data(), nm, x0, pos_c.data(), neg_c.data(), c_pos.data(), c_neg.data(),
                      bound.data(), rho, min_iter, max_iter, x);
}

}  // extern "C"
