// Host side of the legalizer: the design-level entry lg_run_design (replaces
// legalizer::Legalizer::run, Legalizer.cpp:197-228) and the C++ class
// legalizer::Legalizer with the reference's setters.
//
// Import, pixel grid and preplace are host work in the reference too
// (DatabaseIO.cpp, Grid.cpp, Preplace.cpp).  Here the grid is kept as flat
// x-major arrays (no per-pixel objects) and preplace runs on all host cores.
#include <algorithm>
#include <array>
#include <cctype>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

#include "../../include/skyline_lg.h"
#include "lg_grid.h"
#include "../../include/skyline/Legalizer.hpp"

namespace {

thread_local std::string g_host_err;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kIntMax = INT_MAX;
constexpr int kGroupDefault = -1;   // Legalizer.h:33
constexpr int kGroupBlockage = -2;  // Legalizer.h:34

enum RegionKind : uint8_t { kDefaultRegion, kFenceRegion, kBlockageRegion };

struct Rgn {
  int lx, ly, ux, uy, group;
  RegionKind kind;
};

// Allocator whose resize() leaves new elements uninitialised, so that a
// pixel map of tens of millions of entries is filled once, in parallel.
template <typename T>
struct NoInit : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInit<U>;
  };
  NoInit() = default;
  template <typename U>
  NoInit(const NoInit<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
using PixelMap = std::vector<int32_t, NoInit<int32_t>>;

// Pixel grid over the core, flat and x-major (Grid.h:51-105).
class PixelGrid {
 public:
  int lx = 0, ly = 0, sw = 1, rh = 1, xs = 0, ys = 0;
  PixelMap valid, group, region;
  std::vector<int32_t> hov, vov;  // fence overlap bookkeeping (only if fences exist)
  std::vector<int32_t> vdd;

  size_t at(int x, int y) const { return (size_t)x * ys + y; }
  int start_x(int d) const { return (d - lx) / sw; }
  int start_y(int d) const { return (d - ly) / rh; }
  int end_x(int d) const {
    int q = (d - lx) / sw, r = (d - lx) % sw;
    return std::clamp(r == 0 ? q - 1 : q, 0, xs - 1);
  }
  int end_y(int d) const {
    int q = (d - ly) / rh, r = (d - ly) % rh;
    return std::clamp(r == 0 ? q - 1 : q, 0, ys - 1);
  }
  int dbu_x(int g) const { return g * sw + lx; }
  int dbu_y(int g) const { return g * rh + ly; }
  std::array<int, 4> expanded(int a, int b, int c, int d) const {
    return {start_x(a), start_y(b), end_x(c), end_y(d)};
  }
  std::array<int, 4> shrunk(int a, int b, int c, int d) const {
    return {end_x(a + sw), end_y(b + rh), start_x(c - sw), start_y(d - rh)};
  }
};

struct MovCell {
  int lx, ly, w, h, ilx, ily, group, gnd, inst;
};

struct Session {
  LgParams p;
  const LgDesign* d;
  int core_lx, core_ly, core_ux, core_uy;
  PixelGrid grid;
  std::vector<MovCell> cells;
  std::vector<Rgn> regions;
  std::unordered_map<int, std::vector<int>> by_group;  // group -> region ids
  int max_height = 0;
  std::vector<std::array<int, 2>> dirs;  // default candidate moves
  int preplace_fail = 0;
  std::map<int, int> hdist;  // cell height (dbu) -> movable cells (height_distribution_)
  std::vector<std::vector<std::pair<int, int>>> spans;  // per grid row: the rows' x-spans (markDbRow)
  bool any_fence = false, fences_disjoint = true;
  float max_util = 0.f, max_disp = 0.f;  // constraint file: percent, dbu
};

// tokenize (Util.cpp:22-47): split at the delimiters and at white space
std::vector<std::string> tokens_of(const std::string& line, const char* dels) {
  std::vector<std::string> out;
  std::string t;
  for (char ch : line) {
    if (std::strchr(dels, ch) || std::isspace((unsigned char)ch)) {
      if (!t.empty()) out.push_back(std::move(t));
      t.clear();
    } else {
      t.push_back(ch);
    }
  }
  if (!t.empty()) out.push_back(std::move(t));
  return out;
}

// parseSizeFile (Util.cpp:220-234): "name=width_in_sites=height_in_rows"
int read_size_file(const char* path, std::unordered_map<std::string, std::pair<int, int>>& out) {
  std::ifstream in(path);
  if (!in) { g_host_err = std::string("cannot open size file ") + path; return LG_ERR_ARG; }
  std::string line;
  while (std::getline(in, line)) {
    const auto t = tokens_of(line, "=");
    if (t.empty()) continue;
    if (t.size() < 3) { g_host_err = "size file: bad line '" + line + "'"; return LG_ERR_ARG; }
    char* e1 = nullptr;
    char* e2 = nullptr;
    const long w = std::strtol(t[1].c_str(), &e1, 10), h = std::strtol(t[2].c_str(), &e2, 10);
    if (*e1 || *e2 || w <= 0 || h <= 0) { g_host_err = "size file: bad size in '" + line + "'"; return LG_ERR_ARG; }
    out[t[0]] = {(int)w, (int)h};
  }
  return LG_OK;
}

// parseICCAD17Constraint (Util.cpp:190-217): maximum_utilization=U% and
// maximum_movement=Mrows (rows -> dbu through the row height)
int read_constraints(const char* path, int row_height, float& util, float& disp) {
  std::ifstream in(path);
  if (!in) { g_host_err = std::string("cannot open constraint file ") + path; return LG_ERR_ARG; }
  std::string line;
  while (std::getline(in, line)) {
    const auto t = tokens_of(line, "=");
    if (t.size() < 2) continue;
    const std::string& v = t.back();
    if (t.front() == "maximum_utilization" && v.size() > 1) util = std::strtof(v.substr(0, v.size() - 1).c_str(), nullptr);
    if (t.front() == "maximum_movement" && v.size() > 4)
      disp = std::strtof(v.substr(0, v.size() - 4).c_str(), nullptr) * (float)row_height;
  }
  return LG_OK;
}

// computeHpwlFromDB (Legalizer.cpp:402-437) on the given instance positions
// and flip flags (FS mirrors a pin's y, dbITerm::getAvgXY); nets without
// pins add nothing.
int64_t design_hpwl(const LgDesign* d, const int32_t* lx, const int32_t* ly, const uint8_t* flipped) {
  int64_t sum = 0;
  for (int k = 0; k < d->num_nets; k++) {
    int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
    for (int p = d->net_start[k]; p < d->net_start[k + 1]; p++) {
      const int i = d->pin_inst[p];
      int x, y;
      if (i < 0) {
        x = d->pin_dx[p];
        y = d->pin_dy[p];
      } else {
        x = lx[i] + d->pin_dx[p];
        y = (flipped && flipped[i]) ? ly[i] + d->inst_dy[i] - d->pin_dy[p] : ly[i] + d->pin_dy[p];
      }
      x0 = std::min(x0, x); x1 = std::max(x1, x);
      y0 = std::min(y0, y); y1 = std::max(y1, y);
    }
    if (x0 <= x1) sum += (int64_t)(x1 - x0) + (int64_t)(y1 - y0);
  }
  return sum;
}

int check_netlist(const LgDesign* d) {
  if (d->num_nets < 0 || (d->num_nets > 0 && (!d->net_start || !d->pin_inst || !d->pin_dx || !d->pin_dy))) {
    g_host_err = "netlist arrays missing";
    return LG_ERR_ARG;
  }
  for (int k = 0; k < d->num_nets; k++)
    if (d->net_start[k] > d->net_start[k + 1] || d->net_start[k] < 0) {
      g_host_err = "net_start is not a non-decreasing offset array";
      return LG_ERR_ARG;
    }
  const int np = d->num_nets > 0 ? d->net_start[d->num_nets] : 0;
  for (int p = 0; p < np; p++)
    if (d->pin_inst[p] < -1 || d->pin_inst[p] >= d->num_insts) {
      g_host_err = "pin instance index out of range";
      return LG_ERR_ARG;
    }
  return LG_OK;
}

void add_region(Session& S, const Rgn& r) {
  S.by_group[r.group].push_back((int)S.regions.size());
  S.regions.push_back(r);
}

// The pixel maps (x-major): markDbRow's valid pixels from the row spans
// (Grid.cpp:316-342), then the regions in order (markBlockageRegion,
// markGroupRegion, markNonGroupRegion: Grid.cpp:184-314) and the fence edge
// tests (handleRegionSegments, DatabaseIO.cpp:661-681).
void paint_pixels(Session& S) {
  PixelGrid& g = S.grid;
  const size_t P = (size_t)g.xs * g.ys;
  const int nt = std::max(1, S.p.num_cpu_threads);
  const bool pixels = true;
  const bool any_fence = S.any_fence, fences_disjoint = S.fences_disjoint;
  g.valid.resize(P);
  g.group.resize(P);
  g.region.resize(P);
  const auto& spans = S.spans;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int x = 0; x < g.xs; x++)
    for (int y = 0; y < g.ys; y++) {
      const size_t q = g.at(x, y);
      int v = 0;
      for (const auto& sp : spans[y])
        if (sp.first <= x && x <= sp.second) { v = 1; break; }
      g.valid[q] = v;
      g.group[q] = kGroupDefault;
      g.region[q] = -1;
    }
  if (pixels && any_fence && !fences_disjoint) { g.hov.assign(P, 0); g.vov.assign(P, 0); }
  for (int id = 0; id < (int)S.regions.size() && pixels; id++) {
    const Rgn& r = S.regions[id];
    if (r.kind == kBlockageRegion) {  // Grid::markBlockageRegion (Grid.cpp:286-314)
      auto [x1, y1, x2, y2] = g.expanded(r.lx, r.ly, r.ux, r.uy);
      x1 = std::max(0, x1); y1 = std::max(0, y1); x2 = std::min(g.xs - 1, x2); y2 = std::min(g.ys - 1, y2);
      for (int x = x1; x <= x2; x++)
        for (int y = y1; y <= y2; y++) {
          size_t p = g.at(x, y);
          g.valid[p] = 0; g.group[p] = r.group; g.region[p] = id;
        }
    } else if (r.kind == kFenceRegion) {  // Grid::markGroupRegion (Grid.cpp:184-262)
      auto [ex1, ey1, ex2, ey2] = g.expanded(r.lx, r.ly, r.ux, r.uy);
      auto [sx1, sy1, sx2, sy2] = g.shrunk(r.lx, r.ly, r.ux, r.uy);
      for (int x = ex1; x <= ex2; x++)
        for (int y = ey1; y <= ey2; y++) {
          size_t p = g.at(x, y);
          g.group[p] = r.group; g.region[p] = id; g.valid[p] = 1;
        }
      if (fences_disjoint) continue;  // edges tested below
      for (int x = ex1; x <= ex2; x++)
        for (int y = ey1; y <= ey2; y++) {
          size_t p = g.at(x, y);
          g.hov[p] += g.sw; g.vov[p] += g.rh;
        }
      if (ex1 == sx1 - 1)
        for (int y = ey1; y <= ey2; y++) g.hov[g.at(ex1, y)] += (sx1 * g.sw - r.lx) - g.sw;
      if (ex2 == sx2 + 1)
        for (int y = ey1; y <= ey2; y++) g.hov[g.at(ex2, y)] += (r.ux - ex2 * g.sw) - g.sw;
      if (ey1 == sy1 - 1)
        for (int x = ex1; x <= ex2; x++) g.vov[g.at(x, ey1)] += (sy1 * g.rh - r.ly) - g.rh;
      if (ey2 == sy2 + 1)
        for (int x = ex1; x <= ex2; x++) g.vov[g.at(x, ey2)] += (r.uy - ey2 * g.rh) - g.rh;
    } else {  // Grid::markNonGroupRegion (Grid.cpp:264-284); columns in parallel
      auto [x1, y1, x2, y2] = g.shrunk(r.lx, r.ly, r.ux, r.uy);
      x1 = std::clamp(x1, 0, g.xs - 1); y1 = std::clamp(y1, 0, g.ys - 1);
      x2 = std::clamp(x2, 0, g.xs - 1); y2 = std::clamp(y2, 0, g.ys - 1);
      const long long area = (long long)(x2 - x1 + 1) * (y2 - y1 + 1);
#pragma omp parallel for num_threads(nt) schedule(static) if (area > (1 << 16))
      for (int x = x1; x <= x2; x++)
        for (int y = y1; y <= y2; y++) g.region[g.at(x, y)] = id;
    }
  }
  if (pixels && any_fence && !fences_disjoint) {  // handleRegionSegments (DatabaseIO.cpp:661-681)
#pragma omp parallel for num_threads(nt) schedule(static)
    for (long long p = 0; p < (long long)P; p++)
      if ((g.hov[p] > 0 || g.vov[p] > 0) && (g.hov[p] < g.sw || g.vov[p] < g.rh)) g.valid[p] = 0;
  }
  if (pixels && any_fence && fences_disjoint) {
    // the same test on the edge columns and rows of each fence's box, with
    // the counts markGroupRegion gives them (Grid.cpp:184-262)
    for (const Rgn& r : S.regions) {
      if (r.kind != kFenceRegion) continue;
      auto [ex1, ey1, ex2, ey2] = g.expanded(r.lx, r.ly, r.ux, r.uy);
      auto [sx1, sy1, sx2, sy2] = g.shrunk(r.lx, r.ly, r.ux, r.uy);
      auto hov = [&](int x) {
        int v = g.sw;
        if (x == ex1 && ex1 == sx1 - 1) v += (sx1 * g.sw - r.lx) - g.sw;
        if (x == ex2 && ex2 == sx2 + 1) v += (r.ux - ex2 * g.sw) - g.sw;
        return v;
      };
      auto vov = [&](int y) {
        int v = g.rh;
        if (y == ey1 && ey1 == sy1 - 1) v += (sy1 * g.rh - r.ly) - g.rh;
        if (y == ey2 && ey2 == sy2 + 1) v += (r.uy - ey2 * g.rh) - g.rh;
        return v;
      };
      auto test = [&](int x, int y) {
        const int h = hov(x), v = vov(y);
        if ((h > 0 || v > 0) && (h < g.sw || v < g.rh)) g.valid[g.at(x, y)] = 0;
      };
      for (int y = ey1; y <= ey2; y++) { test(ex1, y); test(ex2, y); }
      for (int x = ex1; x <= ex2; x++) { test(x, ey1); test(x, ey2); }
    }
  }
}

// importCells (DatabaseIO.cpp:79-169) + Grid ctor (Grid.cpp:24-67) +
// region import (DatabaseIO.cpp:210-379) + region marking (Grid.cpp:184-314,
// DatabaseIO.cpp:661-681) + adjustYHint (Legalizer.cpp:230-247)
// With pixels == false only the cells, regions and sizes are built (the
// shape query of lg_import_and_preplace).
int import_design(Session& S, bool pixels = true) {
  const LgDesign* d = S.d;
  if (!d || d->num_rows <= 0) { g_host_err = "design has no rows"; return LG_ERR_ARG; }
  S.core_lx = d->core_lx; S.core_ly = d->core_ly; S.core_ux = d->core_ux; S.core_uy = d->core_uy;
  const int row_h = d->row_dy[d->num_rows - 1];
  std::vector<std::array<int, 4>> blk_macro, blk_std;
  std::map<int, int>& hdist = S.hdist;
  hdist.clear();
  // modified-ISPD2015 cell sizes (setSizeByFile)
  std::unordered_map<std::string, std::pair<int, int>> sizes;
  const bool resize = d->size_file && d->size_file[0];
  if (resize) {
    if (!d->inst_name) { g_host_err = "size_file needs inst_name"; return LG_ERR_ARG; }
    int st = read_size_file(d->size_file, sizes);
    if (st) return st;
  }
  const int site_w = d->row_site_width[d->num_rows - 1];
  S.cells.reserve((size_t)d->num_insts);
  // height_distribution_ counted in a short list of (height, count) pairs
  // (a handful of distinct heights), moved into the map afterwards
  std::vector<std::pair<int, int>> hcount;
  int last_h = -1, last_k = -1;
  for (int i = 0; i < d->num_insts; i++) {
    if (d->inst_dx[i] == 0 || d->inst_dy[i] < row_h) continue;
    std::array<int, 4> r = {d->inst_lx[i], d->inst_ly[i], d->inst_lx[i] + d->inst_dx[i],
                            d->inst_ly[i] + d->inst_dy[i]};
    if (d->inst_macro[i]) {
      if (!d->inst_fixed[i]) { g_host_err = "movable macros are not supported"; return LG_ERR_DESIGN; }
      blk_macro.push_back(r);
    } else if (d->inst_fixed[i]) {
      blk_std.push_back(r);
    } else {
      MovCell c{r[0], r[1], d->inst_dx[i], d->inst_dy[i], r[0], r[1],
                d->inst_group ? d->inst_group[i] : -1, d->inst_ground_at_bottom[i] ? 1 : 0, i};
      if (resize && d->inst_name[i]) {
        auto it = sizes.find(d->inst_name[i]);
        if (it != sizes.end()) {
          c.w = it->second.first * site_w;
          c.h = it->second.second * row_h;
        }
      }
      S.cells.push_back(c);
      if (c.h != last_h) {
        last_k = -1;
        for (int k = 0; k < (int)hcount.size(); k++)
          if (hcount[k].first == c.h) { last_k = k; break; }
        if (last_k < 0) { hcount.push_back({c.h, 0}); last_k = (int)hcount.size() - 1; }
        last_h = c.h;
      }
      hcount[last_k].second++;
    }
  }
  for (auto& [hh, k] : hcount) hdist[hh] += k;
  PixelGrid& g = S.grid;
  g.lx = d->core_lx; g.ly = d->core_ly;
  int min_rh = INT_MAX;
  for (int r = 0; r < d->num_rows; r++) min_rh = std::min(min_rh, d->row_dy[r]);
  g.rh = min_rh;
  g.sw = d->row_site_width[d->num_rows - 1];
  if (g.sw <= 0 || g.rh <= 0) { g_host_err = "bad site width / row height"; return LG_ERR_ARG; }
  g.xs = (d->core_ux - d->core_lx) / g.sw;
  g.ys = (d->core_uy - d->core_ly) / g.rh;
  if (g.xs <= 0 || g.ys <= 0) { g_host_err = "empty core"; return LG_ERR_ARG; }
  // markDbRow (Grid.cpp:316-342), rows first: the rows' x-spans per grid
  // row and the rail of each row; the pixels follow in paint_pixels
  S.spans.assign(g.ys, {});
  g.vdd.assign(g.ys, 0);
  for (int r = 0; r < d->num_rows; r++) {
    auto [x1, y1, x2, y2] = g.shrunk(d->row_lx[r], d->row_ly[r], d->row_lx[r] + d->row_dx[r],
                                     d->row_ly[r] + d->row_dy[r]);
    if (y1 > y2 || x1 > x2) continue;
    for (int y = y1; y <= y2; y++) S.spans[y].push_back({x1, x2});
    int o = d->row_orient[r];
    g.vdd[y1] = (o == LG_ORIENT_N || o == LG_ORIENT_FN) ? 1 : 0;
  }
  auto clip = [&](const std::array<int, 4>& r) {
    return std::array<int, 4>{std::max(S.core_lx, r[0]), std::max(S.core_ly, r[1]),
                              std::min(S.core_ux, r[2]), std::min(S.core_uy, r[3])};
  };
  for (auto& r : blk_macro) { auto c = clip(r); add_region(S, {c[0], c[1], c[2], c[3], kGroupBlockage, kBlockageRegion}); }
  for (auto& r : blk_std) { auto c = clip(r); add_region(S, {c[0], c[1], c[2], c[3], kGroupBlockage, kBlockageRegion}); }
  for (int b = 0; b < d->num_region_boxes; b++) {
    auto c = clip({d->box_lx[b], d->box_ly[b], d->box_ux[b], d->box_uy[b]});
    add_region(S, {c[0], c[1], c[2], c[3], d->box_group[b], kFenceRegion});
  }
  // default regions: sweep the fence/blockage rectangles bottom-up
  // (importNonGroupRegions, DatabaseIO.cpp:272-379)
  {
    std::vector<int> obst;
    for (int i = 0; i < (int)S.regions.size(); i++) obst.push_back(i);
    struct Ev { int y, id; bool enter; };
    std::vector<Ev> ev;
    for (int k = 0; k < (int)obst.size(); k++) {
      ev.push_back({S.regions[obst[k]].ly, k, true});
      ev.push_back({S.regions[obst[k]].uy, k, false});
    }
    std::sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) {
      return a.y != b.y ? a.y < b.y : (!a.enter && b.enter);
    });
    std::set<int> live;
    std::unordered_map<int, std::vector<int>> at_lx;
    auto emit = [&](int lx, int ly, int ux, int uy) {
      if (ux - lx < g.sw || uy - ly < g.rh) return;
      for (int id : at_lx[lx]) {
        Rgn& r = S.regions[id];
        if (r.ux - r.lx == ux - lx && r.uy == ly) { r.uy = uy; return; }
      }
      at_lx[lx].push_back((int)S.regions.size());
      add_region(S, {lx, ly, ux, uy, kGroupDefault, kDefaultRegion});
    };
    int y0 = S.core_ly;
    for (const Ev& e : ev) {
      std::vector<std::pair<int, int>> spans;
      for (int k : live) spans.push_back({S.regions[obst[k]].lx, S.regions[obst[k]].ux});
      std::sort(spans.begin(), spans.end());
      int x = S.core_lx;
      for (auto& [a, b] : spans) { emit(x, y0, a, e.y); x = b; }
      emit(x, y0, S.core_ux, e.y);
      if (e.enter) live.insert(e.id); else live.erase(e.id);
      y0 = e.y;
    }
    emit(S.core_lx, y0, S.core_ux, S.core_uy);
  }
  bool& any_fence = S.any_fence;
  any_fence = false;
  for (auto& r : S.regions) any_fence |= (r.kind == kFenceRegion);
  // Fences whose pixel boxes do not overlap: a pixel inside one box gets
  // exactly that fence's overlap counts, full (sw, rh) except on the box's
  // edge columns / rows, so only the edges need the test of
  // handleRegionSegments and no per-pixel counts are kept.  Otherwise the
  // counts are accumulated over the whole grid.
  bool& fences_disjoint = S.fences_disjoint;
  fences_disjoint = true;
  {
    std::vector<std::array<int, 4>> fb;
    for (auto& r : S.regions)
      if (r.kind == kFenceRegion) fb.push_back(g.expanded(r.lx, r.ly, r.ux, r.uy));
    if (fb.size() > 256) fences_disjoint = false;
    for (size_t i = 0; i < fb.size() && fences_disjoint; i++)
      for (size_t j = i + 1; j < fb.size() && fences_disjoint; j++)
        if (fb[i][0] <= fb[j][2] && fb[j][0] <= fb[i][2] && fb[i][1] <= fb[j][3] && fb[j][1] <= fb[i][3])
          fences_disjoint = false;
  }
  if (pixels) paint_pixels(S);
  for (auto& [hh, n] : hdist) S.max_height = std::max(S.max_height, hh);
  int mh = S.max_height / g.rh;
  if (S.p.y_hint < 3 * mh) S.p.y_hint = 3 * mh;
  for (int y = -6; y <= 6; y++) {  // default directions (Legalizer.cpp:252-275)
    static const int xr[7] = {40, 20, 10, 5, 2, 1, 0};
    for (int x = -xr[std::abs(y)]; x <= xr[std::abs(y)]; x++) S.dirs.push_back({x, y});
  }
  return LG_OK;
}

// ---- preplace (Preplace.cpp) ----------------------------------------------

// Grid::checkPixelStatus (Grid.cpp:365-383): true if every covered pixel is
// valid and belongs to the cell's group.
bool pixels_ok(const Session& S, const MovCell& c) {
  const PixelGrid& g = S.grid;
  auto [x1, y1, x2, y2] = g.expanded(c.lx, c.ly, c.lx + c.w, c.ly + c.h);
  for (int x = x1; x <= x2; x++)
    for (int y = y1; y <= y2; y++) {
      size_t p = g.at(x, y);
      if (!g.valid[p] || g.group[p] != c.group) return false;
    }
  return true;
}

struct Cand { int lx, ly, disp; };

bool outside(const Rgn& r, const MovCell& c, int lx, int ly) {  // Preplace.cpp:244-265
  return !(lx >= r.lx && ly >= r.ly && lx + c.w <= r.ux && ly + c.h <= r.uy);
}

// findProjectionPoint (Preplace.cpp:267-353)
Cand project(const Session& S, const MovCell& c, const Rgn& r) {
  const PixelGrid& g = S.grid;
  const int wg = c.w / g.sw, hg = c.h / g.rh;
  auto [gx1, gy1, gx2, gy2] = g.shrunk(r.lx, r.ly, r.ux, r.uy);
  if (gy1 > gy2 || gx1 > gx2) return {-1, -1, kIntMax};
  const int rlx = g.dbu_x(gx1), rly = g.dbu_y(gy1), rux = g.dbu_x(gx2), ruy = g.dbu_y(gy2);
  int nx = c.lx, ny = c.ly;
  if (c.lx + c.w > rux) nx = g.dbu_x(gx2 - wg + 1);
  else if (c.lx < rlx) nx = g.dbu_x(gx1);
  if (c.ly + c.h > ruy) ny = g.dbu_y(gy2 - hg + 1);
  else if (c.ly < rly) ny = g.dbu_y(gy1);
  const bool even = (hg % 2) == 0;
  if (even && g.vdd[g.start_y(ny)] != c.gnd) {
    int best = kIntMax, by = -1;
    for (int yc : {ny - g.rh, ny + g.rh}) {
      int dd = std::abs(nx - c.ilx) + std::abs(yc - c.ily);
      if (!outside(r, c, nx, yc) && dd < best) { best = dd; by = yc; }
    }
    ny = by;
  }
  const bool out = outside(r, c, nx, ny);
  bool bad = false;
  if (even && !out) {
    int x0 = g.start_x(nx), y0 = g.start_y(ny);
    for (int x = x0; x < x0 + wg && !bad; x++)
      for (int y = y0; y < y0 + hg; y++)
        if (!g.valid[g.at(x, y)]) { bad = true; break; }
  }
  if (out || bad) return {-1, -1, kIntMax};
  return {nx, ny, std::abs(c.ilx - nx) + std::abs(c.ily - ny)};
}

bool project_group(const Session& S, MovCell& c) {
  auto it = S.by_group.find(c.group);
  if (it == S.by_group.end() || it->second.empty()) return false;
  Cand best = project(S, c, S.regions[it->second[0]]);
  for (size_t k = 1; k < it->second.size(); k++) {
    Cand q = project(S, c, S.regions[it->second[k]]);
    if (q.disp < best.disp) best = q;
  }
  if (best.disp >= kIntMax) return false;
  c.lx = best.lx; c.ly = best.ly;
  return true;
}

// roundToNearsetPoint (Preplace.cpp:111-157)
void snap(const Session& S, MovCell& c) {
  const PixelGrid& g = S.grid;
  const int wg = c.w / g.sw, hg = c.h / g.rh;
  int gx = std::min(g.xs - wg, std::max(0, g.start_x(c.lx)));
  int ax = g.dbu_x(gx), bx = g.dbu_x(std::min(g.xs - wg, gx + 1));
  int nx = std::abs(c.lx - ax) < std::abs(c.lx - bx) ? ax : bx;
  int gy = std::min(g.ys - hg, std::max(0, g.start_y(c.ly)));
  int ay = g.dbu_y(gy), by = g.dbu_y(std::min(g.ys - hg, gy + 1));
  if (ay == by) by = g.dbu_y(std::max(0, gy - 1));
  int ny = std::abs(c.ly - ay) < std::abs(c.ly - by) ? ay : by;
  if (hg % 2 == 0 && g.vdd[g.start_y(ny)] != c.gnd) {
    if (ny == ay) ny = (by + c.h <= S.core_uy) ? by : ay - g.rh;
    else if (ny == by) ny = (ay >= S.core_ly) ? ay : by + g.rh;
  }
  c.lx = nx; c.ly = ny;
}

// checkHopeless (Preplace.cpp:159-207) over the 245 default moves
bool hopeless(const Session& S, const MovCell& c0) {
  const PixelGrid& g = S.grid;
  MovCell c = c0;
  const int gx = g.start_x(c0.lx), gy = g.start_y(c0.ly);
  for (auto& d : S.dirs) {
    int nx = gx + d[0], ny = gy + d[1];
    c.lx = g.dbu_x(nx); c.ly = g.dbu_y(ny);
    if (c.lx + c.w > S.core_ux || c.ly + c.h > S.core_uy || c.lx < S.core_lx || c.ly < S.core_ly) continue;
    if (g.vdd[ny] == c.gnd && pixels_ok(S, c)) return false;
  }
  return true;
}

// initialLegalize (Preplace.cpp:18-72)
void preplace(Session& S) {
  const int n = (int)S.cells.size();
  const int nt = std::max(1, S.p.num_cpu_threads);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (int i = 0; i < n; i++) {
    MovCell& c = S.cells[i];
    snap(S, c);
    if (!pixels_ok(S, c)) project_group(S, c);
  }
  int fails = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024) reduction(+ : fails)
  for (int i = 0; i < n; i++) {
    MovCell& c = S.cells[i];
    if (hopeless(S, c) && !project_group(S, c)) fails++;
  }
  S.preplace_fail = fails;
}

// The device path of the import: pixel codes built and cells preplaced on
// the GPU (lg_grid.hip) from the rows, the regions and the cells that
// import_design(S, false) collected.  Available without the adaptive
// search (which reads host region maps) and when the fence boxes do not
// overlap (else the overlap counts need the whole grid): *used = false then.
// Grid-level ICCAD 2017 technology of lg_create from the design's
// (Legalizer::importTechInfo, DatabaseIO.cpp:381-572).
struct Tech {
  bool on = false;
  std::vector<int32_t> pmetal;                 // [xs * ys] x-major, -1: none
  std::vector<int32_t> rules;                  // index a + b
  std::vector<int32_t> ledge, redge, macro;    // per movable cell
  std::vector<int32_t> mstart, pbgn, pend, play;
};

int import_tech(const Session& S, Tech& T) {
  const LgDesign* d = S.d;
  if (d->num_masters <= 0) return LG_OK;
  const PixelGrid& g = S.grid;
  const int M = d->num_masters;
  if (!d->inst_master || !d->master_ledge || !d->master_redge || !d->master_is_block || !d->master_pin_start) {
    g_host_err = "technology: master arrays missing";
    return LG_ERR_ARG;
  }
  T.on = true;
  // importEdgeSpcingRules (DatabaseIO.cpp:416-441): spacing in dbu with the
  // dbu as float, truncated, then in whole sites
  int max_type = 0;
  for (int k = 0; k < d->num_edge_spacings; k++) max_type = std::max({max_type, d->es_type1[k], d->es_type2[k]});
  T.rules.assign(max_type * 2 + 1, 0);
  for (int k = 0; k < d->num_edge_spacings; k++) {
    const int dbu_sp = static_cast<int>(d->es_spacing_um[k] * static_cast<float>(d->dbu_per_micron));
    T.rules[d->es_type1[k] + d->es_type2[k]] = dbu_sp / g.sw;
  }
  // importCellMacro (DatabaseIO.cpp:450-512): a tech info per master that is
  // no BLOCK, its signal pins' bounding boxes in whole sites / rows (upper
  // corner clipped to the master), flattened x * h + y
  std::vector<int> tech_id(M, -1), first(M, -1);
  // the master's size in grid units, from its first movable instance
  for (const MovCell& c : S.cells) {
    const int m = d->inst_master[c.inst];
    if (m >= 0 && m < M && first[m] < 0) first[m] = c.inst;
  }
  T.mstart.assign(1, 0);
  for (int m = 0, id = 0; m < M; m++) {
    if (d->master_is_block[m]) continue;
    tech_id[m] = id++;
    const int wg = first[m] >= 0 ? d->inst_dx[first[m]] / g.sw : 0;
    const int hg = first[m] >= 0 ? d->inst_dy[first[m]] / g.rh : 0;
    for (int i = d->master_pin_start[m]; i < d->master_pin_start[m + 1]; i++) {
      if (wg == 0 || hg == 0) break;  // no movable instance: its pins are never read
      const int x0 = d->mpin_lx[i] / g.sw, y0 = d->mpin_ly[i] / g.rh;
      const int x1 = std::min(wg - 1, d->mpin_ux[i] / g.sw), y1 = std::min(hg - 1, d->mpin_uy[i] / g.rh);
      if (x0 < 0 || y0 < 0 || x0 > x1 || y0 > y1) {
        g_host_err = "technology: a pin box outside its master";
        return LG_ERR_DESIGN;
      }
      T.pbgn.push_back(x0 * hg + y0);
      T.pend.push_back(x1 * hg + y1);
      T.play.push_back(d->mpin_layer[i]);
    }
    T.mstart.push_back((int32_t)T.pbgn.size());
  }
  const int N = (int)S.cells.size();
  T.ledge.resize(N); T.redge.resize(N); T.macro.resize(N);
  for (int i = 0; i < N; i++) {
    const int m = d->inst_master[S.cells[i].inst];
    if (m < 0 || m >= M || tech_id[m] < 0) {
      g_host_err = "technology: a movable instance without a cell master";
      return LG_ERR_DESIGN;
    }
    T.ledge[i] = d->master_ledge[m];
    T.redge[i] = d->master_redge[m];
    T.macro[i] = tech_id[m];
  }
  // importPowerMetals (DatabaseIO.cpp:528-572) + markPowerMetalLayer
  // (Grid.cpp:345-362): wires on routing layers 1 and 2, widened by half
  // their width and clipped to the core; a pixel keeps the lowest layer
  T.pmetal.assign((size_t)g.xs * g.ys, -1);
  for (int k = 0; k < d->num_pwires; k++) {
    const int layer = d->pwire_layer[k];
    if (layer < 0 || layer == 0 || layer > 2) continue;
    int x1 = d->pwire_x1[k], x2 = d->pwire_x2[k], y1 = d->pwire_y1[k], y2 = d->pwire_y2[k];
    if (x1 > x2) std::swap(x1, x2);
    if (y1 > y2) std::swap(y1, y2);
    if (x1 == x2 && y1 == y2) continue;  // a via
    const int hw = d->pwire_width[k] / 2;
    const int lx = std::max(S.core_lx, x1 - hw), ux = std::min(S.core_ux, x2 + hw);
    const int ly = std::max(S.core_ly, y1 - hw), uy = std::min(S.core_uy, y2 + hw);
    auto [a, b, c, e] = g.expanded(lx, ly, ux, uy);
    for (int x = std::max(a, 0); x <= std::min(c, g.xs - 1); x++)
      for (int y = std::max(b, 0); y <= std::min(e, g.ys - 1); y++) {
        int32_t& v = T.pmetal[g.at(x, y)];
        v = v == -1 ? layer : std::min(layer, v);
      }
  }
  return LG_OK;
}

int device_grid(Session& S, void** d_pcode, bool* used) {
  *used = false;
  *d_pcode = nullptr;
  const PixelGrid& g = S.grid;
  int num_paint = 0;
  for (const Rgn& r : S.regions) num_paint += r.kind != kDefaultRegion;
  int max_group = -1;
  for (const Rgn& r : S.regions) max_group = std::max(max_group, r.group);
  if (S.p.use_adaptive_search || (S.any_fence && !S.fences_disjoint) || num_paint > 65535 ||
      max_group + 2 > 0x7fff || g.xs > 65535 || g.ys > 65535)
    return LG_OK;
  LgGridBuild in;
  in.xs = g.xs; in.ys = g.ys; in.sw = g.sw; in.rh = g.rh; in.lx = g.lx; in.ly = g.ly;
  in.core_lx = S.core_lx; in.core_ly = S.core_ly; in.core_ux = S.core_ux; in.core_uy = S.core_uy;
  std::vector<int32_t> ss(g.ys + 1, 0), sp, paint, fence, cell, reg, gs, gl, dirs;
  for (int y = 0; y < g.ys; y++) {
    ss[y + 1] = ss[y] + (int)S.spans[y].size();
    for (auto& [a, b] : S.spans[y]) { sp.push_back(a); sp.push_back(b); }
  }
  for (const Rgn& r : S.regions) {
    if (r.kind == kBlockageRegion) {  // markBlockageRegion: clamped, invalid
      auto [x1, y1, x2, y2] = g.expanded(r.lx, r.ly, r.ux, r.uy);
      paint.insert(paint.end(), {std::max(0, x1), std::max(0, y1), std::min(g.xs - 1, x2), std::min(g.ys - 1, y2), 0});
    } else if (r.kind == kFenceRegion) {  // markGroupRegion: valid, the fence's group
      auto [x1, y1, x2, y2] = g.expanded(r.lx, r.ly, r.ux, r.uy);
      paint.insert(paint.end(), {x1, y1, x2, y2, r.group + 2});
      auto [sx1, sy1, sx2, sy2] = g.shrunk(r.lx, r.ly, r.ux, r.uy);
      fence.insert(fence.end(), {x1, y1, x2, y2, sx1, sy1, sx2, sy2, r.lx, r.ly, r.ux, r.uy});
    }
    reg.insert(reg.end(), {r.lx, r.ly, r.ux, r.uy});
  }
  for (const MovCell& c : S.cells) cell.insert(cell.end(), {c.lx, c.ly, c.w, c.h, c.ilx, c.ily, c.group, c.gnd});
  int gmin = 0, gmax = -1;
  for (auto& [k, v] : S.by_group) { gmin = std::min(gmin, k); gmax = std::max(gmax, k); }
  for (const MovCell& c : S.cells) { gmin = std::min(gmin, c.group); gmax = std::max(gmax, c.group); }
  gs.assign(gmax - gmin + 2, 0);
  for (int k = gmin; k <= gmax; k++) {
    auto it = S.by_group.find(k);
    if (it != S.by_group.end()) gl.insert(gl.end(), it->second.begin(), it->second.end());
    gs[k - gmin + 1] = (int)gl.size();
  }
  for (auto& d : S.dirs) { dirs.push_back(d[0]); dirs.push_back(d[1]); }
  in.vdd = g.vdd.data();
  in.span_start = ss.data(); in.span = sp.data();
  in.num_paint = (int)paint.size() / 5; in.paint = paint.data();
  in.num_fences = S.any_fence ? (int)fence.size() / 12 : 0; in.fence = fence.data();
  in.num_cells = (int)S.cells.size(); in.cell = cell.data();
  in.num_regions = (int)S.regions.size(); in.region = reg.data();
  in.group_min = gmin; in.num_group_lists = gmax - gmin + 1; in.glist_start = gs.data(); in.glist = gl.data();
  in.num_dirs = (int)S.dirs.size(); in.dirs = dirs.data();
  std::vector<int32_t> lx(S.cells.size()), ly(S.cells.size());
  int fails = 0;
  std::string err;
  const int st = lg_grid_preplace_device(in, d_pcode, lx.data(), ly.data(), &fails, nullptr, err);
  if (st) { g_host_err = err; return st; }
  for (size_t i = 0; i < S.cells.size(); i++) { S.cells[i].lx = lx[i]; S.cells[i].ly = ly[i]; }
  S.preplace_fail = fails;
  *used = true;
  return LG_OK;
}

}  // namespace

extern "C" {

void lg_default_params(LgParams* p) {  // Legalizer::initParameters (Legalizer.cpp:47-93)
  std::memset(p, 0, sizeof(*p));
  p->flag_iccad17 = 0;
  p->standard_admm = 0;
  p->log_freq = 100;
  p->num_cpu_threads = 16;
  p->max_disp_in_row_height = 3;
  p->max_admm_iter = 25000;
  p->admm_terminate_thr = 0;
  p->x_hint = 250;
  p->y_hint = 12;
  p->iter_update_partition = 1;
  p->max_disp_coeff = 40.0f;
  p->tech_penalty = 1.0f;
  p->edge_spacing_penalty = 10000.0f;
  p->iter_threshold_dual_update = 300;
  p->iter_update_dual_step = 100;
  p->iter_perturbation_on = 2;
  p->coeff_to_init_rho = 3.0f;
  p->coeff_to_init_dual = 0.5f;
  p->coeff_to_update_dual_step = 0.2f;
  p->qp_admm_on = 0;
  p->qp_admm_dp_site_align = 1;
  p->qp_admm_max_iter = 10000;
  p->qp_admm_min_iter = 5;
  p->qp_admm_rho = 2.0f;
  p->use_adaptive_search = 0;
  p->stagnation_window = 10;
  p->stagnation_threshold = 0.01f;
  p->min_iter_to_call_adaptive_search = 300;
  p->cooltime_to_call_adaptive_search = 100;
  p->ovf_ratio_to_call_adaptive_search = 1.05f;
  p->coeff_to_neglect_displace = 20.0f;
  p->random_sorting_seed = 0;
  p->sorting_policy = LG_SORT_POLICY_OVERFLOW;
  p->sorting_algorithm = LG_SORT_M_BITONIC;
  p->verbose = 0;
}

// Import + preplace only: pixel maps (x-major) and preplaced movable-cell
// positions (dbu, movable-cell order).  Exposed for parity tests.
int lg_import_and_preplace(const LgDesign* d, const LgParams* p, int32_t* shape, int32_t* valid,
                           int32_t* group, int32_t* region, int32_t* vdd_up, int32_t* region_bounds,
                           int32_t* cell_lx, int32_t* cell_ly) {
  if (!d || !p || !shape) { g_host_err = "null argument"; return LG_ERR_ARG; }
  Session S;
  S.p = *p;
  S.d = d;
  const double t0 = now_s();
  int st = import_design(S, valid != nullptr);  // a shape query builds no pixels
  if (st) return st;
  const PixelGrid& g = S.grid;
  shape[0] = g.xs; shape[1] = g.ys;
  shape[2] = (int)S.cells.size();
  shape[3] = (int)S.regions.size();
  if (!valid) return LG_OK;  // shape query
  const long long P = (long long)g.xs * g.ys;
  const int nt = std::max(1, S.p.num_cpu_threads);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (long long i = 0; i < P; i++) { valid[i] = g.valid[i]; group[i] = g.group[i]; region[i] = g.region[i]; }
  for (int y = 0; y < g.ys; y++) vdd_up[y] = g.vdd[y];
  for (size_t k = 0; k < S.regions.size(); k++) {
    const Rgn& rg = S.regions[k];
    auto [a, b, c, e] = g.shrunk(rg.lx, rg.ly, rg.ux, rg.uy);
    region_bounds[4 * k] = a; region_bounds[4 * k + 1] = b;
    region_bounds[4 * k + 2] = c; region_bounds[4 * k + 3] = e;
  }
  const double t1 = now_s();
  preplace(S);
  for (size_t i = 0; i < S.cells.size(); i++) { cell_lx[i] = S.cells[i].lx; cell_ly[i] = S.cells[i].ly; }
  if (S.p.verbose)
    std::fprintf(stderr, "[lg] import %.3f s (pixel maps out included), preplace %.3f s\n", t1 - t0, now_s() - t1);
  return LG_OK;
}

int lg_import_and_preplace_device(const LgDesign* d, const LgParams* p, int32_t* shape, int32_t* valid,
                                  int32_t* group, int32_t* cell_lx, int32_t* cell_ly) {
  if (!d || !p || !shape) { g_host_err = "null argument"; return LG_ERR_ARG; }
  Session S;
  S.p = *p;
  S.d = d;
  int st = import_design(S, false);
  if (st) return st;
  const PixelGrid& g = S.grid;
  shape[0] = g.xs; shape[1] = g.ys;
  shape[2] = (int)S.cells.size();
  shape[3] = (int)S.regions.size();
  if (!valid) return LG_OK;
  if (!group || !cell_lx || !cell_ly) { g_host_err = "null output array"; return LG_ERR_ARG; }
  void* d_pcode = nullptr;
  bool used = false;
  if ((st = device_grid(S, &d_pcode, &used))) return st;
  if (!used) { g_host_err = "no device path for this design (adaptive search or overlapping fences)"; return LG_ERR_STATE; }
  const size_t P = (size_t)g.xs * g.ys;
  std::vector<int16_t> pc(P);
  std::string err;
  st = lg_grid_pcode_to_host(d_pcode, P, pc.data(), err);
  lg_grid_free(d_pcode);
  if (st) { g_host_err = err; return st; }
  for (int y = 0; y < g.ys; y++)
    for (int x = 0; x < g.xs; x++) {
      const int v = pc[(size_t)y * g.xs + x];
      valid[g.at(x, y)] = v != 0;
      group[g.at(x, y)] = v != 0 ? v - 2 : -1;
    }
  for (size_t i = 0; i < S.cells.size(); i++) { cell_lx[i] = S.cells[i].lx; cell_ly[i] = S.cells[i].ly; }
  return LG_OK;
}

int lg_run_design(const LgDesign* d, const LgParams* p, LgResult* r) {
  if (!d || !p || !r) { g_host_err = "null argument"; return LG_ERR_ARG; }
  const double t0 = now_s();
  Session S;
  S.p = *p;
  S.d = d;
  int st = check_netlist(d);
  if (st) return st;
  st = import_design(S, false);
  if (st) return st;
  const int row_h = d->row_dy[d->num_rows - 1];
  const bool constrained = d->constraint_file && d->constraint_file[0];
  if (constrained && (st = read_constraints(d->constraint_file, row_h, S.max_util, S.max_disp))) return st;
  // original_hpwl_ = computeHpwlFromDB() at import (DatabaseIO.cpp:68)
  const int64_t hpwl0 = design_hpwl(d, d->inst_lx, d->inst_ly, nullptr);
  const double t1 = now_s();
  // pixel grid and preplace on the device where possible, else on the host
  void* d_pcode = nullptr;
  bool on_device = false;
  if ((st = device_grid(S, &d_pcode, &on_device))) return st;
  if (!on_device) {
    paint_pixels(S);
    preplace(S);
  }
  const double t2 = now_s();
  const PixelGrid& g = S.grid;
  const int N = (int)S.cells.size();
  std::vector<int32_t> rx0, ry0, rx1, ry1;
  for (auto& rg : S.regions) {
    auto [a, b, c, e] = g.shrunk(rg.lx, rg.ly, rg.ux, rg.uy);
    rx0.push_back(a); ry0.push_back(b); rx1.push_back(c); ry1.push_back(e);
  }
  std::vector<int32_t> ox(N), oy(N), lx(N), ly(N), w(N), h(N), grp(N), gnd(N);
  for (int i = 0; i < N; i++) {
    const MovCell& c = S.cells[i];
    ox[i] = c.ilx; oy[i] = c.ily; lx[i] = c.lx; ly[i] = c.ly;
    w[i] = c.w / g.sw; h[i] = c.h / g.rh; grp[i] = c.group; gnd[i] = c.gnd;
  }
  LgGridDesc gd{g.xs, g.ys, g.sw, g.rh, g.lx, g.ly, g.valid.data(), g.group.data(), g.region.data(),
                g.vdd.data(), (int)S.regions.size(), rx0.data(), ry0.data(), rx1.data(), ry1.data()};
  LgCellsSoA cs{N, ox.data(), oy.data(), lx.data(), ly.data(), w.data(), h.data(), grp.data(), gnd.data()};
  Tech T;
  if ((st = import_tech(S, T))) {
    if (d_pcode) lg_grid_free(d_pcode);
    return st;
  }
  if (T.on) {
    gd.pixel_power_layer = T.pmetal.data();
    gd.num_edge_rules = (int32_t)T.rules.size();
    gd.edge_spacing_rules = T.rules.data();
    cs.ledge_type = T.ledge.data();
    cs.redge_type = T.redge.data();
    if (N > 0 && T.mstart.size() > 1) {
      cs.macro_id = T.macro.data();
      cs.num_macros = (int32_t)T.mstart.size() - 1;
      cs.macro_pin_start = T.mstart.data();
      cs.pin_bgn = T.pbgn.data();
      cs.pin_end = T.pend.data();
      cs.pin_layer = T.play.data();
    }
  }
  LgHandle* hd = nullptr;
  if ((st = on_device ? lg_create_device_grid(&gd, &cs, &S.p, d_pcode, &hd) : lg_create(&gd, &cs, &S.p, &hd)))
    return st;
  auto done = [&](int code) { lg_destroy(hd); return code; };
  if ((st = lg_solve(hd, nullptr))) return done(st);
  const double t3 = now_s();
  if ((st = lg_postprocess(hd, nullptr))) return done(st);
  const double t4 = now_s();
  std::vector<int32_t> gx(N), gy(N), rot(N);
  if ((st = lg_copy_out(hd, gx.data(), gy.data(), rot.data()))) return done(st);
  LgStats ds;
  if ((st = lg_get_stats(hd, &ds))) return done(st);
  int32_t ntr = 0;
  if (r->ovf_trace && r->ovf_trace_cap > 0) lg_get_ovf_trace(hd, r->ovf_trace, r->ovf_trace_cap, &ntr);
  lg_destroy(hd);
  // copyToHostCells + exportDB + computeStat (CudaDatabase.cu:274-309,
  // DatabaseIO.cpp:581-596, Legalizer.cpp:439-476)
  for (int i = 0; i < d->num_insts; i++) { r->out_lx[i] = d->inst_lx[i]; r->out_ly[i] = d->inst_ly[i]; r->out_flipped[i] = 0; }
  int64_t tot = 0, mx = 0;
  for (int i = 0; i < N; i++) {
    const MovCell& c = S.cells[i];
    int nx = gx[i] * g.sw + g.lx, ny = gy[i] * g.rh + g.ly;
    r->out_lx[c.inst] = nx;
    r->out_ly[c.inst] = ny;
    r->out_flipped[c.inst] = (uint8_t)rot[i];
    int64_t dd = std::abs(c.ilx - nx) + std::abs(c.ily - ny);
    tot += dd;
    mx = std::max(mx, dd);
  }
  LgStats& s = r->stats;
  s = ds;
  s.total_disp_dbu = tot;
  s.max_disp_dbu = mx;
  const double dbu = d->dbu_per_micron > 0 ? d->dbu_per_micron : 1;
  s.total_disp_um = double(tot) / dbu;
  s.max_disp_um = double(mx) / dbu;
  s.avg_disp_um = s.total_disp_um / double(std::max(d->num_insts, 1));
  s.t_import_s = t1 - t0;
  s.t_preplace_s = t2 - t1;
  s.t_admm_s = t3 - t2;
  s.t_post_s = t4 - t3;
  // computeStat's HPWL (Legalizer.cpp:465-472) after exportDB set the
  // positions and the FS flips
  const int64_t hpwl1 = design_hpwl(d, r->out_lx, r->out_ly, r->out_flipped);
  s.hpwl_before_um = double(hpwl0) / dbu;
  s.hpwl_after_um = double(hpwl1) / dbu;
  s.delta_hpwl_pct = hpwl0 > 0 ? s.hpwl_after_um / s.hpwl_before_um * 100.0 - 100.0 : 0.0;
  // printICCAD17Metric (Util.cpp:141-188)
  {
    std::map<int, double> sum_sw;
    int64_t mx_dbu = 0;
    for (int i = 0; i < N; i++) {
      const MovCell& c = S.cells[i];
      const int64_t dd = std::abs(c.ilx - r->out_lx[c.inst]) + std::abs(c.ily - r->out_ly[c.inst]);
      mx_dbu = std::max(mx_dbu, dd);
      sum_sw[c.h] += double(dd) / double(g.sw);
    }
    double tot = 0.0;
    for (auto& [hh, v] : sum_sw) tot += v / double(std::max(1, S.hdist[hh]));
    const double k = double(std::max<size_t>(S.hdist.size(), 1));
    s.iccad17_s_am = tot / k / (double(row_h) / double(g.sw));
    s.iccad17_m_max = double(mx_dbu) / double(row_h);
    const int np = s.vio.num_pin_short_vio + s.vio.num_pin_access_vio, ne = s.vio.num_edge_spacing_vio;
    const double drv = double(np + ne) / double(std::max(N, 1));
    s.iccad17_score = (1.0 + s.delta_hpwl_pct / 100.0 + drv) * (1.0 + s.iccad17_m_max / 100.0) * s.iccad17_s_am;
    s.max_util_constraint = S.max_util;
    s.max_disp_constraint = S.max_disp;
  }
  s.t_total_s = now_s() - t0;
  if (p->verbose) {  // Legalizer::printStats (Util.cpp:107-139)
    printf("------------------------------------------\n");
    printf("Legalization Time : %12.1f s\n", s.t_total_s);
    printf("Sum Displacement  : %12.1f um\n", s.total_disp_um);
    printf("Avg Displacement  : %12.1f um\n", s.avg_disp_um);
    printf("Max Displacement  : %12.1f um\n", s.max_disp_um);
    printf("\n");
    printf("HPWL Before       : %12.1f um\n", s.hpwl_before_um);
    printf("HPWL After        : %12.1f um\n", s.hpwl_after_um);
    printf("Delta HPWL        : %12.2f %%\n", s.delta_hpwl_pct);
    printf("\n");
    const double dbu_d = dbu;
    printf("Avg Displacement  : %12.2f (rows) \n", s.avg_disp_um * dbu_d / row_h);
    printf("Avg Displacement  : %12.2f (sites)\n", s.avg_disp_um * dbu_d / g.sw);
    printf("Max Displacement  : %12.2f (rows) \n", s.max_disp_um * dbu_d / row_h);
    printf("Max Displacement  : %12.2f (sites)\n", s.max_disp_um * dbu_d / g.sw);
    printf("Sum Displacement  : %12.2f (sites)\n", s.total_disp_um * dbu_d / g.sw);
    printf("\n");
    printf("NumOverlap        : %12d\n", s.vio.total_ovf);
    printf("NumRowOrientVio   : %12d\n", s.vio.num_row_orient_vio);
    if (constrained) {
      printf("\nICCAD 17 Contest Metric\n");
      printf("N_p   : %6d  \n", s.vio.num_pin_short_vio + s.vio.num_pin_access_vio);
      printf("N_e   : %6d  \n", s.vio.num_edge_spacing_vio);
      printf("M_max : %6.3f\n", s.iccad17_m_max);
      printf("S_am  : %6.3f\n", s.iccad17_s_am);
      printf("S     : %6.3f\n", s.iccad17_score);
    }
    printf("preplace           finished (takes %5.2f s)\n", s.t_preplace_s);
    printf("admmLegalize       finished (takes %5.2f s)\n", s.t_admm_s);
    printf("postProcess        finished (takes %5.2f s)\n", s.t_post_s);
    printf("------------------------------------------\n");
  }
  return LG_OK;
}

}  // extern "C"

// ---- the reference's public C++ class (include/legalizer/Legalizer.h:66-97)
namespace legalizer {

Legalizer::Legalizer(const LgDesign* design) : design_(design) { lg_default_params(&params_); }

void Legalizer::setTechPenalty(float v) { params_.tech_penalty = v; }
void Legalizer::setStandardAdmm(bool v) { params_.standard_admm = v ? 1 : 0; }
void Legalizer::setQpRefine(bool v) { params_.qp_admm_on = v ? 1 : 0; }
void Legalizer::setCoeffToInitRho(float v) { params_.coeff_to_init_rho = v; }
void Legalizer::setCoeffToInitDual(float v) { params_.coeff_to_init_dual = v; }
void Legalizer::setMaxAdmmIter(int v) { params_.max_admm_iter = v; }
void Legalizer::setFlagICCAD2017(bool v) { params_.flag_iccad17 = v ? 1 : 0; }
void Legalizer::setAdaptiveSearch(bool v) { params_.use_adaptive_search = v ? 1 : 0; }
void Legalizer::setMinIterToCallAdaptiveSearch(int v) { params_.min_iter_to_call_adaptive_search = v; }
void Legalizer::setOvfRatioToCallAdaptiveSearch(float v) { params_.ovf_ratio_to_call_adaptive_search = v; }
void Legalizer::setCoeffToNeglectDisplace(float v) { params_.coeff_to_neglect_displace = v; }
void Legalizer::setRandomSorting(bool v) {
  params_.sorting_policy = v ? LG_SORT_POLICY_RANDOM : LG_SORT_POLICY_OVERFLOW;
}
void Legalizer::setRandomSortingSeed(int v) { params_.random_sorting_seed = v; }

// Legalizer::setSortingPolicy / setSortingAlgorithm (Legalizer.cpp:161-189):
// unknown names fall back to the defaults
void Legalizer::setSortingPolicy(std::string_view s) {
  params_.sorting_policy = s == "RANDOM" ? LG_SORT_POLICY_RANDOM
                           : s == "SIZE" ? LG_SORT_POLICY_SIZE
                                         : LG_SORT_POLICY_OVERFLOW;
}
void Legalizer::setSortingAlgorithm(std::string_view s) {
  params_.sorting_algorithm = s == "S_BITONIC" ? LG_SORT_S_BITONIC
                              : s == "S_MERGE" ? LG_SORT_S_MERGE
                              : s == "M_MERGE" ? LG_SORT_M_MERGE
                              : s == "CUB"     ? LG_SORT_CUB
                                               : LG_SORT_M_BITONIC;
}

void Legalizer::setICCAD17Constraint(const std::string& file) { constraint_ = file; }
void Legalizer::setSizeByFile(const std::string& file) { size_file_ = file; }

int Legalizer::run(LgResult* result) {
  if (!design_) return lg_run_design(nullptr, &params_, result);
  LgDesign d = *design_;
  if (!constraint_.empty()) d.constraint_file = constraint_.c_str();
  if (!size_file_.empty()) d.size_file = size_file_.c_str();
  return lg_run_design(&d, &params_, result);
}

}  // namespace legalizer
