// Pixel grid build and preplace for gfx950 (lg_grid.h).
//
// Replaces the host loops of the design import for lg_run_design: markDbRow
// (Grid.cpp:316-342), markBlockageRegion / markGroupRegion (Grid.cpp:184-314)
// with the fence edge tests of handleRegionSegments (DatabaseIO.cpp:661-681)
// for fences whose boxes do not overlap, and initialLegalize
// (Preplace.cpp:18-72) with roundToNearsetPoint, findProjectionPoint and
// checkHopeless (Preplace.cpp:111-353).  The arithmetic is the host
// import's (integer), so the maps and positions are the same; only the
// pixel code the legalizer keeps (valid ? group + 2 : 0) is built.
//
// Kernels (P pixels, row-major):
//   k_grid_rows      one thread per pixel: valid from the row spans of its row
//   k_grid_owner     regions painted by atomicMax of their order index, so
//                    that the last region in order owns a pixel (as the host
//                    loop's overwrites); one workgroup row per region
//   k_grid_apply     one thread per pixel: the owner's pixel code
//   k_grid_edges     the edge columns / rows of each fence box
//   k_preplace       one thread per cell
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "lg_grid.h"

namespace {

constexpr int kT = 256;

struct G {
  int xs, ys, sw, rh, lx, ly;
  int core_lx, core_ly, core_ux, core_uy;
};

// the host import's grid helpers (Grid.h:51-105)
__device__ __forceinline__ int start_x(const G& g, int d) { return (d - g.lx) / g.sw; }
__device__ __forceinline__ int start_y(const G& g, int d) { return (d - g.ly) / g.rh; }
__device__ __forceinline__ int end_x(const G& g, int d) {
  const int q = (d - g.lx) / g.sw, r = (d - g.lx) % g.sw;
  return min(max(r == 0 ? q - 1 : q, 0), g.xs - 1);
}
__device__ __forceinline__ int end_y(const G& g, int d) {
  const int q = (d - g.ly) / g.rh, r = (d - g.ly) % g.rh;
  return min(max(r == 0 ? q - 1 : q, 0), g.ys - 1);
}
__device__ __forceinline__ int dbu_x(const G& g, int x) { return x * g.sw + g.lx; }
__device__ __forceinline__ int dbu_y(const G& g, int y) { return y * g.rh + g.ly; }

__global__ __launch_bounds__(kT) void k_grid_rows(G g, const int32_t* span_start, const int32_t* span,
                                                  int16_t* pcode) {
  const int x = blockIdx.x * kT + threadIdx.x, y = blockIdx.y;
  if (x >= g.xs) return;
  int v = 0;
  for (int s = span_start[y]; s < span_start[y + 1]; s++)
    if (span[2 * s] <= x && x <= span[2 * s + 1]) {
      v = 1;  // valid, default group (-1): code -1 + 2
      break;
    }
  pcode[(size_t)y * g.xs + x] = (int16_t)v;
}

// owner[p] = 1 + the largest index of a painted region covering p
__global__ __launch_bounds__(kT) void k_grid_owner(G g, const int32_t* paint, int32_t* owner) {
  const int k = blockIdx.y;
  const int x1 = paint[5 * k], y1 = paint[5 * k + 1], x2 = paint[5 * k + 2], y2 = paint[5 * k + 3];
  if (x2 < x1 || y2 < y1) return;
  const long long wdt = x2 - x1 + 1, n = wdt * (y2 - y1 + 1);
  for (long long q = blockIdx.x * (long long)kT + threadIdx.x; q < n; q += (long long)gridDim.x * kT) {
    const int x = x1 + (int)(q % wdt), y = y1 + (int)(q / wdt);
    if (x < 0 || y < 0 || x >= g.xs || y >= g.ys) continue;
    atomicMax(&owner[(size_t)y * g.xs + x], k + 1);
  }
}

__global__ __launch_bounds__(kT) void k_grid_apply(size_t P, const int32_t* owner, const int32_t* paint,
                                                   int16_t* pcode) {
  for (size_t p = blockIdx.x * (size_t)kT + threadIdx.x; p < P; p += (size_t)gridDim.x * kT) {
    const int o = owner[p];
    if (o > 0) pcode[p] = (int16_t)paint[5 * (o - 1) + 4];
  }
}

// the edge test of a fence box pixel with markGroupRegion's overlap counts
__device__ __forceinline__ void fence_edge(const G& g, const int32_t* f, int x, int y, int16_t* pcode) {
  const int ex1 = f[0], ey1 = f[1], ex2 = f[2], ey2 = f[3];
  const int sx1 = f[4], sy1 = f[5], sx2 = f[6], sy2 = f[7];
  const int rlx = f[8], rly = f[9], rux = f[10], ruy = f[11];
  int h = g.sw, v = g.rh;
  if (x == ex1 && ex1 == sx1 - 1) h += (sx1 * g.sw - rlx) - g.sw;
  if (x == ex2 && ex2 == sx2 + 1) h += (rux - ex2 * g.sw) - g.sw;
  if (y == ey1 && ey1 == sy1 - 1) v += (sy1 * g.rh - rly) - g.rh;
  if (y == ey2 && ey2 == sy2 + 1) v += (ruy - ey2 * g.rh) - g.rh;
  if ((h > 0 || v > 0) && (h < g.sw || v < g.rh) && x >= 0 && y >= 0 && x < g.xs && y < g.ys)
    pcode[(size_t)y * g.xs + x] = 0;
}

__global__ __launch_bounds__(kT) void k_grid_edges(G g, const int32_t* fence, int16_t* pcode) {
  const int32_t* f = fence + 12 * blockIdx.x;
  const int ex1 = f[0], ey1 = f[1], ex2 = f[2], ey2 = f[3];
  for (int y = ey1 + (int)threadIdx.x; y <= ey2; y += kT) {
    fence_edge(g, f, ex1, y, pcode);
    fence_edge(g, f, ex2, y, pcode);
  }
  for (int x = ex1 + (int)threadIdx.x; x <= ex2; x += kT) {
    fence_edge(g, f, x, ey1, pcode);
    fence_edge(g, f, x, ey2, pcode);
  }
}

struct Cell {
  int lx, ly, w, h, ilx, ily, group, gnd;
};

struct Pre {
  G g;
  const int16_t* pcode;
  const int32_t* vdd;
  const int32_t* region;
  int group_min, num_lists;
  const int32_t* glist_start;
  const int32_t* glist;
  const int32_t* dirs;
  int num_dirs;
};

// checkPixelStatus (Grid.cpp:365-383): every covered pixel valid and of the
// cell's group (a pixel outside the grid fails)
__device__ bool pixels_ok(const Pre& a, const Cell& c) {
  const G& g = a.g;
  const int x1 = start_x(g, c.lx), y1 = start_y(g, c.ly);
  const int x2 = end_x(g, c.lx + c.w), y2 = end_y(g, c.ly + c.h);
  const int want = c.group + 2;
  for (int y = y1; y <= y2; y++)
    for (int x = x1; x <= x2; x++) {
      if (x < 0 || y < 0 || x >= g.xs || y >= g.ys) return false;
      if (a.pcode[(size_t)y * g.xs + x] != want) return false;
    }
  return true;
}

__device__ __forceinline__ bool outside(const int32_t* r, const Cell& c, int lx, int ly) {  // Preplace.cpp:244-265
  return !(lx >= r[0] && ly >= r[1] && lx + c.w <= r[2] && ly + c.h <= r[3]);
}

// findProjectionPoint (Preplace.cpp:267-353): (lx, ly, disp), disp INT_MAX if none
__device__ int3 project(const Pre& a, const Cell& c, const int32_t* r) {
  const G& g = a.g;
  const int wg = c.w / g.sw, hg = c.h / g.rh;
  const int gx1 = end_x(g, r[0] + g.sw), gy1 = end_y(g, r[1] + g.rh);
  const int gx2 = start_x(g, r[2] - g.sw), gy2 = start_y(g, r[3] - g.rh);
  if (gy1 > gy2 || gx1 > gx2) return make_int3(-1, -1, INT_MAX);
  const int rlx = dbu_x(g, gx1), rly = dbu_y(g, gy1), rux = dbu_x(g, gx2), ruy = dbu_y(g, gy2);
  int nx = c.lx, ny = c.ly;
  if (c.lx + c.w > rux) nx = dbu_x(g, gx2 - wg + 1);
  else if (c.lx < rlx) nx = dbu_x(g, gx1);
  if (c.ly + c.h > ruy) ny = dbu_y(g, gy2 - hg + 1);
  else if (c.ly < rly) ny = dbu_y(g, gy1);
  const bool even = (hg % 2) == 0;
  if (even) {
    const int sy = start_y(g, ny);
    if (sy < 0 || sy >= g.ys || a.vdd[sy] != c.gnd) {
      int best = INT_MAX, by = -1;
      for (int k = 0; k < 2; k++) {
        const int yc = k == 0 ? ny - g.rh : ny + g.rh;
        const int dd = abs(nx - c.ilx) + abs(yc - c.ily);
        if (!outside(r, c, nx, yc) && dd < best) {
          best = dd;
          by = yc;
        }
      }
      ny = by;
    }
  }
  const bool out = outside(r, c, nx, ny);
  bool bad = false;
  if (even && !out) {
    const int x0 = start_x(g, nx), y0 = start_y(g, ny);
    for (int x = x0; x < x0 + wg && !bad; x++)
      for (int y = y0; y < y0 + hg; y++) {
        if (x < 0 || y < 0 || x >= g.xs || y >= g.ys || a.pcode[(size_t)y * g.xs + x] == 0) {
          bad = true;
          break;
        }
      }
  }
  if (out || bad) return make_int3(-1, -1, INT_MAX);
  return make_int3(nx, ny, abs(c.ilx - nx) + abs(c.ily - ny));
}

__device__ bool project_group(const Pre& a, Cell& c) {
  const int li = c.group - a.group_min;
  if (li < 0 || li >= a.num_lists) return false;
  const int b = a.glist_start[li], e = a.glist_start[li + 1];
  if (b >= e) return false;
  int3 best = project(a, c, a.region + 4 * a.glist[b]);
  for (int k = b + 1; k < e; k++) {
    const int3 q = project(a, c, a.region + 4 * a.glist[k]);
    if (q.z < best.z) best = q;
  }
  if (best.z >= INT_MAX) return false;
  c.lx = best.x;
  c.ly = best.y;
  return true;
}

// roundToNearsetPoint (Preplace.cpp:111-157)
__device__ void snap(const Pre& a, Cell& c) {
  const G& g = a.g;
  const int wg = c.w / g.sw, hg = c.h / g.rh;
  const int gx = min(g.xs - wg, max(0, start_x(g, c.lx)));
  const int ax = dbu_x(g, gx), bx = dbu_x(g, min(g.xs - wg, gx + 1));
  const int nx = abs(c.lx - ax) < abs(c.lx - bx) ? ax : bx;
  const int gy = min(g.ys - hg, max(0, start_y(g, c.ly)));
  int ay = dbu_y(g, gy), by = dbu_y(g, min(g.ys - hg, gy + 1));
  if (ay == by) by = dbu_y(g, max(0, gy - 1));
  int ny = abs(c.ly - ay) < abs(c.ly - by) ? ay : by;
  if (hg % 2 == 0) {
    const int sy = start_y(g, ny);
    if (sy >= 0 && sy < g.ys && a.vdd[sy] != c.gnd) {
      if (ny == ay) ny = (by + c.h <= g.core_uy) ? by : ay - g.rh;
      else if (ny == by) ny = (ay >= g.core_ly) ? ay : by + g.rh;
    }
  }
  c.lx = nx;
  c.ly = ny;
}

// checkHopeless (Preplace.cpp:159-207) over the default moves
__device__ bool hopeless(const Pre& a, const Cell& c0) {
  const G& g = a.g;
  Cell c = c0;
  const int gx = start_x(g, c0.lx), gy = start_y(g, c0.ly);
  for (int k = 0; k < a.num_dirs; k++) {
    const int nx = gx + a.dirs[2 * k], ny = gy + a.dirs[2 * k + 1];
    c.lx = dbu_x(g, nx);
    c.ly = dbu_y(g, ny);
    if (c.lx + c.w > g.core_ux || c.ly + c.h > g.core_uy || c.lx < g.core_lx || c.ly < g.core_ly) continue;
    if (ny < 0 || ny >= g.ys) continue;
    if (a.vdd[ny] == c.gnd && pixels_ok(a, c)) return false;
  }
  return true;
}

__global__ __launch_bounds__(kT) void k_preplace(Pre a, int n, const int32_t* cells, int32_t* out_lx, int32_t* out_ly,
                                                 int* fails) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const int32_t* q = cells + 8 * (size_t)i;
  Cell c{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]};
  snap(a, c);
  if (!pixels_ok(a, c)) project_group(a, c);
  if (hopeless(a, c) && !project_group(a, c)) atomicAdd(fails, 1);
  out_lx[i] = c.lx;
  out_ly[i] = c.ly;
}

template <typename T>
int dup(std::vector<void*>& bufs, T** d, const T* src, size_t n, hipStream_t s, std::string& err) {
  if (hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    err = "hipMalloc";
    return LG_ERR_MEMORY;
  }
  bufs.push_back(*d);
  if (n && src && hipMemcpyAsync(*d, src, n * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) {
    err = "hipMemcpyAsync";
    return LG_ERR_DEVICE;
  }
  return LG_OK;
}

}  // namespace

int lg_grid_preplace_device(const LgGridBuild& in, void** d_pcode_out, int32_t* out_lx, int32_t* out_ly, int* fails,
                            void* stream, std::string& err) {
  hipStream_t s = (hipStream_t)stream;
  *d_pcode_out = nullptr;
  std::vector<void*> tmp;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(s);
    for (void* b : tmp) (void)hipFree(b);
    return code;
  };
  const G g{in.xs, in.ys, in.sw, in.rh, in.lx, in.ly, in.core_lx, in.core_ly, in.core_ux, in.core_uy};
  const size_t P = (size_t)in.xs * in.ys;
  int16_t* pcode = nullptr;
  if (hipMalloc((void**)&pcode, std::max<size_t>(P, 1) * sizeof(int16_t)) != hipSuccess) {
    err = "hipMalloc of the pixel codes";
    return LG_ERR_MEMORY;
  }
  const int nspan = in.span_start[in.ys];
  int32_t *d_ss, *d_sp, *d_paint = nullptr, *d_owner = nullptr, *d_fence = nullptr, *d_vdd, *d_cell, *d_reg, *d_gs,
          *d_gl, *d_dirs, *d_olx, *d_oly;
  int* d_fail;
  int st;
  auto fail_all = [&](int code) {
    (void)hipFree(pcode);
    return done(code);
  };
  if ((st = dup(tmp, &d_ss, in.span_start, (size_t)in.ys + 1, s, err)) ||
      (st = dup(tmp, &d_sp, in.span, 2 * (size_t)nspan, s, err)) ||
      (st = dup(tmp, &d_vdd, in.vdd, (size_t)in.ys, s, err)) ||
      (st = dup(tmp, &d_cell, in.cell, 8 * (size_t)in.num_cells, s, err)) ||
      (st = dup(tmp, &d_reg, in.region, 4 * (size_t)in.num_regions, s, err)) ||
      (st = dup(tmp, &d_gs, in.glist_start, (size_t)in.num_group_lists + 1, s, err)) ||
      (st = dup(tmp, &d_gl, in.glist, (size_t)std::max(in.glist_start[in.num_group_lists], 0), s, err)) ||
      (st = dup(tmp, &d_dirs, in.dirs, 2 * (size_t)in.num_dirs, s, err)) ||
      (st = dup(tmp, &d_olx, (const int32_t*)nullptr, (size_t)in.num_cells, s, err)) ||
      (st = dup(tmp, &d_oly, (const int32_t*)nullptr, (size_t)in.num_cells, s, err)) ||
      (st = dup(tmp, &d_fail, (const int*)nullptr, 1, s, err)))
    return fail_all(st);
  if (P > 0) hipLaunchKernelGGL(k_grid_rows, dim3((in.xs + kT - 1) / kT, in.ys), dim3(kT), 0, s, g, d_ss, d_sp, pcode);
  if (in.num_paint > 0) {
    if ((st = dup(tmp, &d_paint, in.paint, 5 * (size_t)in.num_paint, s, err)) ||
        (st = dup(tmp, &d_owner, (const int32_t*)nullptr, P, s, err)))
      return fail_all(st);
    if (hipMemsetAsync(d_owner, 0, P * sizeof(int32_t), s) != hipSuccess) return fail_all(LG_ERR_DEVICE);
    hipLaunchKernelGGL(k_grid_owner, dim3(64, in.num_paint), dim3(kT), 0, s, g, d_paint, d_owner);
    hipLaunchKernelGGL(k_grid_apply, dim3((unsigned)std::min<size_t>((P + kT - 1) / kT, 8192)), dim3(kT), 0, s, P,
                       d_owner, d_paint, pcode);
  }
  if (in.num_fences > 0) {
    if ((st = dup(tmp, &d_fence, in.fence, 12 * (size_t)in.num_fences, s, err))) return fail_all(st);
    hipLaunchKernelGGL(k_grid_edges, dim3(in.num_fences), dim3(kT), 0, s, g, d_fence, pcode);
  }
  if (hipMemsetAsync(d_fail, 0, sizeof(int), s) != hipSuccess) return fail_all(LG_ERR_DEVICE);
  if (in.num_cells > 0) {
    Pre a{g, pcode, d_vdd, d_reg, in.group_min, in.num_group_lists, d_gs, d_gl, d_dirs, in.num_dirs};
    hipLaunchKernelGGL(k_preplace, dim3((in.num_cells + kT - 1) / kT), dim3(kT), 0, s, a, in.num_cells, d_cell,
                       d_olx, d_oly, d_fail);
  }
  if (hipGetLastError() != hipSuccess) {
    err = "grid / preplace kernel launch";
    return fail_all(LG_ERR_DEVICE);
  }
  const size_t nb = (size_t)in.num_cells * sizeof(int32_t);
  if ((nb && (hipMemcpyAsync(out_lx, d_olx, nb, hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(out_ly, d_oly, nb, hipMemcpyDeviceToHost, s) != hipSuccess)) ||
      hipMemcpyAsync(fails, d_fail, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "grid / preplace copies";
    return fail_all(LG_ERR_DEVICE);
  }
  *d_pcode_out = pcode;
  return done(LG_OK);
}

int lg_grid_pcode_to_host(const void* d_pcode, size_t num_pixels, int16_t* out, std::string& err) {
  if (num_pixels && hipMemcpy(out, d_pcode, num_pixels * sizeof(int16_t), hipMemcpyDeviceToHost) != hipSuccess) {
    err = "hipMemcpy of the pixel codes";
    return LG_ERR_DEVICE;
  }
  return LG_OK;
}

void lg_grid_free(void* d_pcode) {
  if (d_pcode) (void)hipFree(d_pcode);
}
