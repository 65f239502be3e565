// Host setup of the density path: bins, density sizes, fixed area
// (SkyPlaceDB::createBins / updateDensitySize / updateFixedOverlapArea,
// SkyPlaceDB.cpp:446-549, 821-852, 909-938), in FP32 as the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/skyline_density.h"

namespace densdev {
int fail(int code, const std::string& msg);  // sets dens_last_error()
}

namespace {

int prep_fail(const std::string& m) { return densdev::fail(DENS_ERR_ARG, m); }

// bin walk of createBins (SkyPlaceDB.cpp:504-546): nominal width die/num,
// the last column (row) ends at the die edge, bounds accumulated in FP32
void walk(float lo, float hi, int n, float size, std::vector<float>& l, std::vector<float>& u) {
  l.resize(n);
  u.resize(n);
  float x = lo;
  for (int c = 0; c < n; c++) {
    float w = size;
    if (c == n - 1 || x + size > hi) w = hi - x;
    l[c] = x;
    u[c] = x + w;
    x += w;
  }
}

// createBins (SkyPlaceDB.cpp:446-499): power-of-two bin count per side from the
// average movable area, at most 1024 (x the aspect ratio 1, 2 or 4)
void auto_counts(const DensDesign* d, float sum_movable_area, int num_movable, int& nbx, int& nby) {
  const float die_area = (d->die_ux - d->die_lx) * (d->die_uy - d->die_ly);
  const float avg = sum_movable_area / (float)num_movable;
  const float ideal_area = avg / d->target_density;
  const float ideal_count = die_area / ideal_area;
  const int ideal = (int)std::ceil(std::sqrt(ideal_count));
  const float dx = d->die_ux - d->die_lx, dy = d->die_uy - d->die_ly;
  const float aspect = std::max(dx / dy, dy / dx);
  const int ratio = aspect >= 4.0f ? 4 : (aspect >= 2.0f ? 2 : 1);
  int nb = 4;
  while (nb * nb * ratio < ideal * ideal) {
    nb *= 2;
    if (nb >= 1024) break;
  }
  const int e1 = std::abs(ideal * ideal - nb * nb * ratio);
  const int e2 = std::abs(ideal * ideal - nb * nb * ratio / 4);
  if (e2 < e1) nb /= 2;
  if (dx > dy) {
    nbx = nb * ratio;
    nby = nb;
  } else {
    nbx = nb;
    nby = nb * ratio;
  }
}

}  // namespace

extern "C" int dens_prepare(const DensDesign* d, int32_t* shape, float* bin_lx, float* bin_ux, float* bin_ly,
                            float* bin_uy, float* fixed_area, float* scaled_bin_area, float* cell_dw,
                            float* cell_dh, float* cell_scale, float* scalars) {
  if (!d || !shape) return prep_fail("null argument");
  if (d->num_movable <= 0 || !d->cell_dx || !d->cell_dy) return prep_fail("no movable cells");
  if (!(d->die_ux > d->die_lx) || !(d->die_uy > d->die_ly)) return prep_fail("empty die");
  if (!(d->target_density > 0.f)) return prep_fail("target density must be positive");
  const int N = d->num_movable;
  // sumMovableArea_ (SkyPlaceDB.cpp:197), FP32 accumulation, over the cells
  // that are not fillers (importDB runs before createFillers)
  float sum_area = 0.f;
  int n_real = 0;
  for (int i = 0; i < N; i++) {
    if (d->is_filler && d->is_filler[i]) continue;
    sum_area += d->cell_dx[i] * d->cell_dy[i];
    n_real++;
  }
  if (n_real == 0) return prep_fail("no movable cells besides fillers");
  int nbx, nby;
  if (d->num_bin_x > 0 && d->num_bin_y > 0) {
    nbx = d->num_bin_x;
    nby = d->num_bin_y;
  } else {
    auto_counts(d, sum_area, n_real, nbx, nby);
  }
  shape[0] = nbx;
  shape[1] = nby;
  if (!bin_lx) return DENS_OK;
  if (!bin_ux || !bin_ly || !bin_uy || !fixed_area || !scaled_bin_area || !cell_dw || !cell_dh || !cell_scale ||
      !scalars)
    return prep_fail("null output array");
  const float bw = (d->die_ux - d->die_lx) / (float)nbx;
  const float bh = (d->die_uy - d->die_ly) / (float)nby;
  std::vector<float> lx, ux, ly, uy;
  walk(d->die_lx, d->die_ux, nbx, bw, lx, ux);
  walk(d->die_ly, d->die_uy, nby, bh, ly, uy);
  std::copy(lx.begin(), lx.end(), bin_lx);
  std::copy(ux.begin(), ux.end(), bin_ux);
  std::copy(ly.begin(), ly.end(), bin_ly);
  std::copy(uy.begin(), uy.end(), bin_uy);
  const size_t B = (size_t)nbx * nby;
  for (int y = 0; y < nby; y++)
    for (int x = 0; x < nbx; x++)
      scaled_bin_area[(size_t)y * nbx + x] = (ux[x] - lx[x]) * (uy[y] - ly[y]) * d->target_density;
  // updateFixedOverlapArea (SkyPlaceDB.cpp:909-938) with findBin (665-690)
  std::fill(fixed_area, fixed_area + B, 0.f);
  for (int f = 0; f < d->num_fixed; f++) {
    const float flx = d->fixed_lx[f], fly = d->fixed_ly[f], fux = d->fixed_ux[f], fuy = d->fixed_uy[f];
    if (fux <= d->die_lx || flx >= d->die_ux || fuy <= d->die_ly || fly >= d->die_uy) continue;
    int x0 = std::max((int)std::floor((flx - d->die_lx) / bw), 0);
    int x1 = std::min((int)std::ceil((fux - d->die_lx) / bw), nbx);
    int y0 = std::max((int)std::floor((fly - d->die_ly) / bh), 0);
    int y1 = std::min((int)std::ceil((fuy - d->die_ly) / bh), nby);
    for (int x = x0; x < x1; x++)
      for (int y = y0; y < y1; y++) {
        float rlx = std::max(lx[x], flx), rly = std::max(ly[y], fly);
        float rux = std::min(ux[x], fux), ruy = std::min(uy[y], fuy);
        if (rlx >= rux || rly >= ruy) continue;
        fixed_area[(size_t)y * nbx + x] += (rux - rlx) * (ruy - rly) * d->target_density;
      }
  }
  // updateDensitySize (SkyPlaceDB.cpp:821-852); DENSITY_SCALE is the long
  // double literal sqrt(2), so the threshold and the ratio are long double
  const long double sqrt2 = 1.414213562373095048801L;
  const long double tx = sqrt2 * bw, ty = sqrt2 * bh;
  for (int i = 0; i < N; i++) {
    float sx, sy, w, h;
    if (d->cell_dx[i] > tx) { sx = 1.f; w = d->cell_dx[i]; }
    else { sx = (float)(d->cell_dx[i] / tx); w = (float)tx; }
    if (d->cell_dy[i] > ty) { sy = 1.f; h = d->cell_dy[i]; }
    else { sy = (float)(d->cell_dy[i] / ty); h = (float)ty; }
    cell_dw[i] = w;
    cell_dh[i] = h;
    cell_scale[i] = sx * sy;
  }
  scalars[0] = bw;
  scalars[1] = bh;
  scalars[2] = sum_area;
  return DENS_OK;
}
