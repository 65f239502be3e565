// A C++ caller of the device layer of include/skyline_lg.h, as the
// reference's Legalizer drives CudaDatabase / CudaAdmmLegalizer /
// CudaPostProcess (Legalizer.cpp:197-228): host import and preplace
// (lg_import_and_preplace), then lg_create, lg_solve, lg_postprocess,
// lg_check, lg_copy_out, lg_get_stats and lg_destroy on the GPU.
//
//   device_caller <design.lgd> <out.bin>
//
// out.bin: int32 n, then the n movable cells' grid x, grid y and rotation
// flags (int32 each).  Prints one line of statistics.
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../../admm_legalization_amd/csrc/lgd_file.hpp"

int main(int argc, char** argv) {
  if (argc != 3) { std::printf("usage: %s <design.lgd> <out.bin>\n", argv[0]); return 1; }
  lgd::DesignFile f;
  std::string err;
  if (!lgd::read(argv[1], f, err)) { std::printf("%s\n", err.c_str()); return 1; }
  const LgDesign& d = f.d;
  LgParams p;
  lg_default_params(&p);
  int32_t shape[4];
  if (lg_import_and_preplace(&d, &p, shape, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)) {
    std::printf("import: %s\n", lg_last_error());
    return 1;
  }
  const int xs = shape[0], ys = shape[1], n = shape[2], nr = shape[3];
  const size_t P = (size_t)xs * ys;
  std::vector<int32_t> valid(P), group(P), region(P), vdd(ys), rb(4 * std::max(nr, 1)), clx(std::max(n, 1)),
      cly(std::max(n, 1));
  if (lg_import_and_preplace(&d, &p, shape, valid.data(), group.data(), region.data(), vdd.data(), rb.data(),
                             clx.data(), cly.data())) {
    std::printf("import: %s\n", lg_last_error());
    return 1;
  }
  // the movable cells, in instance order (importCells, DatabaseIO.cpp:100-149)
  const int row_h = d.row_dy[d.num_rows - 1], sw = d.row_site_width[d.num_rows - 1];
  int rh = row_h;
  for (int r = 0; r < d.num_rows; r++) rh = std::min(rh, d.row_dy[r]);
  std::vector<int32_t> ox, oy, w, h, grp, gnd;
  int max_h = 1;
  for (int i = 0; i < d.num_insts; i++) {
    if (d.inst_dx[i] == 0 || d.inst_dy[i] < row_h || d.inst_macro[i] || d.inst_fixed[i]) continue;
    ox.push_back(d.inst_lx[i]);
    oy.push_back(d.inst_ly[i]);
    w.push_back(d.inst_dx[i] / sw);
    h.push_back(d.inst_dy[i] / rh);
    max_h = std::max(max_h, d.inst_dy[i] / rh);
    grp.push_back(d.inst_group ? d.inst_group[i] : -1);
    gnd.push_back(d.inst_ground_at_bottom[i] ? 1 : 0);
  }
  if ((int)w.size() != n) { std::printf("movable cell count mismatch\n"); return 1; }
  p.y_hint = std::max(p.y_hint, 3 * max_h);  // adjustYHint (Legalizer.cpp:230-247)
  std::vector<int32_t> rx0(nr), ry0(nr), rx1(nr), ry1(nr);
  for (int k = 0; k < nr; k++) { rx0[k] = rb[4 * k]; ry0[k] = rb[4 * k + 1]; rx1[k] = rb[4 * k + 2]; ry1[k] = rb[4 * k + 3]; }
  LgGridDesc g{xs, ys, sw, rh, d.core_lx, d.core_ly, valid.data(), group.data(), region.data(), vdd.data(), nr,
               rx0.data(), ry0.data(), rx1.data(), ry1.data()};
  LgCellsSoA c{n, ox.data(), oy.data(), clx.data(), cly.data(), w.data(), h.data(), grp.data(), gnd.data()};
  LgHandle* hd = nullptr;
  LgVioStat vio{};
  LgStats st{};
  std::vector<int32_t> gx(std::max(n, 1)), gy(std::max(n, 1)), rot(std::max(n, 1));
  int e = lg_create(&g, &c, &p, &hd);
  if (!e) e = lg_solve(hd, nullptr);
  if (!e) e = lg_postprocess(hd, nullptr);
  if (!e) e = lg_check(hd, &vio);
  if (!e) e = lg_copy_out(hd, gx.data(), gy.data(), rot.data());
  if (!e) e = lg_get_stats(hd, &st);
  if (e) {
    std::printf("device layer: error %d: %s\n", e, lg_last_error());
    lg_destroy(hd);
    return 1;
  }
  lg_destroy(hd);
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 1;
  std::fwrite(&n, 4, 1, out);
  std::fwrite(gx.data(), 4, n, out);
  std::fwrite(gy.data(), 4, n, out);
  std::fwrite(rot.data(), 4, n, out);
  std::fclose(out);
  std::printf("cells=%d iterations=%d converged=%d overlap=%d site_vio=%d fence_vio=%d total_disp_dbu=%lld\n", n,
              st.iterations, st.converged, vio.total_ovf, vio.num_site_vio, vio.num_fence_vio,
              (long long)st.total_disp_dbu);
  return 0;
}
