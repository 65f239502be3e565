// Binary design file of the command driver (skyline_cmd) and the C++ caller
// tests: an LgDesign (include/skyline_lg.h) on disk, so that a command
// sequence can run outside Python.  Written by
// admm_legalization_amd.design.Design.save_lgd.  No counterpart in the
// reference, which reads LEF/DEF into its db::dbDatabase.
//
// Layout (little endian): "LGD1", then int32 dbu_per_micron, core lx ly ux
// uy; int32 num_rows and six int32 row arrays (lx ly dx dy site_width
// orient); int32 num_insts and the instance arrays lx ly dx dy (int32),
// fixed macro (uint8), group (int32), ground_at_bottom (uint8); int32
// num_groups, num_boxes and five int32 box arrays (group lx ly ux uy); int32
// num_nets, net_start[num_nets + 1], int32 num_pins, pin_inst pin_dx pin_dy;
// int32 has_names and, if 1, per instance an int32 length and the bytes.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skyline_lg.h"

namespace lgd {

struct DesignFile {
  LgDesign d{};
  std::vector<int32_t> row_lx, row_ly, row_dx, row_dy, row_sw, row_or;
  std::vector<int32_t> lx, ly, dx, dy, group;
  std::vector<uint8_t> fixed, macro, gnd;
  std::vector<int32_t> box_group, box_lx, box_ly, box_ux, box_uy;
  std::vector<int32_t> net_start, pin_inst, pin_dx, pin_dy;
  std::vector<std::string> names;
  std::vector<const char*> name_ptrs;
};

class Reader {
 public:
  explicit Reader(FILE* f) : f_(f) {}
  bool ok() const { return ok_; }
  int32_t i32() {
    int32_t v = 0;
    ok_ = ok_ && std::fread(&v, 4, 1, f_) == 1;
    return v;
  }
  template <typename T>
  void arr(std::vector<T>& v, int64_t n) {
    if (n < 0 || n > (int64_t)1 << 31) ok_ = false;
    if (!ok_) return;
    v.resize((size_t)n);
    if (n) ok_ = std::fread(v.data(), sizeof(T), (size_t)n, f_) == (size_t)n;
  }

 private:
  FILE* f_;
  bool ok_ = true;
};

// Reads `path` into `out` (its LgDesign points into its own vectors).
inline bool read(const char* path, DesignFile& out, std::string& err) {
  FILE* f = std::fopen(path, "rb");
  if (!f) { err = std::string("cannot open ") + path; return false; }
  char magic[4];
  if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "LGD1", 4) != 0) {
    std::fclose(f);
    err = std::string(path) + ": not an LGD1 design file";
    return false;
  }
  Reader r(f);
  DesignFile& o = out;
  LgDesign& d = o.d;
  d.dbu_per_micron = r.i32();
  d.core_lx = r.i32(); d.core_ly = r.i32(); d.core_ux = r.i32(); d.core_uy = r.i32();
  d.num_rows = r.i32();
  r.arr(o.row_lx, d.num_rows); r.arr(o.row_ly, d.num_rows); r.arr(o.row_dx, d.num_rows);
  r.arr(o.row_dy, d.num_rows); r.arr(o.row_sw, d.num_rows); r.arr(o.row_or, d.num_rows);
  d.num_insts = r.i32();
  r.arr(o.lx, d.num_insts); r.arr(o.ly, d.num_insts); r.arr(o.dx, d.num_insts); r.arr(o.dy, d.num_insts);
  r.arr(o.fixed, d.num_insts); r.arr(o.macro, d.num_insts); r.arr(o.group, d.num_insts);
  r.arr(o.gnd, d.num_insts);
  d.num_groups = r.i32();
  d.num_region_boxes = r.i32();
  r.arr(o.box_group, d.num_region_boxes); r.arr(o.box_lx, d.num_region_boxes);
  r.arr(o.box_ly, d.num_region_boxes); r.arr(o.box_ux, d.num_region_boxes);
  r.arr(o.box_uy, d.num_region_boxes);
  d.num_nets = r.i32();
  r.arr(o.net_start, (int64_t)d.num_nets + 1);
  const int32_t np = r.i32();
  r.arr(o.pin_inst, np); r.arr(o.pin_dx, np); r.arr(o.pin_dy, np);
  const int32_t has_names = r.i32();
  if (r.ok() && has_names) {
    o.names.resize((size_t)d.num_insts);
    for (int i = 0; i < d.num_insts && r.ok(); i++) {
      const int32_t n = r.i32();
      std::vector<char> b;
      r.arr(b, n);
      o.names[i].assign(b.begin(), b.end());
    }
  }
  std::fclose(f);
  if (!r.ok()) { err = std::string(path) + ": truncated or corrupt"; return false; }
  d.row_lx = o.row_lx.data(); d.row_ly = o.row_ly.data(); d.row_dx = o.row_dx.data();
  d.row_dy = o.row_dy.data(); d.row_site_width = o.row_sw.data(); d.row_orient = o.row_or.data();
  d.inst_lx = o.lx.data(); d.inst_ly = o.ly.data(); d.inst_dx = o.dx.data(); d.inst_dy = o.dy.data();
  d.inst_fixed = o.fixed.data(); d.inst_macro = o.macro.data(); d.inst_group = o.group.data();
  d.inst_ground_at_bottom = o.gnd.data();
  d.box_group = o.box_group.data(); d.box_lx = o.box_lx.data(); d.box_ly = o.box_ly.data();
  d.box_ux = o.box_ux.data(); d.box_uy = o.box_uy.data();
  d.net_start = o.net_start.data(); d.pin_inst = o.pin_inst.data();
  d.pin_dx = o.pin_dx.data(); d.pin_dy = o.pin_dy.data();
  if (!o.names.empty()) {
    o.name_ptrs.resize(o.names.size());
    for (size_t i = 0; i < o.names.size(); i++) o.name_ptrs[i] = o.names[i].c_str();
    d.inst_name = o.name_ptrs.data();
  }
  return true;
}

// Result file "LGR1": int32 num_insts, then lx, ly (int32) and flipped
// (uint8) per instance.
inline bool write_result(const char* path, int n, const int32_t* lx, const int32_t* ly, const uint8_t* fl) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite("LGR1", 1, 4, f) == 4 && std::fwrite(&n, 4, 1, f) == 1;
  if (n > 0)
    ok = ok && std::fwrite(lx, 4, n, f) == (size_t)n && std::fwrite(ly, 4, n, f) == (size_t)n &&
         std::fwrite(fl, 1, n, f) == (size_t)n;
  return std::fclose(f) == 0 && ok;
}

}  // namespace lgd
