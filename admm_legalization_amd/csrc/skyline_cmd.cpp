// skyline_cmd: a plain C++ command driver with the reference's Tcl command
// surface for the legalization flow, for hosts without Tcl.  It reads a
// script of commands (one per line or separated by ';', '#' starts a
// comment) and runs them in order:
//
//   read_lef <file.lef>            CmdReadLef: technology / cell LEF (repeatable)
//   read_def <file.def>            CmdReadDef: the design (after the LEF files)
//   write_def <file.def>           CmdWriteDef: the DEF read, COMPONENTS at the
//                                  legalized (else the current) positions
//   read_design <file.lgd>         design in the LGD1 format (lgd_file.hpp)
//   global_place [options]         CmdGlobalPlace (tclCmd/CmdGlobalPlace.hpp:11-93):
//                                  the placed positions replace the design's
//   legalize [options]             CmdLegalize (tclCmd/CmdLegalize.hpp:14-152)
//   write_result <file.lgr>        legalized positions and flips (LGR1)
//
// global_place takes the reference's options -target_density <f>
// -target_overflow <f> -initial_place <random> -initial_place_only
// -initial_lambda <f> -initial_gamma_inv <f> -solver <nesterov|adam> -plot
// <dir> (accepted, nothing drawn), and these, which have no counterpart in
// the reference's command: -bins <n> (n x n bins), -scaling_factor <f>
// (coordinates divided by f), -ref_hpwl <f>, -ref_hpwl_ratio <f>,
// -precision <BF16|SPLIT_BF16|F32>, -max_iter <n>, -seed <n>.
//
// legalize takes the reference's options verbatim: -constraint <file>
// -tech_penalty <f> -max_iter <i> -std_admm -coeff_to_init_rho <f>
// -coeff_to_init_dual <f> -qp_refine -size_file <file> -use_adaptive_search
// -min_iter_to_call_adaptive_search <i> -ovf_ratio_to_call_adaptive_search <f>
// -coeff_to_neglect_displace <f> -random_sorting_seed <i> -sorting_algorithm
// <S_MERGE|M_MERGE|S_BITONIC|M_BITONIC|CUB> -sorting_policy
// <RANDOM|SIZE|OVERFLOW>.  The option grammar is CmdCore's (CmdCore.cpp:
// 35-84): an option begins with '-', takes at most one argument, and an
// unknown option or an option where an argument is due is an error.
//
// Usage: skyline_cmd <script> | skyline_cmd -c "<commands>"
// Exit status 0 when every command succeeded, else 1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/skyline/Legalizer.hpp"
#include "../../include/skyline/SkyPlace.hpp"
#include "../../include/skyline_lefdef.h"
#include "lgd_file.hpp"

namespace {

struct Option {
  int nargs = 0;
  bool valid = false;
  std::string arg;
};

// A command's option table and the CmdCore parsing rules.
struct Command {
  std::map<std::string, Option> opts;
  std::vector<std::string> args;
  void add(const char* name, int nargs) { opts[name] = Option{nargs, false, {}}; }
  bool parse(const std::vector<std::string>& tok, std::string& err) {
    Option* cur = nullptr;
    bool want_arg = false;
    for (size_t i = 1; i < tok.size(); i++) {
      const std::string& t = tok[i];
      if (!t.empty() && t[0] == '-') {
        auto it = opts.find(t);
        if (it == opts.end()) { err = "Unknown Option : " + t; return false; }
        if (want_arg) { err = "Syntax error in option argument"; return false; }
        cur = &it->second;
        cur->valid = true;
        want_arg = cur->nargs > 0;
      } else if (want_arg) {
        cur->arg = t;
        cur = nullptr;
        want_arg = false;
      } else {
        args.push_back(t);
      }
    }
    if (want_arg) { err = "Syntax error in command " + tok[0]; return false; }
    return true;
  }
  const Option* get(const char* name) const {
    auto it = opts.find(name);
    return it != opts.end() && it->second.valid ? &it->second : nullptr;
  }
};

// A working copy of a design read from LEF / DEF: the core arrays in the
// DesignFile's own vectors (global_place replaces the positions), the
// names and technology arrays pointing into the LEF / DEF database.
void copy_design(const LgDesign& s, lgd::DesignFile& f) {
  auto cp = [](auto* p, int n, auto& v) { v.assign(p, p + n); };
  const int nr = s.num_rows, n = s.num_insts, nb = s.num_region_boxes;
  cp(s.row_lx, nr, f.row_lx); cp(s.row_ly, nr, f.row_ly); cp(s.row_dx, nr, f.row_dx);
  cp(s.row_dy, nr, f.row_dy); cp(s.row_site_width, nr, f.row_sw); cp(s.row_orient, nr, f.row_or);
  cp(s.inst_lx, n, f.lx); cp(s.inst_ly, n, f.ly); cp(s.inst_dx, n, f.dx); cp(s.inst_dy, n, f.dy);
  cp(s.inst_group, n, f.group); cp(s.inst_fixed, n, f.fixed); cp(s.inst_macro, n, f.macro);
  cp(s.inst_ground_at_bottom, n, f.gnd);
  cp(s.box_group, nb, f.box_group); cp(s.box_lx, nb, f.box_lx); cp(s.box_ly, nb, f.box_ly);
  cp(s.box_ux, nb, f.box_ux); cp(s.box_uy, nb, f.box_uy);
  cp(s.net_start, s.num_nets + 1, f.net_start);
  const int np = s.num_nets > 0 ? s.net_start[s.num_nets] : 0;
  cp(s.pin_inst, np, f.pin_inst); cp(s.pin_dx, np, f.pin_dx); cp(s.pin_dy, np, f.pin_dy);
  f.d = s;  // scalars, names, technology
  f.d.row_lx = f.row_lx.data(); f.d.row_ly = f.row_ly.data(); f.d.row_dx = f.row_dx.data();
  f.d.row_dy = f.row_dy.data(); f.d.row_site_width = f.row_sw.data(); f.d.row_orient = f.row_or.data();
  f.d.inst_lx = f.lx.data(); f.d.inst_ly = f.ly.data(); f.d.inst_dx = f.dx.data(); f.d.inst_dy = f.dy.data();
  f.d.inst_group = f.group.data(); f.d.inst_fixed = f.fixed.data(); f.d.inst_macro = f.macro.data();
  f.d.inst_ground_at_bottom = f.gnd.data();
  f.d.box_group = f.box_group.data(); f.d.box_lx = f.box_lx.data(); f.d.box_ly = f.box_ly.data();
  f.d.box_ux = f.box_ux.data(); f.d.box_uy = f.box_uy.data();
  f.d.net_start = f.net_start.data(); f.d.pin_inst = f.pin_inst.data();
  f.d.pin_dx = f.pin_dx.data(); f.d.pin_dy = f.pin_dy.data();
}

struct Driver {
  std::unique_ptr<lgd::DesignFile> design;
  LdDb* ld = nullptr;  // LEF / DEF database (read_lef / read_def)
  std::vector<int32_t> lx, ly;
  std::vector<uint8_t> fl;
  bool legalized = false;

  ~Driver() {
    if (ld) ld_destroy(ld);
  }

  // CmdReadLef / CmdReadDef / CmdWriteDef (tclCmd/src/CmdCore.cpp:144-150)
  bool read_lef(const std::vector<std::string>& tok) {
    if (tok.size() != 2) { std::printf("read_lef <file.lef>\n"); return false; }
    if (!ld && ld_create(&ld)) { std::printf("read_lef: %s\n", ld_last_error()); return false; }
    if (ld_read_lef(ld, tok[1].c_str())) { std::printf("read_lef: %s\n", ld_last_error()); return false; }
    return true;
  }

  bool read_def(const std::vector<std::string>& tok) {
    if (tok.size() != 2) { std::printf("read_def <file.def>\n"); return false; }
    if (!ld) { std::printf("read_def: read_lef first\n"); return false; }
    if (ld_read_def(ld, tok[1].c_str())) { std::printf("read_def: %s\n", ld_last_error()); return false; }
    LgDesign s;
    if (ld_get_design(ld, &s)) { std::printf("read_def: %s\n", ld_last_error()); return false; }
    auto f = std::make_unique<lgd::DesignFile>();
    copy_design(s, *f);
    std::printf("read_def: %d instances, %d rows, %d nets, %d masters\n", s.num_insts, s.num_rows, s.num_nets,
                s.num_masters);
    design = std::move(f);
    legalized = false;
    return true;
  }

  bool write_def(const std::vector<std::string>& tok) {
    if (tok.size() != 2) { std::printf("write_def <file.def>\n"); return false; }
    if (!ld || !design) { std::printf("write_def: no DEF read\n"); return false; }
    // the legalized positions, else the design's current (global-placed) ones
    const int32_t* x = legalized ? lx.data() : design->lx.data();
    const int32_t* y = legalized ? ly.data() : design->ly.data();
    if (ld_write_def(ld, tok[1].c_str(), x, y, legalized ? fl.data() : nullptr)) {
      std::printf("write_def: %s\n", ld_last_error());
      return false;
    }
    return true;
  }

  bool read_design(const std::vector<std::string>& tok) {
    if (tok.size() != 2) { std::printf("read_design <file.lgd>\n"); return false; }
    auto f = std::make_unique<lgd::DesignFile>();
    std::string err;
    if (!lgd::read(tok[1].c_str(), *f, err)) { std::printf("read_design: %s\n", err.c_str()); return false; }
    std::printf("read_design: %d instances, %d rows, %d nets\n", f->d.num_insts, f->d.num_rows, f->d.num_nets);
    design = std::move(f);
    legalized = false;
    return true;
  }

  // CmdGlobalPlace::execute (CmdGlobalPlace.hpp:27-93)
  bool global_place(const std::vector<std::string>& tok) {
    Command c;
    for (const char* o : {"-target_density", "-target_overflow", "-initial_place", "-initial_lambda",
                          "-initial_gamma_inv", "-solver", "-plot", "-bins", "-scaling_factor", "-ref_hpwl",
                          "-ref_hpwl_ratio", "-precision", "-max_iter", "-seed"})
      c.add(o, 1);
    c.add("-initial_place_only", 0);
    std::string err;
    if (!c.parse(tok, err)) { std::printf("%s\n", err.c_str()); return false; }
    if (!design) { std::printf("global_place: no design (read_design first)\n"); return false; }
    skyplace::SkyPlace sp(&design->d);
    if (auto o = c.get("-target_overflow")) sp.setTargetOverflow(std::stof(o->arg));
    if (auto o = c.get("-target_density")) sp.setTargetDensity(std::stof(o->arg));
    if (auto o = c.get("-initial_place")) sp.setInitialPlaceMethod(o->arg);
    if (c.get("-initial_place_only")) sp.setInitialPlaceOnly();
    if (auto o = c.get("-initial_lambda")) sp.setInitLambda(std::stof(o->arg));
    if (auto o = c.get("-initial_gamma_inv")) sp.setInitGammaInv(std::stof(o->arg));
    if (auto o = c.get("-solver")) sp.setSolver(o->arg);
    if (auto o = c.get("-plot")) sp.setPlotMode(o->arg);
    GpParams& p = sp.params();
    if (auto o = c.get("-bins")) p.num_bin_x = p.num_bin_y = std::stoi(o->arg);
    if (auto o = c.get("-scaling_factor")) p.scaling_factor = std::stof(o->arg);
    if (auto o = c.get("-ref_hpwl")) sp.setRefHpwl(std::stof(o->arg));
    if (auto o = c.get("-ref_hpwl_ratio")) p.reference_hpwl_ratio = std::stof(o->arg);
    if (auto o = c.get("-max_iter")) p.max_opt_iter = std::stoi(o->arg);
    if (auto o = c.get("-seed")) p.seed = std::stoull(o->arg);
    if (auto o = c.get("-precision")) {
      if (o->arg == "BF16") p.precision = 0;
      else if (o->arg == "SPLIT_BF16") p.precision = 1;
      else if (o->arg == "F32") p.precision = 2;
      else { std::printf("global_place: unknown precision %s\n", o->arg.c_str()); return false; }
    }
    p.verbose = 1;
    const int n = design->d.num_insts;
    std::vector<int32_t> lx(n), ly(n);
    const int st = sp.run(lx.data(), ly.data());
    if (st) { std::printf("global_place: error %d: %s\n", st, gp_last_error()); return false; }
    // exportDB: the design now holds the global placement
    std::copy(lx.begin(), lx.end(), design->lx.begin());
    std::copy(ly.begin(), ly.end(), design->ly.begin());
    legalized = false;
    const GpStats& s = sp.stats();
    std::printf("[gp_summary] converged=%d iterations=%d overflow=%.6f hpwl=%.6f fillers=%d bins=%dx%d\n",
                s.converged, s.iterations, s.overflow, s.hpwl_db, s.num_fillers, s.num_bin_x, s.num_bin_y);
    return true;
  }

  // CmdLegalize::execute (CmdLegalize.hpp:36-152)
  bool legalize(const std::vector<std::string>& tok) {
    Command c;
    for (const char* o : {"-constraint", "-tech_penalty", "-max_iter", "-coeff_to_init_rho", "-coeff_to_init_dual",
                          "-size_file", "-min_iter_to_call_adaptive_search", "-ovf_ratio_to_call_adaptive_search",
                          "-coeff_to_neglect_displace", "-random_sorting_seed", "-sorting_algorithm",
                          "-sorting_policy"})
      c.add(o, 1);
    for (const char* o : {"-std_admm", "-qp_refine", "-use_adaptive_search"}) c.add(o, 0);
    std::string err;
    if (!c.parse(tok, err)) { std::printf("%s\n", err.c_str()); return false; }
    if (!design) { std::printf("legalize: no design (read_design first)\n"); return false; }
    legalizer::Legalizer lg(&design->d);
    if (auto o = c.get("-constraint")) {
      lg.setICCAD17Constraint(o->arg);
      lg.setFlagICCAD2017(true);
    }
    if (auto o = c.get("-tech_penalty")) lg.setTechPenalty(std::stof(o->arg));
    if (auto o = c.get("-max_iter")) lg.setMaxAdmmIter(std::stoi(o->arg));
    if (c.get("-std_admm")) lg.setStandardAdmm(true);
    if (auto o = c.get("-coeff_to_init_rho")) lg.setCoeffToInitRho(std::stof(o->arg));
    if (auto o = c.get("-coeff_to_init_dual")) lg.setCoeffToInitDual(std::stof(o->arg));
    if (c.get("-use_adaptive_search")) lg.setAdaptiveSearch(true);
    if (auto o = c.get("-min_iter_to_call_adaptive_search")) lg.setMinIterToCallAdaptiveSearch(std::stoi(o->arg));
    if (auto o = c.get("-ovf_ratio_to_call_adaptive_search")) lg.setOvfRatioToCallAdaptiveSearch(std::stof(o->arg));
    if (auto o = c.get("-coeff_to_neglect_displace")) lg.setCoeffToNeglectDisplace(std::stof(o->arg));
    if (c.get("-qp_refine")) lg.setQpRefine(true);
    if (auto o = c.get("-sorting_algorithm")) lg.setSortingAlgorithm(o->arg);
    if (auto o = c.get("-sorting_policy")) lg.setSortingPolicy(o->arg);
    if (auto o = c.get("-random_sorting_seed")) lg.setRandomSortingSeed(std::stoi(o->arg));
    if (auto o = c.get("-size_file")) lg.setSizeByFile(o->arg);
    lg.params().verbose = 1;
    const int n = design->d.num_insts;
    lx.assign(n, 0);
    ly.assign(n, 0);
    fl.assign(n, 0);
    LgResult r{};
    r.out_lx = lx.data();
    r.out_ly = ly.data();
    r.out_flipped = fl.data();
    const int st = lg.run(&r);
    if (st) { std::printf("legalize: error %d: %s\n", st, lg_last_error()); return false; }
    const LgStats& s = r.stats;
    // one machine-readable summary line after the reference-style report
    std::printf("[summary] converged=%d iterations=%d total_disp_dbu=%lld max_disp_dbu=%lld overlap=%d "
                "row_orient_vio=%d site_vio=%d fence_vio=%d core_vio=%d rail_vio=%d pin_short_vio=%d "
                "pin_access_vio=%d edge_spacing_vio=%d hpwl_before_um=%.6f hpwl_after_um=%.6f iccad17_score=%.9f\n",
                s.converged, s.iterations, (long long)s.total_disp_dbu, (long long)s.max_disp_dbu, s.vio.total_ovf,
                s.vio.num_row_orient_vio, s.vio.num_site_vio, s.vio.num_fence_vio, s.vio.num_core_vio,
                s.vio.num_rail_vio, s.vio.num_pin_short_vio, s.vio.num_pin_access_vio, s.vio.num_edge_spacing_vio,
                s.hpwl_before_um, s.hpwl_after_um, s.iccad17_score);
    legalized = true;
    return true;
  }

  bool write_result(const std::vector<std::string>& tok) {
    if (tok.size() != 2) { std::printf("write_result <file.lgr>\n"); return false; }
    if (!legalized) { std::printf("write_result: nothing legalized yet\n"); return false; }
    if (!lgd::write_result(tok[1].c_str(), (int)lx.size(), lx.data(), ly.data(), fl.data())) {
      std::printf("write_result: cannot write %s\n", tok[1].c_str());
      return false;
    }
    return true;
  }

  bool run(const std::vector<std::string>& tok) {
    if (tok[0] == "read_lef") return read_lef(tok);
    if (tok[0] == "read_def") return read_def(tok);
    if (tok[0] == "write_def") return write_def(tok);
    if (tok[0] == "read_design") return read_design(tok);
    if (tok[0] == "legalize") return legalize(tok);
    if (tok[0] == "global_place") return global_place(tok);
    if (tok[0] == "write_result") return write_result(tok);
    std::printf("invalid command name \"%s\"\n", tok[0].c_str());
    return false;
  }
};

std::vector<std::vector<std::string>> commands_of(const std::string& text) {
  std::vector<std::vector<std::string>> out;
  std::string line;
  std::istringstream in(text);
  while (std::getline(in, line)) {
    line = line.substr(0, line.find('#'));
    std::stringstream parts(line);
    std::string cmd;
    while (std::getline(parts, cmd, ';')) {
      std::istringstream ws(cmd);
      std::vector<std::string> tok;
      for (std::string t; ws >> t;) tok.push_back(t);
      if (!tok.empty()) out.push_back(tok);
    }
  }
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  std::string text;
  if (argc == 3 && std::string(argv[1]) == "-c") {
    text = argv[2];
  } else if (argc == 2) {
    std::ifstream f(argv[1]);
    if (!f) { std::printf("cannot open %s\n", argv[1]); return 1; }
    std::stringstream ss;
    ss << f.rdbuf();
    text = ss.str();
  } else {
    std::printf("usage: %s <script> | %s -c \"<commands>\"\n", argv[0], argv[0]);
    return 1;
  }
  Driver drv;
  for (const auto& tok : commands_of(text)) {
    std::printf("skyline> ");
    for (const auto& t : tok) std::printf("%s ", t.c_str());
    std::printf("\n");
    std::fflush(stdout);
    if (!drv.run(tok)) return 1;
    std::fflush(stdout);
  }
  return 0;
}
