"""The C ABI: the shared library loads and exports every function the public
headers declare (no device calls; runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from admm_legalization_amd import build, density, gp, lefdef, lg, wirelength

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# every C header of include/ (the C++ class headers under include/skyline/
# are checked by the caller test below)
HEADERS = sorted(os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))
                 if h.endswith(".h"))
BY_NAME = {os.path.basename(h): h for h in HEADERS}


def declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(\w+)\s*\(", text, flags=re.M)) - {"if", "while"}


@pytest.fixture(scope="module")
def library():
    if not os.path.exists(build.LIB):
        build.build()
    return build.LIB


def test_headers_declare_the_expected_entry_points():
    assert set(BY_NAME) == {"skyline_lg.h", "skyline_density.h", "skyline_wl.h", "skyline_gp.h", "skyline_lefdef.h"}
    lg_syms, dens_syms, wl_syms, gp_syms, ld_syms = (
        declared(BY_NAME[h]) for h in ("skyline_lg.h", "skyline_density.h", "skyline_wl.h", "skyline_gp.h",
                                       "skyline_lefdef.h"))
    assert {"lg_create", "lg_solve", "lg_postprocess", "lg_check", "lg_copy_out", "lg_destroy"} <= lg_syms
    assert {"dens_create", "dens_compute_grad", "poisson_create", "poisson_solve"} <= dens_syms
    assert {"wl_create", "wl_compute_grad", "wl_update_pin_coordinates", "tf_precondition"} <= wl_syms
    assert {"ld_read_lef", "ld_read_def", "ld_write_def", "ld_get_design"} <= ld_syms
    assert set(lg.EXPORTED) == lg_syms
    assert set(density.EXPORTED) == dens_syms
    assert set(wirelength.EXPORTED) == wl_syms
    assert set(gp.EXPORTED) == gp_syms
    assert set(lefdef.EXPORTED) == ld_syms


def test_library_exports_every_declared_symbol(library):
    out = subprocess.run(["nm", "-D", "--defined-only", library], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for h in HEADERS:
        missing = declared(h) - exported
        assert not missing, f"{os.path.basename(h)}: not exported {sorted(missing)}"


def test_library_loads_with_ctypes(library):
    L = C.CDLL(library)
    for h in HEADERS:
        for name in declared(h):
            assert getattr(L, name) is not None


def test_default_params_and_error_text(library):
    p = lg.default_params()
    assert p.x_hint == 250 and p.max_admm_iter > 0
    L = lg.lib()
    assert isinstance(L.lg_last_error(), bytes)


def test_server_config_mirror_matches_the_library(library):
    """lg_srv_default_config fills the Python mirror of LgSrvConfig field for
    field, and lg_srv_create refuses a bad configuration before any device
    call."""
    from admm_legalization_amd import server
    c = server.default_config()
    assert C.sizeof(server.LgSrvConfig) == 40
    assert (c.max_slots, c.iters_per_sync, c.mem_budget_bytes, c.pool_chunk_bytes, c.use_graphs) == (
        128, 16, 200 * 10**9, 256 << 20, 0)
    assert c.short_cells >= 0 and c.max_overtakes >= 0 and c.reserved == 0
    L = server._bind()
    h = C.c_void_p()
    for field, bad in (("short_cells", -1), ("max_overtakes", -5), ("max_slots", 0)):
        b = server.default_config()
        setattr(b, field, bad)
        assert L.lg_srv_create(C.byref(b), C.byref(h)) == 1  # LG_ERR_ARG
        assert b"configuration" in L.lg_last_error()


def test_null_arguments_are_rejected_without_a_device(library):
    L = lg.lib()
    assert L.lg_create(None, None, None, None) != 0
    assert b"null" in L.lg_last_error()
    D = density.lib()
    assert D.dens_create(None, None) != 0
    Wl = wirelength.lib()
    assert Wl.wl_create(None, None) == 1 and b"null" in Wl.wl_last_error()
    assert Wl.tf_precondition(-1, 0.0, 1.0, *([None] * 8), None) == 1
    assert D.poisson_solve(None, None, None, None, None, 1, None) != 0
    P = gp.lib()
    h = C.c_void_p()
    assert P.gp_create(None, None, C.byref(h)) == 1 and b"null" in P.gp_last_error()
    assert P.gp_iterate(None, 1, None) == 1


def test_placer_default_params_are_the_reference_hyper_parameters(library):
    """HyperParam (HyperParam.h:23-48) and SkyPlaceDB's target density 1.0."""
    p = gp.default_params()
    assert (p.max_opt_iter, p.max_backtrack_iter, p.log_freq) == (2500, 10, 50)
    for k, v in dict(init_lambda=1e-4, init_gamma_inv_coef=0.01, target_overflow=0.07, min_phi_coef=0.95,
                     max_phi_coef=1.05, min_precond=1.0, init_opt_coef=100.0, reference_hpwl=200000.0,
                     min_step_length=1.0, adam_alpha=100.0, adam_beta1=0.9, adam_beta2=0.999,
                     target_density=1.0).items():
        assert getattr(p, k) == pytest.approx(v, rel=1e-7), k
    assert p.solver == gp.SOLVER_NESTEROV and p.num_bin_x == 0 and p.precision == density.PREC_BF16


def test_cpp_headers_compile_and_link(library, tmp_path):
    """A C++ caller of the public headers builds against the library and runs
    its host-only entry points (parameters, density bin setup)."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "skyline/Legalizer.hpp"
#include "skyline/DensityGradient.hpp"
#include "skyline/SkyPlace.hpp"
int main() {
  LgParams p;
  lg_default_params(&p);
  LgDesign empty{};
  skyplace::SkyPlace sp(&empty);
  sp.setTargetDensity(1.5f);  // out of range: 1.0
  sp.setSolver("adam");
  if (sp.params().target_density != 1.0f || sp.params().solver != GP_SOLVER_ADAM) return 3;
  if (sp.params().max_opt_iter != 2500) return 4;
  int32_t lx[1], ly[1];
  if (sp.run(lx, ly) != GP_ERR_ARG) return 5;  // no instances: refused before any device call
  std::vector<float> dx(100, 1.5f), dy(100, 2.0f);
  DensDesign d{};
  d.die_ux = d.die_uy = 64.f;
  d.target_density = 1.f;
  d.num_movable = 100;
  d.cell_dx = dx.data();
  d.cell_dy = dy.data();
  int32_t shape[2];
  if (dens_prepare(&d, shape, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr) != DENS_OK) return 2;
  std::printf("%d %d %d\n", p.x_hint, shape[0], shape[1]);
  return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(library)
    subprocess.run(["g++", "-std=c++17", str(src), "-I", os.path.join(ROOT, "include"), "-L", libdir,
                    "-lskyline_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert "Invalid target density (150.00%). 100.00% will be used." in res  # SkyPlace.cpp:105-108
    out = res.splitlines()[-1].split()
    assert out[0] == "250" and int(out[1]) >= 4 and out[1] == out[2]


def test_server_rejects_bad_arguments_without_a_device(library):
    from admm_legalization_amd import server
    L = server._bind()
    c = server.default_config()
    assert (c.max_slots, c.iters_per_sync) == (128, 16)
    assert c.mem_budget_bytes == 200 * 10 ** 9 and c.pool_chunk_bytes == 256 << 20
    h = C.c_void_p()
    assert L.lg_srv_create(None, C.byref(h)) == 1  # LG_ERR_ARG
    for bad in (server.LgSrvConfig(0, 16, 1 << 30, 1 << 20), server.LgSrvConfig(8, 0, 1 << 30, 1 << 20),
                server.LgSrvConfig(8, 16, 0, 1 << 20), server.LgSrvConfig(2000, 16, 1 << 30, 1 << 20)):
        assert L.lg_srv_create(C.byref(bad), C.byref(h)) == 1
        assert b"configuration" in L.lg_last_error()
    assert L.lg_srv_submit(None, None, None, None, None) == 1
    assert L.lg_srv_wait(None, 1, None) == 1
    assert L.lg_srv_poll(None, 1, None) == 1
    assert L.lg_srv_destroy(None) == 0


def test_create_rejects_bad_regions_and_window_without_a_device(library):
    """With the adaptive search on, pixel region ids index the region
    bounds: ids outside [-1, num_regions) and a stagnation window above the
    device's 64 entries are argument errors, found before any device call."""
    from admm_legalization_amd import synth
    d = synth.tiny(num_cells=60, x_size=48, y_size=12, seed=2)
    grid, cells, p = lg.device_inputs(d)
    p.use_adaptive_search = 1
    L = lg.lib()
    bad = dict(grid)
    bad["region"] = grid["region"].copy()
    bad["region"][5] = len(grid["region_x_min"]) + 3
    g, c, keep = lg.marshal_device_inputs(bad, cells)
    h = C.c_void_p()
    assert L.lg_create(C.byref(g), C.byref(c), C.byref(p), C.byref(h)) == 1
    assert b"region id" in L.lg_last_error()
    g, c, keep = lg.marshal_device_inputs(grid, cells)
    p.stagnation_window = 65
    assert L.lg_create(C.byref(g), C.byref(c), C.byref(p), C.byref(h)) == 1
    assert b"stagnation_window" in L.lg_last_error()


def test_create_refuses_cells_across_every_partition_column(library):
    """A cell whose footprint can span every partition column would, through
    the colouring's wrap-around, be handed to a partition whose tile does not
    hold it: lg_create refuses it (LG_ERR_DESIGN) before any device call.
    Narrower cells wider than x_hint are accepted (they never move)."""
    from admm_legalization_amd import synth
    d = synth.tiny(num_cells=60, x_size=48, y_size=12, seed=2)
    grid, cells, p = lg.device_inputs(d)
    p.x_hint = 20  # three partition columns on 48 sites
    c2 = {k: v.copy() for k, v in cells.items()}
    c2["w_grid"][0] = 42
    c2["lx_dbu"][0] = grid["core_lx"]
    g, c, keep = lg.marshal_device_inputs(grid, c2)
    h = C.c_void_p()
    L = lg.lib()
    assert L.lg_create(C.byref(g), C.byref(c), C.byref(p), C.byref(h)) == 3  # LG_ERR_DESIGN
    assert b"every partition column" in L.lg_last_error()
