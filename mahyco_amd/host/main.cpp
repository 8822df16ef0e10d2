// mahyco_run -- command-line driver of the GPU path (the counterpart of src/main.cc):
// builds one named case through MahycoModule, runs the time loop through the
// entry points (or fused cycles) and prints the final time and field checksums.
//   mahyco_run <case> [--cycles N] [--fused] [--graph K]
// cases: bisod_x  bisod_x_csts  sod2d_lagrange  sedov2d  bisod3d:<nx>x<ny>x<nz>
//        monorider_rotation  rider_deformation_reverse (advection of the RIDER bubble)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mahyco_module.h"

using namespace mahyco;

static std::unique_ptr<MahycoModule> make_case(const std::string& name) {
  MhMesh mesh;
  std::memset(&mesh, 0, sizeof(mesh));
  mesh.coord_rule = MH_COORD_DIV;
  MahycoOptions opt;
  std::vector<EnvironmentConfig> envs(2);
  auto pg = [](double gamma) {
    PerfectGasOptions o;
    o.adiabatic_cst = gamma;
    o.specific_heat = 2.4;
    return std::make_shared<PerfectGasEOS>(o);
  };
  RemapADIOptions ro;
  auto walls = [&](int dim) {
    for (int s = 0; s < 2 * dim; ++s) opt.boundary_condition.push_back({s, s / 2, 0.0});
  };
  auto fictif = [] {
    FictifOptions o;
    o.tfin_pression = 10.;
    return std::make_shared<FictifEOS>(o);
  };
  if (name == "monorider_rotation" || name == "rider_deformation_reverse") {
    // NONREGRESSION/CAS_MonoRiderRotation, CAS_RiderDeformationTimeReverse: sans-lagrange
    mesh.dim = 3;
    mesh.nx = 50; mesh.ny = 50; mesh.nz = 1;
    mesh.length[0] = 1.; mesh.length[1] = 1.; mesh.length[2] = .1;
    envs[0] = {"Vide", 0., 0., {0, 0, 0}, 0.5, 1.2, fictif()};
    envs[1] = {"Bulle", 1., 0., {0, 0, 0}, 0.5, 1.2, fictif()};
    opt.cas_model = MH_CASE_RIDER;
    opt.sans_lagrange = true;
    opt.deltat_init = 1e-3;
    opt.deltat_max = 1e-4;
    if (name == "monorider_rotation") {
      opt.cas_test = 26;
      opt.final_time = .5;
    } else {
      opt.cas_test = 38;
      opt.final_time = 1.;
      opt.cas.reverse_option = true;
      opt.cas.parametre = 1.;
    }
    return std::unique_ptr<MahycoModule>(new MahycoModule(mesh, envs, opt, std::make_shared<RemapADI>(ro)));
  }
  if (name == "bisod_x" || name == "bisod_x_csts" || name.rfind("bisod3d", 0) == 0) {  // NONREGRESSION/CAS_BiSodCaseX
    opt.schema_csts = name == "bisod_x_csts";
    mesh.dim = 3;
    mesh.nx = 100; mesh.ny = 5; mesh.nz = 5;
    mesh.length[0] = 1.; mesh.length[1] = .1; mesh.length[2] = .1;
    double g1 = 1.4;
    if (name.rfind("bisod3d", 0) == 0) {  // 3D bi-material tube on [0,1]^3, superbee
      if (sscanf(name.c_str(), "bisod3d:%dx%dx%d", &mesh.nx, &mesh.ny, &mesh.nz) != 3) return nullptr;
      mesh.length[1] = mesh.length[2] = 1.;
      g1 = 1.67;
      ro.projection_limiteur_id = MH_SUPERBEE;
      opt.final_time = 1e30;
      opt.deltat_min = 1e-12;
    }
    envs[0] = {"ZG", 1., 1., {0, 0, 0}, 0.5, 1.2, pg(1.4)};
    envs[1] = {"ZD", .125, .1, {0, 0, 0}, 0.5, 1.2, pg(g1)};
    opt.cas_model = MH_CASE_SOD;
    opt.cas_test = 13;
    walls(3);
  } else if (name == "sod2d_lagrange") {  // 2D Sod 200x10, single material, Lagrange only
    mesh.dim = 2;
    mesh.nx = 200; mesh.ny = 10;
    mesh.length[0] = 1.; mesh.length[1] = .1;
    envs.resize(1);
    envs[0] = {"ZG", 1., 1., {0, 0, 0}, 0.5, 1.2, pg(1.4)};
    opt.cas_model = MH_CASE_SOD;
    opt.cas_test = 3;
    opt.with_projection = false;
    opt.deltat_max = 1e-4;
    walls(3);
  } else if (name == "sedov2d") {  // NONREGRESSION/CAS_2D_Sedov
    mesh.dim = 2;
    mesh.nx = 60; mesh.ny = 60;
    mesh.length[0] = 1.2; mesh.length[1] = 1.2;
    envs[0] = {"ZoneAir", 1., 0.0979264e-5, {0, 0, 0}, 0.5, 1.2, pg(1.4)};
    envs[1] = {"ZoneInit", 1., 1., {0, 0, 0}, 0.5, 1.2, pg(1.4)};
    opt.cas_model = MH_CASE_SEDOV;
    opt.cas_test = 1;
    opt.final_time = 1.;
    opt.boundary_condition = {{MH_XMIN, MH_BC_VX, 0.}, {MH_YMIN, MH_BC_VY, 0.}};
  } else {
    return nullptr;
  }
  return std::unique_ptr<MahycoModule>(new MahycoModule(mesh, envs, opt, std::make_shared<RemapADI>(ro)));
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: mahyco_run <case> [--cycles N] [--fused] [--graph K]\n");
    return 2;
  }
  int cycles = 1 << 30, graph = 0;
  bool fused = false;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--cycles") && i + 1 < argc) cycles = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--fused")) fused = true;
    else if (!strcmp(argv[i], "--graph") && i + 1 < argc) graph = atoi(argv[++i]);
  }
  try {
    auto mod = make_case(argv[1]);
    if (!mod) {
      fprintf(stderr, "unknown case %s\n", argv[1]);
      return 2;
    }
    mod->hydroStartInit();
    mod->computeCellMass();
    mod->computeNodeMass();
    if (fused) {
      // queued in chunks, each followed by a status check: iterations after the stop are
      // no-ops on the device, but an unbounded request must not queue them all
      const int chunk = graph > 0 ? 10 * graph : 60;
      for (int done = 0; done < cycles;) {
        const int k = std::min(chunk, cycles - done);
        if (mod->cycle(k, graph) != MH_OK) break;
        done += k;
      }
    } else {
      for (int it = 0; it < cycles; ++it)
        if (!mod->iterate()) break;
    }
    const std::vector<double> rho = mod->field(MH_F_DENSITY);
    const std::vector<double> e = mod->field(MH_F_INTERNAL_ENERGY);
    double srho = 0., se = 0.;
    for (double v : rho) srho += v;
    for (double v : e) se += v;
    printf("case=%s iteration=%d time=%.17g deltat=%.17g sum_rho=%.17g sum_e=%.17g\n", argv[1],
           mod->globalIteration(), mod->globalTime(), mod->globalDeltat(), srho, se);
  } catch (const MahycoError& err) {
    fprintf(stderr, "mahyco_run: %s\n", err.what());
    return 1;  // the reference exits with status 1 at these sites
  }
  return 0;
}
