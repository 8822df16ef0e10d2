// mh_output.h -- on-disk outputs of a case, in the formats the reference writes and its
// non-regression suite compares with diff -r (NONREGRESSION/launch_test.sh:74-96):
//
//  * EnSight Gold ASCII post-processing (the `<output>` block of a case's Donnees.arc):
//    <dir>/ensight.case, <dir>/ensight.geo/ensight.geoNNNNNN and <dir>/<Var>/<Var>NNNNNN,
//    one part "AllCells", values "%12.5e", integers "%10d", hexa8 / quad4 connectivity in
//    the Cartesian generator's node order, 3 coordinate / vector components also in 2D;
//  * gnuplot time histories of EnvSummationService (time_history/EnvSummationService.cc:
//    77-169): <dir>/<Var>_<environment>, one "%.16E %.16E" line per recorded iteration.
//
// Host code: values are read back with mahyco_get_field / mahyco_history_read.
#ifndef MH_OUTPUT_H
#define MH_OUTPUT_H

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mahyco.h"

namespace mh_output {

struct VarSpec {
  const char* name;
  int field;
  bool node;    // per node (else per element)
  bool vector;  // 3 components
};

// Reference variable names (Mahyco.axl) of the fields the library holds.
inline const VarSpec* find_var(const char* name) {
  static const VarSpec specs[] = {
      {"Density", MH_F_DENSITY, false, false},
      {"Pressure", MH_F_PRESSURE, false, false},
      {"InternalEnergy", MH_F_INTERNAL_ENERGY, false, false},
      {"SoundSpeed", MH_F_SOUND_SPEED, false, false},
      {"PseudoViscosity", MH_F_PSEUDO_VISCOSITY, false, false},
      {"CellMass", MH_F_CELL_MASS, false, false},
      {"CellVolume", MH_F_CELL_VOLUME, false, false},
      {"FracVol", MH_F_FRACVOL, false, false},
      {"CaracteristicLength", MH_F_CARACTERISTIC_LENGTH, false, false},
      {"Materiau", MH_F_MATERIAU, false, false},
      {"Velocity", MH_N_VELOCITY, true, true},
      {"NodeMass", MH_N_NODE_MASS, true, false},
  };
  for (const auto& s : specs)
    if (!strcmp(s.name, name)) return &s;
  return nullptr;
}

// mkdir -p
inline bool make_dir(const std::string& d) {
  for (size_t p = d.find('/', 1); p != std::string::npos; p = d.find('/', p + 1)) {
    const std::string part = d.substr(0, p);
    if (mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return false;
  }
  return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST;
}

inline std::string step_name(int step) {
  char b[16];
  snprintf(b, sizeof b, "%06d", step);
  return b;
}

struct File {
  FILE* f = nullptr;
  explicit File(const std::string& path) : f(fopen(path.c_str(), "w")) {}
  ~File() {
    if (f) fclose(f);
  }
};

// Geometry of one step: all nodes (x, then y, then z) and all cells.
inline int write_geometry(MhHandle* h, int dim, const int* n, const std::string& path) {
  const long nc = mahyco_nb_cells(h), nn = mahyco_nb_nodes(h);
  File o(path);
  if (!o.f) return MH_ERR_INVALID;
  fprintf(o.f, "Output ensight test\nFile description\nnode id assign\nelement id assign\npart\n%10d\nAllCells\n", 1);
  fprintf(o.f, "coordinates\n%10ld\n", nn);
  std::vector<double> v(nn);
  for (int c = 0; c < 3; ++c) {
    if (c < dim) {
      const int st = mahyco_get_field(h, MH_N_COORD, -1, c, v.data());
      if (st != MH_OK) return st;
    } else {
      std::fill(v.begin(), v.end(), 0.);
    }
    for (long i = 0; i < nn; ++i) fprintf(o.f, "%12.5e\n", v[i]);
  }
  const long nnx = n[0] + 1, nnxy = nnx * (n[1] + 1);
  fprintf(o.f, "%s\n%10ld\n", dim == 3 ? "hexa8" : "quad4", nc);
  for (long c = 0; c < nc; ++c) {
    const long i = c % n[0], r = c / n[0], j = r % n[1], k = r / n[1];
    const long b = 1 + i + nnx * j + nnxy * k;  // 1-based
    if (dim == 3)
      fprintf(o.f, "%10ld%10ld%10ld%10ld%10ld%10ld%10ld%10ld\n", b, b + 1, b + 1 + nnx, b + nnx, b + nnxy,
              b + 1 + nnxy, b + 1 + nnx + nnxy, b + nnx + nnxy);
    else
      fprintf(o.f, "%10ld%10ld%10ld%10ld\n", b, b + 1, b + 1 + nnx, b + nnx);
  }
  return ferror(o.f) ? MH_ERR_INVALID : MH_OK;
}

inline int write_variable(MhHandle* h, int dim, const VarSpec& s, const std::string& path) {
  const long n = s.node ? mahyco_nb_nodes(h) : mahyco_nb_cells(h);
  File o(path);
  if (!o.f) return MH_ERR_INVALID;
  fprintf(o.f, "%s\npart\n%10d\n%s\n", s.name, 1, s.node ? "coordinates" : (dim == 3 ? "hexa8" : "quad4"));
  std::vector<double> v(n);
  const int ncomp = s.vector ? 3 : 1;
  for (int c = 0; c < ncomp; ++c) {
    if (c < dim || !s.vector) {
      const int st = mahyco_get_field(h, s.field, -1, c, v.data());
      if (st != MH_OK) return st;
    } else {
      std::fill(v.begin(), v.end(), 0.);
    }
    for (long i = 0; i < n; ++i) fprintf(o.f, "%12.5e\n", v[i]);
  }
  return ferror(o.f) ? MH_ERR_INVALID : MH_OK;
}

inline int write_case(const std::string& dir, const std::vector<const VarSpec*>& vars, const std::vector<double>& times) {
  File o(dir + "/ensight.case");
  if (!o.f) return MH_ERR_INVALID;
  fprintf(o.f, "FORMAT\ntype: ensight gold\n\nGEOMETRY\nmodel: 1 ensight.geo/ensight.geo******\n\nVARIABLE\n");
  for (const VarSpec* s : vars) {
    const char* kind = s->vector ? "vector per node:" : s->node ? "scalar per node:" : "scalar per element:";
    fprintf(o.f, "%-23s1   %s %s/%s******\n", kind, s->name, s->name, s->name);
  }
  fprintf(o.f, "\nTIME\n%-23s1\n%-23s%d\n%-23s0\n%-23s1\ntime values:\n", "time set:", "number of steps:",
          (int)times.size(), "filename start number:", "filename increment:");
  for (double t : times) fprintf(o.f, "%.15g\n", t);
  fprintf(o.f, "\n");
  return ferror(o.f) ? MH_ERR_INVALID : MH_OK;
}

// EnvSummation histories: rows of 1 + 5m doubles (time; per environment mean M_k, mean
// p_k, volume-weighted rho_k, mass-weighted e_k, member count).  An environment without
// cells at a recorded iteration adds no line (EnvSummationService.cc:97 skips it), and one
// that never had cells has no file.
inline int write_gnuplot(const std::string& dir, const std::vector<double>& rows, int nrows, int m,
                         const char* const* env_names) {
  static const struct {
    const char* name;
    int col;
  } cols[] = {{"CellMass", 0}, {"Pressure", 1}, {"Density", 2}, {"InternalEnergy", 3}};
  const int w = 1 + 5 * m;
  for (int k = 0; k < m; ++k) {
    bool any = false;
    for (int r = 0; r < nrows; ++r) any |= rows[(size_t)r * w + 1 + 5 * k + 4] > 0.;
    if (!any) continue;
    for (const auto& c : cols) {
      File o(dir + "/" + c.name + "_" + env_names[k]);
      if (!o.f) return MH_ERR_INVALID;
      for (int r = 0; r < nrows; ++r) {
        const double* row = &rows[(size_t)r * w];
        if (row[1 + 5 * k + 4] > 0.) fprintf(o.f, "%.16E %.16E\n", row[0], row[1 + 5 * k + c.col]);
      }
      if (ferror(o.f)) return MH_ERR_INVALID;
    }
  }
  return MH_OK;
}

}  // namespace mh_output

#endif
