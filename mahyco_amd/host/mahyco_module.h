// mahyco_module.h -- C++ host interface mirroring MaHyCo's module and services for the
// Lagrange + ADI-remap path, implemented over the C ABI of include/mahyco.h.
//
// Names and argument meanings follow the reference:
//   IEquationOfState        src/eos/IEquationOfState.h:25-66
//   PerfectGasEOSService    src/eos/perfectgas/PerfectGasEOSService.cc
//   StiffenedGasEOSService  src/eos/stiffenedgas/StiffenedGasEOSService.cc
//   FictifEOSService        src/eos/fictif/FictifEOSService.cc
//   IRemap / RemapADI       src/Remap/IRemap.h:31-64, src/Remap/RemapADIService.cc
//   MahycoModule            src/MahycoModule.h, entry points src/Mahyco.axl:244-261
// An environment is designated by its index.  Where the reference calls exit(1) the
// methods throw MahycoError carrying the MhStatus code.
#ifndef MAHYCO_MODULE_H
#define MAHYCO_MODULE_H

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mahyco.h"

namespace mahyco {

class MahycoError : public std::runtime_error {
 public:
  MahycoError(int status, const std::string& where);
  int status() const { return status_; }

 private:
  int status_;
};

class MahycoModule;

// ---------------------------------------------------------------- IEquationOfState
class IEquationOfState {
 public:
  virtual ~IEquationOfState() = default;
  virtual void initEOS(int env) = 0;   // initial energy and sound speed (done by HydroStartInit)
  virtual void ReinitEOS(int env) = 0; // no state to rebuild on restart
  virtual void applyEOS(int env) = 0;  // pressure and sound speed of the env's cells
  // the same for one cell of the environment (IEquationOfState.h:52; Newton's iterates)
  virtual void applyOneCellEOS(int env, long cell);
  virtual void Endommagement(int env) = 0;  // inert at the default thresholds
  virtual double getAdiabaticCst(int env) = 0;
  virtual double getdensityDamageThresold(int env) = 0;
  virtual MhEnvironment description() const = 0;
  void attach(MahycoModule* m) { module_ = m; }

 protected:
  MahycoModule* module_ = nullptr;
};

struct PerfectGasOptions {  // PerfectGasEOS.axl:157-183
  double adiabatic_cst = 1.4, specific_heat = 0.0;
  double tension_damage_thresold = -1.e12, density_damage_thresold = 0.;
};
struct StiffenedGasOptions {  // StiffenedGasEOS.axl:158-184
  double adiabatic_cst = 1.4, limit_tension = 0.01, specific_heat = 0.0;
  double tension_damage_thresold = -1.e12, density_damage_thresold = 0.;
};

class PerfectGasEOS : public IEquationOfState {
 public:
  explicit PerfectGasEOS(PerfectGasOptions o = PerfectGasOptions()) : o_(o) {}
  void initEOS(int env) override;
  void ReinitEOS(int) override {}
  void applyEOS(int env) override;
  void Endommagement(int env) override;
  double getAdiabaticCst(int) override { return o_.adiabatic_cst; }
  double getdensityDamageThresold(int) override { return o_.density_damage_thresold; }
  MhEnvironment description() const override;

 private:
  PerfectGasOptions o_;
};

class StiffenedGasEOS : public IEquationOfState {
 public:
  explicit StiffenedGasEOS(StiffenedGasOptions o = StiffenedGasOptions()) : o_(o) {}
  void initEOS(int env) override;
  void ReinitEOS(int) override {}
  void applyEOS(int env) override;
  void Endommagement(int env) override;
  double getAdiabaticCst(int) override { return o_.adiabatic_cst; }
  double getdensityDamageThresold(int) override { return o_.density_damage_thresold; }
  MhEnvironment description() const override;

 private:
  StiffenedGasOptions o_;
};

struct FictifOptions {  // FictifEOS.axl: a pressure that depends on time only
  double tdebut_pression = 0., tfin_pression = 0.;
  double valeur_debut_pression = 0., valeur_dependance_temps = 0.;
};

class FictifEOS : public IEquationOfState {
 public:
  explicit FictifEOS(FictifOptions o = FictifOptions()) : o_(o) {}
  void initEOS(int) override {}  // not called in advection mode (MahycoModule.cc:143)
  void ReinitEOS(int) override {}
  void applyEOS(int env) override;
  void Endommagement(int) override {}
  double getAdiabaticCst(int) override { return 1.; }
  double getdensityDamageThresold(int) override { return 0.; }
  MhEnvironment description() const override;

 private:
  FictifOptions o_;
};

// ---------------------------------------------------------------- IRemap
struct RemapADIOptions {  // RemapADI.axl:165-209
  double threshold = 1.e-10;
  std::string is_euler_scheme = "Full";
  int ordre_projection = 2;
  bool conservation_energie_totale = false;
  bool projection_pente_borne = false;
  int projection_pente_borne_debar_fix = 0;
  bool projection_pente_borne_mixte = false;
  bool projection_simple_pente = false;
  int projection_limiteur_id = 0;
  int projection_limiteur_pure_id = 0;
};

class IRemap {
 public:
  virtual ~IRemap() = default;
  virtual void appliRemap(int dimension, int withDualProjection, int nb_vars_to_project, int nb_env) = 0;
  virtual void resizeRemapVariables(int nb_vars_to_project, int nb_env) = 0;
  virtual void synchronizeUremap() = 0;
  virtual int getOrdreProjection() = 0;
  virtual bool hasProjectionPenteBorne() = 0;
  virtual bool hasProjectionSimplePente() = 0;
  virtual bool hasConservationEnergieTotale() = 0;
  virtual std::string isEuler() = 0;
  virtual void remapVariables(int dimension, int withDualProjection, int nb_vars_to_project, int nb_env) = 0;
  void attach(MahycoModule* m) { module_ = m; }

 protected:
  MahycoModule* module_ = nullptr;
};

class RemapADI : public IRemap {
 public:
  explicit RemapADI(RemapADIOptions o = RemapADIOptions()) : o_(o) {}
  // the sweeps, the dual projection and remapVariables run as one device sequence
  void appliRemap(int dimension, int withDualProjection, int nb_vars_to_project, int nb_env) override;
  void resizeRemapVariables(int, int) override {}  // device buffers are sized at creation
  void synchronizeUremap() override {}             // one subdomain: no halo
  int getOrdreProjection() override { return o_.ordre_projection; }
  bool hasProjectionPenteBorne() override { return o_.projection_pente_borne; }
  bool hasProjectionSimplePente() override { return o_.projection_simple_pente; }
  bool hasConservationEnergieTotale() override { return o_.conservation_energie_totale; }
  std::string isEuler() override { return o_.is_euler_scheme; }
  void remapVariables(int, int, int, int) override {}  // part of appliRemap on the device
  const RemapADIOptions& options() const { return o_; }

 private:
  RemapADIOptions o_;
};

// ---------------------------------------------------------------- MahycoModule
struct EnvironmentConfig {  // <environment> of Mahyco.axl
  std::string name;
  double densite_initiale = 0., pression_initiale = 1.e5;
  double vitesse_initiale[3] = {0., 0., 0.};
  double linear_pseudo_coeff = 0.5, quadratic_pseudo_coeff = 1.2;
  std::shared_ptr<IEquationOfState> eos;
  double energie_initiale = 0.;
};

struct CaseOptions {  // options of the RIDER / ADVECTION case services (RIDER.axl, ADVECTION.axl)
  bool reverse_option = false;
  double parametre = 0.;
  double rayon_bulle = 0.15;
  bool densite_interieure_initiale_lineaire = false;
};

struct MahycoOptions {  // Mahyco.axl:263-557 (subset of this path)
  double cfl = 0.3, threshold = 1.e-10;
  double deltat_init = 1e-5, deltat_min = 1e-8, deltat_max = 0.01, final_time = 0.2;
  bool with_projection = true, calcul_vitson_max = true;
  bool schema_csts = false, pseudo_centree = false;
  bool with_newton = false, pression_explicite = false, sans_lagrange = false;
  std::string longueur_caracteristique = "racine-cubique-volume";
  std::vector<MhBoundaryCondition> boundary_condition;
  int cas_model = MH_CASE_SOD, cas_test = 13;
  double interface_position = 0.5;
  int interface_axis = 0;
  double sedov_deposit = 0.244816;
  CaseOptions cas;
};

class MahycoModule {
 public:
  MahycoModule(const MhMesh& mesh, std::vector<EnvironmentConfig> envs, MahycoOptions opt,
               std::shared_ptr<IRemap> remap);
  ~MahycoModule();
  MahycoModule(const MahycoModule&) = delete;
  MahycoModule& operator=(const MahycoModule&) = delete;

  // init entry points
  void hydroStartInit();
  void hydroContinueInit();   // after readCheckpoint: continue as the reference does
  void writeCheckpoint(const std::string& path);
  void readCheckpoint(const std::string& path);
  void computeCellMass() {}    // done by hydroStartInit on the device
  void computeNodeMass() {}    // formed by the first node kernel
  // compute-loop entry points (Mahyco.config:23-36)
  void saveValuesAtN();
  void updateVelocity();
  void applyBoundaryCondition();
  void updatePosition();
  void computeGeometricValues();
  void updateDensity();
  void computeArtificialViscosity();
  void updateElasticityAndPlasticity();
  void updateEnergyAndPressure();
  void remap();
  void computePressionMoyenne();
  void computeDeltaT();

  // one iteration of the time loop in config order; returns false once stopped
  bool iterate();
  // n fused iterations queued at once (no per-entry-point calls)
  int cycle(int n, int cycles_per_graph = 0);

  MhHandle* handle() const { return h_; }
  double globalTime() const;
  double globalDeltat() const;
  int globalIteration() const;
  std::vector<double> field(int field, int env = -1, int comp = 0) const;
  int nbEnv() const { return (int)envs_.size(); }

 private:
  void check(int st, const char* where) const;
  MhMesh mesh_;
  std::vector<EnvironmentConfig> envs_;
  MahycoOptions opt_;
  std::shared_ptr<IRemap> remap_;
  MhHandle* h_ = nullptr;
  friend class PerfectGasEOS;
  friend class StiffenedGasEOS;
  friend class FictifEOS;
  friend class IEquationOfState;
  friend class RemapADI;
};

}  // namespace mahyco

#endif
