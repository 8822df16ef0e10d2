// mahyco_module.cpp -- MahycoModule / IRemap / IEquationOfState over the C ABI.
#include "mahyco_module.h"

#include <cstring>

namespace mahyco {

static const char* status_name(int st) {
  switch (st) {
    case MH_OK: return "MH_OK";
    case MH_STOPPED: return "MH_STOPPED";
    case MH_ERR_NEG_VOLUME: return "negative cell volume (MahycoModule.cc:1164-1191)";
    case MH_ERR_DT_MIN: return "time step below deltat-min (MahycoModule.cc:1816-1833)";
    case MH_ERR_NAN_REMAP: return "NaN density or energy after remap (RemapADIFinal.cc:274-284)";
    case MH_ERR_INVALID: return "invalid argument or call order";
    case MH_ERR_UNSUPPORTED: return "option outside the supported subset";
    case MH_ERR_DEVICE: return "HIP runtime error";
    case MH_ERR_NOMEM: return "device allocation failed";
    default: return "unknown status";
  }
}

MahycoError::MahycoError(int status, const std::string& where)
    : std::runtime_error(where + ": " + status_name(status)), status_(status) {}

// ---------------------------------------------------------------- EOS services
MhEnvironment PerfectGasEOS::description() const {
  MhEnvironment e;
  std::memset(&e, 0, sizeof(e));
  e.eos = MH_EOS_PERFECT_GAS;
  e.adiabatic_cst = o_.adiabatic_cst;
  e.specific_heat = o_.specific_heat;
  return e;
}
MhEnvironment StiffenedGasEOS::description() const {
  MhEnvironment e;
  std::memset(&e, 0, sizeof(e));
  e.eos = MH_EOS_STIFFENED_GAS;
  e.adiabatic_cst = o_.adiabatic_cst;
  e.limit_tension = o_.limit_tension;
  e.specific_heat = o_.specific_heat;
  return e;
}
MhEnvironment FictifEOS::description() const {
  MhEnvironment e;
  std::memset(&e, 0, sizeof(e));
  e.eos = MH_EOS_FICTIF;
  e.adiabatic_cst = 1.;
  e.fictif_tdebut = o_.tdebut_pression;
  e.fictif_tfin = o_.tfin_pression;
  e.fictif_p0 = o_.valeur_debut_pression;
  e.fictif_dpdt = o_.valeur_dependance_temps;
  return e;
}
void FictifEOS::applyEOS(int env) { module_->check(mahyco_apply_eos(module_->h_, env), "FictifEOS::applyEOS"); }
void IEquationOfState::applyOneCellEOS(int env, long cell) {
  module_->check(mahyco_apply_one_cell_eos(module_->h_, env, cell), "IEquationOfState::applyOneCellEOS");
}
// initEOS runs inside mahyco_create (HydroStartInit, MahycoModule.cc:143-152).
void PerfectGasEOS::initEOS(int) {}
void StiffenedGasEOS::initEOS(int) {}
void PerfectGasEOS::applyEOS(int env) {
  module_->check(mahyco_apply_eos(module_->h_, env), "PerfectGasEOS::applyEOS");
}
void StiffenedGasEOS::applyEOS(int env) {
  module_->check(mahyco_apply_eos(module_->h_, env), "StiffenedGasEOS::applyEOS");
}
static void check_damage_defaults(double tension, double density) {
  if (tension != -1.e12 || density != 0.)
    throw MahycoError(MH_ERR_UNSUPPORTED, "Endommagement with non-default thresholds");
}
void PerfectGasEOS::Endommagement(int) {
  check_damage_defaults(o_.tension_damage_thresold, o_.density_damage_thresold);
}
void StiffenedGasEOS::Endommagement(int) {
  check_damage_defaults(o_.tension_damage_thresold, o_.density_damage_thresold);
}

// ---------------------------------------------------------------- RemapADI
// withDualProjection is 0 in advection mode only (PrepareRemap.cc:211), which the device
// sequence knows from the options it was created with.
void RemapADI::appliRemap(int, int, int, int) { module_->check(mahyco_remap(module_->h_), "RemapADI::appliRemap"); }

// ---------------------------------------------------------------- MahycoModule
MahycoModule::MahycoModule(const MhMesh& mesh, std::vector<EnvironmentConfig> envs, MahycoOptions opt,
                           std::shared_ptr<IRemap> remap)
    : mesh_(mesh), envs_(std::move(envs)), opt_(std::move(opt)), remap_(std::move(remap)) {
  if (!remap_) remap_ = std::make_shared<RemapADI>();
  remap_->attach(this);
  for (auto& e : envs_) {
    if (!e.eos) throw MahycoError(MH_ERR_INVALID, "environment " + e.name + " has no eos-model");
    e.eos->attach(this);
  }
}

MahycoModule::~MahycoModule() { mahyco_destroy(h_); }

void MahycoModule::check(int st, const char* where) const {
  if (st < 0) throw MahycoError(st, where);
}

void MahycoModule::hydroStartInit() {
  if (opt_.pseudo_centree) throw MahycoError(MH_ERR_UNSUPPORTED, "pseudo-centree");
  static const char* kLengths[] = {"racine-cubique-volume", "faces-opposees", "monodimX", "monodimY", "monodimZ"};
  int lc = -1;
  for (int i = 0; i < 5; ++i)
    if (opt_.longueur_caracteristique == kLengths[i]) lc = i;
  if (lc < 0) throw MahycoError(MH_ERR_UNSUPPORTED, "longueur-caracteristique " + opt_.longueur_caracteristique);
  RemapADI* adi = dynamic_cast<RemapADI*>(remap_.get());
  if (!adi) throw MahycoError(MH_ERR_UNSUPPORTED, "remap service other than RemapADI");
  if (adi->isEuler() != "Full") throw MahycoError(MH_ERR_UNSUPPORTED, "RemapADI: is-euler-scheme other than Full");
  std::vector<MhEnvironment> ev;
  for (auto& e : envs_) {
    MhEnvironment d = e.eos->description();
    d.densite_initiale = e.densite_initiale;
    d.pression_initiale = e.pression_initiale;
    for (int i = 0; i < 3; ++i) d.vitesse_initiale[i] = e.vitesse_initiale[i];
    d.linear_pseudo_coeff = e.linear_pseudo_coeff;
    d.quadratic_pseudo_coeff = e.quadratic_pseudo_coeff;
    d.energie_initiale = e.energie_initiale;
    ev.push_back(d);
  }
  MhOptions o;
  std::memset(&o, 0, sizeof(o));
  o.cfl = opt_.cfl;
  o.threshold = opt_.threshold;
  o.deltat_init = opt_.deltat_init;
  o.deltat_min = opt_.deltat_min;
  o.deltat_max = opt_.deltat_max;
  o.final_time = opt_.final_time;
  o.with_projection = opt_.with_projection;
  o.calcul_vitson_max = opt_.calcul_vitson_max;
  if (opt_.boundary_condition.size() > MH_MAX_BC) throw MahycoError(MH_ERR_INVALID, "too many boundary conditions");
  o.nb_bc = (int)opt_.boundary_condition.size();
  for (int b = 0; b < o.nb_bc; ++b) o.bc[b] = opt_.boundary_condition[b];
  const RemapADIOptions& ro = adi->options();
  o.ordre_projection = adi->getOrdreProjection();
  o.projection_limiteur_id = ro.projection_limiteur_id;
  o.projection_limiteur_pure_id = ro.projection_limiteur_pure_id;
  o.projection_pente_borne = ro.projection_pente_borne;
  o.pente_borne_debar_fix = ro.projection_pente_borne_debar_fix;
  o.pente_borne_mixte = ro.projection_pente_borne_mixte;
  o.projection_simple_pente = ro.projection_simple_pente;
  o.conservation_energie_totale = adi->hasConservationEnergieTotale();
  o.remap_threshold = ro.threshold;
  o.with_newton = opt_.with_newton;
  o.pression_explicite = opt_.pression_explicite;
  o.schema_csts = opt_.schema_csts;
  o.longueur_caracteristique = lc;
  o.sans_lagrange = opt_.sans_lagrange;
  o.rider_reverse = opt_.cas.reverse_option;
  o.rider_parametre = opt_.cas.parametre;
  o.rider_rayon_bulle = opt_.cas.rayon_bulle;
  o.rider_densite_lineaire = opt_.cas.densite_interieure_initiale_lineaire;
  o.cas_model = opt_.cas_model;
  o.cas_test = opt_.cas_test;
  o.interface_position = opt_.interface_position;
  o.interface_axis = opt_.interface_axis;
  o.sedov_deposit = opt_.sedov_deposit;
  mahyco_destroy(h_);
  h_ = nullptr;
  check(mahyco_create(&mesh_, ev.data(), (int)ev.size(), &o, &h_), "MahycoModule::hydroStartInit");
}

void MahycoModule::hydroContinueInit() { check(mahyco_continue_init(h_), "hydroContinueInit"); }
void MahycoModule::writeCheckpoint(const std::string& path) {
  check(mahyco_checkpoint_write(h_, path.c_str()), "writeCheckpoint");
}
void MahycoModule::readCheckpoint(const std::string& path) {
  check(mahyco_checkpoint_read(h_, path.c_str()), "readCheckpoint");
}
void MahycoModule::saveValuesAtN() { check(mahyco_save_values_at_n(h_), "saveValuesAtN"); }
void MahycoModule::updateVelocity() { check(mahyco_update_velocity(h_), "updateVelocity"); }
void MahycoModule::applyBoundaryCondition() { check(mahyco_apply_boundary_condition(h_), "applyBoundaryCondition"); }
void MahycoModule::updatePosition() { check(mahyco_update_position(h_), "updatePosition"); }
void MahycoModule::computeGeometricValues() { check(mahyco_compute_geometric_values(h_), "computeGeometricValues"); }
void MahycoModule::updateDensity() { check(mahyco_update_density(h_), "updateDensity"); }
void MahycoModule::computeArtificialViscosity() {
  check(mahyco_compute_artificial_viscosity(h_), "computeArtificialViscosity");
}
void MahycoModule::updateElasticityAndPlasticity() {
  check(mahyco_update_elasticity_and_plasticity(h_), "updateElasticityAndPlasticity");
}
void MahycoModule::updateEnergyAndPressure() {
  check(mahyco_update_energy_and_pressure(h_), "updateEnergyAndPressure");
}
// PrepareRemap.cc:203-251: the remap service, then applyEOS and the mean pressure of
// every environment, which the device remap sequence already performs.
void MahycoModule::remap() {
  if (!opt_.with_projection) return;
  remap_->appliRemap(mesh_.dim, 1, 18 * nbEnv(), nbEnv());
}
void MahycoModule::computePressionMoyenne() {
  check(mahyco_compute_pression_moyenne(h_), "computePressionMoyenne");
}
void MahycoModule::computeDeltaT() {
  check(mahyco_compute_delta_t(h_), "computeDeltaT");
  check(mahyco_status(h_), "computeDeltaT");
}

bool MahycoModule::iterate() {
  saveValuesAtN();
  updateVelocity();
  applyBoundaryCondition();
  updatePosition();
  computeGeometricValues();
  updateDensity();
  computeArtificialViscosity();
  updateElasticityAndPlasticity();
  updateEnergyAndPressure();
  remap();
  computeDeltaT();
  return mahyco_status(h_) == MH_OK;
}

int MahycoModule::cycle(int n, int cycles_per_graph) {
  check(cycles_per_graph > 0 ? mahyco_cycle_graph(h_, n, cycles_per_graph) : mahyco_cycle(h_, n),
        "MahycoModule::cycle");
  const int st = mahyco_synchronize(h_);
  check(st, "MahycoModule::cycle");
  return st;
}

double MahycoModule::globalTime() const { return mahyco_time(h_); }
double MahycoModule::globalDeltat() const { return mahyco_deltat(h_); }
int MahycoModule::globalIteration() const { return mahyco_iteration(h_); }

std::vector<double> MahycoModule::field(int f, int env, int comp) const {
  const long n = f >= 100 ? mahyco_nb_nodes(h_) : mahyco_nb_cells(h_);
  std::vector<double> v(n);
  check(mahyco_get_field(h_, f, env, comp, v.data()), "MahycoModule::field");
  return v;
}

}  // namespace mahyco
