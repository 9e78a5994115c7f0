// Host descriptors of the C++ class API (include/ecogen.hpp).
#include <algorithm>
#include <cctype>
#include <cstring>

#include "../../include/ecogen.hpp"

namespace ecogen {

std::unique_ptr<Limiter> makeLimiter(const std::string &name) {
  std::string n = name;
  std::transform(n.begin(), n.end(), n.begin(), [](unsigned char c) { return (char)std::toupper(c); });
  if (n == "MINMOD") return std::unique_ptr<Limiter>(new LimiterMinmod);
  if (n == "VANLEER") return std::unique_ptr<Limiter>(new LimiterVanLeer);
  if (n == "VANALBADA") return std::unique_ptr<Limiter>(new LimiterVanAlbada);
  if (n == "SUPERBEE") return std::unique_ptr<Limiter>(new LimiterSuperBee);
  if (n == "MC") return std::unique_ptr<Limiter>(new LimiterMC);
  if (n == "THINC") return std::unique_ptr<Limiter>(new LimiterTHINC);
  throw EcoError(ECO_ERR_INVALID, "makeLimiter(" + name + ")");
}

static void add_momentum(std::vector<std::string> &v, int dim) {
  const char *c[3] = {"rho_u", "rho_v", "rho_w"};
  for (int d = 0; d < dim; d++) v.push_back(c[d]);
}

std::vector<std::string> FluxEuler::variables(int, int dim) const {
  std::vector<std::string> v = {"rho"};
  add_momentum(v, dim);
  v.push_back("rho_E");
  return v;
}

static std::vector<std::string> multiphase_vars(int nph, int dim) {
  std::vector<std::string> v;
  for (int k = 0; k < nph; k++) {
    v.push_back("alpha_" + std::to_string(k));
    v.push_back("alpha_rho_" + std::to_string(k));
    v.push_back("alpha_rho_e_" + std::to_string(k));
  }
  add_momentum(v, dim);
  v.push_back("rho_E");
  return v;
}
std::vector<std::string> FluxKapila::variables(int nph, int dim) const { return multiphase_vars(nph, dim); }
std::vector<std::string> FluxMultiP::variables(int nph, int dim) const { return multiphase_vars(nph, dim); }
// FluxThermalEq: the phases' masses, momentum, total energy (FluxThermalEq.cpp:113-126)
std::vector<std::string> FluxThermalEq::variables(int nph, int dim) const {
  std::vector<std::string> v;
  for (int k = 0; k < nph; k++) v.push_back("alpha_rho_" + std::to_string(k));
  add_momentum(v, dim);
  v.push_back("rho_E");
  return v;
}
// FluxEulerHomogeneous: the mixture's mass, momentum, total energy (FluxEulerHomogeneous.cpp:70-89)
std::vector<std::string> FluxEulerHomogeneous::variables(int, int dim) const {
  std::vector<std::string> v{"rho"};
  add_momentum(v, dim);
  v.push_back("rho_E");
  return v;
}

eco_config lowerConfig(const Model &model, const std::vector<std::unique_ptr<Eos>> &eos, const CartesianMesh &mesh,
                       double cfl, int order, const Limiter *lim[4], int precision) {
  eco_config c;
  std::memset(&c, 0, sizeof(c));
  c.model = model.kind();
  c.nphase = model.getNumberPhases();
  c.order = order;
  // cascade of defaults of Input::entreeMain (Input.cpp:203-249):
  // interface <- global, global VF <- global, interface VF <- interface
  int g = lim[0] ? lim[0]->kind() : ECO_LIM_NONE;
  int i = lim[1] ? lim[1]->kind() : g;
  int gv = lim[2] ? lim[2]->kind() : g;
  int iv = lim[3] ? lim[3]->kind() : i;
  c.lim[0] = g; c.lim[1] = i; c.lim[2] = gv; c.lim[3] = iv;
  if (order == 1) c.lim[0] = c.lim[1] = c.lim[2] = c.lim[3] = ECO_LIM_NONE;
  if (const ModKapila *k = dynamic_cast<const ModKapila *>(&model)) c.alpha_null = k->alphaNull() ? 1 : 0;
  c.relax_p = 0;
  if (model.kind() == ECO_MODEL_MULTIP)
    for (const auto &r : model.relaxations()) if (r->getType() == "P") c.relax_p = 1;
  for (const auto &r : model.relaxations()) {
    if (r->getType() == "PT") c.relax_pt = 1;
    if (r->getType() == "PTMU") c.relax_ptmu = 1;
  }
  c.liq = 0;
  c.vap = 1;
  if (const ModEulerHomogeneous *h = dynamic_cast<const ModEulerHomogeneous *>(&model)) {
    c.liq = h->getLiq();
    c.vap = h->getVap();
  }
  c.precision = precision;
  for (int d = 0; d < 3; d++) { c.n[d] = mesh.numberCells[d]; c.L[d] = mesh.dimensions[d]; }
  c.cfl = cfl;
  for (int b = 0; b < 6; b++) {
    c.bc[b] = mesh.boundCond[b];
    c.bc_p0[b] = mesh.outflowPressure[b];
    c.bc_inj_m0[b] = mesh.injection[b].m0;
    for (int k = 0; k < 4; k++) {
      c.bc_inj_alpha[b][k] = mesh.injection[b].alpha[k];
      c.bc_inj_rho[b][k] = mesh.injection[b].density[k];
      c.bc_inj_p[b][k] = mesh.injection[b].pressure[k];
      c.bc_tank_frac[b][k] = mesh.tank[b].fraction[k];
    }
    c.bc_tank_p0[b] = mesh.tank[b].p0;
    c.bc_tank_T0[b] = mesh.tank[b].T0;
    c.bc_tank_massfrac[b] = mesh.tank[b].massFractions ? 1 : 0;
    for (int t = 0; t < ECO_MAX_TRANSPORTS; t++) c.bc_tr[b][t] = mesh.transportValue[b][t];
  }
  c.ntransport = model.getNumberTransports();
  for (size_t k = 0; k < eos.size() && k < 4; k++) {
    c.eos[k] = eos[k]->lower();
    c.sref[k] = eos[k]->getSRef();
  }
  return c;
}

Run::Run(std::unique_ptr<Model> model, std::vector<std::unique_ptr<Eos>> eos, const CartesianMesh &mesh, double cfl,
         int order, std::unique_ptr<Limiter> g, std::unique_ptr<Limiter> i, std::unique_ptr<Limiter> gv,
         std::unique_ptr<Limiter> iv, int precision, void *hipStream)
    : m_model(std::move(model)), m_eos(std::move(eos)) {
  m_lim[0] = std::move(g); m_lim[1] = std::move(i); m_lim[2] = std::move(gv); m_lim[3] = std::move(iv);
  const Limiter *l[4] = {m_lim[0].get(), m_lim[1].get(), m_lim[2].get(), m_lim[3].get()};
  m_cfg = lowerConfig(*m_model, m_eos, mesh, cfl, order, l, precision);
  int rc = eco_create(&m_cfg, hipStream, &m_ctx);
  if (rc) throw EcoError(rc, "eco_create");
}

Run::~Run() { eco_destroy(m_ctx); }

long long Run::numberCells() const { return (long long)m_cfg.n[0] * m_cfg.n[1] * m_cfg.n[2]; }

void Run::setState(const std::vector<double> &state) {
  if ((long long)state.size() < ECO_NFIELDS_TR(m_cfg.nphase, m_cfg.ntransport) * numberCells())
    throw EcoError(ECO_ERR_INVALID, "Run::setState");
  int rc = eco_set_state(m_ctx, state.data());
  if (rc) throw EcoError(rc, "eco_set_state");
}

void Run::resumeClock(double physicalTime, double dt) {
  int rc = eco_set_clock(m_ctx, physicalTime, dt);
  if (rc) throw EcoError(rc, "eco_set_clock");
}

std::vector<double> Run::getState() {
  std::vector<double> s((size_t)ECO_NFIELDS_TR(m_cfg.nphase, m_cfg.ntransport) * numberCells());
  int rc = eco_get_state(m_ctx, s.data());
  if (rc) throw EcoError(rc, "eco_get_state");
  return s;
}

static RunStatus to_status(const eco_status &s) {
  RunStatus r;
  r.iteration = s.steps;
  r.physicalTime = s.t;
  r.dt = s.dt_last;
  r.flags = s.flags;
  return r;
}

RunStatus Run::solverIterations(int nbIte) {
  eco_status s;
  int rc = eco_run(m_ctx, nbIte, 0., &s);
  if (rc) throw EcoError(rc, "eco_run");
  return to_status(s);
}

RunStatus Run::solverTime(double finalPhysicalTime) {
  eco_status s;
  int rc = eco_run(m_ctx, 0, finalPhysicalTime, &s);
  if (rc) throw EcoError(rc, "eco_run");
  return to_status(s);
}

RunStatus Run::step() {
  eco_status s;
  int rc = eco_step(m_ctx, &s);
  if (rc) throw EcoError(rc, "eco_step");
  return to_status(s);
}

}  // namespace ecogen
