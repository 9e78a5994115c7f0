// ecogen.hpp -- C++ class API of the MI355X explicit finite-volume step.
//
// Keeps the reference's class names for this path -- Eos, Limiter,
// Relaxation, Flux, Model, and Run with its time loop -- as host-side descriptors that
// are configured like the reference's objects (virtual at configuration time)
// and lower to the POD eco_config of the C ABI (ecogen.h).  The numerical work
// runs in the gfx950 kernels behind that ABI; nothing here computes on the
// host.  Reference interfaces: src/Eos/Eos.h:49-165, src/Ordre2/Limiter.h:44-52,
// src/Relaxations/Relaxation.h:44-60, src/Models/Flux.h:46-131,
// src/Models/Model.h:47-177, src/Run.h:60-120.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ecogen.h"

namespace ecogen {

// --- Eos (src/Eos/Eos.h) -----------------------------------------------------
class Eos {
 public:
  virtual ~Eos() = default;
  virtual std::string getType() const = 0;
  virtual eco_eos lower() const = 0;
  virtual double getGamma() const = 0;
  virtual double getPInf() const { return 0.; }
  virtual double getCv() const = 0;
  virtual double getERef() const = 0;
  // entropyRef (the saturation temperature of Mixture::computeTsat)
  virtual double getSRef() const { return 0.; }
};

// Ideal gas (src/Eos/EosIG.cpp): parameters gamma, cv, energyRef of libEOS/IG_*.xml
class EosIG : public Eos {
 public:
  EosIG(double gamma, double cv, double eref = 0., double sref = 0.) : m_gamma(gamma), m_cv(cv), m_eRef(eref), m_sRef(sref) {}
  std::string getType() const override { return "IG"; }
  eco_eos lower() const override { return eco_eos{ECO_EOS_IG, m_gamma, 0., m_cv, m_eRef}; }
  double getGamma() const override { return m_gamma; }
  double getCv() const override { return m_cv; }
  double getERef() const override { return m_eRef; }
  double getSRef() const override { return m_sRef; }
 private:
  double m_gamma, m_cv, m_eRef, m_sRef;
};

// Stiffened gas (src/Eos/EosSG.cpp): gamma, pInf, cv, energyRef of libEOS/SG_*.xml
class EosSG : public Eos {
 public:
  EosSG(double gamma, double pinf, double cv, double eref = 0., double sref = 0.)
      : m_gamma(gamma), m_pInf(pinf), m_cv(cv), m_eRef(eref), m_sRef(sref) {}
  std::string getType() const override { return "SG"; }
  eco_eos lower() const override { return eco_eos{ECO_EOS_SG, m_gamma, m_pInf, m_cv, m_eRef}; }
  double getGamma() const override { return m_gamma; }
  double getPInf() const override { return m_pInf; }
  double getCv() const override { return m_cv; }
  double getERef() const override { return m_eRef; }
  double getSRef() const override { return m_sRef; }
 private:
  double m_gamma, m_pInf, m_cv, m_eRef, m_sRef;
};

// --- Limiter (src/Ordre2/Limiter.h) -------------------------------------------
class Limiter {
 public:
  virtual ~Limiter() = default;
  virtual int kind() const { return ECO_LIM_NONE; }
  virtual bool AmITHINC() const { return false; }
};
class LimiterMinmod : public Limiter { public: int kind() const override { return ECO_LIM_MINMOD; } };
class LimiterVanLeer : public Limiter { public: int kind() const override { return ECO_LIM_VANLEER; } };
class LimiterVanAlbada : public Limiter { public: int kind() const override { return ECO_LIM_VANALBADA; } };
class LimiterSuperBee : public Limiter { public: int kind() const override { return ECO_LIM_SUPERBEE; } };
class LimiterMC : public Limiter { public: int kind() const override { return ECO_LIM_MC; } };
class LimiterTHINC : public Limiter {
 public:
  int kind() const override { return ECO_LIM_THINC; }
  bool AmITHINC() const override { return true; }
};
// Factory by the names Input::entreeMain accepts (case-insensitive), Input.cpp:187-249.
std::unique_ptr<Limiter> makeLimiter(const std::string &name);

// --- Relaxation (src/Relaxations/Relaxation.h) --------------------------------
class Relaxation {
 public:
  virtual ~Relaxation() = default;
  virtual std::string getType() const = 0;
};
// Stiff mechanical (pressure) relaxation, per-cell FP64 Newton (RelaxationP.cpp:49-111)
class RelaxationP : public Relaxation {
 public:
  std::string getType() const override { return "P"; }
};
// Pressure-temperature relaxation in closed form, two phases (RelaxationPT.cpp:49-141),
// after the pressure relaxation
class RelaxationPT : public Relaxation {
 public:
  std::string getType() const override { return "PT"; }
};
// Relaxation to liquid-vapour saturation, two phases (RelaxationPTMu.cpp:51-157;
// phase 0 the liquid, phase 1 the vapour, as its constructor sets), after
// Kapila's pressure relaxation
class RelaxationPTMu : public Relaxation {
 public:
  std::string getType() const override { return "PTMU"; }
};

// --- Flux (src/Models/Flux.h): conservative vector of one model ---------------
class Flux {
 public:
  virtual ~Flux() = default;
  // names of the conservative variables in the order the device accumulates them
  virtual std::vector<std::string> variables(int numberPhases, int dim) const = 0;
};
class FluxEuler : public Flux {
 public:
  std::vector<std::string> variables(int, int dim) const override;
};
class FluxKapila : public Flux {
 public:
  std::vector<std::string> variables(int numberPhases, int dim) const override;
};
class FluxMultiP : public Flux {
 public:
  std::vector<std::string> variables(int numberPhases, int dim) const override;
};
class FluxThermalEq : public Flux {
 public:
  std::vector<std::string> variables(int numberPhases, int dim) const override;
};
class FluxEulerHomogeneous : public Flux {
 public:
  std::vector<std::string> variables(int numberPhases, int dim) const override;
};

// --- Model (src/Models/Model.h) -----------------------------------------------
class Model {
 public:
  virtual ~Model() = default;
  virtual std::string whoAmI() const = 0;
  virtual int kind() const = 0;
  virtual const Flux &flux() const = 0;
  int getNumberPhases() const { return m_numberPhases; }
  // transport equations of <flowModel numberTransports> (Model(name,
  // numberTransports), src/Models/Model.cpp:45-50; Kapila and MultiP)
  int getNumberTransports() const { return m_numberTransports; }
  void setNumberTransports(int n) { m_numberTransports = n; }
  std::vector<std::unique_ptr<Relaxation>> *getRelaxations() { return &m_relaxations; }
  const std::vector<std::unique_ptr<Relaxation>> &relaxations() const { return m_relaxations; }
 protected:
  explicit Model(int nph) : m_numberPhases(nph) {}
  int m_numberPhases;
  int m_numberTransports = 0;
  std::vector<std::unique_ptr<Relaxation>> m_relaxations;
};
class ModEuler : public Model {
 public:
  ModEuler() : Model(1) {}
  std::string whoAmI() const override { return "EULER"; }
  int kind() const override { return ECO_MODEL_EULER; }
  const Flux &flux() const override { return m_flux; }
 private:
  FluxEuler m_flux;
};
// Kapila's 5-equation model: RelaxationP is imposed by the model (ModKapila.cpp:47-53)
class ModKapila : public Model {
 public:
  explicit ModKapila(int nph, bool alphaNull = false) : Model(nph), m_alphaNull(alphaNull) {
    m_relaxations.emplace_back(new RelaxationP);
  }
  std::string whoAmI() const override { return "KAPILA"; }
  int kind() const override { return ECO_MODEL_KAPILA; }
  const Flux &flux() const override { return m_flux; }
  bool alphaNull() const { return m_alphaNull; }
 private:
  bool m_alphaNull;
  FluxKapila m_flux;
};
// Pressure-disequilibrium model; relaxation only when added (Input.cpp:608-621)
class ModMultiP : public Model {
 public:
  explicit ModMultiP(int nph) : Model(nph) {}
  std::string whoAmI() const override { return "MULTIP"; }
  int kind() const override { return ECO_MODEL_MULTIP; }
  const Flux &flux() const override { return m_flux; }
 private:
  FluxMultiP m_flux;
};

// Two phases in pressure and temperature equilibrium (ModThermalEq)
class ModThermalEq : public Model {
 public:
  ModThermalEq() : Model(2) {}
  std::string whoAmI() const override { return "THERMALEQ"; }
  int kind() const override { return ECO_MODEL_THERMALEQ; }
  const Flux &flux() const override { return m_flux; }
 private:
  FluxThermalEq m_flux;
};
// Homogeneous liquid-vapour mixture at saturation (ModEulerHomogeneous,
// <flowModel name="EulerHomogeneous" liquid vapor>, Input.cpp:479-483); first order
class ModEulerHomogeneous : public Model {
 public:
  ModEulerHomogeneous(int liquid, int vapor) : Model(2), m_liq(liquid), m_vap(vapor) {}
  std::string whoAmI() const override { return "EULERHOMOGENEOUS"; }
  int kind() const override { return ECO_MODEL_EULERHOM; }
  const Flux &flux() const override { return m_flux; }
  int getLiq() const { return m_liq; }
  int getVap() const { return m_vap; }
 private:
  int m_liq, m_vap;
  FluxEulerHomogeneous m_flux;
};

// --- Cartesian mesh and run control -------------------------------------------
struct CartesianMesh {
  int numberCells[3] = {1, 1, 1};
  double dimensions[3] = {1., 1., 1.};
  // boundary condition per side X-, X+, Y-, Y+, Z-, Z+ (initialConditions
  // <boundCond type="abs|wall|outflow|injection|tank">; all but abs run in FP64)
  int boundCond[6] = {ECO_BC_ABS, ECO_BC_ABS, ECO_BC_ABS, ECO_BC_ABS, ECO_BC_ABS, ECO_BC_ABS};
  double outflowPressure[6] = {0., 0., 0., 0., 0., 0.};  // <dataOutflow p0> of ECO_BC_OUTFLOW sides
  // ECO_BC_INJECTION sides (BoundCondInj): <dataInjection m0> and the injected
  // fluid's alpha, density and pressure per phase (<fluidsProp><dataFluid>)
  struct Injection {
    double m0 = 0.;
    double alpha[4] = {1., 0., 0., 0.}, density[4] = {0., 0., 0., 0.}, pressure[4] = {0., 0., 0., 0.};
  } injection[6];
  // ECO_BC_TANK sides (BoundCondTank): <dataTank p0 T0> and, for several
  // phases, the volume fractions (or mass fractions) per phase
  struct Tank {
    double p0 = 0., T0 = 0.;
    double fraction[4] = {1., 0., 0., 0.};
    bool massFractions = false;
  } tank[6];
  // <transport value> of each scalar entering through an outflow, injection
  // or tank side (m_valueTransport of BoundCondOutflow/Inj/Tank)
  double transportValue[6][ECO_MAX_TRANSPORTS] = {};
};

struct RunStatus {
  int iteration = 0;
  double physicalTime = 0.;
  double dt = 0.;
  unsigned flags = 0;
};

// Run: owns one device solve; mirrors Run::initialize / Run::solver /
// Run::finalize (src/Run.cpp:55-160, 204-277, 496-530) for this path.
class Run {
 public:
  Run(std::unique_ptr<Model> model, std::vector<std::unique_ptr<Eos>> eos, const CartesianMesh &mesh,
      double cfl, int order = 1, std::unique_ptr<Limiter> globalLimiter = nullptr,
      std::unique_ptr<Limiter> interfaceLimiter = nullptr,
      std::unique_ptr<Limiter> globalVolumeFractionLimiter = nullptr,
      std::unique_ptr<Limiter> interfaceVolumeFractionLimiter = nullptr,
      int precision = ECO_FP64, void *hipStream = nullptr);
  ~Run();
  Run(const Run &) = delete;
  Run &operator=(const Run &) = delete;

  const eco_config &config() const { return m_cfg; }
  long long numberCells() const;
  // state in the ecogen.h host layout (transported scalars after E)
  void setState(const std::vector<double> &state);
  // Run::resumeSimulation's clock: after setState, start at physicalTime and
  // take dt for the first step (both read from infoCalcul.out)
  void resumeClock(double physicalTime, double dt);
  std::vector<double> getState();
  // iteration control: nbIte steps; physical-time control: until t >= finalTime
  RunStatus solverIterations(int nbIte);
  RunStatus solverTime(double finalPhysicalTime);
  RunStatus step();
  eco_ctx *context() { return m_ctx; }

 private:
  std::unique_ptr<Model> m_model;
  std::vector<std::unique_ptr<Eos>> m_eos;
  std::unique_ptr<Limiter> m_lim[4];
  eco_config m_cfg;
  eco_ctx *m_ctx = nullptr;
};

// Lowers the descriptors to the C ABI configuration.
eco_config lowerConfig(const Model &model, const std::vector<std::unique_ptr<Eos>> &eos, const CartesianMesh &mesh,
                       double cfl, int order, const Limiter *lim[4], int precision);

class EcoError : public std::runtime_error {
 public:
  EcoError(int code, const std::string &where)
      : std::runtime_error(where + ": " + eco_strerror(code)), m_code(code) {}
  int code() const { return m_code; }
 private:
  int m_code;
};

}  // namespace ecogen
