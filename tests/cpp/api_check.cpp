// Checks of the C++ class API (include/ecogen.hpp), built against the
// in-tree libecogen_amd.so by ecogen_amd.build.
//
//   api_check lower            host only: limiter cascade, model and EOS lowering
//   api_check run <out.bin>    GPU: the 1-D Kapila first-order water/air shock
//                              tube (1000 cells, to 240 us) through Run; writes
//                              steps, t and the final state (host layout)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ecogen.hpp"

using namespace ecogen;

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); fails++; } \
  } while (0)

static std::vector<std::unique_ptr<Eos>> air_water() {
  std::vector<std::unique_ptr<Eos>> eos;
  eos.emplace_back(new EosIG(1.4, 800.));
  eos.emplace_back(new EosSG(4.4, 6.e8, 1000.));
  return eos;
}

static int check_lower() {
  // limiter names as Input::entreeMain accepts them, case-insensitive
  CHECK(makeLimiter("minmod")->kind() == ECO_LIM_MINMOD);
  CHECK(makeLimiter("VanLeer")->kind() == ECO_LIM_VANLEER);
  CHECK(makeLimiter("THINC")->AmITHINC());
  CartesianMesh m;
  m.numberCells[0] = 64; m.numberCells[1] = 32;
  m.dimensions[0] = 2.; m.dimensions[1] = 1.;
  auto eos = air_water();
  ModKapila kap(2, true);
  std::unique_ptr<Limiter> g = makeLimiter("vanleer"), iv = makeLimiter("thinc");
  const Limiter *l[4] = {g.get(), nullptr, nullptr, iv.get()};
  eco_config c = lowerConfig(kap, eos, m, 0.8, 2, l, ECO_FP64);
  // cascade: interface <- global, global VF <- global, interface VF given
  CHECK(c.lim[0] == ECO_LIM_VANLEER && c.lim[1] == ECO_LIM_VANLEER);
  CHECK(c.lim[2] == ECO_LIM_VANLEER && c.lim[3] == ECO_LIM_THINC);
  CHECK(c.model == ECO_MODEL_KAPILA && c.nphase == 2 && c.alpha_null == 1 && c.order == 2);
  CHECK(c.n[0] == 64 && c.n[1] == 32 && c.n[2] == 1 && c.L[0] == 2.);
  CHECK(c.eos[0].type == ECO_EOS_IG && c.eos[1].type == ECO_EOS_SG && c.eos[1].pinf == 6.e8);
  CHECK(kap.relaxations().size() == 1 && kap.relaxations()[0]->getType() == "P");
  // first order ignores limiters
  c = lowerConfig(kap, eos, m, 0.8, 1, l, ECO_FP64);
  CHECK(c.lim[0] == ECO_LIM_NONE && c.lim[3] == ECO_LIM_NONE);
  // MultiP relaxes only when a RelaxationP is added
  ModMultiP mp(2);
  c = lowerConfig(mp, eos, m, 0.5, 2, l, ECO_FP32S);
  CHECK(c.relax_p == 0 && c.precision == ECO_FP32S);
  mp.getRelaxations()->emplace_back(new RelaxationP);
  c = lowerConfig(mp, eos, m, 0.5, 2, l, ECO_FP32S);
  CHECK(c.relax_p == 1 && c.relax_pt == 0);
  mp.getRelaxations()->emplace_back(new RelaxationPT);
  c = lowerConfig(mp, eos, m, 0.5, 2, l, ECO_FP64);
  CHECK(c.relax_p == 1 && c.relax_pt == 1);
  // transport equations: the model's count and the sides' entering values
  ModKapila kt(2, false);
  kt.setNumberTransports(2);
  m.boundCond[1] = ECO_BC_OUTFLOW;
  m.transportValue[1][1] = 7.;
  c = lowerConfig(kt, eos, m, 0.8, 2, l, ECO_FP64);
  CHECK(c.ntransport == 2 && c.bc_tr[1][1] == 7. && c.bc_tr[1][0] == 0.);
  // the Euler model has none (Model.h:127): refused before any device call
  try {
    std::unique_ptr<Model> eu(new ModEuler);
    eu->setNumberTransports(1);
    std::vector<std::unique_ptr<Eos>> e1;
    e1.emplace_back(new EosIG(1.4, 800.));
    Run r(std::move(eu), std::move(e1), m, 0.8, 1);
    CHECK(false);
  } catch (const EcoError &e) {
    CHECK(e.code() == ECO_ERR_INVALID);
  }
  // conservative variables in accumulation order
  std::vector<std::string> v = kap.flux().variables(2, 2);
  CHECK(v.size() == 9 && v.front() == "alpha_0" && v.back() == "rho_E");
  std::printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}

static int run_shock_tube(const char *out) {
  const int n = 1000;
  CartesianMesh m;
  m.numberCells[0] = n;
  m.dimensions[0] = 1.; m.dimensions[1] = 0.1; m.dimensions[2] = 0.1;
  Run run(std::unique_ptr<Model>(new ModKapila(2, false)), air_water(), m, 0.8, 1);
  const int nf = ECO_NFIELDS(2);
  std::vector<double> st((size_t)nf * n, 0.);
  for (int i = 0; i < n; i++) {
    const double x = (i + 0.5) * (1. / n);
    const bool hp = x <= 0.7;
    const double aa = hp ? 1e-6 : 1. - 1e-6, p = hp ? 1e9 : 1e5;  // air fraction
    st[0 * n + i] = aa;      st[1 * n + i] = 50.;   st[2 * n + i] = p;
    st[3 * n + i] = 1. - aa; st[4 * n + i] = 1000.; st[5 * n + i] = p;
  }
  run.setState(st);
  // "resume": restart the same state at t = 1e-5 with dt = 3e-8 (Run::resumeClock)
  const bool resume = std::getenv("API_CHECK_RESUME") != nullptr;
  if (resume) run.resumeClock(1e-5, 3e-8);
  RunStatus s = resume ? run.solverIterations(20) : run.solverTime(240e-6);
  std::vector<double> res = run.getState();
  FILE *f = std::fopen(out, "wb");
  if (!f) return 2;
  const double hdr[2] = {(double)s.iteration, s.physicalTime};
  std::fwrite(hdr, sizeof(double), 2, f);
  std::fwrite(res.data(), sizeof(double), res.size(), f);
  std::fclose(f);
  std::printf("steps %d t %.17g flags %u\n", s.iteration, s.physicalTime, s.flags);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "lower") == 0) return check_lower();
  if (argc >= 3 && std::strcmp(argv[1], "run") == 0) {
    try {
      return run_shock_tube(argv[2]);
    } catch (const std::exception &e) {
      std::printf("error: %s\n", e.what());
      return 3;
    }
  }
  std::printf("usage: api_check lower | run <out.bin>\n");
  return 1;
}
