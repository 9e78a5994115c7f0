// Host-side lowering of an eco_config into the kernel parameter blocks,
// shared by the single-solve contexts (eco_capi.cpp) and the pools
// (eco_pool.cpp).
#pragma once
#include <cstddef>
#include <vector>

#include "../../include/ecogen.h"
#include "eco_params.h"

namespace eco {

struct Lowered {
  Geometry geo;             // distance-table pointers set by bind_tables
  SchemeParams sch;
  std::vector<double> tab;  // per-direction distance tables (dcc, dfl, dfr)
  size_t off[3];            // offset of direction d's tables in tab
  int nprim;                // stored primitive fields on the device
  int nlevel;               // fields per time level: nprim, E, and sch.ncache cached values
  int ntr;                  // transported scalars (stored after a level's nlevel fields)
  size_t es;                // bytes per stored value (8 FP64, 4 FP32 storage)
};

// ECO_OK or the eco_status_code the configuration is refused with.
int validate_config(const eco_config *c);
// with_tables = false leaves L.tab empty (offsets are still set)
void lower_config(const eco_config &c, Lowered &L, bool with_tables = true);
void bind_tables(const eco_config &c, Lowered &L, double *dev_tables);
// host field index of each device primitive field (host layout of ecogen.h)
void host_field_map(const eco_config &c, int dim, std::vector<int> &map);
// eco_status of a solve from its control block (the next dt from the last
// step's wave-speed maximum, as Run::solver sets m_dt, src/Run.cpp:270)
void fill_status(const Control &h, double lcfl, double cfl, eco_status *st);

}  // namespace eco
