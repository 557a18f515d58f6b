// rk_model.hpp -- host-side model of a Rockable "box": scalars, group-pair
// parameter tables, shape library, bodies, interactions, breakable
// interfaces, driving controls and body force.  It is what the .conf / .shp
// readers fill and what the conf writer prints; the engine mirrors it in
// device memory.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "rk_math.hpp"

namespace rk {

enum { VV = 0, VE = 1, VF = 2, EE = 3 };
enum { LAW_DEFAULT = 0, LAW_AVALANCHE = 1, LAW_STICKED = 2 };
enum { BF_NONE = 0, BF_ATTRACTING = 1, BF_PREFDIR = 2, BF_VISCOUS = 3 };

// DataTable parameter ids, in the registration order of Rockable.cpp:96-123.
enum {
  P_KN = 0, P_EN2, P_KT, P_MU, P_KR, P_MUR, P_VISCT,
  P_KN_IN, P_KT_IN, P_KR_IN, P_EN2_IN, P_FN0_IN, P_FT0_IN, P_MOM0_IN, P_POW_IN,
  P_KN_OUT, P_KT_OUT, P_KR_OUT, P_EN2_OUT, P_FN0_OUT, P_FT0_OUT, P_MOM0_OUT, P_POW_OUT,
  P_GC_IN, P_GC_OUT,
  P_COUNT
};
extern const char* const kParamName[P_COUNT];
constexpr int kMaxGroups = 256;

// Control codes (DrivingSystem.hpp:52-67).
enum {
  C_X_VEL = 0, C_Y_VEL = 1, C_Z_VEL = 2, C_XROT_VEL = 3, C_YROT_VEL = 4, C_ZROT_VEL = 5,
  C_X_FOR = 6, C_Y_FOR = 7, C_Z_FOR = 8, C_XROT_MOM = 9, C_YROT_MOM = 10, C_ZROT_MOM = 11,
  C_XYZROT_VEL = 100, C_XYZROT_MOM = 101
};

struct Shape {
  std::string name;
  double radius = 0.0, volume = 0.0;
  v3 inertia_mass{0, 0, 0};
  obb_t obb;
  char preCompDone = 'n';
  std::vector<v3> vertex;
  std::vector<int> edge;        // 2 per edge
  std::vector<int> face_start;  // nf + 1
  std::vector<int> face_idx;
  int nv() const { return (int)vertex.size(); }
  int ne() const { return (int)edge.size() / 2; }
  int nf() const { return face_start.empty() ? 0 : (int)face_start.size() - 1; }
};

struct Body {
  std::string shape_name;
  int shape = -1, group = 0, cluster = 0;
  double homothety = 1.0;
  v3 pos{0, 0, 0}, vel{0, 0, 0}, acc{0, 0, 0};
  qt Q{1, 0, 0, 0};
  v3 vrot{0, 0, 0}, arot{0, 0, 0};
  double mass = 0.0;
  v3 inertia{0, 0, 0};
  v3 force{0, 0, 0}, moment{0, 0, 0};
};

// One interaction record; key (i, j, type, isub, jsub) (Interaction.hpp:57-99).
struct Inter {
  int32_t i = 0, j = 0, type = 0, isub = 0, jsub = 0, stick = -1;
  v3 n{0, 0, 0};
  double dn = 0.0;
  v3 pos{0, 0, 0}, vel{0, 0, 0};
  double fn = 0.0;
  v3 ft{0, 0, 0}, mom{0, 0, 0};
  double damp = 0.0;
};

inline bool key_less(const Inter& a, const Inter& b) {
  if (a.i != b.i) return a.i < b.i;
  if (a.j != b.j) return a.j < b.j;
  if (a.type != b.type) return a.type < b.type;
  if (a.isub != b.isub) return a.isub < b.isub;
  return a.jsub < b.jsub;
}

// BreakableInterface (BreakableInterface.hpp:48-77) with its bond keys
// (type, isub, jsub) owned by body i.
struct Interface {
  int i = 0, j = 0;
  double kn = 0, kt = 0, kr = 0, fn0 = 0, ft0 = 0, mom0 = 0, power = 2.0, Gc = 0;
  double dn0 = 0.0, area = 0.0;
  int isInner = 1;
  int alive = 1;
  std::vector<int> bkey;
};

struct Control {
  int type = 0, i = 0;
  double value = 0.0;
  v3 vec{0, 0, 0};
};

struct Model {
  // scalars (Rockable.hpp defaults)
  double t = 0.0, tmax = 1.0, dt = 1e-6, dt_2 = 0.5e-6, dt2 = 1e-12, dt2_2 = 0.5e-12;
  double interVerlet = 0.01, interConf = 0.25, DVerlet = 0.0, dVerlet = 0.0;
  double interVerletC = 0.0, interConfC = 0.0;
  int paramsInInterfaces = 0, dynamicUpdateNL = 0;
  double dispUpdateNL = 1.0, angleUpdateNL = 1.0;
  double numericalDampingCoeff = 0.0, velocityBarrier = 0.0, angularVelocityBarrier = 0.0;
  double velocityBarrierExponent = 1.0, angularVelocityBarrierExponent = 1.0, preventCrossingLength = 0.0;
  v3 gravity{0.0, -9.81, 0.0};
  v3 cellMinSizes{1.0, 1.0, 1.0};
  int boxForLinkCellsOpt = 0;
  int iconf = 0;
  int nDriven = 0;
  int precision = 16;
  int law = LAW_DEFAULT;
  std::string lawName = "Default";
  int nl_linkcells = 0;
  // data table
  std::vector<double> table = std::vector<double>(P_COUNT * kMaxGroups * kMaxGroups, 0.0);
  std::vector<unsigned char> defined = std::vector<unsigned char>(P_COUNT * kMaxGroups * kMaxGroups, 0);
  int ngroup = 0;
  std::vector<double> density = std::vector<double>(kMaxGroups, 0.0);
  int ndensity = 0;
  // library and state
  std::string path;       // directory the conf was read from (m_path)
  std::string shapeFile;  // path of the loaded library
  std::vector<Shape> shapes;
  std::vector<Body> bodies;
  std::vector<Inter> inter;  // loaded interactions, lexicographic order
  std::vector<Interface> ifc;  // sorted by (i, j)
  // driving system
  std::vector<Control> ctrl;
  // ServoFunction (DrivingSystem.cpp:188-496), the last Servo line read:
  // 0 none, 1 tritriIsostaticCompression, 2 tritriBiaxialCompression,
  // 3 tritriCustom, 4 tritriLodeAngle, 5 shaker, 6 triangle_shaker,
  // 7 sawtooth_shaker, 8 ramp
  int servo = 0;
  int servo_icontr = 0, servo_ids[6] = {0, 0, 0, 0, 0, 0};
  double servo_pressure = 0.0, servo_velocity = 0.0;
  double servo_val[6] = {0, 0, 0, 0, 0, 0};
  double lode_a = 0.0, lode_b = 0.0, lode_rate = 0.0;
  v3 shake_dir{0, 0, 0};
  double shake_A = 0.0, shake_omega = 0.0, shake_T = 1.0, shake_tini = 0.0;
  double ramp_vb = 0.0, ramp_ve = 0.0, ramp_tb = 0.0, ramp_te = 0.0, ramp_slope = 0.0;
  bool ramp_slope_set = false;  // the slope is a function-static of the ramp's first call
  std::string driving_path;     // re-read at each conf save (Rockable.cpp:2682-2688)
  // servos 4-8 depend on the time; their steps are launched one by one
  bool servo_timed() const { return servo >= 4; }
  // the controls' values of a time-only servo (5-8) at time t (the lambdas
  // of DrivingSystem.cpp:413-493, evaluated on the host with the C library)
  void servo_time_values(double t, double out[3]);
  // body force
  int bf = BF_NONE;
  v3 bf_point{0, 0, 0};
  double bf_acc = 0.0;
  v3 bf_axisBody{0, 0, 0}, bf_axis{0, 0, 0};
  double bf_momentMax = 0.0, bf_Kr = 0.0, bf_fluidDensity = 0.0;
  double totalBrokenArea = 0.0;

  double param(int id, int g1, int g2) const {
    if (g1 < 0 || g2 < 0 || g1 >= kMaxGroups || g2 >= kMaxGroups) return 0.0;
    return table[(id * kMaxGroups + g1) * kMaxGroups + g2];
  }
  void set_param(int id, int g1, int g2, double v);
};

// .shp reader (Shape::read, Shape.cpp:69-128); returns 0 or -1 with err set.
int load_shapes(Model& m, const std::string& file, std::string& err);
// conf reader (Rockable::loadConf / initParser); does not run any step.
int load_conf(Model& m, const std::string& file, std::string& err);
// drivingSystem.txt reader (DrivingSystem::read); a missing file is not an
// error (the first nDriven bodies then stay fixed).
int load_driving(Model& m, const std::string& file, std::string& err);
// conf writer (Rockable::saveConf): bodies from m.bodies, interactions from
// `active` (already in lexicographic order).
int save_conf(const Model& m, const std::vector<Inter>& active, const std::string& file, std::string& err);
// Damping coefficient of a body pair (Rockable.cpp:1356-1376).
double pair_damp(const Model& m, int i, int j);

// Body ranges of one packing in a merged batch.
struct PackingRange {
  int first_driven = 0, n_driven = 0, first_free = 0, n_free = 0;
};
// rk_batch.cpp: several loaded decks as one model (driven bodies of every
// deck first, then free bodies; groups, shapes, interactions, interfaces and
// controls renumbered).  Returns 0 or -1 with err set.
int merge_models(std::vector<Model>& parts, Model& out, std::vector<PackingRange>& ranges, std::string& err);

}  // namespace rk
