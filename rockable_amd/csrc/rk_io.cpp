// rk_io.cpp -- readers and writer of Rockable's text formats.
//
//   .shp library   Shape::read            (Shape.cpp:69-128)
//   .conf deck     Rockable::initParser   (Rockable.cpp:643-1069), loadConf 1085-1145
//   drivingSystem  DrivingSystem::read    (DrivingSystem.cpp:44-268)
//   conf output    Rockable::saveConf     (Rockable.cpp:440-610)
//
// Keywords are read with the kwParser conventions: a token starting with
// '/', '#' or '!' comments out the rest of its line; unknown keywords are
// skipped.
#include <fenv.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>

#include "rk_model.hpp"

namespace rk {

const char* const kParamName[P_COUNT] = {
    "knContact",    "en2Contact",   "ktContact",    "muContact",     "krContact",    "murContact",
    "viscTContact", "knInnerBond",  "ktInnerBond",  "krInnerBond",   "en2InnerBond", "fn0InnerBond",
    "ft0InnerBond", "mom0InnerBond", "powInnerBond", "knOuterBond",  "ktOuterBond",  "krOuterBond",
    "en2OuterBond", "fn0OuterBond", "ft0OuterBond", "mom0OuterBond", "powOuterBond", "gcInnerBond",
    "gcOuterBond"};

// order in which saveConf writes the law data (Rockable.cpp:513-541)
static const int kParamWriteOrder[P_COUNT] = {
    P_KN,     P_EN2,     P_KT,     P_MU,      P_KR,       P_MUR,      P_VISCT,   P_KN_IN,  P_KT_IN,
    P_KR_IN,  P_FN0_IN,  P_FT0_IN, P_MOM0_IN, P_POW_IN,   P_EN2_IN,   P_KN_OUT,  P_KT_OUT, P_KR_OUT,
    P_FN0_OUT, P_FT0_OUT, P_MOM0_OUT, P_POW_OUT, P_EN2_OUT, P_GC_IN, P_GC_OUT};

void Model::set_param(int id, int g1, int g2, double v) {
  if (g1 < 0 || g2 < 0 || g1 >= kMaxGroups || g2 >= kMaxGroups) return;
  table[(id * kMaxGroups + g1) * kMaxGroups + g2] = v;
  table[(id * kMaxGroups + g2) * kMaxGroups + g1] = v;
  defined[(id * kMaxGroups + g1) * kMaxGroups + g2] = 1;
  defined[(id * kMaxGroups + g2) * kMaxGroups + g1] = 1;
  ngroup = std::max(ngroup, std::max(g1, g2) + 1);
}

namespace {

// The whole file in memory, NUL-terminated; numbers are parsed in place
// (strtod / strtol stop at the whitespace that ends the token, so the value
// is the one they give for the token alone).
struct Reader {
  std::string buf;
  size_t pos = 0;
  bool open(const std::string& file) {
    FILE* f = fopen(file.c_str(), "rb");
    if (!f) return false;
    buf.clear();
    char chunk[1 << 16];
    if (fseek(f, 0, SEEK_END) == 0) {
      const long n = ftell(f);
      if (n > 0) buf.reserve((size_t)n + 1);
      rewind(f);
    }
    size_t k;
    while ((k = fread(chunk, 1, sizeof chunk, f)) > 0) buf.append(chunk, k);
    const bool ok = !ferror(f);
    fclose(f);
    pos = 0;
    return ok;
  }
  static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
  // start of the next token, or false at the end of the file
  bool skip() {
    while (pos < buf.size() && ws(buf[pos])) pos++;
    return pos < buf.size();
  }
  void past() {
    while (pos < buf.size() && !ws(buf[pos])) pos++;
  }
  bool tok(std::string& out) {
    if (!skip()) return false;
    size_t b = pos;
    past();
    out.assign(buf, b, pos - b);
    return true;
  }
  void skipline() {
    while (pos < buf.size() && buf[pos] != '\n') pos++;
  }
  double d() {
    if (!skip()) return 0.0;
    const double v = strtod(buf.c_str() + pos, nullptr);
    past();
    return v;
  }
  long l() {
    if (!skip()) return 0;
    const long v = strtol(buf.c_str() + pos, nullptr, 10);
    past();
    return v;
  }
  v3 vec() {
    v3 v;
    v.x = d();
    v.y = d();
    v.z = d();
    return v;
  }
  qt quat() {
    qt q;
    q.w = d();
    q.x = d();
    q.y = d();
    q.z = d();
    return q;
  }
};

bool is_comment(const std::string& t) { return !t.empty() && (t[0] == '/' || t[0] == '#' || t[0] == '!'); }

std::string dirname_of(const std::string& p) {
  size_t k = p.rfind('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

obb_t default_obb() {
  obb_t o;
  o.c = V(0, 0, 0);
  o.e[0] = V(1, 0, 0);
  o.e[1] = V(0, 1, 0);
  o.e[2] = V(0, 0, 1);
  o.ext = V(0, 0, 0);
  return o;
}

}  // namespace

int load_shapes(Model& m, const std::string& file, std::string& err) {
  Reader r;
  if (!r.open(file)) {
    err = "shape library '" + file + "' has not been found";
    return -1;
  }
  m.shapes.clear();
  std::string t;
  while (r.tok(t)) {
    if (t != "<") continue;
    Shape sh;
    sh.obb = default_obb();
    while (r.tok(t)) {
      if (is_comment(t)) {
        r.skipline();
        continue;
      }
      if (t == ">") break;
      if (t == "name") r.tok(sh.name);
      else if (t == "radius") sh.radius = r.d();
      else if (t == "preCompDone") {
        std::string c;
        r.tok(c);
        sh.preCompDone = c.empty() ? 'n' : c[0];
      } else if (t == "volume") sh.volume = r.d();
      else if (t == "I/m") sh.inertia_mass = r.vec();
      else if (t == "obb.extent") sh.obb.ext = r.vec();
      else if (t == "obb.center") sh.obb.c = r.vec();
      else if (t == "obb.e1") sh.obb.e[0] = r.vec();
      else if (t == "obb.e2") sh.obb.e[1] = r.vec();
      else if (t == "obb.e3") sh.obb.e[2] = r.vec();
      else if (t == "OBBtreeLevel" || t == "MCnstep" || t == "fibObbOption") r.d();
      else if (t == "position") r.vec();
      else if (t == "orientation") r.quat();
      else if (t == "isSurface") {
      } else if (t == "nv") {
        long n = r.l();
        sh.vertex.resize(n > 0 ? n : 0);
        for (long v = 0; v < n; v++) sh.vertex[v] = r.vec();
      } else if (t == "ne") {
        long n = r.l();
        sh.edge.resize(2 * (n > 0 ? n : 0));
        for (long e = 0; e < n; e++) {
          sh.edge[2 * e] = (int)r.l();
          sh.edge[2 * e + 1] = (int)r.l();
        }
      } else if (t == "nf") {
        long n = r.l();
        sh.face_start.assign(1, 0);
        for (long f = 0; f < n; f++) {
          long k = r.l();
          for (long q = 0; q < k; q++) sh.face_idx.push_back((int)r.l());
          sh.face_start.push_back((int)sh.face_idx.size());
        }
      }
    }
    if (sh.preCompDone == 'n') {
      err = "shape '" + sh.name + "': preCompDone n (Monte-Carlo mass properties) is not supported";
      return -1;
    }
    if (sh.face_start.empty()) sh.face_start.assign(1, 0);
    // validate the tables the kernels index
    for (int e : sh.edge)
      if (e < 0 || e >= sh.nv()) {
        err = "shape '" + sh.name + "': edge vertex out of range";
        return -1;
      }
    for (int f = 0; f < sh.nf(); f++) {
      if (sh.face_start[f + 1] - sh.face_start[f] < 3) {
        err = "shape '" + sh.name + "': a face has fewer than 3 vertices";
        return -1;
      }
    }
    for (int v : sh.face_idx)
      if (v < 0 || v >= sh.nv()) {
        err = "shape '" + sh.name + "': face vertex out of range";
        return -1;
      }
    if (sh.nv() < 1) {
      err = "shape '" + sh.name + "' has no vertex";
      return -1;
    }
    m.shapes.push_back(std::move(sh));
  }
  return 0;
}

static int shape_index(const Model& m, const std::string& name) {
  for (size_t k = 0; k < m.shapes.size(); k++)
    if (m.shapes[k].name == name) return (int)k;
  return -1;
}

static int parse_conf(Model& m, const std::string& fullpath, std::string& err) {
  Reader r;
  if (!r.open(fullpath)) {
    err = "cannot read " + fullpath;
    return -1;
  }
  std::string t;
  if (!r.tok(t)) {
    err = "empty conf file " + fullpath;
    return -1;
  }
  if (t == "redirection") {  // Rockable.cpp:1096-1105
    std::string dir, name;
    r.tok(dir);
    r.tok(name);
    m.path = dirname_of(fullpath) + dir;
    return parse_conf(m, m.path + name, err);
  }
  r.tok(t);  // version date
  while (r.tok(t)) {
    if (is_comment(t)) {
      r.skipline();
      continue;
    }
    if (t == "t") m.t = r.d();
    else if (t == "tmax") m.tmax = r.d();
    else if (t == "dt") {
      m.dt = r.d();
      m.dt_2 = 0.5 * m.dt;
      m.dt2 = m.dt * m.dt;
      m.dt2_2 = 0.5 * m.dt2;
    } else if (t == "interVerlet") m.interVerlet = r.d();
    else if (t == "interConf") m.interConf = r.d();
    else if (t == "DVerlet") m.DVerlet = r.d();
    else if (t == "dVerlet") m.dVerlet = r.d();
    else if (t == "parallel_mode") r.tok(t);
    else if (t == "density") {
      long g = r.l();
      double d = r.d();
      if (g >= 0 && g < kMaxGroups) {
        m.density[g] = d;
        m.ndensity = std::max(m.ndensity, (int)g + 1);
      }
    } else if (t == "gravity") m.gravity = r.vec();
    else if (t == "ParamsInInterfaces") m.paramsInInterfaces = (int)r.l();
    else if (t == "dynamicUpdateNL") m.dynamicUpdateNL = (int)r.l();
    else if (t == "dispUpdateNL") m.dispUpdateNL = r.d();
    else if (t == "angleUpdateNL") m.angleUpdateNL = r.d() * (M_PI / 180.0);
    else if (t == "numericalDampingCoeff") m.numericalDampingCoeff = r.d();
    else if (t == "VelocityBarrier") m.velocityBarrier = r.d();
    else if (t == "AngularVelocityBarrier") m.angularVelocityBarrier = r.d();
    else if (t == "VelocityBarrierExponent") m.velocityBarrierExponent = r.d();
    else if (t == "AngularVelocityBarrierExponent") m.angularVelocityBarrierExponent = r.d();
    else if (t == "preventCrossingLength") m.preventCrossingLength = r.d();
    else if (t == "forceLaw") {
      std::string name;
      r.tok(name);
      // unknown names fall back to Default (Rockable.cpp:763-781)
      if (name == "Avalanche") {
        m.law = LAW_AVALANCHE;
        m.lawName = name;
      } else if (name == "StickedLinks") {
        m.law = LAW_STICKED;
        m.lawName = name;
      } else {
        m.law = LAW_DEFAULT;
        m.lawName = "Default";
      }
    } else if (t == "AddOrRemoveInteractions") r.tok(t);
    else if (t == "UpdateNL") {
      std::string name;
      r.tok(name);
      if (name == "linkCells") m.nl_linkcells = 1;
      else if (name == "bruteForce") m.nl_linkcells = 0;
    } else if (t == "Integrator") {
      std::string name;
      r.tok(name);
      if (name != "velocityVerlet") {
        err = "Integrator " + name + " is not supported (velocityVerlet only)";
        return -1;
      }
    } else if (t == "cellMinSizes") m.cellMinSizes = r.vec();
    else if (t == "boxForLinkCellsOpt") m.boxForLinkCellsOpt = (int)r.l();
    else if (t == "ContactPartnership") {
      std::string name;
      r.tok(name);
      if (name != "None") {
        err = "ContactPartnership " + name + " is not supported";
        return -1;
      }
    } else if (t == "iconf") m.iconf = (int)r.l();
    else if (t == "nDriven") m.nDriven = (int)r.l();
    else if (t == "precision") m.precision = (int)r.l();
    else if (t == "shapeFile") {
      std::string name;
      r.tok(name);
      std::string want = m.path + name;
      if (want != m.shapeFile) {
        m.shapeFile = want;
        if (load_shapes(m, want, err) != 0) return -1;
      }
    } else if (t == "Particles") {
      long nb = r.l();
      m.bodies.clear();
      m.bodies.reserve(nb > 0 ? nb : 0);
      for (long i = 0; i < nb; i++) {
        std::string nm;
        if (!r.tok(nm)) break;
        if (nm[0] == '#') {
          r.skipline();
          --i;
          continue;
        }
        if (nm[0] == '!') {
          r.skipline();
          --i;
          --nb;
          continue;
        }
        Body P;
        P.shape_name = nm;
        const long group = r.l();
        if (group < 0 || group >= kMaxGroups) {
          err = "body " + std::to_string(i) + ": group " + std::to_string(group) + " outside [0, " +
                std::to_string(kMaxGroups) + ")";
          return -1;
        }
        P.group = (int)group;
        P.cluster = (int)r.l();
        P.homothety = r.d();
        P.pos = r.vec();
        P.vel = r.vec();
        P.acc = r.vec();
        P.Q = r.quat();
        P.vrot = r.vec();
        P.arot = r.vec();
        P.shape = shape_index(m, nm);
        if (P.shape < 0) {
          err = "unknown shape '" + nm + "'";
          return -1;
        }
        double h = P.homothety;
        const double dens = m.density[P.group];
        P.mass = (h * h * h * m.shapes[P.shape].volume) * dens;
        P.inertia = vsc(P.mass, vsc(h * h, m.shapes[P.shape].inertia_mass));
        m.bodies.push_back(P);
      }
    } else if (t == "Interactions") {
      long n = r.l();
      m.inter.clear();
      m.inter.reserve(n > 0 ? n : 0);
      for (long k = 0; k < n; k++) {
        Inter I;
        I.i = (int)r.l();
        I.j = (int)r.l();
        I.type = (int)r.l();
        I.isub = (int)r.l();
        I.jsub = (int)r.l();
        I.n = r.vec();
        I.dn = r.d();
        I.pos = r.vec();
        I.vel = r.vec();
        I.fn = r.d();
        I.ft = r.vec();
        I.mom = r.vec();
        I.damp = r.d();
        I.stick = -1;
        m.inter.push_back(I);
      }
      // std::set semantics: lexicographic order, first of equal keys kept
      std::stable_sort(m.inter.begin(), m.inter.end(), key_less);
      m.inter.erase(std::unique(m.inter.begin(), m.inter.end(),
                                [](const Inter& a, const Inter& b) { return !key_less(a, b) && !key_less(b, a); }),
                    m.inter.end());
    } else if (t == "Interfaces") {
      long n = r.l();
      m.ifc.clear();
      for (long k = 0; k < n; k++) {
        Interface BI;
        BI.i = (int)r.l();
        BI.j = (int)r.l();
        long nb = r.l();
        BI.dn0 = r.d();
        if (m.paramsInInterfaces == 1) {
          BI.kn = r.d();
          BI.kt = r.d();
          BI.kr = r.d();
          BI.fn0 = r.d();
          BI.ft0 = r.d();
          BI.mom0 = r.d();
          BI.power = r.d();
          BI.Gc = r.d();
        }
        for (long u = 0; u < nb; u++) {
          BI.bkey.push_back((int)r.l());
          BI.bkey.push_back((int)r.l());
          BI.bkey.push_back((int)r.l());
        }
        m.ifc.push_back(BI);
      }
    } else if (t == "BodyForce") {
      std::string name;
      r.tok(name);
      if (name == "AttractingPoint") {
        m.bf = BF_ATTRACTING;
        m.bf_point = r.vec();
        m.bf_acc = r.d();
      } else if (name == "PreferredDirection") {
        m.bf = BF_PREFDIR;
        m.bf_axisBody = r.vec();
        m.bf_axis = r.vec();
        m.bf_momentMax = r.d();
        vnormalize(&m.bf_axisBody);
        vnormalize(&m.bf_axis);
        m.bf_Kr = 2.0 * m.bf_momentMax / M_PI;
      } else if (name == "ViscousFluid") {
        m.bf = BF_VISCOUS;
        m.bf_fluidDensity = r.d();
      }
      // unknown names: warning in the reference, no body force
    } else {
      int id = -1;
      for (int p = 0; p < P_COUNT; p++)
        if (t == kParamName[p]) id = p;
      if (id >= 0) {
        int g1 = (int)r.l(), g2 = (int)r.l();
        double v = r.d();
        m.set_param(id, g1, g2, v);
      }
    }
  }
  return 0;
}

int load_conf(Model& m, const std::string& file, std::string& err) {
  m.path = dirname_of(file);
  if (parse_conf(m, file, err) != 0) return -1;
  if (m.numericalDampingCoeff > 0.0 || m.velocityBarrier > 0.0 || m.angularVelocityBarrier > 0.0) {
    err = "numerical damping and velocity barriers are not supported";
    return -1;
  }
  const int nb = (int)m.bodies.size();
  for (const Inter& I : m.inter)
    if (I.i < 0 || I.i >= nb || I.j < 0 || I.j >= nb || I.type < 0 || I.type > 3) {
      err = "interaction record out of range";
      return -1;
    }
  for (const Inter& I : m.inter) {
    const Shape& si = m.shapes[m.bodies[I.i].shape];
    const Shape& sj = m.shapes[m.bodies[I.j].shape];
    int ni = (I.type == EE) ? si.ne() : si.nv();
    int nj = (I.type == VV) ? sj.nv() : (I.type == VE || I.type == EE) ? sj.ne() : sj.nf();
    if (I.isub < 0 || I.isub >= ni || I.jsub < 0 || I.jsub >= nj) {
      err = "interaction sub-element index out of range";
      return -1;
    }
  }
  // Link interfaces to their bonds (Rockable.cpp:956-1002); interfaces whose
  // bonds are not all listed, or that repeat an (i, j), are dropped.
  // (the pairs already kept are looked up in a hash set: decks hold
  // hundreds of thousands of interfaces)
  std::vector<Interface> kept;
  std::unordered_set<uint64_t> kept_pairs;
  kept_pairs.reserve(2 * m.ifc.size());
  for (Interface& BI : m.ifc) {
    if (BI.i < 0 || BI.i >= nb || BI.j < 0 || BI.j >= nb) {
      err = "interface body index out of range";
      return -1;
    }
    bool missed = false;
    const uint64_t pair = ((uint64_t)(uint32_t)BI.i << 32) | (uint32_t)BI.j;
    const bool dup = kept_pairs.count(pair) != 0;
    for (size_t u = 0; u < BI.bkey.size(); u += 3) {
      Inter key;
      key.i = BI.i;
      key.j = BI.j;
      key.type = BI.bkey[u];
      key.isub = BI.bkey[u + 1];
      key.jsub = BI.bkey[u + 2];
      auto it = std::lower_bound(m.inter.begin(), m.inter.end(), key, key_less);
      if (it == m.inter.end() || key_less(key, *it)) missed = true;
    }
    if (!missed && !dup) {
      kept.push_back(BI);
      kept_pairs.insert(pair);
    }
  }
  std::sort(kept.begin(), kept.end(),
            [](const Interface& a, const Interface& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
  m.ifc = kept;
  for (Inter& I : m.inter) I.stick = -1;
  for (size_t id = 0; id < m.ifc.size(); id++) {
    const Interface& BI = m.ifc[id];
    for (size_t u = 0; u < BI.bkey.size(); u += 3) {
      Inter key;
      key.i = BI.i;
      key.j = BI.j;
      key.type = BI.bkey[u];
      key.isub = BI.bkey[u + 1];
      key.jsub = BI.bkey[u + 2];
      auto it = std::lower_bound(m.inter.begin(), m.inter.end(), key, key_less);
      it->stick = (int)id;
    }
  }
  return 0;
}

int load_driving(Model& m, const std::string& file, std::string& err) {
  Reader r;
  m.driving_path = file;  // read again at each conf save, even if absent now
  if (!r.open(file)) return 0;
  static const char* names[] = {"_x_Vel_",  "_y_Vel_",   "_z_Vel_",    "_xrot_Vel_", "_yrot_Vel_",
                                "_zrot_Vel_", "_x_For_", "_y_For_",    "_z_For_",    "_xrot_Mom_",
                                "_yrot_Mom_", "_zrot_Mom_", "_xyzrot_Vel_", "_xyzrot_Mom_"};
  static const int codes[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 100, 101};
  m.ctrl.clear();  // the servo function stays until a Servo line replaces it
  std::string t;
  while (r.tok(t)) {
    if (is_comment(t)) {
      r.skipline();
      continue;
    }
    if (t == "Control") {
      std::string ty;
      r.tok(ty);
      Control c;
      for (int k = 0; k < 14; k++)
        if (ty == names[k]) c.type = codes[k];
      c.i = (int)r.l();
      if (c.type >= 100)
        c.vec = r.vec();
      else
        c.value = r.d();
      m.ctrl.push_back(c);
    } else if (t == "Servo") {
      std::string nm;
      r.tok(nm);
      if (nm.compare(0, 6, "tritri") == 0) {  // DrivingSystem.cpp:191-392
        int id[6];
        for (int k = 0; k < 6; k++) id[k] = (int)r.l();
        const int icontr = (int)m.ctrl.size();
        const int types[6] = {C_X_VEL, C_X_FOR, C_Y_VEL, C_Y_FOR, C_Z_VEL, C_Z_FOR};
        for (int k = 0; k < 6; k++) {
          Control c;
          c.type = types[k];
          c.i = id[k];
          m.ctrl.push_back(c);
        }
        int kind = 0;
        if (nm == "tritriIsostaticCompression") {
          kind = 1;
          m.servo_pressure = r.d();
        } else if (nm == "tritriBiaxialCompression") {
          kind = 2;
          m.servo_pressure = r.d();
          m.servo_velocity = r.d();
          m.ctrl[icontr + 3].type = C_Y_VEL;
          m.ctrl[icontr + 3].value = -m.servo_velocity;
        } else if (nm == "tritriCustom") {
          kind = 3;
          const int vel[3] = {C_X_VEL, C_Y_VEL, C_Z_VEL}, frc[3] = {C_X_FOR, C_Y_FOR, C_Z_FOR};
          for (int k = 0; k < 6; k++) {
            const long ty = r.l();
            m.servo_val[k] = r.d();
            m.ctrl[icontr + k].type = ty == 0 ? vel[k / 2] : frc[k / 2];
            m.ctrl[icontr + k].value = (k & 1) ? -m.servo_val[k] : m.servo_val[k];
          }
        } else if (nm == "tritriLodeAngle") {
          kind = 4;
          m.servo_pressure = r.d();
          const double lode = r.d();
          m.lode_rate = r.d();
          const double rad = lode * M_PI / 180.0;
          m.lode_a = (3.0 * tan(rad) - sqrt(3.0)) / (2.0 * sqrt(3.0));
          m.lode_b = -(1.0 + m.lode_a);
        }
        if (kind) {  // an unknown tritri name keeps its controls and adds no function
          m.servo = kind;
          m.servo_icontr = icontr;
          for (int k = 0; k < 6; k++) m.servo_ids[k] = id[k];
        }
      } else if (nm == "shaker" || nm == "triangle_shaker" || nm == "sawtooth_shaker") {  // 397-477
        Control c;
        c.i = (int)r.l();
        v3 dir = r.vec();
        const double A = r.d(), freq = r.d();
        const double tini = nm == "sawtooth_shaker" ? r.d() : 0.0;
        vnormalize(&dir);
        m.shake_dir = dir;
        m.shake_A = A;
        m.shake_omega = 2.0 * M_PI * freq;
        m.shake_T = 1.0 / freq;
        m.shake_tini = tini;
        m.servo_icontr = (int)m.ctrl.size();
        m.servo = nm == "shaker" ? 5 : nm == "triangle_shaker" ? 6 : 7;
        const int types[3] = {C_X_VEL, C_Y_VEL, C_Z_VEL};
        for (int k = 0; k < 3; k++) {
          c.type = types[k];
          m.ctrl.push_back(c);
        }
      } else if (nm == "ramp") {  // 479-494
        std::string ty;
        r.tok(ty);
        Control c;
        c.type = 0;  // an unknown control name maps to _x_Vel_ (std::map default)
        for (int k = 0; k < 14; k++)
          if (ty == names[k]) c.type = codes[k];
        c.i = (int)r.l();
        m.ramp_vb = r.d();
        m.ramp_ve = r.d();
        m.ramp_tb = r.d();
        m.ramp_te = r.d();
        c.value = m.ramp_vb;
        m.servo_icontr = (int)m.ctrl.size();
        m.servo = 8;
        m.ctrl.push_back(c);
      }
      // other servo names: ignored, as the reference's chain of names ignores them
    }
  }
  for (const Control& c : m.ctrl)
    if (c.i < 0 || c.i >= m.nDriven) {
      err = "controls must address driven bodies (index < nDriven)";
      return -1;
    }
  if (m.servo >= 1 && m.servo <= 4)
    for (int k = 0; k < 6; k++)
      if (m.servo_ids[k] < 0 || m.servo_ids[k] >= (int)m.bodies.size()) {
        err = "servo wall index out of range";
        return -1;
      }
  return 0;
}

void Model::servo_time_values(double t, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  if (servo >= 5 && servo <= 7) {
    v3 v;
    if (servo == 5) {  // DrivingSystem.cpp:413-418
      v = vsc(shake_A * shake_omega * cos(shake_omega * t), shake_dir);
    } else {  // 436-444, 467-475
      const double T = shake_T;
      v = vsc(4.0 * shake_A / T, shake_dir);
      double phase = servo == 6 ? t - std::floor(t / T) * T : t - std::floor((t - shake_tini) / T) * T;
      phase /= T;
      if (phase >= 0.25 && phase < 0.75) v = vsc(-1.0, v);
    }
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
  } else if (servo == 8) {  // 485-493
    if (!ramp_slope_set) {
      ramp_slope = (ramp_ve - ramp_vb) / (ramp_te - ramp_tb);
      ramp_slope_set = true;
    }
    if (t <= ramp_tb)
      out[0] = ramp_vb;
    else if (t >= ramp_te)
      out[0] = ramp_ve;
    else
      out[0] = ramp_vb + ramp_slope * (t - ramp_tb);
  }
}

double pair_damp(const Model& m, int i, int j) {
  const Body& bi = m.bodies[i];
  const Body& bj = m.bodies[j];
  double en2 = m.param(P_EN2, bi.group, bj.group);
  double kn = m.param(P_KN, bi.group, bj.group);
  double meff;
  if (i < m.nDriven)
    meff = bj.mass;
  else if (j < m.nDriven)
    meff = bi.mass;
  else
    meff = (bi.mass * bj.mass) / (bi.mass + bj.mass);
  if (en2 > 0.0 && en2 < 1.0) {
    double logen = 0.5 * log(en2);
    double dampRate = -logen / sqrt(logen * logen + M_PI * M_PI);
    return dampRate * 2.0 * sqrt(kn * meff);
  } else if (en2 <= 0.0) {
    return 2.0 * sqrt(kn * meff);
  }
  return 0.0;
}

namespace {
void put_e(FILE* f, int prec, double x) { fprintf(f, "%.*e", prec, x); }
void put_v(FILE* f, int prec, v3 v) {
  put_e(f, prec, v.x);
  fputc(' ', f);
  put_e(f, prec, v.y);
  fputc(' ', f);
  put_e(f, prec, v.z);
}
}  // namespace

int save_conf(const Model& m, const std::vector<Inter>& active, const std::string& file, std::string& err) {
  FILE* f = fopen(file.c_str(), "w");
  if (!f) {
    err = "cannot write " + file;
    return -1;
  }
  fprintf(f, "Rockable 21-08-2022\n");
  fprintf(f, "AddOrRemoveInteractions bruteForce\n");
  fprintf(f, "Integrator velocityVerlet\n");
  fprintf(f, "UpdateNL %s\n", m.nl_linkcells ? "linkCells" : "bruteForce");
  fprintf(f, "forceLaw %s\n", m.lawName.c_str());
  fprintf(f, "t %g\n", m.t);
  fprintf(f, "tmax %g\n", m.tmax);
  fprintf(f, "dt %g\n", m.dt);
  fprintf(f, "interVerlet %g\n", m.interVerlet);
  fprintf(f, "interConf %g\n", m.interConf);
  fprintf(f, "DVerlet %g\n", m.DVerlet);
  fprintf(f, "dVerlet %g\n", m.dVerlet);
  for (int g = 0; g < m.ndensity; g++)
    if (m.density[g] > 0.0) fprintf(f, "density %d %g\n", g, m.density[g]);
  fprintf(f, "gravity %g %g %g\n", m.gravity.x, m.gravity.y, m.gravity.z);
  if (m.bf == BF_ATTRACTING)
    fprintf(f, "BodyForce AttractingPoint %g %g %g %g\n", m.bf_point.x, m.bf_point.y, m.bf_point.z, m.bf_acc);
  else if (m.bf == BF_PREFDIR)
    fprintf(f, "BodyForce PreferredDirection %g %g %g %g %g %g %g\n", m.bf_axisBody.x, m.bf_axisBody.y,
            m.bf_axisBody.z, m.bf_axis.x, m.bf_axis.y, m.bf_axis.z, m.bf_momentMax);
  else if (m.bf == BF_VISCOUS)
    fprintf(f, "BodyForce ViscousFluid %g\n", m.bf_fluidDensity);
  fprintf(f, "ParamsInInterfaces %d\n", m.paramsInInterfaces);
  fprintf(f, "dynamicUpdateNL %d\n", m.dynamicUpdateNL);
  if (m.dynamicUpdateNL != 0) {
    fprintf(f, "dispUpdateNL %g\n", m.dispUpdateNL);
    fprintf(f, "angleUpdateNL %g\n", m.angleUpdateNL * (180.0 / M_PI));
  }
  fprintf(f, "numericalDampingCoeff %g\n", m.numericalDampingCoeff);
  fprintf(f, "VelocityBarrier %g\n", m.velocityBarrier);
  fprintf(f, "AngularVelocityBarrier %g\n", m.angularVelocityBarrier);
  fprintf(f, "VelocityBarrierExponent %g\n", m.velocityBarrierExponent);
  fprintf(f, "AngularVelocityBarrierExponent %g\n", m.angularVelocityBarrierExponent);
  for (int q = 0; q < P_COUNT; q++) {
    int id = kParamWriteOrder[q];
    for (int g1 = 0; g1 < m.ngroup; g1++)
      for (int g2 = g1; g2 < m.ngroup; g2++)
        if (m.defined[(id * kMaxGroups + g1) * kMaxGroups + g2])
          fprintf(f, "%s %d %d %g\n", kParamName[id], g1, g2, m.param(id, g1, g2));
  }
  fprintf(f, "ContactPartnership None\n");
  fprintf(f, "cellMinSizes %g %g %g\n", m.cellMinSizes.x, m.cellMinSizes.y, m.cellMinSizes.z);
  fprintf(f, "boxForLinkCellsOpt %d\n", m.boxForLinkCellsOpt);
  fprintf(f, "iconf %d\n", m.iconf);
  fprintf(f, "nDriven %d\n", m.nDriven);
  {
    size_t k = m.shapeFile.rfind('/');
    fprintf(f, "shapeFile %s\n", k == std::string::npos ? m.shapeFile.c_str() : m.shapeFile.c_str() + k + 1);
  }
  const int prec = m.precision;
  fprintf(f, "precision %d\n", prec);
  fprintf(f, "Particles %zu\n", m.bodies.size());
  fprintf(f,
          "#name group cluster homothety pos.x pos.y pos.z vel.x vel.y vel.z acc.x acc.y acc.z Q.w Q.x Q.y Q.z "
          "vrot.x vrot.y vrot.z arot.x arot.y arot.z\n");
  for (const Body& p : m.bodies) {
    fprintf(f, "%s %d %d ", m.shapes[p.shape].name.c_str(), p.group, p.cluster);
    put_e(f, prec, p.homothety);
    fputc(' ', f);
    put_v(f, prec, p.pos);
    fputc(' ', f);
    put_v(f, prec, p.vel);
    fputc(' ', f);
    put_v(f, prec, p.acc);
    fputc(' ', f);
    put_e(f, prec, p.Q.w);
    fputc(' ', f);
    put_e(f, prec, p.Q.x);
    fputc(' ', f);
    put_e(f, prec, p.Q.y);
    fputc(' ', f);
    put_e(f, prec, p.Q.z);
    fputc(' ', f);
    put_v(f, prec, p.vrot);
    fputc(' ', f);
    put_v(f, prec, p.arot);
    fputc('\n', f);
  }
  fprintf(f, "Interactions %zu\n", active.size());
  const int prev = fegetround();
  for (const Inter& I : active) {
    fprintf(f, "%d %d %d %d %d ", I.i, I.j, I.type, I.isub, I.jsub);
    fflush(f);
    fesetround(FE_TOWARDZERO);  // n is truncated to the given precision
    put_v(f, prec, I.n);
    fflush(f);
    fesetround(prev);
    fputc(' ', f);
    put_e(f, prec, I.dn);
    fputc(' ', f);
    put_v(f, prec, I.pos);
    fputc(' ', f);
    put_v(f, prec, I.vel);
    fputc(' ', f);
    put_e(f, prec, I.fn);
    fputc(' ', f);
    put_v(f, prec, I.ft);
    fputc(' ', f);
    put_v(f, prec, I.mom);
    fputc(' ', f);
    put_e(f, prec, I.damp);
    fputc('\n', f);
  }
  size_t nalive = 0;
  for (const Interface& BI : m.ifc) nalive += BI.alive ? 1 : 0;
  fprintf(f, "Interfaces %zu\n", nalive);
  for (const Interface& BI : m.ifc) {
    if (!BI.alive) continue;
    fprintf(f, "%d %d  %zu %.6f", BI.i, BI.j, BI.bkey.size() / 3, BI.dn0);
    if (m.paramsInInterfaces == 1)
      fprintf(f, " %.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f ", BI.kn, BI.kt, BI.kr, BI.fn0, BI.ft0, BI.mom0,
              BI.power, BI.Gc);
    for (size_t u = 0; u < BI.bkey.size(); u += 3)
      fprintf(f, "  %d %d %d", BI.bkey[u], BI.bkey[u + 1], BI.bkey[u + 2]);
    fputc('\n', f);
  }
  fclose(f);
  return 0;
}

}  // namespace rk
