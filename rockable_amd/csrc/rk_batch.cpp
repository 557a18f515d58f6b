// rk_batch.cpp -- several decks (independent packings) in one model.
//
// The packings keep their own bodies, groups, parameters, interactions and
// interfaces; only indices are renumbered.  Rockable requires the driven
// bodies to come first (nDriven is a prefix of the body list,
// Rockable.cpp:1979-2158), so the merged list holds every packing's driven
// bodies, packing after packing, then every packing's free bodies.  Both maps
// are increasing within a packing, so the key order of its interactions and
// the serial order of its per-body sums are its own: each packing of the
// batch computes what it computes alone, provided no body of another
// packing comes near it (the packings must be apart in space).
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "rk_model.hpp"

namespace rk {

namespace {

bool same_shape(const Shape& a, const Shape& b) {
  if (a.radius != b.radius || a.volume != b.volume || a.vertex.size() != b.vertex.size() || a.edge != b.edge ||
      a.face_start != b.face_start || a.face_idx != b.face_idx)
    return false;
  if (a.inertia_mass.x != b.inertia_mass.x || a.inertia_mass.y != b.inertia_mass.y ||
      a.inertia_mass.z != b.inertia_mass.z)
    return false;
  for (size_t k = 0; k < a.vertex.size(); k++)
    if (a.vertex[k].x != b.vertex[k].x || a.vertex[k].y != b.vertex[k].y || a.vertex[k].z != b.vertex[k].z)
      return false;
  return std::memcmp(&a.obb, &b.obb, sizeof(obb_t)) == 0;
}

bool same_v3(const v3& a, const v3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

}  // namespace

int merge_models(std::vector<Model>& parts, Model& out, std::vector<PackingRange>& ranges, std::string& err) {
  if (parts.empty()) {
    err = "batch: no deck";
    return -1;
  }
  const Model& m0 = parts[0];
  for (size_t p = 1; p < parts.size(); p++) {
    const Model& m = parts[p];
    const char* what = nullptr;
    if (m.law != m0.law) what = "forceLaw";
    else if (m.dt != m0.dt) what = "dt";
    else if (m.interVerlet != m0.interVerlet) what = "interVerlet";
    else if (m.interConf != m0.interConf) what = "interConf";
    else if (m.DVerlet != m0.DVerlet || m.dVerlet != m0.dVerlet) what = "DVerlet / dVerlet";
    else if (m.dynamicUpdateNL != m0.dynamicUpdateNL || m.dispUpdateNL != m0.dispUpdateNL ||
             m.angleUpdateNL != m0.angleUpdateNL)
      what = "dynamicUpdateNL";
    else if (m.paramsInInterfaces != m0.paramsInInterfaces) what = "ParamsInInterfaces";
    else if (m.nl_linkcells != m0.nl_linkcells) what = "UpdateNL";
    else if (!same_v3(m.cellMinSizes, m0.cellMinSizes) || m.boxForLinkCellsOpt != m0.boxForLinkCellsOpt)
      what = "cellMinSizes / boxForLinkCellsOpt";
    else if (!same_v3(m.gravity, m0.gravity)) what = "gravity";
    else if (m.bf != m0.bf) what = "BodyForce";
    else if (m.t != m0.t || m.tmax != m0.tmax) what = "t / tmax";
    if (what) {
      err = std::string("batch: deck ") + std::to_string(p) + " differs from deck 0 in " + what;
      return -1;
    }
  }
  for (const Model& m : parts)
    if (m.servo != 0) {
      err = "batch: servo controls are not supported in a batch";
      return -1;
    }

  out = Model();
  // scalars, body force and options of deck 0 (equal in every deck)
  out.t = m0.t;
  out.tmax = m0.tmax;
  out.dt = m0.dt;
  out.dt_2 = m0.dt_2;
  out.dt2 = m0.dt2;
  out.dt2_2 = m0.dt2_2;
  out.interVerlet = m0.interVerlet;
  out.interConf = m0.interConf;
  out.DVerlet = m0.DVerlet;
  out.dVerlet = m0.dVerlet;
  out.paramsInInterfaces = m0.paramsInInterfaces;
  out.dynamicUpdateNL = m0.dynamicUpdateNL;
  out.dispUpdateNL = m0.dispUpdateNL;
  out.angleUpdateNL = m0.angleUpdateNL;
  out.gravity = m0.gravity;
  out.cellMinSizes = m0.cellMinSizes;
  out.boxForLinkCellsOpt = m0.boxForLinkCellsOpt;
  out.iconf = m0.iconf;
  out.precision = m0.precision;
  out.law = m0.law;
  out.lawName = m0.lawName;
  out.nl_linkcells = m0.nl_linkcells;
  out.path = m0.path;
  out.shapeFile = m0.shapeFile;
  out.bf = m0.bf;
  out.bf_point = m0.bf_point;
  out.bf_acc = m0.bf_acc;
  out.bf_axisBody = m0.bf_axisBody;
  out.bf_axis = m0.bf_axis;
  out.bf_momentMax = m0.bf_momentMax;
  out.bf_Kr = m0.bf_Kr;
  out.bf_fluidDensity = m0.bf_fluidDensity;

  // shapes: one library, a name standing for one shape
  std::map<std::string, int> shape_id;
  std::vector<std::vector<int>> shape_map(parts.size());
  for (size_t p = 0; p < parts.size(); p++)
    for (const Shape& s : parts[p].shapes) {
      auto it = shape_id.find(s.name);
      if (it == shape_id.end()) {
        shape_id[s.name] = (int)out.shapes.size();
        shape_map[p].push_back((int)out.shapes.size());
        out.shapes.push_back(s);
      } else {
        if (!same_shape(out.shapes[it->second], s)) {
          err = "batch: shape '" + s.name + "' differs between decks";
          return -1;
        }
        shape_map[p].push_back(it->second);
      }
    }

  // groups: the groups each deck's bodies use, renumbered in order
  std::vector<std::map<int, int>> group_map(parts.size());
  int ng = 0;
  for (size_t p = 0; p < parts.size(); p++) {
    std::vector<int> used;
    for (const Body& b : parts[p].bodies) used.push_back(b.group);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    for (int g : used) {
      if (ng >= kMaxGroups) {
        err = "batch: more than " + std::to_string(kMaxGroups) + " groups in all";
        return -1;
      }
      group_map[p][g] = ng++;
    }
  }
  out.ndensity = 0;
  for (size_t p = 0; p < parts.size(); p++) {
    const Model& m = parts[p];
    for (const auto& a : group_map[p]) {
      // the loader accepts only groups in [0, kMaxGroups)
      if (a.first < 0 || a.first >= kMaxGroups) {
        err = "batch: group " + std::to_string(a.first) + " out of range";
        return -1;
      }
      out.density[a.second] = m.density[a.first];
      out.ndensity = std::max(out.ndensity, a.second + 1);
      for (const auto& b : group_map[p])
        for (int id = 0; id < P_COUNT; id++) {
          const size_t k = ((size_t)id * kMaxGroups + a.first) * kMaxGroups + b.first;
          if (m.defined[k]) out.set_param(id, a.second, b.second, m.table[k]);
        }
    }
  }
  out.ngroup = std::max(out.ngroup, ng);

  // bodies: driven of every deck, then free of every deck
  int nd = 0, nf = 0;
  for (const Model& m : parts) {
    nd += m.nDriven;
    nf += (int)m.bodies.size() - m.nDriven;
  }
  out.nDriven = nd;
  out.bodies.resize((size_t)nd + nf);
  ranges.assign(parts.size(), PackingRange());
  std::vector<std::vector<int>> body_map(parts.size());
  int bd = 0, bf = nd;
  for (size_t p = 0; p < parts.size(); p++) {
    const Model& m = parts[p];
    ranges[p].first_driven = bd;
    ranges[p].n_driven = m.nDriven;
    ranges[p].first_free = bf;
    ranges[p].n_free = (int)m.bodies.size() - m.nDriven;
    body_map[p].resize(m.bodies.size());
    for (size_t b = 0; b < m.bodies.size(); b++) {
      const int nb = (int)b < m.nDriven ? bd++ : bf++;
      body_map[p][b] = nb;
      Body B = m.bodies[b];
      B.shape = shape_map[p][B.shape];
      B.group = group_map[p][B.group];
      out.bodies[nb] = B;
    }
    out.totalBrokenArea += m.totalBrokenArea;
  }

  // interactions, interfaces and controls renumbered, then in key order
  for (size_t p = 0; p < parts.size(); p++) {
    for (Inter I : parts[p].inter) {
      I.i = body_map[p][I.i];
      I.j = body_map[p][I.j];
      I.stick = -1;
      out.inter.push_back(I);
    }
    for (Interface BI : parts[p].ifc) {
      BI.i = body_map[p][BI.i];
      BI.j = body_map[p][BI.j];
      out.ifc.push_back(BI);
    }
    for (Control C : parts[p].ctrl) {
      C.i = body_map[p][C.i];
      out.ctrl.push_back(C);
    }
  }
  std::sort(out.inter.begin(), out.inter.end(), key_less);
  std::sort(out.ifc.begin(), out.ifc.end(),
            [](const Interface& a, const Interface& b) { return a.i != b.i ? a.i < b.i : a.j < b.j; });
  for (size_t id = 0; id < out.ifc.size(); id++) {
    const Interface& BI = out.ifc[id];
    for (size_t u = 0; u < BI.bkey.size(); u += 3) {
      Inter key;
      key.i = BI.i;
      key.j = BI.j;
      key.type = BI.bkey[u];
      key.isub = BI.bkey[u + 1];
      key.jsub = BI.bkey[u + 2];
      auto it = std::lower_bound(out.inter.begin(), out.inter.end(), key, key_less);
      if (it != out.inter.end() && !key_less(key, *it)) it->stick = (int)id;
    }
  }
  return 0;
}

}  // namespace rk
