// rk_engine.cpp -- the engine: host orchestration of the device step and the
// exported C ABI (include/rockable_amd.h).
//
// The host keeps the scalar time-loop arithmetic of Rockable::integrate
// (t, interConfC, interVerletC accumulated in double exactly as
// Rockable.cpp:2584-2715 does) and decides conf saves and neighbour-list
// rebuilds from it; all per-body and per-interaction arithmetic runs on the
// device.  There is no CPU fallback: without a usable gfx950 device every
// entry point fails.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rockable_amd.h"
#include "rk_device.hpp"
#include "rk_model.hpp"
#include "rk_nl.hpp"

using namespace rk;

struct KernelTimer {
  int64_t launches = 0;
  double ms = 0.0;
};

struct rk_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Model m;
  bool loaded = false;
  int record_geometry = 1;
  int profile = 0;
  std::map<std::string, KernelTimer> timers;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int law_group = 1, geom_block = 256, gather_waves = 4, body_block = 256, fuse_law = 1, gather_lanes = 16;
  // 0 / -1: chosen from the packing's size (large packings: 64 bodies per
  // resultant wavefront, global vertices written by the predictor per body;
  // smaller ones: by the predictor's workgroups of 64 bodies)
  int gather_groups = 0, fuse_glob = -1;
  int type_sort = 1;
  int persist = 0;
  // host wall time of the rebuilds while the autotuner runs (it synchronises
  // after each rebuild then)
  double* stat_part = nullptr;
  double* stat_out = nullptr;
  size_t stat_cap = 0;
  int64_t loaded_interfaces = 0;
  std::vector<PackingRange> packings;  // one per deck of a batch (one for a single deck)
  bool tuning = false;
  double nl_wall_s = 0.0;
  // profile mode: start/end event pairs recorded on the engine stream and
  // resolved lazily, so that timing adds no host synchronisation
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<std::string, int>> ev_pending;  // (kernel, index of start event)
  size_t ev_used = 0;
  int tic_index = -1;
  // device buffers
  BodyKin* kin = nullptr;
  double *acc = nullptr, *arot = nullptr, *force = nullptr, *moment = nullptr;
  double *mass = nullptr, *inertia = nullptr, *bf_surface = nullptr;
  double *cbound = nullptr, *rmax = nullptr;  // far-entry skip: per-body motion bounds
  int skip_far = 1;
  int skip_seg = 1024;  // entries tested and compacted per workgroup under the skip (256: no compaction)
  unsigned long long* skip_count = nullptr;  // count_skips: the geometry pass's counts
  int32_t* cluster = nullptr;
  int32_t *sh_nv = nullptr, *sh_ne = nullptr, *sh_nf = nullptr, *sh_voff = nullptr, *sh_eoff = nullptr,
          *sh_foff = nullptr, *sh_edge = nullptr, *sh_fstart = nullptr, *sh_fidx = nullptr;
  double *sh_vtx = nullptr, *sh_radius = nullptr, *sh_obb = nullptr, *sh_reach = nullptr;
  double *ptab = nullptr, *damp_rate = nullptr;
  int32_t* damp_kind = nullptr;
  uint64_t* ifc_key = nullptr;
  double* ifc_par = nullptr;
  int32_t *ifc_inner = nullptr, *ifc_alive = nullptr, *ifc_bond_off = nullptr, *ifc_bond_key = nullptr;
  int32_t* ifc_istart = nullptr;  // per body: its first interface as the lower index
  uint8_t* ifc_brk = nullptr;
  int32_t *ifc_broken = nullptr, *ifc_nbroken = nullptr;  // k_breakage's list of broken interfaces
  DevControl* ctrl = nullptr;
  int32_t *ctrl_off = nullptr, *ctrl_idx = nullptr;
  int ctrl_parallel = 0;  // every control acts on a driven body
  int32_t* flags = nullptr;
  uint8_t* passive = nullptr;
  double* gv = nullptr;
  int32_t *bvoff = nullptr, *bvbody = nullptr, *beoff = nullptr, *bfoff = nullptr;
  int netot = 0, nftot = 0;
  int nvtot = 0, nvdriven = 0;
  int32_t* wvstart = nullptr;  // free bodies' vertices by wavefronts of whole bodies
  int nwv = 0;
  // global vertices: of no body, of the listed bodies (those the force phase
  // reads), or of every body (the neighbour list reads them) current
  enum { GV_STALE, GV_LISTED, GV_ALL } gv_state = GV_STALE;
  // device guard: the displacement check (dynamicUpdateNL) and the freeze
  // of the steps after a bond breakage (StickedLinks with interfaces)
  DevGuard* guard = nullptr;
  double *gdpos = nullptr, *gdq = nullptr;
  unsigned long long* gpart = nullptr;  // the gather workgroups' guard maxima
  int gpart_cap = 0;                    // pairs
  double wstar = -2.0;
  int guard_batch = 16;  // steps enqueued between two reads of the guard
  // while bonds live: the batch length, from 2 up to guard_batch, twice the
  // steps a batch ended by a freeze executed, doubled after one that was not
  // (breakages come in bursts, and each frozen step still costs launches)
  int64_t bond_batch = 16;
  // rebuild-interval network (not owned) and its inputs / output on the device
  rk_mlp* model = nullptr;
  float* feat = nullptr;   // 16
  float* pred = nullptr;   // 1
  double* feat_part = nullptr;
  size_t feat_part_cap = 0;
  int64_t pred_batch = 16;
  // proposals are scaled by this: a step enqueued past a rebuild (frozen)
  // costs more than an extra read of the check on large packings
  double model_scale = 0.75;
  // calibration of the network on the running packing: shrinks after a batch
  // that ran past a rebuild, recovers slowly after batches that did not
  double model_trust = 1.0;
  int64_t guard_reads = 0, guard_wasted = 0, guard_polls = 0;
  // asynchronous polls of the device check: pinned flags and their events
  int* poll_host = nullptr;
  // host-mapped progress and freeze words of the check (the device writes
  // them, the host reads them without a copy or an event); note_dev is the
  // device address
  HostNote* note = nullptr;
  HostNote* note_dev = nullptr;
  // pinned staging of the guard read at the end of each batch (the guard
  // record and the breakage flag)
  struct GuardRead {
    DevGuard g;
    int32_t broke;
  };
  GuardRead* guard_host = nullptr;
  hipEvent_t poll_ev[8] = {};
  int64_t last_rebuild_step = 0, last_interval = 0;
  int exact_sums = 0;
  // steps enqueued but not yet launched (deferred so that runs of graph_steps
  // steps go out as one captured graph); with_guard: they end with the
  // displacement check
  int graph_steps = 0;
  int64_t pending = 0;
  bool pending_guard = false;
  struct StepGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    int64_t used = 0;
    int n = 0;
    bool with_guard = false;
  };
  std::vector<StepGraph> graphs;
  int64_t graph_captures = 0, graph_launches = 0, graph_clock = 0;
  EntryArrays E[2];
  int cur = 0;
  NLWork w;
  NLParams nlp;
  Dev d;
  bool needUpdate = false;
  int64_t steps = 0, nl_updates = 0;
  int live_interfaces = 0;
  std::vector<double> ctrl_sig;  // the driving file's content last uploaded
  // integrate()'s output files in its output directory (perf.txt,
  // kineticEnergy.txt, staticBalance.txt) and the data extractors
  FILE *perf_file = nullptr, *ke_file = nullptr, *sb_file = nullptr;
  struct Extractor {
    int kind;  // 1 dnStat, 2 TrackDamage
    int nrec;
    FILE* f;
  };
  std::vector<Extractor> extractors;
  double inAreaInterfaces = 0.0;
  int64_t ext_step = 0;  // integrate()'s step counter (Rockable.cpp:2578, 2742)
  std::chrono::steady_clock::time_point conf_clock;
  double nl_since_conf_s = 0.0, force_since_conf_s = 0.0;
};

namespace {

constexpr int kPollSlots = 8;  // ring of pinned flags and events
// bodies from which the per-body kernels favour whole wavefronts of bodies
// over more wavefronts in flight (about 4 wavefronts per SIMD of 64 bodies)
constexpr int kLargePacking = 1 << 18;
constexpr int kPollAhead = 2;  // polls in flight before the host waits

int fail(rk_engine* e, const std::string& msg) {
  e->err = msg;
  return -1;
}

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) return fail(eng, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int upload(rk_engine* eng, T*& dst, const std::vector<T>& src) {
  if (dst) (void)hipFree(dst);
  dst = nullptr;
  size_t n = src.size() > 0 ? src.size() : 1;
  HIPCHK(hipMalloc(&dst, n * sizeof(T)));
  if (!src.empty()) HIPCHK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, eng->stream));
  return 0;
}

void free_graphs(rk_engine* eng) {
  for (auto& g : eng->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  eng->graphs.clear();
}

void free_device(rk_engine* eng) {
  free_graphs(eng);
  void* ps[] = {eng->kin,      eng->acc,       eng->arot,      eng->force,     eng->moment,   eng->mass,
                eng->inertia,  eng->bf_surface, eng->cluster,  eng->sh_nv,     eng->sh_ne,    eng->sh_nf,
                eng->sh_voff,  eng->sh_eoff,   eng->sh_foff,   eng->sh_edge,   eng->sh_fstart, eng->sh_fidx,
                eng->sh_vtx,   eng->sh_radius, eng->sh_obb,    eng->sh_reach,  eng->ptab,      eng->damp_rate, eng->damp_kind,
                eng->ifc_key,  eng->ifc_istart, eng->ifc_par,  eng->ifc_inner, eng->ifc_alive, eng->ifc_bond_off, eng->ifc_bond_key,
                eng->ifc_brk,  eng->ifc_broken, eng->ifc_nbroken, eng->ctrl,      eng->ctrl_off,  eng->ctrl_idx,  eng->flags,    eng->passive,
                eng->gv,       eng->bvoff,     eng->bvbody,    eng->guard,     eng->gdpos,    eng->gdq,      eng->gpart,
                eng->beoff,    eng->bfoff,     eng->cbound,    eng->rmax,
                eng->feat,     eng->pred,      eng->feat_part, eng->stat_part, eng->stat_out};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  eng->kin = nullptr;
  eng->acc = eng->arot = eng->force = eng->moment = eng->mass = eng->inertia = eng->bf_surface = nullptr;
  eng->cbound = eng->rmax = nullptr;
  eng->cluster = nullptr;
  eng->sh_nv = eng->sh_ne = eng->sh_nf = eng->sh_voff = eng->sh_eoff = eng->sh_foff = eng->sh_edge = nullptr;
  eng->sh_fstart = eng->sh_fidx = nullptr;
  eng->sh_vtx = eng->sh_radius = eng->sh_obb = eng->sh_reach = nullptr;
  eng->ptab = eng->damp_rate = nullptr;
  eng->damp_kind = nullptr;
  eng->ifc_key = nullptr;
  eng->ifc_istart = nullptr;
  eng->ifc_par = nullptr;
  eng->ifc_inner = eng->ifc_alive = eng->ifc_bond_off = eng->ifc_bond_key = nullptr;
  eng->ifc_brk = nullptr;
  eng->ifc_broken = eng->ifc_nbroken = nullptr;
  eng->ctrl = nullptr;
  eng->ctrl_off = eng->ctrl_idx = nullptr;
  eng->flags = nullptr;
  eng->passive = nullptr;
  eng->gv = nullptr;
  eng->bvoff = eng->bvbody = eng->beoff = eng->bfoff = nullptr;
  eng->netot = eng->nftot = 0;
  eng->guard = nullptr;
  eng->gdpos = eng->gdq = nullptr;
  eng->gpart = nullptr;
  eng->gpart_cap = 0;
  eng->feat = eng->pred = nullptr;
  eng->feat_part = nullptr;
  eng->feat_part_cap = 0;
  eng->stat_part = eng->stat_out = nullptr;
  eng->stat_cap = 0;
  eng->nvtot = eng->nvdriven = 0;
  if (eng->wvstart) (void)hipFree(eng->wvstart);
  eng->wvstart = nullptr;
  eng->nwv = 0;
  for (auto& E : eng->E) {
    entries_free(E);
    if (E.istart) (void)hipFree(E.istart);
    if (E.jstart) (void)hipFree(E.jstart);
    if (E.jlist) (void)hipFree(E.jlist);
    if (E.soff) (void)hipFree(E.soff);
    if (E.heavy) (void)hipFree(E.heavy);
    if (E.hch) (void)hipFree(E.hch);
    if (E.hch_off) (void)hipFree(E.hch_off);
    if (E.hpart) (void)hipFree(E.hpart);
    if (E.edesc) (void)hipFree(E.edesc);
    E = EntryArrays();
  }
  nl_work_free(eng->w);
}

// 1 for a driven body whose force and moment feed nothing back into the
// motion: no force or moment control acts on it (a fixed or velocity-driven
// wall).  Its resultant is only reported.
std::vector<uint8_t> passive_flags(const Model& m) {
  const int nb = (int)m.bodies.size();
  std::vector<uint8_t> p(nb, 0);
  for (int b = 0; b < std::min(m.nDriven, nb); b++) p[b] = 1;
  for (const Control& c : m.ctrl) {
    const bool fm = c.type == C_X_FOR || c.type == C_Y_FOR || c.type == C_Z_FOR || c.type == C_XROT_MOM ||
                    c.type == C_YROT_MOM || c.type == C_ZROT_MOM || c.type == C_XYZROT_MOM;
    if (fm && c.i >= 0 && c.i < nb) p[c.i] = 0;
  }
  return p;
}

// The controls of each driven body, in list order (off: nDriven + 1 starts
// into idx).  Returns 1 when every control acts on a driven body (each body's
// controls can then run on a thread of their own), else 0.
int control_csr(const Model& m, std::vector<int32_t>& off, std::vector<int32_t>& idx) {
  const int nd = std::max(0, std::min(m.nDriven, (int)m.bodies.size()));
  off.assign(nd + 1, 0);
  for (const Control& k : m.ctrl)
    if (k.i >= 0 && k.i < nd) off[k.i + 1]++;
  for (int b = 0; b < nd; b++) off[b + 1] += off[b];
  idx.assign(off[nd], 0);
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  for (size_t q = 0; q < m.ctrl.size(); q++) {
    const int b = m.ctrl[q].i;
    if (b >= 0 && b < nd) idx[fill[b]++] = (int)q;
  }
  return off[nd] == (int)m.ctrl.size() ? 1 : 0;
}

// Initial far-entry bounds: 0, or +inf for a driven body that a control
// moves or turns (its entries are then never skipped).
std::vector<double> initial_bounds(const Model& m) {
  std::vector<double> c(std::max<size_t>(1, m.bodies.size()), 0.0);
  for (const Control& k : m.ctrl)
    if (k.i >= 0 && k.i < (int)m.bodies.size()) c[k.i] = INFINITY;
  return c;
}

// Room in gpart for the gather's workgroups under any launch shape (their
// count grows with the heavy bodies' chunks, so after each list change).
int ensure_gpart(rk_engine* eng) {
  if (!eng->guard || eng->m.dynamicUpdateNL == 0) return 0;
  const int need = gather_blocks_max((int)eng->m.bodies.size(), eng->E[eng->cur].nhch);
  if (need <= eng->gpart_cap) return 0;
  if (eng->gpart) (void)hipFree(eng->gpart);
  eng->gpart = nullptr;
  eng->gpart_cap = 0;
  HIPCHK(hipMalloc(&eng->gpart, 2 * sizeof(unsigned long long) * (size_t)need));
  eng->gpart_cap = need;
  return 0;
}

void assemble_dev(rk_engine* eng) {
  Dev& d = eng->d;
  memset(&d, 0, sizeof d);  // the captured-step cache compares Dev bytewise
  const Model& m = eng->m;
  EntryArrays& E = eng->E[eng->cur];
  d.nb = (int)m.bodies.size();
  d.nDriven = m.nDriven;
  d.kin = eng->kin;
  d.gv = eng->gv;
  d.bvoff = eng->bvoff;
  d.beoff = eng->beoff;
  d.bfoff = eng->bfoff;
  d.netot = eng->netot;
  d.nftot = eng->nftot;
  d.bvbody = eng->bvbody;
  d.nvtot = eng->nvtot;
  d.nvdriven = eng->nvdriven;
  d.wvstart = eng->wvstart;
  d.nwv = eng->nwv;
  d.acc = eng->acc;
  d.arot = eng->arot;
  d.force = eng->force;
  d.moment = eng->moment;
  d.mass = eng->mass;
  d.inertia = eng->inertia;
  d.cluster = eng->cluster;
  d.sh.nv = eng->sh_nv;
  d.sh.ne = eng->sh_ne;
  d.sh.nf = eng->sh_nf;
  d.sh.voff = eng->sh_voff;
  d.sh.eoff = eng->sh_eoff;
  d.sh.foff = eng->sh_foff;
  d.sh.vtx = eng->sh_vtx;
  d.sh.edge = eng->sh_edge;
  d.sh.fstart = eng->sh_fstart;
  d.sh.fidx = eng->sh_fidx;
  d.sh.radius = eng->sh_radius;
  d.sh.obb = eng->sh_obb;
  d.sh.reach = eng->sh_reach;
  d.par.tab = eng->ptab;
  d.par.damp_rate = eng->damp_rate;
  d.par.damp_kind = eng->damp_kind;
  d.par.G = std::max(1, m.ngroup);
  d.ne = E.n;
  d.ei = E.ei;
  d.ej = E.ej;
  d.esub = E.esub;
  d.edesc = E.edesc;
  d.est = E.est;
  d.edn = E.edn;
  d.egeo = E.egeo;
  d.edamp = E.edamp;
  d.estick = E.estick;
  d.eact = E.eact;
  d.ekey = E.ekey;
  d.cbound = eng->cbound;
  d.rmax = eng->rmax;
  d.skip_far = eng->skip_far;
  d.skip_count = eng->skip_count;
  d.skip_seg = eng->skip_seg;
  d.spos = E.spos;
  d.soff = E.soff;
  d.sact = E.sact;
  d.sval = E.sval;
  d.heavy = E.heavy;
  d.hch = E.hch;
  d.hch_off = E.hch_off;
  d.hpart = E.hpart;
  d.nhch = E.nhch;
  d.nheavy_exact = eng->exact_sums ? E.nheavy : E.nheavy_np;
  d.nheavy = E.nheavy;
  d.heavy_min = eng->w.heavy_min;
  d.passive = eng->passive;
  d.akin = E.akin;
  d.guard_on = eng->guard != nullptr;
  d.guard = eng->guard;
  d.note = eng->note_dev;
  d.gdpos = eng->gdpos;
  d.gdq = eng->gdq;
  d.gpart = eng->gpart;
  d.disp = eng->m.dispUpdateNL;
  d.wstar = eng->wstar;
  d.disp_check = eng->m.dynamicUpdateNL != 0 ? 1 : 0;
  d.acount = E.acount;
  d.exact_sums = eng->exact_sums;
  d.law_group = eng->law_group;
  d.geom_block = eng->geom_block;
  d.gather_waves = eng->gather_waves;
  d.body_block = eng->body_block;
  d.fuse_law = eng->fuse_law;
  d.gather_lanes = eng->gather_lanes;
  d.type_sort = eng->type_sort;
  {
    const bool large = d.nb >= kLargePacking;
    const int g = eng->gather_groups > 0 ? eng->gather_groups : large ? eng->gather_lanes : 1;
    d.gather_groups = std::min(g, eng->gather_lanes);
    d.fuse_glob = eng->fuse_glob >= 0 ? eng->fuse_glob : large ? 0 : 3;
  }
  d.persist = eng->persist;
  d.istart = E.istart;
  d.jstart = E.jstart;
  d.jlist = E.jlist;
  d.nctrl = (int)m.ctrl.size();
  d.ctrl = eng->ctrl;
  d.ctrl_off = eng->ctrl_off;
  d.ctrl_idx = eng->ctrl_idx;
  d.ctrl_parallel = eng->ctrl_parallel;
  d.servo = m.servo;
  d.servo_icontr = m.servo_icontr;
  for (int k = 0; k < 6; k++) d.servo_ids[k] = m.servo_ids[k];
  d.servo_pressure = m.servo_pressure;
  for (int k = 0; k < 6; k++) d.servo_val[k] = m.servo_val[k];
  d.lode_a = m.lode_a;
  d.lode_b = m.lode_b;
  d.lode_rate = m.lode_rate;
  d.bf = m.bf;
  d.bf_point[0] = m.bf_point.x;
  d.bf_point[1] = m.bf_point.y;
  d.bf_point[2] = m.bf_point.z;
  d.bf_acc = m.bf_acc;
  d.bf_axisBody[0] = m.bf_axisBody.x;
  d.bf_axisBody[1] = m.bf_axisBody.y;
  d.bf_axisBody[2] = m.bf_axisBody.z;
  d.bf_axis[0] = m.bf_axis.x;
  d.bf_axis[1] = m.bf_axis.y;
  d.bf_axis[2] = m.bf_axis.z;
  d.bf_Kr = m.bf_Kr;
  d.bf_fluidDensity = m.bf_fluidDensity;
  d.bf_surface = eng->bf_surface;
  d.ifc.n = (int)m.ifc.size();
  d.ifc.key = eng->ifc_key;
  d.ifc.istart = eng->ifc_istart;
  d.ifc.par = eng->ifc_par;
  d.ifc.isInner = eng->ifc_inner;
  d.ifc.alive = eng->ifc_alive;
  d.ifc.brk = eng->ifc_brk;
  d.ifc.broken = eng->ifc_broken;
  d.ifc.nbroken = eng->ifc_nbroken;
  d.ifc.bond_off = eng->ifc_bond_off;
  d.ifc.bond_key = eng->ifc_bond_key;
  d.paramsInInterfaces = m.paramsInInterfaces;
  d.flags = eng->flags;
}

StepParams step_params(rk_engine* eng) {
  StepParams sp;
  memset(&sp, 0, sizeof sp);  // the captured-step cache compares it bytewise
  sp.dt = eng->m.dt;
  sp.dt_2 = eng->m.dt_2;
  sp.dt2_2 = eng->m.dt2_2;
  sp.gx = eng->m.gravity.x;
  sp.gy = eng->m.gravity.y;
  sp.gz = eng->m.gravity.z;
  sp.law = eng->m.law;
  if (eng->m.servo_timed()) {  // such steps are launched one by one, at their time
    sp.t = eng->m.t;
    eng->m.servo_time_values(eng->m.t, sp.servo_v);
  }
  return sp;
}

// The largest w in [-1, 1] with 2 acos(w) >= angle (glibc acos, as the
// oracle's get_angle); 2 acos is non-increasing in w, so the device compares
// the minimum w of the accumulated rotations with it.  -2 when no w qualifies.
double wstar_for(double angle) {
  auto ok = [&](double w) { return 2.0 * acos(w) >= angle; };
  if (ok(1.0)) return 1.0;
  if (!ok(-1.0)) return -2.0;
  auto key = [](double w) {
    uint64_t u;
    memcpy(&u, &w, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
  };
  auto unkey = [](uint64_t k) {
    uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double w;
    memcpy(&w, &u, 8);
    return w;
  };
  uint64_t lo = key(-1.0), hi = key(1.0);  // ok(lo), !ok(hi)
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (ok(unkey(mid)))
      lo = mid;
    else
      hi = mid;
  }
  return unkey(lo);
}

// How far a shape's near tests (Particle.cpp:88-209) reach beyond the convex
// hull of its vertices, at homothety 1: a vertex-edge test accepts the
// vertex's foot on the edge's line up to dVerlet edge lengths past either
// end (so the longest edge), and a vertex-face test accepts a foot inside the
// face's outline projected on the plane of its first two and last vertices
// (so the largest distance of a face vertex from that plane).  A degenerate
// face (no plane) gives +inf, which turns the narrow phase's separation test
// off for the shape.
using v2d = std::pair<double, double>;
v2d shape_reach(const Shape& s) {
  double lmax = 0.0, off = 0.0;
  for (int e = 0; e < s.ne(); e++) {
    const v3 E = vsub(s.vertex[s.edge[2 * e + 1]], s.vertex[s.edge[2 * e]]);
    lmax = std::max(lmax, std::sqrt(vdot(E, E)));
  }
  for (int f = 0; f < s.nf(); f++) {
    const int f0 = s.face_start[f], f1 = s.face_start[f + 1];
    if (f1 - f0 < 3) {
      off = INFINITY;
      continue;
    }
    const v3 A = s.vertex[s.face_idx[f0]];
    v3 v1 = vsub(s.vertex[s.face_idx[f0 + 1]], A);
    v1 = vsc(1.0 / std::sqrt(vdot(v1, v1)), v1);
    v3 n = vcross(v1, vsub(s.vertex[s.face_idx[f1 - 1]], A));
    n = vsc(1.0 / std::sqrt(vdot(n, n)), n);
    for (int k = f0; k < f1; k++) {
      const double h = std::fabs(vdot(vsub(s.vertex[s.face_idx[k]], A), n));
      off = (h == h) ? std::max(off, h) : INFINITY;
    }
  }
  if (!(lmax == lmax)) lmax = INFINITY;
  return {lmax, off};
}

// upload everything the model holds (after a conf load)
int upload_model(rk_engine* eng) {
  Model& m = eng->m;
  const int nb = (int)m.bodies.size();
  // shapes
  std::vector<int32_t> nv, ne, nf, voff, eoff, foff, edge, fstart, fidx;
  std::vector<double> vtx, radius, obb, reach;
  int maxv = 1;
  int maxne = 0, maxnf = 0;
  for (const Shape& s : m.shapes) {
    voff.push_back((int)vtx.size() / 3);
    eoff.push_back((int)edge.size() / 2);
    foff.push_back((int)fstart.size());
    nv.push_back(s.nv());
    ne.push_back(s.ne());
    nf.push_back(s.nf());
    maxv = std::max(maxv, s.nv());
    maxne = std::max(maxne, s.ne());
    maxnf = std::max(maxnf, s.nf());
    for (const v3& v : s.vertex) {
      vtx.push_back(v.x);
      vtx.push_back(v.y);
      vtx.push_back(v.z);
    }
    for (int x : s.edge) edge.push_back(x);
    const int base = (int)fidx.size();
    for (int f = 0; f < s.nf(); f++) fstart.push_back(base + s.face_start[f]);
    for (int x : s.face_idx) fidx.push_back(x);
    radius.push_back(s.radius);
    const obb_t& o = s.obb;
    const double ob[15] = {o.c.x,    o.c.y,    o.c.z,    o.e[0].x, o.e[0].y, o.e[0].z, o.e[1].x, o.e[1].y,
                           o.e[1].z, o.e[2].x, o.e[2].y, o.e[2].z, o.ext.x,  o.ext.y,  o.ext.z};
    obb.insert(obb.end(), ob, ob + 15);
    const v2d r = shape_reach(s);
    reach.push_back(r.first);
    reach.push_back(r.second);
  }
  fstart.push_back((int)fidx.size());
  if (maxv > 2048 || maxne > 8192 || maxnf > 8192) return fail(eng, "shape too large for the narrow phase");
  const long maxT = (long)maxv * (maxv + maxne + maxnf) + (long)maxne * maxne + (long)maxv * (maxne + maxnf);
  eng->nlp.maxv = maxv;
  eng->nlp.maxe = maxne;
  eng->nlp.maxf = maxnf;
  eng->nlp.maxwords = (int)((maxT + 63) / 64);
  eng->nlp.has_points = 0;
  for (const Shape& sh : m.shapes)
    if (sh.nv() == 1 && sh.edge.empty() && sh.nf() == 0) eng->nlp.has_points = 1;
  const size_t lds = narrow_lds_bytes(eng->nlp);
  if (lds > 65536) return fail(eng, "shapes have too many sub-elements for the narrow-phase LDS budget");
  if (upload(eng, eng->sh_nv, nv) || upload(eng, eng->sh_ne, ne) || upload(eng, eng->sh_nf, nf) ||
      upload(eng, eng->sh_voff, voff) || upload(eng, eng->sh_eoff, eoff) || upload(eng, eng->sh_foff, foff) ||
      upload(eng, eng->sh_edge, edge) || upload(eng, eng->sh_fstart, fstart) || upload(eng, eng->sh_fidx, fidx) ||
      upload(eng, eng->sh_vtx, vtx) || upload(eng, eng->sh_radius, radius) || upload(eng, eng->sh_obb, obb) ||
      upload(eng, eng->sh_reach, reach))
    return -1;
  // bodies
  std::vector<BodyKin> kin(nb);
  std::vector<double> acc(3 * nb), arot(3 * nb), zero3(3 * (size_t)nb, 0.0), mass(nb), inertia(3 * nb), surf(nb);
  std::vector<int32_t> cl(nb);
  for (int b = 0; b < nb; b++) {
    const Body& p = m.bodies[b];
    BodyKin& k = kin[b];
    k.px = p.pos.x;
    k.py = p.pos.y;
    k.pz = p.pos.z;
    k.qw = p.Q.w;
    k.qx = p.Q.x;
    k.qy = p.Q.y;
    k.qz = p.Q.z;
    k.vx = p.vel.x;
    k.vy = p.vel.y;
    k.vz = p.vel.z;
    k.wx = p.vrot.x;
    k.wy = p.vrot.y;
    k.wz = p.vrot.z;
    k.R = p.homothety * m.shapes[p.shape].radius;
    k.h = p.homothety;
    k.shape = p.shape;
    k.group = p.group;
    acc[3 * b] = p.acc.x;
    acc[3 * b + 1] = p.acc.y;
    acc[3 * b + 2] = p.acc.z;
    arot[3 * b] = p.arot.x;
    arot[3 * b + 1] = p.arot.y;
    arot[3 * b + 2] = p.arot.z;
    mass[b] = p.mass;
    inertia[3 * b] = p.inertia.x;
    inertia[3 * b + 1] = p.inertia.y;
    inertia[3 * b + 2] = p.inertia.z;
    cl[b] = p.cluster;
    const double h = p.homothety;
    surf[b] = pow(m.shapes[p.shape].volume * h * h * h, 2.0 / 3.0);  // BodyForce_ViscousFluid.cpp:58-59
  }
  if (upload(eng, eng->kin, kin) || upload(eng, eng->acc, acc) || upload(eng, eng->arot, arot) ||
      upload(eng, eng->force, zero3) || upload(eng, eng->moment, zero3) || upload(eng, eng->mass, mass) ||
      upload(eng, eng->inertia, inertia) || upload(eng, eng->cluster, cl) || upload(eng, eng->bf_surface, surf))
    return -1;
  {
    // far-entry skip: each body's largest vertex distance from its centre
    std::vector<double> rmax(std::max(1, nb), 0.0);
    for (int b = 0; b < nb; b++) {
      double r2 = 0.0;
      for (const v3& v : m.shapes[m.bodies[b].shape].vertex) r2 = std::max(r2, v.x * v.x + v.y * v.y + v.z * v.z);
      rmax[b] = m.bodies[b].homothety * sqrt(r2);
    }
    if (upload(eng, eng->rmax, rmax) || upload(eng, eng->cbound, initial_bounds(m))) return -1;
  }
  // global-vertex table: body b's vertices from bvoff[b]
  {
    std::vector<int32_t> bvoff(nb + 1, 0), bvbody;
    for (int b = 0; b < nb; b++) {
      const int nvb = m.shapes[m.bodies[b].shape].nv();
      bvoff[b + 1] = bvoff[b] + nvb;
      for (int v = 0; v < nvb; v++) bvbody.push_back(b);
    }
    eng->nvtot = bvoff[nb];
    eng->nvdriven = bvoff[m.nDriven];
    // the free bodies' vertices packed into wavefronts of at most 64, whole
    // bodies each (none if a free body has more than 64 vertices)
    std::vector<int32_t> wv(1, eng->nvdriven);
    bool fits = true;
    for (int b = m.nDriven, used = 0; b < nb; b++) {
      const int nvb = bvoff[b + 1] - bvoff[b];
      if (nvb > 64) fits = false;
      if (used + nvb > 64) {
        wv.push_back(bvoff[b]);
        used = 0;
      }
      used += nvb;
    }
    if (wv.back() != eng->nvtot) wv.push_back(eng->nvtot);
    eng->nwv = fits && eng->nvtot > eng->nvdriven ? (int)wv.size() - 1 : 0;
    if (upload(eng, eng->wvstart, wv)) return -1;
    if (upload(eng, eng->bvoff, bvoff) || upload(eng, eng->bvbody, bvbody) ||
        upload(eng, eng->gv, std::vector<double>(3 * (size_t)std::max(1, eng->nvtot), 0.0)))
      return -1;
    eng->gv_state = rk_engine::GV_STALE;
    // each body's edges and faces (for the narrow phase's cull boxes)
    std::vector<int32_t> beoff(nb + 1, 0), bfoff(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
      const Shape& sh = m.shapes[m.bodies[b].shape];
      beoff[b + 1] = beoff[b] + sh.ne();
      bfoff[b + 1] = bfoff[b] + sh.nf();
    }
    eng->netot = beoff[nb];
    eng->nftot = bfoff[nb];
    if (upload(eng, eng->beoff, beoff) || upload(eng, eng->bfoff, bfoff)) return -1;
  }
  // device guard: with dynamicUpdateNL, or with bonds that may break (the
  // steps after a breakage wait for the rebuild on the device, so bonded
  // runs need no host read per step)
  if (m.dynamicUpdateNL != 0 || (m.law == LAW_STICKED && !m.ifc.empty())) {
    std::vector<double> q(4 * (size_t)std::max(1, nb), 0.0);
    for (int b = 0; b < nb; b++) q[4 * (size_t)b] = 1.0;
    if (upload(eng, eng->gdpos, std::vector<double>(3 * (size_t)std::max(1, nb), 0.0)) || upload(eng, eng->gdq, q))
      return -1;
    DevGuard g0;
    memset(&g0, 0, sizeof g0);
    g0.minw = ~0ull;
    if (upload(eng, eng->guard, std::vector<DevGuard>(1, g0))) return -1;
    if (eng->note) *(volatile int32_t*)&eng->note->frozen = 0;
    eng->wstar = wstar_for(m.angleUpdateNL);
  } else {
    void* ps[] = {eng->guard, eng->gdpos, eng->gdq, eng->gpart};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    eng->guard = nullptr;
    eng->gdpos = eng->gdq = nullptr;
    eng->gpart = nullptr;
    eng->gpart_cap = 0;
  }
  // parameter tables (compact G = ngroup)
  const int G = std::max(1, m.ngroup);
  std::vector<double> tab((size_t)P_COUNT * G * G), rate((size_t)G * G);
  std::vector<int32_t> kind((size_t)G * G);
  for (int id = 0; id < P_COUNT; id++)
    for (int a = 0; a < G; a++)
      for (int b = 0; b < G; b++) tab[((size_t)id * G + a) * G + b] = m.param(id, a, b);
  for (int a = 0; a < G; a++)
    for (int b = 0; b < G; b++) {
      const double en2 = m.param(P_EN2, a, b);
      if (en2 > 0.0 && en2 < 1.0) {
        const double logen = 0.5 * log(en2);
        rate[a * G + b] = -logen / sqrt(logen * logen + M_PI * M_PI);
        kind[a * G + b] = 0;
      } else if (en2 <= 0.0) {
        kind[a * G + b] = 1;
      } else {
        kind[a * G + b] = 2;
      }
    }
  if (upload(eng, eng->ptab, tab) || upload(eng, eng->damp_rate, rate) || upload(eng, eng->damp_kind, kind)) return -1;
  // interfaces
  {
    std::vector<uint64_t> key;
    std::vector<double> par;
    std::vector<int32_t> inner, alive, boff(1, 0), bkey;
    for (const Interface& BI : m.ifc) {
      key.push_back(((uint64_t)(uint32_t)BI.i << 32) | (uint32_t)BI.j);
      const double p8[8] = {BI.kn, BI.kt, BI.kr, BI.fn0, BI.ft0, BI.mom0, BI.power, BI.dn0};
      par.insert(par.end(), p8, p8 + 8);
      inner.push_back(BI.isInner);
      alive.push_back(BI.alive);
      for (int x : BI.bkey) bkey.push_back(x);
      boff.push_back((int)bkey.size() / 3);
    }
    // keys sorted by (i, j): a counting pass gives each body's range
    std::vector<int32_t> istart((size_t)nb + 1, 0);
    for (const Interface& BI : m.ifc) istart[(size_t)BI.i + 1]++;
    for (int b = 0; b < nb; b++) istart[(size_t)b + 1] += istart[b];
    std::vector<uint8_t> brk(m.ifc.size(), 0);
    if (upload(eng, eng->ifc_key, key) || upload(eng, eng->ifc_istart, istart) || upload(eng, eng->ifc_par, par) || upload(eng, eng->ifc_inner, inner) ||
        upload(eng, eng->ifc_alive, alive) || upload(eng, eng->ifc_bond_off, boff) ||
        upload(eng, eng->ifc_bond_key, bkey) || upload(eng, eng->ifc_brk, brk) ||
        upload(eng, eng->ifc_broken, std::vector<int32_t>(std::max<size_t>(1, m.ifc.size()), 0)) ||
        upload(eng, eng->ifc_nbroken, std::vector<int32_t>(1, 0)))
      return -1;
    eng->live_interfaces = 0;
    for (const Interface& BI : m.ifc) eng->live_interfaces += BI.alive;
    eng->loaded_interfaces = eng->live_interfaces;
  }
  // controls: list + per driven body CSR preserving list order
  {
    std::vector<DevControl> c;
    for (const Control& x : m.ctrl) {
      DevControl dc;
      dc.type = x.type;
      dc.i = x.i;
      dc.value = x.value;
      dc.vx = x.vec.x;
      dc.vy = x.vec.y;
      dc.vz = x.vec.z;
      c.push_back(dc);
    }
    std::vector<int32_t> off, idx;
    eng->ctrl_parallel = control_csr(m, off, idx);
    if (upload(eng, eng->ctrl, c) || upload(eng, eng->ctrl_off, off) || upload(eng, eng->ctrl_idx, idx)) return -1;
    if (upload(eng, eng->passive, passive_flags(m))) return -1;
  }
  std::vector<int32_t> fl(4, 0);
  if (upload(eng, eng->flags, fl)) return -1;
  // loaded interactions: already lexicographic
  EntryArrays& E = eng->E[eng->cur];
  const int n = (int)m.inter.size();
  if (entries_reserve(eng->w, E, n) != 0) return fail(eng, eng->w.err);
  E.n = n;
  if (n > 0) {
    std::vector<int32_t> ei(n), ej(n), st(n);
    std::vector<uint32_t> sub(n);
    std::vector<EState> es(n);
    std::vector<double> dn(n);
    std::vector<uint8_t> act(n);
    std::vector<EGeom> eg(n);
    std::vector<double> dp(n);
    for (int k = 0; k < n; k++) {
      const Inter& I = m.inter[k];
      ei[k] = I.i;
      ej[k] = I.j;
      sub[k] = pack_sub(I.type, I.isub, I.jsub);
      st[k] = I.stick;
      EState s;
      dn[k] = I.dn;
      const bool hist = !(I.fn == 0.0 && I.ft.x == 0.0 && I.ft.y == 0.0 && I.ft.z == 0.0 && I.mom.x == 0.0 &&
                          I.mom.y == 0.0 && I.mom.z == 0.0);
      act[k] = ((I.dn < 0.0 || I.stick >= 0) ? 1 : 0) | (hist ? 2 : 0);
      s.fn = I.fn;
      s.ftx = I.ft.x;
      s.fty = I.ft.y;
      s.ftz = I.ft.z;
      s.momx = I.mom.x;
      s.momy = I.mom.y;
      s.momz = I.mom.z;
      es[k] = s;
      EGeom g;
      g.nx = I.n.x;
      g.ny = I.n.y;
      g.nz = I.n.z;
      g.px = I.pos.x;
      g.py = I.pos.y;
      g.pz = I.pos.z;
      g.vx = I.vel.x;
      g.vy = I.vel.y;
      g.vz = I.vel.z;
      eg[k] = g;
      dp[k] = I.damp;
    }
    HIPCHK(hipMemcpyAsync(E.ei, ei.data(), 4 * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.ej, ej.data(), 4 * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.esub, sub.data(), 4 * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.estick, st.data(), 4 * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.est, es.data(), sizeof(EState) * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.edn, dn.data(), sizeof(double) * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.eact, act.data(), n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.egeo, eg.data(), sizeof(EGeom) * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(E.edamp, dp.data(), sizeof(double) * n, hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
  }
  if (nl_build_csr(eng->w, nb, E, eng->passive, eng->stream) != 0) return fail(eng, eng->w.err);
  eng->nlp.DVerlet = m.DVerlet;
  eng->nlp.dVerlet = m.dVerlet;
  eng->nlp.linkcells = m.nl_linkcells;
  eng->nlp.cellMinSizes[0] = m.cellMinSizes.x;
  eng->nlp.cellMinSizes[1] = m.cellMinSizes.y;
  eng->nlp.cellMinSizes[2] = m.cellMinSizes.z;
  eng->nlp.boxForLinkCellsOpt = m.boxForLinkCellsOpt;
  if (ensure_gpart(eng) != 0) return -1;
  assemble_dev(eng);
  launch_entry_desc(eng->d, eng->stream);
  launch_reset_keys(eng->d, eng->stream);
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

hipEvent_t pool_event(rk_engine* eng) {
  if (eng->ev_used == eng->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    eng->ev_pool.push_back(e);
  }
  return eng->ev_pool[eng->ev_used++];
}

void resolve_events(rk_engine* eng) {
  if (eng->ev_pending.empty()) return;
  (void)hipStreamSynchronize(eng->stream);
  for (const auto& pr : eng->ev_pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, eng->ev_pool[pr.second], eng->ev_pool[pr.second + 1]);
    KernelTimer& t = eng->timers[pr.first];
    t.launches++;
    t.ms += ms;
  }
  eng->ev_pending.clear();
  eng->ev_used = 0;
}

// Timing brackets: level 1 brackets the force phase (k_geom, k_law, gather
// kernels, name "force") and the neighbour-list rebuild ("nl"); level 2
// brackets every kernel on its own.
void tic(rk_engine* eng, int level) {
  if (eng->profile == 0 || eng->profile != level) return;
  if (eng->ev_used + 2 > 4096) resolve_events(eng);
  hipEvent_t a = pool_event(eng);
  hipEvent_t b = pool_event(eng);
  if (!a || !b) return;
  eng->tic_index = (int)eng->ev_used - 2;
  (void)hipEventRecord(a, eng->stream);
}
void toc(rk_engine* eng, const char* name, int level) {
  if (eng->profile != level || eng->tic_index < 0) return;
  (void)hipEventRecord(eng->ev_pool[eng->tic_index + 1], eng->stream);
  eng->ev_pending.push_back({name, eng->tic_index});
  eng->tic_index = -1;
}

// Global vertices of the current poses (k_glob), before a kernel reads them:
// of every body (all) or of the bodies with a listed interaction.
void ensure_gv(rk_engine* eng, bool all) {
  if (eng->gv_state == rk_engine::GV_ALL || (!all && eng->gv_state == rk_engine::GV_LISTED)) return;
  launch_glob(eng->d, all ? 1 : 0, eng->stream);
  eng->gv_state = all ? rk_engine::GV_ALL : rk_engine::GV_LISTED;
}

// Ordered accounting of the interfaces broken by the last executed step
// (breakage_of_interfaces, Rockable.cpp:3391-3410): the broken areas add to
// totalBrokenArea in interface order, the host's alive flags and count follow
// the device's, the breakage marks and flag are cleared, a rebuild is due.
// The device lists the interfaces it broke (k_breakage, in no particular
// order): the host reads that short list, not every alive flag, and sorts it.
int account_breakage(rk_engine* eng) {
  int32_t n = 0;
  HIPCHK(hipMemcpy(&n, eng->ifc_nbroken, sizeof n, hipMemcpyDeviceToHost));
  if (n < 0 || (size_t)n > eng->m.ifc.size()) return fail(eng, "breakage: inconsistent count of broken interfaces");
  std::vector<int32_t> broken(n);
  if (n > 0) HIPCHK(hipMemcpy(broken.data(), eng->ifc_broken, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  std::sort(broken.begin(), broken.end());
  for (int32_t k : broken) {
    if (k < 0 || (size_t)k >= eng->m.ifc.size() || !eng->m.ifc[k].alive)
      return fail(eng, "breakage: a listed interface is not a live one");
    eng->m.totalBrokenArea += eng->m.ifc[k].area;
    eng->m.ifc[k].alive = 0;
    eng->live_interfaces--;
  }
  const int32_t z[2] = {0, 0};
  HIPCHK(hipMemcpy(eng->ifc_nbroken, z, sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(eng->flags, z, sizeof(int32_t), hipMemcpyHostToDevice));
  eng->needUpdate = true;
  return 0;
}

// Rockable::accelerations (Rockable.cpp:3562-3719); kick fuses the closing
// half kick of velocityVerletStep.  guard: a step of the guarded loop (the
// guard kernel follows; it turns a breakage into a freeze that the host
// reads with the guard); other calls account a breakage at once.
int run_accelerations(rk_engine* eng, int kick, int write_geom, bool guard = false) {
  const StepParams sp = step_params(eng);
  ensure_gv(eng, false);
  tic(eng, 1);
  tic(eng, 2);
  launch_interact(eng->d, sp, write_geom || eng->record_geometry, eng->stream);
  toc(eng, "interact", 2);
  tic(eng, 2);
  launch_gather(eng->d, sp, kick, guard && eng->d.disp_check ? 1 : 0, eng->stream);
  toc(eng, "gather", 2);
  toc(eng, "force", 1);
  if (eng->m.law == LAW_STICKED && eng->live_interfaces > 0) {
    tic(eng, 2);
    launch_breakage(eng->d, sp, eng->stream);
    toc(eng, "breakage", 2);
    if (!guard) {
      int32_t fl = 0;
      HIPCHK(hipMemcpyAsync(&fl, eng->flags, sizeof(int32_t), hipMemcpyDeviceToHost, eng->stream));
      HIPCHK(hipStreamSynchronize(eng->stream));
      if (fl && account_breakage(eng) != 0) return -1;
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int do_update_nl(rk_engine* eng) {
  if (eng->tuning) HIPCHK(hipStreamSynchronize(eng->stream));
  const auto t0 = std::chrono::steady_clock::now();
  struct NlClock {  // host wall time of the rebuild (it ends with a synchronisation), for perf.txt
    rk_engine* e;
    std::chrono::steady_clock::time_point t0;
    ~NlClock() { e->nl_since_conf_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  } nl_clock{eng, t0};
  ensure_gv(eng, true);
  // the entries the last step skipped get their gaps before the merge
  // copies them into the new list
  launch_refresh_gaps(eng->d, eng->stream);
  tic(eng, eng->profile);
  const int nxt = 1 - eng->cur;
  if (nl_rebuild(eng->w, eng->d, eng->nlp, eng->E[eng->cur], eng->E[nxt], eng->stream) != 0)
    return fail(eng, std::string("neighbour list: ") + eng->w.err);
  eng->cur = nxt;
  if (ensure_gpart(eng) != 0) return -1;
  assemble_dev(eng);
  launch_entry_desc(eng->d, eng->stream);
  eng->nl_updates++;
  eng->last_interval = eng->steps - eng->last_rebuild_step;
  eng->last_rebuild_step = eng->steps;
  toc(eng, "nl", eng->profile);
  if (eng->tuning) {
    HIPCHK(hipStreamSynchronize(eng->stream));
    eng->nl_wall_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return 0;
}

// guard: the step ends with the guard kernel (the displacement check of
// dynamicUpdateNL, its accumulation fused into the resultant kernels, and
// the freeze after a bond breakage)
int vv_step(rk_engine* eng, int write_geom, bool guard) {
  const StepParams sp = step_params(eng);
  if (guard && !eng->guard) return fail(eng, "device guard: not allocated");
  if (guard && eng->d.disp_check && eng->gpart_cap < gather_blocks_max(eng->d.nb, eng->d.nhch))
    return fail(eng, "device guard: no room for the gather's maxima");
  tic(eng, 2);
  launch_predict(eng->d, sp, 1, eng->stream);
  toc(eng, "predict", 2);
  // the predictor leaves the listed bodies' global vertices current (every
  // body's while there is no list)
  eng->gv_state = eng->d.soff ? rk_engine::GV_LISTED : rk_engine::GV_ALL;
  if (run_accelerations(eng, 1, write_geom, guard) != 0) return -1;
  if (guard) {
    tic(eng, 2);
    launch_guard(eng->d, eng->stream);
    toc(eng, "guard", 2);
  }
  return 0;
}

// One step of the loop as the device runs it (write_geom 0): the
// velocity-Verlet step, and the guard kernel in the guarded loop.
int launch_one_step(rk_engine* eng, bool with_guard) {
  return vv_step(eng, 0, with_guard);
}

// What a captured run of steps depends on: the device view (pointers and
// sizes), the step parameters and the launch options.
std::vector<unsigned char> graph_key(rk_engine* eng, int n, bool with_guard) {
  const StepParams sp = step_params(eng);
  const double spv[6] = {sp.dt, sp.dt_2, sp.dt2_2, sp.gx, sp.gy, sp.gz};
  const int iv[5] = {n, with_guard ? 1 : 0, sp.law, eng->record_geometry, eng->live_interfaces};
  std::vector<unsigned char> k(sizeof(Dev) + sizeof spv + sizeof iv);
  memcpy(k.data(), &eng->d, sizeof(Dev));
  memcpy(k.data() + sizeof(Dev), spv, sizeof spv);
  memcpy(k.data() + sizeof(Dev) + sizeof spv, iv, sizeof iv);
  return k;
}

// A graph of n steps for the current device view, from the cache (the last
// few views; the two entry buffers alternate at each rebuild) or captured
// now.  Capturing launches nothing.
int step_graph(rk_engine* eng, int n, bool with_guard, hipGraphExec_t* out) {
  std::vector<unsigned char> key = graph_key(eng, n, with_guard);
  eng->graph_clock++;
  for (auto& g : eng->graphs)
    if (g.key == key) {
      g.used = eng->graph_clock;
      *out = g.exec;
      return 0;
    }
  HIPCHK(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < n && rc == 0; k++) rc = launch_one_step(eng, with_guard);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(eng->stream, &graph);
  eng->gv_state = rk_engine::GV_LISTED;  // each step of the graph ends with the listed bodies' fresh
  if (rc != 0 || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(eng, rc != 0 ? eng->err : std::string("step capture: ") + hipGetErrorString(ec));
  }
  eng->graph_captures++;
  // after a rebuild only kernel arguments and grid sizes change: update the
  // least recently used graph of the same length in place, which is much
  // cheaper than instantiating a new one
  int lru = -1;
  for (size_t q = 0; q < eng->graphs.size(); q++)
    if (eng->graphs[q].n == n && eng->graphs[q].with_guard == with_guard &&
        (lru < 0 || eng->graphs[q].used < eng->graphs[lru].used))
      lru = (int)q;
  if (lru >= 0 && eng->graphs.size() >= 2) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(eng->graphs[lru].exec, graph, &bad, &res) == hipSuccess &&
        res == hipGraphExecUpdateSuccess) {
      (void)hipGraphDestroy(graph);
      eng->graphs[lru].key = std::move(key);
      eng->graphs[lru].used = eng->graph_clock;
      *out = eng->graphs[lru].exec;
      return 0;
    }
    (void)hipGetLastError();
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(eng, std::string("step graph: ") + hipGetErrorString(ei));
  if (eng->graphs.size() >= 4) {
    size_t old = 0;
    for (size_t q = 1; q < eng->graphs.size(); q++)
      if (eng->graphs[q].used < eng->graphs[old].used) old = q;
    (void)hipGraphExecDestroy(eng->graphs[old].exec);
    eng->graphs.erase(eng->graphs.begin() + old);
  }
  rk_engine::StepGraph g;
  g.key = std::move(key);
  g.exec = exec;
  g.used = eng->graph_clock;
  g.n = n;
  g.with_guard = with_guard;
  eng->graphs.push_back(std::move(g));
  *out = exec;
  return 0;
}

// Launch the deferred steps: whole runs of graph_steps as captured graphs,
// the rest one by one.  Every host action that reads or changes the device
// state (rebuild, conf file, guard read, counts) comes after a flush.
int flush_steps(rk_engine* eng) {
  int64_t n = eng->pending;
  const bool with_guard = eng->pending_guard;
  eng->pending = 0;
  const int G = eng->graph_steps;
  if (G > 1 && n >= G && eng->profile == 0) {
    hipGraphExec_t exec = nullptr;
    if (step_graph(eng, G, with_guard, &exec) != 0) return -1;
    for (; n >= G; n -= G) {
      HIPCHK(hipGraphLaunch(exec, eng->stream));
      eng->graph_launches++;
    }
    // a replayed step leaves only the listed bodies' global vertices current
    eng->gv_state = rk_engine::GV_LISTED;
  }
  for (; n > 0; n--)
    if (launch_one_step(eng, with_guard) != 0) return -1;
  return 0;
}

// One step of the loop from the host's side: deferred when nothing on the
// host depends on it (no conf file written at it, no bonds outside the
// guarded loop, whose breakage the host accounts after the step), launched
// now otherwise.
int enqueue_step(rk_engine* eng, bool write_geom, bool with_guard) {
  const bool bonds = !with_guard && eng->m.law == LAW_STICKED && eng->live_interfaces > 0;
  if (eng->pending > 0 && eng->pending_guard != with_guard && flush_steps(eng) != 0) return -1;
  if (write_geom || bonds || eng->graph_steps <= 1 || eng->m.servo_timed()) {
    if (flush_steps(eng) != 0) return -1;
    return vv_step(eng, write_geom ? 1 : 0, with_guard);
  }
  eng->pending_guard = with_guard;
  eng->pending++;
  return 0;
}

int download_bodies(rk_engine* eng) {
  Model& m = eng->m;
  const int nb = (int)m.bodies.size();
  std::vector<BodyKin> kin(nb);
  std::vector<double> acc(3 * (size_t)nb), arot(3 * (size_t)nb), F(3 * (size_t)nb), M(3 * (size_t)nb);
  if (nb == 0) return 0;
  HIPCHK(hipStreamSynchronize(eng->stream));
  HIPCHK(hipMemcpy(kin.data(), eng->kin, sizeof(BodyKin) * nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(acc.data(), eng->acc, 8 * 3 * (size_t)nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(arot.data(), eng->arot, 8 * 3 * (size_t)nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(F.data(), eng->force, 8 * 3 * (size_t)nb, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(M.data(), eng->moment, 8 * 3 * (size_t)nb, hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; b++) {
    Body& p = m.bodies[b];
    const BodyKin& k = kin[b];
    p.pos = V(k.px, k.py, k.pz);
    p.Q.w = k.qw;
    p.Q.x = k.qx;
    p.Q.y = k.qy;
    p.Q.z = k.qz;
    p.vel = V(k.vx, k.vy, k.vz);
    p.vrot = V(k.wx, k.wy, k.wz);
    p.acc = V(acc[3 * b], acc[3 * b + 1], acc[3 * b + 2]);
    p.arot = V(arot[3 * b], arot[3 * b + 1], arot[3 * b + 2]);
    p.force = V(F[3 * b], F[3 * b + 1], F[3 * b + 2]);
    p.moment = V(M[3 * b], M[3 * b + 1], M[3 * b + 2]);
  }
  return 0;
}

int download_interactions(rk_engine* eng, int active_only, std::vector<Inter>& out) {
  const EntryArrays& E = eng->E[eng->cur];
  const int n = E.n;
  out.clear();
  if (n == 0) return 0;
  if (!active_only) {
    // gaps of the entries the far-entry skip left untouched
    if (flush_steps(eng) != 0) return -1;
    ensure_gv(eng, false);
    launch_refresh_gaps(eng->d, eng->stream);
    HIPCHK(hipGetLastError());
  }
  std::vector<int32_t> ei(n), ej(n), st(n);
  std::vector<uint32_t> sub(n);
  std::vector<EState> es(n);
  std::vector<EGeom> eg(n);
  std::vector<double> dp(n);
  std::vector<uint8_t> act(n);
  HIPCHK(hipStreamSynchronize(eng->stream));
  HIPCHK(hipMemcpy(ei.data(), E.ei, 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ej.data(), E.ej, 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sub.data(), E.esub, 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(st.data(), E.estick, 4 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(es.data(), E.est, sizeof(EState) * n, hipMemcpyDeviceToHost));
  std::vector<double> dn(n);
  HIPCHK(hipMemcpy(dn.data(), E.edn, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(eg.data(), E.egeo, sizeof(EGeom) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(dp.data(), E.edamp, 8 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(act.data(), E.eact, n, hipMemcpyDeviceToHost));
  out.reserve(n);
  for (int k = 0; k < n; k++) {
    if (active_only && !(act[k] & 1)) continue;
    Inter I;
    I.i = ei[k];
    I.j = ej[k];
    I.type = (int)(sub[k] >> 30);
    I.isub = (int)((sub[k] >> 15) & 0x7fffu);
    I.jsub = (int)(sub[k] & 0x7fffu);
    I.stick = st[k];
    I.n = V(eg[k].nx, eg[k].ny, eg[k].nz);
    I.pos = V(eg[k].px, eg[k].py, eg[k].pz);
    I.vel = V(eg[k].vx, eg[k].vy, eg[k].vz);
    I.dn = dn[k];
    I.fn = es[k].fn;
    I.ft = V(es[k].ftx, es[k].fty, es[k].ftz);
    I.mom = V(es[k].momx, es[k].momy, es[k].momz);
    I.damp = dp[k];
    out.push_back(I);
  }
  return 0;
}

// getKineticEnergy over every body (Rockable.cpp:3958-3970, first = 0).
void kinetic_energy(const Model& m, double* Et, double* Er) {
  double et = 0.0, er = 0.0;
  for (const Body& p : m.bodies) {
    et += p.mass * vdot(p.vel, p.vel);
    const v3 w = qrot(qconj(p.Q), p.vrot);
    er += (p.inertia.x * w.x * w.x + p.inertia.y * w.y * w.y + p.inertia.z * w.z * w.z);
  }
  *Et = et * 0.5;
  *Er = er * 0.5;
}

// getResultantQuickStats from nDriven (Rockable.cpp:4145-4206) and
// getInteractionQuickStats (4217-4252) over the active interactions, in
// their serial order.
void quick_stats(const Model& m, const std::vector<Inter>& act, double* Fmax, double* F_fnmax, double* fnMaxOut,
                 double* fnMeanOut) {
  const int nb = (int)m.bodies.size();
  std::vector<double> fnMax(nb + 1, 0.0);
  std::vector<size_t> nbCtc(nb + 1, 0);
  for (const Inter& I : act) {
    nbCtc[I.i] += 1;
    nbCtc[I.j] += 1;
    if (fnMax[I.i] < I.fn) fnMax[I.i] = I.fn;
    if (fnMax[I.j] < I.fn) fnMax[I.j] = I.fn;
  }
  const int first = m.nDriven, last = nb - 1;
  double F = first <= last ? sqrt(vnorm2(m.bodies[first].force)) : 0.0;
  *Fmax = F;
  *F_fnmax = 0.0;
  for (int i = first; i <= last; i++) {
    if (nbCtc[i] <= 1) continue;
    F = sqrt(vnorm2(m.bodies[i].force));
    if (F > *Fmax) *Fmax = F;
    if (fnMax[i] > 0.0) {
      F /= fnMax[i];
      if (F > *F_fnmax) *F_fnmax = F;
    }
  }
  if (act.empty()) {
    *fnMaxOut = *fnMeanOut = 0.0;
    return;
  }
  double fn = act[0].fn, mx = fn, mean = 0.0;
  size_t n = 1;
  for (size_t k = 1; k < act.size(); k++) {
    n += 1;
    fn = act[k].fn;
    mean += fn;
    if (fn > mx) mx = fn;
  }
  if (n > 1) mean /= (double)(n - 1);
  *fnMaxOut = mx;
  *fnMeanOut = mean;
}

// The records integrate() writes at each conf save (Rockable.cpp:2600-2668):
// perf.txt (t, simulated interConf per wall second, % of that wall time in
// neighbour-list updates and in force phases, as the host saw them),
// staticBalance.txt, kineticEnergy.txt; numbers as an std::ostream prints them.
void conf_outputs(rk_engine* eng, const std::vector<Inter>& act) {
  const Model& m = eng->m;
  if (eng->perf_file) {
    const auto now = std::chrono::steady_clock::now();
    const double el = std::chrono::duration<double>(now - eng->conf_clock).count();
    const double nlp = el > 0.0 ? std::round(1000.0 * eng->nl_since_conf_s / el) / 10.0 : 0.0;
    const double fp = el > 0.0 ? std::round(1000.0 * eng->force_since_conf_s / el) / 10.0 : 0.0;
    fprintf(eng->perf_file, "%g %g %g %g\n", m.t, el > 0.0 ? m.interConf / el : 0.0, nlp, fp);
    fflush(eng->perf_file);
    eng->conf_clock = now;
    eng->nl_since_conf_s = eng->force_since_conf_s = 0.0;
  }
  double Fmax, F_fnmax, fnMax, fnMean;
  quick_stats(m, act, &Fmax, &F_fnmax, &fnMax, &fnMean);
  if (eng->sb_file && fnMax != 0.0 && fnMean != 0.0) {
    fprintf(eng->sb_file, "%g %g %g\n", m.t, F_fnmax, fabs(Fmax / fnMean));
    fflush(eng->sb_file);
  }
  if (eng->ke_file) {
    double Et, Er;
    kinetic_energy(m, &Et, &Er);
    fprintf(eng->ke_file, "%g %g %g\n", m.t, Et, Er);
    fflush(eng->ke_file);
  }
}

int save_conf_now(rk_engine* eng, const std::string& path, bool outputs = false) {
  if (download_bodies(eng) != 0) return -1;
  std::vector<Inter> act;
  if (download_interactions(eng, 1, act) != 0) return -1;
  if (outputs) conf_outputs(eng, act);
  std::string err;
  if (save_conf(eng->m, act, path, err) != 0) return fail(eng, err);
  return 0;
}

// A record of the data extractors is due at integrate()'s current step.
bool record_due(const rk_engine* eng, bool* needs_geometry) {
  bool due = false, geo = false;
  for (const auto& x : eng->extractors)
    if (x.nrec > 0 && eng->ext_step % x.nrec == 0) {
      due = true;
      geo = geo || x.kind == 1;
    }
  if (needs_geometry) *needs_geometry = geo;
  return due;
}

// DataExtractor::record of dnStat (DataExtractor_dnStat.cpp:69-105) and
// TrackDamage (DataExtractor_TrackDamage.cpp:47-50) at integrate()'s step.
int extractors_record(rk_engine* eng) {
  if (!record_due(eng, nullptr)) return 0;
  std::vector<Inter> act;
  bool need_act = false;
  for (const auto& x : eng->extractors)
    if (x.kind == 1 && x.nrec > 0 && eng->ext_step % x.nrec == 0) need_act = true;
  if (need_act && download_interactions(eng, 1, act) != 0) return -1;
  const double t = eng->m.t;
  for (const auto& x : eng->extractors) {
    if (x.nrec <= 0 || eng->ext_step % x.nrec != 0) continue;
    if (x.kind == 1) {
      double dnMin = 1e20, dnMax = -1e20, dnMeanNeg = 0.0, dnMeanPos = 0.0, nbNeg = 0.0, nbPos = 0.0, dnMean = 0.0;
      v3 posMin = V(0.0, 0.0, 0.0);
      for (const Inter& I : act) {
        const double dn = I.dn;
        dnMean += dn;
        if (dn < 0.0) {
          dnMeanNeg += dn;
          nbNeg += 1.0;
        } else {
          dnMeanPos += dn;
          nbPos += 1.0;
        }
        if (dn < dnMin) {
          dnMin = dn;
          posMin = I.pos;
        }
        if (dn > dnMax) dnMax = dn;
      }
      if (nbNeg > 0.0) dnMeanNeg /= nbNeg;
      if (nbPos > 0.0) dnMeanPos /= nbPos;
      if (nbNeg > 0.0 || nbPos > 0.0) dnMean /= (nbNeg + nbPos);
      fprintf(x.f, "%g %g %g %g %g %g %d %d %g %g %g\n", t, dnMin, dnMax, dnMean, dnMeanNeg, dnMeanPos, (int)nbNeg,
              (int)nbPos, posMin.x, posMin.y, posMin.z);
    } else {
      fprintf(x.f, "%g %g\n", t, eng->m.totalBrokenArea / eng->inAreaInterfaces);
    }
    fflush(x.f);
  }
  return 0;
}

void close_outputs(rk_engine* eng) {
  for (FILE** f : {&eng->perf_file, &eng->ke_file, &eng->sb_file})
    if (*f) {
      fclose(*f);
      *f = nullptr;
    }
}

void close_extractors(rk_engine* eng) {
  for (auto& x : eng->extractors)
    if (x.f) fclose(x.f);
  eng->extractors.clear();
}

int check_loaded(rk_engine* eng) {
  if (!eng) return -1;
  if (!eng->loaded) return fail(eng, "no configuration loaded");
  return 0;
}

// reset: the start of integrate(), which zeroes the dynamicCheckUpdateNL
// accumulations (Rockable.cpp:2552-2560); a continued run keeps them.
int integrate_prepare(rk_engine* eng, bool reset) {
  Model& m = eng->m;
  m.dt_2 = 0.5 * m.dt;
  m.dt2_2 = m.dt_2 * m.dt;
  m.dt2 = m.dt * m.dt;
  if (reset && eng->guard) {
    // the accumulations restart at the next step, the maxima are cleared
    DevGuard g;
    HIPCHK(hipMemcpy(&g, eng->guard, sizeof g, hipMemcpyDeviceToHost));
    g.reset = 1;
    g.maxpos2 = 0ull;
    g.minw = ~0ull;
    HIPCHK(hipMemcpy(eng->guard, &g, sizeof g, hipMemcpyHostToDevice));
  }
  return 0;
}

int upload_controls(rk_engine* eng);

// The controls and servo parameters as read from the driving file: equal
// signatures mean that a re-read changed nothing.
std::vector<double> control_signature(const Model& m) {
  std::vector<double> v;
  for (const Control& c : m.ctrl) {
    const double x[6] = {(double)c.type, (double)c.i, c.value, c.vec.x, c.vec.y, c.vec.z};
    v.insert(v.end(), x, x + 6);
  }
  const double sv[] = {(double)m.servo, (double)m.servo_icontr, m.servo_pressure, m.servo_velocity,
                       m.servo_val[0], m.servo_val[1], m.servo_val[2], m.servo_val[3], m.servo_val[4],
                       m.servo_val[5], m.lode_a, m.lode_b, m.lode_rate, m.shake_dir.x, m.shake_dir.y,
                       m.shake_dir.z, m.shake_A, m.shake_omega, m.shake_T, m.shake_tini, m.ramp_vb, m.ramp_ve,
                       m.ramp_tb, m.ramp_te};
  v.insert(v.end(), sv, sv + sizeof sv / sizeof sv[0]);
  for (int k = 0; k < 6; k++) v.push_back(m.servo_ids[k]);
  return v;
}

// At each conf save the reference reads its driving file again and applies
// the servo (Rockable.cpp:2682-2688): new controls take effect from the next
// step.  The device copy is refreshed only when the file's content changed.
int reread_driving(rk_engine* eng) {
  if (eng->m.driving_path.empty()) return 0;
  if (flush_steps(eng) != 0) return -1;
  std::string err;
  if (load_driving(eng->m, eng->m.driving_path, err) != 0) return fail(eng, err);
  if (control_signature(eng->m) == eng->ctrl_sig) return 0;
  return upload_controls(eng);
}

// one iteration of the integrate() loop (Rockable.cpp:2584-2745)
int loop_body(rk_engine* eng, const char* outdir) {
  Model& m = eng->m;
  m.t += m.dt;
  m.interConfC += m.dt;
  m.interVerletC += m.dt;
  const bool conf_due = m.interConfC >= m.interConf - m.dt_2;
  bool rec_geo = false;
  const bool rec = outdir && record_due(eng, &rec_geo);
  if (enqueue_step(eng, (conf_due && outdir) || rec_geo, false) != 0) return -1;
  eng->steps++;
  if (conf_due) {
    m.iconf++;
    if (outdir) {
      char fn[4096];
      snprintf(fn, sizeof fn, "%s/conf%d", outdir, m.iconf);
      if (save_conf_now(eng, fn, true) != 0) return -1;
    }
    if (reread_driving(eng) != 0) return -1;
    m.interConfC = 0.0;
  }
  if (eng->needUpdate || m.interVerletC >= m.interVerlet - m.dt_2) {
    if (flush_steps(eng) != 0) return -1;
    if (do_update_nl(eng) != 0) return -1;
    if (!eng->needUpdate)
      m.interVerletC = 0.0;
    else
      eng->needUpdate = false;
  }
  if (rec) {
    if (flush_steps(eng) != 0) return -1;
    if (extractors_record(eng) != 0) return -1;
  }
  eng->ext_step++;
  return 0;
}

// The loop scalars of one iteration (the host keeps them; the step itself
// runs on the device).
struct LoopScalars {
  double t, interConfC, interVerletC;
  int iconf;
  int64_t steps, ext_step;
};
LoopScalars loop_scalars(const rk_engine* eng) {
  return {eng->m.t, eng->m.interConfC, eng->m.interVerletC, eng->m.iconf, eng->steps, eng->ext_step};
}
void set_loop_scalars(rk_engine* eng, const LoopScalars& s) {
  eng->m.t = s.t;
  eng->m.interConfC = s.interConfC;
  eng->m.interVerletC = s.interVerletC;
  eng->m.iconf = s.iconf;
  eng->steps = s.steps;
  eng->ext_step = s.ext_step;
}

// The 16 network inputs of the current state into eng->feat (device).
int compute_features(rk_engine* eng) {
  const int np = std::max(1, ((int)eng->m.bodies.size() + 255) / 256);
  if (!eng->feat) {
    HIPCHK(hipMalloc(&eng->feat, 16 * sizeof(float)));
    HIPCHK(hipMalloc(&eng->pred, sizeof(float)));
  }
  if (eng->feat_part_cap < (size_t)np * 9) {
    if (eng->feat_part) (void)hipFree(eng->feat_part);
    eng->feat_part = nullptr;
    HIPCHK(hipMalloc(&eng->feat_part, (size_t)np * 9 * sizeof(double)));
    eng->feat_part_cap = (size_t)np * 9;
  }
  launch_guard_features(eng->d, step_params(eng), eng->m.dispUpdateNL, eng->m.angleUpdateNL,
                        (long long)(eng->steps - eng->last_rebuild_step), (long long)eng->last_interval,
                        eng->feat_part, eng->feat, eng->stream);
  HIPCHK(hipGetLastError());
  return 0;
}

// The next batch size proposed by the network: exp(y) - 1 iterations.
int propose_batch(rk_engine* eng) {
  if (compute_features(eng) != 0) return -1;
  if (rk_mlp_forward_device(eng->model, eng->feat, eng->pred, eng->stream) != 0) return fail(eng, "network forward failed");
  float y = 0.0f;
  HIPCHK(hipMemcpyAsync(&y, eng->pred, sizeof y, hipMemcpyDeviceToHost, eng->stream));
  HIPCHK(hipStreamSynchronize(eng->stream));
  double b = eng->model_scale * eng->model_trust * std::expm1((double)y);
  if (!(b == b)) b = 1.0;
  eng->pred_batch = (int64_t)std::max(1.0, std::min(b, 65536.0));
  return 0;
}

// Waits until the device has completed target steps (the check's progress
// word) or frozen the steps; 1 if frozen, 0 if not, -1 on an error.  The
// deferred steps are launched first; a drained stream ends the wait (the
// words are then final).
int wait_progress(rk_engine* eng, int64_t target) {
  if (flush_steps(eng) != 0) return -1;
  const volatile HostNote* n = eng->note;
  for (unsigned spin = 0;; spin++) {
    if (n->frozen) return 1;
    if (n->done >= target) return 0;
    if ((spin & 255u) == 255u) {
      const hipError_t q = hipStreamQuery(eng->stream);
      if (q == hipSuccess) return n->frozen ? 1 : 0;
      if (q != hipErrorNotReady) HIPCHK(q);
    }
    std::this_thread::yield();
  }
}

// The integrate() loop with the device guard (dynamicUpdateNL, its
// displacement check on the device, or bonds that may break).  Iterations
// are enqueued in batches without host synchronisation; a batch ends where
// the host has work (a conf file, a scheduled rebuild), after guard_batch
// iterations (no network), or once the host has seen the check freeze the
// steps (with a network).  When the device finds a rebuild due after
// iteration s (the check fired the iteration before, or a bond broke in s),
// the step kernels of the later iterations do nothing: the host reads the
// guard at the end of the batch, keeps the loop scalars of iteration s,
// accounts the broken interfaces, rebuilds and continues.  The schedule of
// rebuilds is exactly the reference's.
int run_guarded(rk_engine* eng, const char* outdir, int64_t nmax, bool until_tmax, int64_t* ndone) {
  Model& m = eng->m;
  int64_t done = 0;
  std::vector<LoopScalars> hist;
  std::vector<char> conf_at;
  while (done < nmax && (!until_tmax || m.t <= m.tmax)) {
    const bool bonds = m.law == LAW_STICKED && eng->live_interfaces > 0;
    // With the network the batch runs on until a host action or a rebuild is
    // seen.  The check writes its progress and its freeze to host-mapped
    // words, which the host reads after every step it enqueues; it lets the
    // device fall at most a lag of steps behind (waiting on the progress
    // word), the lag following the network's estimate of the steps to the
    // next rebuild: long far from it, short near it, so that few steps are
    // enqueued past a freeze.  Without the mapped words: a copy of the frozen
    // flag to pinned memory and an event after each chunk of steps of that
    // length, at most kPollAhead chunks in flight.
    const bool pipe = eng->model != nullptr && eng->poll_host != nullptr;
    const int64_t want = pipe ? nmax - done
                         : eng->model ? eng->pred_batch
                         : bonds ? std::min<int64_t>(eng->bond_batch, eng->guard_batch) : eng->guard_batch;
    const int64_t B = std::max<int64_t>(1, std::min<int64_t>(want, nmax - done));
    hist.assign(1, loop_scalars(eng));
    conf_at.assign(1, 0);
    launch_guard_set_step(eng->d, (long long)eng->steps, eng->stream);
    int64_t k = 0;
    int head = 0, inflight = 0;
    bool seen = false;
    auto chunk_for = [&](int64_t taken) -> int64_t {
      const int64_t r = eng->pred_batch - taken;
      return std::max<int64_t>(2, std::min<int64_t>(32, r / 2));
    };
    int64_t chunk_left = pipe ? chunk_for(0) : B;
    while (k < B) {
      m.t += m.dt;
      m.interConfC += m.dt;
      m.interVerletC += m.dt;
      const bool conf_due = m.interConfC >= m.interConf - m.dt_2;
      bool rec_geo = false;
      const bool rec = outdir && record_due(eng, &rec_geo);
      if (enqueue_step(eng, (conf_due && outdir) || rec_geo, true) != 0) return -1;
      eng->steps++;
      if (conf_due) {
        m.iconf++;
        m.interConfC = 0.0;
      }
      eng->ext_step++;
      hist.push_back(loop_scalars(eng));
      conf_at.push_back(conf_due ? 1 : 0);
      k++;
      if ((conf_due && (outdir || !m.driving_path.empty())) || rec || eng->needUpdate ||
          m.interVerletC >= m.interVerlet - m.dt_2 || (until_tmax && m.t > m.tmax))
        break;
      if (pipe && eng->note) {
        const volatile HostNote* n = eng->note;
        if (n->frozen) break;
        const int64_t target = (int64_t)eng->steps - chunk_for(k);
        if (n->done < target) {
          eng->guard_polls++;
          const int r = wait_progress(eng, target);
          if (r < 0) return -1;
          if (r == 1) break;
        }
        continue;
      }
      if (pipe && --chunk_left == 0) {
        if (flush_steps(eng) != 0) return -1;
        const int slot = (head + inflight) % kPollSlots;
        HIPCHK(hipMemcpyAsync(eng->poll_host + slot, &eng->guard->frozen, sizeof(int), hipMemcpyDeviceToHost,
                              eng->stream));
        HIPCHK(hipEventRecord(eng->poll_ev[slot], eng->stream));
        inflight++;
        while (inflight > 0) {
          const hipError_t q =
              inflight > kPollAhead ? hipEventSynchronize(eng->poll_ev[head]) : hipEventQuery(eng->poll_ev[head]);
          if (q == hipErrorNotReady) break;
          HIPCHK(q);
          if (eng->poll_host[head]) seen = true;
          head = (head + 1) % kPollSlots;
          inflight--;
        }
        eng->guard_polls++;
        if (seen) break;
        chunk_left = chunk_for(k);
      }
    }
    if (flush_steps(eng) != 0) return -1;
    DevGuard g;
    int32_t broke = 0;
    rk_engine::GuardRead* gh = eng->guard_host;
    HIPCHK(hipMemcpyAsync(gh ? &gh->g : &g, eng->guard, sizeof g, hipMemcpyDeviceToHost, eng->stream));
    if (bonds)
      HIPCHK(hipMemcpyAsync(gh ? &gh->broke : &broke, eng->flags, sizeof broke, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    if (gh) {
      g = gh->g;
      if (bonds) broke = gh->broke;
    }
    eng->guard_reads++;
    int64_t executed = k;
    bool need = eng->needUpdate;
    if (g.frozen) {
      executed = g.frozen_step - hist[0].steps + 1;
      if (executed < 1 || executed > k) return fail(eng, "device guard: inconsistent frozen step");
      set_loop_scalars(eng, hist[executed]);
      eng->guard_wasted += k - executed;
      if (eng->model && k - executed > 1) eng->model_trust = std::max(0.05, eng->model_trust * 0.7);
      need = true;
      const int zero = 0;
      HIPCHK(hipMemcpy(&eng->guard->frozen, &zero, sizeof zero, hipMemcpyHostToDevice));
      if (eng->note) *(volatile int32_t*)&eng->note->frozen = 0;
    }
    if (broke) {
      // a breakage freezes the steps after its own
      if (!g.frozen) return fail(eng, "device guard: a bond broke without a freeze");
      if (account_breakage(eng) != 0) return -1;
    }
    if (conf_at[executed] && outdir) {
      char fn[4096];
      snprintf(fn, sizeof fn, "%s/conf%d", outdir, m.iconf);
      if (save_conf_now(eng, fn, true) != 0) return -1;
    }
    if (conf_at[executed] && reread_driving(eng) != 0) return -1;
    if (need || m.interVerletC >= m.interVerlet - m.dt_2) {
      if (do_update_nl(eng) != 0) return -1;
      if (!need) m.interVerletC = 0.0;
      eng->needUpdate = false;
    }
    if (outdir) {
      // the data extractors of the last executed step (its counter was
      // advanced with it)
      eng->ext_step--;
      if (extractors_record(eng) != 0) return -1;
      eng->ext_step++;
    }
    done += executed;
    if (eng->model && !g.frozen) eng->model_trust = std::min(1.0, eng->model_trust * 1.1);
    if (bonds)
      eng->bond_batch = g.frozen ? std::max<int64_t>(2, 2 * executed)
                                 : std::min<int64_t>(eng->guard_batch, 2 * std::max<int64_t>(1, eng->bond_batch));
    if (eng->model && propose_batch(eng) != 0) return -1;
  }
  if (ndone) *ndone = done;
  return 0;
}

}  // namespace

extern "C" {

int rk_inspect_conf(const char* path, rk_conf_info* out, char* err, size_t errlen) {
  auto report = [&](const std::string& msg) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", msg.c_str());
    return -1;
  };
  if (!path) return report("no path given");
  if (!out) return report("no output record given");
  Model m;
  std::string e;
  if (load_conf(m, path, e) != 0) return report(e);
  out->nbodies = (int64_t)m.bodies.size();
  out->ndriven = m.nDriven;
  out->nshapes = (int64_t)m.shapes.size();
  out->ninteractions = (int64_t)m.inter.size();
  out->ninterfaces = (int64_t)m.ifc.size();
  out->ngroups = m.ngroup;
  out->law = m.law;
  out->linkcells = m.nl_linkcells;
  if (err && errlen > 0) err[0] = 0;
  return 0;
}

int rk_engine_create(rk_engine** out, int device) {
  if (!out) return -1;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return -2;
  if (hipSetDevice(device) != hipSuccess) return -2;
  rk_engine* eng = new rk_engine();
  eng->device = device;
  if (hipStreamCreate(&eng->stream) != hipSuccess || hipEventCreate(&eng->ev0) != hipSuccess ||
      hipEventCreate(&eng->ev1) != hipSuccess) {
    delete eng;
    return -2;
  }
  // pinned flags and events of the asynchronous guard polls (without them
  // the guarded loop reads the check synchronously)
  if (hipHostMalloc((void**)&eng->poll_host, sizeof(int) * kPollSlots, hipHostMallocDefault) == hipSuccess) {
    for (int k = 0; k < kPollSlots; k++)
      if (hipEventCreateWithFlags(&eng->poll_ev[k], hipEventDisableTiming) != hipSuccess) {
        (void)hipHostFree(eng->poll_host);
        eng->poll_host = nullptr;
        break;
      }
  } else {
    eng->poll_host = nullptr;
  }
  if (hipHostMalloc((void**)&eng->guard_host, sizeof(rk_engine::GuardRead), hipHostMallocDefault) != hipSuccess)
    eng->guard_host = nullptr;
  if (hipHostMalloc((void**)&eng->note, sizeof(HostNote), hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
    memset(eng->note, 0, sizeof(HostNote));
    if (hipHostGetDevicePointer((void**)&eng->note_dev, eng->note, 0) != hipSuccess) {
      (void)hipHostFree(eng->note);
      eng->note = eng->note_dev = nullptr;
    }
  } else {
    eng->note = nullptr;
  }
  *out = eng;
  return 0;
}

void rk_engine_destroy(rk_engine* eng) {
  if (!eng) return;
  close_outputs(eng);
  close_extractors(eng);
  (void)hipSetDevice(eng->device);
  if (eng->stream) (void)hipStreamSynchronize(eng->stream);
  free_device(eng);
  if (eng->skip_count) (void)hipFree(eng->skip_count);
  for (hipEvent_t e : eng->ev_pool) (void)hipEventDestroy(e);
  if (eng->ev0) (void)hipEventDestroy(eng->ev0);
  if (eng->ev1) (void)hipEventDestroy(eng->ev1);
  for (hipEvent_t e : eng->poll_ev)
    if (e) (void)hipEventDestroy(e);
  if (eng->poll_host) (void)hipHostFree(eng->poll_host);
  if (eng->guard_host) (void)hipHostFree(eng->guard_host);
  if (eng->note) (void)hipHostFree(eng->note);
  if (eng->stream) (void)hipStreamDestroy(eng->stream);
  delete eng;
}

const char* rk_last_error(const rk_engine* eng) { return eng ? eng->err.c_str() : "null engine"; }

namespace {
void reset_engine(rk_engine* eng) {
  (void)hipSetDevice(eng->device);
  close_outputs(eng);
  close_extractors(eng);
  free_device(eng);
  eng->m = Model();
  eng->loaded = false;
  eng->cur = 0;
  eng->needUpdate = false;
  eng->steps = 0;
  eng->nl_updates = 0;
  eng->packings.clear();
}
}  // namespace

int rk_load_batch(rk_engine* eng, const char* const* conf_paths, const char* const* driving_paths, int n) {
  if (!eng || !conf_paths || n < 1) return -1;
  reset_engine(eng);
  std::vector<Model> parts(n);
  std::string err;
  for (int p = 0; p < n; p++) {
    if (!conf_paths[p]) return fail(eng, "batch: null deck path");
    if (load_conf(parts[p], conf_paths[p], err) != 0) return fail(eng, std::string(conf_paths[p]) + ": " + err);
    if (driving_paths && driving_paths[p] && load_driving(parts[p], driving_paths[p], err) != 0)
      return fail(eng, std::string(driving_paths[p]) + ": " + err);
  }
  if (merge_models(parts, eng->m, eng->packings, err) != 0) return fail(eng, err);
  parts.clear();
  if (upload_model(eng) != 0) return -1;
  // (upload_model uploads the decks' controls with the rest)
  eng->loaded = true;
  // the "fake step" of loadConf (Rockable.cpp:1133-1138), every deck at once
  if (run_accelerations(eng, 0, 1) != 0) return -1;
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

int rk_batch_packing(rk_engine* eng, int p, int64_t* out) {
  if (check_loaded(eng) || !out) return -1;
  if (p < 0 || p >= (int)eng->packings.size()) return fail(eng, "no packing " + std::to_string(p));
  const PackingRange& r = eng->packings[p];
  out[0] = r.first_driven;
  out[1] = r.n_driven;
  out[2] = r.first_free;
  out[3] = r.n_free;
  return 0;
}

int rk_load_conf(rk_engine* eng, const char* path) {
  if (!eng || !path) return -1;
  reset_engine(eng);
  std::string err;
  if (load_conf(eng->m, path, err) != 0) return fail(eng, err);
  PackingRange one;
  one.n_driven = eng->m.nDriven;
  one.first_free = eng->m.nDriven;
  one.n_free = (int)eng->m.bodies.size() - eng->m.nDriven;
  eng->packings.assign(1, one);
  if (upload_model(eng) != 0) return -1;
  eng->loaded = true;
  // the "fake step" of loadConf (Rockable.cpp:1133-1138)
  if (run_accelerations(eng, 0, 1) != 0) return -1;
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

int rk_load_driving_system(rk_engine* eng, const char* path) {
  if (check_loaded(eng)) return -1;
  if (!path) return fail(eng, "no path given");
  (void)hipSetDevice(eng->device);
  if (flush_steps(eng) != 0) return -1;
  std::string err;
  if (load_driving(eng->m, path, err) != 0) return fail(eng, err);
  return upload_controls(eng);
}

namespace {
int upload_controls(rk_engine* eng) {
  std::vector<DevControl> c;
  for (const Control& x : eng->m.ctrl) {
    DevControl dc;
    dc.type = x.type;
    dc.i = x.i;
    dc.value = x.value;
    dc.vx = x.vec.x;
    dc.vy = x.vec.y;
    dc.vz = x.vec.z;
    c.push_back(dc);
  }
  const int nd = std::max(0, std::min(eng->m.nDriven, (int)eng->m.bodies.size()));
  std::vector<int32_t> off, idx;
  eng->ctrl_parallel = control_csr(eng->m, off, idx);
  if (upload(eng, eng->ctrl, c) || upload(eng, eng->ctrl_off, off) || upload(eng, eng->ctrl_idx, idx)) return -1;
  if (upload(eng, eng->passive, passive_flags(eng->m))) return -1;
  if (nd > 0) {
    // the driven bodies' far-entry bounds follow their controls; every entry
    // is evaluated at the next step
    const std::vector<double> cb = initial_bounds(eng->m);
    HIPCHK(hipMemcpyAsync(eng->cbound, cb.data(), sizeof(double) * nd, hipMemcpyHostToDevice, eng->stream));
  }
  launch_reset_keys(eng->d, eng->stream);
  // the heavy bodies' summation paths depend on the passive flags
  if (nl_build_heavy(eng->w, (int)eng->m.bodies.size(), eng->E[eng->cur], eng->passive, eng->stream) != 0)
    return fail(eng, eng->w.err);
  if (ensure_gpart(eng) != 0) return -1;
  assemble_dev(eng);
  HIPCHK(hipStreamSynchronize(eng->stream));
  eng->ctrl_sig = control_signature(eng->m);
  return 0;
}
}  // namespace

int rk_load_data_extractors(rk_engine* eng, const char* path) {
  if (check_loaded(eng)) return -1;
  if (!path) return fail(eng, "no path given");
  std::ifstream is(path);
  if (!is) return 0;  // no file: no extractor (Rockable.cpp:1157-1159)
  close_extractors(eng);
  const std::string p(path);
  const size_t slash = p.rfind('/');
  const std::string dir = slash == std::string::npos ? std::string() : p.substr(0, slash + 1);
  std::string tok;
  while (is >> tok) {
    if (tok[0] == '/' || tok[0] == '#' || tok[0] == '!') {
      std::getline(is, tok);
      continue;
    }
    const int kind = tok == "dnStat" ? 1 : tok == "TrackDamage" ? 2 : 0;
    if (!kind) continue;  // unknown: the reference warns, and reads on token by token
    std::string name;
    int nrec = 0;
    is >> name >> nrec;
    const std::string full = (!name.empty() && name[0] == '/') ? name : dir + name;
    FILE* f = fopen(full.c_str(), "w");
    if (!f) return fail(eng, "cannot write " + full);
    eng->extractors.push_back({kind, nrec, f});
    if (kind == 2) {  // TrackDamage::init: the live interfaces' total area
      eng->inAreaInterfaces = 0.0;
      for (const Interface& BI : eng->m.ifc)
        if (BI.alive) eng->inAreaInterfaces += BI.area;
    }
  }
  return 0;
}

int rk_set_force_law(rk_engine* eng, const char* name) {
  if (check_loaded(eng)) return -1;
  std::string n = name ? name : "";
  int rc = 0;
  if (n == "Avalanche") {
    eng->m.law = LAW_AVALANCHE;
    eng->m.lawName = n;
  } else if (n == "StickedLinks") {
    eng->m.law = LAW_STICKED;
    eng->m.lawName = n;
  } else {
    if (n != "Default") {
      eng->err = "The ForceLaw named " + n + " is unknown! -> set to Default";
      rc = 1;
    }
    eng->m.law = LAW_DEFAULT;
    eng->m.lawName = "Default";
  }
  return rc;
}

int rk_set_body_force(rk_engine* eng, const char* name, const double* a, int nargs) {
  if (check_loaded(eng)) return -1;
  if (nargs < 0 || (nargs > 0 && !a)) return fail(eng, "body force arguments missing");
  std::string n = name ? name : "";
  Model& m = eng->m;
  if (n.empty()) {
    m.bf = BF_NONE;
  } else if (n == "AttractingPoint" && nargs >= 4) {
    m.bf = BF_ATTRACTING;
    m.bf_point = V(a[0], a[1], a[2]);
    m.bf_acc = a[3];
  } else if (n == "PreferredDirection" && nargs >= 7) {
    m.bf = BF_PREFDIR;
    m.bf_axisBody = V(a[0], a[1], a[2]);
    m.bf_axis = V(a[3], a[4], a[5]);
    m.bf_momentMax = a[6];
    vnormalize(&m.bf_axisBody);
    vnormalize(&m.bf_axis);
    m.bf_Kr = 2.0 * m.bf_momentMax / M_PI;
  } else if (n == "ViscousFluid" && nargs >= 1) {
    m.bf = BF_VISCOUS;
    m.bf_fluidDensity = a[0];
  } else {
    return fail(eng, "unknown BodyForce '" + n + "' or wrong number of arguments");
  }
  assemble_dev(eng);
  return 0;
}

int rk_set_option(rk_engine* eng, const char* name, double value) {
  if (!eng || !name) return -1;
  std::string n = name;
  if (n == "record_geometry")
    eng->record_geometry = value != 0.0;
  else if (n == "profile") {
    if (!(value == 0.0 || value == 1.0 || value == 2.0)) return fail(eng, "profile must be 0, 1 or 2");
    resolve_events(eng);
    eng->profile = (int)value;
  }
  else if (n == "exact_sums") {
    eng->exact_sums = value != 0.0;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "skip_far") {
    if (!(value == 0.0 || value == 1.0)) return fail(eng, "skip_far must be 0 or 1");
    if (eng->loaded) {
      (void)hipSetDevice(eng->device);
      if (flush_steps(eng) != 0) return -1;
      ensure_gv(eng, false);
      launch_refresh_gaps(eng->d, eng->stream);  // the gaps the last step skipped
    }
    eng->skip_far = (int)value;
    if (eng->loaded) {
      assemble_dev(eng);
      launch_reset_keys(eng->d, eng->stream);  // keys are kept only while skipping is on
    }
  } else if (n == "skip_segment") {
    if (!(value == 256.0 || value == 512.0 || value == 1024.0 || value == 2048.0 || value == 4096.0))
      return fail(eng, "skip_segment must be 256, 512, 1024, 2048 or 4096");
    if (eng->loaded) {
      (void)hipSetDevice(eng->device);
      if (flush_steps(eng) != 0) return -1;
    }
    eng->skip_seg = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "count_skips") {
    // 1: the geometry pass counts the entries it skips, the inactive ones with
    // zero history and all it looks at (three integer atomics per wavefront;
    // read back as skipped_entries, zero_history_entries, entries_looked_at);
    // setting it again clears the counts; 0 stops counting
    if (!(value == 0.0 || value == 1.0)) return fail(eng, "count_skips must be 0 or 1");
    if (eng->loaded) {
      (void)hipSetDevice(eng->device);
      if (flush_steps(eng) != 0) return -1;
      HIPCHK(hipStreamSynchronize(eng->stream));
    }
    if (value == 0.0 && eng->skip_count) {
      (void)hipFree(eng->skip_count);
      eng->skip_count = nullptr;
    } else if (value == 1.0) {
      if (!eng->skip_count) HIPCHK(hipMalloc(&eng->skip_count, 3 * sizeof(unsigned long long)));
      HIPCHK(hipMemset(eng->skip_count, 0, 3 * sizeof(unsigned long long)));
    }
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "gravity_x" || n == "gravity_y" || n == "gravity_z") {
    // the conf keyword "gravity" (Rockable.cpp:978-980), changed between steps
    if (check_loaded(eng)) return -1;
    (n == "gravity_x" ? eng->m.gravity.x : n == "gravity_y" ? eng->m.gravity.y : eng->m.gravity.z) = value;
  } else if (n == "model_scale") {
    if (!(value > 0.0 && value <= 4.0)) return fail(eng, "model_scale must be in (0, 4]");
    eng->model_scale = value;
  } else if (n == "persistent") {
    if (!(value == 0.0 || (value >= 8.0 && value <= 1048576.0 && value == (double)(int)value)))
      return fail(eng, "persistent must be 0 or an integer in [8, 1048576]");
    eng->persist = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "gather_lanes") {
    if (!(value == 8.0 || value == 16.0)) return fail(eng, "gather_lanes must be 8 or 16");
    eng->gather_lanes = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "type_sort") {
    if (!(value == 0.0 || value == 1.0)) return fail(eng, "type_sort must be 0 or 1");
    eng->type_sort = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "narrow_pairs") {
    if (!(value == 0.0 || value == 1.0 || value == 2.0))
      return fail(eng, "narrow_pairs must be 0 (by shape size), 1 or 2");
    eng->w.narrow_pairs = (int)value;
  } else if (n == "gather_groups") {
    if (!(value == 0.0 || value == 1.0 || value == 2.0 || value == 4.0 || value == 8.0 || value == 16.0))
      return fail(eng, "gather_groups must be 0 (by size), 1, 2, 4, 8 or 16");
    eng->gather_groups = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "fuse_glob") {
    if (!(value == -1.0 || value == 0.0 || value == 1.0 || value == 2.0 || value == 3.0))
      return fail(eng, "fuse_glob must be -1 (by size), 0, 1, 2 or 3");
    eng->fuse_glob = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "fuse_law") {
    if (!(value == 0.0 || value == 1.0)) return fail(eng, "fuse_law must be 0 or 1");
    eng->fuse_law = (int)value;
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "law_group" || n == "geom_block" || n == "gather_waves" || n == "body_block") {
    const int v = (int)value;
    if ((double)v != value) return fail(eng, n + " must be an integer");
    if (n == "law_group") {
      if (v != 1 && v != 2 && v != 4 && v != 8) return fail(eng, "law_group must be 1, 2, 4 or 8");
      eng->law_group = v;
    } else if (n == "geom_block") {
      if (v != 64 && v != 128 && v != 256 && v != 512) return fail(eng, "geom_block must be 64, 128, 256 or 512");
      eng->geom_block = v;
    } else if (n == "gather_waves") {
      if (v != 1 && v != 2 && v != 4) return fail(eng, "gather_waves must be 1, 2 or 4");
      eng->gather_waves = v;
    } else {
      if (v != 64 && v != 128 && v != 256) return fail(eng, "body_block must be 64, 128 or 256");
      eng->body_block = v;
    }
    if (eng->loaded) assemble_dev(eng);
  } else if (n == "graph_steps") {
    if (!(value >= 0.0 && value <= 4096.0)) return fail(eng, "graph_steps must be in [0, 4096]");
    if (eng->pending > 0 && eng->loaded) {
      (void)hipSetDevice(eng->device);
      if (flush_steps(eng) != 0) return -1;
    }
    eng->graph_steps = (int)value;
    free_graphs(eng);
  } else if (n == "guard_batch") {
    if (!(value >= 1.0 && value <= 1e6)) return fail(eng, "guard_batch must be in [1, 1e6]");
    eng->guard_batch = (int)value;
  } else if (n == "heavy_min") {
    if (!(value >= 0.0 && value < 2.0e9)) return fail(eng, "heavy_min must be in [0, 2e9)");
    eng->w.heavy_min = (int)value;
    if (eng->loaded) {
      (void)hipSetDevice(eng->device);
      if (nl_build_heavy(eng->w, (int)eng->m.bodies.size(), eng->E[eng->cur], eng->passive, eng->stream) != 0)
        return fail(eng, eng->w.err);
      if (ensure_gpart(eng) != 0) return -1;
      assemble_dev(eng);
    }
  } else
    return fail(eng, "unknown option " + n);
  return 0;
}

int rk_update_nl(rk_engine* eng) {
  if (check_loaded(eng)) return -1;
  (void)hipSetDevice(eng->device);
  return do_update_nl(eng);
}

int rk_accelerations(rk_engine* eng) {
  if (check_loaded(eng)) return -1;
  (void)hipSetDevice(eng->device);
  if (run_accelerations(eng, 0, 1) != 0) return -1;
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

int rk_step(rk_engine* eng, int64_t nsteps) {
  if (check_loaded(eng)) return -1;
  (void)hipSetDevice(eng->device);
  if (integrate_prepare(eng, false) != 0) return -1;
  if (eng->guard) return run_guarded(eng, nullptr, nsteps, false, nullptr);
  for (int64_t k = 0; k < nsteps; k++)
    if (loop_body(eng, nullptr) != 0) return -1;
  return flush_steps(eng);
}

namespace {
int* tuned_option(rk_engine* eng, const std::string& n) {
  if (n == "law_group") return &eng->law_group;
  if (n == "geom_block") return &eng->geom_block;
  if (n == "gather_waves") return &eng->gather_waves;
  if (n == "body_block") return &eng->body_block;
  if (n == "graph_steps") return &eng->graph_steps;
  if (n == "fuse_law") return &eng->fuse_law;
  if (n == "gather_lanes") return &eng->gather_lanes;
  if (n == "gather_groups") return &eng->gather_groups;
  if (n == "narrow_pairs") return &eng->w.narrow_pairs;
  if (n == "type_sort") return &eng->type_sort;
  if (n == "fuse_glob") return &eng->fuse_glob;
  if (n == "persistent") return &eng->persist;
  return nullptr;
}
}  // namespace

int rk_autotune(rk_engine* eng, int steps_per_trial, int64_t* steps_done) {
  if (check_loaded(eng)) return -1;
  if (steps_per_trial < 1) return fail(eng, "steps_per_trial must be positive");
  (void)hipSetDevice(eng->device);
  if (eng->profile != 0) return fail(eng, "autotune needs profile 0");
  struct Knob {
    const char* name;
    std::vector<int> values;
  };
  const Knob knobs[] = {{"fuse_law", {0, 1}},
                        {"type_sort", {1, 0}},
                        {"skip_segment", {1024, 2048, 512, 4096, 256}},
                        {"law_group", {1, 2, 4, 8}},
                        {"geom_block", {64, 128, 256, 512}},
                        {"gather_lanes", {8, 16}},
                        {"gather_groups", {1, 4, 16}},
                        {"fuse_glob", {0, 1, 2, 3}},
                        {"gather_waves", {1, 2, 4}},
                        {"body_block", {64, 128, 256}},
                        {"persistent", {0, 2048, 8192}},
                        {"graph_steps", {0, 8, 32}}};
  int64_t done = 0;
  eng->tuning = true;
  // wall seconds per step of one trial, the rebuilds excluded
  auto trial = [&](double* per_step) -> int {
    HIPCHK(hipStreamSynchronize(eng->stream));
    const double nl0 = eng->nl_wall_s;
    const auto t0 = std::chrono::steady_clock::now();
    if (rk_step(eng, steps_per_trial) != 0) return -1;
    HIPCHK(hipStreamSynchronize(eng->stream));
    const double w = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    done += steps_per_trial;
    *per_step = (w - (eng->nl_wall_s - nl0)) / steps_per_trial;
    return 0;
  };
  // Per-kernel events (profile level 2) time the kernel-shape knobs: the sum of
  // the step kernels' durations per step, which a trial's host work, guard
  // reads and pipeline drains around a rebuild do not disturb.  The knobs of
  // the launches themselves (persistent grids, graph length) are timed by
  // wall time.  The caller's timers are kept aside meanwhile.
  auto trial_kernels = [&](double* per_step) -> int {
    resolve_events(eng);
    auto saved = eng->timers;
    eng->timers.clear();
    eng->profile = 2;
    const int r = rk_step(eng, steps_per_trial);
    resolve_events(eng);
    eng->profile = 0;
    double ms = 0.0;
    for (const char* n : {"predict", "interact", "gather", "guard"}) {
      auto it = eng->timers.find(n);
      if (it != eng->timers.end()) ms += it->second.ms;
    }
    eng->timers = saved;
    if (r != 0) return -1;
    done += steps_per_trial;
    *per_step = 1e-3 * ms / steps_per_trial;
    return 0;
  };
  int rc = 0;
  // the starting shapes (the defaults or the caller's), timed first: if the
  // search ends slower than them (its trials are short and one knob at a
  // time), they are kept
  std::vector<int> start;
  for (const Knob& k : knobs) {
    double v = 0.0;
    if (rk_get_option(eng, k.name, &v) != 0) rc = -1;
    start.push_back((int)v);
  }
  double t_start = 1e30;
  for (int rep = 0; rep < 2 && rc == 0; rep++) {
    double t = 0.0;
    const int warm = std::max(1, eng->graph_steps);
    if (rk_step(eng, warm) != 0 || trial(&t) != 0) rc = -1;
    done += warm;
    t_start = std::min(t_start, t);
  }
  double t_last = 1e30;
  for (const Knob& k : knobs) {
    if (rc != 0) break;
    std::vector<double> best(k.values.size(), 1e30);
    for (int rep = 0; rep < 2 && rc == 0; rep++)
      for (size_t c = 0; c < k.values.size() && rc == 0; c++) {
        if (rk_set_option(eng, k.name, k.values[c]) != 0) {
          rc = -1;
          break;
        }
        double t = 0.0;
        // a first short run captures the graphs of the new shape, then the trial
        const int warm = std::max(1, eng->graph_steps);
        const bool launch_knob = !strcmp(k.name, "persistent") || !strcmp(k.name, "graph_steps");
        if (rk_step(eng, warm) != 0 || (launch_knob ? trial(&t) : trial_kernels(&t)) != 0) {
          rc = -1;
          break;
        }
        done += warm;
        best[c] = std::min(best[c], t);
      }
    if (rc != 0) break;
    size_t b = 0;
    for (size_t c = 1; c < best.size(); c++)
      if (best[c] < best[b]) b = c;
    t_last = best[b];
    if (rk_set_option(eng, k.name, k.values[b]) != 0) {
      rc = -1;
      break;
    }
  }
  if (rc == 0) {
    // the chosen shapes timed again against the starting ones
    double t_end = 1e30;
    for (int rep = 0; rep < 2 && rc == 0; rep++) {
      double t = 0.0;
      const int warm = std::max(1, eng->graph_steps);
      if (rk_step(eng, warm) != 0 || trial(&t) != 0) rc = -1;
      done += warm;
      t_end = std::min(t_end, t);
    }
    (void)t_last;
    if (rc == 0 && t_start < t_end)
      for (size_t q = 0; q < start.size() && rc == 0; q++)
        if (rk_set_option(eng, knobs[q].name, start[q]) != 0) rc = -1;
  }
  eng->tuning = false;
  if (steps_done) *steps_done = done;
  return rc;
}

int rk_get_option(rk_engine* eng, const char* name, double* value) {
  if (!eng || !name || !value) return -1;
  const std::string n = name;
  if (int* p = tuned_option(eng, n))
    *value = *p;
  else if (n == "record_geometry")
    *value = eng->record_geometry;
  else if (n == "profile")
    *value = eng->profile;
  else if (n == "exact_sums")
    *value = eng->exact_sums;
  else if (n == "skip_far")
    *value = eng->skip_far;
  else if (n == "count_skips")
    *value = eng->skip_count != nullptr;
  else if (n == "skip_segment")
    *value = eng->skip_seg;
  else if (n == "skipped_entries" || n == "zero_history_entries" || n == "entries_looked_at") {
    unsigned long long c[3] = {0ull, 0ull, 0ull};
    if (eng->skip_count) {
      if (flush_steps(eng) != 0) return -1;
      HIPCHK(hipMemcpyAsync(c, eng->skip_count, sizeof c, hipMemcpyDeviceToHost, eng->stream));
      HIPCHK(hipStreamSynchronize(eng->stream));
    }
    *value = (double)c[n == "skipped_entries" ? 0 : n == "zero_history_entries" ? 1 : 2];
  }
  else if (n == "heavy_min")
    *value = eng->w.heavy_min;
  else if (n == "guard_batch")
    *value = eng->guard_batch;
  else if (n == "model_scale")
    *value = eng->model_scale;
  else if (n == "gravity_x")
    *value = eng->m.gravity.x;
  else if (n == "gravity_y")
    *value = eng->m.gravity.y;
  else if (n == "gravity_z")
    *value = eng->m.gravity.z;
  else if (n == "dt")  // the deck's scalars, read-only
    *value = eng->m.dt;
  else if (n == "interVerlet")
    *value = eng->m.interVerlet;
  else if (n == "interConf")
    *value = eng->m.interConf;
  else if (n == "DVerlet")
    *value = eng->m.DVerlet;
  else if (n == "dVerlet")
    *value = eng->m.dVerlet;
  else if (n == "tmax")
    *value = eng->m.tmax;
  else
    return fail(eng, "unknown option " + n);
  return 0;
}

int rk_integrate(rk_engine* eng, const char* outdir, int64_t max_steps, int64_t* steps_done) {
  if (check_loaded(eng)) return -1;
  (void)hipSetDevice(eng->device);
  if (integrate_prepare(eng, true) != 0) return -1;
  if (outdir) {
    char fn[4096];
    snprintf(fn, sizeof fn, "%s/conf%d", outdir, eng->m.iconf);
    if (save_conf_now(eng, fn) != 0) return -1;
    // the output files of initOutputFiles (Rockable.cpp:215-222), in the output directory
    close_outputs(eng);
    const char* names[3] = {"perf.txt", "kineticEnergy.txt", "staticBalance.txt"};
    FILE** fs[3] = {&eng->perf_file, &eng->ke_file, &eng->sb_file};
    for (int k = 0; k < 3; k++) {
      snprintf(fn, sizeof fn, "%s/%s", outdir, names[k]);
      *fs[k] = fopen(fn, "w");
      if (!*fs[k]) return fail(eng, std::string("cannot write ") + fn);
    }
    eng->conf_clock = std::chrono::steady_clock::now();
    eng->nl_since_conf_s = eng->force_since_conf_s = 0.0;
  }
  eng->ext_step = (int64_t)(eng->m.t / eng->m.dt);  // Rockable.cpp:2578
  int64_t n = 0;
  if (eng->guard) {
    if (run_guarded(eng, outdir, max_steps < 0 ? INT64_MAX : max_steps, true, &n) != 0) return -1;
  } else {
    while (eng->m.t <= eng->m.tmax && (max_steps < 0 || n < max_steps)) {
      if (loop_body(eng, outdir) != 0) return -1;
      n++;
    }
    if (flush_steps(eng) != 0) return -1;
  }
  if (steps_done) *steps_done = n;
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

int rk_save_conf(rk_engine* eng, const char* path) {
  if (check_loaded(eng)) return -1;
  if (!path) return fail(eng, "no path given");
  return save_conf_now(eng, path);
}

int rk_synchronize(rk_engine* eng) {
  if (!eng) return -1;
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

int rk_get_stats(rk_engine* eng, rk_stats* out) {
  if (check_loaded(eng) || !out) return -1;
  (void)hipSetDevice(eng->device);
  const size_t need = (size_t)stat_parts(eng->d);
  if (eng->stat_cap < need) {
    if (eng->stat_part) (void)hipFree(eng->stat_part);
    eng->stat_part = nullptr;
    HIPCHK(hipMalloc(&eng->stat_part, need * sizeof(double)));
    eng->stat_cap = need;
  }
  if (!eng->stat_out) HIPCHK(hipMalloc(&eng->stat_out, 16 * sizeof(double)));
  launch_stats(eng->d, eng->stat_part, eng->stat_out, eng->stream);
  HIPCHK(hipGetLastError());
  double v[9];
  HIPCHK(hipMemcpyAsync(v, eng->stat_out, sizeof v, hipMemcpyDeviceToHost, eng->stream));
  HIPCHK(hipStreamSynchronize(eng->stream));
  out->kinetic_translational = v[0];
  out->kinetic_rotational = v[1];
  out->vmax = v[2];
  out->nfree = (int64_t)v[3];
  out->nactive = (int64_t)v[4];
  out->ncontacts = (int64_t)v[5];
  out->active_pairs = (int64_t)v[6];
  out->coordination = v[3] > 0.0 ? 2.0 * v[6] / v[3] : 0.0;
  out->dn_mean = v[5] > 0.0 ? v[7] / v[5] : 0.0;
  out->dn_min = v[5] > 0.0 ? v[8] : 0.0;
  out->ninterfaces = eng->live_interfaces;
  out->broken_interfaces = eng->loaded_interfaces - eng->live_interfaces;
  out->total_broken_area = eng->m.totalBrokenArea;
  return 0;
}

int rk_get_stats_noalloc(rk_engine* eng, rk_stats* out) {
  if (check_loaded(eng) || !out) return -1;
  HIPCHK(hipSetDevice(eng->device));
  const long need = (long)stat_parts(eng->d) + 16;
  if (need > kStatScratchDoubles)
    return fail(eng, "rk_get_stats_noalloc: the packing needs " + std::to_string(need) +
                         " doubles of scratch, more than the fixed " + std::to_string(kStatScratchDoubles) +
                         " (use rk_get_stats)");
  double* base = nullptr;
  char msg[256];
  if (stat_scratch(&base, msg, sizeof msg) != 0) return fail(eng, msg);
  double* part = base;
  double* res = base + (kStatScratchDoubles - 16);
  launch_stats(eng->d, part, res, eng->stream);
  HIPCHK(hipGetLastError());
  double v[9];
  HIPCHK(hipMemcpyAsync(v, res, sizeof v, hipMemcpyDeviceToHost, eng->stream));
  HIPCHK(hipStreamSynchronize(eng->stream));
  out->kinetic_translational = v[0];
  out->kinetic_rotational = v[1];
  out->vmax = v[2];
  out->nfree = (int64_t)v[3];
  out->nactive = (int64_t)v[4];
  out->ncontacts = (int64_t)v[5];
  out->active_pairs = (int64_t)v[6];
  out->coordination = v[3] > 0.0 ? 2.0 * v[6] / v[3] : 0.0;
  out->dn_mean = v[5] > 0.0 ? v[7] / v[5] : 0.0;
  out->dn_min = v[5] > 0.0 ? v[8] : 0.0;
  out->ninterfaces = eng->live_interfaces;
  out->broken_interfaces = eng->loaded_interfaces - eng->live_interfaces;
  out->total_broken_area = eng->m.totalBrokenArea;
  return 0;
}

int rk_get_counts(rk_engine* eng, rk_counts* out) {
  if (check_loaded(eng) || !out) return -1;
  const EntryArrays& E = eng->E[eng->cur];
  out->nbodies = (int64_t)eng->m.bodies.size();
  out->ndriven = eng->m.nDriven;
  out->nlisted = E.n;
  int64_t na = 0;
  if (E.n > 0) {
    std::vector<uint8_t> act(E.n);
    HIPCHK(hipStreamSynchronize(eng->stream));
    HIPCHK(hipMemcpy(act.data(), E.eact, E.n, hipMemcpyDeviceToHost));
    for (uint8_t a : act) na += a & 1;
  }
  out->nactive = na;
  out->ninterfaces = eng->live_interfaces;
  out->nl_updates = eng->nl_updates;
  out->steps = eng->steps;
  out->t = eng->m.t;
  out->dt = eng->m.dt;
  out->total_broken_area = eng->m.totalBrokenArea;
  out->last_candidates = eng->w.last_candidates;
  out->guard_reads = eng->guard_reads;
  out->guard_wasted = eng->guard_wasted;
  out->graph_launches = eng->graph_launches;
  out->graph_captures = eng->graph_captures;
  out->guard_polls = eng->guard_polls;
  return 0;
}

int rk_get_bodies(rk_engine* eng, double* pos, double* Q, double* vel, double* vrot, double* acc, double* arot,
                  double* force, double* moment, size_t nbodies) {
  if (check_loaded(eng)) return -1;
  if (nbodies < eng->m.bodies.size()) return fail(eng, "output arrays too small");
  if (download_bodies(eng) != 0) return -1;
  for (size_t b = 0; b < eng->m.bodies.size(); b++) {
    const Body& p = eng->m.bodies[b];
    auto put = [&](double* a, v3 v) {
      if (a) {
        a[3 * b] = v.x;
        a[3 * b + 1] = v.y;
        a[3 * b + 2] = v.z;
      }
    };
    put(pos, p.pos);
    put(vel, p.vel);
    put(vrot, p.vrot);
    put(acc, p.acc);
    put(arot, p.arot);
    put(force, p.force);
    put(moment, p.moment);
    if (Q) {
      Q[4 * b] = p.Q.w;
      Q[4 * b + 1] = p.Q.x;
      Q[4 * b + 2] = p.Q.y;
      Q[4 * b + 3] = p.Q.z;
    }
  }
  return 0;
}

int rk_set_bodies(rk_engine* eng, const double* pos, const double* Q, const double* vel, const double* vrot,
                  size_t nbodies) {
  if (check_loaded(eng)) return -1;
  const int nb = (int)eng->m.bodies.size();
  if (nbodies < (size_t)nb) return fail(eng, "input arrays too small");
  std::vector<BodyKin> kin(nb);
  HIPCHK(hipStreamSynchronize(eng->stream));
  HIPCHK(hipMemcpy(kin.data(), eng->kin, sizeof(BodyKin) * nb, hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; b++) {
    BodyKin& k = kin[b];
    if (pos) {
      k.px = pos[3 * b];
      k.py = pos[3 * b + 1];
      k.pz = pos[3 * b + 2];
    }
    if (Q) {
      k.qw = Q[4 * b];
      k.qx = Q[4 * b + 1];
      k.qy = Q[4 * b + 2];
      k.qz = Q[4 * b + 3];
    }
    if (vel) {
      k.vx = vel[3 * b];
      k.vy = vel[3 * b + 1];
      k.vz = vel[3 * b + 2];
    }
    if (vrot) {
      k.wx = vrot[3 * b];
      k.wy = vrot[3 * b + 1];
      k.wz = vrot[3 * b + 2];
    }
  }
  HIPCHK(hipMemcpy(eng->kin, kin.data(), sizeof(BodyKin) * nb, hipMemcpyHostToDevice));
  eng->gv_state = rk_engine::GV_STALE;
  launch_reset_keys(eng->d, eng->stream);  // poses set from outside: no motion bound holds
  HIPCHK(hipGetLastError());
  return 0;
}

int rk_get_interactions(rk_engine* eng, int active_only, rk_interaction* out, size_t cap, size_t* n) {
  if (check_loaded(eng)) return -1;
  std::vector<Inter> v;
  if (download_interactions(eng, active_only, v) != 0) return -1;
  if (n) *n = v.size();
  if (out) {
    const size_t k = std::min(cap, v.size());
    for (size_t q = 0; q < k; q++) {
      const Inter& I = v[q];
      rk_interaction& o = out[q];
      o.i = I.i;
      o.j = I.j;
      o.type = I.type;
      o.isub = I.isub;
      o.jsub = I.jsub;
      o.stick = I.stick;
      o.n[0] = I.n.x;
      o.n[1] = I.n.y;
      o.n[2] = I.n.z;
      o.dn = I.dn;
      o.pos[0] = I.pos.x;
      o.pos[1] = I.pos.y;
      o.pos[2] = I.pos.z;
      o.vel[0] = I.vel.x;
      o.vel[1] = I.vel.y;
      o.vel[2] = I.vel.z;
      o.fn = I.fn;
      o.ft[0] = I.ft.x;
      o.ft[1] = I.ft.y;
      o.ft[2] = I.ft.z;
      o.mom[0] = I.mom.x;
      o.mom[1] = I.mom.y;
      o.mom[2] = I.mom.z;
      o.damp = I.damp;
    }
  }
  return 0;
}

int rk_get_kernel_time(rk_engine* eng, const char* kernel, int64_t* launches, double* total_ms) {
  if (!eng || !kernel) return -1;
  resolve_events(eng);
  auto it = eng->timers.find(kernel);
  if (launches) *launches = it == eng->timers.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == eng->timers.end() ? 0.0 : it->second.ms;
  return 0;
}

int rk_reset_kernel_time(rk_engine* eng) {
  if (!eng) return -1;
  resolve_events(eng);
  eng->timers.clear();
  return 0;
}

int rk_set_rebuild_model(rk_engine* eng, rk_mlp* mlp) {
  if (!eng) return -1;
  eng->model = mlp;
  eng->pred_batch = eng->guard_batch;
  eng->model_trust = 1.0;
  return 0;
}

int rk_guard_features(rk_engine* eng, float* out16) {
  if (check_loaded(eng) || !out16) return -1;
  (void)hipSetDevice(eng->device);
  ensure_gv(eng, true);
  if (compute_features(eng) != 0) return -1;
  HIPCHK(hipMemcpyAsync(out16, eng->feat, 16 * sizeof(float), hipMemcpyDeviceToHost, eng->stream));
  HIPCHK(hipStreamSynchronize(eng->stream));
  return 0;
}

}  // extern "C"
