// rk_device.hpp -- device-resident state of the DEM engine and the kernel
// entry points.
//
// Layout in HBM (all FP64 unless stated):
//   bodies   BodyKin[N]   128 B record read by every gather of the step
//                         (pos, Q, vel, vrot, Minkowski radius, homothety,
//                          shape id, group id)
//            gv[3 NV]     global vertex positions, body b's from bvoff[b],
//                         recomputed once per step
//            acc/arot/force/moment  3 doubles per body (xyz interleaved)
//            mass[N], inertia[3N], cluster[N] (int)
//   shapes   vertex table (3 doubles / vertex), edge pairs (2 int), face CSR
//   entries  one per listed interaction, sorted by (i, j, type, isub, jsub):
//            ei, ej (int), esub (uint32: type<<30 | isub<<15 | jsub)
//            edn     normal distance dn                8 B, every step
//            EState  {fn, ft.xyz, mom.xyz}           56 B history record
//            EGeom   {n.xyz, pos.xyz, vel.xyz}        72 B geometry record
//            damp, stick (int, interface index or -1), act (uint8: bit 0
//            active at the last force phase, bit 1 history maybe non-zero)
//            spos[2n]     the entry's two resultant slots (i side, j side)
//   CSR      istart[N+1]  entries owned by body b are [istart[b], istart[b+1])
//            jstart[N+1], jlist[]  entries whose j == b, ascending index
//   slots    body b's incident entries in the serial order of the reference's
//            resultant loop (entries with j == b below b's own range, b's own
//            range, entries with j == b above it) are the slots
//            [soff[b], soff[b+1]); per slot an activity byte sact and the
//            six signed terms sval {F.xyz, M.xyz} (48 B) that the entry adds
//            to that body: i side (f, mi), j side (-f, mj)
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rk_math.hpp"

namespace rk {

// A word pair in host-mapped memory that the displacement check writes for
// the host to read without a copy or an event: done, the steps the device
// has completed; frozen, set when it freezes the later steps.
struct HostNote {
  int32_t frozen;
  int32_t pad;
  long long done;
};

struct alignas(16) BodyKin {
  double px, py, pz;
  double qw, qx, qy, qz;
  double vx, vy, vz;
  double wx, wy, wz;
  double R;  // homothety * shape radius (Particle::MinkowskiRadius)
  double h;  // homothety
  int32_t shape;
  int32_t group;
};
static_assert(sizeof(BodyKin) == 128, "BodyKin must stay one 128-byte record");

// Force-law history of an entry (fn, ft, mom); its normal distance dn is
// kept apart (edn) because every step rewrites dn of every listed entry while
// only the active ones touch this record.
struct EState {
  double fn;
  double ftx, fty, ftz;
  double momx, momy, momz;
};
static_assert(sizeof(EState) == 56, "EState is 56 bytes");

struct EGeom {
  double nx, ny, nz;
  double px, py, pz;
  double vx, vy, vz;
};

// An active entry handed from the geometry kernel to the force-law kernel.
struct alignas(16) ActiveKin {
  double nx, ny, nz;
  double px, py, pz;
  double dn, prev_dn;
  int32_t e, deact;
  int32_t i, j;  // the entry's bodies (saves the force law a dependent load)
};
static_assert(sizeof(ActiveKin) == 80, "ActiveKin is 80 bytes");

// dynamicCheckUpdateNL on the device (Rockable.cpp:1571-1596).  The running
// maxima are kept as order-preserving integer keys so that atomicMax / atomicMin
// (integer, order-independent) give the serial loop's result exactly:
// maxDeltaPos = sqrt(max |dx|^2), and maxDeltaRot = 2 acos(min w) is compared
// through w <= wstar, wstar being the largest w with 2 acos(w) >= angleUpdateNL.
struct DevGuard {
  unsigned long long maxpos2;  // bits of max |dx|^2 (a non-negative double)
  unsigned long long minw;     // order key of min w of the accumulated dQ
  int reset;                   // the accumulations restart at the next step
  int nu;                      // needUpdate set by the last check
  int frozen;                  // a rebuild is due: later step kernels do nothing
  int pad;
  long long frozen_step;       // the step after which the rebuild is due
  long long step;              // index of the step the next check ends
};

struct DevControl {
  int32_t type, i;
  double value;
  double vx, vy, vz;
};

struct DevShapes {
  const int32_t* nv;
  const int32_t* ne;
  const int32_t* nf;
  const int32_t* voff;    // first vertex of the shape in vtx
  const int32_t* eoff;    // first edge of the shape in edge
  const int32_t* foff;    // first face of the shape in fstart
  const double* vtx;      // 3 per vertex
  const int32_t* edge;    // 2 per edge (local vertex indices)
  const int32_t* fstart;  // per global face: start in fidx; one trailing entry
  const int32_t* fidx;    // local vertex indices
  const double* radius;
  const double* obb;      // 15 per shape: c, e0, e1, e2, ext
  // 2 per shape, at homothety 1: the longest edge and the largest distance
  // of a face's vertex from the plane of its first two and last vertices
  // (+inf for a degenerate face); they bound how far the narrow phase's
  // near tests reach beyond the vertices' convex hull
  const double* reach;
};

// Group-pair parameter table: tab[(id * G + g1) * G + g2].
struct DevParams {
  const double* tab;
  const double* damp_rate;  // per group pair: Rockable.cpp:1368-1370 rate
  const int32_t* damp_kind; // 0 rate form, 1 en2 <= 0, 2 zero
  int G;
};

struct DevInterfaces {
  int n;
  const uint64_t* key;  // (i << 32) | j, sorted
  const int32_t* istart;  // the interfaces whose first body is i: [istart[i], istart[i + 1])
  const double* par;    // 8 per interface: kn kt kr fn0 ft0 mom0 power dn0
  const int32_t* isInner;
  int32_t* alive;
  uint8_t* brk;         // set by the force law when the yield criterion is met (cleared by k_breakage)
  int32_t* broken;      // the interfaces k_breakage broke since the host last accounted them
  int32_t* nbroken;     // their count
  const int32_t* bond_off;  // bonds of interface k: bond_key[bond_off[k] .. bond_off[k+1])
  const int32_t* bond_key;  // 3 per bond: type isub jsub
};

struct Dev {
  int nb, nDriven;
  BodyKin* kin;
  // global vertices: body b's are gv[3 * bvoff[b] ...], bvbody[k] owns slot k
  double* gv;
  const int32_t* bvoff;
  const int32_t* beoff;   // body b's edges (as sub-elements) from beoff[b], its faces from bfoff[b]
  const int32_t* bfoff;
  int netot, nftot;       // edges and faces of all bodies
  const int32_t* bvbody;
  int nvtot;
  int nvdriven;           // vertices of the driven bodies (the first ones)
  const int32_t* wvstart; // free bodies' vertices packed by whole bodies into wavefronts of 64: [wvstart[w], wvstart[w+1])
  int nwv;                // such wavefronts (0: a free body has more than 64 vertices)
  double* acc;
  double* arot;
  double* force;
  double* moment;
  const double* mass;
  const double* inertia;
  const int32_t* cluster;
  DevShapes sh;
  DevParams par;
  // entries
  int ne;
  int32_t* ei;
  int32_t* ej;
  uint32_t* esub;
  int4* edesc;            // per entry: resolved vertex indices of its geometry (k_entry_desc)
  EState* est;
  double* edn;
  EGeom* egeo;
  double* edamp;
  int32_t* estick;
  uint8_t* eact;
  // Far-entry skip.  cbound[b] accumulates a bound on how far any vertex of
  // body b has moved since the list was built (|dx| + 2 |dQ| rmax[b] per step,
  // rmax the largest vertex distance from the body's centre; +inf for a
  // driven body with a control).  An inactive entry with zero history keeps
  // ekey = dn + cbound[i] + cbound[j] of the step it was last evaluated: its
  // gap cannot have fallen by more than the growth of cbound[i] + cbound[j]
  // since, so while ekey - cbound[i] - cbound[j] > kSkipMargin it is still
  // positive and the entry still inactive, and the step leaves it alone.
  double* ekey;
  double* cbound;
  const double* rmax;
  int skip_far;           // 1: skip such entries (only when the step writes no geometry of inactive entries)
  int skip_seg;           // entries a workgroup tests and compacts before evaluating (512 ... 4096; else none)
  // with the count_skips option: [0] entries skipped, [1] inactive entries
  // with zero history, [2] entries looked at, summed over the steps
  unsigned long long* skip_count;
  const int32_t* istart;
  const int32_t* jstart;
  const int32_t* jlist;
  // resultant slots
  const int32_t* spos;
  const int32_t* soff;
  uint8_t* sact;
  double* sval;
  const int32_t* heavy;   // bodies with more than heavy_min slots
  int nheavy, heavy_min;
  // passive heavy bodies: their slots in chunks of 64 (chunk c: slots
  // [hch[c].x, hch[c].y)), heavy body h's chunks [hch_off[h], hch_off[h+1]),
  // six partial sums per chunk in hpart
  const int2* hch;
  const int32_t* hch_off;
  double* hpart;
  int nhch;
  int nheavy_exact;       // heavy bodies summed in serial order (k_gather_heavy)
  // per body: 1 if its force and moment feed nothing back into the motion
  // (a driven body without force or moment controls)
  const uint8_t* passive;
  int exact_sums;         // 1: serial sums for passive bodies too
  // launch shapes (any choice gives the same results)
  int law_group;          // geometry blocks per k_law wavefront (1..8)
  int geom_block;         // k_geom workgroup size (64, 128, 256, 512)
  int gather_waves;       // k_gather wavefronts per workgroup (1, 2, 4)
  int gather_lanes;       // k_gather lanes per body (8, 16)
  int gather_groups;      // k_gather slot groups per body wavefront (1 .. gather_lanes)
  int fuse_glob;          // the free bodies' global vertices: 0 k_glob, 1 k_predict per body, 2 k_predict_vtx, 3 k_predict_blk
  int body_block;         // workgroup size of the per-body kernels (64, 128, 256)
  int fuse_law;           // 1: geometry and law in one kernel (k_interact)
  int type_sort;          // 1: k_interact takes a block's entries by contact type
  int persist;            // step kernels' grid cap (0: one tile per workgroup)
  // device displacement guard (dynamicUpdateNL)
  int guard_on;
  DevGuard* guard;
  HostNote* note;  // host-mapped (may be null): the check's progress and freeze, for the host
  double* gdpos;          // 3 per body
  double* gdq;            // 4 per body
  unsigned long long* gpart;  // per gather workgroup: max |dx|^2 key, min w key
  double disp, wstar;
  int disp_check;         // the guard runs the displacement check (dynamicUpdateNL); without it, it
                          // only turns a bond breakage (flags[0]) into a rebuild due after the step
  // active entries of the current force phase: geometry block b's in
  // akin[b * geom_block, b * geom_block + acount[b])
  ActiveKin* akin;
  int* acount;
  // driving
  int nctrl;
  DevControl* ctrl;
  const int32_t* ctrl_off;   // per driven body, [ctrl_off[b], ctrl_off[b+1]) into ctrl_idx
  int ctrl_parallel;         // every control acts on a driven body: one thread per driven body
  const int32_t* ctrl_idx;
  // ServoFunction: 0 none, 1 tritriIsostaticCompression, 2 tritriBiaxialCompression,
  // 3 tritriCustom, 4 tritriLodeAngle (the walls' gaps, on the device);
  // 5-8 shakers and ramp (functions of t only: the host evaluates them and
  // passes the values in StepParams.servo_v)
  int servo, servo_icontr, servo_ids[6];
  double servo_pressure;
  double servo_val[6];        // tritriCustom values (xmin xmax ymin ymax zmin zmax)
  double lode_a, lode_b, lode_rate;
  // body force
  int bf;
  double bf_point[3], bf_acc, bf_axisBody[3], bf_axis[3], bf_Kr, bf_fluidDensity;
  const double* bf_surface;  // per body pow(volume h^3, 2/3) for ViscousFluid
  DevInterfaces ifc;
  int paramsInInterfaces;
  int32_t* flags;  // [0] needUpdate
};

// Margin of the far-entry skip (m): covers the rounding of the gaps and of
// the motion bounds, far below any gap of interest.
constexpr double kSkipMargin = 1e-9;

struct StepParams {
  double dt, dt_2, dt2_2;
  double gx, gy, gz;
  int law;
  double t;           // box time of the step (after its increment)
  double servo_v[3];  // values of a time-only servo's controls at t
};

RK_HD uint32_t pack_sub(int type, int isub, int jsub) {
  return ((uint32_t)type << 30) | ((uint32_t)isub << 15) | (uint32_t)jsub;
}

// Workgroups are handed to the 8 XCDs round-robin in dispatch order (XCD =
// block mod 8), and each XCD has its own L2.  xcd_block maps the blocks one
// XCD runs onto one contiguous range of [0, n) (a bijection), so that the
// bodies a block's work reads are mostly those its XCD's L2 already holds
// (entries and candidate pairs are sorted by body, and neighbouring indices
// are near in space).
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int x = b & 7, k = b >> 3;
  const int q = n >> 3, r = n & 7;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Persistent launches: a grid of G workgroups covers T tiles.  The workgroups
// on XCD x (block mod 8) take XCD x's contiguous share of [0, T) -- the tile
// xcd_block gives them when G == T -- each every (workgroups on x)-th tile of
// it, so the locality is the same whatever G.  G is T, or a multiple of 8
// below it (persistent_grid).
struct TileLoop {
  int t, end, step;
};
__device__ __forceinline__ TileLoop xcd_tiles(int T) {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b & 7, k = b >> 3;
  const int q = T >> 3, r = T & 7;
  const int start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  TileLoop L;
  L.t = start + k;
  L.end = start + q + (x < r ? 1 : 0);
  L.step = (G - x + 7) >> 3;
  return L;
}
inline int persistent_grid(int T, int persist) {
  if (persist <= 0 || T <= persist) return T;
  return std::max(8, persist & ~7);
}

// rk_step.hip
// the entries' resolved vertex indices, after every change of the list
void launch_entry_desc(const Dev& d, hipStream_t s);
// the predictor; then the global vertices of the new poses: of the listed
// bodies (glob), else of every body
void launch_predict(const Dev& d, const StepParams& sp, int glob, hipStream_t s);
// all: every body's vertices; else those of the bodies with listed slots
void launch_glob(const Dev& d, int all, hipStream_t s);
void launch_interact(const Dev& d, const StepParams& sp, int write_geom_all, hipStream_t s);
// The gap of every inactive entry with zero history from the current poses,
// as the last step computed it (an entry the step skipped keeps an older gap
// until then); before the list is read out.
void launch_refresh_gaps(const Dev& d, hipStream_t s);
// Every entry evaluated at the next step (after the poses were set from outside).
void launch_reset_keys(const Dev& d, hipStream_t s);
// kick: the closing half kick; guard: also the displacement check's
// accumulation (k_guard_check then follows)
void launch_gather(const Dev& d, const StepParams& sp, int kick, int guard, hipStream_t s);
// capacity of gpart for any launch shape
int gather_blocks_max(int nb, int nhch);
void launch_breakage(const Dev& d, const StepParams& sp, hipStream_t s);
// statistics: part holds stat_parts(d) doubles, out 9 doubles (KE translation,
// KE rotation, max |v|, free bodies, active, contacts, active pairs, sum dn,
// min dn over contacts)
int stat_parts(const Dev& d);
void launch_stats(const Dev& d, double* part, double* out, hipStream_t s);
// The statistics' fixed scratch (no allocation): its capacity in doubles
// (partials and the 9 outputs), and its device address on the current
// device, looked up once per device (0, or -1 with err set).
constexpr int kStatScratchDoubles = 1 << 20;
int stat_scratch(double** ptr, char* err, size_t errlen);
void launch_guard(const Dev& d, hipStream_t s);
// the step index the next guard check ends (the checks count on from it)
void launch_guard_set_step(const Dev& d, long long step, hipStream_t s);
// part: (nb / 256 + 1) * 9 doubles; feat: 16 floats
void launch_guard_features(const Dev& d, const StepParams& sp, double disp, double angle, long long since,
                           long long last, double* part, float* feat, hipStream_t s);

}  // namespace rk

// rk_mlp.hip: y[0] = network(x[0..15]) for a feature vector on the device
struct rk_mlp;
int rk_mlp_forward_device(rk_mlp* mlp, const float* x_dev, float* y_dev, hipStream_t s);
