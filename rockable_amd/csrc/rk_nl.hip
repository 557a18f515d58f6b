// rk_nl.hip -- neighbour-list rebuild on the GPU.
//
// Restates UpdateNL_bruteForce (Rockable.cpp:1614-1662), UpdateNL_linkCells
// (1667-1809) and AddOrRemoveInteractions_bruteForce (1309-1432) with the
// sub-element predicates of Particle.cpp:88-209, keeping the reference's
// set semantics exactly:
//   * a key that is tested and near is kept (with its history) or created,
//   * a key that is tested and not near is erased,
//   * a key that is not tested (pair not a candidate, vertex gated out, or a
//     type never tested in that direction) is retained unchanged.
//
// Pipeline (each step a kernel; counts come back to the host between steps):
//   k_obb            Particle::updateObb (+ AABB for the cell grid)
//   k_broad_*        candidate body pairs (i < j) whose enlarged OBBs overlap
//   groups           directed (owner, other) groups = candidates x 2 directions
//                    merged with the groups of the current list (radix sort)
//   k_narrow         one wavefront per candidate pair: global vertices and
//                    vertex gates in LDS, every predicate of both directions
//                    evaluated 64 at a time, near bits kept as ballot masks,
//                    then written as keys in (type, isub, jsub) order with a
//                    wavefront prefix scan (no atomics per key)
//   k_merge          per directed group, merge of the sorted old entries and
//                    the sorted near keys (count pass, scan, write pass of the
//                    keys and of each entry's source); the output is again
//                    sorted by (i, j, type, isub, jsub)
//   k_merge_fill     one thread per new entry: its state record, copied from
//                    the old entry it keeps or started from zero
//   CSR              istart by owner; jlist = entries sorted (stably) by j
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <utility>
#include <vector>

#include "rk_device.hpp"
#include "rk_model.hpp"
#include "rk_nl.hpp"

namespace rk {

namespace {

constexpr int kBlock = 256;

#define RK_CHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      snprintf(w.err, sizeof(w.err), "%s failed: %s", #x, hipGetErrorString(e_)); \
      return -1;                                                                    \
    }                                                                               \
  } while (0)

__device__ __forceinline__ v3 kpos(const BodyKin& b) { return V(b.px, b.py, b.pz); }
__device__ __forceinline__ qt kq(const BodyKin& b) {
  qt q;
  q.w = b.qw;
  q.x = b.qx;
  q.y = b.qy;
  q.z = b.qz;
  return q;
}
__device__ __forceinline__ obb_t load_obb(const double* o) {
  obb_t r;
  r.c = V(o[0], o[1], o[2]);
  r.e[0] = V(o[3], o[4], o[5]);
  r.e[1] = V(o[6], o[7], o[8]);
  r.e[2] = V(o[9], o[10], o[11]);
  r.ext = V(o[12], o[13], o[14]);
  return r;
}
__device__ __forceinline__ void store_obb(double* o, const obb_t& r) {
  o[0] = r.c.x;
  o[1] = r.c.y;
  o[2] = r.c.z;
  o[3] = r.e[0].x;
  o[4] = r.e[0].y;
  o[5] = r.e[0].z;
  o[6] = r.e[1].x;
  o[7] = r.e[1].y;
  o[8] = r.e[1].z;
  o[9] = r.e[2].x;
  o[10] = r.e[2].y;
  o[11] = r.e[2].z;
  o[12] = r.ext.x;
  o[13] = r.ext.y;
  o[14] = r.ext.z;
}

// Particle::updateObb (Particle.cpp:78-86) and the body AABB of
// Rockable::computeAABB (Rockable.cpp:3934-3941).
__global__ void k_obb(Dev d, double* bobb, double* baabb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.nb) return;
  const BodyKin P = d.kin[b];
  obb_t o = load_obb(d.sh.obb + 15 * P.shape);
  const qt Q = kq(P);
  o.e[0] = qrot(Q, o.e[0]);
  o.e[1] = qrot(Q, o.e[1]);
  o.e[2] = qrot(Q, o.e[2]);
  o.c = qrot(Q, o.c);
  o.ext = vsc(P.h, o.ext);
  o.c = vsc(P.h, o.c);
  o.c = vadd(o.c, kpos(P));
  store_obb(bobb + 15 * (size_t)b, o);
  if (baabb) {
    const double* gb = d.gv + 3 * (size_t)d.bvoff[b];
    v3 mn = V(gb[0], gb[1], gb[2]), mx = mn;
    const int nv = d.sh.nv[P.shape];
    for (int v = 1; v < nv; v++) {
      v3 g = V(gb[3 * v], gb[3 * v + 1], gb[3 * v + 2]);
      mn.x = g.x < mn.x ? g.x : mn.x;
      mn.y = g.y < mn.y ? g.y : mn.y;
      mn.z = g.z < mn.z ? g.z : mn.z;
      mx.x = g.x > mx.x ? g.x : mx.x;
      mx.y = g.y > mx.y ? g.y : mx.y;
      mx.z = g.z > mx.z ? g.z : mx.z;
    }
    const double r = P.R;
    double* a = baabb + 6 * (size_t)b;
    a[0] = mn.x - r;
    a[1] = mn.y - r;
    a[2] = mn.z - r;
    a[3] = mx.x + r;
    a[4] = mx.y + r;
    a[5] = mx.z + r;
  }
}

// Interfaces[i].find((i, j)) on the sorted live-interface keys.
// Whether (i, j) has a live interface: the first key not below (i, j) among
// i's interfaces (their keys are sorted by j), as a binary search over all
// keys would find it, without its chain of dependent loads.
// IfcRange holds i's range, loaded once for all of i's pairs.
struct IfcRange {
  int lo, hi;
};
__device__ __forceinline__ IfcRange ifc_range(const DevInterfaces& f, int i) {
  IfcRange r;
  r.lo = f.n ? f.istart[i] : 0;
  r.hi = f.n ? f.istart[i + 1] : 0;
  return r;
}
__device__ __forceinline__ bool has_live_interface(const DevInterfaces& f, IfcRange r, int j) {
  const uint32_t uj = (uint32_t)j;
  for (int k = r.lo; k < r.hi; k++) {
    const uint32_t kj = (uint32_t)(f.key[k] & 0xffffffffull);
    if (kj >= uj) return kj == uj && f.alive[k];
  }
  return false;
}

// Output counters are striped: stripe s owns counter cnt[s * kStripeStride]
// (its own 128-byte line) and the region [s * rcap, (s + 1) * rcap) of the
// output; a workgroup appends to stripe (its index mod kStripes), one atomic
// per wavefront and emission point.  A single shared counter serialises at
// one L2 channel (~10 ns per atomic, milliseconds per rebuild).
constexpr int kStripes = 256;
constexpr int kStripeStride = 32;

__device__ __forceinline__ int stripe_of_block() {
  return (int)((blockIdx.x + blockIdx.y * gridDim.x) & (kStripes - 1));
}

// every active lane of the wavefront calls this; lanes with emit append i, j
__device__ __forceinline__ void emit_pair(bool emit, uint64_t* cand, int* cnt, int rcap, int i, int j) {
  const unsigned long long m = __ballot(emit);
  if (m == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  const int s = stripe_of_block();
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt + s * kStripeStride, __popcll(m));
  base = __shfl(base, leader, 64);
  if (emit) {
    const int k = base + __popcll(m & ((1ull << lane) - 1ull));
    if (k < rcap) cand[(size_t)s * rcap + k] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
  }
}

// one workgroup of kStripes threads: exclusive offsets of the stripes, their
// total and their maximum
__global__ void k_stripe_scan(const int* cnt, int* off, int* total_max) {
  __shared__ int sc[kStripes], mx[kStripes];
  const int t = threadIdx.x;
  const int v = cnt[t * kStripeStride];
  sc[t] = v;
  mx[t] = v;
  __syncthreads();
  for (int o = 1; o < kStripes; o <<= 1) {
    const int x = t >= o ? sc[t - o] : 0;
    __syncthreads();
    sc[t] += x;
    __syncthreads();
  }
  for (int h = kStripes / 2; h > 0; h >>= 1) {
    if (t < h) mx[t] = max(mx[t], mx[t + h]);
    __syncthreads();
  }
  off[t] = sc[t] - v;
  if (t == kStripes - 1) total_max[0] = sc[t];
  if (t == 0) total_max[1] = mx[0];
}

// the stripes, concatenated
__global__ void k_stripe_gather(const uint64_t* src, int rcap, const int* cnt, const int* off, uint64_t* dst) {
  const int s = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = min(cnt[s * kStripeStride], rcap);
  if (k < n) dst[off[s] + k] = src[(size_t)s * rcap + k];
}

__device__ __forceinline__ bool pair_overlap(const double* bobb, int i, int j, double half) {
  obb_t oi = load_obb(bobb + 15 * (size_t)i);
  obb_enlarge(&oi, half);
  obb_t oj = load_obb(bobb + 15 * (size_t)j);
  obb_enlarge(&oj, half);
  return obb_intersect(oi, oj);
}

// UpdateNL_bruteForce pair loop (Rockable.cpp:1622-1658): i < j, j >= nDriven
constexpr int kBruteChunk = 128;
__global__ void k_broad_brute(Dev d, const double* bobb, double half, uint64_t* cand, int* cnt, int rcap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nb) return;
  int j0 = i + 1;
  if (j0 < d.nDriven) j0 = d.nDriven;
  const int jb = j0 + (int)blockIdx.y * kBruteChunk;
  const int je = min(jb + kBruteChunk, d.nb);
  const IfcRange ir = ifc_range(d.ifc, i);
  for (int j = jb; j < je; j++) {
    const bool emit = pair_overlap(bobb, i, j, half) && !has_live_interface(d.ifc, ir, j);
    emit_pair(emit, cand, cnt, rcap, i, j);
  }
}
// (the OBB test before the interface search in every broad-phase kernel: both
// are pure, and the search costs a chain of dependent loads per pair)

// --- linkCells broad phase ---------------------------------------------
__global__ void k_box_partial(int first, int nb, const double* baabb, double* part) {
  __shared__ double red[6][kBlock];
  const int t = threadIdx.x;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = first + blockIdx.x * blockDim.x + t; b < nb; b += gridDim.x * blockDim.x) {
    const double* a = baabb + 6 * (size_t)b;
    for (int k = 0; k < 3; k++) {
      v[k] = a[k] < v[k] ? a[k] : v[k];
      v[3 + k] = a[3 + k] > v[3 + k] ? a[3 + k] : v[3 + k];
    }
  }
  for (int k = 0; k < 6; k++) red[k][t] = v[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 3; k++) {
        red[k][t] = red[k][t + s] < red[k][t] ? red[k][t + s] : red[k][t];
        red[3 + k][t] = red[3 + k][t + s] > red[3 + k][t] ? red[3 + k][t + s] : red[3 + k][t];
      }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = red[k][0];
}

struct Grid {
  int N[3];
  double mn[3], sz[3];
};

// Largest distance, per dimension, from a body's position to a point of its
// OBB enlarged by `half` (a celled body sits in the cell of its position).
__device__ __forceinline__ void obb_reach(const double* o, const double p[3], double half, double r[3]) {
  for (int k = 0; k < 3; k++)
    r[k] = fabs(o[k] - p[k]) + fabs(o[3 + k]) * (o[12] + half) + fabs(o[6 + k]) * (o[13] + half) +
           fabs(o[9 + k]) * (o[14] + half);
}

// Cell of each body (oversized bodies to their own list) and, in `reach`
// (bit patterns of non-negative doubles, which order as integers), the
// largest obb_reach over the celled bodies.
__global__ void k_cell_assign(Dev d, const double* bobb, const double* baabb, double half, Grid g, int* cell_of,
                              int* ccount, int* over, int* nover, unsigned long long* reach) {
  __shared__ double red[3][kBlock];
  const int t = threadIdx.x;
  const int b = blockIdx.x * blockDim.x + t;
  double rr[3] = {0.0, 0.0, 0.0};
  if (b < d.nb) {
    const double* a = baabb + 6 * (size_t)b;
    const bool oversized = (a[3] - a[0] > g.sz[0]) || (a[4] - a[1] > g.sz[1]) || (a[5] - a[2] > g.sz[2]);
    if (oversized) {
      cell_of[b] = -1;
      over[atomicAdd(nover, 1)] = b;
    } else {
      const BodyKin& P = d.kin[b];
      const double p[3] = {P.px, P.py, P.pz};
      int c3[3];
      for (int k = 0; k < 3; k++) {
        int q = (int)floor((p[k] - g.mn[k]) / g.sz[k]);
        q = q < 0 ? 0 : q;
        q = q >= g.N[k] ? g.N[k] - 1 : q;
        c3[k] = q;
      }
      const int c = (c3[0] * g.N[1] + c3[1]) * g.N[2] + c3[2];
      cell_of[b] = c;
      atomicAdd(ccount + c, 1);
      obb_reach(bobb + 15 * (size_t)b, p, half, rr);
    }
  }
  for (int k = 0; k < 3; k++) red[k][t] = rr[k];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 3; k++) red[k][t] = red[k][t + s] > red[k][t] ? red[k][t + s] : red[k][t];
    __syncthreads();
  }
  if (t < 3) atomicMax(reach + t, (unsigned long long)__double_as_longlong(red[t][0]));
}

__global__ void k_cell_fill(int nb, const int* cell_of, const int* cstart, int* cfill, int* cbody) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int c = cell_of[b];
  if (c < 0) return;
  cbody[cstart[c] + atomicAdd(cfill + c, 1)] = b;
}

// main loop of UpdateNL_linkCells (Rockable.cpp:1697-1751): 32 lanes per
// body.  Body i's 27 neighbour cells are 9 runs of consecutive cells (the
// cells of a column along z are consecutive in the cell order), so its
// candidates are the bodies of 9 ranges of cbody; lane q < 9 finds range q,
// and the 32 lanes then take the ranges' bodies in turn, so that a lane is
// idle only past the last one (a lane per cell, looping over that cell's
// bodies, leaves the lanes of emptier cells idle while the fullest is
// walked).  The candidates are sorted afterwards, so the order in which
// they are found does not matter.
constexpr int kCellLanes = 32;
__global__ void k_broad_cells(Dev d, const double* bobb, double half, Grid g, const int* cell_of, const int* cstart,
                              const int* cbody, uint64_t* cand, int* cnt, int rcap) {
  constexpr int L = kCellLanes;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / L), q = (int)(t % L);
  const int c = i < d.nb ? cell_of[i] : -1;
  int r0 = 0, len = 0;
  if (c >= 0 && q < 9) {
    const int cz = c % g.N[2], cy = (c / g.N[2]) % g.N[1], cx = c / (g.N[2] * g.N[1]);
    const int nx = cx + q / 3 - 1, ny = cy + q % 3 - 1;
    if (nx >= 0 && ny >= 0 && nx < g.N[0] && ny < g.N[1]) {
      const int base = (nx * g.N[1] + ny) * g.N[2];
      const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz + 1 < g.N[2] ? cz + 1 : g.N[2] - 1;
      r0 = cstart[base + z0];
      len = cstart[base + z1 + 1] - r0;
    }
  }
  // inclusive prefix of the range lengths over lanes 0-8 of the body
  int pre = len;
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
    const int x = __shfl_up(pre, o, L);
    if (q >= o) pre += x;
  }
  // every lane: range k's end in the concatenation and its shift into cbody
  int pend[9], poff[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    pend[k] = __shfl(pre, k, L);
    poff[k] = __shfl(r0 - (pre - len), k, L);
  }
  const int total = pend[8];
  const bool own = c >= 0;
  const IfcRange ir = own ? ifc_range(d.ifc, i) : IfcRange{0, 0};
  for (int base = 0; __any(own && base < total); base += L) {
    const int x = base + q;
    bool emit = false;
    int j = 0;
    if (own && x < total) {
      int off = poff[8];
#pragma unroll
      for (int k = 7; k >= 0; k--)
        if (x < pend[k]) off = poff[k];
      j = cbody[x + off];
      emit = !(j < d.nDriven && i < d.nDriven) && j > i && pair_overlap(bobb, i, j, half) &&
             !has_live_interface(d.ifc, ir, j);
    }
    emit_pair(emit, cand, cnt, rcap, i, j);
  }
}

// Oversized bodies against the bodies of the cells (Rockable.cpp:1755-1805):
// workgroup (q, y) takes oversized body over[q] and every gridDim.y-th
// (x, y) column of cells.  The reference visits every cell; a celled body
// whose enlarged OBB meets the oversized body's lies within `reach` of that
// OBB's box, so only the cells covering the box plus reach (one more on each
// side against rounding) can hold one, and the others are skipped.
__global__ void k_broad_over(Dev d, const double* bobb, double half, Grid g, const int* over, int nover,
                             const int* cstart, const int* cbody, const unsigned long long* reach, uint64_t* cand,
                             int* cnt, int rcap) {
  const int q = blockIdx.x;
  if (q >= nover) return;
  const int i = over[q];
  const double* o = bobb + 15 * (size_t)i;
  int lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    const double ext =
        fabs(o[3 + k]) * (o[12] + half) + fabs(o[6 + k]) * (o[13] + half) + fabs(o[9 + k]) * (o[14] + half);
    const double rch = __longlong_as_double((long long)reach[k]);
    const double top = (double)(g.N[k] - 1);
    const double a = floor((o[k] - ext - rch - g.mn[k]) / g.sz[k]) - 1.0;
    const double b = floor((o[k] + ext + rch - g.mn[k]) / g.sz[k]) + 1.0;
    lo[k] = (int)fmin(fmax(a, 0.0), top);
    hi[k] = (int)fmin(fmax(b, 0.0), top);
  }
  const int ny = hi[1] - lo[1] + 1;
  const int ncol = (hi[0] - lo[0] + 1) * ny;
  const IfcRange ir = ifc_range(d.ifc, i);
  for (int col = blockIdx.y; col < ncol; col += gridDim.y) {
    const int cx = lo[0] + col / ny, cy = lo[1] + col % ny;
    const int cbase = (cx * g.N[1] + cy) * g.N[2];
    const int r0 = cstart[cbase + lo[2]], r1 = cstart[cbase + hi[2] + 1];
    for (int rb = r0; rb < r1; rb += blockDim.x) {
      const int r = rb + (int)threadIdx.x;
      bool emit = false;
      int j = 0;
      if (r < r1) {
        j = cbody[r];
        emit = !(j < d.nDriven && i < d.nDriven) && j > i && pair_overlap(bobb, i, j, half) &&
               !has_live_interface(d.ifc, ir, j);
      }
      emit_pair(emit, cand, cnt, rcap, i, j);
    }
  }
}

// --- directed groups -------------------------------------------------------
// head of each (i, j) run of the current (sorted) entry list
__global__ void k_old_heads(int n, const int32_t* ei, const int32_t* ej, int32_t* flag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  flag[e] = (e == 0 || ei[e] != ei[e - 1] || ej[e] != ej[e - 1]) ? 1 : 0;
}
__global__ void k_old_compact(int n, const int32_t* ei, const int32_t* ej, const int32_t* flag, const int32_t* pos,
                              uint64_t* okey, int32_t* ostart) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || !flag[e]) return;
  okey[pos[e]] = ((uint64_t)(uint32_t)ei[e] << 32) | (uint32_t)ej[e];
  ostart[pos[e]] = e;
}
// combined key list: candidates in both directions, then old groups
constexpr uint32_t kOldTag = 0x80000000u;
// The group keys (owner, other) packed as owner << bb | other, bb the bits
// of a body index, so that the radix sort runs over 2 bb bits instead of 64
// (the order is the same: owner first, then other).
__device__ __forceinline__ uint64_t pack_pair(uint32_t i, uint32_t j, int bb) { return ((uint64_t)i << bb) | j; }
__global__ void k_fill_groups(int nc, const uint64_t* cand, int ng_old, const uint64_t* okey, uint64_t* gk,
                              uint32_t* gv, int bb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nc) {
    const uint64_t k = cand[t];
    const uint32_t i = (uint32_t)(k >> 32), j = (uint32_t)(k & 0xffffffffu);
    gk[2 * t] = pack_pair(i, j, bb);
    gv[2 * t] = (uint32_t)t << 1;
    gk[2 * t + 1] = pack_pair(j, i, bb);
    gv[2 * t + 1] = ((uint32_t)t << 1) | 1u;
  } else if (t < nc + ng_old) {
    const int g = t - nc;
    const uint64_t k = okey[g];
    gk[2 * nc + g] = pack_pair((uint32_t)(k >> 32), (uint32_t)(k & 0xffffffffu), bb);
    gv[2 * nc + g] = kOldTag | (uint32_t)g;
  }
}
__global__ void k_group_heads(int n, const uint64_t* gk, int32_t* flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  flag[t] = (t == 0 || gk[t] != gk[t - 1]) ? 1 : 0;
}
__global__ void k_group_compact(int n, const uint64_t* gk, const uint32_t* gv, const int32_t* flag,
                                const int32_t* pos, uint64_t* ukey, int32_t* ucand, int32_t* uold, int bb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n || !flag[t]) return;
  const int u = pos[t];
  // back to (owner << 32) | other
  ukey[u] = ((gk[t] >> bb) << 32) | (gk[t] & ((1ull << bb) - 1ull));
  int32_t c = -1, o = -1;
  for (int q = t; q < n && q < t + 2 && gk[q] == gk[t]; q++) {
    if (gv[q] & kOldTag)
      o = (int32_t)(gv[q] & ~kOldTag);
    else
      c = (int32_t)gv[q];
  }
  ucand[u] = c;
  uold[u] = o;
}

// --- narrow phase ------------------------------------------------------
struct NarrowCfg {
  int maxv;       // max vertices per shape
  int maxe;       // max edges per shape
  int maxf;       // max faces per shape
  int maxwords;   // max 64-bit bitmask words per pair
  double dmax;    // dVerlet
  // the culls' boxes of every body's edges and faces (k_gate_boxes)
  const float* gedge;  // kGate floats per edge / face
  const float* gface;
};

__device__ __forceinline__ v3 lds_v(const double* a, int k) { return V(a[3 * k], a[3 * k + 1], a[3 * k + 2]); }

// Particle::VertexIsNearVertex/Edge/Face, EdgeIsNearEdge (Particle.cpp:88-209)
// with the global vertices of both bodies read from LDS.
__device__ bool near_vv(const double* gvi, const double* gvj, double Ri, double Rj, int isub, int jsub, double dmax) {
  v3 pi = lds_v(gvi, isub), pj = lds_v(gvj, jsub);
  double sum = Ri + Rj + dmax;
  return vnorm2(vsub(pj, pi)) <= sum * sum;
}
__device__ bool near_ve(const DevShapes& sh, int shj, const double* gvi, const double* gvj, double Ri, double Rj,
                        int isub, int jsub, double dmax) {
  const int* ed = sh.edge + 2 * (sh.eoff[shj] + jsub);
  v3 pi = lds_v(gvi, isub), p0 = lds_v(gvj, ed[0]), p1 = lds_v(gvj, ed[1]);
  v3 E = vsub(p1, p0);
  v3 v = vsub(pi, p0);
  double r = vdot(v, E) / vdot(E, E);
  if (r < -dmax) return false;
  if (r > 1.0 + dmax) return false;
  double sum = Ri + Rj + dmax;
  return vnorm2(vsub(pi, vadd(p0, vsc(r, E)))) <= sum * sum;
}
__device__ bool near_vf(const DevShapes& sh, int shj, const double* gvi, const double* gvj, double Ri, double Rj,
                        int isub, int jsub, double dmax) {
  const int gface = sh.foff[shj] + jsub;
  const int fs = sh.fstart[gface];
  const int nbv = sh.fstart[gface + 1] - fs;
  const int* fv = sh.fidx + fs;
  v3 A = lds_v(gvj, fv[0]), B = lds_v(gvj, fv[1]), C = lds_v(gvj, fv[nbv - 1]);
  v3 pi = lds_v(gvi, isub);
  v3 v = vsub(pi, A);
  v3 v1 = vsub(B, A);
  vnormalize(&v1);
  v3 v2 = vsub(C, A);
  v3 n = vcross(v1, v2);
  vnormalize(&n);
  double dist = vdot(v, n);
  if (dist < 0.0) {
    n = vneg(n);
    dist = -dist;
  }
  v3 P = vsub(pi, vsc(dist, n));
  v2 = vcross(n, v1);
  int odd = 0;
  double ori1 = vdot(P, v1), ori2 = vdot(P, v2);
  for (int iva = 0; iva < nbv; iva++) {
    int ivb = iva + 1;
    if (ivb == nbv) ivb = 0;
    v3 a = lds_v(gvj, fv[iva]), b = lds_v(gvj, fv[ivb]);
    double pa1 = vdot(a, v1), pb1 = vdot(b, v1);
    double pa2 = vdot(a, v2), pb2 = vdot(b, v2);
    if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
      if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
    }
  }
  if (odd == 1) {
    dist -= (Ri + Rj);
    if (dist <= dmax) return true;
  }
  return false;
}
__device__ bool near_ee(const DevShapes& sh, int shi, int shj, const double* gvi, const double* gvj, double Ri,
                        double Rj, int isub, int jsub, double dmax) {
  const int* edj = sh.edge + 2 * (sh.eoff[shj] + jsub);
  const int* edi = sh.edge + 2 * (sh.eoff[shi] + isub);
  v3 p0j = lds_v(gvj, edj[0]), p1j = lds_v(gvj, edj[1]);
  v3 p0i = lds_v(gvi, edi[0]), p1i = lds_v(gvi, edi[1]);
  v3 Ei = vsub(p1i, p0i);
  v3 Ej = vsub(p1j, p0j);
  v3 v = vsub(p0i, p0j);
  double c = vdot(Ei, Ei), d = vdot(Ej, Ej), e = vdot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    double a = vdot(Ei, v), b = vdot(Ej, v);
    double s = (e * b - a * d) * f;
    double t = (c * b - e * a) * f;
    if (s >= 0.0 && s <= 1.0 && t >= 0.0 && t <= 1.0) {
      v3 n = vsub(vadd(p0i, vsc(s, Ei)), vadd(p0j, vsc(t, Ej)));
      double sum = Ri + Rj + dmax;
      return vnorm2(n) <= sum * sum;
    }
  }
  return false;
}

// vertex gate of AddOrRemoveInteractions_bruteForce (Rockable.cpp:1385-1393):
// obb_other enlarged by dmax against an axis-aligned box around the vertex
__device__ __forceinline__ bool vertex_gate(v3 gv, double Rowner, const obb_t& other, double dmax) {
  obb_t sub;
  sub.c = gv;
  sub.e[0] = V(1, 0, 0);
  sub.e[1] = V(0, 1, 0);
  sub.e[2] = V(0, 0, 1);
  sub.ext = V(0, 0, 0);
  obb_enlarge(&sub, Rowner + dmax);
  obb_t o = other;
  obb_enlarge(&o, dmax);
  return obb_intersect(o, sub);
}

// Sub-element culls.  They only skip tests whose predicate is false, so the
// add / erase / retain outcome is the reference's:
//  * EdgeIsNearEdge is true only if points of the two segments lie within
//    sum = Ri + Rj + dmax of each other, so each edge's AABB grown by sum
//    meets the vertex AABB of the other body;
//  * VertexIsNearEdge accepts r in [-dmax, 1 + dmax] (r is the fraction
//    along the edge), so the point on the line is within dmax |E|_1 of the
//    edge's AABB;
//  * VertexIsNearFace needs the vertex within sum of its projection P on the
//    face plane with P inside the projected polygon, so P lies within the
//    face's thickness along its normal of the face's AABB.
// Every grown box also gets a margin of 1e-9 (1 + |x|max) for rounding.
__device__ __forceinline__ bool box_meets(const double lo[3], const double hi[3], double grow, const double* blo,
                                          const double* bhi) {
  return lo[0] - grow <= bhi[0] && hi[0] + grow >= blo[0] && lo[1] - grow <= bhi[1] && hi[1] + grow >= blo[1] &&
         lo[2] - grow <= bhi[2] && hi[2] + grow >= blo[2];
}
// The parts of the edge and face culls that depend on the body alone, once
// per body and rebuild instead of once per candidate pair: an edge's box and
// the L1 length of its box; a face's box and its thickness along its normal
// (the face normal from its first two edges, as the predicates take it).
// Stored as 8 floats each (lo[3], hi[3], length or thickness, unused),
// rounded outwards: lo down, hi and the growth terms up, so that the culls
// stay conservative at half the bytes (the narrow phase reads them once per
// candidate pair).  Body b's edges from beoff[b], its faces from bfoff[b].
constexpr int kGate = 8;
__device__ __forceinline__ void store_gate(float* o, const double lo[3], const double hi[3], double x) {
  float4 a, b;
  a.x = __double2float_rd(lo[0]);
  a.y = __double2float_rd(lo[1]);
  a.z = __double2float_rd(lo[2]);
  a.w = __double2float_ru(hi[0]);
  b.x = __double2float_ru(hi[1]);
  b.y = __double2float_ru(hi[2]);
  b.z = __double2float_ru(x);
  b.w = 0.0f;
  reinterpret_cast<float4*>(o)[0] = a;
  reinterpret_cast<float4*>(o)[1] = b;
}
// 2^lg lanes per body (lg from 3 to 6: the shapes' largest edge plus face
// count rounded up), each lane taking every 2^lg-th of its body's elements.
__global__ void k_gate_boxes(Dev d, float* gedge, float* gface, int lg) {
  const unsigned t = blockIdx.x * (unsigned)blockDim.x + threadIdx.x;
  const int b = (int)(t >> lg);
  const int lane = (int)(t & ((1u << lg) - 1u));
  if (b >= d.nb) return;
  const int shp = d.kin[b].shape;
  const int ne = d.sh.ne[shp], nf = d.sh.nf[shp];
  const double* gv = d.gv + 3 * (size_t)d.bvoff[b];
  for (int k = lane; k < ne + nf; k += 1 << lg) {
    double lo[3], hi[3], x;
    if (k < ne) {
      const int* ed = d.sh.edge + 2 * (d.sh.eoff[shp] + k);
      const v3 a = lds_v(gv, ed[0]), c = lds_v(gv, ed[1]);
      lo[0] = fmin(a.x, c.x);
      lo[1] = fmin(a.y, c.y);
      lo[2] = fmin(a.z, c.z);
      hi[0] = fmax(a.x, c.x);
      hi[1] = fmax(a.y, c.y);
      hi[2] = fmax(a.z, c.z);
      x = (hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]);
    } else {
      const int gfc = d.sh.foff[shp] + (k - ne);
      const int fs = d.sh.fstart[gfc];
      const int nbv = d.sh.fstart[gfc + 1] - fs;
      const int* fv = d.sh.fidx + fs;
      const v3 A = lds_v(gv, fv[0]), B = lds_v(gv, fv[1]), C = lds_v(gv, fv[nbv - 1]);
      v3 v1 = vsub(B, A);
      vnormalize(&v1);
      v3 n = vcross(v1, vsub(C, A));
      vnormalize(&n);
      lo[0] = hi[0] = A.x;
      lo[1] = hi[1] = A.y;
      lo[2] = hi[2] = A.z;
      x = 0.0;
      for (int q = 1; q < nbv; q++) {
        const v3 p = lds_v(gv, fv[q]);
        lo[0] = fmin(lo[0], p.x);
        lo[1] = fmin(lo[1], p.y);
        lo[2] = fmin(lo[2], p.z);
        hi[0] = fmax(hi[0], p.x);
        hi[1] = fmax(hi[1], p.y);
        hi[2] = fmax(hi[2], p.z);
        x = fmax(x, fabs(vdot(vsub(p, A), n)));
      }
    }
    store_gate(k < ne ? gedge + kGate * ((size_t)d.beoff[b] + k) : gface + kGate * ((size_t)d.bfoff[b] + (k - ne)),
               lo, hi, x);
  }
}

// The culls (they skip only tests whose predicate is provably false):
// EdgeIsNearEdge needs points of both segments within Ri + Rj + dmax, so an
// edge whose box grown by that (and dmax times its L1 length, for
// VertexIsNearEdge's r in [-dmax, 1 + dmax]) misses the other body's box has
// no test to pass; VertexIsNearFace needs the projection inside the face's
// prism (box grown by twice the face's thickness).
__device__ __forceinline__ bool edge_gate_pre(const float* g, double sum, double dmax, double margin,
                                              const double* blo, const double* bhi) {
  const float4 a = reinterpret_cast<const float4*>(g)[0], b = reinterpret_cast<const float4*>(g)[1];
  const double lo[3] = {a.x, a.y, a.z}, hi[3] = {a.w, b.x, b.y};
  return box_meets(lo, hi, sum + dmax * (double)b.z + margin, blo, bhi);
}
__device__ __forceinline__ bool face_gate_pre(const float* g, double sum, double margin, const double* blo,
                                              const double* bhi) {
  const float4 a = reinterpret_cast<const float4*>(g)[0], b = reinterpret_cast<const float4*>(g)[1];
  const double thick = b.z;
  if (!(thick == thick)) return true;  // degenerate face: no cull
  const double lo[3] = {a.x, a.y, a.z}, hi[3] = {a.w, b.x, b.y};
  return box_meets(lo, hi, sum + 2.0 * thick + margin, blo, bhi);
}



// The largest (hi) or smallest projection on axis a of a point of an edge's
// or face's box (the gate data: lo[3], hi[3] as floats).
__device__ __forceinline__ double gate_box_reach(const float* g, v3 a, bool hi) {
  const float4 p = reinterpret_cast<const float4*>(g)[0], q = reinterpret_cast<const float4*>(g)[1];
  const double lo[3] = {p.x, p.y, p.z}, up[3] = {p.w, q.x, q.y}, ak[3] = {a.x, a.y, a.z};
  double s = 0.0;
  for (int k = 0; k < 3; k++) {
    const double x = ak[k] * lo[k], y = ak[k] * up[k];
    s += hi ? fmax(x, y) : fmin(x, y);
  }
  return s;
}

// floor(k / n) for 0 <= k < 2^20 and n >= 1, with inv = 1 / n in FP32:
// (k + 0.5) / n lies at least 0.5 / n from an integer, and the FP32 product
// is within 3e-7 of it relative, less than that for k < 2^20.  Larger k
// (shapes of thousands of sub-elements) take the integer division.
__device__ __forceinline__ int quot_small(int k, int n, float inv) {
  return k < (1 << 20) ? (int)(((float)k + 0.5f) * inv) : k / n;
}

// Compaction of the indices v < n with pred(v) into list (in ascending
// order) by the SW lanes of one pair (sl: lane within them, g: the pair's
// group of lanes in the wavefront); returns the count.
template <int SW, class F>
__device__ __forceinline__ int group_compact(int n, int sl, int g, int* list, F pred) {
  int base = 0;
  for (int v0 = 0; v0 < n; v0 += SW) {
    const int v = v0 + sl;
    const bool f = v < n && pred(v);
    unsigned long long m = __ballot(f);
    if constexpr (SW < 64) m = (m >> (g * SW)) & ((1ull << SW) - 1ull);
    if (f) list[base + __popcll(m & ((1ull << sl) - 1ull))] = v;
    base += __popcll(m);
  }
  return base;
}

template <int SW>
__device__ __forceinline__ int group_sum(int x) {
  for (int off = SW / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, SW);
  return x;
}
template <int SW>
__device__ __forceinline__ double group_max(double x) {
  for (int off = SW / 2; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, SW));
  return x;
}
template <int SW>
__device__ __forceinline__ double group_min(double x) {
  for (int off = SW / 2; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off, SW));
  return x;
}

// Candidate pairs of two point shapes (one vertex, no edge or face: the
// spheres), one pair per lane.  Their only test is vertex-vertex, so their
// result is k_narrow's computed directly: the key (VV, 0, 0) when i's vertex
// passes the reference's gate against j's OBB, j's vertex lies within sum of
// i's box and VertexIsNearVertex holds.  A wavefront's near keys go to one
// stripe (one atomic); rest[c] flags the other candidates for k_narrow.
__global__ void __launch_bounds__(kBlock) k_narrow_points(Dev d, const double* bobb, const double* baabb,
                                                          NarrowCfg cfg, int nc, const uint64_t* cand,
                                                          uint32_t* nearbuf, int* near_cnt, int near_rcap,
                                                          int32_t* cbase, int32_t* cnf, int32_t* cnb, int32_t* rest) {
  const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int lane = threadIdx.x & 63;
  bool point = false, near = false;
  if (c < nc) {
    const uint64_t key = cand[c];
    const int i = (int)(key >> 32), j = (int)(key & 0xffffffffu);
    const int shi = d.kin[i].shape, shj = d.kin[j].shape;
    point = d.sh.nv[shi] == 1 && d.sh.ne[shi] == 0 && d.sh.nf[shi] == 0 && d.sh.nv[shj] == 1 && d.sh.ne[shj] == 0 &&
            d.sh.nf[shj] == 0;
    rest[c] = point ? 0 : 1;
    if (point) {
      const double Ri = d.kin[i].R, Rj = d.kin[j].R;
      const double *bxi = baabb + 6 * (size_t)i, *bxj = baabb + 6 * (size_t)j;
      double xmax = 0.0;
      for (int k = 0; k < 6; k++) xmax = fmax(fmax(xmax, fabs(bxi[k])), fabs(bxj[k]));
      const double dmax = cfg.dmax;
      const double sum = Ri + Rj + dmax;
      const double margin = 1e-9 * (1.0 + xmax + sum);
      const double* gvi = d.gv + 3 * (size_t)d.bvoff[i];
      const double* gvj = d.gv + 3 * (size_t)d.bvoff[j];
      const double gvv = sum + margin;
      const v3 pj = lds_v(gvj, 0);
      const bool jnear = pj.x >= bxi[0] - gvv && pj.x <= bxi[3] + gvv && pj.y >= bxi[1] - gvv &&
                         pj.y <= bxi[4] + gvv && pj.z >= bxi[2] - gvv && pj.z <= bxi[5] + gvv;
      near = jnear && vertex_gate(lds_v(gvi, 0), Ri, load_obb(bobb + 15 * (size_t)j), dmax) &&
             near_vv(gvi, gvj, Ri, Rj, 0, 0, dmax);
    }
  }
  const unsigned long long m = __ballot(near);
  const int stripe = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & (kStripes - 1);
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(near_cnt + stripe * kStripeStride, __popcll(m));
  base = __shfl(base, 0, 64);
  if (point) {
    const int local = base + __popcll(m & ((1ull << lane) - 1ull));
    cbase[c] = stripe * near_rcap + local;
    cnf[c] = near ? 1 : 0;
    cnb[c] = 0;
    // over capacity: the host re-runs with a larger buffer
    if (near && local < near_rcap) nearbuf[stripe * near_rcap + local] = pack_sub(VV, 0, 0);
  }
}

// P candidate pairs per wavefront (64 / P lanes each, LDS stride `stride`
// bytes per pair): the work of a pair is mostly a fixed sequence of
// wavefront instructions (gates, compactions, one pass per predicate, the
// key output), which the pairs of a wavefront then share.  Occupancy asked
// of the compiler: 6 wavefronts per SIMD with one pair (80 VGPRs, no
// spill), 3 with two.
template <int P>
__global__ void __launch_bounds__(64, P == 1 ? 6 : 3) k_narrow(Dev d, const double* bobb, const double* baabb, NarrowCfg cfg, int nc,
                                               const uint64_t* cand, uint32_t* nearbuf, int* near_cnt,
                                               int near_rcap, int32_t* cbase, int32_t* cnf, int32_t* cnb,
                                               int stride, const int32_t* clist) {
  constexpr int SW = 64 / P;
  extern __shared__ double lds_all[];
  const int lane = threadIdx.x;
  // (with one pair per wavefront the pair's values stay wavefront-uniform,
  // in scalar registers)
  const int g = P == 1 ? 0 : lane / SW, sl = P == 1 ? lane : lane % SW;
  // the idx-th pair to test: candidate clist[idx] (all candidates in order
  // without a list)
  const int idx = xcd_block((int)blockIdx.x, (int)gridDim.x) * P + g;
  const bool valid = idx < nc;
  if (P == 1 && !valid) return;
  const int c = valid ? (clist ? clist[idx] : idx) : 0;
  double* lds = lds_all + (size_t)g * (stride / 8);
  uint64_t key = 0;
  if (valid) key = cand[c];
  const int i = (int)(key >> 32), j = (int)(key & 0xffffffffu);
  const int shi = d.kin[i].shape, shj = d.kin[j].shape;
  const double Ri = d.kin[i].R, Rj = d.kin[j].R;
  // an invalid group (past the last pair) takes body 0 twice with no
  // sub-element, so that it runs every pass with nothing to do
  const int nvi = valid ? d.sh.nv[shi] : 0, nvj = valid ? d.sh.nv[shj] : 0;
  const int nei = valid ? d.sh.ne[shi] : 0, nej = valid ? d.sh.ne[shj] : 0;
  const int nfi = valid ? d.sh.nf[shi] : 0, nfj = valid ? d.sh.nf[shj] : 0;
  double* pbox = lds;                // 48 doubles: both OBBs and both body boxes (several pairs per wavefront)
  double* gvi = lds + 48;
  double* gvj = gvi + 3 * cfg.maxv;
  unsigned long long* bits = (unsigned long long*)(gvj + 3 * cfg.maxv);
  int* lst = (int*)(bits + cfg.maxwords);
  int* vg_l = lst;                  // gated vertices: i's as v, j's as maxv + v
  int* vjn_l = vg_l + 2 * cfg.maxv; // vertices of j near i's box (vertex-vertex)
  int* ef_l = vjn_l + cfg.maxv;     // gated edges and faces, tagged (see below)
  int* wpre = ef_l + 2 * cfg.maxe + 2 * cfg.maxf + 8;  // prefix of the bit counts per word

  // test order of the reference (AddOrRemoveInteractions_bruteForce): the
  // bit of test (segment s, own, oth) is segstart[s] + own * n_oth[s] + oth
  const int s1 = nvi * nvj, s2 = s1 + nvi * nej, s3 = s2 + nvi * nfj, s4 = s3 + nei * nej, s5 = s4 + nvj * nei;
  const int T = s5 + nvj * nfi, Tf = s4;
  // (nwords <= maxwords: the shapes are checked at load)
  const int nwords = min((T + 63) / 64, cfg.maxwords);
  for (int w = sl; w < nwords; w += SW) bits[w] = 0ull;
  {
    const double* src_i = d.gv + 3 * (size_t)d.bvoff[i];
    const double* src_j = d.gv + 3 * (size_t)d.bvoff[j];
    for (int q = sl; q < 3 * nvi; q += SW) gvi[q] = src_i[q];
    for (int q = sl; q < 3 * nvj; q += SW) gvj[q] = src_j[q];
  }
  // the OBBs and the body boxes (vertices +- Minkowski radius, from k_obb)
  // of both bodies: read where they lie when the pair's values are uniform
  // over the wavefront (scalar loads), else from copies in the pair's LDS
  // (vector registers would hold them for every lane)
  const double *obi = bobb + 15 * (size_t)i, *obj = bobb + 15 * (size_t)j;
  const double *bxi = baabb + 6 * (size_t)i, *bxj = baabb + 6 * (size_t)j;
  if (P > 1) {
    for (int q = sl; q < 42; q += SW)
      pbox[q] = q < 15 ? obi[q] : q < 30 ? obj[q - 15] : q < 36 ? bxi[q - 30] : bxj[q - 36];
    obi = pbox;
    obj = pbox + 15;
    bxi = pbox + 30;
    bxj = pbox + 36;
  }
  __syncthreads();
  double xmax = 0.0;
  for (int k = 0; k < 6; k++) xmax = fmax(fmax(xmax, fabs(bxi[k])), fabs(bxj[k]));
  const double dmax = cfg.dmax;
  const double sum = Ri + Rj + dmax;
  const double margin = 1e-9 * (1.0 + xmax + sum);

  // Every near test of the pair is false when a plane separates the two
  // bodies' vertices by more than the tests reach: a vertex-vertex or
  // edge-edge test compares the distance of two points of the vertices'
  // convex hulls with sum; a vertex-edge test takes a foot on the edge's line
  // up to dVerlet edge lengths past its ends; a vertex-face test takes a
  // foot in the face's outline projected on a plane the face's vertices lie
  // within reach[1] of.  The planes tried are those normal to the line of
  // centres, whose projection is 1-Lipschitz, so a gap along it above sum +
  // those allowances + margin (rounding) leaves no test true.  Such pairs are
  // most candidates in a settled packing; they get zero counts without a
  // gate or test.  A NaN coordinate (or a zero axis) turns the test off.
  // The same bound culls single elements of the other pairs: a test of an
  // element of i needs its hull's projection to reach lo_j - slack (lim_i),
  // one of j's to reach down to hi_i + slack (lim_j); the culls apply it to
  // the vertices and to the edges' and faces' boxes, which hold them.
  bool apart = false;
  v3 ax = V(0.0, 0.0, 0.0);
  double lim_i = -INFINITY, lim_j = INFINITY;
  // bit v: vertex v within the limit (all set where not computed, shapes of
  // more than SW vertices)
  unsigned long long keep_i = ~0ull, keep_j = ~0ull;
  if (valid) {
    const BodyKin& Ki = d.kin[i];
    const BodyKin& Kj = d.kin[j];
    ax = V(Kj.px - Ki.px, Kj.py - Ki.py, Kj.pz - Ki.pz);
    const double a2 = vdot(ax, ax);
    if (a2 > 0.0) {
      ax = vsc(1.0 / sqrt(a2), ax);
      double hi_i = -INFINITY, lo_j = INFINITY, qi = 0.0, qj = 0.0;
      for (int v = sl; v < nvi; v += SW) {
        qi = vdot(ax, lds_v(gvi, v));
        hi_i = qi == qi ? fmax(hi_i, qi) : INFINITY;
      }
      for (int v = sl; v < nvj; v += SW) {
        qj = vdot(ax, lds_v(gvj, v));
        lo_j = qj == qj ? fmin(lo_j, qj) : -INFINITY;
      }
      hi_i = group_max<SW>(hi_i);
      lo_j = group_min<SW>(lo_j);
      const double* ri = d.sh.reach + 2 * shi;
      const double* rj = d.sh.reach + 2 * shj;
      const double reach = dmax * fmax(Ki.h * ri[0], Kj.h * rj[0]) + fmax(Ki.h * ri[1], Kj.h * rj[1]);
      const double slack = sum + reach + margin;
      apart = lo_j - hi_i > slack;
      const double li = lo_j - slack, lj = hi_i + slack;
      if (li == li) lim_i = li;
      if (lj == lj) lim_j = lj;
      if (nvi <= SW && nvj <= SW) {
        // (qi, qj: the lane's own vertex; a NaN one is dropped, as its tests are false)
        unsigned long long bi = __ballot(sl < nvi && qi >= lim_i), bj = __ballot(sl < nvj && qj <= lim_j);
        if constexpr (SW < 64) {
          bi = (bi >> (g * SW)) & ((1ull << SW) - 1ull);
          bj = (bj >> (g * SW)) & ((1ull << SW) - 1ull);
        }
        keep_i = bi;
        keep_j = bj;
      }
    } else {
      ax = V(0.0, 0.0, 0.0);
    }
  }
  // (with several pairs per wavefront, when every pair of it is apart; the
  // workgroup is this one wavefront)
  if (P == 1 ? apart : __all(apart || !valid)) {
    if (valid && sl == 0) {
      cbase[c] = 0;
      cnf[c] = 0;
      cnb[c] = 0;
    }
    return;
  }

  // edges and faces of both bodies in one pass, index q over i's edges, j's
  // edges, i's faces, j's faces (before the vertices, so that ax and the
  // limits are dead in the vertex gates' registers)
  const int ne2 = nei + nej, nef = ne2 + nfi + nfj;
  const float* gei = cfg.gedge + kGate * (size_t)d.beoff[i];
  const float* gej = cfg.gedge + kGate * (size_t)d.beoff[j];
  const float* gfi = cfg.gface + kGate * (size_t)d.bfoff[i];
  const float* gfj = cfg.gface + kGate * (size_t)d.bfoff[j];
  const int nefg = group_compact<SW>(nef, sl, g, ef_l, [&](int q) {
    if (apart) return false;
    if (q < nei) {
      const float* gb = gei + kGate * q;
      return gate_box_reach(gb, ax, true) >= lim_i && edge_gate_pre(gb, sum, dmax, margin, bxj, bxj + 3);
    }
    if (q < ne2) {
      const float* gb = gej + kGate * (q - nei);
      return gate_box_reach(gb, ax, false) <= lim_j && edge_gate_pre(gb, sum, dmax, margin, bxi, bxi + 3);
    }
    if (q < ne2 + nfi) {
      const float* gb = gfi + kGate * (q - ne2);
      return gate_box_reach(gb, ax, true) >= lim_i && face_gate_pre(gb, sum, margin, bxj, bxj + 3);
    }
    const float* gb = gfj + kGate * (q - ne2 - nfi);
    return gate_box_reach(gb, ax, false) <= lim_j && face_gate_pre(gb, sum, margin, bxi, bxi + 3);
  });
  // reference vertex gates of both bodies in one pass: i's vertices against
  // j's OBB, j's against i's
  const int nvg = group_compact<SW>(nvi + nvj, sl, g, vg_l, [&](int v) {
    if (apart) return false;
    if (v < nvi) return (v >= 64 || ((keep_i >> v) & 1ull)) && vertex_gate(lds_v(gvi, v), Ri, load_obb(obj), dmax);
    const int u = v - nvi;
    return (u >= 64 || ((keep_j >> u) & 1ull)) && vertex_gate(lds_v(gvj, u), Rj, load_obb(obi), dmax);
  });
  // vertex-vertex partners: vertices of j within sum of i's box
  const double gvv = sum + margin;
  const int nvj_n = group_compact<SW>(nvj, sl, g, vjn_l, [&](int v) {
    const v3 p = lds_v(gvj, v);
    return p.x >= bxi[0] - gvv && p.x <= bxi[3] + gvv && p.y >= bxi[1] - gvv && p.y <= bxi[4] + gvv &&
           p.z >= bxi[2] - gvv && p.z <= bxi[5] + gvv && (v >= 64 || ((keep_j >> v) & 1ull));
  });
  // split the compacted lists by kind (they are in ascending q, so each kind
  // is a contiguous run)
  int nvi_g = 0, nei_g = 0, nej_g = 0, nfi_g = 0;
  for (int k = sl; k < nvg; k += SW) nvi_g += vg_l[k] < nvi ? 1 : 0;
  for (int k = sl; k < nefg; k += SW) {
    const int q = ef_l[k];
    nei_g += q < nei ? 1 : 0;
    nej_g += (q >= nei && q < ne2) ? 1 : 0;
    nfi_g += (q >= ne2 && q < ne2 + nfi) ? 1 : 0;
  }
  nvi_g = group_sum<SW>(nvi_g);
  nei_g = group_sum<SW>(nei_g);
  nej_g = group_sum<SW>(nej_g);
  nfi_g = group_sum<SW>(nfi_g);
  const int nvj_g = nvg - nvi_g, nfj_g = nefg - nei_g - nej_g - nfi_g;
  const int* vi_l = vg_l;            // values v
  const int* vj_l = vg_l + nvi_g;    // values nvi + v
  const int* ei_l = ef_l;            // values e
  const int* ej_l = ef_l + nei_g;    // values nei + e
  const int* fi_l = ej_l + nej_g;    // values ne2 + f
  const int* fj_l = fi_l + nfi_g;    // values ne2 + nfi + f
  __syncthreads();

  // one loop per segment, so that a wavefront runs one predicate at a time
  int n = nvi_g * nvj_n;
  float inv = 1.0f / (float)max(nvj_n, 1);
  for (int k = sl; k < n; k += SW) {
    const int q = quot_small(k, nvj_n, inv);
    const int a = vi_l[q], b = vjn_l[k - q * nvj_n];
    if (near_vv(gvi, gvj, Ri, Rj, a, b, dmax)) {
      const int bit = a * nvj + b;
      atomicOr(bits + (bit >> 6), 1ull << (bit & 63));
    }
  }
  // vertex-edge both ways in one pass (i's vertices against j's edges, then
  // j's against i's, the same predicate with the roles swapped per lane),
  // and vertex-face likewise: a wavefront runs each predicate once where
  // the two directions' tests fit its lanes together
  {
    const int n1 = nvi_g * nej_g, n2 = nvj_g * nei_g;
    const float inv1 = 1.0f / (float)max(nej_g, 1), inv2 = 1.0f / (float)max(nei_g, 1);
    for (int k = sl; k < n1 + n2; k += SW) {
      const bool fw = k < n1;
      const int kk = fw ? k : k - n1, no = fw ? nej_g : nei_g;
      const int q = quot_small(kk, no, fw ? inv1 : inv2);
      const int a = fw ? vi_l[q] : vj_l[q] - nvi;
      const int b = fw ? ej_l[kk - q * no] - nei : ei_l[kk - q * no];
      if (near_ve(d.sh, fw ? shj : shi, fw ? gvi : gvj, fw ? gvj : gvi, fw ? Ri : Rj, fw ? Rj : Ri, a, b, dmax)) {
        const int bit = fw ? s1 + a * nej + b : s4 + a * nei + b;
        atomicOr(bits + (bit >> 6), 1ull << (bit & 63));
      }
    }
  }
  {
    const int n1 = nvi_g * nfj_g, n2 = nvj_g * nfi_g;
    const float inv1 = 1.0f / (float)max(nfj_g, 1), inv2 = 1.0f / (float)max(nfi_g, 1);
    for (int k = sl; k < n1 + n2; k += SW) {
      const bool fw = k < n1;
      const int kk = fw ? k : k - n1, no = fw ? nfj_g : nfi_g;
      const int q = quot_small(kk, no, fw ? inv1 : inv2);
      const int a = fw ? vi_l[q] : vj_l[q] - nvi;
      const int b = fw ? fj_l[kk - q * no] - ne2 - nfi : fi_l[kk - q * no] - ne2;
      if (near_vf(d.sh, fw ? shj : shi, fw ? gvi : gvj, fw ? gvj : gvi, fw ? Ri : Rj, fw ? Rj : Ri, a, b, dmax)) {
        const int bit = fw ? s2 + a * nfj + b : s5 + a * nfi + b;
        atomicOr(bits + (bit >> 6), 1ull << (bit & 63));
      }
    }
  }
  n = nei_g * nej_g;
  inv = 1.0f / (float)max(nej_g, 1);
  for (int k = sl; k < n; k += SW) {
    const int q = quot_small(k, nej_g, inv);
    const int a = ei_l[q], b = ej_l[k - q * nej_g] - nei;
    if (near_ee(d.sh, shi, shj, gvi, gvj, Ri, Rj, a, b, dmax)) {
      const int bit = s3 + a * nej + b;
      atomicOr(bits + (bit >> 6), 1ull << (bit & 63));
    }
  }
  __syncthreads();
  // counts per direction
  int cf = 0, cb = 0;
  for (int w = sl; w < nwords; w += SW) {
    const unsigned long long m = bits[w];
    const int lo = w * 64;
    if (lo + 64 <= Tf)
      cf += __popcll(m);
    else if (lo >= Tf)
      cb += __popcll(m);
    else {
      const int k = Tf - lo;
      const unsigned long long fm = (k >= 64) ? ~0ull : ((1ull << k) - 1ull);
      cf += __popcll(m & fm);
      cb += __popcll(m & ~fm);
    }
  }
  cf = group_sum<SW>(cf);
  cb = group_sum<SW>(cb);
  const int stripe = c & (kStripes - 1);
  int local = 0;
  if (valid && sl == 0) local = atomicAdd(near_cnt + stripe * kStripeStride, cf + cb);
  local = __shfl(local, 0, SW);
  const int base_out = stripe * near_rcap + local;
  if (valid && sl == 0) {
    cbase[c] = base_out;
    cnf[c] = cf;
    cnb[c] = cb;
  }
  // over capacity: the host re-runs with a larger buffer
  const bool room = valid && local + cf + cb <= near_rcap;
  // write the keys in test order (output slot = rank of the set bit), one
  // key per lane: the exclusive prefix of the words' bit counts in LDS, then
  // key q's word by a binary search over it and its bit as the
  // (q - prefix)-th set bit of the word (a binary search over popcounts)
  {
    int running = 0;
    for (int w0 = 0; w0 < nwords; w0 += SW) {
      const int w = w0 + sl;
      const int pc = w < nwords ? __popcll(bits[w]) : 0;
      int incl = pc;
      for (int off = 1; off < SW; off <<= 1) {
        const int y = __shfl_up(incl, off, SW);
        if (sl >= off) incl += y;
      }
      if (w < nwords) wpre[w] = running + incl - pc;
      running += __shfl(incl, SW - 1, SW);
    }
  }
  __syncthreads();
  const int ntot = room ? cf + cb : 0;
  for (int q = sl; q < ntot; q += SW) {
    int lo = 0, hi = nwords - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (wpre[mid] <= q)
        lo = mid;
      else
        hi = mid - 1;
    }
    unsigned long long m = bits[lo];
    int r = q - wpre[lo], bit = 0;
    for (int width = 32; width > 0; width >>= 1) {
      const int cnt = __popcll(m & ((1ull << width) - 1ull));
      if (r >= cnt) {
        r -= cnt;
        m >>= width;
        bit += width;
      }
    }
    const int t = lo * 64 + bit;
    // the segment of test t: its type, first bit and count of "other" elements
    int type, first, noth;
    if (t < s1) {
      type = VV, first = 0, noth = nvj;
    } else if (t < s2) {
      type = VE, first = s1, noth = nej;
    } else if (t < s3) {
      type = VF, first = s2, noth = nfj;
    } else if (t < s4) {
      type = EE, first = s3, noth = nej;
    } else if (t < s5) {
      type = VE, first = s4, noth = nei;
    } else {
      type = VF, first = s5, noth = nfi;
    }
    const int rt = t - first;
    const int own = quot_small(rt, noth, 1.0f / (float)noth);
    nearbuf[base_out + q] = pack_sub(type, own, rt - own * noth);
  }
}

// --- merge ---------------------------------------------------------------
struct MergeIn {
  int ng;                    // groups (count pass with ng_pos: the groups' upper bound)
  const int32_t* ng_pos;     // count pass: the group count is *ng_pos + *ng_flag (the
  const int32_t* ng_flag;    // heads' scan and flag at the last key), else null
  const uint64_t* ukey;
  const int32_t* ucand;
  const int32_t* uold;
  const int32_t* ostart;  // old group start, size ng_old + 1 (trailing = n_old)
  const int32_t* cbase;
  const int32_t* cnf;
  const int32_t* cnb;
  const uint32_t* nearbuf;
  const double* bobb;
  double dmax;
  // old entries
  const int32_t* oi;
  const int32_t* oj;
  const uint32_t* osub;
  const EState* oest;
  const double* oedn;
  const EGeom* ogeo;
  const double* odamp;
  const int32_t* ostick;
  const uint8_t* oeact;
  const double* oekey;
};
struct MergeOut {
  int32_t* count;   // per group (count pass)
  const int32_t* offset;  // per group (write pass)
  int32_t* src;     // per new entry (write pass): the old entry kept, -1 for a fresh one
  int32_t* ei;
  int32_t* ej;
  uint32_t* esub;
  EState* est;
  double* edn;
  EGeom* egeo;
  double* edamp;
  int32_t* estick;
  uint8_t* eact;
  double* ekey;
};

// Damp of the body pair (Rockable.cpp:1356-1376), pair (i < j)
__device__ double pair_damp(const Dev& d, int i, int j) {
  const int gi = d.kin[i].group, gj = d.kin[j].group;
  if (gi < 0 || gj < 0 || gi >= d.par.G || gj >= d.par.G) return 0.0;
  const int gp = gi * d.par.G + gj;
  const int kind = d.par.damp_kind[gp];
  if (kind == 2) return 0.0;
  const double kn = d.par.tab[(P_KN * d.par.G + gi) * d.par.G + gj];
  double meff;
  if (i < d.nDriven)
    meff = d.mass[j];
  else if (j < d.nDriven)
    meff = d.mass[i];
  else
    meff = (d.mass[i] * d.mass[j]) / (d.mass[i] + d.mass[j]);
  if (kind == 0) return d.par.damp_rate[gp] * 2.0 * sqrt(kn * meff);
  return 2.0 * sqrt(kn * meff);
}

// The group count and the merged list's length (the count scan at it), for
// one read with the narrow phase's totals.
__global__ void k_merge_total(const int32_t* pos_last, const int32_t* flag_last, const int32_t* goff, int32_t* out) {
  const int ng = *pos_last + *flag_last;
  out[0] = ng;
  out[1] = goff[ng];
}

template <bool WRITE>
__global__ void k_merge(Dev d, MergeIn in, MergeOut out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  // the count pass runs over the groups' upper bound (+1, the scan's
  // trailing zero) and zeroes the counts past the last group
  const int ng = (!WRITE && in.ng_pos) ? *in.ng_pos + *in.ng_flag : in.ng;
  if (g >= ng) {
    if (!WRITE && g <= in.ng) out.count[g] = 0;
    return;
  }
  const uint64_t key = in.ukey[g];
  const int owner = (int)(key >> 32), other = (int)(key & 0xffffffffu);
  const int cref = in.ucand[g], oref = in.uold[g];
  int os = 0, on = 0;
  if (oref >= 0) {
    os = in.ostart[oref];
    on = in.ostart[oref + 1] - os;
  }
  int ns = 0, nn = 0;
  bool fwd = true;
  if (cref >= 0) {
    const int c = cref >> 1;
    fwd = (cref & 1) == 0;
    ns = in.cbase[c] + (fwd ? 0 : in.cnf[c]);
    nn = fwd ? in.cnf[c] : in.cnb[c];
  }
  const bool candidate = cref >= 0;
  const BodyKin Po = d.kin[owner];
  obb_t oth;
  if (candidate) oth = load_obb(in.bobb + 15 * (size_t)other);
  int cnt = WRITE ? out.offset[g] : 0;
  int a = 0, b = 0;
  while (a < on || b < nn) {
    const uint32_t sa = (a < on) ? in.osub[os + a] : 0xffffffffu;
    const uint32_t sb = (b < nn) ? in.nearbuf[ns + b] : 0xffffffffu;
    int take;  // 0 old kept, 1 fresh near, 2 old matched by near, -1 old erased
    if (b < nn && (a >= on || sb < sa)) {
      take = 1;
    } else if (a < on && (b >= nn || sa < sb)) {
      bool tested = false;
      if (candidate) {
        const int type = (int)(sa >> 30), isub = (int)((sa >> 15) & 0x7fffu);
        if (fwd && type == EE)
          tested = true;
        else if (type == VV || type == VE || type == VF) {
          if (fwd || type != VV) {
            const double* g = d.gv + 3 * ((size_t)d.bvoff[owner] + isub);
            tested = vertex_gate(V(g[0], g[1], g[2]), Po.R, oth, in.dmax);
          }
        }
      }
      take = tested ? -1 : 0;
    } else {
      take = 2;
    }
    if (take != -1) {
      if (WRITE) {
        // Records are filled by k_merge_fill, one thread per entry: written
        // here, a group's records would be strided stores across the wavefront.
        const int e = cnt;
        out.ei[e] = owner;
        out.ej[e] = other;
        out.esub[e] = take == 1 ? sb : in.osub[os + a];
        out.src[e] = take == 1 ? -1 : os + a;
      }
      cnt++;
    }
    if (take == 1)
      b++;
    else if (take == 2) {
      a++;
      b++;
    } else
      a++;
  }
  if (!WRITE) out.count[g] = cnt;
}

// Records of the merged entries: a fresh entry starts from zero state with the
// damp of its body pair; a kept one carries its old state, active as it was at
// the last accelerations() (Rockable.cpp:1356-1376 for the damp).
__global__ void k_merge_fill(Dev d, int n, MergeIn in, MergeOut out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int o = out.src[e];
  if (o < 0) {
    const int i = out.ei[e], j = out.ej[e];
    EState s;
    s.fn = 0.0;
    s.ftx = s.fty = s.ftz = 0.0;
    s.momx = s.momy = s.momz = 0.0;
    out.est[e] = s;
    out.edn[e] = 0.0;
    EGeom q;
    q.nx = q.ny = q.nz = q.px = q.py = q.pz = q.vx = q.vy = q.vz = 0.0;
    out.egeo[e] = q;
    out.edamp[e] = pair_damp(d, i < j ? i : j, i < j ? j : i);
    out.estick[e] = -1;
    out.eact[e] = 0;
  } else {
    const double dn = in.oedn[o];
    const int stick = in.ostick[o];
    const EState st = in.oest[o];
    out.est[e] = st;
    out.edn[e] = dn;
    out.egeo[e] = in.ogeo[o];
    out.edamp[e] = in.odamp[o];
    out.estick[e] = stick;
    // bit 1: the history is not all zero (a NaN counts as non-zero); bit 0
    // is kept: the entries active at the last force phase stay the active
    // set until the next one (activeInteractions holds its pointers through
    // UpdateNL, Rockable.cpp:3005-3029), a bond broken in that phase included
    const bool hist = !(st.fn == 0.0 && st.ftx == 0.0 && st.fty == 0.0 && st.ftz == 0.0 && st.momx == 0.0 &&
                        st.momy == 0.0 && st.momz == 0.0);
    out.eact[e] = (in.oeact[o] & 1) | (hist ? 2 : 0);
  }
  // the far-entry key: a kept entry keeps its own (the key bounds the
  // entry's gap through the bodies' accumulated motion, which a rebuild does
  // not reset); a new entry is evaluated at its first step
  out.ekey[e] = o >= 0 ? in.oekey[o] : -INFINITY;
}

// --- CSR -----------------------------------------------------------------
// Counts of a sorted key array without atomics (a wall owning thousands of
// entries would serialise them): the last entry of each run writes the run's
// length.  Its first index is found by 8-way searches, each round's 8 loads
// independent: probes 1..8 spans back from the end, the span growing 8-fold
// until a probe leaves the run, then the interval holding the run's start
// cut into 8 per round (a few rounds of load latency for a run of tens of
// entries, against a binary search's chain of log2(n) dependent loads).
__global__ void k_count_sorted(int n, const int32_t* key, int32_t* cnt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int k = key[e];
  if (e + 1 < n && key[e + 1] == k) return;
  // invariant: key[hi] == k, and lo < hi is -1 or an index whose key is not k
  int hi = e, lo = -1;
  for (long span = 1;; span *= 8) {
    int out = 9;
#pragma unroll
    for (int q = 8; q >= 1; q--) {
      const long x = (long)e - span * q;
      if (x < 0 || key[x] != k) out = q;
    }
    if (out <= 8) {
      hi = (int)((long)e - span * (out - 1));
      const long l = (long)e - span * out;
      lo = l < -1 ? -1 : (int)l;
      break;
    }
  }
  while (hi - lo > 1) {
    const int st = (hi - lo + 7) / 8;
    int nlo = lo, nhi = hi;
#pragma unroll
    for (int q = 1; q < 8; q++) {
      const int x = lo + st * q;
      if (x < hi) {
        if (key[x] == k)
          nhi = x < nhi ? x : nhi;
        else
          nlo = x > nlo ? x : nlo;
      }
    }
    lo = nlo;
    hi = nhi;
  }
  cnt[k] = e - hi + 1;
}
__global__ void k_iota(int n, int32_t* v) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) v[e] = e;
}

// --- resultant slots -------------------------------------------------------
// compute_resultants (Rockable.cpp:3267-3272) walks the bodies in index order
// and each body's interactions in key order, adding to Particles[i] and
// subtracting from Particles[j]; so body b receives, in this order, the
// entries (a, b) with a < b, then its own entries (b, c), then the entries
// (a, b) with a > b.  Body b's slots are [soff[b], soff[b+1]) with
// soff[b] = istart[b] + jstart[b]; nlow[b] counts b's j-side entries whose
// index is below b's own range.
__global__ void k_slot_body(int nb, const int32_t* istart, const int32_t* jstart, const int32_t* jlist,
                            int32_t* soff, int32_t* nlow) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  soff[b] = istart[b] + jstart[b];
  if (b == nb) return;
  int lo = jstart[b], hi = jstart[b + 1];
  const int is = istart[b];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (jlist[mid] < is)
      lo = mid + 1;
    else
      hi = mid;
  }
  nlow[b] = lo - jstart[b];
}
__global__ void k_slot_i(int n, const int32_t* ei, const int32_t* istart, const int32_t* soff, const int32_t* nlow,
                         int32_t* spos) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int b = ei[e];
  spos[2 * e] = soff[b] + nlow[b] + (e - istart[b]);
}
__global__ void k_slot_j(int n, const int32_t* ej, const int32_t* jlist, const int32_t* jstart,
                         const int32_t* istart, const int32_t* soff, int32_t* spos) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int e = jlist[q];
  const int b = ej[e];
  const int above = e >= istart[b] ? istart[b + 1] - istart[b] : 0;
  spos[2 * e + 1] = soff[b] + (q - jstart[b]) + above;
}
// slot activity of a fresh list: the force phase rewrites it only on change
__global__ void k_slot_act(int n, const int32_t* spos, const uint8_t* eact, uint8_t* sact) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint8_t a = eact[e] & 1;
  sact[spos[2 * e]] = a;
  sact[spos[2 * e + 1]] = a;
}
__global__ void k_heavy_flag(int nb, const int32_t* soff, int heavy_min, int32_t* flag) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) flag[b] = (soff[b + 1] - soff[b]) > heavy_min ? 1 : 0;
}
__global__ void k_heavy_list(int nb, const int32_t* flag, const int32_t* pos, int32_t* heavy) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && flag[b]) heavy[pos[b]] = b;
}

// Chunks of 64 slots of the heavy bodies, and the count of heavy bodies that
// are not passive.  One workgroup: thread 0 lays out the chunk offsets
// (a few thousand heavy bodies at most), then the threads fill the chunks.
// The heavy count is the flag scan's last value plus the last flag; out
// receives the chunk count, the count of heavy bodies that are not passive
// and the heavy count.
__global__ void k_heavy_chunks(const int32_t* heavy, const int32_t* pos_last, const int32_t* flag_last,
                               const int32_t* soff, const uint8_t* passive, int32_t* hch_off, int2* hch, int cap,
                               int32_t* out) {
  const int nheavy = *pos_last + *flag_last;
  if (threadIdx.x == 0) {
    int off = 0, np = 0;
    for (int h = 0; h < nheavy; h++) {
      const int b = heavy[h];
      hch_off[h] = off;
      off += (soff[b + 1] - soff[b] + 63) / 64;
      np += passive[b] ? 0 : 1;
    }
    hch_off[nheavy] = off;
    out[0] = off;
    out[1] = np;
    out[2] = nheavy;
  }
  __syncthreads();
  for (int h = threadIdx.x; h < nheavy; h += blockDim.x) {
    const int b = heavy[h];
    const int k0 = soff[b], k1 = soff[b + 1];
    for (int c = hch_off[h], k = k0; c < hch_off[h + 1] && c < cap; c++, k += 64) hch[c] = make_int2(k, min(k + 64, k1));
  }
}

inline int nblk(long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <class T>
static int ensure(NLWork& w, T*& p, size_t& cap, size_t n) {
  if (n <= cap && p) return 0;
  if (p) (void)hipFree(p);
  size_t c = n + n / 2 + 256;
  if (hipMalloc(&p, c * sizeof(T)) != hipSuccess) {
    snprintf(w.err, sizeof(w.err), "hipMalloc of %zu bytes failed", c * sizeof(T));
    p = nullptr;
    cap = 0;
    return -1;
  }
  cap = c;
  return 0;
}

static int scan_exclusive(NLWork& w, const int32_t* in, int32_t* out, int n, hipStream_t s) {
  size_t bytes = 0;
  RK_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), s));
  if (ensure(w, w.tmp, w.tmp_cap, bytes) != 0) return -1;
  RK_CHECK(rocprim::exclusive_scan((void*)w.tmp, bytes, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), s));
  return 0;
}

template <class K, class Vv>
static int sort_pairs(NLWork& w, const K* kin, K* kout, const Vv* vin, Vv* vout, int n, int bits, hipStream_t s) {
  size_t bytes = 0;
  RK_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
  if (ensure(w, w.tmp, w.tmp_cap, bytes) != 0) return -1;
  RK_CHECK(rocprim::radix_sort_pairs((void*)w.tmp, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)bits, s));
  return 0;
}

static int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

int nl_build_csr(NLWork& w, int nb, EntryArrays& E, const uint8_t* passive, hipStream_t s) {
  const int n = E.n;
  if (ensure(w, E.istart, E.istart_cap, (size_t)nb + 1) || ensure(w, E.jstart, E.jstart_cap, (size_t)nb + 1) ||
      ensure(w, E.jlist, E.jlist_cap, (size_t)(n > 0 ? n : 1)) || ensure(w, w.i32a, w.i32a_cap, (size_t)nb + 1) ||
      ensure(w, w.i32b, w.i32b_cap, (size_t)(n > 0 ? n : 1)) || ensure(w, w.i32c, w.i32c_cap, (size_t)(n > 0 ? n : 1)) ||
      ensure(w, E.edesc, E.edesc_cap, (size_t)(n > 0 ? n : 1)))
    return -1;
  // istart: counts by owner, exclusive scan over nb + 1 (trailing zero)
  RK_CHECK(hipMemsetAsync(w.i32a, 0, sizeof(int32_t) * (nb + 1), s));
  if (n > 0) hipLaunchKernelGGL(k_count_sorted, dim3(nblk(n)), dim3(kBlock), 0, s, n, E.ei, w.i32a);
  if (scan_exclusive(w, w.i32a, E.istart, nb + 1, s)) return -1;
  // jlist: stable radix sort of entry indices by j
  if (n > 0) {
    hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(kBlock), 0, s, n, w.i32b);
    RK_CHECK(hipMemcpyAsync(w.i32c, E.ej, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    if (ensure(w, w.kout, w.kout_cap, (size_t)n)) return -1;
    if (sort_pairs(w, (const uint32_t*)w.i32c, (uint32_t*)w.kout, w.i32b, E.jlist, n,
                   bits_for((uint64_t)(nb > 1 ? nb - 1 : 1)), s))
      return -1;
  }
  // jstart: counts of the sorted j keys (a wall is the j of thousands of
  // entries: counted by atomics, they would serialise on its counter)
  RK_CHECK(hipMemsetAsync(w.i32a, 0, sizeof(int32_t) * (nb + 1), s));
  if (n > 0) hipLaunchKernelGGL(k_count_sorted, dim3(nblk(n)), dim3(kBlock), 0, s, n, (const int32_t*)w.kout, w.i32a);
  if (scan_exclusive(w, w.i32a, E.jstart, nb + 1, s)) return -1;
  // resultant slots
  if (ensure(w, E.soff, E.soff_cap, (size_t)nb + 1)) return -1;
  hipLaunchKernelGGL(k_slot_body, dim3(nblk(nb + 1)), dim3(kBlock), 0, s, nb, E.istart, E.jstart, E.jlist, E.soff,
                     w.i32a);
  if (n > 0) {
    hipLaunchKernelGGL(k_slot_i, dim3(nblk(n)), dim3(kBlock), 0, s, n, E.ei, E.istart, E.soff, w.i32a, E.spos);
    hipLaunchKernelGGL(k_slot_j, dim3(nblk(n)), dim3(kBlock), 0, s, n, E.ej, E.jlist, E.jstart, E.istart, E.soff,
                       E.spos);
    hipLaunchKernelGGL(k_slot_act, dim3(nblk(n)), dim3(kBlock), 0, s, n, E.spos, E.eact, E.sact);
  }
  RK_CHECK(hipGetLastError());
  return nl_build_heavy(w, nb, E, passive, s);
}

// Reads of small device values through the pinned staging buffer: copy_in
// queues the copies (byte offsets into the buffer), wait synchronises the
// stream and copy_out takes the values.
constexpr size_t kPinBytes = 8192;
static char* pin_buf(NLWork& w) {
  if (!w.pin && hipHostMalloc((void**)&w.pin, kPinBytes, hipHostMallocDefault) != hipSuccess) w.pin = nullptr;
  return w.pin;
}
static int read_small(NLWork& w, hipStream_t s, std::initializer_list<std::pair<void*, std::pair<const void*, size_t>>> xs) {
  char* P = pin_buf(w);
  size_t off = 0;
  for (const auto& x : xs) {
    if (P && off + x.second.second <= kPinBytes)
      RK_CHECK(hipMemcpyAsync(P + off, x.second.first, x.second.second, hipMemcpyDeviceToHost, s));
    else
      RK_CHECK(hipMemcpyAsync(x.first, x.second.first, x.second.second, hipMemcpyDeviceToHost, s));
    off += (x.second.second + 7) & ~(size_t)7;
  }
  RK_CHECK(hipStreamSynchronize(s));
  off = 0;
  for (const auto& x : xs) {
    if (P && off + x.second.second <= kPinBytes) memcpy(x.first, P + off, x.second.second);
    off += (x.second.second + 7) & ~(size_t)7;
  }
  return 0;
}

int nl_build_heavy(NLWork& w, int nb, EntryArrays& E, const uint8_t* passive, hipStream_t s) {
  E.nheavy = 0;
  E.nhch = 0;
  E.nheavy_np = 0;
  if (nb <= 0 || !E.soff) return 0;
  if (ensure(w, E.heavy, E.heavy_cap, (size_t)nb) || ensure(w, w.i32a, w.i32a_cap, (size_t)nb + 1) ||
      ensure(w, w.i32c, w.i32c_cap, (size_t)nb + 1) || ensure(w, w.counter, w.counter_cap, 4))
    return -1;
  hipLaunchKernelGGL(k_heavy_flag, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, E.soff, w.heavy_min, w.i32a);
  if (scan_exclusive(w, w.i32a, w.i32c, nb, s)) return -1;
  hipLaunchKernelGGL(k_heavy_list, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, w.i32a, w.i32c, E.heavy);
  // chunk table sized as if every body were heavy: at most (slots of all
  // bodies) / 64 + one per body, so that the heavy count is read once, with
  // the chunks'
  const size_t cap = (size_t)E.cap * 2 / 64 + (size_t)nb + 1;
  if (ensure(w, E.hch, E.hch_cap, cap) || ensure(w, E.hch_off, E.hch_off_cap, (size_t)nb + 1)) return -1;
  if (E.hpart_cap < cap) {
    if (E.hpart) (void)hipFree(E.hpart);
    E.hpart = nullptr;
    E.hpart_cap = 0;
    if (hipMalloc(&E.hpart, cap * 6 * sizeof(double)) != hipSuccess) {
      snprintf(w.err, sizeof(w.err), "out of device memory (heavy-body partial sums)");
      return -1;
    }
    E.hpart_cap = cap;
  }
  hipLaunchKernelGGL(k_heavy_chunks, dim3(1), dim3(kBlock), 0, s, E.heavy, w.i32c + nb - 1, w.i32a + nb - 1, E.soff,
                     passive, E.hch_off, E.hch, (int)cap, w.counter);
  int32_t cnt[3] = {0, 0, 0};
  if (read_small(w, s, {{cnt, {w.counter, 3 * sizeof(int32_t)}}})) return -1;
  if ((size_t)cnt[0] > cap) {
    snprintf(w.err, sizeof(w.err), "heavy-body chunk table overflow");
    return -1;
  }
  E.nhch = cnt[0];
  E.nheavy_np = cnt[1];
  E.nheavy = cnt[2];
  return 0;
}

int nl_rebuild(NLWork& w, Dev& d, const NLParams& p, EntryArrays& cur, EntryArrays& nxt, hipStream_t s) {
  const int nb = d.nb;
  if (nb == 0) return 0;
  const bool cells = p.linkcells != 0;
  if (ensure(w, w.bobb, w.bobb_cap, (size_t)15 * nb)) return -1;
  if (ensure(w, w.baabb, w.baabb_cap, (size_t)6 * nb)) return -1;
  hipLaunchKernelGGL(k_obb, dim3(nblk(nb)), dim3(kBlock), 0, s, d, w.bobb, w.baabb);
  const double half = 0.5 * p.DVerlet;

  // ---- broad phase: candidate pairs
  if (ensure(w, w.counter, w.counter_cap, 4) || ensure(w, w.scnt, w.scnt_cap, (size_t)kStripes * kStripeStride) ||
      ensure(w, w.soff, w.soff_cap, (size_t)kStripes))
    return -1;
  int nc = 0;
  for (int attempt = 0; attempt < 3; attempt++) {
    // striped output: kStripes regions of rcap pairs
    const int rcap = std::max(64, (w.cand_hint + kStripes - 1) / kStripes);
    if (ensure(w, w.cstripe, w.cstripe_cap, (size_t)kStripes * rcap)) return -1;
    RK_CHECK(hipMemsetAsync(w.counter, 0, sizeof(int) * 4, s));
    RK_CHECK(hipMemsetAsync(w.scnt, 0, sizeof(int) * kStripes * kStripeStride, s));
    if (!cells) {
      const int jmax = nb - std::min(nb, std::max(1, d.nDriven));
      const int chunks = std::max(1, (nb + kBruteChunk - 1) / kBruteChunk);
      (void)jmax;
      hipLaunchKernelGGL(k_broad_brute, dim3(nblk(nb), chunks), dim3(kBlock), 0, s, d, w.bobb, half, w.cstripe,
                         w.scnt, rcap);
    } else {
      // bounding box of the bodies [first, nb) (Rockable.cpp:1683-1688, 3943-3947)
      int first = (p.boxForLinkCellsOpt == 1) ? d.nDriven : 0;
      if (first >= nb) first = 0;
      const int pb = 64;
      if (ensure(w, w.f64a, w.f64a_cap, (size_t)6 * pb)) return -1;
      hipLaunchKernelGGL(k_box_partial, dim3(pb), dim3(kBlock), 0, s, first, nb, w.baabb, w.f64a);
      std::vector<double> part(6 * pb);
      if (read_small(w, s, {{part.data(), {w.f64a, sizeof(double) * 6 * pb}}})) return -1;
      double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int q = 0; q < pb; q++)
        for (int k = 0; k < 3; k++) {
          mn[k] = std::min(mn[k], part[6 * q + k]);
          mx[k] = std::max(mx[k], part[6 * q + 3 + k]);
        }
      Grid g;
      const double ms[3] = {p.cellMinSizes[0], p.cellMinSizes[1], p.cellMinSizes[2]};
      long ncell = 1;
      for (int k = 0; k < 3; k++) {
        const double len = mx[k] - mn[k];
        int n = (ms[k] > 0.0) ? (int)floor(len / ms[k]) : 1;
        g.N[k] = n < 1 ? 1 : n;
        g.mn[k] = mn[k];
        g.sz[k] = len / g.N[k];
        ncell *= g.N[k];
      }
      if (ncell > (1L << 28)) {
        snprintf(w.err, sizeof(w.err), "linkCells: %ld cells (cellMinSizes too small for the domain)", ncell);
        return -1;
      }
      if (ensure(w, w.cell_of, w.cell_of_cap, (size_t)nb) || ensure(w, w.ccount, w.ccount_cap, (size_t)ncell + 1) ||
          ensure(w, w.cstart, w.cstart_cap, (size_t)ncell + 1) || ensure(w, w.cfill, w.cfill_cap, (size_t)ncell) ||
          ensure(w, w.cbody, w.cbody_cap, (size_t)nb) || ensure(w, w.over, w.over_cap, (size_t)nb))
        return -1;
      RK_CHECK(hipMemsetAsync(w.ccount, 0, sizeof(int32_t) * (ncell + 1), s));
      RK_CHECK(hipMemsetAsync(w.cfill, 0, sizeof(int32_t) * ncell, s));
      unsigned long long* reach = reinterpret_cast<unsigned long long*>(w.f64a);
      RK_CHECK(hipMemsetAsync(reach, 0, sizeof(unsigned long long) * 3, s));
      hipLaunchKernelGGL(k_cell_assign, dim3(nblk(nb)), dim3(kBlock), 0, s, d, w.bobb, w.baabb, half, g, w.cell_of,
                         w.ccount, w.over, w.counter + 1, reach);
      if (scan_exclusive(w, w.ccount, w.cstart, (int)ncell + 1, s)) return -1;
      hipLaunchKernelGGL(k_cell_fill, dim3(nblk(nb)), dim3(kBlock), 0, s, nb, w.cell_of, w.cstart, w.cfill, w.cbody);
      int hv[2];
      if (read_small(w, s, {{hv, {w.counter + 1, sizeof(int)}}})) return -1;
      const int nover = hv[0];
      const int ncb = nb - nover;
      hipLaunchKernelGGL(k_broad_cells, dim3(nblk((long)kCellLanes * nb)), dim3(kBlock), 0, s, d, w.bobb, half, g, w.cell_of, w.cstart,
                         w.cbody, w.cstripe, w.scnt, rcap);
      if (nover > 0 && ncb > 0) {
        // enough workgroups to fill the chip when one body (a wall or a plane
        // over the whole domain) meets every cell: each takes every split-th
        // column of cells its box covers
        const int split = std::max(1, std::min(4096, (8192 + nover - 1) / nover));
        hipLaunchKernelGGL(k_broad_over, dim3(nover, split), dim3(kBlock), 0, s, d, w.bobb, half, g, w.over, nover,
                           w.cstart, w.cbody, reach, w.cstripe, w.scnt, rcap);
      }
    }
    hipLaunchKernelGGL(k_stripe_scan, dim3(1), dim3(kStripes), 0, s, w.scnt, w.soff, w.counter + 2);
    int tm[2] = {0, 0};
    if (read_small(w, s, {{tm, {w.counter + 2, sizeof(int) * 2}}})) return -1;
    nc = tm[0];
    if (tm[1] <= rcap) {
      if (ensure(w, w.cand, w.cand_cap, (size_t)std::max(nc, 1))) return -1;
      hipLaunchKernelGGL(k_stripe_gather, dim3((rcap + kBlock - 1) / kBlock, kStripes), dim3(kBlock), 0, s, w.cstripe,
                         rcap, w.scnt, w.soff, w.cand);
      break;
    }
    w.cand_hint = kStripes * (tm[1] + tm[1] / 4 + 64);
    if (attempt == 2) {
      snprintf(w.err, sizeof(w.err), "broad phase overflow");
      return -1;
    }
  }
  w.last_candidates = nc;

  // ---- narrow phase
  if (ensure(w, w.cbase, w.cbase_cap, (size_t)std::max(nc, 1)) || ensure(w, w.cnf, w.cnf_cap, (size_t)std::max(nc, 1)) ||
      ensure(w, w.cnb, w.cnb_cap, (size_t)std::max(nc, 1)))
    return -1;
  std::function<int(int)> launch_narrow;
  int narrow_rcap = 0;
  if (nc > 0) {
    NarrowCfg cfg;
    cfg.maxv = p.maxv;
    cfg.maxe = p.maxe;
    cfg.maxf = p.maxf;
    cfg.maxwords = p.maxwords;
    cfg.dmax = p.dVerlet;
    // the bodies' edge and face boxes for the culls
    if (ensure(w, w.gedge, w.gedge_cap, (size_t)kGate * std::max(d.netot, 1)) ||
        ensure(w, w.gface, w.gface_cap, (size_t)kGate * std::max(d.nftot, 1)))
      return -1;
    int glg = 3;
    while (glg < 6 && (1 << glg) < p.maxe + p.maxf) glg++;
    hipLaunchKernelGGL(k_gate_boxes, dim3((int)((((long)d.nb << glg) + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, d,
                       w.gedge, w.gface, glg);
    cfg.gedge = w.gedge;
    cfg.gface = w.gface;
    // candidate pairs per wavefront: by default 2 for shapes of up to 8
    // vertices, else 1 (measured: 2 pairs are faster for sphero-cubes, 1 for
    // hulls of 8-14 vertices, whose passes fill a wavefront; more pairs cost
    // occupancy, their values being per lane), then as many as the LDS holds
    const size_t stride = (narrow_lds_bytes(p) + 7) & ~(size_t)7;
    int P = w.narrow_pairs;
    if (P <= 0) P = 2 * p.maxv <= 16 ? 2 : 1;
    while (P > 1 && (size_t)P * stride > 65536) P >>= 1;
    const size_t lds = (size_t)P * stride;
    // one run of the narrow phase into stripes of rcap keys; the stripe
    // totals (all, largest) land in counter[2..3]
    launch_narrow = [&, cfg, P, stride, lds](int rcap) -> int {
      if (ensure(w, w.nearbuf, w.nearbuf_cap, (size_t)kStripes * rcap)) return -1;
      RK_CHECK(hipMemsetAsync(w.scnt, 0, sizeof(int) * kStripes * kStripeStride, s));
      // pairs of point shapes one per lane; the others (all of them without
      // point shapes) one or two per wavefront
      int nrest = nc;
      const int32_t* clist = nullptr;
      if (p.has_points) {
        if (ensure(w, w.pflag, w.pflag_cap, (size_t)nc) || ensure(w, w.ppos, w.ppos_cap, (size_t)nc) ||
            ensure(w, w.plist, w.plist_cap, (size_t)nc))
          return -1;
        hipLaunchKernelGGL(k_narrow_points, dim3(nblk(nc)), dim3(kBlock), 0, s, d, w.bobb, w.baabb, cfg, nc, w.cand,
                           w.nearbuf, w.scnt, rcap, w.cbase, w.cnf, w.cnb, w.pflag);
        if (scan_exclusive(w, w.pflag, w.ppos, nc, s)) return -1;
        hipLaunchKernelGGL(k_heavy_list, dim3(nblk(nc)), dim3(kBlock), 0, s, nc, w.pflag, w.ppos, w.plist);
        int last[2] = {0, 0};
        if (read_small(w, s, {{&last[0], {w.ppos + nc - 1, sizeof(int)}}, {&last[1], {w.pflag + nc - 1, sizeof(int)}}}))
          return -1;
        nrest = last[0] + last[1];
        clist = w.plist;
      }
      const dim3 grid((nrest + P - 1) / P);
      if (nrest > 0 && P == 2)
        hipLaunchKernelGGL(k_narrow<2>, grid, dim3(64), lds, s, d, w.bobb, w.baabb, cfg, nrest, w.cand, w.nearbuf,
                           w.scnt, rcap, w.cbase, w.cnf, w.cnb, (int)stride, clist);
      else if (nrest > 0)
        hipLaunchKernelGGL(k_narrow<1>, grid, dim3(64), lds, s, d, w.bobb, w.baabb, cfg, nrest, w.cand, w.nearbuf,
                           w.scnt, rcap, w.cbase, w.cnf, w.cnb, (int)stride, clist);
      RK_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_stripe_scan, dim3(1), dim3(kStripes), 0, s, w.scnt, w.soff, w.counter + 2);
      return 0;
    };
    // the stripes' overflow is checked with the merge's totals (one read);
    // on overflow the phase runs again with room, before the merge counts
    narrow_rcap = std::max(64, (w.near_hint + kStripes - 1) / kStripes);
    if (launch_narrow(narrow_rcap)) return -1;
  }

  // ---- old groups
  const int nold = cur.n;
  int ng_old = 0;
  if (ensure(w, w.i32b, w.i32b_cap, (size_t)std::max(nold, 1) + 1) ||
      ensure(w, w.i32c, w.i32c_cap, (size_t)std::max(nold, 1) + 1) ||
      ensure(w, w.okey, w.okey_cap, (size_t)std::max(nold, 1)) ||
      ensure(w, w.ostart, w.ostart_cap, (size_t)std::max(nold, 1) + 1))
    return -1;
  if (nold > 0) {
    hipLaunchKernelGGL(k_old_heads, dim3(nblk(nold)), dim3(kBlock), 0, s, nold, cur.ei, cur.ej, w.i32b);
    if (scan_exclusive(w, w.i32b, w.i32c, nold, s)) return -1;
    hipLaunchKernelGGL(k_old_compact, dim3(nblk(nold)), dim3(kBlock), 0, s, nold, cur.ei, cur.ej, w.i32b, w.i32c,
                       w.okey, w.ostart);
    int last[2];
    if (read_small(w, s, {{&last[0], {w.i32c + nold - 1, sizeof(int)}}, {&last[1], {w.i32b + nold - 1, sizeof(int)}}}))
      return -1;
    ng_old = last[0] + last[1];
  }
  RK_CHECK(hipMemcpyAsync(w.ostart + ng_old, &nold, sizeof(int), hipMemcpyHostToDevice, s));

  // ---- unique directed groups
  const int ngk = 2 * nc + ng_old;
  int ng = 0;
  if (ensure(w, w.gk, w.gk_cap, (size_t)std::max(ngk, 1)) || ensure(w, w.gk2, w.gk2_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.gv, w.gv_cap, (size_t)std::max(ngk, 1)) || ensure(w, w.gv2, w.gv2_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.gflag, w.gflag_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.gpos, w.gpos_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.ukey, w.ukey_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.ucand, w.ucand_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.uold, w.uold_cap, (size_t)std::max(ngk, 1)) ||
      ensure(w, w.gcount, w.gcount_cap, (size_t)std::max(ngk, 1) + 1) ||
      ensure(w, w.goff, w.goff_cap, (size_t)std::max(ngk, 1) + 1))
    return -1;
  if (ngk > 0) {
    int bb = 1;
    while ((1ll << bb) < (long long)nb) bb++;
    hipLaunchKernelGGL(k_fill_groups, dim3(nblk(nc + ng_old)), dim3(kBlock), 0, s, nc, w.cand, ng_old, w.okey, w.gk,
                       w.gv, bb);
    if (sort_pairs(w, w.gk, w.gk2, w.gv, w.gv2, ngk, 2 * bb, s)) return -1;
    hipLaunchKernelGGL(k_group_heads, dim3(nblk(ngk)), dim3(kBlock), 0, s, ngk, w.gk2, w.gflag);
    if (scan_exclusive(w, w.gflag, w.gpos, ngk, s)) return -1;
    hipLaunchKernelGGL(k_group_compact, dim3(nblk(ngk)), dim3(kBlock), 0, s, ngk, w.gk2, w.gv2, w.gflag, w.gpos,
                       w.ukey, w.ucand, w.uold, bb);
  }

  // ---- merge: count, scan, allocate, write
  MergeIn in;
  in.ng = ngk;
  in.ng_pos = ngk > 0 ? w.gpos + ngk - 1 : nullptr;
  in.ng_flag = ngk > 0 ? w.gflag + ngk - 1 : nullptr;
  in.ukey = w.ukey;
  in.ucand = w.ucand;
  in.uold = w.uold;
  in.ostart = w.ostart;
  in.cbase = w.cbase;
  in.cnf = w.cnf;
  in.cnb = w.cnb;
  in.nearbuf = w.nearbuf;
  in.bobb = w.bobb;
  in.dmax = p.dVerlet;
  in.oi = cur.ei;
  in.oj = cur.ej;
  in.osub = cur.esub;
  in.oest = cur.est;
  in.oedn = cur.edn;
  in.ogeo = cur.egeo;
  in.odamp = cur.edamp;
  in.ostick = cur.estick;
  in.oeact = cur.eact;
  in.oekey = cur.ekey;
  MergeOut out;
  out.count = w.gcount;
  out.offset = w.goff;
  int nnew = 0;
  if (ngk > 0) {
    // count pass over the groups' upper bound, its scan, then one read: the
    // group count, the merged length and the narrow phase's stripe totals
    auto count_pass = [&]() -> int {
      hipLaunchKernelGGL(k_merge<false>, dim3(nblk((long)ngk + 1)), dim3(kBlock), 0, s, d, in, out);
      if (scan_exclusive(w, w.gcount, w.goff, ngk + 1, s)) return -1;
      hipLaunchKernelGGL(k_merge_total, dim3(1), dim3(1), 0, s, in.ng_pos, in.ng_flag, w.goff, w.counter);
      return 0;
    };
    if (count_pass()) return -1;
    int rd[4] = {0, 0, 0, 0};
    if (read_small(w, s, {{rd, {w.counter, 4 * sizeof(int)}}})) return -1;
    if (launch_narrow && rd[3] > narrow_rcap) {
      for (int attempt = 0;; attempt++) {
        if (attempt == 2) {
          snprintf(w.err, sizeof(w.err), "narrow phase overflow");
          return -1;
        }
        w.near_hint = kStripes * (rd[3] + rd[3] / 4 + 64);
        narrow_rcap = std::max(64, (w.near_hint + kStripes - 1) / kStripes);
        if (launch_narrow(narrow_rcap)) return -1;
        if (read_small(w, s, {{rd + 2, {w.counter + 2, 2 * sizeof(int)}}})) return -1;
        if (rd[3] <= narrow_rcap) break;
      }
      in.nearbuf = w.nearbuf;
      if (count_pass()) return -1;
      if (read_small(w, s, {{rd, {w.counter, 2 * sizeof(int)}}})) return -1;
    }
    ng = rd[0];
    nnew = rd[1];
  }
  in.ng = ng;
  in.ng_pos = in.ng_flag = nullptr;
  if (entries_reserve(w, nxt, nnew) != 0) return -1;
  nxt.n = nnew;
  if (ng > 0 && nnew > 0) {
    out.ei = nxt.ei;
    out.ej = nxt.ej;
    out.esub = nxt.esub;
    out.est = nxt.est;
    out.edn = nxt.edn;
    out.egeo = nxt.egeo;
    out.edamp = nxt.edamp;
    out.estick = nxt.estick;
    out.eact = nxt.eact;
    out.ekey = nxt.ekey;
    if (ensure(w, w.msrc, w.msrc_cap, (size_t)nnew)) return -1;
    out.src = w.msrc;
    hipLaunchKernelGGL(k_merge<true>, dim3(nblk(ng)), dim3(kBlock), 0, s, d, in, out);
    hipLaunchKernelGGL(k_merge_fill, dim3(nblk(nnew)), dim3(kBlock), 0, s, d, nnew, in, out);
  }
  RK_CHECK(hipGetLastError());
  if (nl_build_csr(w, nb, nxt, d.passive, s) != 0) return -1;
  RK_CHECK(hipStreamSynchronize(s));
  return 0;
}

int entries_reserve(NLWork& w, EntryArrays& E, int n) {
  size_t need = (size_t)(n > 0 ? n : 1);
  if (need <= E.cap && E.ei) return 0;
  entries_free(E);
  size_t c = need + need / 2 + 1024;
  bool ok = hipMalloc(&E.ei, c * sizeof(int32_t)) == hipSuccess && hipMalloc(&E.ej, c * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&E.esub, c * sizeof(uint32_t)) == hipSuccess && hipMalloc(&E.est, c * sizeof(EState)) == hipSuccess &&
            hipMalloc(&E.edn, c * sizeof(double)) == hipSuccess &&
            hipMalloc(&E.egeo, c * sizeof(EGeom)) == hipSuccess && hipMalloc(&E.edamp, c * sizeof(double)) == hipSuccess &&
            hipMalloc(&E.estick, c * sizeof(int32_t)) == hipSuccess && hipMalloc(&E.eact, c * sizeof(uint8_t)) == hipSuccess &&
            hipMalloc(&E.ekey, c * sizeof(double)) == hipSuccess &&
            hipMalloc(&E.spos, c * 2 * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&E.sact, c * 2 * sizeof(uint8_t)) == hipSuccess &&
            hipMalloc(&E.sval, c * 12 * sizeof(double)) == hipSuccess &&
            hipMalloc(&E.akin, (c + 512) * sizeof(ActiveKin)) == hipSuccess &&
            hipMalloc(&E.acount, (c / 64 + 2) * sizeof(int32_t)) == hipSuccess;
  if (!ok) {
    snprintf(w.err, sizeof(w.err), "hipMalloc of %zu interaction entries failed", c);
    entries_free(E);
    return -1;
  }
  E.cap = c;
  return 0;
}

void entries_free(EntryArrays& E) {
  void* ps[] = {E.ei, E.ej, E.esub, E.est, E.edn, E.egeo, E.edamp, E.estick, E.eact, E.ekey, E.spos, E.sact, E.sval,
                E.akin, E.acount};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  E.ekey = nullptr;
  E.ei = E.ej = nullptr;
  E.esub = nullptr;
  E.est = nullptr;
  E.edn = nullptr;
  E.egeo = nullptr;
  E.edamp = nullptr;
  E.estick = nullptr;
  E.eact = nullptr;
  E.spos = nullptr;
  E.sact = nullptr;
  E.sval = nullptr;
  E.akin = nullptr;
  E.acount = nullptr;
  E.cap = 0;
}

void nl_work_free(NLWork& w) {
  void* ps[] = {w.pflag, w.ppos, w.plist,
                w.cstripe, w.scnt, w.soff, w.bobb,  w.baabb, w.counter, w.cand,  w.cbase, w.cnf,   w.cnb,   w.nearbuf, w.okey,  w.ostart,
                w.gk,    w.gk2,   w.gv,      w.gv2,   w.gflag, w.gpos,  w.ukey,  w.ucand,   w.uold,  w.gcount,
                w.goff,  w.tmp,   w.i32a,    w.i32b,  w.i32c,  w.f64a,  w.cell_of, w.ccount, w.cstart, w.cfill,
                w.cbody, w.over, w.gedge, w.gface, w.kout, w.msrc};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (w.pin) (void)hipHostFree(w.pin);
  NLWork z;
  w = z;
}

}  // namespace rk
