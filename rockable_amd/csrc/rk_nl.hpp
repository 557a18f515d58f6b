// rk_nl.hpp -- neighbour-list rebuild: entry arrays and work buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "rk_device.hpp"

namespace rk {

// One sorted interaction list (see rk_device.hpp for the layout).
struct EntryArrays {
  int n = 0;
  size_t cap = 0;
  int32_t* ei = nullptr;
  int32_t* ej = nullptr;
  uint32_t* esub = nullptr;
  EState* est = nullptr;
  double* edn = nullptr;
  EGeom* egeo = nullptr;
  double* edamp = nullptr;
  int32_t* estick = nullptr;
  uint8_t* eact = nullptr;
  // far-entry skip: for an inactive entry with zero history, its last gap
  // plus both bodies' motion bounds at that step (-inf: evaluate)
  double* ekey = nullptr;
  int32_t* spos = nullptr;   // 2 per entry
  uint8_t* sact = nullptr;   // 2 per entry
  double* sval = nullptr;    // 12 per entry
  ActiveKin* akin = nullptr; // 1 per entry (+512)
  int32_t* acount = nullptr; // 1 per geometry block (64 entries or more)
  int4* edesc = nullptr;     // 1 per entry: vertex indices of its geometry
  size_t edesc_cap = 0;
  int32_t* istart = nullptr;
  size_t istart_cap = 0;
  int32_t* jstart = nullptr;
  size_t jstart_cap = 0;
  int32_t* jlist = nullptr;
  size_t jlist_cap = 0;
  int32_t* soff = nullptr;
  size_t soff_cap = 0;
  int32_t* heavy = nullptr;
  size_t heavy_cap = 0;
  int nheavy = 0;
  // heavy bodies' slots in chunks of 64: chunk c covers slots [hch[c].x,
  // hch[c].y), heavy body h its chunks [hch_off[h], hch_off[h + 1]); hpart
  // holds each chunk's six partial sums
  int2* hch = nullptr;
  int32_t* hch_off = nullptr;
  double* hpart = nullptr;
  size_t hch_cap = 0, hch_off_cap = 0, hpart_cap = 0;
  int nhch = 0;
  int nheavy_np = 0;  // heavy bodies that are not passive
};

struct NLParams {
  double DVerlet = 0.0, dVerlet = 0.0;
  int linkcells = 0;
  double cellMinSizes[3] = {1.0, 1.0, 1.0};
  int boxForLinkCellsOpt = 0;
  int maxv = 1;      // max vertices per shape
  int maxe = 0;      // max edges per shape
  int maxf = 0;      // max faces per shape
  int maxwords = 1;  // max 64-bit words of predicate bits per body pair
  int has_points = 0;  // some shape is a point (one vertex, no edge or face)
};

// LDS of one narrow-phase pair: both bodies' OBBs and boxes, the global
// vertices of both bodies, the predicate bits, the lists of gated vertices, edges and faces, the prefix
// of the bit counts per word.
inline size_t narrow_lds_bytes(const NLParams& p) {
  return sizeof(double) * (48 + 6 * p.maxv) + sizeof(unsigned long long) * p.maxwords +
         sizeof(int) * (3 * p.maxv + 2 * p.maxe + 2 * p.maxf + 8 + p.maxwords);
}

struct NLWork {
  char err[256] = {0};
  // pinned host staging of the small device-to-host reads (counts, the box
  // partials): a pageable destination makes the runtime stage each copy
  char* pin = nullptr;
  int cand_hint = 0, near_hint = 0, last_candidates = 0;
  int heavy_min = 256;  // slot count above which a body's sum runs block-wide
  int narrow_pairs = 0;  // candidate pairs per narrow-phase wavefront (1, 2; 0: by shape size)
#define RK_BUF(T, name) \
  T* name = nullptr;    \
  size_t name##_cap = 0;
  RK_BUF(double, bobb)
  RK_BUF(double, baabb)
  RK_BUF(int32_t, pflag)  // narrow phase: candidates that are not point pairs, their prefix, their list
  RK_BUF(int32_t, ppos)
  RK_BUF(int32_t, plist)
  RK_BUF(float, gedge)
  RK_BUF(float, gface)
  RK_BUF(int32_t, kout)
  RK_BUF(int, counter)
  RK_BUF(uint64_t, cand)
  RK_BUF(uint64_t, cstripe)
  RK_BUF(int32_t, scnt)
  RK_BUF(int32_t, soff)
  RK_BUF(int32_t, cbase)
  RK_BUF(int32_t, cnf)
  RK_BUF(int32_t, cnb)
  RK_BUF(uint32_t, nearbuf)
  RK_BUF(uint64_t, okey)
  RK_BUF(int32_t, ostart)
  RK_BUF(uint64_t, gk)
  RK_BUF(uint64_t, gk2)
  RK_BUF(uint32_t, gv)
  RK_BUF(uint32_t, gv2)
  RK_BUF(int32_t, gflag)
  RK_BUF(int32_t, gpos)
  RK_BUF(uint64_t, ukey)
  RK_BUF(int32_t, ucand)
  RK_BUF(int32_t, uold)
  RK_BUF(int32_t, gcount)
  RK_BUF(int32_t, goff)
  RK_BUF(int32_t, msrc)  // merge: per new entry, the old entry it keeps (-1 for a fresh one)
  RK_BUF(unsigned char, tmp)
  RK_BUF(int32_t, i32a)
  RK_BUF(int32_t, i32b)
  RK_BUF(int32_t, i32c)
  RK_BUF(double, f64a)
  RK_BUF(int32_t, cell_of)
  RK_BUF(int32_t, ccount)
  RK_BUF(int32_t, cstart)
  RK_BUF(int32_t, cfill)
  RK_BUF(int32_t, cbody)
  RK_BUF(int32_t, over)
#undef RK_BUF
};

// Rebuild the list: reads `cur` (sorted), writes `nxt` (sorted, with CSR).
int nl_rebuild(NLWork& w, Dev& d, const NLParams& p, EntryArrays& cur, EntryArrays& nxt, hipStream_t s);
// istart / jstart / jlist, resultant slots and the heavy-body list of a
// sorted list.
int nl_build_csr(NLWork& w, int nb, EntryArrays& E, const uint8_t* passive, hipStream_t s);
// The heavy-body list alone (after a change of w.heavy_min).
int nl_build_heavy(NLWork& w, int nb, EntryArrays& E, const uint8_t* passive, hipStream_t s);
int entries_reserve(NLWork& w, EntryArrays& E, int n);
void entries_free(EntryArrays& E);
void nl_work_free(NLWork& w);

}  // namespace rk
