// rk_step.hip -- the per-step kernels of the velocity-Verlet DEM step on
// CDNA4 (gfx950), FP64 throughout, no floating-point atomics.
//
//   k_predict   Rockable::velocityVerletStep predictor (Rockable.cpp:1979-2055)
//               + ServoFunction (DrivingSystem.cpp:242-268) and
//               velocityControlledDrive (Rockable.cpp:1856-1900)
//   k_geom      update_interactions (Rockable.cpp:2968-2974 ->
//               Interaction.cpp:147-326) fused with build_activeInteractions
//               (3005-3029): contact geometry of every listed entry; the
//               active ones go to a compact list
//   k_law       ForceLaw::computeInteraction (3035-3046, ForceLaw_*.cpp) on
//               the active list; writes each active entry's terms for body i
//               and body j (incrementResultants, 2752-2789) into the two
//               resultant slots of the entry
//   k_gather    initialise_particle_forces_and_moments (2869-2929) +
//               compute_resultants (3267-3272) as a per-body ordered sum over
//               the body's slots, which hold its incident entries in the
//               serial order of the reference, + compute_accelerations_from_
//               resultants (3419-3552) + the closing kick (2120-2155)
//   k_gather_heavy  the same for bodies with many slots, one workgroup each
//   k_breakage  breakage_of_interfaces (3350-3413)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>

#include "rk_device.hpp"
#include "rk_model.hpp"

namespace rk {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ v3 kpos(const BodyKin& b) { return V(b.px, b.py, b.pz); }
__device__ __forceinline__ v3 kvel(const BodyKin& b) { return V(b.vx, b.vy, b.vz); }
__device__ __forceinline__ v3 kvrot(const BodyKin& b) { return V(b.wx, b.wy, b.wz); }
__device__ __forceinline__ qt kq(const BodyKin& b) {
  qt q;
  q.w = b.qw;
  q.x = b.qx;
  q.y = b.qy;
  q.z = b.qz;
  return q;
}
__device__ __forceinline__ v3 ld3(const double* p) { return V(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(double* p, v3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// A rebuild found due by the device guard suspends the step kernels until
// the host has rebuilt the list (every thread reads the same flag).
__device__ __forceinline__ bool frozen(const Dev& d) {
  return d.guard_on && *(volatile const int*)&d.guard->frozen;
}

// Particle::GlobVertex (Particle.cpp:55-57): pos + Q * (homothety * vertex)
__device__ __forceinline__ v3 glob_vertex(const DevShapes& sh, const BodyKin& b, int v) {
  const double* p = sh.vtx + 3 * (sh.voff[b.shape] + v);
  v3 hv = vsc(b.h, V(p[0], p[1], p[2]));
  return vadd(kpos(b), qrot(kq(b), hv));
}

// v(Qj) - v(Qi) (Interaction.cpp:162)
__device__ __forceinline__ v3 rel_vel(v3 pos, const BodyKin& Pi, const BodyKin& Pj) {
  v3 a = vsub(kvel(Pj), vcross(vsub(pos, kpos(Pj)), kvrot(Pj)));
  v3 b = vsub(kvel(Pi), vcross(vsub(pos, kpos(Pi)), kvrot(Pi)));
  return vsub(a, b);
}

__device__ __forceinline__ double prm(const DevParams& p, int id, int g1, int g2) {
  if (g1 < 0 || g2 < 0 || g1 >= p.G || g2 >= p.G) return 0.0;
  return p.tab[(id * p.G + g1) * p.G + g2];
}

struct Kin {
  v3 n, pos, vel;
  double dn, prev_dn;
  bool deact;
};

// The global vertices of a body, computed once per step by k_glob.
__device__ __forceinline__ v3 gvtx(const double* g, int v) { return V(g[3 * v], g[3 * v + 1], g[3 * v + 2]); }

// even-odd crossing test of P against a face of body j (Interaction.cpp:236-253):
// its nbv vertices are gv[bj + fidx[fs + k]]
__device__ bool face_crossing(const Dev& d, int fs, int nbv, int bj, v3 P, v3 v1, v3 v2) {
  int odd = 0;
  double ori1 = vdot(P, v1), ori2 = vdot(P, v2);
  const int* fv = d.sh.fidx + fs;
  v3 A = gvtx(d.gv, bj + fv[0]);
  for (int iva = 0; iva < nbv; iva++) {
    int ivb = iva + 1;
    if (ivb == nbv) ivb = 0;
    v3 B = gvtx(d.gv, bj + fv[ivb]);
    double pa1 = vdot(A, v1), pb1 = vdot(B, v1);
    double pa2 = vdot(A, v2), pb2 = vdot(B, v2);
    if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
      if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
    }
    A = B;
  }
  return odd != 0;
}

// Per entry, the global-vertex indices (and face range) its contact geometry
// reads, resolved once per list change rather than through the bodies'
// shapes at every step (which costs a chain of dependent loads):
//   VV (vi, vj, -, -); VE (vi, j's edge ends); EE (i's edge ends, j's edge
//   ends); VF (vi, start of j's face in sh.fidx, its vertex count, bvoff[j]).
__global__ void __launch_bounds__(kBlock) k_entry_desc(Dev d) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.ne; e += gridDim.x * blockDim.x) {
    const int i = d.ei[e], j = d.ej[e];
    const uint32_t sub = d.esub[e];
    const int type = (int)(sub >> 30), isub = (int)((sub >> 15) & 0x7fffu), jsub = (int)(sub & 0x7fffu);
    const int bi = d.bvoff[i], bj = d.bvoff[j];
    const int shi = d.kin[i].shape, shj = d.kin[j].shape;
    int4 q = make_int4(bi + isub, 0, 0, 0);
    if (type == VV) {
      q.y = bj + jsub;
    } else if (type == VE) {
      const int* ed = d.sh.edge + 2 * (d.sh.eoff[shj] + jsub);
      q.y = bj + ed[0];
      q.z = bj + ed[1];
    } else if (type == VF) {
      const int gface = d.sh.foff[shj] + jsub;
      q.y = d.sh.fstart[gface];
      q.z = d.sh.fstart[gface + 1] - q.y;
      q.w = bj;
    } else {
      const int* edi = d.sh.edge + 2 * (d.sh.eoff[shi] + isub);
      const int* edj = d.sh.edge + 2 * (d.sh.eoff[shj] + jsub);
      q = make_int4(bi + edi[0], bi + edi[1], bj + edj[0], bj + edj[1]);
    }
    d.edesc[e] = q;
  }
}

// Interaction::UpdateDispatcher (Interaction.cpp:147-326), on the entry's
// resolved vertex indices q (k_entry_desc).
// WITH_VEL: also the relative velocity at the contact point (Interaction.cpp:162).
template <bool WITH_VEL>
__device__ void update_kinematics(const Dev& d, int type, int4 q, const BodyKin& Pi, const BodyKin& Pj, Kin& k) {
  const double Ri = Pi.R, Rj = Pj.R;
  if (type == VV) {
    v3 posi = gvtx(d.gv, q.x);
    v3 posj = gvtx(d.gv, q.y);
    k.n = vsub(posi, posj);
    k.dn = vnormalize(&k.n) - (Ri + Rj);
    k.pos = vsub(posi, vsc(Ri + 0.5 * k.dn, k.n));
    if (WITH_VEL) k.vel = rel_vel(k.pos, Pi, Pj);
  } else if (type == VE) {
    v3 posj1 = gvtx(d.gv, q.y);
    v3 posj2 = gvtx(d.gv, q.z);
    v3 posi = gvtx(d.gv, q.x);
    v3 E = vsub(posj2, posj1);
    v3 v = vsub(posi, posj1);
    double r = vdot(v, E) / vdot(E, E);
    if (r <= 0.0 || r >= 1.0) {
      k.deact = true;
      return;
    }
    k.n = vsub(posi, vadd(posj1, vsc(r, E)));
    k.dn = vnormalize(&k.n) - (Ri + Rj);
    k.pos = vsub(posi, vsc(Ri + 0.5 * k.dn, k.n));
    if (WITH_VEL) k.vel = rel_vel(k.pos, Pi, Pj);
  } else if (type == VF) {
    const int fs = q.y, nbv = q.z, bj = q.w;
    const int* fv = d.sh.fidx + fs;
    v3 A = gvtx(d.gv, bj + fv[0]);
    v3 B = gvtx(d.gv, bj + fv[1]);
    v3 C = gvtx(d.gv, bj + fv[nbv - 1]);
    v3 posi = gvtx(d.gv, q.x);
    v3 v = vsub(posi, A);
    v3 v1 = vsub(B, A);
    vnormalize(&v1);
    v3 v2 = vsub(C, A);
    k.n = vcross(v1, v2);
    vnormalize(&k.n);
    double dist = vdot(v, k.n);
    if (dist < 0.0) {
      k.n = vneg(k.n);
      dist = -dist;
    }
    v3 P = vsub(posi, vsc(dist, k.n));
    v2 = vcross(k.n, v1);
    if (face_crossing(d, fs, nbv, bj, P, v1, v2)) {
      k.dn = dist - (Ri + Rj);
      k.pos = vsub(posi, vsc(Ri + 0.5 * k.dn, k.n));
      if (WITH_VEL) k.vel = rel_vel(k.pos, Pi, Pj);
    } else {
      k.deact = true;  // n keeps the new face normal, pos and vel stay stale
    }
  } else {
    v3 posi1 = gvtx(d.gv, q.x);
    v3 posi2 = gvtx(d.gv, q.y);
    v3 posj1 = gvtx(d.gv, q.z);
    v3 posj2 = gvtx(d.gv, q.w);
    v3 Ei = vsub(posi2, posi1);
    v3 Ej = vsub(posj2, posj1);
    v3 v = vsub(posi1, posj1);
    double c = vdot(Ei, Ei), dd = vdot(Ej, Ej), e = vdot(Ei, Ej);
    double f = (c * dd) - (e * e);
    if (fabs(f) > 1.0e-12) {
      f = 1.0 / f;
      double a = vdot(Ei, v), b = vdot(Ej, v);
      double s = (e * b - a * dd) * f;
      double t = (c * b - e * a) * f;
      if (s <= 0.0 || s >= 1.0 || t <= 0.0 || t >= 1.0) {
        k.deact = true;
        return;
      }
      k.n = vsub(vadd(posi1, vsc(s, Ei)), vadd(posj1, vsc(t, Ej)));
      k.dn = vnormalize(&k.n) - (Ri + Rj);
      k.pos = vsub(vadd(posi1, vsc(s, Ei)), vsc(Ri + 0.5 * k.dn, k.n));
      if (WITH_VEL) k.vel = rel_vel(k.pos, Pi, Pj);
    } else {
      k.deact = true;
    }
  }
}

struct Hist {
  double fn;
  v3 ft, mom;
};

// ForceLaw_Default::computeInteraction (ForceLaw_Default.cpp:57-131), FT_CORR off
__device__ void law_default(const Dev& d, const StepParams& sp, const Kin& k, double damp, const BodyKin& Pi,
                            const BodyKin& Pj, Hist& h) {
  if (k.dn > 0.0) {
    h.fn = 0.0;
    h.ft = V(0, 0, 0);
    h.mom = V(0, 0, 0);
    return;
  }
  const int g1 = Pi.group, g2 = Pj.group;
  const double kn = prm(d.par, P_KN, g1, g2), kt = prm(d.par, P_KT, g1, g2), mu = prm(d.par, P_MU, g1, g2);
  const double kr = prm(d.par, P_KR, g1, g2), mur = prm(d.par, P_MUR, g1, g2);
  double vn = vdot(k.vel, k.n);
  double fne = -kn * k.dn;
  double fnv = damp * vn;
  h.fn = fne + fnv;
  if (h.fn < 0.0) h.fn = 0.0;
  v3 vt = vsub(k.vel, vsc(vn, k.n));
  h.ft = vadd(h.ft, vsc(kt, vsc(sp.dt, vt)));
  double threshold = fabs(mu * fne);
  double ft_square = vdot(h.ft, h.ft);
  if (ft_square > 0.0 && ft_square >= threshold * threshold) h.ft = vsc(threshold / sqrt(ft_square), h.ft);
  if (kr > 0.0) {
    h.mom = vadd(h.mom, vsc(sp.dt, vsc(kr, vsub(kvrot(Pj), kvrot(Pi)))));
    double threshold_mom = fabs(mur * h.fn);
    double mom_square = vdot(h.mom, h.mom);
    if (mom_square > 0.0 && mom_square >= threshold_mom * threshold_mom)
      h.mom = vsc(threshold_mom / sqrt(mom_square), h.mom);
  }
}

// ForceLaw_Avalanche::computeInteraction (ForceLaw_Avalanche.cpp:57-136)
__device__ void law_avalanche(const Dev& d, const StepParams& sp, const Kin& k, int i, const BodyKin& Pi,
                              const BodyKin& Pj, Hist& h) {
  if (k.dn > 0.0) {
    h.fn = 0.0;
    h.ft = V(0, 0, 0);
    h.mom = V(0, 0, 0);
    return;
  }
  const int g1 = Pi.group, g2 = Pj.group;
  const double kn = prm(d.par, P_KN, g1, g2), en2 = prm(d.par, P_EN2, g1, g2), kt = prm(d.par, P_KT, g1, g2);
  const double mu = prm(d.par, P_MU, g1, g2), kr = prm(d.par, P_KR, g1, g2), mur = prm(d.par, P_MUR, g1, g2);
  double prev_dn = k.prev_dn;
  if (prev_dn > 0.0) prev_dn = 0.0;
  double delta_Dn = k.dn - prev_dn;
  if (delta_Dn > 0.0)
    h.fn = -kn * en2 * k.dn;
  else if (delta_Dn < 0.0)
    h.fn += -kn * delta_Dn;
  if (h.fn < 0.0) h.fn = 0.0;
  v3 vt = vsub(k.vel, vsc(vdot(k.vel, k.n), k.n));
  h.ft = vadd(h.ft, vsc(kt, vsc(sp.dt, vt)));
  double threshold_ft = fabs(mu * h.fn);
  double ft_square = vdot(h.ft, h.ft);
  if (ft_square > 0.0 && ft_square > threshold_ft * threshold_ft)
    h.ft = vsc(threshold_ft / sqrt(ft_square), h.ft);
  h.mom = vadd(h.mom, vsc(sp.dt, vsc(kr, vsub(kvrot(Pj), kvrot(Pi)))));
  const BodyKin& P = (i < d.nDriven) ? Pj : Pi;
  v3 branch = vsub(k.pos, kpos(P));
  double r = vdot(branch, kvrot(P)) / vdot(kvrot(P), kvrot(P));
  branch = vsub(branch, vsc(r, kvrot(P)));
  double threshold_mom = fabs(mur * sqrt(vnorm2(branch)) * h.fn);
  double mom_square = vdot(h.mom, h.mom);
  if (mom_square > 0.0 && mom_square > threshold_mom * threshold_mom)
    h.mom = vsc(threshold_mom / sqrt(mom_square), h.mom);
}

// ForceLaw_StickedLinks::computeInteraction (ForceLaw_StickedLinks.cpp:75-215)
__device__ void law_sticked(const Dev& d, const StepParams& sp, const Kin& k, double damp, int stick, int i, int j,
                            const BodyKin& Pi, const BodyKin& Pj, Hist& h, bool paramsInInterfaces) {
  if (stick >= 0) {
    double kn = 0.0, kt = 0.0, kr = 0.0, fn0 = 1.0, ft0 = 1.0, mom0 = 1.0, power = 1.0;
    const double* bp = d.ifc.par + 8 * stick;
    double dn0 = bp[7];
    const int g1 = Pi.group, g2 = Pj.group;
    if (paramsInInterfaces) {
      kn = bp[0];
      kt = bp[1];
      kr = bp[2];
      fn0 = bp[3];
      ft0 = bp[4];
      mom0 = bp[5];
      power = bp[6];
    } else if (d.cluster[i] == d.cluster[j]) {
      kn = prm(d.par, P_KN_IN, g1, g2);
      kt = prm(d.par, P_KT_IN, g1, g2);
      kr = prm(d.par, P_KR_IN, g1, g2);
      fn0 = prm(d.par, P_FN0_IN, g1, g2);
      ft0 = prm(d.par, P_FT0_IN, g1, g2);
      mom0 = prm(d.par, P_MOM0_IN, g1, g2);
      power = prm(d.par, P_POW_IN, g1, g2);
      dn0 = 0.0;
    } else {
      kn = prm(d.par, P_KN_OUT, g1, g2);
      kt = prm(d.par, P_KT_OUT, g1, g2);
      kr = prm(d.par, P_KR_OUT, g1, g2);
      fn0 = prm(d.par, P_FN0_OUT, g1, g2);
      ft0 = prm(d.par, P_FT0_OUT, g1, g2);
      mom0 = prm(d.par, P_MOM0_OUT, g1, g2);
      power = prm(d.par, P_POW_OUT, g1, g2);
    }
    double vn = vdot(k.vel, k.n);
    double fne = -kn * (k.dn - dn0);
    double fnv = damp * vn;
    h.fn = fne + fnv;
    v3 vt = vsub(k.vel, vsc(vn, k.n));
    h.ft = vadd(h.ft, vsc(kt, vsc(sp.dt, vt)));
    v3 ftv = vsc(damp, vt);
    h.ft = vadd(h.ft, ftv);
    h.mom = vadd(h.mom, vsc(sp.dt, vsc(kr, vsub(kvrot(Pj), kvrot(Pi)))));
    // power 2 (every generated deck, the interfaces' default): x * x, which
    // is what glibc's pow(x, 2.0) returns for every x (exhaustive sampling
    // in tests/test_pow_square.py), so the break decision is the host's bit
    // for bit; another power goes through the device pow (within an ulp of
    // glibc's, the decision differing only for a yield within an ulp of 0)
    const double rt = sqrt(vnorm2(h.ft)) / ft0, rm = sqrt(vnorm2(h.mom)) / mom0;
    const double pt = power == 2.0 ? rt * rt : pow(rt, power);
    const double pm = power == 2.0 ? rm * rm : pow(rm, power);
    double yield = pt + pm - h.fn / fn0 - 1.0;
    if (yield > 0.0) d.ifc.brk[stick] = 1;  // idempotent byte store, consumed after the resultants
  } else {
    if (k.dn > 0.0) {
      h.fn = 0.0;
      h.ft = V(0, 0, 0);
      h.mom = V(0, 0, 0);
      return;
    }
    const int g1 = Pi.group, g2 = Pj.group;
    const double kn = prm(d.par, P_KN, g1, g2), kt = prm(d.par, P_KT, g1, g2), mu = prm(d.par, P_MU, g1, g2);
    const double kr = prm(d.par, P_KR, g1, g2), mur = prm(d.par, P_MUR, g1, g2);
    double vn = vdot(k.vel, k.n);
    double fne = -kn * k.dn;
    double fnv = damp * vn;
    h.fn = fne + fnv;
    if (h.fn < 0.0) h.fn = 0.0;
    v3 vt = vsub(k.vel, vsc(vn, k.n));
    h.ft = vadd(h.ft, vsc(kt, vsc(sp.dt, vt)));
    double threshold = fabs(mu * h.fn);
    double ft_square = vdot(h.ft, h.ft);
    if (ft_square > 0.0 && ft_square >= threshold * threshold) h.ft = vsc(threshold / sqrt(ft_square), h.ft);
    h.mom = vadd(h.mom, vsc(sp.dt, vsc(kr, vsub(kvrot(Pj), kvrot(Pi)))));
    double threshold_mom = fabs(mur * h.fn);
    double mom_square = vdot(h.mom, h.mom);
    if (mom_square > 0.0 && mom_square >= threshold_mom * threshold_mom)
      h.mom = vsc(threshold_mom / sqrt(mom_square), h.mom);
  }
}

// The geometry pass of one listed entry e (Interaction::UpdateDispatcher,
// Rockable.cpp:2934-2977): its gap, and for an active entry its kinematics in
// a (returns true); an inactive entry gets its gap and activity byte here,
// and with skipping its far-entry key.
template <bool GEOM_ALL>
__device__ __forceinline__ bool geom_entry(const Dev& d, int e, bool skipping, ActiveKin& a) {
  bool act = false;
  const int i = d.ei[e], j = d.ej[e];
  const int type = (int)(d.esub[e] >> 30);
  const BodyKin& Pi = d.kin[i];
  const BodyKin& Pj = d.kin[j];
  const int stick = d.ifc.n > 0 ? d.estick[e] : -1;  // bonds exist only with interfaces
  Kin k;
  k.deact = false;
  k.prev_dn = 0.0;  // the previous dn is read by the law, for active entries only
  if (stick >= 0) {  // a bond may run on stale geometry after a deactivation
    const EGeom& g = d.egeo[e];
    k.n = V(g.nx, g.ny, g.nz);
    k.pos = V(g.px, g.py, g.pz);
  }
  update_kinematics<false>(d, type, d.edesc[e], Pi, Pj, k);
  if (k.deact) k.dn = 0.0;  // Interaction::deactivate (Interaction.cpp:123-128)
  act = (k.dn < 0.0) || (stick >= 0);
  if (act) {
    a.nx = k.n.x;
    a.ny = k.n.y;
    a.nz = k.n.z;
    a.px = k.pos.x;
    a.py = k.pos.y;
    a.pz = k.pos.z;
    a.dn = k.dn;
    a.prev_dn = 0.0;  // (unused: the law reads the previous dn itself)
    a.e = e;
    a.deact = k.deact ? 1 : 0;
    a.i = i;
    a.j = j;
  } else {
    // activity byte: bit 0 active, bit 1 history maybe non-zero; the
    // history and the slot flags are written only when they change
    const uint8_t old = d.eact[e];
    uint8_t now = old & 2;
    if (k.deact) {
      if (old & 2) {
        EState o;
        o.fn = 0.0;
        o.ftx = o.fty = o.ftz = 0.0;
        o.momx = o.momy = o.momz = 0.0;
        d.est[e] = o;
        now = 0;
      }
      d.edn[e] = 0.0;
      if (skipping) d.ekey[e] = -INFINITY;
    } else {
      d.edn[e] = k.dn;  // stale fn/ft/mom are kept (Rockable.cpp:3022)
      if (skipping && now == 0) {
        // the key of a zero-history entry (a body without a finite bound
        // never lets its entries be skipped)
        const double ci = d.cbound[i], cj = d.cbound[j];
        d.ekey[e] = (ci < INFINITY && cj < INFINITY) ? k.dn + ci + cj : -INFINITY;
      }
      if (GEOM_ALL) {
        const v3 vel = rel_vel(k.pos, Pi, Pj);
        EGeom g;
        g.nx = k.n.x;
        g.ny = k.n.y;
        g.nz = k.n.z;
        g.px = k.pos.x;
        g.py = k.pos.y;
        g.pz = k.pos.z;
        g.vx = vel.x;
        g.vy = vel.y;
        g.vz = vel.z;
        d.egeo[e] = g;
      }
    }
    if (now != old) {
      d.eact[e] = now;
      if (old & 1) {
        const int2 slots = reinterpret_cast<const int2*>(d.spos)[e];
        d.sact[slots.x] = 0;
        d.sact[slots.y] = 0;
      }
    }
  }
  return act;
}

// The force phase: geometry, then law, as two kernels or fused in one.
//  geometry  every listed entry: contact geometry from the body poses
//            (Interaction::UpdateDispatcher without the relative velocity).
//            An inactive entry gets its new dn (or the zeroed history of a
//            deactivation) and cleared activity flags; an active one (dn < 0
//            or bonded) keeps its geometry for the law.
//  law       every active entry: relative velocity, the force law, the
//            history and the two resultant slots.
// Two kernels: k_geom writes the active entries of geometry block b to
// akin[b * BS, b * BS + acount[b]) (ballot + block prefix, no atomics) and
// k_law gives a wavefront to each block's list.  Fused: k_interact keeps
// them in LDS.
//
// The geometry pass of one block of BS entries (block index blk): returns
// whether this thread's entry is active, with its kinematics in a.  Every
// thread of the block calls it (it synchronises the block).
template <bool GEOM_ALL, int BS, bool SORT = true>
__device__ __forceinline__ bool geom_pass(const Dev& d, int blk, int (*wtype)[5], short* perm, ActiveKin& a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // SORT: the block's entries are taken in the order of their contact type
  // (a counting sort in LDS; entries past the end last), so that a wavefront
  // runs one branch of update_kinematics instead of all four; without it
  // each thread takes its own entry, one dependent load and two barriers
  // sooner, its wavefront running the branches of the types it holds.
  // The far-entry skip (only when no geometry of inactive entries is
  // written): an inactive entry with zero history whose gap is proven
  // positive is left as it is; it sorts past the end.
  const bool skipping = !GEOM_ALL && d.skip_far;
  if (SORT) {
    const int e0 = blk * BS + (int)threadIdx.x;
    int ty = e0 < d.ne ? (int)(d.esub[e0] >> 30) : 4;
    if (skipping && e0 < d.ne && d.eact[e0] == 0 &&
        d.ekey[e0] - d.cbound[d.ei[e0]] - d.cbound[d.ej[e0]] > kSkipMargin)
      ty = 4;
    unsigned long long mine = 0ull;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const unsigned long long m = __ballot(ty == t);
      if (ty == t) mine = m;
      if (lane == 0) wtype[wave][t] = __popcll(m);
    }
    __syncthreads();
    int rank = 0;
    for (int t = 0; t < 5; t++)
      for (int w = 0; w < BS / 64; w++) rank += (t < ty || (t == ty && w < wave)) ? wtype[w][t] : 0;
    rank += __popcll(mine & ((1ull << lane) - 1ull));
    perm[rank] = (short)threadIdx.x;
    __syncthreads();
  }
  const int e = blk * BS + (SORT ? (int)perm[threadIdx.x] : (int)threadIdx.x);
  bool act = false;
  bool skip = false;
  if (skipping && e < d.ne && d.eact[e] == 0)
    skip = d.ekey[e] - d.cbound[d.ei[e]] - d.cbound[d.ej[e]] > kSkipMargin;
  if (d.skip_count) {  // counts for the measurement of the skip (off in production)
    const unsigned long long ms = __ballot(skip), mz = __ballot(e < d.ne && d.eact[e] == 0),
                             ml = __ballot(e < d.ne);
    if (lane == 0) {
      atomicAdd(d.skip_count, (unsigned long long)__popcll(ms));
      atomicAdd(d.skip_count + 1, (unsigned long long)__popcll(mz));
      atomicAdd(d.skip_count + 2, (unsigned long long)__popcll(ml));
    }
  }
  if (e < d.ne && !skip) act = geom_entry<GEOM_ALL>(d, e, skipping, a);
  return act;
}

// Rank of this thread's active entry among the block's (ballot + LDS
// prefix, no atomics) and the block's count of active entries.
template <int BS>
__device__ __forceinline__ int active_rank(bool act, int* wcount, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(act);
  if (lane == 0) wcount[wave] = __popcll(m);
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < BS / 64; w++) {
    base += w < wave ? wcount[w] : 0;
    tot += wcount[w];
  }
  *total = tot;
  return base + __popcll(m & ((1ull << lane) - 1ull));
}

template <bool GEOM_ALL, int BS>
__global__ void __launch_bounds__(BS) k_geom(Dev d) {
  __shared__ int wcount[BS / 64];
  __shared__ int wtype[BS / 64][5];
  __shared__ short perm[BS];
  if (frozen(d)) return;
  for (TileLoop L = xcd_tiles((d.ne + BS - 1) / BS); L.t < L.end; L.t += L.step) {
    const int blk = L.t;
    ActiveKin a;
    const bool act = geom_pass<GEOM_ALL, BS>(d, blk, wtype, perm, a);
    int total = 0;
    const int r = active_rank<BS>(act, wcount, &total);
    if (act) d.akin[(size_t)blk * BS + r] = a;
    if (threadIdx.x == 0) d.acount[blk] = total;
    __syncthreads();
  }
}

// The force law of one active entry (kinematics in a): history, both
// resultant slots, activity flags.
template <int LAW>
__device__ __forceinline__ void law_item(const Dev& d, const StepParams& sp, int write_geom_all, int pii,
                                         const ActiveKin& a) {
  const int e = a.e;
  const int i = a.i, j = a.j;
  const BodyKin Pi = d.kin[i];
  const BodyKin Pj = d.kin[j];
  const int stick = d.estick[e];
  Kin k;
  k.n = V(a.nx, a.ny, a.nz);
  k.pos = V(a.px, a.py, a.pz);
  k.dn = a.dn;
  // the previous step's dn (the Avalanche law's only; a deactivation zeroes it)
  k.prev_dn = (LAW == LAW_AVALANCHE && !a.deact) ? d.edn[e] : 0.0;
  k.deact = a.deact != 0;
  Hist h;
  if (k.deact) {  // a bond keeps its stale velocity; the history restarts
    const EGeom& g = d.egeo[e];
    k.vel = V(g.vx, g.vy, g.vz);
    h.fn = 0.0;
    h.ft = V(0, 0, 0);
    h.mom = V(0, 0, 0);
  } else {
    k.vel = rel_vel(k.pos, Pi, Pj);
    const EState s = d.est[e];
    h.fn = s.fn;
    h.ft = V(s.ftx, s.fty, s.ftz);
    h.mom = V(s.momx, s.momy, s.momz);
  }
  const double damp = d.edamp[e];
  if (LAW == LAW_AVALANCHE)
    law_avalanche(d, sp, k, i, Pi, Pj, h);
  else if (LAW == LAW_STICKED)
    law_sticked(d, sp, k, damp, stick, i, j, Pi, Pj, h, pii != 0);
  else
    law_default(d, sp, k, damp, Pi, Pj, h);
  EState o;
  o.fn = h.fn;
  o.ftx = h.ft.x;
  o.fty = h.ft.y;
  o.ftz = h.ft.z;
  o.momx = h.mom.x;
  o.momy = h.mom.y;
  o.momz = h.mom.z;
  d.est[e] = o;
  d.edn[e] = k.dn;
  // incrementResultants (Rockable.cpp:2756-2764), both sides
  v3 f = vadd(vsc(h.fn, k.n), h.ft);
  v3 Ci = vsub(k.pos, kpos(Pi));
  v3 Cj = vsub(k.pos, vadd(kpos(Pj), V(0.0, 0.0, 0.0)));
  v3 mi = vadd(vcross(Ci, f), h.mom);
  v3 mj = vsub(vcross(Cj, vneg(f)), h.mom);
  // the two resultant slots: body i adds (f, mi), body j adds (-f, mj);
  // x + (-y) is x - y exactly, so the serial sums are unchanged
  const int2 slots = reinterpret_cast<const int2*>(d.spos)[e];
  double* ci = d.sval + 6 * (size_t)slots.x;
  double* cj = d.sval + 6 * (size_t)slots.y;
  st3(ci, f);
  st3(ci + 3, mi);
  st3(cj, vneg(f));
  st3(cj + 3, mj);
  // Geometry is persistent for bonds (a deactivated bond keeps its stale n,
  // pos, vel); for other entries it is written when an output is due.
  if (stick >= 0 || (write_geom_all && !k.deact)) {
    EGeom g;
    g.nx = k.n.x;
    g.ny = k.n.y;
    g.nz = k.n.z;
    g.px = k.pos.x;
    g.py = k.pos.y;
    g.pz = k.pos.z;
    g.vx = k.vel.x;
    g.vy = k.vel.y;
    g.vz = k.vel.z;
    d.egeo[e] = g;
  }
  const uint8_t old = d.eact[e];
  if (old != 3) {
    d.eact[e] = 3;
    if (!(old & 1)) {
      d.sact[slots.x] = 1;
      d.sact[slots.y] = 1;
    }
  }
}

// Each wavefront takes W (at most kLawGroup) consecutive geometry blocks
// and spreads their active entries over its lanes, so that lanes are not left
// idle by blocks with few active entries.
constexpr int kLawGroup = 8;

template <int LAW>
__global__ void __launch_bounds__(kBlock, LAW == LAW_STICKED ? 2 : 4) k_law(Dev d, StepParams sp, int write_geom_all, int pii, int ngeom, int W) {
  if (frozen(d)) return;
  const int waves = (ngeom + W - 1) / W;
  for (TileLoop L = xcd_tiles((waves + kBlock / 64 - 1) / (kBlock / 64)); L.t < L.end; L.t += L.step) {
  const int g0 = (int)((L.t * (unsigned)blockDim.x + threadIdx.x) >> 6) * W;  // first geometry block
  if (g0 >= ngeom) continue;
  const int lane = threadIdx.x & 63;
  // inclusive prefix sums of the active counts of the W blocks (lanes < W)
  int incl = (lane < W && g0 + lane < ngeom) ? d.acount[g0 + lane] : 0;
  for (int o = 1; o < kLawGroup; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  int pre[kLawGroup];
#pragma unroll
  for (int m = 0; m < kLawGroup; m++) pre[m] = __shfl(incl, m, 64);
  const int total = pre[kLawGroup - 1];
  for (int t = lane; t < total; t += 64) {
    int gb = 0, before = 0;
#pragma unroll
    for (int m = 0; m < kLawGroup; m++)
      if (pre[m] <= t) {
        gb = m + 1;
        before = pre[m];
      }
    const ActiveKin a = d.akin[(size_t)(g0 + gb) * d.geom_block + (t - before)];
    law_item<LAW>(d, sp, write_geom_all, pii, a);
  }
  }
}

// Geometry and force law fused: the block's active entries go to LDS instead
// of the global active list, and the block's first threads apply the law to
// them after the geometry pass (no second launch, no round trip of the
// active kinematics through memory).
template <int LAW, bool GEOM_ALL, bool SORT>
__global__ void __launch_bounds__(kBlock, LAW == LAW_STICKED ? 2 : 4) k_interact(Dev d, StepParams sp, int write_geom_all, int pii) {
  __shared__ int wcount[kBlock / 64];
  __shared__ int wtype[kBlock / 64][5];
  __shared__ short perm[kBlock];
  __shared__ ActiveKin sa[kBlock];
  if (frozen(d)) return;
  for (TileLoop L = xcd_tiles((d.ne + kBlock - 1) / kBlock); L.t < L.end; L.t += L.step) {
    ActiveKin a;
    const bool act = geom_pass<GEOM_ALL, kBlock, SORT>(d, L.t, wtype, perm, a);
    int total = 0;
    const int r = active_rank<kBlock>(act, wcount, &total);
    if (act) sa[r] = a;
    __syncthreads();
    if ((int)threadIdx.x < total) law_item<LAW>(d, sp, write_geom_all, pii, sa[threadIdx.x]);
    __syncthreads();
  }
}

// The far-entry skip with compaction: a workgroup takes SEG consecutive
// entries, tests each against the motion bounds (coalesced reads of the
// activity byte, the key and the pair; the bounds are per body and stay in
// L2), gathers the entries still to evaluate into an LDS list, and runs the
// geometry and the force law on that list, each lane taking the law of its
// own entry.  In a settled packing about one entry in ten is evaluated and
// most of those are active: the list fills whole wavefronts, where the
// per-entry test of k_interact leaves most lanes of every wavefront idle for
// the length of the busiest lane's chain.  The
// entries are independent (each writes its own state and its own two
// resultant slots), so the order of the list changes no result.
template <int LAW, int SEG>
__global__ void __launch_bounds__(kBlock, LAW == LAW_STICKED ? 2 : 4) k_interact_seg(Dev d, StepParams sp, int pii) {
  constexpr int M = SEG / kBlock;
  __shared__ int wcnt[M][kBlock / 64];
  __shared__ int list[SEG];
  if (frozen(d)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long cnt[3] = {0ull, 0ull, 0ull};  // the wavefront's counts (count_skips)
  for (TileLoop L = xcd_tiles((d.ne + SEG - 1) / SEG); L.t < L.end; L.t += L.step) {
    const int e0 = L.t * SEG + (int)threadIdx.x;
    bool need[M];
#pragma unroll
    for (int r = 0; r < M; r++) {
      // the activity byte, key and pair together (one round trip), then the bounds
      const int e = e0 + r * kBlock;
      bool nd = e < d.ne;
      if (nd) {
        const uint8_t ea = d.eact[e];
        const double key = d.ekey[e];
        const int bi = d.ei[e], bj = d.ej[e];
        nd = !(ea == 0 && key - d.cbound[bi] - d.cbound[bj] > kSkipMargin);
      }
      need[r] = nd;
    }
    if (d.skip_count) {  // counts for the measurement of the skip (off in production)
#pragma unroll
      for (int r = 0; r < M; r++) {
        const int e = e0 + r * kBlock;
        cnt[0] += __popcll(__ballot(e < d.ne && !need[r]));
        cnt[1] += __popcll(__ballot(e < d.ne && d.eact[e] == 0));
        cnt[2] += __popcll(__ballot(e < d.ne));
      }
    }
    // rank the entries to evaluate, in entry order: every wavefront's counts
    // for the M rows first, one barrier, then each thread its own places
    unsigned long long mask[M];
#pragma unroll
    for (int r = 0; r < M; r++) {
      mask[r] = __ballot(need[r]);
      if (lane == 0) wcnt[r][wave] = __popcll(mask[r]);
    }
    __syncthreads();
    int n = 0;
#pragma unroll
    for (int r = 0; r < M; r++) {
      int before = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; w++) {
        const int c = wcnt[r][w];
        before += w < wave ? c : 0;
        tot += c;
      }
      if (need[r]) list[n + before + __popcll(mask[r] & ((1ull << lane) - 1ull))] = e0 + r * kBlock;
      n += tot;
    }
    __syncthreads();
    // the evaluated entries are dense in the list and most of them are
    // active: each lane applies the law to its own entry right after its
    // geometry (the body records still in registers), no second ranking
    for (int c = (int)threadIdx.x; c < n; c += kBlock) {
      ActiveKin a;
      if (geom_entry<false>(d, list[c], true, a)) law_item<LAW>(d, sp, 0, pii, a);
    }
    __syncthreads();  // the next tile's ranking rewrites wcnt and list
  }
  if (d.skip_count) {  // one atomic per workgroup and count (the counts summed in LDS first)
    __shared__ unsigned long long wsum[3][kBlock / 64];
    if (lane == 0)
      for (int k = 0; k < 3; k++) wsum[k][wave] = cnt[k];
    __syncthreads();
    if (threadIdx.x < 3) {
      unsigned long long t = 0ull;
      for (int w = 0; w < kBlock / 64; w++) t += wsum[threadIdx.x][w];
      atomicAdd(d.skip_count + threadIdx.x, t);
    }
  }
}

// The gaps of the entries the last step skipped, from the current poses:
// what that step would have computed for them (Interaction::UpdateDispatcher,
// deactivate() giving 0).  The test is the skip's own with the bounds of the
// last step, so it also takes entries that step evaluated with a positive
// gap (recomputed to the same bits) and never an entry not evaluated since
// its key was cleared (a new list's entries keep the gap the merge gave).
__global__ void __launch_bounds__(kBlock) k_refresh_gaps(Dev d) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.ne; e += gridDim.x * blockDim.x) {
    if (d.eact[e] != 0 || !(d.ekey[e] - d.cbound[d.ei[e]] - d.cbound[d.ej[e]] > kSkipMargin)) continue;
    Kin k;
    k.deact = false;
    k.prev_dn = 0.0;
    update_kinematics<false>(d, (int)(d.esub[e] >> 30), d.edesc[e], d.kin[d.ei[e]], d.kin[d.ej[e]], k);
    d.edn[e] = k.deact ? 0.0 : k.dn;
  }
}

__global__ void __launch_bounds__(kBlock) k_reset_keys(Dev d) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.ne; e += gridDim.x * blockDim.x) d.ekey[e] = -INFINITY;
}

__device__ __forceinline__ v3 euler_domega(v3 I, v3 w, v3 M) {
  return V((M.x - (I.z - I.y) * w.y * w.z) / I.x, (M.y - (I.x - I.z) * w.z * w.x) / I.y,
           (M.z - (I.y - I.x) * w.x * w.y) / I.z);
}

// The body force of a free body as (F, M), returned by value: reference
// outputs merged across the branches compile to a stack copy (scratch).
struct FM {
  v3 F, M;
};
__device__ __forceinline__ FM body_force(const Dev& d, int b, const BodyKin& P) {
  FM r;
  r.F = V(0.0, 0.0, 0.0);
  r.M = V(0.0, 0.0, 0.0);
  if (d.bf == BF_ATTRACTING) {  // BodyForce_AttractingPoint.cpp:51-59
    v3 dir = vsub(V(d.bf_point[0], d.bf_point[1], d.bf_point[2]), kpos(P));
    double l2 = vnorm2(dir);
    if (l2 > 0.0) {
      double inv = 1.0 / sqrt(l2);
      dir = V(dir.x * inv, dir.y * inv, dir.z * inv);
      r.F = vsc(d.mass[b] * d.bf_acc, dir);
    }
  } else if (d.bf == BF_PREFDIR) {  // BodyForce_PreferredDirection.cpp:54-75
    v3 axis = V(d.bf_axis[0], d.bf_axis[1], d.bf_axis[2]);
    v3 axisB = qrot(kq(P), V(d.bf_axisBody[0], d.bf_axisBody[1], d.bf_axisBody[2]));
    vnormalize(&axisB);
    v3 u = vcross(axis, axisB);
    vnormalize(&u);
    double cosa = vdot(axisB, axis);
    double angle = acos(cosa);
    double sgn = -1.0;
    if (cosa < 0.0) {
      sgn = 1.0;
      angle = M_PI - angle;
    }
    r.M = vsc(angle, vsc(d.bf_Kr * sgn, u));
  } else if (d.bf == BF_VISCOUS) {  // BodyForce_ViscousFluid.cpp:51-62
    v3 velSqr = V(P.vx * P.vx, P.vy * P.vy, P.vz * P.vz);
    r.F = vsc(0.5 * 0.47 * d.bf_fluidDensity * d.bf_surface[b], velSqr);
  }
  return r;
}

// initialise_particle_forces_and_moments (Rockable.cpp:2869-2929): body force
// for free bodies, imposed force and moment values for driven ones.
// BF false: the deck has no body force (the kernel then holds none of its code).
template <bool BF = true>
__device__ __forceinline__ void init_resultants(const Dev& d, int b, const BodyKin& P, bool driven, v3& F, v3& M,
                                                int& c0, int& c1) {
  F = V(0.0, 0.0, 0.0);
  M = V(0.0, 0.0, 0.0);
  c0 = c1 = 0;
  if (BF && !driven && d.bf != BF_NONE) {
    const FM bf = body_force(d, b, P);
    F = vadd(F, bf.F);
    M = vadd(M, bf.M);
  }
  if (driven && d.nctrl > 0) {
    c0 = d.ctrl_off[b];
    c1 = d.ctrl_off[b + 1];
    // selects rather than a switch: a switch over components compiles to
    // an indexed store into a stack copy of (F, M), i.e. scratch memory
    for (int q = c0; q < c1; q++) {
      const DevControl& C = d.ctrl[d.ctrl_idx[q]];
      const int ty = C.type;
      const double v = C.value;
      F.x = ty == C_X_FOR ? v : F.x;
      F.y = ty == C_Y_FOR ? v : F.y;
      F.z = ty == C_Z_FOR ? v : F.z;
      M.x = ty == C_XROT_MOM ? v : ty == C_XYZROT_MOM ? C.vx : M.x;
      M.y = ty == C_YROT_MOM ? v : ty == C_XYZROT_MOM ? C.vy : M.y;
      M.z = ty == C_ZROT_MOM ? v : ty == C_XYZROT_MOM ? C.vz : M.z;
    }
  }
}

__device__ __forceinline__ double pick(const v3& F, const v3& M, int c) {
  return c == 0 ? F.x : c == 1 ? F.y : c == 2 ? F.z : c == 3 ? M.x : c == 4 ? M.y : M.z;
}

// Stores the resultants, compute_accelerations_from_resultants
// (Rockable.cpp:3419-3552) and the closing half kick (2120-2155).
// m, I: the body's mass and principal inertia (loaded by the caller, early)
__device__ __forceinline__ void finish_body(const Dev& d, const StepParams& sp, int kick, int b, BodyKin& P,
                                            bool driven, v3 F, v3 M, int c0, int c1, double m, v3 I) {
  st3(d.force + 3 * (size_t)b, F);
  st3(d.moment + 3 * (size_t)b, M);
  const qt Q = kq(P);
  v3 acc, arot;
  if (driven) {
    acc = V(0.0, 0.0, 0.0);
    arot = V(0.0, 0.0, 0.0);
    for (int qq = c0; qq < c1; qq++) {
      const DevControl& C = d.ctrl[d.ctrl_idx[qq]];
      const int ty = C.type;
      acc.x = ty == C_X_FOR ? F.x / m : acc.x;
      acc.y = ty == C_Y_FOR ? F.y / m : acc.y;
      acc.z = ty == C_Z_FOR ? F.z / m : acc.z;
      // selects rather than branches on the component (branches compile to
      // an indexed store into a stack copy of arot, i.e. scratch memory)
      const bool all = ty == C_XYZROT_MOM;
      const bool rx = all || ty == C_XROT_MOM, ry = all || ty == C_YROT_MOM, rz = all || ty == C_ZROT_MOM;
      if (rx || ry || rz) {
        const qt Qinv = qconj(Q);
        const v3 w = V(rx ? P.wx : 0.0, ry ? P.wy : 0.0, rz ? P.wz : 0.0);
        const v3 omega = qrot(Qinv, w);
        const v3 Mb = qrot(Qinv, M);
        const v3 dw = qrot(Q, euler_domega(I, omega, Mb));
        arot.x = rx ? dw.x : arot.x;
        arot.y = ry ? dw.y : arot.y;
        arot.z = rz ? dw.z : arot.z;
      }
    }
  } else {
    acc = vadd(vdiv(F, m), V(sp.gx, sp.gy, sp.gz));
    qt Qinv = qconj(Q);
    v3 omega = qrot(Qinv, kvrot(P));
    v3 Mb = qrot(Qinv, M);
    arot = qrot(Q, euler_domega(I, omega, Mb));
  }
  st3(d.acc + 3 * (size_t)b, acc);
  st3(d.arot + 3 * (size_t)b, arot);
  if (!kick) return;
  if (driven) {
    for (int qq = c0; qq < c1; qq++) {
      const DevControl& C = d.ctrl[d.ctrl_idx[qq]];
      const int ty = C.type;
      const v3 w = vadd(kvrot(P), vsc(sp.dt_2, arot));
      P.vx = ty == C_X_FOR ? P.vx + sp.dt_2 * acc.x : P.vx;
      P.vy = ty == C_Y_FOR ? P.vy + sp.dt_2 * acc.y : P.vy;
      P.vz = ty == C_Z_FOR ? P.vz + sp.dt_2 * acc.z : P.vz;
      P.wx = ty == C_XROT_MOM || ty == C_XYZROT_MOM ? w.x : P.wx;
      P.wy = ty == C_YROT_MOM || ty == C_XYZROT_MOM ? w.y : P.wy;
      P.wz = ty == C_ZROT_MOM || ty == C_XYZROT_MOM ? w.z : P.wz;
    }
  } else {
    v3 v = vadd(kvel(P), vsc(sp.dt_2, acc));
    v3 w = vadd(kvrot(P), vsc(sp.dt_2, arot));
    P.vx = v.x;
    P.vy = v.y;
    P.vz = v.z;
    P.wx = w.x;
    P.wy = w.y;
    P.wz = w.z;
  }
  BodyKin* out = d.kin + b;
  out->vx = P.vx;
  out->vy = P.vy;
  out->vz = P.vz;
  out->wx = P.wx;
  out->wy = P.wy;
  out->wz = P.wz;
}

// dynamicCheckUpdateNL (Rockable.cpp:1571-1596), accumulation for one body
// at the end of its step: deltaPos += vel dt, deltaQ += Q.dot(vrot) dt, then
// the running max of |deltaPos|^2 and min of deltaQ.w as order keys (NaNs are
// ignored as the serial comparisons ignore them).  rst: the accumulations
// restart (a rebuild at the previous step).
__device__ __forceinline__ unsigned long long wkey(double w) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(w);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
struct GuardAcc {
  unsigned long long kp, kw;  // max key of |deltaPos|^2, min key of deltaQ.w
};
__device__ __forceinline__ void guard_body(const Dev& d, const StepParams& sp, int b, const BodyKin& P, bool rst,
                                           GuardAcc& a) {
  v3 dp = rst ? V(0.0, 0.0, 0.0) : ld3(d.gdpos + 3 * (size_t)b);
  dp = vadd(dp, vsc(sp.dt, kvel(P)));
  st3(d.gdpos + 3 * (size_t)b, dp);
  const double tmp1 = vnorm2(dp);
  if (tmp1 == tmp1) {
    const unsigned long long k1 = (unsigned long long)__double_as_longlong(tmp1);
    a.kp = k1 > a.kp ? k1 : a.kp;
  }
  const qt dq = qdot(kq(P), kvrot(P));
  double* g = d.gdq + 4 * (size_t)b;
  qt q;
  if (rst) {
    q.w = 1.0;
    q.x = q.y = q.z = 0.0;
  } else {
    q.w = g[0];
    q.x = g[1];
    q.y = g[2];
    q.z = g[3];
  }
  q.w += dq.w * sp.dt;
  q.x += dq.x * sp.dt;
  q.y += dq.y * sp.dt;
  q.z += dq.z * sp.dt;
  g[0] = q.w;
  g[1] = q.x;
  g[2] = q.y;
  g[3] = q.z;
  double w = q.w;
  if (w > 1.0) w = 1.0;
  if (w < -1.0) w = -1.0;
  if (w == w) {
    const unsigned long long k2 = wkey(w);
    a.kw = k2 < a.kw ? k2 : a.kw;
  }
}
__device__ __forceinline__ bool guard_reset(const Dev& d) { return *(volatile const int*)&d.guard->reset != 0; }

// compute_resultants (Rockable.cpp:3267-3272).  L lanes per body (8 or 16;
// 64 / L bodies per wavefront).  The lanes read the slot activity bytes 64
// at a time (64 / L each) and join them into one mask; lanes 0-5, one per
// component of (F, M), then fetch the terms of the active slots eight at a
// time (independent loads) and add them in slot order.  Called by every
// lane of the wavefront; returns the updated sum in lanes gl < 6.
template <int L>  // lanes per body: 8 or 16 (windows of 64 slots, 64 / L per lane)
__device__ __forceinline__ double sum_slots(const Dev& d, bool run, int k0, int k1, int gl, double s) {
  constexpr int B = 64 / L;  // activity bytes read per lane and window
  const double* val = d.sval + (gl < 6 ? gl : 0);
  for (int it = 0;; it++) {
    const int w = k0 + 64 * it;
    const bool live = run && w < k1;
    if (!__any(live)) break;
    unsigned mb = 0u;
    if (live) {
#pragma unroll
      for (int q = 0; q < B; q++) {
        const int k = w + B * gl + q;
        if (k < k1 && d.sact[k]) mb |= 1u << q;
      }
    }
    unsigned long long rest = (unsigned long long)mb << (B * gl);
    for (int o = 1; o < L; o <<= 1) rest |= __shfl_xor(rest, o, L);
    if (gl < 6) {
      while (rest) {
        double x[8];
        int n = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) {
          x[t] = 0.0;
          if (rest) {
            const int pbit = __ffsll((long long)rest) - 1;
            rest &= rest - 1ull;
            x[t] = val[6 * (size_t)(w + pbit)];
            n = t + 1;
          }
        }
#pragma unroll
        for (int t = 0; t < 8; t++)
          if (t < n) s += x[t];
      }
    }
  }
  return s;
}

// Bodies with at most heavy_min slots: the serial sum, so every component is
// the reference's bit for bit.  A body wavefront takes G groups of 64 / L
// bodies (G from 1 to L, so 64 / L to 64 bodies), one per lane: each lane
// loads its body's slot range and initial force and moment, the slots are
// summed group after group (L lanes per body; groups with no slot are
// skipped), and each lane then finishes its own body.  Few groups keep more
// wavefronts in flight (small packings), many use every lane for the
// per-body work (large ones).  The
// wavefronts after the bodies' take the 64-slot chunks of the heavy bodies
// and store each chunk's partial sums (from -0.0, the identity of addition)
// for k_heavy_finish.  With guard, the displacement check accumulates each
// finished body and every workgroup stores its maxima in gpart[2 blockIdx.x].
// Occupancy asked of the compiler: 5 wavefronts per SIMD without a body force
// (96 VGPRs, no spill), 3 with one (its code needs more registers).
template <int L, bool BF>  // BF: the deck has a body force
__global__ void __launch_bounds__(kBlock, BF ? 3 : 5) k_gather(Dev d, StepParams sp, int kick, int guard) {
  constexpr int kPer = 64 / L;  // bodies per slot group (chunks per chunk wavefront)
  __shared__ unsigned long long red[2][kBlock / 64];
  if (frozen(d)) return;
  const int lane = threadIdx.x & 63;
  const int gl = lane & (L - 1), sub = lane / L;
  const int wpb = (int)(blockDim.x >> 6);
  const int G = d.gather_groups;  // slot groups per body wavefront (1..L)
  const int NB = kPer * G;        // bodies per body wavefront
  const int body_waves = (d.nb + NB - 1) / NB;
  const int waves = body_waves + (d.nhch + kPer - 1) / kPer;
  const bool rst = guard && guard_reset(d);
  GuardAcc ga;
  ga.kp = 0ull;
  ga.kw = ~0ull;
  for (TileLoop T = xcd_tiles((waves + wpb - 1) / wpb); T.t < T.end; T.t += T.step) {
    const int wave = (int)((T.t * (unsigned)blockDim.x + threadIdx.x) >> 6);
    if (wave >= body_waves) {
      const int c = (wave - body_waves) * kPer + sub;
      const bool run = c < d.nhch;
      int2 r = make_int2(0, 0);
      if (run) r = d.hch[c];
      const double sc = sum_slots<L>(d, run, r.x, r.y, gl, -0.0);
      if (run && gl < 6) d.hpart[6 * (size_t)c + gl] = sc;
      continue;
    }
    const int b = wave * NB + lane;
    int k0 = 0, k1 = 0, c0 = 0, c1 = 0;
    bool run = false, driven = false;
    if (lane < NB && b < d.nb) {
      k0 = d.soff[b];
      k1 = d.soff[b + 1];
      run = (k1 - k0) <= d.heavy_min;
    }
    v3 F = V(0.0, 0.0, 0.0), M = V(0.0, 0.0, 0.0);
    BodyKin P;
    double mb = 0.0;
    v3 Ib = V(0.0, 0.0, 0.0);
    if (run) {
      // the body's state, mass and inertia are loaded before the sums, so
      // that their latency overlaps the slot loads
      P = d.kin[b];
      mb = d.mass[b];
      Ib = ld3(d.inertia + 3 * (size_t)b);
      driven = b < d.nDriven;
      init_resultants<BF>(d, b, P, driven, F, M, c0, c1);
    }
    const unsigned long long busy = __ballot(run && k1 > k0);
    for (int g = 0; g < G; g++) {
      if (((busy >> (g * kPer)) & ((1ull << kPer) - 1ull)) == 0ull) continue;  // uniform
      // lane (sub, gl) sums component gl of body lane g kPer + sub
      const int src = g * kPer + sub;
      const int s0 = __shfl(k0, src), s1 = __shfl(k1, src);
      const bool r2 = __shfl(run ? 1 : 0, src) != 0;
      const v3 Fs = V(__shfl(F.x, src), __shfl(F.y, src), __shfl(F.z, src));
      const v3 Ms = V(__shfl(M.x, src), __shfl(M.y, src), __shfl(M.z, src));
      const double sg = sum_slots<L>(d, r2, s0, s1, gl, pick(Fs, Ms, gl < 6 ? gl : 0));
      // back to the body lanes of group g: component c of body lane l is in
      // lane (l - g kPer) L + c
      const int base = (lane & (kPer - 1)) * L;
      const double fx = __shfl(sg, base), fy = __shfl(sg, base + 1), fz = __shfl(sg, base + 2);
      const double mx = __shfl(sg, base + 3), my = __shfl(sg, base + 4), mz = __shfl(sg, base + 5);
      if (lane / kPer == g) {
        F = V(fx, fy, fz);
        M = V(mx, my, mz);
      }
    }
    if (run) {
      finish_body(d, sp, kick, b, P, driven, F, M, c0, c1, mb, Ib);
      if (guard) guard_body(d, sp, b, P, rst, ga);
    }
  }
  if (!guard) return;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long op = __shfl_xor(ga.kp, off, 64), ow = __shfl_xor(ga.kw, off, 64);
    ga.kp = op > ga.kp ? op : ga.kp;
    ga.kw = ow < ga.kw ? ow : ga.kw;
  }
  const int wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = ga.kp;
    red[1][wv] = ga.kw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < wpb; k++) {
      ga.kp = red[0][k] > ga.kp ? red[0][k] : ga.kp;
      ga.kw = red[1][k] < ga.kw ? red[1][k] : ga.kw;
    }
    d.gpart[2 * (size_t)blockIdx.x] = ga.kp;
    d.gpart[2 * (size_t)blockIdx.x + 1] = ga.kw;
  }
}

// A passive heavy body (its force and moment feed nothing back into the
// motion; exact_sums off): the initial force and moment plus its chunks'
// partial sums, reduced by one workgroup of four wavefronts in a fixed order
// (thread t adds chunks t, t + 256, ... in turn, a butterfly over each
// wavefront's lanes, then (w0 + w1) + (w2 + w3)).  Deterministic; it
// differs from the serial sum only by rounding.
__device__ __forceinline__ void guard_heavy(const Dev& d, const StepParams& sp, int b, const BodyKin& P) {
  GuardAcc a;
  a.kp = 0ull;
  a.kw = ~0ull;
  guard_body(d, sp, b, P, guard_reset(d), a);
  if (a.kp) atomicMax(&d.guard->maxpos2, a.kp);
  if (a.kw != ~0ull) atomicMin(&d.guard->minw, a.kw);
}

static_assert(kBlock == 256, "k_heavy_finish combines four wavefronts");
__global__ void __launch_bounds__(kBlock) k_heavy_finish(Dev d, StepParams sp, int kick, int guard) {
  __shared__ double red[kBlock / 64][6];
  if (frozen(d)) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (TileLoop T = xcd_tiles(d.nheavy); T.t < T.end; T.t += T.step) {
    const int h = T.t;
    const int b = d.heavy[h];
    if (!d.passive[b] || d.exact_sums) continue;  // the same for the whole workgroup
    double a[6] = {-0.0, -0.0, -0.0, -0.0, -0.0, -0.0};
    for (int c = d.hch_off[h] + t; c < d.hch_off[h + 1]; c += kBlock) {
      const double* p = d.hpart + 6 * (size_t)c;
#pragma unroll
      for (int q = 0; q < 6; q++) a[q] += p[q];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int q = 0; q < 6; q++) a[q] += __shfl_xor(a[q], off, 64);
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 6; q++) red[wv][q] = a[q];
    __syncthreads();
    if (t == 0) {
#pragma unroll
      for (int q = 0; q < 6; q++) a[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
      BodyKin P = d.kin[b];
      const bool driven = b < d.nDriven;
      v3 F, M;
      int c0, c1;
      init_resultants(d, b, P, driven, F, M, c0, c1);
      finish_body(d, sp, kick, b, P, driven, vadd(F, V(a[0], a[1], a[2])), vadd(M, V(a[3], a[4], a[5])), c0, c1,
                  d.mass[b], ld3(d.inertia + 3 * (size_t)b));
      if (guard) guard_heavy(d, sp, b, P);
    }
    __syncthreads();
  }
}

// compute_resultants for a heavy body that is not passive (or any heavy
// body with exact_sums): the serial sum, one workgroup per body.  Waves 1-3
// stage chunks of kHeavyChunk slots into LDS (inactive slots as -0.0, the
// identity of addition: x + (-0.0) == x for every x), double-buffered, while
// lanes 0-5 of wave 0 add them in order, one lane per component.
constexpr int kHeavyChunk = 512;

__global__ void __launch_bounds__(kBlock) k_gather_heavy(Dev d, StepParams sp, int kick, int guard) {
  __shared__ double lds[2 * kHeavyChunk * 6];
  if (frozen(d)) return;
  const int b = d.heavy[blockIdx.x];
  const int t = threadIdx.x;
  const int k0 = d.soff[b], k1 = d.soff[b + 1];
  if (d.passive[b] && !d.exact_sums) return;  // summed by chunks
  BodyKin P = d.kin[b];
  const bool driven = b < d.nDriven;
  v3 F, M;
  int c0, c1;
  init_resultants(d, b, P, driven, F, M, c0, c1);
  const uint8_t* act = d.sact;
  const double* val = d.sval;
  const int len = k1 - k0;
  const int nchunks = (len + kHeavyChunk - 1) / kHeavyChunk;
  const int wave = t >> 6, lane = t & 63;
  double s = pick(F, M, lane < 6 ? lane : 0);
  auto stage = [&](int c) {
    double* buf = lds + (c & 1) * kHeavyChunk * 6;
    const int kb = k0 + c * kHeavyChunk;
    const int n = min(kHeavyChunk, k1 - kb);
    for (int idx = t - 64; idx < n; idx += kBlock - 64) {
      const int k = kb + idx;
      const bool on = act[k] != 0;
      const double* v = val + 6 * (size_t)k;
#pragma unroll
      for (int q = 0; q < 6; q++) buf[6 * idx + q] = on ? v[q] : -0.0;
    }
  };
  if (wave > 0 && nchunks > 0) stage(0);
  __syncthreads();
  for (int c = 0; c < nchunks; c++) {
    if (wave > 0) {
      if (c + 1 < nchunks) stage(c + 1);
    } else if (lane < 6) {
      const double* buf = lds + (c & 1) * kHeavyChunk * 6 + lane;
      const int n = min(kHeavyChunk, k1 - (k0 + c * kHeavyChunk));
      for (int idx = 0; idx < n; idx++) s += buf[6 * idx];
    }
    __syncthreads();
  }
  if (wave == 0) {
    const double fx = __shfl(s, 0), fy = __shfl(s, 1), fz = __shfl(s, 2);
    const double mx = __shfl(s, 3), my = __shfl(s, 4), mz = __shfl(s, 5);
    if (lane == 0) {
      finish_body(d, sp, kick, b, P, driven, V(fx, fy, fz), V(mx, my, mz), c0, c1, d.mass[b],
                  ld3(d.inertia + 3 * (size_t)b));
      if (guard) guard_heavy(d, sp, b, P);
    }
  }
}

// the end of a step of the integrate() loop (Rockable.cpp:2690-2740): a
// needUpdate left by the previous check makes the rebuild due now (the step
// kernels then wait for the host); then the check itself, on the maxima of
// the gather's workgroups (nparts pairs in gpart; max and min do not depend
// on the order).  The step index lives on the device, so that a captured
// sequence of steps replays as is.
constexpr int kCheckBlock = 1024;
__global__ void __launch_bounds__(kCheckBlock) k_guard_check(Dev d, int nparts) {
  __shared__ unsigned long long red[2][kCheckBlock / 64];
  if (frozen(d)) return;
  unsigned long long kp = 0ull, kw = ~0ull;
  for (int k = threadIdx.x; k < nparts; k += kCheckBlock) {
    const unsigned long long p = d.gpart[2 * (size_t)k], w = d.gpart[2 * (size_t)k + 1];
    kp = p > kp ? p : kp;
    kw = w < kw ? w : kw;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long op = __shfl_xor(kp, off, 64), ow = __shfl_xor(kw, off, 64);
    kp = op > kp ? op : kp;
    kw = ow < kw ? ow : kw;
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = kp;
    red[1][threadIdx.x >> 6] = kw;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < kCheckBlock / 64; k++) {
    kp = red[0][k] > kp ? red[0][k] : kp;
    kw = red[1][k] < kw ? red[1][k] : kw;
  }
  DevGuard* G = d.guard;
  const long long step = G->step;
  G->step = step + 1;
  if (d.note) *(volatile long long*)&d.note->done = step + 1;
  const int rebuild_now = G->nu;
  G->nu = 0;
  if (d.disp_check) {
    if (kp > G->maxpos2) G->maxpos2 = kp;
    if (kw < G->minw) G->minw = kw;
    G->reset = 0;
    const double mp = sqrt(__longlong_as_double((long long)G->maxpos2));
    const bool rot = G->minw != ~0ull && G->minw <= wkey(d.wstar);
    if (mp >= d.disp || rot) {
      G->nu = 1;
      G->reset = 1;
      G->maxpos2 = 0ull;
      G->minw = ~0ull;
    }
  }
  // a rebuild is due after this step: the check of the step before asked
  // for it, or an interface broke in this step (breakage_of_interfaces sets
  // needUpdate, Rockable.cpp:3410, and the loop rebuilds at once, 2693)
  if (rebuild_now || *(volatile const int*)d.flags != 0) {
    G->frozen = 1;
    G->frozen_step = step;
    if (d.note) *(volatile int32_t*)&d.note->frozen = 1;
  }
}

// The 16 statistics the rebuild-interval network reads (per-body speed, spin,
// acceleration and accumulated displacement, and the guard's running
// maxima), scaled by the guard's thresholds.  Pass 1: per-block partials;
// pass 2: one workgroup reduces them in a fixed order (no atomics).
constexpr int kFeatPart = 9;
__global__ void __launch_bounds__(kBlock) k_guard_feat_part(Dev d, StepParams sp, double disp, double* part) {
  __shared__ double red[kFeatPart][kBlock];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = threadIdx.x;
  double v[kFeatPart] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (b < d.nb) {
    const BodyKin& P = d.kin[b];
    const double vn = sqrt(vnorm2(kvel(P))), wn = sqrt(vnorm2(kvrot(P))), an = sqrt(vnorm2(ld3(d.acc + 3 * (size_t)b)));
    const double dx = d.gdpos ? sqrt(vnorm2(ld3(d.gdpos + 3 * (size_t)b))) : 0.0;
    v[0] = vn;                                     // max |v|
    v[1] = vn;                                     // sum |v|
    v[2] = vn * vn;                                // sum |v|^2
    v[3] = wn;                                     // max |w|
    v[4] = wn;                                     // sum |w|
    v[5] = an;                                     // max |a|
    v[6] = an;                                     // sum |a|
    v[7] = dx;                                     // sum |dx|
    v[8] = (vn * sp.dt > 0.01 * disp) ? 1.0 : 0.0;  // moving bodies
  }
  for (int k = 0; k < kFeatPart; k++) red[k][t] = (v[k] == v[k]) ? v[k] : 0.0;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int k = 0; k < kFeatPart; k++) {
        const bool mx = (k == 0 || k == 3 || k == 5);
        red[k][t] = mx ? fmax(red[k][t], red[k][t + h]) : red[k][t] + red[k][t + h];
      }
    __syncthreads();
  }
  if (t < kFeatPart) part[(size_t)blockIdx.x * kFeatPart + t] = red[t][0];
}

__global__ void __launch_bounds__(kBlock) k_guard_feat_final(Dev d, StepParams sp, const double* part, int nparts,
                                                             double disp, double angle, long long since,
                                                             long long last, float* feat) {
  // thread t takes the partials t, t + 256, ... in turn, then a fixed tree
  // over the threads (the same order whatever the scheduling)
  __shared__ double red[kFeatPart][kBlock];
  const int t = threadIdx.x;
  double a[kFeatPart] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = t; q < nparts; q += kBlock)
#pragma unroll
    for (int k = 0; k < kFeatPart; k++) {
      const double x = part[(size_t)q * kFeatPart + k];
      const bool mx = (k == 0 || k == 3 || k == 5);
      a[k] = mx ? fmax(a[k], x) : a[k] + x;
    }
#pragma unroll
  for (int k = 0; k < kFeatPart; k++) red[k][t] = a[k];
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    __syncthreads();
    if (t < h)
#pragma unroll
      for (int k = 0; k < kFeatPart; k++) {
        const bool mx = (k == 0 || k == 3 || k == 5);
        red[k][t] = mx ? fmax(red[k][t], red[k][t + h]) : red[k][t] + red[k][t + h];
      }
  }
  if (t != 0) return;
#pragma unroll
  for (int k = 0; k < kFeatPart; k++) a[k] = red[k][0];
  const double nb = d.nb > 0 ? (double)d.nb : 1.0;
  const double ds = disp > 0.0 ? disp : 1.0, an = angle > 0.0 ? angle : 1.0, dt = sp.dt;
  double dxmax = 0.0, rot = 0.0;
  if (d.guard) {
    dxmax = sqrt(__longlong_as_double((long long)d.guard->maxpos2));
    if (d.guard->minw != ~0ull && d.wstar < 1.0 && d.wstar >= -1.0) {
      const unsigned long long k = d.guard->minw;
      const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
      const double w = __longlong_as_double((long long)u);
      rot = (1.0 - w) / (1.0 - d.wstar);
    }
  }
  double f[16];
  f[0] = dxmax / ds;
  f[1] = rot;
  f[2] = a[0] * dt / ds;
  f[3] = a[1] / nb * dt / ds;
  f[4] = sqrt(a[2] / nb) * dt / ds;
  f[5] = a[3] * dt / an;
  f[6] = a[4] / nb * dt / an;
  f[7] = a[5] * dt * dt / ds;
  f[8] = a[6] / nb * dt * dt / ds;
  f[9] = log1p(fmax(0.0, (1.0 - f[0]) / fmax(f[2], 1e-12))) / 10.0;
  f[10] = log1p((double)since) / 10.0;
  f[11] = log1p((double)last) / 10.0;
  f[12] = a[8] / nb;
  f[13] = a[7] / nb / ds;
  f[14] = fmin(f[2] * (double)since, 10.0);
  f[15] = 1.0;
  for (int k = 0; k < 16; k++) {
    double x = f[k] == f[k] ? f[k] : 0.0;
    x = x > 10.0 ? 10.0 : (x < -10.0 ? -10.0 : x);
    feat[k] = (float)x;
  }
}

// Particle::GlobVertex (Particle.cpp:55-57): one thread per vertex of the
// first nv vertices (every body's, or the driven bodies', which come first),
// of every body (all) or of the bodies with slots in the current list (the
// only ones the force phase reads; the list builder reads all).
__global__ void __launch_bounds__(kBlock) k_glob(Dev d, int all, int nv) {
  if (frozen(d)) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nv; k += gridDim.x * blockDim.x) {
    const int b = d.bvbody[k];
    if (!all && d.soff[b + 1] == d.soff[b]) continue;
    const BodyKin& P = d.kin[b];
    st3(d.gv + 3 * (size_t)k, glob_vertex(d.sh, P, k - d.bvoff[b]));
  }
}

// ServoFunction (DrivingSystem.cpp:251-494), then the controlled bodies in
// list order (Rockable.cpp:1856-1900, 1988-2033), all by one thread.  The
// tritri servos read the walls' inner faces here; the shakers' and the
// ramp's values (functions of t only) come from the host in sp.servo_v.
// The servo of the driving system (DrivingSystem.cpp:242-496): the control
// values it imposes, from the walls' positions before any control moves them.
__device__ void drive_servo(const Dev& d, const StepParams& sp) {
  const int ic = d.servo_icontr;
  if (d.servo >= 1 && d.servo <= 4 && ic + 5 < d.nctrl) {
    const int* id = d.servo_ids;
    const double Xmin = d.kin[id[0]].px + d.kin[id[0]].R;
    const double Xmax = d.kin[id[1]].px - d.kin[id[1]].R;
    const double Ymin = d.kin[id[2]].py + d.kin[id[2]].R;
    const double Ymax = d.kin[id[3]].py - d.kin[id[3]].R;
    const double Zmin = d.kin[id[4]].pz + d.kin[id[4]].R;
    const double Zmax = d.kin[id[5]].pz - d.kin[id[5]].R;
    const double dX = fabs(Xmax - Xmin), dY = fabs(Ymax - Ymin), dZ = fabs(Zmax - Zmin);
    const double surfX = dY * dZ, surfY = dX * dZ, surfZ = dY * dX;
    const double p = d.servo_pressure;
    DevControl* C = d.ctrl + ic;
    if (d.servo == 1) {  // isostatic (251-268)
      C[1].value = -p * surfX;
      C[3].value = -p * surfY;
      C[5].value = -p * surfZ;
    } else if (d.servo == 2) {  // biaxial (280-298)
      C[1].value = -p * surfX;
      C[5].value = -p * surfZ;
    } else if (d.servo == 3) {  // custom (328-353): force-driven walls only
      const double* v = d.servo_val;
      if (C[0].type == C_X_FOR) C[0].value = v[0] * surfX;
      if (C[1].type == C_X_FOR) C[1].value = -v[1] * surfX;
      if (C[2].type == C_Y_FOR) C[2].value = v[2] * surfY;
      if (C[3].type == C_Y_FOR) C[3].value = -v[3] * surfY;
      if (C[4].type == C_Z_FOR) C[4].value = v[4] * surfZ;
      if (C[5].type == C_Z_FOR) C[5].value = -v[5] * surfZ;
    } else {  // Lode angle (369-392)
      const double dSig = d.lode_rate * sp.t;
      C[1].value = -(p + dSig) * surfX;
      C[3].value = -(p + d.lode_a * dSig) * surfY;
      C[5].value = -(p + d.lode_b * dSig) * surfZ;
    }
  } else if (d.servo >= 5 && d.servo <= 7 && ic + 2 < d.nctrl) {  // shakers
    d.ctrl[ic].value = sp.servo_v[0];
    d.ctrl[ic + 1].value = sp.servo_v[1];
    d.ctrl[ic + 2].value = sp.servo_v[2];
  } else if (d.servo == 8 && ic < d.nctrl) {  // ramp
    d.ctrl[ic].value = sp.servo_v[0];
  }
}

// The controls of the predictor (Rockable.cpp:1979-2035): first the
// velocity controls of every control in list order, then the force and
// moment ones.  A control reads and writes only its own body, so the two
// passes restricted to one body give that body's state whatever the other
// bodies' controls do.
__device__ __forceinline__ void control_vel(const Dev& d, const StepParams& sp, const DevControl& C) {
  BodyKin* p = d.kin + C.i;
  switch (C.type) {
    case C_X_VEL:
      p->vx = C.value;
      p->px += sp.dt * C.value;
      break;
    case C_Y_VEL:
      p->vy = C.value;
      p->py += sp.dt * C.value;
      break;
    case C_Z_VEL:
      p->vz = C.value;
      p->pz += sp.dt * C.value;
      break;
    case C_XROT_VEL:
    case C_YROT_VEL:
    case C_ZROT_VEL:
    case C_XYZROT_VEL: {
      v3 w;
      if (C.type == C_XROT_VEL)
        w = V(C.value, 0.0, 0.0);
      else if (C.type == C_YROT_VEL)
        w = V(0.0, C.value, 0.0);
      else if (C.type == C_ZROT_VEL)
        w = V(0.0, 0.0, C.value);
      else
        w = V(C.vx, C.vy, C.vz);
      qt Q = kq(*p);
      Q = qadd_scaled(Q, qdot(Q, w), sp.dt);
      qnormalize(&Q);
      p->qw = Q.w;
      p->qx = Q.x;
      p->qy = Q.y;
      p->qz = Q.z;
    } break;
    default: break;
  }
}
__device__ __forceinline__ void control_force(const Dev& d, const StepParams& sp, const DevControl& C) {
  BodyKin* p = d.kin + C.i;
  const double* a = d.acc + 3 * (size_t)C.i;
  const double* ar = d.arot + 3 * (size_t)C.i;
  switch (C.type) {
    case C_X_FOR:
      p->px += sp.dt * p->vx + sp.dt2_2 * a[0];
      p->vx += sp.dt_2 * a[0];
      break;
    case C_Y_FOR:
      p->py += sp.dt * p->vy + sp.dt2_2 * a[1];
      p->vy += sp.dt_2 * a[1];
      break;
    case C_Z_FOR:
      p->pz += sp.dt * p->vz + sp.dt2_2 * a[2];
      p->vz += sp.dt_2 * a[2];
      break;
    case C_XROT_MOM:
    case C_YROT_MOM:
    case C_ZROT_MOM:
    case C_XYZROT_MOM: {
      v3 w;
      if (C.type == C_XROT_MOM)
        w = V(p->wx * 1.0, p->wx * 0.0, p->wx * 0.0);
      else if (C.type == C_YROT_MOM)
        w = V(p->wy * 0.0, p->wy * 1.0, p->wy * 0.0);
      else if (C.type == C_ZROT_MOM)
        w = V(p->wz * 0.0, p->wz * 0.0, p->wz * 1.0);
      else
        w = V(p->wx, p->wy, p->wz);
      qt Q = kq(*p);
      Q = qadd_scaled(Q, qdot(Q, w), sp.dt);
      qnormalize(&Q);
      p->qw = Q.w;
      p->qx = Q.x;
      p->qy = Q.y;
      p->qz = Q.z;
      if (C.type == C_XROT_MOM)
        p->wx += sp.dt_2 * ar[0];
      else if (C.type == C_YROT_MOM)
        p->wy += sp.dt_2 * ar[1];
      else if (C.type == C_ZROT_MOM)
        p->wz += sp.dt_2 * ar[2];
      else {
        v3 nw = vadd(V(p->wx, p->wy, p->wz), vsc(sp.dt_2, V(ar[0], ar[1], ar[2])));
        p->wx = nw.x;
        p->wy = nw.y;
        p->wz = nw.z;
      }
    } break;
    default: break;
  }
}

// All controls by one thread in list order (the servo first).
__device__ void drive_controls(const Dev& d, const StepParams& sp) {
  drive_servo(d, sp);
  for (int c = 0; c < d.nctrl; c++) control_vel(d, sp, d.ctrl[c]);
  for (int c = 0; c < d.nctrl; c++) control_force(d, sp, d.ctrl[c]);
}

// The controls by the threads of one workgroup (every thread calls it): the
// servo by thread 0, then each driven body's controls in list order by one
// thread (ctrl_parallel: every control acts on a driven body), else all of
// them by thread 0.
__device__ void drive_controls_block(const Dev& d, const StepParams& sp) {
  if (d.nctrl <= 0) return;
  if (!d.ctrl_parallel) {
    if (threadIdx.x == 0) drive_controls(d, sp);
    return;
  }
  if (threadIdx.x == 0) drive_servo(d, sp);
  __syncthreads();  // the servo's values are read by the other threads
  for (int b = threadIdx.x; b < d.nDriven; b += blockDim.x) {
    const int q0 = d.ctrl_off[b], q1 = d.ctrl_off[b + 1];
    for (int q = q0; q < q1; q++) control_vel(d, sp, d.ctrl[d.ctrl_idx[q]]);
    for (int q = q0; q < q1; q++) control_force(d, sp, d.ctrl[d.ctrl_idx[q]]);
  }
}

// predict_free: the new pose and velocities of free body b (Rockable.cpp:
// 2036-2055) in P; store_predicted writes them back
__device__ __forceinline__ void predict_free_with(const StepParams& sp, BodyKin& P, v3 a, v3 ar) {
  v3 pos = vadd(kpos(P), vadd(vsc(sp.dt, kvel(P)), vsc(sp.dt2_2, a)));
  v3 vel = vadd(kvel(P), vsc(sp.dt_2, a));
  qt Q = kq(P);
  const v3 w = kvrot(P);
  Q = qadd_scaled(Q, qdot(Q, w), sp.dt);
  Q = qadd_scaled(Q, qddot(Q, w, ar), sp.dt2_2);
  qnormalize(&Q);
  v3 w2 = vadd(w, vsc(sp.dt_2, ar));
  P.px = pos.x;
  P.py = pos.y;
  P.pz = pos.z;
  P.qw = Q.w;
  P.qx = Q.x;
  P.qy = Q.y;
  P.qz = Q.z;
  P.vx = vel.x;
  P.vy = vel.y;
  P.vz = vel.z;
  P.wx = w2.x;
  P.wy = w2.y;
  P.wz = w2.z;
}
__device__ __forceinline__ void predict_free(const Dev& d, const StepParams& sp, int b, BodyKin& P) {
  predict_free_with(sp, P, ld3(d.acc + 3 * (size_t)b), ld3(d.arot + 3 * (size_t)b));
}
// The far-entry skip's motion bound of one predicted body: no vertex of it
// moved by more than |dx| + 2 |dQ| rmax in this step (for unit quaternions
// |Q' v Q'* - Q v Q*| <= 2 |Q' - Q| |v|), plus a rounding allowance.
__device__ __forceinline__ void add_motion_bound_with(const Dev& d, int b, const BodyKin& old, const BodyKin& P,
                                                      double cb, double rmax) {
  if (!d.skip_far) return;
  const double dx = P.px - old.px, dy = P.py - old.py, dz = P.pz - old.pz;
  const double qw = P.qw - old.qw, qx = P.qx - old.qx, qy = P.qy - old.qy, qz = P.qz - old.qz;
  const double step = sqrt(dx * dx + dy * dy + dz * dz) + 2.0 * sqrt(qw * qw + qx * qx + qy * qy + qz * qz) * rmax;
  const double mag = fabs(P.px) + fabs(P.py) + fabs(P.pz);
  d.cbound[b] = cb + (step + 1e-15 * (1.0 + mag));
}
__device__ __forceinline__ void add_motion_bound(const Dev& d, int b, const BodyKin& old, const BodyKin& P) {
  if (!d.skip_far) return;
  add_motion_bound_with(d, b, old, P, d.cbound[b], d.rmax[b]);
}

__device__ __forceinline__ void store_predicted(const Dev& d, int b, const BodyKin& P) {
  BodyKin* out = d.kin + b;
  out->px = P.px;
  out->py = P.py;
  out->pz = P.pz;
  out->qw = P.qw;
  out->qx = P.qx;
  out->qy = P.qy;
  out->qz = P.qz;
  out->vx = P.vx;
  out->vy = P.vy;
  out->vz = P.vz;
  out->wx = P.wx;
  out->wy = P.wy;
  out->wz = P.wz;
}

// glob: also the global vertices of the free bodies with listed slots, from
// the new pose (the driven bodies' follow in k_glob)
__global__ void __launch_bounds__(kBlock) k_predict(Dev d, StepParams sp, int glob) {
  if (frozen(d)) return;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0) drive_controls_block(d, sp);
  // free bodies (Rockable.cpp:2036-2055), a grid-stride loop
  for (int b = gid; b < d.nb; b += gridDim.x * blockDim.x) {
    if (b < d.nDriven) continue;
    BodyKin P = d.kin[b];
    const BodyKin old = P;
    predict_free(d, sp, b, P);
    store_predicted(d, b, P);
    add_motion_bound(d, b, old, P);
    if (glob && d.soff[b + 1] != d.soff[b]) {
      const int v0 = d.bvoff[b], nv = d.bvoff[b + 1] - v0;
      for (int v = 0; v < nv; v++) st3(d.gv + 3 * (size_t)(v0 + v), glob_vertex(d.sh, P, v));
    }
  }
}

// The predictor by workgroups of 64 free bodies.  The tile's 64 body records
// (128 bytes each) are read and written whole by the 256 threads in
// contiguous doubles through LDS (17 doubles per record there, so that the
// per-body reads spread over the banks); one thread per body computes its
// new state in between and keeps the pose in LDS; then the 256 threads share
// those bodies' vertices (global vertices of the listed ones).  Rewriting a
// record's unchanged fields (R, h, shape, group) stores the bits it read.
// The driven bodies are moved by thread 0 of workgroup 0 (controls); with
// driven_glob, workgroup 0 then writes their listed global vertices (else
// k_glob does after this kernel).
__global__ void __launch_bounds__(kBlock) k_predict_blk(Dev d, StepParams sp, int driven_glob) {
  constexpr int kRec = 17;          // LDS doubles per record
  __shared__ double rec[64 * kRec];
  __shared__ double pose[64][8];    // px py pz qw qx qy qz h
  __shared__ int shp[64];           // shape, or -1 if the body is not listed
  static_assert(sizeof(BodyKin) == 16 * sizeof(double), "a body record is 16 doubles");
  if (frozen(d)) return;
  const int t = threadIdx.x;
  if (blockIdx.x == 0) {
    drive_controls_block(d, sp);
    if (driven_glob) {
      __syncthreads();  // the driven poses, written by thread 0, are visible to the workgroup
      for (int k = t; k < d.nvdriven; k += blockDim.x) {
        const int b = d.bvbody[k];
        if (d.soff[b + 1] == d.soff[b]) continue;
        const BodyKin& P = d.kin[b];
        st3(d.gv + 3 * (size_t)k, glob_vertex(d.sh, P, k - d.bvoff[b]));
      }
    }
  }
  const int nfree = d.nb - d.nDriven;
  for (int tile = blockIdx.x; tile * 64 < nfree; tile += gridDim.x) {
    const int b0 = d.nDriven + tile * 64;
    const int bend = min(b0 + 64, d.nb);
    const int nd = 16 * (bend - b0);
    double* kd = reinterpret_cast<double*>(d.kin + b0);
    for (int k = t; k < nd; k += blockDim.x) rec[(k >> 4) * kRec + (k & 15)] = kd[k];
    // the body's other inputs loaded with the records, so that their
    // latencies overlap
    v3 acc = V(0.0, 0.0, 0.0), arot = acc;
    double cb = 0.0, rmx = 0.0;
    bool listed = false;
    if (t < 64 && b0 + t < bend) {
      const int b = b0 + t;
      acc = ld3(d.acc + 3 * (size_t)b);
      arot = ld3(d.arot + 3 * (size_t)b);
      if (d.skip_far) {
        cb = d.cbound[b];
        rmx = d.rmax[b];
      }
      listed = d.soff[b + 1] != d.soff[b];
    }
    __syncthreads();
    if (t < 64) {
      const int b = b0 + t;
      shp[t] = -1;
      if (b < bend) {
        double* r = rec + t * kRec;
        BodyKin P;
        P.px = r[0];
        P.py = r[1];
        P.pz = r[2];
        P.qw = r[3];
        P.qx = r[4];
        P.qy = r[5];
        P.qz = r[6];
        P.vx = r[7];
        P.vy = r[8];
        P.vz = r[9];
        P.wx = r[10];
        P.wy = r[11];
        P.wz = r[12];
        P.R = r[13];
        P.h = r[14];
        P.shape = __double2loint(r[15]);
        P.group = __double2hiint(r[15]);
        const BodyKin old = P;
        predict_free_with(sp, P, acc, arot);
        r[0] = P.px;
        r[1] = P.py;
        r[2] = P.pz;
        r[3] = P.qw;
        r[4] = P.qx;
        r[5] = P.qy;
        r[6] = P.qz;
        r[7] = P.vx;
        r[8] = P.vy;
        r[9] = P.vz;
        r[10] = P.wx;
        r[11] = P.wy;
        r[12] = P.wz;
        add_motion_bound_with(d, b, old, P, cb, rmx);
        if (listed) {
          pose[t][0] = P.px;
          pose[t][1] = P.py;
          pose[t][2] = P.pz;
          pose[t][3] = P.qw;
          pose[t][4] = P.qx;
          pose[t][5] = P.qy;
          pose[t][6] = P.qz;
          pose[t][7] = P.h;
          shp[t] = P.shape;
        }
      }
    }
    __syncthreads();
    for (int k = t; k < nd; k += blockDim.x) kd[k] = rec[(k >> 4) * kRec + (k & 15)];
    for (int k = d.bvoff[b0] + t; k < d.bvoff[bend]; k += blockDim.x) {
      const int b = d.bvbody[k], lb = b - b0;
      if (shp[lb] < 0) continue;
      BodyKin P;
      P.px = pose[lb][0];
      P.py = pose[lb][1];
      P.pz = pose[lb][2];
      P.qw = pose[lb][3];
      P.qx = pose[lb][4];
      P.qy = pose[lb][5];
      P.qz = pose[lb][6];
      P.h = pose[lb][7];
      P.shape = shp[lb];
      st3(d.gv + 3 * (size_t)k, glob_vertex(d.sh, P, k - d.bvoff[b]));
    }
    __syncthreads();
  }
}

// The predictor with one thread per vertex of the free bodies, each
// wavefront holding whole bodies (wavefront w: vertices [wvstart[w],
// wvstart[w + 1])): every lane computes its body's new state, the lane of
// vertex 0 stores it and every lane of a listed body its global vertex.  The
// lanes of a body read its old state in the same load instruction, before
// any of them stores, so the body is read whole before it is overwritten.
// The driven bodies (controls by thread 0) and their vertices are left to
// k_glob after it.
__global__ void __launch_bounds__(kBlock) k_predict_vtx(Dev d, StepParams sp) {
  if (frozen(d)) return;
  if (blockIdx.x == 0) drive_controls_block(d, sp);
  const int lane = threadIdx.x & 63;
  for (int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); wv < d.nwv;
       wv += (int)((gridDim.x * blockDim.x) >> 6)) {
    const int k = d.wvstart[wv] + lane;
    if (k >= d.wvstart[wv + 1]) continue;
    const int b = d.bvbody[k];
    const int v = k - d.bvoff[b];
    const bool listed = d.soff[b + 1] != d.soff[b];
    if (v != 0 && !listed) continue;
    BodyKin P = d.kin[b];
    const BodyKin old = P;
    predict_free(d, sp, b, P);
    if (v == 0) {
      store_predicted(d, b, P);
      add_motion_bound(d, b, old, P);
    }
    if (listed) st3(d.gv + 3 * (size_t)k, glob_vertex(d.sh, P, v));
  }
}

// locate entry (owner, other, sub) in the owner's sorted run
__device__ int find_entry(const Dev& d, int owner, int other, uint32_t sub) {
  int lo = d.istart[owner], hi = d.istart[owner + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    bool less = (d.ej[mid] < other) || (d.ej[mid] == other && d.esub[mid] < sub);
    if (less)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < d.istart[owner + 1] && d.ej[lo] == other && d.esub[lo] == sub) return lo;
  return -1;
}

__global__ void k_breakage(Dev d) {
  if (frozen(d)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.ifc.n || !d.ifc.brk[k]) return;
  d.ifc.brk[k] = 0;
  if (!d.ifc.alive[k]) return;
  const uint64_t key = d.ifc.key[k];
  const int bi = (int)(key >> 32), bj = (int)(key & 0xffffffffu);
  const int g1 = d.kin[bi].group, g2 = d.kin[bj].group;
  for (int q = d.ifc.bond_off[k]; q < d.ifc.bond_off[k + 1]; q++) {
    const int* bk = d.ifc.bond_key + 3 * q;
    const int e = find_entry(d, bi, bj, pack_sub(bk[0], bk[1], bk[2]));
    if (e < 0) continue;
    d.estick[e] = -1;
    EState s = d.est[e];
    if (d.edn[e] >= 0.0) {
      s.fn = 0.0;
      s.ftx = s.fty = s.ftz = 0.0;
    } else {
      v3 T = V(s.ftx, s.fty, s.ftz);
      double FT = vnormalize(&T);
      double mu = prm(d.par, P_MU, g1, g2);
      double thr = fabs(mu * s.fn);
      if (FT > thr) {
        v3 ft = vsc(thr, T);
        s.ftx = ft.x;
        s.fty = ft.y;
        s.ftz = ft.z;
      }
    }
    d.est[e] = s;
  }
  d.ifc.alive[k] = 0;
  // for the host's accounting (it sorts them: interface order); an interface
  // breaks once, so the list holds at most n
  d.ifc.broken[atomicAdd(d.ifc.nbroken, 1)] = k;
  d.flags[0] = 1;  // needUpdate
}

}  // namespace

int body_block(const Dev& d) { return d.body_block == 64 || d.body_block == 128 ? d.body_block : 256; }

// workgroups for T tiles (persistent kernels loop over their tiles)
dim3 grid_for(const Dev& d, long T) { return dim3(persistent_grid((int)std::max(1L, T), d.persist)); }

void launch_glob(const Dev& d, int all, int nv, hipStream_t s) {
  if (nv <= 0) return;
  const int bs = body_block(d);
  hipLaunchKernelGGL(k_glob, grid_for(d, (nv + bs - 1) / bs), dim3(bs), 0, s, d, all || !d.soff ? 1 : 0, nv);
}

void launch_glob(const Dev& d, int all, hipStream_t s) { launch_glob(d, all, d.nvtot, s); }

void launch_refresh_gaps(const Dev& d, hipStream_t s) {
  if (d.ne <= 0 || !d.skip_far) return;
  hipLaunchKernelGGL(k_refresh_gaps, dim3(std::min((d.ne + kBlock - 1) / kBlock, 65536)), dim3(kBlock), 0, s, d);
}

void launch_reset_keys(const Dev& d, hipStream_t s) {
  if (d.ne <= 0 || !d.ekey) return;
  hipLaunchKernelGGL(k_reset_keys, dim3(std::min((d.ne + kBlock - 1) / kBlock, 65536)), dim3(kBlock), 0, s, d);
}

void launch_entry_desc(const Dev& d, hipStream_t s) {
  if (d.ne <= 0 || !d.edesc) return;
  hipLaunchKernelGGL(k_entry_desc, dim3((d.ne + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d);
}

void launch_predict(const Dev& d, const StepParams& sp, int glob, hipStream_t s) {
  const int n = d.nb > 0 ? d.nb : 1, bs = body_block(d);
  const bool listed = glob && d.soff;  // else every body's vertices
  if (listed && d.fuse_glob == 3 && d.nb > d.nDriven) {
    // workgroups of 64 free bodies, their vertices from the poses in LDS; a
    // few driven bodies' vertices by workgroup 0 after the controls
    const int tiles = (d.nb - d.nDriven + 63) / 64;
    const int driven_glob = d.nvdriven <= 1024 ? 1 : 0;
    hipLaunchKernelGGL(k_predict_blk, grid_for(d, tiles), dim3(kBlock), 0, s, d, sp, driven_glob);
    if (!driven_glob) launch_glob(d, 0, d.nvdriven, s);
    return;
  }
  if (listed && d.fuse_glob == 2 && d.nwv > 0) {
    // one thread per vertex of the free bodies, wavefronts of whole bodies
    const int wpb = kBlock / 64;
    hipLaunchKernelGGL(k_predict_vtx, grid_for(d, (d.nwv + wpb - 1) / wpb), dim3(kBlock), 0, s, d, sp);
    launch_glob(d, 0, d.nvdriven, s);
    return;
  }
  const int fused = listed && d.fuse_glob != 0;  // the free bodies' in k_predict
  hipLaunchKernelGGL(k_predict, grid_for(d, (n + bs - 1) / bs), dim3(bs), 0, s, d, sp, fused);
  launch_glob(d, listed ? 0 : 1, fused ? d.nvdriven : d.nvtot, s);
}

template <int BS>
void launch_geom(const Dev& d, int write_geom_all, hipStream_t s) {
  const dim3 g = grid_for(d, (d.ne + BS - 1) / BS), bl(BS);
  if (write_geom_all)
    hipLaunchKernelGGL((k_geom<true, BS>), g, bl, 0, s, d);
  else
    hipLaunchKernelGGL((k_geom<false, BS>), g, bl, 0, s, d);
}

template <bool GEOM_ALL, bool SORT>
void launch_fused_sorted(const Dev& d, const StepParams& sp, int write_geom_all, hipStream_t s) {
  const dim3 g = grid_for(d, (d.ne + kBlock - 1) / kBlock), bl(kBlock);
  if (sp.law == LAW_AVALANCHE)
    hipLaunchKernelGGL((k_interact<LAW_AVALANCHE, GEOM_ALL, SORT>), g, bl, 0, s, d, sp, write_geom_all, 0);
  else if (sp.law == LAW_STICKED)
    hipLaunchKernelGGL((k_interact<LAW_STICKED, GEOM_ALL, SORT>), g, bl, 0, s, d, sp, write_geom_all,
                       d.paramsInInterfaces);
  else
    hipLaunchKernelGGL((k_interact<LAW_DEFAULT, GEOM_ALL, SORT>), g, bl, 0, s, d, sp, write_geom_all, 0);
}

template <int SEG>
void launch_seg(const Dev& d, const StepParams& sp, hipStream_t s) {
  const dim3 g = grid_for(d, (d.ne + SEG - 1) / SEG), bl(kBlock);
  if (sp.law == LAW_AVALANCHE)
    hipLaunchKernelGGL((k_interact_seg<LAW_AVALANCHE, SEG>), g, bl, 0, s, d, sp, 0);
  else if (sp.law == LAW_STICKED)
    hipLaunchKernelGGL((k_interact_seg<LAW_STICKED, SEG>), g, bl, 0, s, d, sp, d.paramsInInterfaces);
  else
    hipLaunchKernelGGL((k_interact_seg<LAW_DEFAULT, SEG>), g, bl, 0, s, d, sp, 0);
}

template <bool GEOM_ALL>
void launch_fused(const Dev& d, const StepParams& sp, int write_geom_all, hipStream_t s) {
  if (!GEOM_ALL && d.skip_far && d.skip_seg >= 512) {
    switch (d.skip_seg) {
      case 512: launch_seg<512>(d, sp, s); return;
      case 2048: launch_seg<2048>(d, sp, s); return;
      case 4096: launch_seg<4096>(d, sp, s); return;
      default: launch_seg<1024>(d, sp, s); return;
    }
  }
  if (d.type_sort)
    launch_fused_sorted<GEOM_ALL, true>(d, sp, write_geom_all, s);
  else
    launch_fused_sorted<GEOM_ALL, false>(d, sp, write_geom_all, s);
}

void launch_interact(const Dev& d, const StepParams& sp, int write_geom_all, hipStream_t s) {
  if (d.ne <= 0) return;
  if (d.fuse_law) {
    if (write_geom_all)
      launch_fused<true>(d, sp, write_geom_all, s);
    else
      launch_fused<false>(d, sp, write_geom_all, s);
    return;
  }
  switch (d.geom_block) {
    case 64: launch_geom<64>(d, write_geom_all, s); break;
    case 128: launch_geom<128>(d, write_geom_all, s); break;
    case 512: launch_geom<512>(d, write_geom_all, s); break;
    default: launch_geom<256>(d, write_geom_all, s); break;
  }
  const dim3 bl(kBlock);
  // W geometry blocks per wavefront
  const int gbs = (d.geom_block == 64 || d.geom_block == 128 || d.geom_block == 512) ? d.geom_block : 256;
  const int ngeom = (d.ne + gbs - 1) / gbs;
  const int W = std::min(std::max(d.law_group, 1), kLawGroup);
  const int waves = (ngeom + W - 1) / W;
  const dim3 gl = grid_for(d, (waves + kBlock / 64 - 1) / (kBlock / 64));
  if (sp.law == LAW_AVALANCHE)
    hipLaunchKernelGGL(k_law<LAW_AVALANCHE>, gl, bl, 0, s, d, sp, write_geom_all, 0, ngeom, W);
  else if (sp.law == LAW_STICKED)
    hipLaunchKernelGGL(k_law<LAW_STICKED>, gl, bl, 0, s, d, sp, write_geom_all, d.paramsInInterfaces, ngeom, W);
  else
    hipLaunchKernelGGL(k_law<LAW_DEFAULT>, gl, bl, 0, s, d, sp, write_geom_all, 0, ngeom, W);
}

// k_gather's workgroups: 64 / lanes x gather_groups bodies per body
// wavefront, 64 / lanes heavy chunks per chunk wavefront, gather_waves
// wavefronts per workgroup
int gather_blocks(const Dev& d) {
  const int per = d.gather_lanes == 8 ? 8 : 4;
  const long nbw = (long)per * d.gather_groups;
  const long waves = ((long)d.nb + nbw - 1) / nbw + ((long)d.nhch + per - 1) / per;
  const int wpb = d.gather_waves == 1 || d.gather_waves == 2 ? d.gather_waves : 4;
  return (int)grid_for(d, (waves + wpb - 1) / wpb).x;
}

int gather_blocks_max(int nb, int nhch) { return (nb + 3) / 4 + (nhch + 3) / 4 + 8; }

void launch_gather(const Dev& d, const StepParams& sp, int kick, int guard, hipStream_t s) {
  if (d.nb <= 0) return;
  guard = guard && kick;
  const int wpb = d.gather_waves == 1 || d.gather_waves == 2 ? d.gather_waves : 4;
  const dim3 blocks(gather_blocks(d));
  const bool bf = d.bf != BF_NONE;
  if (d.gather_lanes == 8 && bf)
    hipLaunchKernelGGL((k_gather<8, true>), blocks, dim3(64 * wpb), 0, s, d, sp, kick, guard);
  else if (d.gather_lanes == 8)
    hipLaunchKernelGGL((k_gather<8, false>), blocks, dim3(64 * wpb), 0, s, d, sp, kick, guard);
  else if (bf)
    hipLaunchKernelGGL((k_gather<16, true>), blocks, dim3(64 * wpb), 0, s, d, sp, kick, guard);
  else
    hipLaunchKernelGGL((k_gather<16, false>), blocks, dim3(64 * wpb), 0, s, d, sp, kick, guard);
  if (d.nheavy > 0) {
    hipLaunchKernelGGL(k_heavy_finish, grid_for(d, d.nheavy), dim3(kBlock), 0, s, d, sp, kick, guard);
    if (d.nheavy_exact > 0)
      hipLaunchKernelGGL(k_gather_heavy, dim3(d.nheavy), dim3(kBlock), 0, s, d, sp, kick, guard);
  }
}

// --- statistics of the packing (DataExtractor-style quantities) -----------
// Partial sums per workgroup (a fixed tree in LDS), then one workgroup
// reduces the partials in a fixed order: the results do not depend on
// scheduling.
constexpr int kStatBody = 4;   // KE translation, KE rotation, max |v|, free bodies
constexpr int kStatEntry = 5;  // active, contacts (dn < 0), active pairs, sum dn, min dn

__device__ __forceinline__ void stat_block_reduce(double (*red)[kBlock], int nval, const int* is_max, int t) {
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int k = 0; k < nval; k++)
        red[k][t] = is_max[k] == 1   ? fmax(red[k][t], red[k][t + h])
                    : is_max[k] == 2 ? fmin(red[k][t], red[k][t + h])
                                     : red[k][t] + red[k][t + h];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_stat_bodies(Dev d, double* part) {
  __shared__ double red[kStatBody][kBlock];
  const int t = threadIdx.x;
  const int b = blockIdx.x * kBlock + t;
  double v[kStatBody] = {0.0, 0.0, 0.0, 0.0};
  if (b < d.nb && b >= d.nDriven) {
    const BodyKin P = d.kin[b];
    const double m = d.mass[b];
    const v3 I = ld3(d.inertia + 3 * (size_t)b);
    const v3 vel = kvel(P);
    const v3 wb = qrot(qconj(kq(P)), kvrot(P));
    const double v2 = vdot(vel, vel);
    v[0] = 0.5 * m * v2;
    v[1] = 0.5 * (I.x * wb.x * wb.x + I.y * wb.y * wb.y + I.z * wb.z * wb.z);
    v[2] = sqrt(v2);
    v[3] = 1.0;
  }
  for (int k = 0; k < kStatBody; k++) red[k][t] = v[k];
  const int ops[kStatBody] = {0, 0, 1, 0};
  stat_block_reduce(red, kStatBody, ops, t);
  if (t < kStatBody) part[(size_t)blockIdx.x * kStatBody + t] = red[t][0];
}

__global__ void __launch_bounds__(kBlock) k_stat_entries(Dev d, double* part) {
  __shared__ double red[kStatEntry][kBlock];
  const int t = threadIdx.x;
  const int e = blockIdx.x * kBlock + t;
  double v[kStatEntry] = {0.0, 0.0, 0.0, 0.0, INFINITY};
  if (e < d.ne && (d.eact[e] & 1)) {
    const double dn = d.edn[e];
    v[0] = 1.0;
    if (dn < 0.0) {
      v[1] = 1.0;
      v[3] = dn;
      v[4] = dn;
    }
    // the first active entry of its body pair (entries are sorted by pair)
    bool head = true;
    const int i = d.ei[e], j = d.ej[e];
    for (int k = e - 1; k >= 0 && d.ei[k] == i && d.ej[k] == j; k--)
      if (d.eact[k] & 1) {
        head = false;
        break;
      }
    v[2] = head ? 1.0 : 0.0;
  }
  for (int k = 0; k < kStatEntry; k++) red[k][t] = v[k];
  const int ops[kStatEntry] = {0, 0, 0, 0, 2};
  stat_block_reduce(red, kStatEntry, ops, t);
  if (t < kStatEntry) part[(size_t)blockIdx.x * kStatEntry + t] = red[t][0];
}

// out[0..3] body statistics, out[4..8] entry statistics
__global__ void __launch_bounds__(kBlock) k_stat_final(const double* pb, int nbp, const double* pe, int nep,
                                                       double* out) {
  // thread t takes the partials t, t + 256, ... in turn, then a fixed tree
  // over the threads (the same order whatever the scheduling)
  constexpr int NS = kStatBody + kStatEntry;
  __shared__ double red[NS][kBlock];
  const int t = threadIdx.x;
  double s[NS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, INFINITY};
  for (int q = t; q < nbp; q += kBlock) {
    s[0] += pb[kStatBody * q];
    s[1] += pb[kStatBody * q + 1];
    s[2] = fmax(s[2], pb[kStatBody * q + 2]);
    s[3] += pb[kStatBody * q + 3];
  }
  for (int q = t; q < nep; q += kBlock) {
    for (int k = 0; k < 4; k++) s[kStatBody + k] += pe[kStatEntry * q + k];
    s[kStatBody + 4] = fmin(s[kStatBody + 4], pe[kStatEntry * q + 4]);
  }
#pragma unroll
  for (int k = 0; k < NS; k++) red[k][t] = s[k];
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    __syncthreads();
    if (t < h)
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const double a = red[k][t], b = red[k][t + h];
        red[k][t] = k == 2 ? fmax(a, b) : k == NS - 1 ? fmin(a, b) : a + b;
      }
  }
  __syncthreads();
  if (t < NS) out[t] = red[t][0];
}

int stat_parts(const Dev& d) {
  return std::max(1, (d.nb + kBlock - 1) / kBlock) * kStatBody + std::max(1, (d.ne + kBlock - 1) / kBlock) * kStatEntry;
}

void launch_stats(const Dev& d, double* part, double* out, hipStream_t s) {
  const int nbp = std::max(1, (d.nb + kBlock - 1) / kBlock), nep = std::max(1, (d.ne + kBlock - 1) / kBlock);
  double* pe = part + (size_t)nbp * kStatBody;
  hipLaunchKernelGGL(k_stat_bodies, dim3(nbp), dim3(kBlock), 0, s, d, part);
  hipLaunchKernelGGL(k_stat_entries, dim3(nep), dim3(kBlock), 0, s, d, pe);
  hipLaunchKernelGGL(k_stat_final, dim3(1), dim3(kBlock), 0, s, part, nbp, pe, nep, out);
}

// Scratch of rk_get_stats_noalloc: partial sums and outputs of the
// statistics kernels in a fixed array of each device (8 MiB, held until the
// process ends), so that the call allocates nothing.  Host code never uses
// the array's name as an address: stat_scratch() looks the device address up
// once per device (hipGetSymbolAddress, with that device current) and the
// kernels receive that address.  Every call on one device shares the array;
// calls that overlap on one device race.
static __device__ __align__(256) double scratch[kStatScratchDoubles];

int stat_scratch(double** ptr, char* err, size_t errlen) {
  static std::mutex mu;
  static double* addr[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    snprintf(err, errlen, "stat_scratch: no current device");
    return -1;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      snprintf(err, errlen, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      return -1;
    }
    addr[dev] = static_cast<double*>(p);
  }
  *ptr = addr[dev];
  return 0;
}

__global__ void k_guard_set_step(Dev d, long long step) {
  if (threadIdx.x == 0) d.guard->step = step;
}

void launch_guard(const Dev& d, hipStream_t s) {
  const int nparts = (d.disp_check && d.nb > 0) ? gather_blocks(d) : 0;
  hipLaunchKernelGGL(k_guard_check, dim3(1), dim3(kCheckBlock), 0, s, d, nparts);
}

void launch_guard_set_step(const Dev& d, long long step, hipStream_t s) {
  hipLaunchKernelGGL(k_guard_set_step, dim3(1), dim3(64), 0, s, d, step);
}

void launch_guard_features(const Dev& d, const StepParams& sp, double disp, double angle, long long since,
                           long long last, double* part, float* feat, hipStream_t s) {
  const int np = std::max(1, (d.nb + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_guard_feat_part, dim3(np), dim3(kBlock), 0, s, d, sp, disp, part);
  hipLaunchKernelGGL(k_guard_feat_final, dim3(1), dim3(kBlock), 0, s, d, sp, part, np, disp, angle, since, last, feat);
}

void launch_breakage(const Dev& d, const StepParams& sp, hipStream_t s) {
  (void)sp;
  if (d.ifc.n <= 0) return;
  hipLaunchKernelGGL(k_breakage, dim3((d.ifc.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d);
}

}  // namespace rk
