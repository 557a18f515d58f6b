/*
 * rk_oracle.c -- CPU oracle for the Rockable velocity-Verlet DEM step.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).  The product
 * library never links or calls this file.  It restates, in plain C99 with
 * OpenMP, Rockable's DefaultParallelMode path:
 *
 *   integrate()                    Rockable.cpp:2538-2746
 *   velocityVerletStep()           Rockable.cpp:1979-2158
 *   velocityControlledDrive()      Rockable.cpp:1856-1900
 *   accelerations()                Rockable.cpp:3562-3719
 *   initialise_particle_forces...  Rockable.cpp:2869-2929
 *   update_interactions()          Rockable.cpp:2934-2977 -> Interaction.cpp:147-326
 *   build_activeInteractions()     Rockable.cpp:3005-3029
 *   compute_forces_and_moments()   Rockable.cpp:3035-3046 -> ForceLaw_*.cpp
 *   compute_resultants()           Rockable.cpp:3220-3275, incrementResultants 2752-2789
 *   breakage_of_interfaces()       Rockable.cpp:3350-3413
 *   compute_accelerations_from...  Rockable.cpp:3419-3552
 *   UpdateNL_bruteForce()          Rockable.cpp:1614-1662
 *   UpdateNL_linkCells()           Rockable.cpp:1667-1809
 *   AddOrRemoveInteractions_bruteForce  Rockable.cpp:1309-1432, Particle.cpp:88-209
 *   dynamicCheckUpdateNL()         Rockable.cpp:1571-1596
 *   loadConf/saveConf/initParser   Rockable.cpp:1085-1145, 440-610, 643-1069
 *   Shape::read                    Shape.cpp:69-128
 *   DrivingSystem::read            DrivingSystem.cpp:44-268
 *
 * Floating-point forms of the toofus primitives (absent from the reference
 * tree) are selected by rko_set_variant(); the defaults are the forms that
 * reproduce the reference's golden conf files line for line.
 *
 * Interactions are kept, as in the reference, in one ordered container per
 * owner body (Rockable.hpp:107: std::vector<std::set<Interaction>>); here an
 * array sorted by (j, type, isub, jsub) per owner.
 */
#define _GNU_SOURCE
#include "rk_oracle.h"

#include <fenv.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* toofus floating-point forms                                               */
/* ------------------------------------------------------------------------- */
enum {
  VAR_VDIV = 0,      /* vec3 / double: 0 = per-component divide, 1 = multiply by reciprocal */
  VAR_VNORM = 1,     /* vec3::normalize: 0 = divide, 1 = multiply by reciprocal */
  VAR_QROT = 2,      /* quat * vec3: 0 = matrix form, 1 = Rodrigues t=2(u x v), 2 = Hamilton products */
  VAR_QDOT = 3,      /* quat::dot(w): term order variants */
  VAR_QDDOT = 4,     /* quat::ddot(w, dw) variants */
  VAR_QNORM = 5,     /* quat::normalize: 0 = divide, 1 = multiply by reciprocal */
  VAR_OBBEPS = 6,    /* OBB::intersect epsilon: index into eps_tab */
  VAR_MUTANT = 7,    /* planted mistakes for the pinning tests (0 = none) */
  VAR_COUNT = 8
};
static int g_var[VAR_COUNT] = {1, 1, 1, 0, 2, 3, 0, 0};
void rko_set_variant(int which, int value) {
  if (which >= 0 && which < VAR_COUNT) g_var[which] = value;
}
int rko_get_variant(int which) { return (which >= 0 && which < VAR_COUNT) ? g_var[which] : -1; }

typedef struct {
  double x, y, z;
} v3;
typedef struct {
  double w, x, y, z;
} qt;

static inline v3 V(double x, double y, double z) {
  v3 r = {x, y, z};
  return r;
}
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vsc(double s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vnorm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 vdiv(v3 a, double s) {
  if (g_var[VAR_VDIV] == 1) {
    double r = 1.0 / s;
    return V(a.x * r, a.y * r, a.z * r);
  }
  return V(a.x / s, a.y / s, a.z / s);
}
/* vec3::normalize() returns the former length (used at Interaction.cpp:157). */
static inline double vnormalize(v3* a) {
  double n = sqrt(vnorm2(*a));
  if (g_var[VAR_VNORM] == 1) {
    double r = 1.0 / n;
    a->x *= r;
    a->y *= r;
    a->z *= r;
  } else {
    a->x /= n;
    a->y /= n;
    a->z /= n;
  }
  return n;
}

/* Hamilton product a (x) b = (a.s b.s - a.v.b.v, a.s b.v + b.s a.v + a.v x b.v) */
static inline qt qmul(qt a, qt b) {
  v3 av = V(a.x, a.y, a.z), bv = V(b.x, b.y, b.z);
  v3 c = vcross(av, bv);
  qt r;
  r.w = a.w * b.w - vdot(av, bv);
  r.x = (a.w * bv.x + b.w * av.x) + c.x;
  r.y = (a.w * bv.y + b.w * av.y) + c.y;
  r.z = (a.w * bv.z + b.w * av.z) + c.z;
  return r;
}
/* quat * vec3: rotation of a body-frame vector into the global frame. */
static inline v3 qrot(qt q, v3 v) {
  switch (g_var[VAR_QROT]) {
    case 1: { /* v + w t + u x t, t = 2 (u x v) */
      v3 u = V(q.x, q.y, q.z);
      v3 t = vcross(u, v);
      t = V(2.0 * t.x, 2.0 * t.y, 2.0 * t.z);
      v3 c = vcross(u, t);
      return V(v.x + q.w * t.x + c.x, v.y + q.w * t.y + c.y, v.z + q.w * t.z + c.z);
    }
    case 2: { /* (q * (0,v)) * conj(q) with Hamilton products */
      double pw = -q.x * v.x - q.y * v.y - q.z * v.z;
      double px = q.w * v.x + q.y * v.z - q.z * v.y;
      double py = q.w * v.y + q.z * v.x - q.x * v.z;
      double pz = q.w * v.z + q.x * v.y - q.y * v.x;
      return V(-pw * q.x + px * q.w - py * q.z + pz * q.y, -pw * q.y + py * q.w - pz * q.x + px * q.z,
               -pw * q.z + pz * q.w - px * q.y + py * q.x);
    }
    case 3: { /* (q (x) (0,v)) (x) conj(q) with vector-form Hamilton products */
      qt p = {0.0, v.x, v.y, v.z};
      qt t = qmul(q, p);
      qt c = {q.w, -q.x, -q.y, -q.z};
      qt r = qmul(t, c);
      return V(r.x, r.y, r.z);
    }
    case 4: { /* v + 2 w (u x v) + 2 u x (u x v) */
      v3 u = V(q.x, q.y, q.z);
      v3 c1 = vcross(u, v);
      v3 c2 = vcross(u, c1);
      return V(v.x + 2.0 * q.w * c1.x + 2.0 * c2.x, v.y + 2.0 * q.w * c1.y + 2.0 * c2.y,
               v.z + 2.0 * q.w * c1.z + 2.0 * c2.z);
    }
    case 5: { /* v + (2w)(u x v) + 2 (u x (u x v)) with t = u x v */
      v3 u = V(q.x, q.y, q.z);
      v3 c1 = vcross(u, v);
      v3 c2 = vcross(u, c1);
      double w2 = 2.0 * q.w;
      return V(v.x + (w2 * c1.x + 2.0 * c2.x), v.y + (w2 * c1.y + 2.0 * c2.y), v.z + (w2 * c1.z + 2.0 * c2.z));
    }
    default: { /* rotation-matrix form */
      double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
      double xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
      double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
      return V((1.0 - 2.0 * (yy + zz)) * v.x + 2.0 * (xy - wz) * v.y + 2.0 * (xz + wy) * v.z,
               2.0 * (xy + wz) * v.x + (1.0 - 2.0 * (xx + zz)) * v.y + 2.0 * (yz - wx) * v.z,
               2.0 * (xz - wy) * v.x + 2.0 * (yz + wx) * v.y + (1.0 - 2.0 * (xx + yy)) * v.z);
    }
  }
}
static inline qt qconj(qt q) {
  qt r = {q.w, -q.x, -q.y, -q.z};
  return r;
}
/* dQ/dt for a global-frame angular velocity: 1/2 (0,w) (x) Q, i.e.
 * 1/2 (-(v.w), s w + w x v) with the cross product formed first. */
static inline qt qdot(qt q, v3 w) {
  qt r;
  v3 c = vcross(w, V(q.x, q.y, q.z));
  if (g_var[VAR_QDOT] == 1) {
    r.w = -0.5 * (w.x * q.x + w.y * q.y + w.z * q.z);
    r.x = 0.5 * (w.x * q.w + w.y * q.z - w.z * q.y);
    r.y = 0.5 * (w.y * q.w + w.z * q.x - w.x * q.z);
    r.z = 0.5 * (w.z * q.w + w.x * q.y - w.y * q.x);
  } else {
    r.w = 0.5 * (-(q.x * w.x + q.y * w.y + q.z * w.z));
    r.x = 0.5 * (q.w * w.x + c.x);
    r.y = 0.5 * (q.w * w.y + c.y);
    r.z = 0.5 * (q.w * w.z + c.z);
  }
  return r;
}
/* d2Q/dt2 = 1/2 (0,dw) (x) Q + 1/2 (0,w) (x) dQ/dt. */
static inline qt qddot(qt q, v3 w, v3 dw) {
  qt r;
  switch (g_var[VAR_QDDOT]) {
    case 1: {
      double w2 = w.x * w.x + w.y * w.y + w.z * w.z;
      qt a = qdot(q, dw);
      r.w = a.w - 0.25 * w2 * q.w;
      r.x = a.x - 0.25 * w2 * q.x;
      r.y = a.y - 0.25 * w2 * q.y;
      r.z = a.z - 0.25 * w2 * q.z;
    } break;
    case 2: { /* 1/2 [ (0,dw) Q + (0,w) dQ ] with Hamilton products */
      qt d = qdot(q, w);
      qt zdw = {0.0, dw.x, dw.y, dw.z}, zw = {0.0, w.x, w.y, w.z};
      qt a = qmul(zdw, q), b = qmul(zw, d);
      r.w = 0.5 * (a.w + b.w);
      r.x = 0.5 * (a.x + b.x);
      r.y = 0.5 * (a.y + b.y);
      r.z = 0.5 * (a.z + b.z);
    } break;
    case 3: { /* dot(dw) + 1/2 (0,w) dQ */
      qt d = qdot(q, w);
      qt zw = {0.0, w.x, w.y, w.z};
      qt a = qdot(q, dw), b = qmul(zw, d);
      r.w = a.w + 0.5 * b.w;
      r.x = a.x + 0.5 * b.x;
      r.y = a.y + 0.5 * b.y;
      r.z = a.z + 0.5 * b.z;
    } break;
    default: {
      qt a = qdot(q, dw);
      qt d = qdot(q, w);
      qt b = qdot(d, w);
      r.w = a.w + b.w;
      r.x = a.x + b.x;
      r.y = a.y + b.y;
      r.z = a.z + b.z;
    }
  }
  return r;
}
static inline void qnormalize(qt* q) {
  double n2;
  int v = g_var[VAR_QNORM];
  if (v == 2 || v == 3)
    n2 = q->w * q->w + (q->x * q->x + q->y * q->y + q->z * q->z);
  else if (v == 4 || v == 5)
    n2 = (q->x * q->x + q->y * q->y + q->z * q->z) + q->w * q->w;
  else
    n2 = q->w * q->w + q->x * q->x + q->y * q->y + q->z * q->z;
  double n = sqrt(n2);
  if (v == 1 || v == 3 || v == 5) {
    double r = 1.0 / n;
    q->w *= r;
    q->x *= r;
    q->y *= r;
    q->z *= r;
  } else {
    q->w /= n;
    q->x /= n;
    q->y /= n;
    q->z /= n;
  }
}
static inline qt qadd_scaled(qt q, qt d, double s) {
  /* Q += (d *= s) */
  qt r = {q.w + d.w * s, q.x + d.x * s, q.y + d.y * s, q.z + d.z * s};
  return r;
}
/* quat::get_angle() (used by dynamicCheckUpdateNL, Rockable.cpp:1581). */
static inline double qangle(qt q) {
  double w = q.w;
  if (w > 1.0) w = 1.0;
  if (w < -1.0) w = -1.0;
  return 2.0 * acos(w);
}

/* ------------------------------------------------------------------------- */
/* OBB (toofus OBB): center, 3 unit axes, 3 half extents                     */
/* ------------------------------------------------------------------------- */
typedef struct {
  v3 c;
  v3 e[3];
  v3 ext;
} obb_t;

static inline obb_t obb_default(void) {
  obb_t o;
  o.c = V(0, 0, 0);
  o.e[0] = V(1, 0, 0);
  o.e[1] = V(0, 1, 0);
  o.e[2] = V(0, 0, 1);
  o.ext = V(0, 0, 0);
  return o;
}
static inline void obb_enlarge(obb_t* o, double d) {
  o->ext.x += d;
  o->ext.y += d;
  o->ext.z += d;
}
/* Separating-axis test over the 15 candidate axes (Gottschalk, Lin, Manocha
 * 1996, "OBBTree"); true when the boxes overlap (touching counts). */
static int obb_intersect(const obb_t* a, const obb_t* b) {
  static const double eps_tab[6] = {0.0, 1e-12, 1e-6, 2.220446049250313e-16, 1.1920928955078125e-07, 1e-9};
  const double eps = eps_tab[(g_var[VAR_OBBEPS] >= 0 && g_var[VAR_OBBEPS] < 6) ? g_var[VAR_OBBEPS] : 0];
  double R[3][3], AR[3][3];
  v3 d = vsub(b->c, a->c);
  double t[3] = {vdot(d, a->e[0]), vdot(d, a->e[1]), vdot(d, a->e[2])};
  double ea[3] = {a->ext.x, a->ext.y, a->ext.z};
  double eb[3] = {b->ext.x, b->ext.y, b->ext.z};
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) {
      R[i][k] = vdot(a->e[i], b->e[k]);
      AR[i][k] = fabs(R[i][k]) + eps;
    }
  double ra, rb;
  for (int i = 0; i < 3; i++) {
    ra = ea[i];
    rb = eb[0] * AR[i][0] + eb[1] * AR[i][1] + eb[2] * AR[i][2];
    if (fabs(t[i]) > ra + rb) return 0;
  }
  for (int k = 0; k < 3; k++) {
    ra = ea[0] * AR[0][k] + ea[1] * AR[1][k] + ea[2] * AR[2][k];
    rb = eb[k];
    if (fabs(t[0] * R[0][k] + t[1] * R[1][k] + t[2] * R[2][k]) > ra + rb) return 0;
  }
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int k = 0; k < 3; k++) {
      int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      ra = ea[i1] * AR[i2][k] + ea[i2] * AR[i1][k];
      rb = eb[k1] * AR[i][k2] + eb[k2] * AR[i][k1];
      if (fabs(t[i2] * R[i1][k] - t[i1] * R[i2][k]) > ra + rb) return 0;
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------- */
/* Data model                                                                */
/* ------------------------------------------------------------------------- */
enum { VV = 0, VE = 1, VF = 2, EE = 3 };

typedef struct {
  char name[128];
  double radius, volume;
  v3 inertia_mass;
  obb_t obb;
  char preCompDone;
  int nv, ne, nf;
  v3* vertex;
  int* edge;       /* 2 per edge */
  int* face_start; /* nf + 1 */
  int* face_idx;
} shape_t;

typedef struct {
  int group, cluster;
  double homothety;
  v3 pos, vel, acc;
  qt Q;
  v3 vrot, arot;
  int shape;
  double mass;
  v3 inertia;
  v3 force, moment;
  obb_t obb;
} body_t;

typedef struct {
  int i, j, type, isub, jsub;
  v3 prev_n, n;
  double dn, prev_dn;
  v3 pos, vel;
  double fn;
  v3 ft, mom;
  double damp;
  int stick; /* index of the BreakableInterface, -1 when none */
  int act;   /* in activeInteractions of the last accelerations() */
} inter_t;

typedef struct {
  inter_t* a;
  int n, cap;
} ilist_t;

typedef struct {
  int i, j;
  double kn, kt, kr, fn0, ft0, mom0, power, Gc;
  double dn0, area;
  int isInner;
  int alive;
  int nb;
  int* bkey; /* 3 per bond: type isub jsub (owner i, other j) */
} iface_t;

enum {
  CTRL_X_VEL = 0,
  CTRL_Y_VEL = 1,
  CTRL_Z_VEL = 2,
  CTRL_XROT_VEL = 3,
  CTRL_YROT_VEL = 4,
  CTRL_ZROT_VEL = 5,
  CTRL_X_FOR = 6,
  CTRL_Y_FOR = 7,
  CTRL_Z_FOR = 8,
  CTRL_XROT_MOM = 9,
  CTRL_YROT_MOM = 10,
  CTRL_ZROT_MOM = 11,
  CTRL_XYZROT_VEL = 100,
  CTRL_XYZROT_MOM = 101
};
typedef struct {
  int type;
  int i;
  double value;
  v3 vec_value;
} control_t;

/* DataTable parameter ids (Rockable.cpp:96-123) */
enum {
  P_KN = 0, P_EN2, P_KT, P_MU, P_KR, P_MUR, P_VISCT,
  P_KN_IN, P_KT_IN, P_KR_IN, P_EN2_IN, P_FN0_IN, P_FT0_IN, P_MOM0_IN, P_POW_IN,
  P_KN_OUT, P_KT_OUT, P_KR_OUT, P_EN2_OUT, P_FN0_OUT, P_FT0_OUT, P_MOM0_OUT, P_POW_OUT,
  P_GC_IN, P_GC_OUT,
  P_COUNT
};
static const char* PNAME[P_COUNT] = {
    "knContact", "en2Contact", "ktContact", "muContact", "krContact", "murContact", "viscTContact",
    "knInnerBond", "ktInnerBond", "krInnerBond", "en2InnerBond", "fn0InnerBond", "ft0InnerBond", "mom0InnerBond",
    "powInnerBond", "knOuterBond", "ktOuterBond", "krOuterBond", "en2OuterBond", "fn0OuterBond", "ft0OuterBond",
    "mom0OuterBond", "powOuterBond", "gcInnerBond", "gcOuterBond"};
/* order in which saveConf writes them (Rockable.cpp:513-541) */
static const int PWRITE[P_COUNT] = {P_KN,     P_EN2,     P_KT,     P_MU,      P_KR,       P_MUR,      P_VISCT,
                                    P_KN_IN,  P_KT_IN,   P_KR_IN,  P_FN0_IN,  P_FT0_IN,   P_MOM0_IN,  P_POW_IN,
                                    P_EN2_IN, P_KN_OUT,  P_KT_OUT, P_KR_OUT,  P_FN0_OUT,  P_FT0_OUT,  P_MOM0_OUT,
                                    P_POW_OUT, P_EN2_OUT, P_GC_IN, P_GC_OUT};

#define MAXG 256
enum { LAW_DEFAULT = 0, LAW_AVALANCHE = 1, LAW_STICKED = 2 };
enum { BF_NONE = 0, BF_ATTRACTING = 1, BF_PREFDIR = 2, BF_VISCOUS = 3 };

struct rko_sim {
  char err[512];
  int verbose;
  int nthreads;
  /* scalars */
  double t, tmax, dt, dt_2, dt2, dt2_2;
  double interVerlet, interConf, DVerlet, dVerlet;
  double interVerletC, interConfC;
  int paramsInInterfaces, dynamicUpdateNL;
  double dispUpdateNL, angleUpdateNL;
  double numericalDampingCoeff, velocityBarrier, angularVelocityBarrier;
  double velocityBarrierExponent, angularVelocityBarrierExponent, preventCrossingLength;
  v3 gravity;
  v3 cellMinSizes;
  int boxForLinkCellsOpt;
  int iconf;
  int nDriven;
  int precision;
  int law;
  char lawName[64];
  int nl_linkcells;
  int needUpdate;
  double totalBrokenArea;
  int64_t nl_updates;
  /* data table */
  double table[P_COUNT][MAXG][MAXG];
  unsigned char defined[P_COUNT][MAXG][MAXG];
  int ngroup;
  double density[MAXG];
  int ndensity;
  /* shapes */
  char shapeFile[1024];
  char path[512];
  shape_t* shapes;
  int nshapes;
  /* bodies */
  body_t* b;
  int nb;
  /* interactions */
  ilist_t* inter; /* per owner */
  inter_t** active;
  size_t nactive, cap_active;
  omp_lock_t* locks;
  /* interfaces */
  iface_t* ifc;
  int nifc;
  unsigned char* ifc_break; /* flag per interface, set by the force law */
  /* driving system */
  control_t* ctrl;
  int nctrl;
  /* ServoFunction (DrivingSystem.cpp:188-496): 0 none, 1 tritriIsostaticCompression,
   * 2 tritriBiaxialCompression, 3 tritriCustom, 4 tritriLodeAngle, 5 shaker,
   * 6 triangle_shaker, 7 sawtooth_shaker, 8 ramp; the last Servo line read wins */
  int servo;
  int servo_icontr, servo_ids[6];
  double servo_pressure, servo_velocity;
  double servo_val[6];                 /* tritriCustom: xmin xmax ymin ymax zmin zmax values */
  double lode_a, lode_b, lode_rate;    /* tritriLodeAngle */
  v3 shake_dir;                        /* shakers: normalised direction, amplitude, */
  double shake_A, shake_omega, shake_T, shake_tini; /* angular frequency, period, start */
  double ramp_vb, ramp_ve, ramp_tb, ramp_te, ramp_slope;
  int ramp_slope_set;                  /* the slope is a function-static of the first call */
  char driving_path[1024];             /* re-read at each conf save */
  /* integrate()'s output files (Rockable.cpp:215-222) and the DataExtractors
     of a dataExtractors.txt file (Rockable.cpp:1152-1180; dnStat, TrackDamage) */
  FILE *perf_file, *ke_file, *sb_file;
  int nextr;
  struct {
    int kind; /* 1 dnStat, 2 TrackDamage */
    int nrec;
    FILE* f;
  } extr[16];
  double inAreaInterfaces;
  int64_t ext_step; /* integrate()'s step counter (Rockable.cpp:2578, 2742) */
  /* body force */
  int bf;
  v3 bf_point;
  double bf_acc;
  v3 bf_axisBody, bf_axis;
  double bf_momentMax, bf_Kr, bf_fluidDensity;
  /* dynamic NL check */
  v3* deltaPos;
  qt* deltaQ;
  double maxDeltaPos, maxDeltaRot;
};

#ifndef _OPENMP
typedef int omp_lock_t;
static void omp_init_lock(omp_lock_t* l) { (void)l; }
static void omp_destroy_lock(omp_lock_t* l) { (void)l; }
static void omp_set_lock(omp_lock_t* l) { (void)l; }
static void omp_unset_lock(omp_lock_t* l) { (void)l; }
#endif

static int seterr(rko_sim* s, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->err, sizeof(s->err), fmt, ap);
  va_end(ap);
  if (s->verbose) fprintf(stderr, "[oracle] %s\n", s->err);
  return -1;
}

/* ------------------------------------------------------------------------- */
/* DataTable (toofus): symmetric group-pair tables, 0.0 when undefined        */
/* ------------------------------------------------------------------------- */
static void dt_set(rko_sim* s, int id, int g1, int g2, double v) {
  if (g1 < 0 || g2 < 0 || g1 >= MAXG || g2 >= MAXG) return;
  s->table[id][g1][g2] = v;
  s->table[id][g2][g1] = v;
  s->defined[id][g1][g2] = s->defined[id][g2][g1] = 1;
  if (g1 + 1 > s->ngroup) s->ngroup = g1 + 1;
  if (g2 + 1 > s->ngroup) s->ngroup = g2 + 1;
}
static inline double dt_get(const rko_sim* s, int id, int g1, int g2) {
  if (g1 < 0 || g2 < 0 || g1 >= MAXG || g2 >= MAXG) return 0.0;
  return s->table[id][g1][g2];
}

/* ------------------------------------------------------------------------- */
/* Text reader with the kwParser conventions                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
  char* buf;
  size_t len, pos;
} reader_t;

static int rd_open(reader_t* r, const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  r->buf = (char*)malloc((size_t)n + 1);
  r->len = fread(r->buf, 1, (size_t)n, f);
  r->buf[r->len] = 0;
  r->pos = 0;
  fclose(f);
  return 0;
}
static void rd_close(reader_t* r) { free(r->buf); }
/* Next whitespace-separated token; returns 0 at end of input. */
static int rd_tok(reader_t* r, char* out, size_t cap) {
  while (r->pos < r->len && (r->buf[r->pos] == ' ' || r->buf[r->pos] == '\t' || r->buf[r->pos] == '\n' ||
                             r->buf[r->pos] == '\r'))
    r->pos++;
  if (r->pos >= r->len) return 0;
  size_t k = 0;
  while (r->pos < r->len && !(r->buf[r->pos] == ' ' || r->buf[r->pos] == '\t' || r->buf[r->pos] == '\n' ||
                              r->buf[r->pos] == '\r')) {
    if (k + 1 < cap) out[k++] = r->buf[r->pos];
    r->pos++;
  }
  out[k] = 0;
  return 1;
}
static void rd_skipline(reader_t* r) {
  while (r->pos < r->len && r->buf[r->pos] != '\n') r->pos++;
}
static double rd_d(reader_t* r) {
  char t[128];
  if (!rd_tok(r, t, sizeof t)) return 0.0;
  return strtod(t, NULL);
}
static long rd_l(reader_t* r) {
  char t[128];
  if (!rd_tok(r, t, sizeof t)) return 0;
  return strtol(t, NULL, 10);
}
static v3 rd_v3(reader_t* r) {
  v3 v;
  v.x = rd_d(r);
  v.y = rd_d(r);
  v.z = rd_d(r);
  return v;
}
static qt rd_qt(reader_t* r) {
  qt q;
  q.w = rd_d(r);
  q.x = rd_d(r);
  q.y = rd_d(r);
  q.z = rd_d(r);
  return q;
}

/* ------------------------------------------------------------------------- */
/* Shapes (Shape.cpp:69-128; Rockable.cpp:1187-1225)                          */
/* ------------------------------------------------------------------------- */
static void shape_free(shape_t* sh) {
  free(sh->vertex);
  free(sh->edge);
  free(sh->face_start);
  free(sh->face_idx);
}

static int load_shapes(rko_sim* s, const char* path) {
  reader_t r;
  if (rd_open(&r, path) != 0) return seterr(s, "shape library '%s' has not been found", path);
  for (int k = 0; k < s->nshapes; k++) shape_free(&s->shapes[k]);
  free(s->shapes);
  s->shapes = NULL;
  s->nshapes = 0;
  int cap = 0;
  char tok[256];
  while (rd_tok(&r, tok, sizeof tok)) {
    if (strcmp(tok, "<") != 0) continue;
    shape_t sh;
    memset(&sh, 0, sizeof sh);
    sh.obb = obb_default();
    sh.preCompDone = 'n';
    while (rd_tok(&r, tok, sizeof tok)) {
      if (tok[0] == '/' || tok[0] == '#' || tok[0] == '!') {
        rd_skipline(&r);
        continue;
      }
      if (strcmp(tok, ">") == 0) break;
      if (strcmp(tok, "name") == 0) {
        rd_tok(&r, sh.name, sizeof sh.name);
      } else if (strcmp(tok, "radius") == 0) {
        sh.radius = rd_d(&r);
      } else if (strcmp(tok, "preCompDone") == 0) {
        char c[16];
        rd_tok(&r, c, sizeof c);
        sh.preCompDone = c[0];
      } else if (strcmp(tok, "volume") == 0) {
        sh.volume = rd_d(&r);
      } else if (strcmp(tok, "I/m") == 0) {
        sh.inertia_mass = rd_v3(&r);
      } else if (strcmp(tok, "obb.extent") == 0) {
        sh.obb.ext = rd_v3(&r);
      } else if (strcmp(tok, "obb.center") == 0) {
        sh.obb.c = rd_v3(&r);
      } else if (strcmp(tok, "obb.e1") == 0) {
        sh.obb.e[0] = rd_v3(&r);
      } else if (strcmp(tok, "obb.e2") == 0) {
        sh.obb.e[1] = rd_v3(&r);
      } else if (strcmp(tok, "obb.e3") == 0) {
        sh.obb.e[2] = rd_v3(&r);
      } else if (strcmp(tok, "OBBtreeLevel") == 0 || strcmp(tok, "MCnstep") == 0 ||
                 strcmp(tok, "fibObbOption") == 0) {
        rd_d(&r);
      } else if (strcmp(tok, "position") == 0) {
        rd_v3(&r);
      } else if (strcmp(tok, "orientation") == 0) {
        rd_qt(&r);
      } else if (strcmp(tok, "isSurface") == 0) {
      } else if (strcmp(tok, "nv") == 0) {
        sh.nv = (int)rd_l(&r);
        sh.vertex = (v3*)malloc(sizeof(v3) * (size_t)(sh.nv > 0 ? sh.nv : 1));
        for (int v = 0; v < sh.nv; v++) sh.vertex[v] = rd_v3(&r);
      } else if (strcmp(tok, "ne") == 0) {
        sh.ne = (int)rd_l(&r);
        sh.edge = (int*)malloc(sizeof(int) * 2 * (size_t)(sh.ne > 0 ? sh.ne : 1));
        for (int e = 0; e < sh.ne; e++) {
          sh.edge[2 * e] = (int)rd_l(&r);
          sh.edge[2 * e + 1] = (int)rd_l(&r);
        }
      } else if (strcmp(tok, "nf") == 0) {
        sh.nf = (int)rd_l(&r);
        sh.face_start = (int*)malloc(sizeof(int) * (size_t)(sh.nf + 1));
        int fcap = 16, fn = 0;
        sh.face_idx = (int*)malloc(sizeof(int) * (size_t)fcap);
        for (int f = 0; f < sh.nf; f++) {
          int k = (int)rd_l(&r);
          sh.face_start[f] = fn;
          for (int q = 0; q < k; q++) {
            if (fn == fcap) {
              fcap *= 2;
              sh.face_idx = (int*)realloc(sh.face_idx, sizeof(int) * (size_t)fcap);
            }
            sh.face_idx[fn++] = (int)rd_l(&r);
          }
        }
        sh.face_start[sh.nf] = fn;
      } else {
        if (s->verbose) fprintf(stderr, "[oracle] unknown shape keyword '%s'\n", tok);
      }
    }
    if (sh.preCompDone == 'n') {
      rd_close(&r);
      return seterr(s, "shape '%s': preCompDone n (Monte-Carlo mass properties) is not supported", sh.name);
    }
    if (s->nshapes == cap) {
      cap = cap ? 2 * cap : 8;
      s->shapes = (shape_t*)realloc(s->shapes, sizeof(shape_t) * (size_t)cap);
    }
    s->shapes[s->nshapes++] = sh;
  }
  rd_close(&r);
  return 0;
}

static int shape_id(const rko_sim* s, const char* name) {
  for (int k = 0; k < s->nshapes; k++)
    if (strcmp(s->shapes[k].name, name) == 0) return k;
  return -1;
}

/* ------------------------------------------------------------------------- */
/* Body geometry (Particle.cpp:49-86)                                        */
/* ------------------------------------------------------------------------- */
static inline double mink(const rko_sim* s, const body_t* p) { return p->homothety * s->shapes[p->shape].radius; }
static inline v3 glob_vertex(const rko_sim* s, const body_t* p, int v) {
  v3 hv = vsc(p->homothety, s->shapes[p->shape].vertex[v]);
  return vadd(p->pos, qrot(p->Q, hv));
}
static inline v3 glob_face_vertex(const rko_sim* s, const body_t* p, int f, int k) {
  const shape_t* sh = &s->shapes[p->shape];
  return glob_vertex(s, p, sh->face_idx[sh->face_start[f] + k]);
}
static void update_obb(rko_sim* s, body_t* p) {
  obb_t o = s->shapes[p->shape].obb;
  o.e[0] = qrot(p->Q, o.e[0]);
  o.e[1] = qrot(p->Q, o.e[1]);
  o.e[2] = qrot(p->Q, o.e[2]);
  o.c = qrot(p->Q, o.c);
  o.ext = vsc(p->homothety, o.ext);
  o.c = vsc(p->homothety, o.c);
  o.c = vadd(o.c, p->pos);
  p->obb = o;
}

/* ------------------------------------------------------------------------- */
/* Interaction containers                                                     */
/* ------------------------------------------------------------------------- */
static inline int key_cmp(const inter_t* a, int j, int type, int isub, int jsub) {
  if (a->j != j) return a->j < j ? -1 : 1;
  if (a->type != type) return a->type < type ? -1 : 1;
  if (a->isub != isub) return a->isub < isub ? -1 : 1;
  if (a->jsub != jsub) return a->jsub < jsub ? -1 : 1;
  return 0;
}
/* lower bound within one owner list */
static int il_find(const ilist_t* L, int j, int type, int isub, int jsub, int* found) {
  int lo = 0, hi = L->n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key_cmp(&L->a[mid], j, type, isub, jsub) < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  *found = (lo < L->n && key_cmp(&L->a[lo], j, type, isub, jsub) == 0);
  return lo;
}
static void il_insert_at(ilist_t* L, int at, const inter_t* I) {
  if (L->n == L->cap) {
    L->cap = L->cap ? 2 * L->cap : 8;
    L->a = (inter_t*)realloc(L->a, sizeof(inter_t) * (size_t)L->cap);
  }
  memmove(&L->a[at + 1], &L->a[at], sizeof(inter_t) * (size_t)(L->n - at));
  L->a[at] = *I;
  L->n++;
}
static void il_erase_at(ilist_t* L, int at) {
  memmove(&L->a[at], &L->a[at + 1], sizeof(inter_t) * (size_t)(L->n - at - 1));
  L->n--;
}
static inter_t inter_new(int i, int j, int type, int isub, int jsub, double damp, int stick) {
  inter_t I;
  memset(&I, 0, sizeof I);
  I.i = i;
  I.j = j;
  I.type = type;
  I.isub = isub;
  I.jsub = jsub;
  I.damp = damp;
  I.stick = stick;
  return I;
}
/* Interaction::deactivate (Interaction.cpp:123-128) */
static inline void deactivate(inter_t* I) {
  I->dn = I->prev_dn = 0.0;
  I->fn = 0.0;
  I->ft = V(0, 0, 0);
  I->mom = V(0, 0, 0);
}

/* ------------------------------------------------------------------------- */
/* Contact kinematics: Interaction::UpdateDispatcher (Interaction.cpp:147-326) */
/* ------------------------------------------------------------------------- */
static inline v3 rel_vel(const inter_t* I, const body_t* Pi, const body_t* Pj) {
  /* v(Qj) - v(Qi) (Interaction.cpp:162) */
  if (g_var[VAR_MUTANT] == 2) {
    v3 a = vsub(Pj->vel, vcross(Pj->vrot, vsub(I->pos, Pj->pos)));
    v3 b = vsub(Pi->vel, vcross(Pi->vrot, vsub(I->pos, Pi->pos)));
    return vsub(a, b);
  }
  v3 a = vsub(Pj->vel, vcross(vsub(I->pos, Pj->pos), Pj->vrot));
  v3 b = vsub(Pi->vel, vcross(vsub(I->pos, Pi->pos), Pi->vrot));
  return vsub(a, b);
}

static void update_vv(const rko_sim* s, inter_t* I, const body_t* Pi, const body_t* Pj) {
  v3 posi = glob_vertex(s, Pi, I->isub);
  v3 posj = glob_vertex(s, Pj, I->jsub);
  double Ri = mink(s, Pi), Rj = mink(s, Pj);
  if (g_var[VAR_MUTANT] == 3) Rj = Ri;
  I->prev_n = I->n;
  I->n = vsub(posi, posj);
  I->prev_dn = I->dn;
  I->dn = vnormalize(&I->n) - (Ri + Rj);
  I->pos = vsub(posi, vsc(Ri + 0.5 * I->dn, I->n));
  I->vel = rel_vel(I, Pi, Pj);
}

static void update_ve(const rko_sim* s, inter_t* I, const body_t* Pi, const body_t* Pj) {
  const shape_t* shj = &s->shapes[Pj->shape];
  v3 posj1 = glob_vertex(s, Pj, shj->edge[2 * I->jsub]);
  v3 posj2 = glob_vertex(s, Pj, shj->edge[2 * I->jsub + 1]);
  v3 posi = glob_vertex(s, Pi, I->isub);
  v3 E = vsub(posj2, posj1);
  v3 v = vsub(posi, posj1);
  double r = vdot(v, E) / vdot(E, E);
  if (r <= 0.0 || r >= 1.0) {
    deactivate(I);
    return;
  }
  double Ri = mink(s, Pi), Rj = mink(s, Pj);
  I->prev_n = I->n;
  I->n = vsub(posi, vadd(posj1, vsc(r, E)));
  I->prev_dn = I->dn;
  I->dn = vnormalize(&I->n) - (Ri + Rj);
  I->pos = vsub(posi, vsc(Ri + 0.5 * I->dn, I->n));
  I->vel = rel_vel(I, Pi, Pj);
}

/* even-odd crossing test of the projection P against face f of body j */
static int face_crossing(const rko_sim* s, const body_t* Pj, int f, int nbv, v3 P, v3 v1, v3 v2) {
  int odd = 0;
  double ori1 = vdot(P, v1), ori2 = vdot(P, v2);
  for (int iva = 0; iva < nbv; iva++) {
    int ivb = iva + 1;
    if (ivb == nbv) ivb = 0;
    v3 A = glob_face_vertex(s, Pj, f, iva);
    v3 B = glob_face_vertex(s, Pj, f, ivb);
    double pa1 = vdot(A, v1), pb1 = vdot(B, v1);
    double pa2 = vdot(A, v2), pb2 = vdot(B, v2);
    if ((pa2 < ori2 && pb2 >= ori2) || (pb2 < ori2 && pa2 >= ori2)) {
      if (pa1 + (ori2 - pa2) / (pb2 - pa2) * (pb1 - pa1) < ori1) odd = 1 - odd;
    }
  }
  return odd;
}

static void update_vf(const rko_sim* s, inter_t* I, const body_t* Pi, const body_t* Pj) {
  const shape_t* shj = &s->shapes[Pj->shape];
  int nbv = shj->face_start[I->jsub + 1] - shj->face_start[I->jsub];
  v3 A = glob_face_vertex(s, Pj, I->jsub, 0);
  v3 B = glob_face_vertex(s, Pj, I->jsub, 1);
  v3 C = glob_face_vertex(s, Pj, I->jsub, nbv - 1);
  v3 posi = glob_vertex(s, Pi, I->isub);
  v3 v = vsub(posi, A);
  v3 v1 = vsub(B, A);
  vnormalize(&v1);
  v3 v2 = vsub(C, A);
  I->prev_n = I->n;
  I->n = vcross(v1, v2);
  vnormalize(&I->n);
  double dist = vdot(v, I->n);
  if (dist < 0.0) {
    I->n = vneg(I->n);
    dist = -dist;
  }
  v3 P = vsub(posi, vsc(dist, I->n));
  v2 = vcross(I->n, v1);
  if (face_crossing(s, Pj, I->jsub, nbv, P, v1, v2)) {
    double Ri = mink(s, Pi), Rj = mink(s, Pj);
    I->prev_dn = I->dn;
    I->dn = dist - (Ri + Rj);
    I->pos = vsub(posi, vsc(((g_var[VAR_MUTANT] == 7) ? Rj : Ri) + 0.5 * I->dn, I->n));
    I->vel = rel_vel(I, Pi, Pj);
    return;
  }
  deactivate(I);
}

static void update_ee(const rko_sim* s, inter_t* I, const body_t* Pi, const body_t* Pj) {
  const shape_t* shi = &s->shapes[Pi->shape];
  const shape_t* shj = &s->shapes[Pj->shape];
  v3 posi1 = glob_vertex(s, Pi, shi->edge[2 * I->isub]);
  v3 posi2 = glob_vertex(s, Pi, shi->edge[2 * I->isub + 1]);
  v3 posj1 = glob_vertex(s, Pj, shj->edge[2 * I->jsub]);
  v3 posj2 = glob_vertex(s, Pj, shj->edge[2 * I->jsub + 1]);
  v3 Ei = vsub(posi2, posi1);
  v3 Ej = vsub(posj2, posj1);
  v3 v = vsub(posi1, posj1);
  double c = vdot(Ei, Ei), d = vdot(Ej, Ej), e = vdot(Ei, Ej);
  double f = (c * d) - (e * e);
  double sp, tp;
  if (fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    double a = vdot(Ei, v), b = vdot(Ej, v);
    sp = (e * b - a * d) * f;
    tp = (c * b - e * a) * f;
    if (sp <= 0.0 || sp >= 1.0 || tp <= 0.0 || tp >= 1.0) {
      deactivate(I);
      return;
    }
    I->prev_n = I->n;
    I->n = vsub(vadd(posi1, vsc(sp, Ei)), vadd(posj1, vsc(tp, Ej)));
  } else {
    deactivate(I);
    return;
  }
  double Ri = mink(s, Pi), Rj = mink(s, Pj);
  I->prev_dn = I->dn;
  I->dn = vnormalize(&I->n) - (Ri + Rj);
  I->pos = vsub(vadd(posi1, vsc(sp, Ei)), vsc(Ri + 0.5 * I->dn, I->n));
  I->vel = rel_vel(I, Pi, Pj);
}

static void update_interaction(const rko_sim* s, inter_t* I) {
  const body_t* Pi = &s->b[I->i];
  const body_t* Pj = &s->b[I->j];
  switch (I->type) {
    case VV: update_vv(s, I, Pi, Pj); break;
    case VE: update_ve(s, I, Pi, Pj); break;
    case VF: update_vf(s, I, Pi, Pj); break;
    default: update_ee(s, I, Pi, Pj); break;
  }
}

/* ------------------------------------------------------------------------- */
/* Sub-element proximity predicates (Particle.cpp:88-209)                    */
/* ------------------------------------------------------------------------- */
static int near_vv(const rko_sim* s, const body_t* Pi, const body_t* Pj, int isub, int jsub, double dmax) {
  v3 pi = glob_vertex(s, Pi, isub);
  v3 pj = glob_vertex(s, Pj, jsub);
  double sum = mink(s, Pi) + mink(s, Pj) + dmax;
  return vnorm2(vsub(pj, pi)) <= sum * sum;
}
static int near_ve(const rko_sim* s, const body_t* Pi, const body_t* Pj, int isub, int jsub, double dmax) {
  const shape_t* shj = &s->shapes[Pj->shape];
  v3 pi = glob_vertex(s, Pi, isub);
  v3 p0 = glob_vertex(s, Pj, shj->edge[2 * jsub]);
  v3 p1 = glob_vertex(s, Pj, shj->edge[2 * jsub + 1]);
  v3 E = vsub(p1, p0);
  v3 v = vsub(pi, p0);
  double r = vdot(v, E) / vdot(E, E);
  if (r < -dmax) return 0;
  if (r > 1.0 + dmax) return 0;
  double sum = mink(s, Pi) + mink(s, Pj) + dmax;
  return vnorm2(vsub(pi, vadd(p0, vsc(r, E)))) <= sum * sum;
}
static int near_vf(const rko_sim* s, const body_t* Pi, const body_t* Pj, int isub, int jsub, double dmax) {
  const shape_t* shj = &s->shapes[Pj->shape];
  int nbv = shj->face_start[jsub + 1] - shj->face_start[jsub];
  v3 A = glob_face_vertex(s, Pj, jsub, 0);
  v3 B = glob_face_vertex(s, Pj, jsub, 1);
  v3 C = glob_face_vertex(s, Pj, jsub, nbv - 1);
  v3 pi = glob_vertex(s, Pi, isub);
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
  if (face_crossing(s, Pj, jsub, nbv, P, v1, v2) == 1) {
    dist -= (mink(s, Pi) + mink(s, Pj));
    if (dist <= dmax) return 1;
  }
  return 0;
}
static int near_ee(const rko_sim* s, const body_t* Pi, const body_t* Pj, int isub, int jsub, double dmax) {
  const shape_t* shi = &s->shapes[Pi->shape];
  const shape_t* shj = &s->shapes[Pj->shape];
  v3 p0j = glob_vertex(s, Pj, shj->edge[2 * jsub]);
  v3 p1j = glob_vertex(s, Pj, shj->edge[2 * jsub + 1]);
  v3 p0i = glob_vertex(s, Pi, shi->edge[2 * isub]);
  v3 p1i = glob_vertex(s, Pi, shi->edge[2 * isub + 1]);
  v3 Ei = vsub(p1i, p0i);
  v3 Ej = vsub(p1j, p0j);
  v3 v = vsub(p0i, p0j);
  double c = vdot(Ei, Ei), d = vdot(Ej, Ej), e = vdot(Ei, Ej);
  double f = (c * d) - (e * e);
  if (fabs(f) > 1.0e-12) {
    f = 1.0 / f;
    double a = vdot(Ei, v), b = vdot(Ej, v);
    double sp = (e * b - a * d) * f;
    double tp = (c * b - e * a) * f;
    if (sp >= 0.0 && sp <= 1.0 && tp >= 0.0 && tp <= 1.0) {
      v3 n = vsub(vadd(p0i, vsc(sp, Ei)), vadd(p0j, vsc(tp, Ej)));
      double sum = mink(s, Pi) + mink(s, Pj) + dmax;
      return vnorm2(n) <= sum * sum;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Force laws (src/ForceLaws)                                                 */
/* ------------------------------------------------------------------------- */
/* ForceLaw_Default::computeInteraction (ForceLaw_Default.cpp:57-131), FT_CORR off */
static void law_default(rko_sim* s, inter_t* I) {
  if (I->dn > 0.0) {
    I->fn = 0.0;
    I->ft = V(0, 0, 0);
    I->mom = V(0, 0, 0);
    return;
  }
  int g1 = s->b[I->i].group, g2 = s->b[I->j].group;
  double kn = dt_get(s, P_KN, g1, g2), kt = dt_get(s, P_KT, g1, g2), mu = dt_get(s, P_MU, g1, g2);
  double kr = dt_get(s, P_KR, g1, g2), mur = dt_get(s, P_MUR, g1, g2);
  double damp = I->damp;
  double vn = vdot(I->vel, I->n);
  double fne = -kn * I->dn;
  double fnv = damp * vn;
  if (g_var[VAR_MUTANT] == 1) fnv = 0.0;
  I->fn = fne + fnv;
  if (I->fn < 0.0) I->fn = 0.0;
  v3 vt = vsub(I->vel, vsc(vn, I->n));
  I->ft = vadd(I->ft, vsc(kt, vsc(s->dt, vt)));
  double threshold = fabs(mu * ((g_var[VAR_MUTANT] == 8) ? I->fn : fne));
  double ft_square = vdot(I->ft, I->ft);
  if (ft_square > 0.0 && ft_square >= threshold * threshold) I->ft = vsc(threshold / sqrt(ft_square), I->ft);
  if (kr > 0.0) {
    I->mom = vadd(I->mom, vsc(s->dt, vsc(kr, vsub(s->b[I->j].vrot, s->b[I->i].vrot))));
    double threshold_mom = fabs(mur * I->fn);
    double mom_square = vdot(I->mom, I->mom);
    if (mom_square > 0.0 && mom_square >= threshold_mom * threshold_mom)
      I->mom = vsc(threshold_mom / sqrt(mom_square), I->mom);
  }
}

/* ForceLaw_Avalanche::computeInteraction (ForceLaw_Avalanche.cpp:57-136) */
static void law_avalanche(rko_sim* s, inter_t* I) {
  if (I->dn > 0.0) {
    I->fn = 0.0;
    I->ft = V(0, 0, 0);
    I->mom = V(0, 0, 0);
    return;
  }
  int g1 = s->b[I->i].group, g2 = s->b[I->j].group;
  double kn = dt_get(s, P_KN, g1, g2), en2 = dt_get(s, P_EN2, g1, g2), kt = dt_get(s, P_KT, g1, g2);
  double mu = dt_get(s, P_MU, g1, g2), kr = dt_get(s, P_KR, g1, g2), mur = dt_get(s, P_MUR, g1, g2);
  if (I->prev_dn > 0.0) I->prev_dn = 0.0;
  double delta_Dn = I->dn - I->prev_dn;
  if (delta_Dn > 0.0)
    I->fn = -kn * en2 * I->dn;
  else if (delta_Dn < 0.0)
    I->fn += -kn * delta_Dn;
  if (I->fn < 0.0) I->fn = 0.0;
  v3 vt = vsub(I->vel, vsc(vdot(I->vel, I->n), I->n));
  I->ft = vadd(I->ft, vsc(kt, vsc(s->dt, vt)));
  double threshold_ft = fabs(mu * I->fn);
  double ft_square = vdot(I->ft, I->ft);
  if (ft_square > 0.0 && ft_square > threshold_ft * threshold_ft)
    I->ft = vsc(threshold_ft / sqrt(ft_square), I->ft);
  I->mom = vadd(I->mom, vsc(s->dt, vsc(kr, vsub(s->b[I->j].vrot, s->b[I->i].vrot))));
  v3 branch;
  const body_t* P = (I->i < s->nDriven) ? &s->b[I->j] : &s->b[I->i];
  branch = vsub(I->pos, P->pos);
  double r = vdot(branch, P->vrot) / vdot(P->vrot, P->vrot);
  if (g_var[VAR_MUTANT] != 10) branch = vsub(branch, vsc(r, P->vrot));
  double threshold_mom = fabs(mur * sqrt(vnorm2(branch)) * I->fn);
  double mom_square = vdot(I->mom, I->mom);
  if (mom_square > 0.0 && mom_square > threshold_mom * threshold_mom)
    I->mom = vsc(threshold_mom / sqrt(mom_square), I->mom);
}

/* ForceLaw_StickedLinks::computeInteraction (ForceLaw_StickedLinks.cpp:75-215) */
static void law_sticked(rko_sim* s, inter_t* I) {
  if (I->stick >= 0) {
    const iface_t* BI = &s->ifc[I->stick];
    double kn = 0.0, kt = 0.0, kr = 0.0, fn0 = 1.0, ft0 = 1.0, mom0 = 1.0, power = 1.0;
    double dn0 = BI->dn0;
    int isInner = (s->b[I->i].cluster == s->b[I->j].cluster);
    if (s->paramsInInterfaces == 1) {
      isInner = BI->isInner;
      kn = BI->kn;
      kt = BI->kt;
      kr = BI->kr;
      fn0 = BI->fn0;
      ft0 = BI->ft0;
      mom0 = BI->mom0;
      power = BI->power;
    } else {
      int g1 = s->b[I->i].group, g2 = s->b[I->j].group;
      if (isInner) {
        kn = dt_get(s, P_KN_IN, g1, g2);
        kt = dt_get(s, P_KT_IN, g1, g2);
        kr = dt_get(s, P_KR_IN, g1, g2);
        fn0 = dt_get(s, P_FN0_IN, g1, g2);
        ft0 = dt_get(s, P_FT0_IN, g1, g2);
        mom0 = dt_get(s, P_MOM0_IN, g1, g2);
        power = dt_get(s, P_POW_IN, g1, g2);
        dn0 = 0.0;
      } else {
        kn = dt_get(s, P_KN_OUT, g1, g2);
        kt = dt_get(s, P_KT_OUT, g1, g2);
        kr = dt_get(s, P_KR_OUT, g1, g2);
        fn0 = dt_get(s, P_FN0_OUT, g1, g2);
        ft0 = dt_get(s, P_FT0_OUT, g1, g2);
        mom0 = dt_get(s, P_MOM0_OUT, g1, g2);
        power = dt_get(s, P_POW_OUT, g1, g2);
      }
    }
    (void)isInner;
    double vn = vdot(I->vel, I->n);
    double fne = -kn * (I->dn - dn0);
    double fnv = I->damp * vn;
    I->fn = fne + fnv;
    v3 vt = vsub(I->vel, vsc(vn, I->n));
    I->ft = vadd(I->ft, vsc(kt, vsc(s->dt, vt)));
    v3 ftv = vsc(I->damp, vt);
    if (g_var[VAR_MUTANT] != 11) I->ft = vadd(I->ft, ftv);
    I->mom = vadd(I->mom, vsc(s->dt, vsc(kr, vsub(s->b[I->j].vrot, s->b[I->i].vrot))));
    double yield = pow(sqrt(vnorm2(I->ft)) / ft0, power) +
                   ((g_var[VAR_MUTANT] == 12) ? 0.0 : pow(sqrt(vnorm2(I->mom)) / mom0, power)) - I->fn / fn0 - 1.0;
    if (yield > 0.0) s->ifc_break[I->stick] = 1; /* idempotent flag, read after the resultants */
  } else {
    if (I->dn > 0.0) {
      I->fn = 0.0;
      I->ft = V(0, 0, 0);
      I->mom = V(0, 0, 0);
      return;
    }
    int g1 = s->b[I->i].group, g2 = s->b[I->j].group;
    double kn = dt_get(s, P_KN, g1, g2), kt = dt_get(s, P_KT, g1, g2), mu = dt_get(s, P_MU, g1, g2);
    double kr = dt_get(s, P_KR, g1, g2), mur = dt_get(s, P_MUR, g1, g2);
    double damp = I->damp;
    double vn = vdot(I->vel, I->n);
    double fne = -kn * I->dn;
    double fnv = damp * vn;
    I->fn = fne + fnv;
    if (I->fn < 0.0) I->fn = 0.0;
    v3 vt = vsub(I->vel, vsc(vn, I->n));
    I->ft = vadd(I->ft, vsc(kt, vsc(s->dt, vt)));
    double threshold = fabs(mu * ((g_var[VAR_MUTANT] == 16) ? fne : I->fn));
    double ft_square = vdot(I->ft, I->ft);
    if (ft_square > 0.0 && ft_square >= threshold * threshold) I->ft = vsc(threshold / sqrt(ft_square), I->ft);
    I->mom = vadd(I->mom, vsc(s->dt, vsc(kr, vsub(s->b[I->j].vrot, s->b[I->i].vrot))));
    double threshold_mom = fabs(mur * I->fn);
    double mom_square = vdot(I->mom, I->mom);
    if (mom_square > 0.0 && mom_square >= threshold_mom * threshold_mom)
      I->mom = vsc(threshold_mom / sqrt(mom_square), I->mom);
  }
}

/* ------------------------------------------------------------------------- */
/* Body forces (src/BodyForces)                                               */
/* ------------------------------------------------------------------------- */
static void body_force(const rko_sim* s, int i, v3* F, v3* M) {
  const body_t* p = &s->b[i];
  if (s->bf == BF_ATTRACTING) { /* BodyForce_AttractingPoint.cpp:51-59 */
    v3 dir = vsub(s->bf_point, p->pos);
    double l2 = vnorm2(dir);
    if (l2 > 0.0) {
      double inv = 1.0 / sqrt(l2);
      dir = V(dir.x * inv, dir.y * inv, dir.z * inv);
      *F = vsc(p->mass * s->bf_acc, dir);
    }
  } else if (s->bf == BF_PREFDIR) { /* BodyForce_PreferredDirection.cpp:54-75 */
    v3 axisB = qrot(p->Q, s->bf_axisBody);
    vnormalize(&axisB);
    v3 u = vcross(s->bf_axis, axisB);
    vnormalize(&u);
    double cosa = vdot(axisB, s->bf_axis);
    double angle = acos(cosa);
    double sgn = -1.0;
    if (cosa < 0.0 && g_var[VAR_MUTANT] != 13) {
      sgn = 1.0;
      angle = M_PI - angle;
    }
    *M = vsc(angle, vsc(s->bf_Kr * sgn, u));
  } else if (s->bf == BF_VISCOUS) { /* BodyForce_ViscousFluid.cpp:51-62 */
    v3 velSqr = V(p->vel.x * p->vel.x, p->vel.y * p->vel.y, p->vel.z * p->vel.z);
    double h = p->homothety;
    double pVolume = s->shapes[p->shape].volume * h * h * h;
    double pSurface = pow(pVolume, 2.0 / 3.0);
    *F = vsc(0.5 * 0.47 * s->bf_fluidDensity * pSurface, velSqr);
  }
}

/* ------------------------------------------------------------------------- */
/* accelerations() and its parts                                              */
/* ------------------------------------------------------------------------- */
static void initialise_forces(rko_sim* s) { /* Rockable.cpp:2869-2929 */
  int nd = s->nDriven < s->nb ? s->nDriven : s->nb;
  for (int i = 0; i < nd; i++) {
    s->b[i].force = V(0, 0, 0);
    s->b[i].acc = V(0, 0, 0);
    s->b[i].moment = V(0, 0, 0);
    s->b[i].arot = V(0, 0, 0);
  }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = nd; i < s->nb; i++) {
    s->b[i].force = V(0, 0, 0);
    s->b[i].moment = V(0, 0, 0);
    if (s->bf != BF_NONE) {
      v3 fm = V(0, 0, 0), mm = V(0, 0, 0);
      body_force(s, i, &fm, &mm);
      s->b[i].force = vadd(s->b[i].force, fm);
      s->b[i].moment = vadd(s->b[i].moment, mm);
    }
  }
  for (int c = 0; c < s->nctrl; c++) {
    control_t* C = &s->ctrl[c];
    body_t* p = &s->b[C->i];
    switch (C->type) {
      case CTRL_X_FOR: p->force.x = C->value; break;
      case CTRL_Y_FOR: p->force.y = C->value; break;
      case CTRL_Z_FOR: p->force.z = C->value; break;
      case CTRL_XROT_MOM: p->moment.x = C->value; break;
      case CTRL_YROT_MOM: p->moment.y = C->value; break;
      case CTRL_ZROT_MOM: p->moment.z = C->value; break;
      case CTRL_XYZROT_MOM: p->moment = C->vec_value; break;
      default: break;
    }
  }
}

static void update_interactions(rko_sim* s) { /* Rockable.cpp:2968-2974 */
#pragma omp parallel for schedule(dynamic, 64) num_threads(s->nthreads)
  for (int k = 0; k < s->nb; k++) {
    ilist_t* L = &s->inter[k];
    for (int q = 0; q < L->n; q++) update_interaction(s, &L->a[q]);
  }
}

/* activeInteractions from the entries' flags, in list order */
static void active_from_flags(rko_sim* s) {
  size_t tot = 0;
  for (int k = 0; k < s->nb; k++) tot += (size_t)s->inter[k].n;
  if (tot > s->cap_active) {
    s->cap_active = tot + tot / 2 + 16;
    s->active = (inter_t**)realloc(s->active, sizeof(inter_t*) * s->cap_active);
  }
  size_t na = 0;
  for (int k = 0; k < s->nb; k++) {
    ilist_t* L = &s->inter[k];
    for (int q = 0; q < L->n; q++)
      if (L->a[q].act) s->active[na++] = &L->a[q];
  }
  s->nactive = na;
}

/* build_activeInteractions (Rockable.cpp:3005-3029): the listed entries with
   dn < 0 or a bond.  The reference keeps the vector's pointers until the next
   call, through UpdateNL, which keeps the interactions it does not erase in
   place (std::set nodes): the active set, a bond broken in this step
   included, holds until the next accelerations().  Each entry carries the
   membership as a flag, which the list updates move with it; the outputs
   (saveConf, the statistics, dnStat) read the set back from the flags. */
static void build_active(rko_sim* s) {
  for (int k = 0; k < s->nb; k++) {
    ilist_t* L = &s->inter[k];
    for (int q = 0; q < L->n; q++) L->a[q].act = (L->a[q].dn < 0.0 || L->a[q].stick >= 0);
  }
  active_from_flags(s);
}

static void compute_forces(rko_sim* s) { /* Rockable.cpp:3035-3046 */
  long na = (long)s->nactive;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (long q = 0; q < na; q++) {
    inter_t* I = s->active[q];
    if (I->dn < 0.0 || I->stick >= 0) {
      if (s->law == LAW_AVALANCHE)
        law_avalanche(s, I);
      else if (s->law == LAW_STICKED)
        law_sticked(s, I);
      else
        law_default(s, I);
    }
  }
}

static void compute_resultants(rko_sim* s) { /* Rockable.cpp:3267-3272, 2752-2789 (serial) */
  for (size_t q = 0; q < s->nactive; q++) {
    inter_t* I = s->active[q];
    if (!(I->dn < 0.0 || I->stick >= 0)) continue;
    body_t* Pi = &s->b[I->i];
    body_t* Pj = &s->b[I->j];
    v3 f = vadd(vsc(I->fn, I->n), I->ft);
    Pi->force = vadd(Pi->force, f);
    Pj->force = vsub(Pj->force, f);
    v3 Ci = vsub(I->pos, Pi->pos);
    v3 Cj = vsub(I->pos, vadd(Pj->pos, V(0.0, 0.0, 0.0))); /* + jPeriodicShift (zero) */
    Pi->moment = vadd(Pi->moment, vadd(vcross(Ci, f), I->mom));
    if (g_var[VAR_MUTANT] == 4)
      Pj->moment = vadd(Pj->moment, vadd(vcross(Cj, vneg(f)), I->mom));
    else
      Pj->moment = vadd(Pj->moment, vsub(vcross(Cj, vneg(f)), I->mom));
  }
}

static inter_t* find_inter(rko_sim* s, int owner, int j, int type, int isub, int jsub) {
  if (owner < 0 || owner >= s->nb) return NULL;
  int found;
  int at = il_find(&s->inter[owner], j, type, isub, jsub, &found);
  return found ? &s->inter[owner].a[at] : NULL;
}

static void breakage(rko_sim* s) { /* Rockable.cpp:3350-3413; interfaces visited in (i, j) order */
  for (int k = 0; k < s->nifc; k++) {
    if (!s->ifc_break[k]) continue;
    s->ifc_break[k] = 0;
    iface_t* BI = &s->ifc[k];
    if (!BI->alive) continue;
    int g1 = s->b[BI->i].group, g2 = s->b[BI->j].group;
    for (int q = 0; q < BI->nb; q++) {
      inter_t* I = find_inter(s, BI->i, BI->j, BI->bkey[3 * q], BI->bkey[3 * q + 1], BI->bkey[3 * q + 2]);
      if (!I) continue;
      I->stick = -1;
      if (I->dn >= 0.0 || g_var[VAR_MUTANT] == 17) {
        I->fn = 0.0;
        I->ft = V(0, 0, 0);
      } else {
        v3 T = I->ft;
        double FT = vnormalize(&T);
        double mu = dt_get(s, P_MU, g1, g2);
        double thr = fabs(mu * I->fn);
        if (FT > thr) I->ft = vsc(thr, T);
      }
    }
    s->totalBrokenArea += BI->area;
    BI->alive = 0;
    s->needUpdate = 1;
  }
}

static inline v3 euler_domega(const body_t* p, v3 omega, v3 M) {
  if (g_var[VAR_MUTANT] == 5)
    return V((M.x - (p->inertia.y - p->inertia.z) * omega.y * omega.z) / p->inertia.x,
             (M.y - (p->inertia.x - p->inertia.z) * omega.z * omega.x) / p->inertia.y,
             (M.z - (p->inertia.y - p->inertia.x) * omega.x * omega.y) / p->inertia.z);
  return V((M.x - (p->inertia.z - p->inertia.y) * omega.y * omega.z) / p->inertia.x,
           (M.y - (p->inertia.x - p->inertia.z) * omega.z * omega.x) / p->inertia.y,
           (M.z - (p->inertia.y - p->inertia.x) * omega.x * omega.y) / p->inertia.z);
}

static void accelerations_from_resultants(rko_sim* s) { /* Rockable.cpp:3419-3552 */
  for (int c = 0; c < s->nctrl; c++) {
    control_t* C = &s->ctrl[c];
    body_t* p = &s->b[C->i];
    switch (C->type) {
      case CTRL_X_FOR: p->acc.x = p->force.x / p->mass; break;
      case CTRL_Y_FOR: p->acc.y = p->force.y / p->mass; break;
      case CTRL_Z_FOR: p->acc.z = p->force.z / p->mass; break;
      case CTRL_XROT_MOM:
      case CTRL_YROT_MOM:
      case CTRL_ZROT_MOM:
      case CTRL_XYZROT_MOM: {
        qt Qinv = qconj(p->Q);
        v3 w;
        if (C->type == CTRL_XROT_MOM)
          w = V(p->vrot.x, 0.0, 0.0);
        else if (C->type == CTRL_YROT_MOM)
          w = V(0.0, p->vrot.y, 0.0);
        else if (C->type == CTRL_ZROT_MOM)
          w = V(0.0, 0.0, p->vrot.z);
        else
          w = p->vrot;
        v3 omega = qrot(Qinv, w);
        v3 M = (g_var[VAR_MUTANT] == 15) ? p->moment : qrot(Qinv, p->moment);
        v3 dw = qrot(p->Q, euler_domega(p, omega, M));
        if (C->type == CTRL_XROT_MOM)
          p->arot.x = dw.x;
        else if (C->type == CTRL_YROT_MOM)
          p->arot.y = dw.y;
        else if (C->type == CTRL_ZROT_MOM)
          p->arot.z = dw.z;
        else
          p->arot = dw;
      } break;
      default: break;
    }
  }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = s->nDriven; i < s->nb; i++) {
    body_t* p = &s->b[i];
    p->acc = vadd(vdiv(p->force, p->mass), s->gravity);
    qt Qinv = qconj(p->Q);
    v3 omega = qrot(Qinv, p->vrot);
    v3 M = qrot(Qinv, p->moment);
    p->arot = qrot(p->Q, euler_domega(p, omega, M));
  }
}

void rko_update_kinematics(rko_sim* s) { update_interactions(s); }

/* Resultants and accelerations from the current interaction forces, without
 * running the force law (used to pin incrementResultants and
 * compute_accelerations_from_resultants against stored conf records). */
void rko_resultants_only(rko_sim* s) {
  initialise_forces(s);
  build_active(s);
  compute_resultants(s);
  accelerations_from_resultants(s);
}

void rko_accelerations(rko_sim* s) { /* Rockable.cpp:3562-3719, DefaultParallelMode */
  initialise_forces(s);
  update_interactions(s);
  build_active(s);
  compute_forces(s);
  compute_resultants(s);
  breakage(s);
  accelerations_from_resultants(s);
}

/* ------------------------------------------------------------------------- */
/* Driving system (DrivingSystem.cpp:242-268) and velocity-Verlet step        */
/* ------------------------------------------------------------------------- */
/* The ServoFunction lambdas of DrivingSystem::read (DrivingSystem.cpp:251-494),
 * called with the box time t at the start of each step and after each conf
 * save.  The tritri servos read the six walls' inner faces; the shakers and
 * the ramp depend on t only. */
static void servo(rko_sim* s) {
  const int ic = s->servo_icontr;
  if (s->servo >= 1 && s->servo <= 4) {
    if (ic + 5 >= s->nctrl) return;
    const int* id = s->servo_ids;
    const double Xmin = s->b[id[0]].pos.x + mink(s, &s->b[id[0]]);
    const double Xmax = s->b[id[1]].pos.x - mink(s, &s->b[id[1]]);
    const double Ymin = s->b[id[2]].pos.y + mink(s, &s->b[id[2]]);
    const double Ymax = s->b[id[3]].pos.y - mink(s, &s->b[id[3]]);
    const double Zmin = s->b[id[4]].pos.z + mink(s, &s->b[id[4]]);
    const double Zmax = s->b[id[5]].pos.z - mink(s, &s->b[id[5]]);
    const double dX = fabs(Xmax - Xmin), dY = fabs(Ymax - Ymin), dZ = fabs(Zmax - Zmin);
    const double surfX = dY * dZ, surfY = dX * dZ, surfZ = dY * dX;
    const double p = s->servo_pressure;
    control_t* C = s->ctrl + ic;
    if (s->servo == 1) { /* 251-268 */
      C[1].value = -p * surfX;
      C[3].value = -p * surfY;
      C[5].value = -p * surfZ;
    } else if (s->servo == 2) { /* 280-298 */
      C[1].value = -p * ((g_var[VAR_MUTANT] == 14) ? surfY : surfX);
      C[5].value = -p * surfZ;
    } else if (s->servo == 3) { /* 328-353 */
      const double* v = s->servo_val;
      if (C[0].type == CTRL_X_FOR) C[0].value = v[0] * surfX;
      if (C[1].type == CTRL_X_FOR) C[1].value = -v[1] * surfX;
      if (C[2].type == CTRL_Y_FOR) C[2].value = v[2] * surfY;
      if (C[3].type == CTRL_Y_FOR) C[3].value = -v[3] * surfY;
      if (C[4].type == CTRL_Z_FOR) C[4].value = v[4] * surfZ;
      if (C[5].type == CTRL_Z_FOR) C[5].value = -v[5] * surfZ;
    } else { /* 369-392 */
      const double dSig = s->lode_rate * s->t;
      C[1].value = -(p + dSig) * surfX;
      C[3].value = -(p + s->lode_a * dSig) * surfY;
      C[5].value = -(p + s->lode_b * dSig) * surfZ;
    }
  } else if (s->servo >= 5 && s->servo <= 7) {
    if (ic + 2 >= s->nctrl) return;
    v3 v;
    if (s->servo == 5) { /* 413-418 */
      const double a = s->shake_A * s->shake_omega * cos(s->shake_omega * s->t);
      v = vsc(a, s->shake_dir);
    } else { /* 436-444, 467-475 */
      const double T = s->shake_T;
      v = vsc(4.0 * s->shake_A / T, s->shake_dir);
      double phase = (s->servo == 6) ? s->t - floor(s->t / T) * T : s->t - floor((s->t - s->shake_tini) / T) * T;
      phase /= T;
      if (phase >= 0.25 && phase < 0.75) v = vsc(-1.0, v);
    }
    s->ctrl[ic].value = v.x;
    s->ctrl[ic + 1].value = v.y;
    s->ctrl[ic + 2].value = v.z;
  } else if (s->servo == 8) { /* 485-493 */
    if (ic >= s->nctrl) return;
    if (!s->ramp_slope_set) {
      s->ramp_slope = (s->ramp_ve - s->ramp_vb) / (s->ramp_te - s->ramp_tb);
      s->ramp_slope_set = 1;
    }
    if (s->t <= s->ramp_tb)
      s->ctrl[ic].value = s->ramp_vb;
    else if (s->t >= s->ramp_te)
      s->ctrl[ic].value = s->ramp_ve;
    else
      s->ctrl[ic].value = s->ramp_vb + s->ramp_slope * (s->t - s->ramp_tb);
  }
}

static void velocity_controlled_drive(rko_sim* s) { /* Rockable.cpp:1856-1900 */
  for (int c = 0; c < s->nctrl; c++) {
    control_t* C = &s->ctrl[c];
    body_t* p = &s->b[C->i];
    switch (C->type) {
      case CTRL_X_VEL:
        p->vel.x = C->value;
        p->pos.x += s->dt * C->value;
        break;
      case CTRL_Y_VEL:
        p->vel.y = C->value;
        p->pos.y += s->dt * C->value;
        break;
      case CTRL_Z_VEL:
        p->vel.z = C->value;
        p->pos.z += s->dt * C->value;
        break;
      case CTRL_XROT_VEL:
      case CTRL_YROT_VEL:
      case CTRL_ZROT_VEL:
      case CTRL_XYZROT_VEL: {
        v3 w;
        if (C->type == CTRL_XROT_VEL)
          w = V(C->value, 0.0, 0.0);
        else if (C->type == CTRL_YROT_VEL)
          w = V(0.0, C->value, 0.0);
        else if (C->type == CTRL_ZROT_VEL)
          w = V(0.0, 0.0, C->value);
        else
          w = C->vec_value;
        p->Q = qadd_scaled(p->Q, qdot(p->Q, w), s->dt);
        qnormalize(&p->Q);
      } break;
      default: break;
    }
  }
}

void rko_vv_step(rko_sim* s) { /* Rockable.cpp:1979-2158 */
  const double dt = s->dt, dt_2 = s->dt_2, dt2_2 = s->dt2_2;
  servo(s);
  velocity_controlled_drive(s);
  for (int c = 0; c < s->nctrl; c++) {
    control_t* C = &s->ctrl[c];
    body_t* p = &s->b[C->i];
    switch (C->type) {
      case CTRL_X_FOR:
        p->pos.x += dt * p->vel.x + dt2_2 * p->acc.x;
        p->vel.x += dt_2 * p->acc.x;
        break;
      case CTRL_Y_FOR:
        p->pos.y += dt * p->vel.y + dt2_2 * p->acc.y;
        p->vel.y += dt_2 * p->acc.y;
        break;
      case CTRL_Z_FOR:
        p->pos.z += dt * p->vel.z + dt2_2 * p->acc.z;
        p->vel.z += dt_2 * p->acc.z;
        break;
      case CTRL_XROT_MOM:
        p->Q = qadd_scaled(p->Q, qdot(p->Q, V(p->vrot.x * 1.0, p->vrot.x * 0.0, p->vrot.x * 0.0)), dt);
        qnormalize(&p->Q);
        p->vrot.x += dt_2 * p->arot.x;
        break;
      case CTRL_YROT_MOM:
        p->Q = qadd_scaled(p->Q, qdot(p->Q, V(p->vrot.y * 0.0, p->vrot.y * 1.0, p->vrot.y * 0.0)), dt);
        qnormalize(&p->Q);
        p->vrot.y += dt_2 * p->arot.y;
        break;
      case CTRL_ZROT_MOM:
        p->Q = qadd_scaled(p->Q, qdot(p->Q, V(p->vrot.z * 0.0, p->vrot.z * 0.0, p->vrot.z * 1.0)), dt);
        qnormalize(&p->Q);
        p->vrot.z += dt_2 * p->arot.z;
        break;
      case CTRL_XYZROT_MOM:
        p->Q = qadd_scaled(p->Q, qdot(p->Q, p->vrot), dt);
        qnormalize(&p->Q);
        p->vrot = vadd(p->vrot, vsc(dt_2, p->arot));
        break;
      default: break;
    }
  }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = s->nDriven; i < s->nb; i++) {
    body_t* p = &s->b[i];
    p->pos = vadd(p->pos, vadd(vsc(dt, p->vel), vsc(dt2_2, p->acc)));
    p->vel = vadd(p->vel, vsc(dt_2, p->acc));
    p->Q = qadd_scaled(p->Q, qdot(p->Q, p->vrot), dt);
    p->Q = qadd_scaled(p->Q, qddot(p->Q, p->vrot, p->arot), dt2_2);
    qnormalize(&p->Q);
    p->vrot = vadd(p->vrot, vsc(dt_2, p->arot));
  }
  rko_accelerations(s);
  for (int c = 0; c < s->nctrl; c++) {
    control_t* C = &s->ctrl[c];
    body_t* p = &s->b[C->i];
    switch (C->type) {
      case CTRL_X_FOR: p->vel.x += dt_2 * p->acc.x; break;
      case CTRL_Y_FOR: p->vel.y += dt_2 * p->acc.y; break;
      case CTRL_Z_FOR: p->vel.z += dt_2 * p->acc.z; break;
      case CTRL_XROT_MOM: p->vrot.x += dt_2 * p->arot.x; break;
      case CTRL_YROT_MOM: p->vrot.y += dt_2 * p->arot.y; break;
      case CTRL_ZROT_MOM: p->vrot.z += dt_2 * p->arot.z; break;
      case CTRL_XYZROT_MOM: p->vrot = vadd(p->vrot, vsc(dt_2, p->arot)); break;
      default: break;
    }
  }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = s->nDriven; i < s->nb; i++) {
    body_t* p = &s->b[i];
    double k = (g_var[VAR_MUTANT] == 6) ? dt : dt_2;
    p->vel = vadd(p->vel, vsc(k, p->acc));
    p->vrot = vadd(p->vrot, vsc(k, p->arot));
  }
}

/* ------------------------------------------------------------------------- */
/* Neighbour list                                                             */
/* ------------------------------------------------------------------------- */
typedef int (*near_fn)(const rko_sim*, const body_t*, const body_t*, int, int, double);

/* addOrRemoveSingleInteraction (Rockable.cpp:1316-1340) */
static void add_or_remove(rko_sim* s, int i_, int j_, int isub, int type, int nbj, near_fn fn, double dmax,
                          double damp) {
  ilist_t* L = &s->inter[i_];
  omp_set_lock(&s->locks[i_]);
  for (int jsub = 0; jsub < nbj; jsub++) {
    int found;
    int at = il_find(L, j_, type, isub, jsub, &found);
    int isnear = fn(s, &s->b[i_], &s->b[j_], isub, jsub, dmax);
    if (isnear && !found) {
      inter_t I = inter_new(i_, j_, type, isub, jsub, damp, -1);
      il_insert_at(L, at, &I);
    } else if (!isnear && found) {
      il_erase_at(L, at);
    }
  }
  omp_unset_lock(&s->locks[i_]);
}

/* Damp for a body pair (Rockable.cpp:1356-1376) */
static double pair_damp(const rko_sim* s, int i, int j) {
  double en2 = dt_get(s, P_EN2, s->b[i].group, s->b[j].group);
  double kn = dt_get(s, P_KN, s->b[i].group, s->b[j].group);
  double meff;
  if (i < s->nDriven)
    meff = s->b[j].mass;
  else if (j < s->nDriven)
    meff = s->b[i].mass;
  else
    meff = (s->b[i].mass * s->b[j].mass) / (s->b[i].mass + s->b[j].mass);
  if (en2 > 0.0 && en2 < 1.0) {
    double logen = 0.5 * log(en2);
    double dampRate = -logen / sqrt(logen * logen + M_PI * M_PI);
    return dampRate * 2.0 * sqrt(kn * meff);
  } else if (en2 <= 0.0) {
    return (g_var[VAR_MUTANT] == 9) ? 0.0 : 2.0 * sqrt(kn * meff);
  }
  return 0.0;
}

/* AddOrRemoveInteractions_bruteForce (Rockable.cpp:1309-1432) */
static void add_or_remove_pair(rko_sim* s, int i, int j, double dmax) {
  const body_t* Pi = &s->b[i];
  const body_t* Pj = &s->b[j];
  const shape_t* shi = &s->shapes[Pi->shape];
  const shape_t* shj = &s->shapes[Pj->shape];
  double damp = pair_damp(s, i, j);
  for (int isub = 0; isub < shi->nv; isub++) {
    obb_t sub = obb_default();
    sub.c = glob_vertex(s, Pi, isub);
    obb_enlarge(&sub, mink(s, Pi) + dmax);
    obb_t oj = Pj->obb;
    obb_enlarge(&oj, dmax);
    if (!obb_intersect(&oj, &sub)) continue;
    add_or_remove(s, i, j, isub, VV, shj->nv, near_vv, dmax, damp);
    add_or_remove(s, i, j, isub, VE, shj->ne, near_ve, dmax, damp);
    add_or_remove(s, i, j, isub, VF, shj->nf, near_vf, dmax, damp);
  }
  for (int jsub = 0; jsub < shj->nv; jsub++) {
    obb_t sub = obb_default();
    sub.c = glob_vertex(s, Pj, jsub);
    obb_enlarge(&sub, mink(s, Pj) + dmax);
    obb_t oi = Pi->obb;
    obb_enlarge(&oi, dmax);
    if (!obb_intersect(&oi, &sub)) continue;
    add_or_remove(s, j, i, jsub, VE, shi->ne, near_ve, dmax, damp);
    add_or_remove(s, j, i, jsub, VF, shi->nf, near_vf, dmax, damp);
  }
  for (int isub = 0; isub < shi->ne; isub++) add_or_remove(s, i, j, isub, EE, shj->ne, near_ee, dmax, damp);
}

static int has_interface(const rko_sim* s, int i, int j) {
  /* Interfaces[i].find((i, j)) -- interfaces are stored under their i */
  for (int k = 0; k < s->nifc; k++)
    if (s->ifc[k].alive && s->ifc[k].i == i && s->ifc[k].j == j) return 1;
  return 0;
}

/* per-body sorted index of live interfaces, rebuilt before a NL update */
typedef struct {
  int* start;
  int* j;
} ifc_index_t;
static ifc_index_t ifc_index_build(const rko_sim* s) {
  ifc_index_t x;
  x.start = (int*)calloc((size_t)s->nb + 1, sizeof(int));
  int n = 0;
  for (int k = 0; k < s->nifc; k++)
    if (s->ifc[k].alive && s->ifc[k].i >= 0 && s->ifc[k].i < s->nb) {
      x.start[s->ifc[k].i + 1]++;
      n++;
    }
  for (int b = 0; b < s->nb; b++) x.start[b + 1] += x.start[b];
  x.j = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int* fill = (int*)calloc((size_t)s->nb, sizeof(int));
  for (int k = 0; k < s->nifc; k++)
    if (s->ifc[k].alive && s->ifc[k].i >= 0 && s->ifc[k].i < s->nb) {
      int b = s->ifc[k].i;
      x.j[x.start[b] + fill[b]++] = s->ifc[k].j;
    }
  free(fill);
  return x;
}
static inline int ifc_index_has(const ifc_index_t* x, int i, int j) {
  for (int q = x->start[i]; q < x->start[i + 1]; q++)
    if (x->j[q] == j) return 1;
  return 0;
}

/* UpdateNL_bruteForce (Rockable.cpp:1614-1662) */
static void update_nl_brute(rko_sim* s) {
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = 0; i < s->nb; i++) update_obb(s, &s->b[i]);
  ifc_index_t xi = ifc_index_build(s);
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
  for (int i = 0; i < s->nb; i++) {
    obb_t oi = s->b[i].obb;
    obb_enlarge(&oi, 0.5 * s->DVerlet);
    int jnext = i + 1;
    if (jnext < s->nDriven) jnext = s->nDriven;
    for (int j = jnext; j < s->nb; j++) {
      if (ifc_index_has(&xi, i, j)) continue;
      obb_t oj = s->b[j].obb;
      obb_enlarge(&oj, 0.5 * s->DVerlet);
      if (obb_intersect(&oi, &oj)) add_or_remove_pair(s, i, j, s->dVerlet);
    }
  }
  free(xi.start);
  free(xi.j);
}

/* body AABB (Rockable.cpp:3925-3948) */
typedef struct {
  v3 mn, mx;
} aabb_t;
static aabb_t body_aabb(const rko_sim* s, const body_t* p) {
  aabb_t a;
  v3 v0 = glob_vertex(s, p, 0);
  a.mn = a.mx = v0;
  int nv = s->shapes[p->shape].nv;
  for (int v = 1; v < nv; v++) {
    v3 g = glob_vertex(s, p, v);
    if (g.x < a.mn.x) a.mn.x = g.x;
    if (g.y < a.mn.y) a.mn.y = g.y;
    if (g.z < a.mn.z) a.mn.z = g.z;
    if (g.x > a.mx.x) a.mx.x = g.x;
    if (g.y > a.mx.y) a.mx.y = g.y;
    if (g.z > a.mx.z) a.mx.z = g.z;
  }
  double r = mink(s, p);
  a.mn = V(a.mn.x - r, a.mn.y - r, a.mn.z - r);
  a.mx = V(a.mx.x + r, a.mx.y + r, a.mx.z + r);
  return a;
}

/* Cell grid of the linkCells broad phase (toofus linkCells re-derived):
 * N = max(1, floor(box size / minSizes)) cells per axis, cell size =
 * box size / N; a body goes to the cell holding its centre unless its AABB is
 * larger than a cell along any axis, in which case it is "oversized"; each
 * cell's neighbourhood is the (up to) 27 cells around it, itself included. */
typedef struct {
  int N[3];
  v3 mn;
  double sz[3];
} grid_t;

static int grid_dim(double len, double minsz) {
  int n = (minsz > 0.0) ? (int)floor(len / minsz) : 1;
  return n < 1 ? 1 : n;
}

/* UpdateNL_linkCells (Rockable.cpp:1667-1809) */
static void update_nl_cells(rko_sim* s) {
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = 0; i < s->nb; i++) update_obb(s, &s->b[i]);
  aabb_t* pa = (aabb_t*)malloc(sizeof(aabb_t) * (size_t)s->nb);
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int i = 0; i < s->nb; i++) pa[i] = body_aabb(s, &s->b[i]);
  int first = (s->boxForLinkCellsOpt == 1) ? s->nDriven : 0;
  if (first >= s->nb) first = 0;
  aabb_t box = pa[first];
  for (int i = first + 1; i < s->nb; i++) {
    if (pa[i].mn.x < box.mn.x) box.mn.x = pa[i].mn.x;
    if (pa[i].mn.y < box.mn.y) box.mn.y = pa[i].mn.y;
    if (pa[i].mn.z < box.mn.z) box.mn.z = pa[i].mn.z;
    if (pa[i].mx.x > box.mx.x) box.mx.x = pa[i].mx.x;
    if (pa[i].mx.y > box.mx.y) box.mx.y = pa[i].mx.y;
    if (pa[i].mx.z > box.mx.z) box.mx.z = pa[i].mx.z;
  }
  grid_t g;
  g.mn = box.mn;
  g.N[0] = grid_dim(box.mx.x - box.mn.x, s->cellMinSizes.x);
  g.N[1] = grid_dim(box.mx.y - box.mn.y, s->cellMinSizes.y);
  g.N[2] = grid_dim(box.mx.z - box.mn.z, s->cellMinSizes.z);
  g.sz[0] = (box.mx.x - box.mn.x) / g.N[0];
  g.sz[1] = (box.mx.y - box.mn.y) / g.N[1];
  g.sz[2] = (box.mx.z - box.mn.z) / g.N[2];
  long ncell = (long)g.N[0] * g.N[1] * g.N[2];
  int* cell_of = (int*)malloc(sizeof(int) * (size_t)s->nb);
  int* cstart = (int*)calloc((size_t)ncell + 1, sizeof(int));
  int noversized = 0;
  for (int i = 0; i < s->nb; i++) {
    int over = (pa[i].mx.x - pa[i].mn.x > g.sz[0]) || (pa[i].mx.y - pa[i].mn.y > g.sz[1]) ||
               (pa[i].mx.z - pa[i].mn.z > g.sz[2]);
    if (over) {
      cell_of[i] = -1;
      noversized++;
      continue;
    }
    int c3[3];
    double pp[3] = {s->b[i].pos.x, s->b[i].pos.y, s->b[i].pos.z};
    double m3[3] = {g.mn.x, g.mn.y, g.mn.z};
    for (int a = 0; a < 3; a++) {
      int k = (int)floor((pp[a] - m3[a]) / g.sz[a]);
      if (k < 0) k = 0;
      if (k >= g.N[a]) k = g.N[a] - 1;
      c3[a] = k;
    }
    cell_of[i] = (c3[0] * g.N[1] + c3[1]) * g.N[2] + c3[2];
    cstart[cell_of[i] + 1]++;
  }
  for (long c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  int* cbody = (int*)malloc(sizeof(int) * (size_t)(s->nb > 0 ? s->nb : 1));
  int* cfill = (int*)calloc((size_t)ncell, sizeof(int));
  int* over = (int*)malloc(sizeof(int) * (size_t)(noversized > 0 ? noversized : 1));
  int no = 0;
  for (int i = 0; i < s->nb; i++) {
    if (cell_of[i] < 0)
      over[no++] = i;
    else
      cbody[cstart[cell_of[i]] + cfill[cell_of[i]]++] = i;
  }
  ifc_index_t xi = ifc_index_build(s);
  /* main loop: every cell, its neighbour cells, i in the cell, j in the neighbour, j > i */
#pragma omp parallel for schedule(dynamic, 16) num_threads(s->nthreads)
  for (long c = 0; c < ncell; c++) {
    int cz = (int)(c % g.N[2]);
    int cy = (int)((c / g.N[2]) % g.N[1]);
    int cx = (int)(c / ((long)g.N[2] * g.N[1]));
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          int nx = cx + dx, ny = cy + dy, nz = cz + dz;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= g.N[0] || ny >= g.N[1] || nz >= g.N[2]) continue;
          long nc = ((long)nx * g.N[1] + ny) * g.N[2] + nz;
          for (int q = cstart[c]; q < cstart[c + 1]; q++) {
            int i = cbody[q];
            obb_t oi = s->b[i].obb;
            obb_enlarge(&oi, 0.5 * s->DVerlet);
            for (int r = cstart[nc]; r < cstart[nc + 1]; r++) {
              int j = cbody[r];
              if (j < s->nDriven && i < s->nDriven) continue;
              if (j <= i) continue;
              if (ifc_index_has(&xi, i, j)) continue;
              obb_t oj = s->b[j].obb;
              obb_enlarge(&oj, 0.5 * s->DVerlet);
              if (obb_intersect(&oi, &oj)) add_or_remove_pair(s, i, j, s->dVerlet);
            }
          }
        }
  }
  /* oversized bodies against the bodies of every cell */
#pragma omp parallel for schedule(dynamic, 16) num_threads(s->nthreads)
  for (long c = 0; c < ncell; c++) {
    for (int q = 0; q < no; q++) {
      int i = over[q];
      obb_t oi = s->b[i].obb;
      obb_enlarge(&oi, 0.5 * s->DVerlet);
      for (int r = cstart[c]; r < cstart[c + 1]; r++) {
        int j = cbody[r];
        if (j < s->nDriven && i < s->nDriven) continue;
        if (j <= i) continue;
        if (ifc_index_has(&xi, i, j)) continue;
        obb_t oj = s->b[j].obb;
        obb_enlarge(&oj, 0.5 * s->DVerlet);
        if (obb_intersect(&oi, &oj)) add_or_remove_pair(s, i, j, s->dVerlet);
      }
    }
  }
  free(xi.start);
  free(xi.j);
  free(pa);
  free(cell_of);
  free(cstart);
  free(cbody);
  free(cfill);
  free(over);
}

void rko_update_nl(rko_sim* s) {
  if (s->nl_linkcells)
    update_nl_cells(s);
  else
    update_nl_brute(s);
  s->nl_updates++;
}

/* dynamicCheckUpdateNL (Rockable.cpp:1571-1596) */
static void dynamic_check(rko_sim* s) {
  for (int i = 0; i < s->nb; i++) {
    s->deltaPos[i] = vadd(s->deltaPos[i], vsc(s->dt, s->b[i].vel));
    double tmp1 = vnorm2(s->deltaPos[i]);
    if (tmp1 > s->maxDeltaPos * s->maxDeltaPos) s->maxDeltaPos = sqrt(tmp1);
    qt d = qdot(s->b[i].Q, s->b[i].vrot);
    s->deltaQ[i].w += d.w * s->dt;
    s->deltaQ[i].x += d.x * s->dt;
    s->deltaQ[i].y += d.y * s->dt;
    s->deltaQ[i].z += d.z * s->dt;
    double tmp2 = qangle(s->deltaQ[i]);
    if (tmp2 > s->maxDeltaRot) s->maxDeltaRot = tmp2;
  }
  if (s->maxDeltaPos >= s->dispUpdateNL || s->maxDeltaRot >= s->angleUpdateNL) {
    s->needUpdate = 1;
    for (int i = 0; i < s->nb; i++) {
      s->deltaPos[i] = V(0, 0, 0);
      s->deltaQ[i].w = 1.0;
      s->deltaQ[i].x = s->deltaQ[i].y = s->deltaQ[i].z = 0.0;
    }
    s->maxDeltaPos = 0.0;
    s->maxDeltaRot = 0.0;
  }
}

/* ------------------------------------------------------------------------- */
/* Conf files                                                                 */
/* ------------------------------------------------------------------------- */
static void dirname_of(const char* path, char* out, size_t cap) {
  const char* slash = strrchr(path, '/');
  if (!slash) {
    out[0] = 0;
    return;
  }
  size_t n = (size_t)(slash - path) + 1;
  if (n >= cap) n = cap - 1;
  memcpy(out, path, n);
  out[n] = 0;
}

static int parse_conf(rko_sim* s, const char* fullpath) {
  reader_t r;
  if (rd_open(&r, fullpath) != 0) return seterr(s, "cannot read %s", fullpath);
  char tok[512];
  if (!rd_tok(&r, tok, sizeof tok)) {
    rd_close(&r);
    return seterr(s, "empty conf file %s", fullpath);
  }
  if (strcmp(tok, "Rockable") != 0) {
    if (strcmp(tok, "redirection") == 0) {
      /* Rockable.cpp:1096-1105: "redirection <dir> <conf>" */
      char dir[512], name[512], base[512], full[1536];
      rd_tok(&r, dir, sizeof dir);
      rd_tok(&r, name, sizeof name);
      rd_close(&r);
      dirname_of(fullpath, base, sizeof base);
      snprintf(s->path, sizeof s->path, "%s%s", base, dir);
      snprintf(full, sizeof full, "%s%s", s->path, name);
      return parse_conf(s, full);
    }
  }
  rd_tok(&r, tok, sizeof tok); /* version date */
  while (rd_tok(&r, tok, sizeof tok)) {
    if (tok[0] == '/' || tok[0] == '#' || tok[0] == '!') {
      rd_skipline(&r);
      continue;
    }
    if (!strcmp(tok, "t")) s->t = rd_d(&r);
    else if (!strcmp(tok, "tmax")) s->tmax = rd_d(&r);
    else if (!strcmp(tok, "dt")) {
      s->dt = rd_d(&r);
      s->dt_2 = 0.5 * s->dt;
      s->dt2 = s->dt * s->dt;
      s->dt2_2 = 0.5 * s->dt2;
    } else if (!strcmp(tok, "interVerlet")) s->interVerlet = rd_d(&r);
    else if (!strcmp(tok, "interConf")) s->interConf = rd_d(&r);
    else if (!strcmp(tok, "DVerlet")) s->DVerlet = rd_d(&r);
    else if (!strcmp(tok, "dVerlet")) s->dVerlet = rd_d(&r);
    else if (!strcmp(tok, "parallel_mode")) rd_tok(&r, tok, sizeof tok);
    else if (!strcmp(tok, "density")) {
      int g = (int)rd_l(&r);
      double d = rd_d(&r);
      if (g >= 0 && g < MAXG) {
        s->density[g] = d;
        if (g + 1 > s->ndensity) s->ndensity = g + 1;
      }
    } else if (!strcmp(tok, "gravity")) s->gravity = rd_v3(&r);
    else if (!strcmp(tok, "ParamsInInterfaces")) s->paramsInInterfaces = (int)rd_l(&r);
    else if (!strcmp(tok, "dynamicUpdateNL")) s->dynamicUpdateNL = (int)rd_l(&r);
    else if (!strcmp(tok, "dispUpdateNL")) s->dispUpdateNL = rd_d(&r);
    else if (!strcmp(tok, "angleUpdateNL")) s->angleUpdateNL = rd_d(&r) * (M_PI / 180.0);
    else if (!strcmp(tok, "numericalDampingCoeff")) s->numericalDampingCoeff = rd_d(&r);
    else if (!strcmp(tok, "VelocityBarrier")) s->velocityBarrier = rd_d(&r);
    else if (!strcmp(tok, "AngularVelocityBarrier")) s->angularVelocityBarrier = rd_d(&r);
    else if (!strcmp(tok, "VelocityBarrierExponent")) s->velocityBarrierExponent = rd_d(&r);
    else if (!strcmp(tok, "AngularVelocityBarrierExponent")) s->angularVelocityBarrierExponent = rd_d(&r);
    else if (!strcmp(tok, "preventCrossingLength")) s->preventCrossingLength = rd_d(&r);
    else if (!strcmp(tok, "forceLaw")) {
      rd_tok(&r, tok, sizeof tok);
      /* unknown names fall back to Default (Rockable.cpp:763-781) */
      if (!strcmp(tok, "Avalanche")) {
        s->law = LAW_AVALANCHE;
        strcpy(s->lawName, "Avalanche");
      } else if (!strcmp(tok, "StickedLinks")) {
        s->law = LAW_STICKED;
        strcpy(s->lawName, "StickedLinks");
      } else {
        s->law = LAW_DEFAULT;
        strcpy(s->lawName, "Default");
      }
    } else if (!strcmp(tok, "AddOrRemoveInteractions")) rd_tok(&r, tok, sizeof tok);
    else if (!strcmp(tok, "UpdateNL")) {
      rd_tok(&r, tok, sizeof tok);
      if (!strcmp(tok, "linkCells")) s->nl_linkcells = 1;
      else if (!strcmp(tok, "bruteForce")) s->nl_linkcells = 0;
    } else if (!strcmp(tok, "Integrator")) {
      rd_tok(&r, tok, sizeof tok);
      if (strcmp(tok, "velocityVerlet") != 0) {
        rd_close(&r);
        return seterr(s, "Integrator %s is not supported (velocityVerlet only)", tok);
      }
    } else if (!strcmp(tok, "cellMinSizes")) s->cellMinSizes = rd_v3(&r);
    else if (!strcmp(tok, "boxForLinkCellsOpt")) s->boxForLinkCellsOpt = (int)rd_l(&r);
    else if (!strcmp(tok, "ContactPartnership")) {
      rd_tok(&r, tok, sizeof tok);
      if (strcmp(tok, "None") != 0) {
        rd_close(&r);
        return seterr(s, "ContactPartnership %s is not supported", tok);
      }
    } else if (!strcmp(tok, "iconf")) s->iconf = (int)rd_l(&r);
    else if (!strcmp(tok, "nDriven")) s->nDriven = (int)rd_l(&r);
    else if (!strcmp(tok, "precision")) s->precision = (int)rd_l(&r);
    else if (!strcmp(tok, "shapeFile")) {
      char name[512];
      rd_tok(&r, name, sizeof name);
      char want[1024];
      snprintf(want, sizeof want, "%s%s", s->path, name);
      if (strcmp(want, s->shapeFile) != 0) {
        snprintf(s->shapeFile, sizeof s->shapeFile, "%s", want);
        if (load_shapes(s, s->shapeFile) != 0) {
          rd_close(&r);
          return -1;
        }
      }
    } else if (!strcmp(tok, "Particles")) {
      long nb = rd_l(&r);
      free(s->b);
      s->b = (body_t*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(body_t));
      s->nb = 0;
      for (long i = 0; i < nb; i++) {
        char nm[256];
        if (!rd_tok(&r, nm, sizeof nm)) break;
        if (nm[0] == '#') {
          rd_skipline(&r);
          --i;
          continue;
        }
        if (nm[0] == '!') {
          rd_skipline(&r);
          --i;
          --nb;
          continue;
        }
        body_t P;
        memset(&P, 0, sizeof P);
        P.group = (int)rd_l(&r);
        P.cluster = (int)rd_l(&r);
        P.homothety = rd_d(&r);
        P.pos = rd_v3(&r);
        P.vel = rd_v3(&r);
        P.acc = rd_v3(&r);
        P.Q = rd_qt(&r);
        P.vrot = rd_v3(&r);
        P.arot = rd_v3(&r);
        P.shape = shape_id(s, nm);
        if (P.shape < 0) {
          rd_close(&r);
          return seterr(s, "unknown shape '%s'", nm);
        }
        double h = P.homothety;
        double dens = (P.group >= 0 && P.group < MAXG) ? s->density[P.group] : 0.0;
        P.mass = (h * h * h * s->shapes[P.shape].volume) * dens;
        P.inertia = vsc(P.mass, vsc(h * h, s->shapes[P.shape].inertia_mass));
        s->b[s->nb++] = P;
      }
      for (int k = 0; k < s->nb; k++) {
        /* (re)size the per-owner containers */
      }
      if (s->inter) {
        free(s->inter);
      }
      s->inter = (ilist_t*)calloc((size_t)(s->nb > 0 ? s->nb : 1), sizeof(ilist_t));
      if (s->locks) free(s->locks);
      s->locks = (omp_lock_t*)malloc(sizeof(omp_lock_t) * (size_t)(s->nb > 0 ? s->nb : 1));
      for (int k = 0; k < s->nb; k++) omp_init_lock(&s->locks[k]);
    } else if (!strcmp(tok, "Interactions")) {
      long n = rd_l(&r);
      for (long k = 0; k < n; k++) {
        inter_t I;
        memset(&I, 0, sizeof I);
        I.i = (int)rd_l(&r);
        I.j = (int)rd_l(&r);
        I.type = (int)rd_l(&r);
        I.isub = (int)rd_l(&r);
        I.jsub = (int)rd_l(&r);
        I.n = rd_v3(&r);
        I.dn = rd_d(&r);
        I.pos = rd_v3(&r);
        I.vel = rd_v3(&r);
        I.fn = rd_d(&r);
        I.ft = rd_v3(&r);
        I.mom = rd_v3(&r);
        I.damp = rd_d(&r);
        I.stick = -1;
        if (I.i < 0 || I.i >= s->nb || I.j < 0 || I.j >= s->nb) continue;
        int found;
        int at = il_find(&s->inter[I.i], I.j, I.type, I.isub, I.jsub, &found);
        if (!found) il_insert_at(&s->inter[I.i], at, &I);
      }
    } else if (!strcmp(tok, "Interfaces")) {
      long n = rd_l(&r);
      for (int k = 0; k < s->nifc; k++) free(s->ifc[k].bkey);
      free(s->ifc);
      s->ifc = (iface_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(iface_t));
      s->nifc = 0;
      for (long k = 0; k < n; k++) {
        iface_t BI;
        memset(&BI, 0, sizeof BI);
        BI.power = 2.0;
        BI.isInner = 1;
        BI.i = (int)rd_l(&r);
        BI.j = (int)rd_l(&r);
        BI.nb = (int)rd_l(&r);
        BI.dn0 = rd_d(&r);
        if (s->paramsInInterfaces == 1) {
          BI.kn = rd_d(&r);
          BI.kt = rd_d(&r);
          BI.kr = rd_d(&r);
          BI.fn0 = rd_d(&r);
          BI.ft0 = rd_d(&r);
          BI.mom0 = rd_d(&r);
          BI.power = rd_d(&r);
          BI.Gc = rd_d(&r);
        }
        BI.bkey = (int*)malloc(sizeof(int) * 3 * (size_t)(BI.nb > 0 ? BI.nb : 1));
        int missed = 0;
        for (int u = 0; u < BI.nb; u++) {
          BI.bkey[3 * u] = (int)rd_l(&r);
          BI.bkey[3 * u + 1] = (int)rd_l(&r);
          BI.bkey[3 * u + 2] = (int)rd_l(&r);
          if (!find_inter(s, BI.i, BI.j, BI.bkey[3 * u], BI.bkey[3 * u + 1], BI.bkey[3 * u + 2])) missed = 1;
        }
        /* duplicates of (i, j) are ignored, as in std::set<BreakableInterface> */
        int dup = 0;
        for (int q = 0; q < s->nifc; q++)
          if (s->ifc[q].i == BI.i && s->ifc[q].j == BI.j) dup = 1;
        if (missed || dup) {
          free(BI.bkey);
          continue;
        }
        BI.alive = 1;
        int id = s->nifc++;
        s->ifc[id] = BI;
        for (int u = 0; u < BI.nb; u++) {
          inter_t* I = find_inter(s, BI.i, BI.j, BI.bkey[3 * u], BI.bkey[3 * u + 1], BI.bkey[3 * u + 2]);
          I->stick = id;
        }
      }
      /* keep interfaces sorted by (i, j): the order breakage visits them */
      for (int a = 1; a < s->nifc; a++) {
        iface_t x = s->ifc[a];
        int q = a - 1;
        while (q >= 0 && (s->ifc[q].i > x.i || (s->ifc[q].i == x.i && s->ifc[q].j > x.j))) {
          s->ifc[q + 1] = s->ifc[q];
          q--;
        }
        s->ifc[q + 1] = x;
      }
      for (int k = 0; k < s->nb; k++)
        for (int q = 0; q < s->inter[k].n; q++) s->inter[k].a[q].stick = -1;
      for (int id = 0; id < s->nifc; id++) {
        iface_t* BI = &s->ifc[id];
        for (int u = 0; u < BI->nb; u++) {
          inter_t* I = find_inter(s, BI->i, BI->j, BI->bkey[3 * u], BI->bkey[3 * u + 1], BI->bkey[3 * u + 2]);
          if (I) I->stick = id;
        }
      }
      free(s->ifc_break);
      s->ifc_break = (unsigned char*)calloc((size_t)(s->nifc > 0 ? s->nifc : 1), 1);
    } else if (!strcmp(tok, "BodyForce")) {
      rd_tok(&r, tok, sizeof tok);
      if (!strcmp(tok, "AttractingPoint")) {
        s->bf = BF_ATTRACTING;
        s->bf_point = rd_v3(&r);
        s->bf_acc = rd_d(&r);
      } else if (!strcmp(tok, "PreferredDirection")) {
        s->bf = BF_PREFDIR;
        s->bf_axisBody = rd_v3(&r);
        s->bf_axis = rd_v3(&r);
        s->bf_momentMax = rd_d(&r);
        vnormalize(&s->bf_axisBody);
        vnormalize(&s->bf_axis);
        s->bf_Kr = 2.0 * s->bf_momentMax / M_PI;
      } else if (!strcmp(tok, "ViscousFluid")) {
        s->bf = BF_VISCOUS;
        s->bf_fluidDensity = rd_d(&r);
      }
    } else {
      int id = -1;
      for (int p = 0; p < P_COUNT; p++)
        if (!strcmp(tok, PNAME[p])) id = p;
      if (id >= 0) {
        int g1 = (int)rd_l(&r), g2 = (int)rd_l(&r);
        double v = rd_d(&r);
        dt_set(s, id, g1, g2, v);
      } else if (s->verbose) {
        fprintf(stderr, "[oracle] unknown keyword '%s'\n", tok);
      }
    }
  }
  rd_close(&r);
  if (!s->ifc_break) s->ifc_break = (unsigned char*)calloc(1, 1);
  return 0;
}

int rko_load_conf(rko_sim* s, const char* path, int run_fake_step) {
  s->path[0] = 0;
  dirname_of(path, s->path, sizeof s->path);
  if (parse_conf(s, path) != 0) return -1;
  if (s->numericalDampingCoeff > 0.0 || s->velocityBarrier > 0.0 || s->angularVelocityBarrier > 0.0)
    return seterr(s, "numerical damping and velocity barriers are not supported");
  if (run_fake_step) rko_accelerations(s);
  return 0;
}

int rko_read_driving(rko_sim* s, const char* path) { /* DrivingSystem.cpp:44-496 */
  reader_t r;
  /* read again at each conf save (Rockable.cpp:2682-2688), even if absent now */
  if (path != s->driving_path) snprintf(s->driving_path, sizeof s->driving_path, "%s", path);
  if (rd_open(&r, s->driving_path) != 0) return 0; /* no file: the first nDriven bodies do not move */
  free(s->ctrl);
  s->ctrl = NULL;
  s->nctrl = 0;
  int cap = 0;
  char tok[256];
#define PUSHC(T, I, VAL)                                                   \
  do {                                                                     \
    if (s->nctrl == cap) {                                                 \
      cap = cap ? 2 * cap : 16;                                            \
      s->ctrl = (control_t*)realloc(s->ctrl, sizeof(control_t) * (size_t)cap); \
    }                                                                      \
    control_t c_;                                                          \
    memset(&c_, 0, sizeof c_);                                             \
    c_.type = (T);                                                         \
    c_.i = (I);                                                            \
    c_.value = (VAL);                                                      \
    s->ctrl[s->nctrl++] = c_;                                              \
  } while (0)
  static const char* names[] = {"_x_Vel_",  "_y_Vel_",  "_z_Vel_",    "_xrot_Vel_", "_yrot_Vel_",
                                "_zrot_Vel_", "_x_For_", "_y_For_",    "_z_For_",    "_xrot_Mom_",
                                "_yrot_Mom_", "_zrot_Mom_", "_xyzrot_Vel_", "_xyzrot_Mom_"};
  static const int codes[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 100, 101};
  while (rd_tok(&r, tok, sizeof tok)) {
    if (tok[0] == '/' || tok[0] == '#' || tok[0] == '!') {
      rd_skipline(&r);
      continue;
    }
    if (!strcmp(tok, "Control")) {
      char ty[64];
      rd_tok(&r, ty, sizeof ty);
      int code = 0;
      for (int k = 0; k < 14; k++)
        if (!strcmp(ty, names[k])) code = codes[k];
      int i = (int)rd_l(&r);
      if (code >= 100) {
        PUSHC(code, i, 0.0);
        s->ctrl[s->nctrl - 1].vec_value = rd_v3(&r);
      } else {
        double v = rd_d(&r);
        PUSHC(code, i, v);
      }
    } else if (!strcmp(tok, "Servo")) {
      char nm[128];
      rd_tok(&r, nm, sizeof nm);
      if (strncmp(nm, "tritri", 6) == 0) { /* DrivingSystem.cpp:191-392 */
        int id[6];
        for (int k = 0; k < 6; k++) id[k] = (int)rd_l(&r);
        int icontr = s->nctrl;
        PUSHC(CTRL_X_VEL, id[0], 0.0);
        PUSHC(CTRL_X_FOR, id[1], 0.0);
        PUSHC(CTRL_Y_VEL, id[2], 0.0);
        PUSHC(CTRL_Y_FOR, id[3], 0.0);
        PUSHC(CTRL_Z_VEL, id[4], 0.0);
        PUSHC(CTRL_Z_FOR, id[5], 0.0);
        int kind = 0;
        if (!strcmp(nm, "tritriIsostaticCompression")) {
          kind = 1;
          s->servo_pressure = rd_d(&r);
        } else if (!strcmp(nm, "tritriBiaxialCompression")) {
          kind = 2;
          s->servo_pressure = rd_d(&r);
          s->servo_velocity = rd_d(&r);
          s->ctrl[icontr + 3].type = CTRL_Y_VEL;
          s->ctrl[icontr + 3].value = -s->servo_velocity;
        } else if (!strcmp(nm, "tritriCustom")) {
          kind = 3;
          static const int vel[3] = {CTRL_X_VEL, CTRL_Y_VEL, CTRL_Z_VEL};
          static const int frc[3] = {CTRL_X_FOR, CTRL_Y_FOR, CTRL_Z_FOR};
          for (int k = 0; k < 6; k++) {
            int ty = (int)rd_l(&r);
            s->servo_val[k] = rd_d(&r);
            s->ctrl[icontr + k].type = (ty == 0) ? vel[k / 2] : frc[k / 2];
            s->ctrl[icontr + k].value = (k & 1) ? -s->servo_val[k] : s->servo_val[k];
          }
        } else if (!strcmp(nm, "tritriLodeAngle")) {
          kind = 4;
          s->servo_pressure = rd_d(&r);
          double lode = rd_d(&r);
          s->lode_rate = rd_d(&r);
          double rad = lode * M_PI / 180.0;
          s->lode_a = (3.0 * tan(rad) - sqrt(3.0)) / (2.0 * sqrt(3.0));
          s->lode_b = -(1.0 + s->lode_a);
        }
        if (kind) { /* an unknown tritri name keeps its controls and no new function */
          s->servo = kind;
          s->servo_icontr = icontr;
          memcpy(s->servo_ids, id, sizeof id);
        }
      } else if (!strcmp(nm, "shaker") || !strcmp(nm, "triangle_shaker") || !strcmp(nm, "sawtooth_shaker")) {
        /* 397-477: velocity of one body along dir */
        int i = (int)rd_l(&r);
        v3 dir = rd_v3(&r);
        double A = rd_d(&r), freq = rd_d(&r);
        double tini = 0.0;
        if (!strcmp(nm, "sawtooth_shaker")) tini = rd_d(&r);
        vnormalize(&dir);
        s->shake_dir = dir;
        s->shake_A = A;
        s->shake_omega = 2.0 * M_PI * freq;
        s->shake_T = 1.0 / freq;
        s->shake_tini = tini;
        s->servo_icontr = s->nctrl;
        s->servo = !strcmp(nm, "shaker") ? 5 : !strcmp(nm, "triangle_shaker") ? 6 : 7;
        PUSHC(CTRL_X_VEL, i, 0.0);
        PUSHC(CTRL_Y_VEL, i, 0.0);
        PUSHC(CTRL_Z_VEL, i, 0.0);
      } else if (!strcmp(nm, "ramp")) { /* 479-494 */
        char ty[64];
        rd_tok(&r, ty, sizeof ty);
        int code = 0;
        for (int k = 0; k < 14; k++)
          if (!strcmp(ty, names[k])) code = codes[k];
        int i = (int)rd_l(&r);
        s->ramp_vb = rd_d(&r);
        s->ramp_ve = rd_d(&r);
        s->ramp_tb = rd_d(&r);
        s->ramp_te = rd_d(&r);
        s->servo_icontr = s->nctrl;
        s->servo = 8;
        PUSHC(code, i, s->ramp_vb);
      }
    }
  }
#undef PUSHC
  rd_close(&r);
  return 0;
}

/* comparison for the lexicographic interaction order (Interaction.hpp:104-122) */
static int cmp_inter_ptr(const void* a, const void* b) {
  const inter_t* x = *(inter_t* const*)a;
  const inter_t* y = *(inter_t* const*)b;
  if (x->i != y->i) return x->i < y->i ? -1 : 1;
  if (x->j != y->j) return x->j < y->j ? -1 : 1;
  if (x->type != y->type) return x->type < y->type ? -1 : 1;
  if (x->isub != y->isub) return x->isub < y->isub ? -1 : 1;
  if (x->jsub != y->jsub) return x->jsub < y->jsub ? -1 : 1;
  return 0;
}

static void fput_e(FILE* f, int prec, double x) { fprintf(f, "%.*e", prec, x); }
static void fput_v3e(FILE* f, int prec, v3 v) {
  fput_e(f, prec, v.x);
  fputc(' ', f);
  fput_e(f, prec, v.y);
  fputc(' ', f);
  fput_e(f, prec, v.z);
}

/* saveConf (Rockable.cpp:440-610) */
int rko_save_conf(rko_sim* s, const char* path) {
  active_from_flags(s); /* the listed entries that were active at the last accelerations() */
  FILE* f = fopen(path, "w");
  if (!f) return seterr(s, "cannot write %s", path);
  fprintf(f, "Rockable 21-08-2022\n");
  fprintf(f, "AddOrRemoveInteractions bruteForce\n");
  fprintf(f, "Integrator velocityVerlet\n");
  fprintf(f, "UpdateNL %s\n", s->nl_linkcells ? "linkCells" : "bruteForce");
  fprintf(f, "forceLaw %s\n", s->lawName);
  fprintf(f, "t %g\n", s->t);
  fprintf(f, "tmax %g\n", s->tmax);
  fprintf(f, "dt %g\n", s->dt);
  fprintf(f, "interVerlet %g\n", s->interVerlet);
  fprintf(f, "interConf %g\n", s->interConf);
  fprintf(f, "DVerlet %g\n", s->DVerlet);
  fprintf(f, "dVerlet %g\n", s->dVerlet);
  for (int g = 0; g < s->ndensity; g++)
    if (s->density[g] > 0.0) fprintf(f, "density %d %g\n", g, s->density[g]);
  fprintf(f, "gravity %g %g %g\n", s->gravity.x, s->gravity.y, s->gravity.z);
  fprintf(f, "ParamsInInterfaces %d\n", s->paramsInInterfaces);
  fprintf(f, "dynamicUpdateNL %d\n", s->dynamicUpdateNL);
  if (s->dynamicUpdateNL != 0) {
    fprintf(f, "dispUpdateNL %g\n", s->dispUpdateNL);
    fprintf(f, "angleUpdateNL %g\n", s->angleUpdateNL * (180.0 / M_PI));
  }
  fprintf(f, "numericalDampingCoeff %g\n", s->numericalDampingCoeff);
  fprintf(f, "VelocityBarrier %g\n", s->velocityBarrier);
  fprintf(f, "AngularVelocityBarrier %g\n", s->angularVelocityBarrier);
  fprintf(f, "VelocityBarrierExponent %g\n", s->velocityBarrierExponent);
  fprintf(f, "AngularVelocityBarrierExponent %g\n", s->angularVelocityBarrierExponent);
  for (int q = 0; q < P_COUNT; q++) {
    int id = PWRITE[q];
    for (int g1 = 0; g1 < s->ngroup; g1++)
      for (int g2 = g1; g2 < s->ngroup; g2++)
        if (s->defined[id][g1][g2]) fprintf(f, "%s %d %d %g\n", PNAME[id], g1, g2, s->table[id][g1][g2]);
  }
  fprintf(f, "ContactPartnership None\n");
  fprintf(f, "cellMinSizes %g %g %g\n", s->cellMinSizes.x, s->cellMinSizes.y, s->cellMinSizes.z);
  fprintf(f, "boxForLinkCellsOpt %d\n", s->boxForLinkCellsOpt);
  fprintf(f, "iconf %d\n", s->iconf);
  fprintf(f, "nDriven %d\n", s->nDriven);
  {
    const char* sf = strrchr(s->shapeFile, '/');
    fprintf(f, "shapeFile %s\n", sf ? sf + 1 : s->shapeFile);
  }
  int prec = s->precision;
  fprintf(f, "precision %d\n", prec);
  fprintf(f, "Particles %d\n", s->nb);
  fprintf(f,
          "#name group cluster homothety pos.x pos.y pos.z vel.x vel.y vel.z acc.x acc.y acc.z Q.w Q.x Q.y Q.z "
          "vrot.x vrot.y vrot.z arot.x arot.y arot.z\n");
  for (int i = 0; i < s->nb; i++) {
    body_t* p = &s->b[i];
    fprintf(f, "%s %d %d ", s->shapes[p->shape].name, p->group, p->cluster);
    fput_e(f, prec, p->homothety);
    fputc(' ', f);
    fput_v3e(f, prec, p->pos);
    fputc(' ', f);
    fput_v3e(f, prec, p->vel);
    fputc(' ', f);
    fput_v3e(f, prec, p->acc);
    fputc(' ', f);
    fput_e(f, prec, p->Q.w);
    fputc(' ', f);
    fput_e(f, prec, p->Q.x);
    fputc(' ', f);
    fput_e(f, prec, p->Q.y);
    fputc(' ', f);
    fput_e(f, prec, p->Q.z);
    fputc(' ', f);
    fput_v3e(f, prec, p->vrot);
    fputc(' ', f);
    fput_v3e(f, prec, p->arot);
    fputc('\n', f);
  }
  inter_t** act = (inter_t**)malloc(sizeof(inter_t*) * (s->nactive > 0 ? s->nactive : 1));
  memcpy(act, s->active, sizeof(inter_t*) * s->nactive);
  qsort(act, s->nactive, sizeof(inter_t*), cmp_inter_ptr);
  fprintf(f, "Interactions %zu\n", s->nactive);
  int prev = fegetround();
  for (size_t q = 0; q < s->nactive; q++) {
    inter_t* I = act[q];
    fprintf(f, "%d %d %d %d %d ", I->i, I->j, I->type, I->isub, I->jsub);
    fflush(f);
    fesetround(FE_TOWARDZERO);
    fput_v3e(f, prec, I->n);
    fflush(f);
    fesetround(prev);
    fputc(' ', f);
    fput_e(f, prec, I->dn);
    fputc(' ', f);
    fput_v3e(f, prec, I->pos);
    fputc(' ', f);
    fput_v3e(f, prec, I->vel);
    fputc(' ', f);
    fput_e(f, prec, I->fn);
    fputc(' ', f);
    fput_v3e(f, prec, I->ft);
    fputc(' ', f);
    fput_v3e(f, prec, I->mom);
    fputc(' ', f);
    fput_e(f, prec, I->damp);
    fputc('\n', f);
  }
  free(act);
  int nalive = 0;
  for (int k = 0; k < s->nifc; k++) nalive += s->ifc[k].alive;
  fprintf(f, "Interfaces %d\n", nalive);
  for (int k = 0; k < s->nifc; k++) {
    iface_t* BI = &s->ifc[k];
    if (!BI->alive) continue;
    fprintf(f, "%d %d  %d %.6f", BI->i, BI->j, BI->nb, BI->dn0);
    if (s->paramsInInterfaces == 1)
      fprintf(f, " %.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f ", BI->kn, BI->kt, BI->kr, BI->fn0, BI->ft0, BI->mom0,
              BI->power, BI->Gc);
    for (int u = 0; u < BI->nb; u++) fprintf(f, "  %d %d %d", BI->bkey[3 * u], BI->bkey[3 * u + 1], BI->bkey[3 * u + 2]);
    fputc('\n', f);
  }
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* integrate() (Rockable.cpp:2538-2746)                                       */
/* ------------------------------------------------------------------------- */
/* reset: the start of integrate(), which zeroes the dynamicCheckUpdateNL
   accumulations; a continued run (rko_run_steps) keeps them. */
static void integrate_prepare(rko_sim* s, int reset) {
  if (s->dynamicUpdateNL != 0 && (reset || !s->deltaPos)) {
    free(s->deltaPos);
    free(s->deltaQ);
    s->deltaPos = (v3*)calloc((size_t)(s->nb > 0 ? s->nb : 1), sizeof(v3));
    s->deltaQ = (qt*)calloc((size_t)(s->nb > 0 ? s->nb : 1), sizeof(qt));
    for (int i = 0; i < s->nb; i++) s->deltaQ[i].w = 1.0;
    s->maxDeltaPos = 0.0;
    s->maxDeltaRot = 0.0;
  }
  s->dt_2 = 0.5 * s->dt;
  s->dt2_2 = s->dt_2 * s->dt;
  s->dt2 = s->dt * s->dt;
}

/* getKineticEnergy over every body (Rockable.cpp:3958-3970, first = 0) */
static void kinetic_energy(const rko_sim* s, double* Et, double* Er) {
  double et = 0.0, er = 0.0;
  for (int i = 0; i < s->nb; i++) {
    const body_t* p = &s->b[i];
    et += p->mass * vdot(p->vel, p->vel);
    v3 w = qrot(qconj(p->Q), p->vrot);
    er += (p->inertia.x * w.x * w.x + p->inertia.y * w.y * w.y + p->inertia.z * w.z * w.z);
  }
  *Et = et * 0.5;
  *Er = er * 0.5;
}

/* getResultantQuickStats from nDriven (Rockable.cpp:4145-4206) and
   getInteractionQuickStats (4217-4252), on the active interactions in their
   serial order */
static void quick_stats(rko_sim* s, double* Fmax, double* F_fnmax, double* fnMax_out, double* fnMean_out) {
  active_from_flags(s);
  double* fnMax = (double*)calloc((size_t)s->nb + 1, sizeof(double));
  size_t* nbCtc = (size_t*)calloc((size_t)s->nb + 1, sizeof(size_t));
  for (size_t k = 0; k < s->nactive; k++) {
    const inter_t* I = s->active[k];
    nbCtc[I->i] += 1;
    nbCtc[I->j] += 1;
    if (fnMax[I->i] < I->fn) fnMax[I->i] = I->fn;
    if (fnMax[I->j] < I->fn) fnMax[I->j] = I->fn;
  }
  const int first = s->nDriven, last = s->nb - 1;
  double F = first <= last ? sqrt(vnorm2(s->b[first].force)) : 0.0;
  *Fmax = F;
  *F_fnmax = 0.0;
  for (int i = first; i <= last; i++) {
    if (nbCtc[i] <= 1) continue;
    F = sqrt(vnorm2(s->b[i].force));
    if (F > *Fmax) *Fmax = F;
    if (fnMax[i] > 0.0) {
      F /= fnMax[i];
      if (F > *F_fnmax) *F_fnmax = F;
    }
  }
  free(fnMax);
  free(nbCtc);
  if (s->nactive == 0) {
    *fnMax_out = *fnMean_out = 0.0;
    return;
  }
  double fn = s->active[0]->fn, mx = fn, mean = 0.0;
  size_t n = 1;
  if (g_var[VAR_MUTANT] == 18) { /* a planted mistake: the mean over every fn */
    mean = fn;
    n = 2;
  }
  for (size_t k = 1; k < s->nactive; k++) {
    n += 1;
    fn = s->active[k]->fn;
    mean += fn;
    if (fn > mx) mx = fn;
  }
  if (n > 1) mean /= (double)(n - 1);
  *fnMax_out = mx;
  *fnMean_out = mean;
}

/* The records integrate() writes at each conf save (Rockable.cpp:2600-2668):
   perf.txt (times: this restatement's own wall clock), staticBalance.txt,
   kineticEnergy.txt; numbers as an std::ostream prints them (%g). */
static void conf_outputs(rko_sim* s, double elapsed) {
  if (s->perf_file) {
    fprintf(s->perf_file, "%g %g %g %g\n", s->t, elapsed > 0.0 ? s->interConf / elapsed : 0.0, 0.0, 0.0);
    fflush(s->perf_file);
  }
  double Fmax, F_fnmax, fnMax, fnMean;
  quick_stats(s, &Fmax, &F_fnmax, &fnMax, &fnMean);
  if (s->sb_file && fnMax != 0.0 && fnMean != 0.0) {
    fprintf(s->sb_file, "%g %g %g\n", s->t, F_fnmax, fabs(Fmax / fnMean));
    fflush(s->sb_file);
  }
  if (s->ke_file) {
    double Et, Er;
    kinetic_energy(s, &Et, &Er);
    fprintf(s->ke_file, "%g %g %g\n", s->t, Et, Er);
    fflush(s->ke_file);
  }
}

/* DataExtractor::record of dnStat (DataExtractor_dnStat.cpp:69-105) and
   TrackDamage (DataExtractor_TrackDamage.cpp:47-50) */
static void extractors_record(rko_sim* s, int64_t step) {
  for (int q = 0; q < s->nextr; q++) {
    if (s->extr[q].nrec <= 0 || step % s->extr[q].nrec != 0) continue;
    FILE* f = s->extr[q].f;
    if (s->extr[q].kind == 1) {
      active_from_flags(s);
      double dnMin = 1e20, dnMax = -1e20, dnMeanNeg = 0.0, dnMeanPos = 0.0, nbNeg = 0.0, nbPos = 0.0, dnMean = 0.0;
      v3 posMin = V(0.0, 0.0, 0.0);
      for (size_t k = 0; k < s->nactive; k++) {
        const inter_t* I = s->active[k];
        const double dn = I->dn;
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
          posMin = I->pos;
        }
        if (dn > dnMax) dnMax = dn;
      }
      if (nbNeg > 0.0) dnMeanNeg /= nbNeg;
      if (nbPos > 0.0) dnMeanPos /= nbPos;
      if (nbNeg > 0.0 || nbPos > 0.0) dnMean /= (nbNeg + nbPos);
      fprintf(f, "%g %g %g %g %g %g %d %d %g %g %g\n", s->t, dnMin, dnMax, dnMean, dnMeanNeg, dnMeanPos, (int)nbNeg,
              (int)nbPos, posMin.x, posMin.y, posMin.z);
    } else {
      fprintf(f, "%g %g\n", s->t, s->totalBrokenArea / s->inAreaInterfaces);
    }
    fflush(f);
  }
}

static void close_outputs(rko_sim* s) {
  FILE** fs[] = {&s->perf_file, &s->ke_file, &s->sb_file};
  for (int k = 0; k < 3; k++)
    if (*fs[k]) {
      fclose(*fs[k]);
      *fs[k] = NULL;
    }
}

/* readDataExtractors (Rockable.cpp:1152-1180): names and their arguments,
   output files relative to the directory of the file; an unknown name is
   skipped token by token, as the reference's stream reading skips it.
   TrackDamage::init sums the live interfaces' areas at this point. */
int rko_load_data_extractors(rko_sim* s, const char* path) {
  reader_t r;
  if (rd_open(&r, path) != 0) return 0;
  char dir[1024] = {0};
  dirname_of(path, dir, sizeof dir);
  char tok[1024];
  while (rd_tok(&r, tok, sizeof tok)) {
    if (tok[0] == '/' || tok[0] == '#' || tok[0] == '!') {
      rd_skipline(&r);
      continue;
    }
    int kind = !strcmp(tok, "dnStat") ? 1 : !strcmp(tok, "TrackDamage") ? 2 : 0;
    if (!kind || s->nextr >= 16) continue;
    char name[1024], full[2300];
    rd_tok(&r, name, sizeof name);
    int nrec = (int)rd_l(&r);
    if (name[0] == '/')
      snprintf(full, sizeof full, "%s", name);
    else
      snprintf(full, sizeof full, "%s%s", dir, name);
    FILE* f = fopen(full, "w");
    if (!f) {
      rd_close(&r);
      return seterr(s, "cannot write %s", full);
    }
    s->extr[s->nextr].kind = kind;
    s->extr[s->nextr].nrec = nrec;
    s->extr[s->nextr].f = f;
    s->nextr++;
    if (kind == 2) {
      s->inAreaInterfaces = 0.0;
      for (int k = 0; k < s->nifc; k++)
        if (s->ifc[k].alive) s->inAreaInterfaces += s->ifc[k].area;
    }
  }
  rd_close(&r);
  return 0;
}

static int loop_body(rko_sim* s, const char* outdir) {
  s->t += s->dt;
  s->interConfC += s->dt;
  s->interVerletC += s->dt;
  rko_vv_step(s);
  if (s->interConfC >= s->interConf - s->dt_2) {
    s->iconf++;
    if (outdir) {
      char fn[1024];
      snprintf(fn, sizeof fn, "%s/conf%d", outdir, s->iconf);
      conf_outputs(s, 0.0);
      if (rko_save_conf(s, fn) != 0) return -1;
    }
    /* re-read the driving file, then the servo on the new controls
       (Rockable.cpp:2682-2688) */
    if (s->driving_path[0] && rko_read_driving(s, s->driving_path) != 0) return -1;
    servo(s);
    s->interConfC = 0.0;
  }
  if (s->needUpdate || s->interVerletC >= s->interVerlet - s->dt_2) {
    rko_update_nl(s);
    if (!s->needUpdate)
      s->interVerletC = 0.0;
    else
      s->needUpdate = 0;
  }
  if (outdir) extractors_record(s, s->ext_step);
  if (s->dynamicUpdateNL != 0) dynamic_check(s);
  s->ext_step++;
  return 0;
}

int64_t rko_integrate(rko_sim* s, const char* outdir, int64_t max_steps) {
  integrate_prepare(s, 1);
  if (outdir) {
    char fn[1024];
    snprintf(fn, sizeof fn, "%s/conf%d", outdir, s->iconf);
    if (rko_save_conf(s, fn) != 0) return -1;
    const char* names[3] = {"perf.txt", "kineticEnergy.txt", "staticBalance.txt"};
    FILE** fs[3] = {&s->perf_file, &s->ke_file, &s->sb_file};
    close_outputs(s);
    for (int k = 0; k < 3; k++) {
      snprintf(fn, sizeof fn, "%s/%s", outdir, names[k]);
      *fs[k] = fopen(fn, "w");
    }
  }
  s->ext_step = (int64_t)(s->t / s->dt); /* Rockable.cpp:2578 */
  int64_t n = 0;
  while (s->t <= s->tmax && (max_steps < 0 || n < max_steps)) {
    if (loop_body(s, outdir) != 0) return -1;
    n++;
  }
  return n;
}

int64_t rko_run_steps(rko_sim* s, int64_t nsteps) {
  integrate_prepare(s, 0);
  for (int64_t n = 0; n < nsteps; n++)
    if (loop_body(s, NULL) != 0) return -1;
  return nsteps;
}

/* ------------------------------------------------------------------------- */
/* Life cycle and accessors                                                   */
/* ------------------------------------------------------------------------- */
rko_sim* rko_create(void) {
  rko_sim* s = (rko_sim*)calloc(1, sizeof(rko_sim));
  s->nthreads = 1;
  s->tmax = 1.0;
  s->dt = 1e-6;
  s->dt_2 = 0.5e-6;
  s->dt2 = 1e-12;
  s->dt2_2 = 0.5e-12;
  s->interVerlet = 0.01;
  s->interConf = 0.25;
  s->dispUpdateNL = 1.0;
  s->angleUpdateNL = 1.0;
  s->velocityBarrierExponent = 1.0;
  s->angularVelocityBarrierExponent = 1.0;
  s->gravity = V(0.0, -9.81, 0.0);
  s->cellMinSizes = V(1.0, 1.0, 1.0);
  s->precision = 16;
  s->law = LAW_DEFAULT;
  strcpy(s->lawName, "Default");
  strcpy(s->shapeFile, "noShapeFile");
  s->maxDeltaPos = 1.0;
  s->maxDeltaRot = 1.0;
  return s;
}

void rko_destroy(rko_sim* s) {
  if (s) {
    close_outputs(s);
    for (int q = 0; q < s->nextr; q++)
      if (s->extr[q].f) fclose(s->extr[q].f);
  }
  if (!s) return;
  for (int k = 0; k < s->nshapes; k++) shape_free(&s->shapes[k]);
  free(s->shapes);
  for (int k = 0; k < s->nb; k++) free(s->inter ? s->inter[k].a : NULL);
  if (s->locks)
    for (int k = 0; k < s->nb; k++) omp_destroy_lock(&s->locks[k]);
  free(s->locks);
  free(s->inter);
  free(s->b);
  free(s->active);
  for (int k = 0; k < s->nifc; k++) free(s->ifc[k].bkey);
  free(s->ifc);
  free(s->ifc_break);
  free(s->ctrl);
  free(s->deltaPos);
  free(s->deltaQ);
  free(s);
}

const char* rko_last_error(rko_sim* s) { return s->err; }
void rko_set_threads(rko_sim* s, int n) { s->nthreads = n > 0 ? n : 1; }
void rko_set_verbose(rko_sim* s, int v) { s->verbose = v; }
size_t rko_nbodies(rko_sim* s) { return (size_t)s->nb; }
size_t rko_ndriven(rko_sim* s) { return (size_t)s->nDriven; }
size_t rko_nlisted(rko_sim* s) {
  size_t n = 0;
  for (int k = 0; k < s->nb; k++) n += (size_t)s->inter[k].n;
  return n;
}
size_t rko_nactive(rko_sim* s) {
  active_from_flags(s);
  return s->nactive;
}
size_t rko_ninterfaces(rko_sim* s) {
  size_t n = 0;
  for (int k = 0; k < s->nifc; k++) n += (size_t)s->ifc[k].alive;
  return n;
}
double rko_time(rko_sim* s) { return s->t; }
int64_t rko_nl_updates(rko_sim* s) { return s->nl_updates; }
double rko_total_broken_area(rko_sim* s) { return s->totalBrokenArea; }

void rko_get_bodies(rko_sim* s, double* pos, double* vel, double* acc, double* Q, double* vrot, double* arot,
                    double* force, double* moment) {
  for (int i = 0; i < s->nb; i++) {
    body_t* p = &s->b[i];
#define PUT3(A, F)          \
  if (A) {                  \
    (A)[3 * i] = p->F.x;     \
    (A)[3 * i + 1] = p->F.y; \
    (A)[3 * i + 2] = p->F.z; \
  }
    PUT3(pos, pos);
    PUT3(vel, vel);
    PUT3(acc, acc);
    PUT3(vrot, vrot);
    PUT3(arot, arot);
    PUT3(force, force);
    PUT3(moment, moment);
#undef PUT3
    if (Q) {
      Q[4 * i] = p->Q.w;
      Q[4 * i + 1] = p->Q.x;
      Q[4 * i + 2] = p->Q.y;
      Q[4 * i + 3] = p->Q.z;
    }
  }
}

void rko_get_mass(rko_sim* s, double* mass, double* inertia) {
  for (int i = 0; i < s->nb; i++) {
    if (mass) mass[i] = s->b[i].mass;
    if (inertia) {
      inertia[3 * i] = s->b[i].inertia.x;
      inertia[3 * i + 1] = s->b[i].inertia.y;
      inertia[3 * i + 2] = s->b[i].inertia.z;
    }
  }
}

void rko_set_bodies(rko_sim* s, const double* pos, const double* vel, const double* Q, const double* vrot) {
  for (int i = 0; i < s->nb; i++) {
    body_t* p = &s->b[i];
    if (pos) p->pos = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    if (vel) p->vel = V(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    if (vrot) p->vrot = V(vrot[3 * i], vrot[3 * i + 1], vrot[3 * i + 2]);
    if (Q) {
      p->Q.w = Q[4 * i];
      p->Q.x = Q[4 * i + 1];
      p->Q.y = Q[4 * i + 2];
      p->Q.z = Q[4 * i + 3];
    }
  }
}

static void to_record(const inter_t* I, rko_interaction* o) {
  o->i = I->i;
  o->j = I->j;
  o->type = I->type;
  o->isub = I->isub;
  o->jsub = I->jsub;
  o->stick = I->stick;
  o->n[0] = I->n.x;
  o->n[1] = I->n.y;
  o->n[2] = I->n.z;
  o->dn = I->dn;
  o->pos[0] = I->pos.x;
  o->pos[1] = I->pos.y;
  o->pos[2] = I->pos.z;
  o->vel[0] = I->vel.x;
  o->vel[1] = I->vel.y;
  o->vel[2] = I->vel.z;
  o->fn = I->fn;
  o->ft[0] = I->ft.x;
  o->ft[1] = I->ft.y;
  o->ft[2] = I->ft.z;
  o->mom[0] = I->mom.x;
  o->mom[1] = I->mom.y;
  o->mom[2] = I->mom.z;
  o->damp = I->damp;
}

size_t rko_get_interactions(rko_sim* s, int active, rko_interaction* out, size_t cap) {
  size_t n = 0;
  for (int k = 0; k < s->nb; k++) {
    ilist_t* L = &s->inter[k];
    for (int q = 0; q < L->n; q++) {
      const inter_t* I = &L->a[q];
      if (active && !I->act) continue;
      if (out && n < cap) to_record(I, &out[n]);
      n++;
    }
  }
  return n;
}
