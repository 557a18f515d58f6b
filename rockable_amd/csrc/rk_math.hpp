// rk_math.hpp -- FP64 vector, quaternion and OBB arithmetic of the DEM step,
// usable from host and device code.
//
// Every expression keeps the operand order of the reference so that results
// are bit-identical to Rockable's CPU path (the whole library is compiled with
// -ffp-contract=off: no fused multiply-add is formed implicitly).  The forms of
// the primitives Rockable takes from its external toofus headers were pinned
// against the reference's regression goldens:
//   * quat * vec3 rotation:  t = 2 (u x v);  v + w t + u x t
//   * vec3::normalize():     multiply by the reciprocal of the length
//   * vec3 / double:         multiply by the reciprocal
//   * quat::dot(w):          1/2 (-(u.w), s w + w x u)
//   * quat::normalize():     n = sqrt(s^2 + u.u), multiply by 1/n
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define RK_HD __host__ __device__ __forceinline__

namespace rk {

struct v3 {
  double x, y, z;
};
struct qt {
  double w, x, y, z;
};

RK_HD v3 V(double x, double y, double z) {
  v3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
RK_HD v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
RK_HD v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
RK_HD v3 vsc(double s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
RK_HD v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
RK_HD double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RK_HD double vnorm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RK_HD v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RK_HD v3 vdiv(v3 a, double s) {
  double r = 1.0 / s;
  return V(a.x * r, a.y * r, a.z * r);
}
// vec3::normalize(): returns the former length (Interaction.cpp:157).
RK_HD double vnormalize(v3* a) {
  double n = sqrt(vnorm2(*a));
  double r = 1.0 / n;
  a->x *= r;
  a->y *= r;
  a->z *= r;
  return n;
}

// Rotation of a body-frame vector into the global frame.
RK_HD v3 qrot(qt q, v3 v) {
  v3 u = V(q.x, q.y, q.z);
  v3 t = vcross(u, v);
  t = V(2.0 * t.x, 2.0 * t.y, 2.0 * t.z);
  v3 c = vcross(u, t);
  return V(v.x + q.w * t.x + c.x, v.y + q.w * t.y + c.y, v.z + q.w * t.z + c.z);
}
RK_HD qt qconj(qt q) {
  qt r;
  r.w = q.w;
  r.x = -q.x;
  r.y = -q.y;
  r.z = -q.z;
  return r;
}
// Hamilton product (a.s b.s - a.u.b.u, a.s b.u + b.s a.u + a.u x b.u)
RK_HD qt qmul(qt a, qt b) {
  v3 av = V(a.x, a.y, a.z), bv = V(b.x, b.y, b.z);
  v3 c = vcross(av, bv);
  qt r;
  r.w = a.w * b.w - vdot(av, bv);
  r.x = (a.w * bv.x + b.w * av.x) + c.x;
  r.y = (a.w * bv.y + b.w * av.y) + c.y;
  r.z = (a.w * bv.z + b.w * av.z) + c.z;
  return r;
}
// dQ/dt for a global-frame angular velocity: 1/2 (0,w) (x) Q.
RK_HD qt qdot(qt q, v3 w) {
  v3 c = vcross(w, V(q.x, q.y, q.z));
  qt r;
  r.w = 0.5 * (-(q.x * w.x + q.y * w.y + q.z * w.z));
  r.x = 0.5 * (q.w * w.x + c.x);
  r.y = 0.5 * (q.w * w.y + c.y);
  r.z = 0.5 * (q.w * w.z + c.z);
  return r;
}
// d2Q/dt2 = 1/2 [ (0,dw) (x) Q + (0,w) (x) dQ/dt ].
RK_HD qt qddot(qt q, v3 w, v3 dw) {
  qt d = qdot(q, w);
  qt zdw, zw;
  zdw.w = 0.0;
  zdw.x = dw.x;
  zdw.y = dw.y;
  zdw.z = dw.z;
  zw.w = 0.0;
  zw.x = w.x;
  zw.y = w.y;
  zw.z = w.z;
  qt a = qmul(zdw, q), b = qmul(zw, d);
  qt r;
  r.w = 0.5 * (a.w + b.w);
  r.x = 0.5 * (a.x + b.x);
  r.y = 0.5 * (a.y + b.y);
  r.z = 0.5 * (a.z + b.z);
  return r;
}
RK_HD void qnormalize(qt* q) {
  double n = sqrt(q->w * q->w + (q->x * q->x + q->y * q->y + q->z * q->z));
  double r = 1.0 / n;
  q->w *= r;
  q->x *= r;
  q->y *= r;
  q->z *= r;
}
// Q += (d *= s)
RK_HD qt qadd_scaled(qt q, qt d, double s) {
  qt r;
  r.w = q.w + d.w * s;
  r.x = q.x + d.x * s;
  r.y = q.y + d.y * s;
  r.z = q.z + d.z * s;
  return r;
}

// Oriented bounding box: centre, three unit axes, three half extents.
struct obb_t {
  v3 c;
  v3 e[3];
  v3 ext;
};

RK_HD void obb_enlarge(obb_t* o, double d) {
  o->ext.x += d;
  o->ext.y += d;
  o->ext.z += d;
}
// Separating-axis test over the 15 candidate axes (Gottschalk, Lin, Manocha
// 1996); touching boxes intersect.
// Row i of the rotation (and t[i]) is computed just before the test on a's
// axis i needs it, so that a pair separated on an early axis skips the rest;
// every value is the same expression as computed all at once.
RK_HD bool obb_intersect(const obb_t& a, const obb_t& b) {
  double R[3][3], AR[3][3], t[3];
  v3 d = vsub(b.c, a.c);
  double ea[3] = {a.ext.x, a.ext.y, a.ext.z};
  double eb[3] = {b.ext.x, b.ext.y, b.ext.z};
  double ra, rb;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    t[i] = vdot(d, a.e[i]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R[i][k] = vdot(a.e[i], b.e[k]);
      AR[i][k] = fabs(R[i][k]);
    }
    ra = ea[i];
    rb = eb[0] * AR[i][0] + eb[1] * AR[i][1] + eb[2] * AR[i][2];
    if (fabs(t[i]) > ra + rb) return false;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ra = ea[0] * AR[0][k] + ea[1] * AR[1][k] + ea[2] * AR[2][k];
    rb = eb[k];
    if (fabs(t[0] * R[0][k] + t[1] * R[1][k] + t[2] * R[2][k]) > ra + rb) return false;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
      ra = ea[i1] * AR[i2][k] + ea[i2] * AR[i1][k];
      rb = eb[k1] * AR[i][k2] + eb[k2] * AR[i][k1];
      if (fabs(t[i2] * R[i1][k] - t[i1] * R[i2][k]) > ra + rb) return false;
    }
  }
  return true;
}

}  // namespace rk
