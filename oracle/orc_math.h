/*
 * orc_math.h -- vector and quaternion helpers of the CPU oracle (test
 * infrastructure only), with the operand order of the reference's
 * Group3/Vector3/Quaternion classes (paths relative to src/Grains3D/Grains/).
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H
#include <math.h>

/* Basic.hh:22-29 */
#define ORC_PI 3.14159265358979323846
#define ORC_EPSILON 1.0e-10

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static inline void st(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
/* Group3.cpp:149-196 */
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double d) { return mk(a.x * d, a.y * d, a.z * d); }
static inline v3 dvd(v3 a, double d) { return mk(a.x / d, a.y / d, a.z / d); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Vector3.cpp:114-119 */
static inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
/* Vector3.cpp:147-151 */
static inline double norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* Group3.cpp:390-395 */
static inline v3 round3(v3 a) {
  return mk(fabs(a.x) < ORC_EPSILON ? 0. : a.x, fabs(a.y) < ORC_EPSILON ? 0. : a.y,
            fabs(a.z) < ORC_EPSILON ? 0. : a.z);
}

/* Quaternion (vector part v, scalar w), Quaternion.hh:250-251 */
typedef struct { v3 v; double w; } quat;
static inline quat qmk(v3 v, double w) { quat q; q.v = v; q.w = w; return q; }
/* Quaternion.cpp:74-77 */
static inline quat qconj(quat q) { return qmk(neg(q.v), q.w); }
/* Quaternion.cpp:95-100 ([0,lhs] x q, reached through operator,(Vector3,Quaternion)) */
static inline quat qleftvec(quat q, v3 lhs) {
  double tmp = dot(neg(lhs), q.v);
  v3 vt = add(cross(lhs, q.v), scl(lhs, q.w));
  return qmk(vt, tmp);
}
/* Quaternion.cpp:108-113 */
static inline v3 qmulv(quat p, quat r) {
  return add(add(cross(p.v, r.v), scl(r.v, p.w)), scl(p.v, r.w));
}
/* Quaternion.cpp:442-448 */
static inline quat qmul(quat p, quat r) {
  double tmp = (p.w * r.w) - dot(p.v, r.v);
  v3 vt = add(add(cross(p.v, r.v), scl(r.v, p.w)), scl(p.v, r.w));
  return qmk(vt, tmp);
}


/* Transform::composeLeftByRotation (Geometry/src/Transform.cpp:140-156):
   basis <- R(q) basis, row-major basis */
static inline void compose_left_rotation(double* mm, quat dq) {
  double qx = dq.v.x, qy = dq.v.y, qz = dq.v.z, qw = dq.w;
  double nb[9];
  for (int i = 0; i < 3; ++i) {
    double px = qy * mm[6 + i] - qz * mm[3 + i] + qw * mm[0 + i];
    double py = qz * mm[0 + i] - qx * mm[6 + i] + qw * mm[3 + i];
    double pz = qx * mm[3 + i] - qy * mm[0 + i] + qw * mm[6 + i];
    double pw = -qx * mm[0 + i] - qy * mm[3 + i] - qz * mm[6 + i];
    nb[0 + i] = qy * pz - qz * py - pw * qx + qw * px;
    nb[3 + i] = qz * px - qx * pz - pw * qy + qw * py;
    nb[6 + i] = qx * py - qy * px - pw * qz + qw * pz;
  }
  for (int k = 0; k < 9; ++k) mm[k] = nb[k];
}

#endif
