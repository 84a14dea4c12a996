/*
 * gdem_oracle.c -- CPU restatement of the Grains3D sphere DEM step.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the timed CPU
 * baseline); see gdem_oracle.h.  Built with -O2 -std=c99 -ffp-contract=off.
 *
 * All paths below are relative to src/Grains3D/Grains/ of the reference.
 * Every expression keeps the reference's operand order so that rounding is
 * identical: a*b*c is (a*b)*c, "double * Vector3" multiplies each component,
 * cross products and norms follow Vector3.cpp / Group3.cpp.
 */
#include "gdem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_math.h"

typedef struct {
  int kind;
  double p[8];
  double beta;      /* HookeContactForceModel.cpp:19 / HertzContactForceModel.cpp:20 */
  double m2sqrt56;  /* HertzContactForceModel.cpp:21 */
} model_t;

/* One entry of a component's contact map (Component.hh:505-508): key = id of
   the other component (the elementary ids are 1 for spheres), active flag,
   cumulative tangential spring, previous normal, rolling spring torque. */
typedef struct {
  int key, active;
  double kd[3], nrm[3], spr[3];
} mem_t;

typedef struct {
  double basis[9], origin[3], ext[3];
  double inv_basis[9], inv_origin[3];
  double bb_centre[3], bb_ext[3];
  double crust, req, corners[8][3];
  double vel[3], omega[3]; /* Obstacle::m_translationalVelocity, m_angularVelocity */
  double load_f[3], load_t[3]; /* torsor of the last force evaluation, about the centre */
  double circ;  /* circumscribed radius (rectangles) */
  int material;
  int kind;     /* ORC_WALL_BOX or ORC_WALL_RECTANGLE (ext = LX/2, LY/2, 0) */
} wall_t;

/* ObstacleImposedVelocity (Component/src/ObstacleImposedVelocity.cpp:54-257):
   one imposed motion of a wall over [t0, t1] */
typedef struct {
  int wall, type, live, about_centre;
  double t0, t1;
  double vel[3], omega[3], unit[3], centre[3];
  double amp, period, phase;
  double ms_amp[3], ms_period[3], ms_phase[3];
} motion_t;

/* ObstacleImposedForce, ConstantTranslation (ObstacleImposedForce.cpp:30-90) */
typedef struct {
  int wall, live, automass;
  double t0, t1, amp[3], dir[3], mass, vmax, vt[3];
} force_t;

/* per-step state of ObstacleKinematicsVelocity (ObstacleKinematicsVelocity.cpp:
   60-177): translation and rotation over the step, rotation quaternion over
   the step, velocities */
typedef struct {
  v3 trans, rot, vt, vr;
  quat q;
} kin_t;

/* CompositeObstacle (Component/src/CompositeObstacle.cpp): members in attach
   order (wall k, or composite c as -1 - c) */
typedef struct {
  int parent, nmem, restrict_mask;
  int mem[ORC_MAX_WALLS + ORC_MAX_COMPOSITES];
  double centre[3], basis[9], vel[3], omega[3];
  kin_t kin;
  double conf_vt[3], conf_trans[3];  /* its ObstacleKinematicsForce */
} comp_t;

struct orc_sys {
  orc_config cfg;
  int nc[3];
  double edge[3], lc_origin[3];
  int ncell;
  int n;
  double *x, *v, *w, *wbf, *q, *basis, *F, *T;
  double *r, *crust, *mass, *inertia, *req;
  int32_t* mat;
  int nwalls;
  wall_t walls[ORC_MAX_WALLS];
  /* simulation time and imposed wall motions, each wall's in the order of
     ObstacleKinematicsVelocity::append; live = not yet completed */
  double time;
  int nmot, active_prev;
  motion_t mot[ORC_MAX_MOTIONS];
  kin_t wall_kin[ORC_MAX_WALLS];
  int wall_parent[ORC_MAX_WALLS];
  int ncomp;
  comp_t comp[ORC_MAX_COMPOSITES];
  /* imposed forces, the controller memory (function statics of
     ObstacleImposedForce::translationalVelocity, :252-257, shared by all of
     them) and each wall's ObstacleKinematicsForce state */
  int nforce;
  force_t force[ORC_MAX_MOTIONS];
  double pid_dforce_nm1[3], pid_dforce_nm2[3], pid_depl_nm1[3];
  int pid_nonzero;
  double conf_vt[ORC_MAX_WALLS][3], conf_trans[ORC_MAX_WALLS][3];
  /* composites' torsors exist once a step has ended (Grains.cpp:381) */
  int comp_loads;
  int wall_restrict[ORC_MAX_WALLS];   /* Obstacle::setRestrictedGeomDirMotion */
  model_t models[ORC_MAX_MATERIALS][ORC_MAX_MATERIALS];
  /* cell lists: doubly linked, head insertion (Cell.cpp:90-93, push_front) */
  int *head, *next, *prev, *cell;
  int stencil[13][3];
  /* contact log */
  int log_cap, log_n;
  int32_t *log_id0, *log_id1;
  double *log_pt, *log_ov, *log_d, *log_f;
  int64_t ncand;
  /* contact statistics of AppCollision (AppCollision.hh:213-219) and the
     function-static counters of addToContactsFeatures (AppCollision.cpp:163,184) */
  double ov_mean, ov_max, ov_time_max, gjk_mean, ov_count, gjk_count;
  /* contact maps of the memory models (one per particle) */
  int has_memory;
  double dt_mem;
  int* mem_n;
  mem_t* mem;     /* [n][ORC_MEM_CAP] */
  /* error */
  int err_code;
  int32_t err_id0, err_id1;
  double err_value, err_limit;
};

static int set_error(orc_sys* s, int code, int id0, int id1, double val, double lim) {
  if (s->err_code == ORC_OK) {
    s->err_code = code;
    s->err_id0 = id0;
    s->err_id1 = id1;
    s->err_value = val;
    s->err_limit = lim;
  }
  return code;
}

/* ---------------------------------------------------------------------- */
/* LinkedCell::set (Base/src/LinkedCell.cpp:47-119) without ghost layers:
   n_d = (int)((L_d + EPSILON) / h), edge_d = L_d / n_d. */
int orc_create(const orc_config* cfg, orc_sys** out) {
  orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
  if (!s) return ORC_ERR_ARG;
  s->cfg = *cfg;
  for (int d = 0; d < 3; ++d) {
    int n = 1;
    double e = cfg->size[d];
    if (cfg->cell_size > ORC_EPSILON) {
      n = (int)((cfg->size[d] + ORC_EPSILON) / cfg->cell_size);
      if (!n) n = 1;
      e = cfg->size[d] / n;
    }
    if (cfg->periodic[d] && n < 3) { free(s); return ORC_ERR_ARG; }
    s->nc[d] = n;
    s->edge[d] = e;
    s->lc_origin[d] = cfg->origin[d];
  }
  s->ncell = s->nc[0] * s->nc[1] * s->nc[2];
  /* LinkedCell.cpp:965-967, offsets (di, dj, dk) per column */
  static const int cel[3][13] = {{-1, 0, +1, +1, -1, 0, +1, -1, 0, +1, -1, 0, +1},
                                 {0, 0, 0, 0, +1, +1, +1, +1, +1, +1, +1, +1, +1},
                                 {+1, +1, +1, 0, +1, +1, +1, 0, 0, 0, -1, -1, -1}};
  for (int j = 0; j < 13; ++j)
    for (int d = 0; d < 3; ++d) s->stencil[j][d] = cel[d][j];
  s->head = (int*)malloc(sizeof(int) * (size_t)s->ncell);
  for (int c = 0; c < s->ncell; ++c) s->head[c] = -1;
  *out = s;
  return ORC_OK;
}

static void free_particles(orc_sys* s) {
  free(s->x); free(s->v); free(s->w); free(s->wbf); free(s->q); free(s->basis);
  free(s->F); free(s->T); free(s->r); free(s->crust); free(s->mass);
  free(s->inertia); free(s->req); free(s->mat); free(s->next); free(s->prev);
  free(s->cell); free(s->mem_n); free(s->mem);
  s->mem_n = NULL;
  s->mem = NULL;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  free_particles(s);
  free(s->head);
  free(s->log_id0); free(s->log_id1); free(s->log_pt); free(s->log_ov);
  free(s->log_d); free(s->log_f);
  free(s);
}

/* Sphere.cpp:152-155 (volume), Sphere.cpp:61-70 (inertia), Particle.cpp:158-164
   (m = rho V, I *= rho), Component.cpp:205-208 (equivalent radius),
   ParticleKinematics.cpp:197-205 (w_bf from w at setup). */
int orc_set_particles(orc_sys* s, int n, const double* x, const double* v,
                      const double* w, const double* q, const double* radius,
                      const double* density, const double* crust,
                      const int32_t* material) {
  if (n < 0) return ORC_ERR_ARG;
  free_particles(s);
  s->n = n;
  size_t n3 = (size_t)n * 3;
  s->x = (double*)malloc(sizeof(double) * n3);
  s->v = (double*)malloc(sizeof(double) * n3);
  s->w = (double*)malloc(sizeof(double) * n3);
  s->wbf = (double*)malloc(sizeof(double) * n3);
  s->q = (double*)malloc(sizeof(double) * (size_t)n * 4);
  s->basis = (double*)malloc(sizeof(double) * (size_t)n * 9);
  s->F = (double*)calloc(n3, sizeof(double));
  s->T = (double*)calloc(n3, sizeof(double));
  s->r = (double*)malloc(sizeof(double) * (size_t)n);
  s->crust = (double*)malloc(sizeof(double) * (size_t)n);
  s->mass = (double*)malloc(sizeof(double) * (size_t)n);
  s->inertia = (double*)malloc(sizeof(double) * (size_t)n);
  s->req = (double*)malloc(sizeof(double) * (size_t)n);
  s->mat = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  s->next = (int*)malloc(sizeof(int) * (size_t)n);
  s->prev = (int*)malloc(sizeof(int) * (size_t)n);
  s->cell = (int*)malloc(sizeof(int) * (size_t)n);
  s->mem_n = (int*)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
  s->mem = (mem_t*)calloc((size_t)(n > 0 ? n : 1) * ORC_MEM_CAP, sizeof(mem_t));
  for (int i = 0; i < n; ++i) {
    double rr = radius[i], rho = density[i];
    double vol = 4.0 * ORC_PI * pow(rr, 3.) / 3.0;
    double in0 = 8.0 * ORC_PI * pow(rr, 5.) / 15.0;
    s->r[i] = rr;
    s->crust[i] = crust[i];
    s->mass[i] = rho * vol;
    s->inertia[i] = in0 * rho;
    s->req[i] = pow((0.75 / ORC_PI) * vol, 1. / 3.);
    s->mat[i] = material ? material[i] : 0;
    memcpy(s->x + 3 * i, x + 3 * i, sizeof(double) * 3);
    memcpy(s->v + 3 * i, v + 3 * i, sizeof(double) * 3);
    memcpy(s->w + 3 * i, w + 3 * i, sizeof(double) * 3);
    memcpy(s->q + 4 * i, q + 4 * i, sizeof(double) * 4);
    quat qq = qmk(mk(q[4 * i], q[4 * i + 1], q[4 * i + 2]), q[4 * i + 3]);
    quat pc = qconj(qq);
    st(s->wbf + 3 * i, qmulv(pc, qleftvec(qq, ld(w + 3 * i))));
    /* basis = rotation matrix of q (orientation given at insertion) */
    double qx = qq.v.x, qy = qq.v.y, qz = qq.v.z, qw = qq.w;
    double* b = s->basis + 9 * i;
    b[0] = 1. - 2. * (qy * qy + qz * qz); b[1] = 2. * (qx * qy - qz * qw); b[2] = 2. * (qx * qz + qy * qw);
    b[3] = 2. * (qx * qy + qz * qw); b[4] = 1. - 2. * (qx * qx + qz * qz); b[5] = 2. * (qy * qz - qx * qw);
    b[6] = 2. * (qx * qz - qy * qw); b[7] = 2. * (qy * qz + qx * qw); b[8] = 1. - 2. * (qx * qx + qy * qy);
    s->cell[i] = -1;
    s->next[i] = s->prev[i] = -1;
  }
  for (int c = 0; c < s->ncell; ++c) s->head[c] = -1;
  return ORC_OK;
}

/* Transform::setToInverseTransform (Geometry/src/Transform.cpp:71-77) for a
   rotation basis; Convex::bbox (Convex.cpp:29-41) + BBox::setValue
   (BBox.cpp:235-239) for the wall bounding box; Box::setCornersFaces
   (Box.cpp:92-111); Box::getVolume (Box.cpp:273-279). */
static void build_wall(wall_t* wl, const double* basis, const double* centre,
                       const double* half_extent, double crust, int material, int kind) {
  memset(wl, 0, sizeof(*wl));
  wl->kind = kind;
  memcpy(wl->basis, basis, sizeof(double) * 9);
  memcpy(wl->origin, centre, sizeof(double) * 3);
  memcpy(wl->ext, half_extent, sizeof(double) * 3);
  wl->crust = crust;
  wl->material = material;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) wl->inv_basis[3 * i + j] = basis[3 * j + i];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(wl->inv_basis[3 * i], wl->inv_basis[3 * i + 1], wl->inv_basis[3 * i + 2]);
    wl->inv_origin[i] = dot(neg(row), ld(centre));
  }
  double e0 = half_extent[0], e1 = half_extent[1], e2 = half_extent[2];
  double mn[3], mx[3];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(basis[3 * i], basis[3 * i + 1], basis[3 * i + 2]);
    v3 nrow = neg(row);
    /* Box::support (Box.cpp:307-316) */
    v3 sm = norm(nrow) < ORC_EPSILON ? mk(0., 0., 0.)
                                      : mk(nrow.x < 0. ? -e0 : e0, nrow.y < 0. ? -e1 : e1,
                                           nrow.z < 0. ? -e2 : e2);
    v3 sp = norm(row) < ORC_EPSILON ? mk(0., 0., 0.)
                                     : mk(row.x < 0. ? -e0 : e0, row.y < 0. ? -e1 : e1,
                                          row.z < 0. ? -e2 : e2);
    mn[i] = centre[i] + dot(row, sm);
    mx[i] = centre[i] + dot(row, sp);
  }
  for (int i = 0; i < 3; ++i) {
    wl->bb_ext[i] = (mx[i] - mn[i]) / 2.;
    wl->bb_centre[i] = mn[i] + wl->bb_ext[i];
  }
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1},
                               {-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}};
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) wl->corners[k][d] = sg[k][d] < 0 ? -half_extent[d] : half_extent[d];
  if (kind == ORC_WALL_RECTANGLE) {
    /* Rectangle::getVolume (Rectangle.cpp:182-185) is the area LX LY;
       Rectangle::computeCircumscribedRadius (:118-121) */
    double lx = 2. * e0, ly = 2. * e1;
    wl->req = pow((0.75 / ORC_PI) * (lx * ly), 1. / 3.);
    wl->circ = sqrt(lx * lx + ly * ly) / 2.;
    return;
  }
  double vol = 8.0 * e0 * e1 * e2;
  wl->req = pow((0.75 / ORC_PI) * vol, 1. / 3.);
}

int orc_add_wall(orc_sys* s, const double* basis, const double* centre,
                 const double* half_extent, double crust, int material) {
  if (s->nwalls >= ORC_MAX_WALLS || material < 0 || material >= ORC_MAX_MATERIALS)
    return ORC_ERR_ARG;
  s->wall_parent[s->nwalls] = -1;
  s->wall_kin[s->nwalls].q.w = 1.;
  build_wall(&s->walls[s->nwalls++], basis, centre, half_extent, crust, material, ORC_WALL_BOX);
  return ORC_OK;
}

/* Rectangle obstacle (Geometry/src/Rectangle.cpp:13-21): sides lx, ly in
   the local xy plane */
int orc_add_rectangle_wall(orc_sys* s, const double* basis, const double* centre, double lx,
                           double ly, double crust, int material) {
  if (s->nwalls >= ORC_MAX_WALLS || material < 0 || material >= ORC_MAX_MATERIALS ||
      !(lx > 0.) || !(ly > 0.))
    return ORC_ERR_ARG;
  double half[3] = {lx / 2., ly / 2., 0.};
  s->wall_parent[s->nwalls] = -1;
  s->wall_kin[s->nwalls].q.w = 1.;
  build_wall(&s->walls[s->nwalls++], basis, centre, half, crust, material, ORC_WALL_RECTANGLE);
  return ORC_OK;
}

/* ---- imposed wall motions ---------------------------------------------- */

/* Vector3::normalize (Geometry/src/Vector3.cpp:67-70, Group3.cpp:261-267) */
static v3 normalized(double x, double y, double z) {
  v3 u = mk(x, y, z);
  return dvd(u, norm(u));
}

/* constructor (ObstacleImposedVelocity.cpp:54-257) and
   ObstacleKinematicsVelocity::append (ObstacleKinematicsVelocity.cpp:30-41) */
static int add_motion(orc_sys* s, int wall, int type, double t0, double t1, const double* p) {
  if (s->nmot >= ORC_MAX_MOTIONS || !(t1 >= t0)) return ORC_ERR_ARG;
  motion_t m;
  memset(&m, 0, sizeof(m));
  m.wall = wall;
  m.type = type;
  m.t0 = t0;
  m.t1 = t1;
  m.live = 1;
  m.about_centre = 1;
  v3 u;
  switch (type) {
    case ORC_MOTION_CONSTANT_TRANSLATION:
      m.vel[0] = p[0]; m.vel[1] = p[1]; m.vel[2] = p[2];
      break;
    case ORC_MOTION_SIN_TRANSLATION:
      u = normalized(p[0], p[1], p[2]);
      st(m.unit, u);
      m.amp = p[3];
      m.period = p[4];
      m.phase = p[5] * ORC_PI / 180.;
      break;
    case ORC_MOTION_CRENEL_TRANSLATION:
      u = normalized(p[0], p[1], p[2]);
      st(m.unit, u);
      m.amp = p[3];
      m.period = p[4];
      break;
    case ORC_MOTION_MULTISIN_TRANSLATION:
      u = normalized(p[0], p[1], p[2]);
      st(m.unit, u);
      for (int i = 0; i < 3; ++i) {
        m.ms_period[i] = p[6 + i];
        m.ms_phase[i] = p[9 + i] * ORC_PI / 180.;
        m.ms_amp[i] = p[3 + i] * m.unit[i]; /* stored times the direction (:249-254) */
      }
      break;
    case ORC_MOTION_CONSTANT_ROTATION:
      m.omega[0] = p[0]; m.omega[1] = p[1]; m.omega[2] = p[2];
      m.centre[0] = p[3]; m.centre[1] = p[4]; m.centre[2] = p[5];
      m.about_centre = p[6] != 0.;
      break;
    default:
      return ORC_ERR_ARG;
  }
  int pos = s->nmot;
  for (int i = 0; i < s->nmot; ++i)
    if (s->mot[i].wall == wall && !(s->mot[i].t0 < t0)) { pos = i; break; }
  memmove(&s->mot[pos + 1], &s->mot[pos], sizeof(motion_t) * (size_t)(s->nmot - pos));
  s->mot[pos] = m;
  s->nmot++;
  return ORC_OK;
}

int orc_add_wall_motion(orc_sys* s, int wall, int type, double t0, double t1, const double* p) {
  if (wall < 0 || wall >= s->nwalls) return ORC_ERR_ARG;
  return add_motion(s, wall, type, t0, t1, p);
}

/* isActif / isCompleted (ObstacleImposedVelocity.cpp:305-325) */
static int mot_active(const motion_t* m, double td, double te, double dt, double* subint) {
  *subint = te - td;
  if (td < m->t0) *subint -= (m->t0 - td);
  if (m->t1 < te) *subint -= (te - m->t1);
  return *subint > dt * 1.e-5;
}
static int mot_completed(const motion_t* m, double t, double dt) {
  return t > m->t1 + dt * 1.e-5;
}

/* translationalVelocity (ObstacleImposedVelocity.cpp:330-370); updates the
   stored velocity as the reference does */
static v3 mot_velocity(motion_t* m, double time, double dt, v3 cg) {
  if (mot_completed(m, time, dt)) {
    m->vel[0] = m->vel[1] = m->vel[2] = 0.;
  } else if (m->type == ORC_MOTION_SIN_TRANSLATION) {
    double a = m->amp * sin(2. * ORC_PI * (time - m->t0) / m->period + m->phase);
    for (int i = 0; i < 3; ++i) m->vel[i] = a * m->unit[i];
  } else if (m->type == ORC_MOTION_CRENEL_TRANSLATION) {
    double a = m->amp * ((int)((time - m->t0) / m->period) % 2 == 0 ? 1 : -1);
    for (int i = 0; i < 3; ++i) m->vel[i] = a * m->unit[i];
  } else if (m->type == ORC_MOTION_MULTISIN_TRANSLATION) {
    for (int i = 0; i < 3; ++i)
      m->vel[i] = m->ms_amp[i] *
                  sin(2. * ORC_PI * (time - m->t0) / m->ms_period[i] + m->ms_phase[i]) *
                  m->unit[i];
  } else if (m->type == ORC_MOTION_CONSTANT_ROTATION && !m->about_centre) {
    st(m->vel, cross(ld(m->omega), sub(cg, ld(m->centre))));
  }
  return ld(m->vel);
}

/* angularVelocity (:375-381) and angularMotion (:449-455) */
static v3 mot_omega(motion_t* m, double time, double dt) {
  if (mot_completed(m, time, dt)) m->omega[0] = m->omega[1] = m->omega[2] = 0.;
  return ld(m->omega);
}
static v3 mot_rotation(const motion_t* m, double subint) { return scl(ld(m->omega), subint); }

/* rotation vector -> quaternion (sin(d/2)/d rota, cos(d/2)), identity for 0
   (ObstacleKinematicsVelocity.cpp:104-110) */
static quat rot_quat(v3 rota) {
  double d = norm(rota);
  if (d != 0.) return qmk(scl(rota, sin(d / 2.) / d), cos(d / 2.));
  return qmk(mk(0., 0., 0.), 1.);
}

/* translationalMotion (ObstacleImposedVelocity.cpp:387-443) */
static v3 mot_translation(const motion_t* m, double time, double dt, double subint, v3 cg) {
  double ti = time - dt < m->t0 ? m->t0 : time - dt;
  double te = time > m->t1 ? m->t1 : time;
  v3 tr = mk(0., 0., 0.);
  if (m->type == ORC_MOTION_CONSTANT_TRANSLATION) {
    tr = scl(ld(m->vel), subint);
  } else if (m->type == ORC_MOTION_CRENEL_TRANSLATION) {
    double a = m->amp * ((int)((0.5 * (ti + te) - m->t0) / m->period) % 2 == 0 ? 1 : -1) * subint;
    tr = scl(ld(m->unit), a);
  } else if (m->type == ORC_MOTION_SIN_TRANSLATION) {
    double a = (m->amp * m->period / (2. * ORC_PI)) *
               (cos(2. * ORC_PI * (ti - m->t0) / m->period + m->phase) -
                cos(2. * ORC_PI * (te - m->t0) / m->period + m->phase));
    tr = scl(ld(m->unit), a);
  } else if (m->type == ORC_MOTION_MULTISIN_TRANSLATION) {
    double c[3];
    for (int i = 0; i < 3; ++i)
      c[i] = (m->ms_amp[i] * m->ms_period[i] / (2. * ORC_PI)) *
             (cos(2. * ORC_PI * (ti - m->t0) / m->ms_period[i] + m->ms_phase[i]) -
              cos(2. * ORC_PI * (te - m->t0) / m->ms_period[i] + m->ms_phase[i])) *
             m->unit[i];
    tr = ld(c);
  } else if (m->type == ORC_MOTION_CONSTANT_ROTATION && !m->about_centre) {
    quat q = rot_quat(mot_rotation(m, subint));
    v3 lever = sub(cg, ld(m->centre));
    /* Quaternion::rotateVector (Geometry/src/Quaternion.cpp:293-298) */
    double nq = norm(q.v);
    v3 r = add(add(scl(lever, q.w * q.w - nq * nq), scl(q.v, 2. * dot(lever, q.v))),
               scl(cross(q.v, lever), 2. * q.w));
    tr = sub(r, lever);
  }
  return tr;
}

/* Quaternion::rotateVector (Geometry/src/Quaternion.cpp:293-298) */
static v3 q_rotate(quat q, v3 v) {
  double nq = norm(q.v);
  return add(add(scl(v, q.w * q.w - nq * nq), scl(q.v, 2. * dot(v, q.v))),
             scl(cross(q.v, v), 2. * q.w));
}

static kin_t kin_zero(void) {
  kin_t k;
  k.trans = k.rot = k.vt = k.vr = mk(0., 0., 0.);
  k.q = qmk(mk(0., 0., 0.), 1.);
  return k;
}

/* ObstacleKinematicsVelocity::Compose (ObstacleKinematicsVelocity.cpp:62-77) */
static void kin_compose(kin_t* k, const kin_t* o, v3 lever) {
  k->rot = add(k->rot, o->rot);
  k->q = qmul(o->q, k->q);
  k->trans = add(o->trans, sub(q_rotate(o->q, lever), lever));
  k->vr = add(k->vr, o->vr);
  k->vt = add(k->vt, add(o->vt, cross(o->vr, lever)));
}

/* ObstacleKinematicsVelocity::ImposedMotion (ObstacleKinematicsVelocity.cpp:
   86-122) of obstacle `target` (wall k, or composite c as -1 - c): its active
   motions added to its kinematics; returns whether they moved it */
static int kin_imposed(orc_sys* s, kin_t* k, int target, double time, double dt, v3 cg) {
  v3 depl = mk(0., 0., 0.), rota = mk(0., 0., 0.), vt = mk(0., 0., 0.), vr = mk(0., 0., 0.);
  double subint = 0.;
  for (int i = 0; i < s->nmot; ++i) {
    motion_t* m = &s->mot[i];
    if (m->wall != target || !mot_active(m, time - dt, time, dt, &subint)) continue;
    depl = add(depl, mot_translation(m, time, dt, subint, cg));
    rota = add(rota, mot_rotation(m, subint));
    vt = add(vt, mot_velocity(m, time, dt, cg));
    vr = add(vr, mot_omega(m, time, dt));
  }
  quat qrot = rot_quat(rota);
  k->trans = add(k->trans, depl);
  k->rot = add(k->rot, rota);
  k->q = qmul(qrot, k->q);
  k->vr = add(k->vr, vr);
  k->vt = add(k->vt, vt);
  return norm(depl) != 0. || norm(rota) != 0.;
}

/* SimpleObstacle::Move (Component/src/SimpleObstacle.cpp:116-175) of wall w:
   translation (Transform::composeLeftByTranslation, Transform.cpp:300-304),
   rotation of the basis (RigidBody::Rotate, RigidBody.cpp:270-273), new
   bounding box, velocities of the kinematics when it moved */
/* ObstacleImposedForce::translationalVelocity (ObstacleImposedForce.cpp:
   227-296), ConstantTranslation */
static void force_velocity(orc_sys* s, force_t* m, double time, double dt, double subint,
                           v3 force, double crust, double* vt, double* trans) {
  if (m->automass) m->mass = dt * dt * norm(ld(m->amp)) * 1.e7 / (2. * crust);
  double Kp = 0.5 * dt * dt / m->mass;
  double Ki = 2. * Kp / dt;
  double Kd = 3. * Kp * dt / 24.;
  if (!s->pid_nonzero) s->pid_nonzero = norm(force) > ORC_EPSILON;
  double dforce[3];
  for (int i = 0; i < 3; ++i) {
    dforce[i] = (i == 0 ? force.x : i == 1 ? force.y : force.z) - m->amp[i];
    dforce[i] *= m->dir[i];
  }
  if (s->pid_nonzero) {
    double A = Kp + Ki * dt + Kd / dt, B = Kp + 2. * Kd / dt, Cc = Kd / dt;
    double depl[3];
    for (int i = 0; i < 3; ++i) {
      depl[i] = s->pid_depl_nm1[i] + A * dforce[i] - B * s->pid_dforce_nm1[i] +
                Cc * s->pid_dforce_nm2[i];
      m->vt[i] = depl[i] / dt;
      trans[i] = (subint / dt) * depl[i];
    }
    memcpy(s->pid_dforce_nm2, s->pid_dforce_nm1, sizeof(dforce));
    memcpy(s->pid_dforce_nm1, dforce, sizeof(dforce));
    memcpy(s->pid_depl_nm1, depl, sizeof(depl));
  } else {
    for (int i = 0; i < 3; ++i) {
      m->vt[i] = -m->vmax * m->dir[i];
      trans[i] = subint * m->vt[i];
    }
  }
  if (time > m->t1 + dt * 1.e-5) m->vt[0] = m->vt[1] = m->vt[2] = 0.;
  memcpy(vt, m->vt, sizeof(double) * 3);
}

static void move_wall(orc_sys* s, int w, double time, double dt, int mother, int mother_force,
                      int displace) {
  wall_t* wl = &s->walls[w];
  kin_t* k = &s->wall_kin[w];
  int moving = kin_imposed(s, k, w, time, dt, ld(wl->origin)) || mother;
  if (moving && displace) {
    double basis[9], origin[3], ext[3], crust = wl->crust;
    int material = wl->material, kind = wl->kind;
    memcpy(basis, wl->basis, sizeof(basis));
    memcpy(ext, wl->ext, sizeof(ext));
    /* SimpleObstacle.cpp:131-136: restricted directions do not move */
    origin[0] = wl->origin[0] + ((s->wall_restrict[w] & 1) ? 0. : k->trans.x);
    origin[1] = wl->origin[1] + ((s->wall_restrict[w] & 2) ? 0. : k->trans.y);
    origin[2] = wl->origin[2] + ((s->wall_restrict[w] & 4) ? 0. : k->trans.z);
    compose_left_rotation(basis, k->q);
    double lf[3], lt[3];
    memcpy(lf, wl->load_f, sizeof(lf));
    memcpy(lt, wl->load_t, sizeof(lt));
    build_wall(wl, basis, origin, ext, crust, material, kind);
    memcpy(wl->load_f, lf, sizeof(lf));
    memcpy(wl->load_t, lt, sizeof(lt));
  }
  /* imposed forces (ObstacleKinematicsForce::ImposedMotion,
     ObstacleKinematicsForce.cpp:40-63: each active one in turn), then the
     translation (SimpleObstacle.cpp:142-152) */
  double* cvt = s->conf_vt[w];
  double* ctr = s->conf_trans[w];
  for (int i = 0; i < s->nforce; ++i) {
    force_t* m = &s->force[i];
    double sub_dt = 0.;
    if (m->wall != w) continue;
    sub_dt = time - (time - dt);
    if (time - dt < m->t0) sub_dt -= (m->t0 - (time - dt));
    if (m->t1 < time) sub_dt -= (time - m->t1);
    if (!(sub_dt > dt * 1.e-5)) continue;
    force_velocity(s, m, time, dt, sub_dt, ld(wl->load_f), wl->crust, cvt, ctr);
  }
  int move_force = norm(ld(cvt)) != 0. || mother_force;
  if (move_force && displace) {
    double basis[9], origin[3], ext[3], crust = wl->crust;
    int material = wl->material, kind = wl->kind;
    double lf[3], lt[3];
    memcpy(basis, wl->basis, sizeof(basis));
    memcpy(ext, wl->ext, sizeof(ext));
    memcpy(lf, wl->load_f, sizeof(lf));
    memcpy(lt, wl->load_t, sizeof(lt));
    for (int d = 0; d < 3; ++d)
      origin[d] = wl->origin[d] + ((s->wall_restrict[w] >> d) & 1 ? 0. : ctr[d]);
    build_wall(wl, basis, origin, ext, crust, material, kind);
    memcpy(wl->load_f, lf, sizeof(lf));
    memcpy(wl->load_t, lt, sizeof(lt));
  }
  if (!moving && !move_force) return;
  /* Obstacle::setVelocity (Obstacle.cpp:635-640): both kinematics */
  st(wl->vel, add(k->vt, ld(cvt)));
  st(wl->omega, k->vr);
}

static void move_members(orc_sys* s, int parent, double time, double dt, int mother,
                         int mother_force, int displace);

/* CompositeObstacle::getTorsor (CompositeObstacle.cpp:650-660), its force:
   zero, then += each member's in attach order */
static v3 composite_load(const orc_sys* s, int c) {
  const comp_t* C = &s->comp[c];
  v3 f = mk(0., 0., 0.);
  for (int i = 0; i < C->nmem; ++i) {
    int mbr = C->mem[i];
    f = add(f, mbr < 0 ? composite_load(s, -1 - mbr) : ld(s->walls[mbr].load_f));
  }
  return f;
}

/* CompositeObstacle::Move (Component/src/CompositeObstacle.cpp:135-220) */
static void move_composite(orc_sys* s, int c, double time, double dt, int mother,
                           int mother_force, int displace) {
  comp_t* C = &s->comp[c];
  int moving = kin_imposed(s, &C->kin, -1 - c, time, dt, ld(C->centre)) || mother;
  if (moving && displace) {
    /* CompositeObstacle.cpp:153-160: the composite's own translation only */
    v3 tr = C->kin.trans;
    if (C->restrict_mask & 1) tr.x = 0.;
    if (C->restrict_mask & 2) tr.y = 0.;
    if (C->restrict_mask & 4) tr.z = 0.;
    st(C->centre, add(ld(C->centre), tr));
    compose_left_rotation(C->basis, C->kin.q);
  }
  if (moving)
    for (int i = 0; i < C->nmem; ++i) {
      int mbr = C->mem[i];
      if (mbr < 0) {
        comp_t* M = &s->comp[-1 - mbr];
        kin_compose(&M->kin, &C->kin, sub(ld(M->centre), ld(C->centre)));
      } else {
        kin_compose(&s->wall_kin[mbr], &C->kin, sub(ld(s->walls[mbr].origin), ld(C->centre)));
      }
    }
  /* imposed forces on the composite's load (CompositeObstacle.cpp:182-203) */
  v3 load = s->comp_loads ? composite_load(s, c) : mk(0., 0., 0.);
  for (int i = 0; i < s->nforce; ++i) {
    force_t* m = &s->force[i];
    double sub_dt;
    if (m->wall != -1 - c) continue;
    sub_dt = time - (time - dt);
    if (time - dt < m->t0) sub_dt -= (m->t0 - (time - dt));
    if (m->t1 < time) sub_dt -= (time - m->t1);
    if (!(sub_dt > dt * 1.e-5)) continue;
    force_velocity(s, m, time, dt, sub_dt, load, 0., C->conf_vt, C->conf_trans);
  }
  int move_force = norm(ld(C->conf_vt)) != 0. || mother_force;
  if (move_force && displace)
    for (int d = 0; d < 3; ++d)
      C->centre[d] = C->centre[d] + ((C->restrict_mask >> d) & 1 ? 0. : C->conf_trans[d]);
  /* ObstacleKinematicsForce::Compose (ObstacleKinematicsForce.cpp:51-55) */
  if (move_force)
    for (int i = 0; i < C->nmem; ++i) {
      int mbr = C->mem[i];
      double* vt = mbr < 0 ? s->comp[-1 - mbr].conf_vt : s->conf_vt[mbr];
      double* tr = mbr < 0 ? s->comp[-1 - mbr].conf_trans : s->conf_trans[mbr];
      st(vt, add(ld(vt), ld(C->conf_vt)));
      st(tr, add(ld(tr), ld(C->conf_trans)));
    }
  move_members(s, c, time, dt, moving, move_force, displace);
  if (moving || move_force) {
    st(C->vel, add(C->kin.vt, ld(C->conf_vt)));
    st(C->omega, C->kin.vr);
  }
}

static void move_members(orc_sys* s, int parent, double time, double dt, int mother,
                         int mother_force, int displace) {
  for (int k = 0; k < s->ncomp; ++k)
    if (s->comp[k].parent == parent) move_composite(s, k, time, dt, mother, mother_force, displace);
  for (int w = 0; w < s->nwalls; ++w)
    if (s->wall_parent[w] == parent) move_wall(s, w, time, dt, mother, mother_force, displace);
}

/* Obstacle::resetKinematics (Obstacle.cpp:160-167), from the root down
   (CompositeObstacle.cpp:449-460) */
static void reset_walls(orc_sys* s) {
  for (int w = 0; w < s->nwalls; ++w) {
    for (int d = 0; d < 3; ++d) s->walls[w].vel[d] = s->walls[w].omega[d] = 0.;
    s->wall_kin[w] = kin_zero();
    /* ObstacleKinematicsForce::reset (ObstacleKinematicsForce.cpp:118-124) */
    for (int d = 0; d < 3; ++d) s->conf_vt[w][d] = s->conf_trans[w][d] = 0.;
  }
  for (int c = 0; c < s->ncomp; ++c) {
    s->comp[c].kin = kin_zero();
    for (int d = 0; d < 3; ++d)
      s->comp[c].vel[d] = s->comp[c].omega[d] = s->comp[c].conf_vt[d] = s->comp[c].conf_trans[d] = 0.;
  }
}

static int any_live(const orc_sys* s) {
  for (int i = 0; i < s->nmot; ++i)
    if (s->mot[i].live) return 1;
  for (int i = 0; i < s->nforce; ++i)
    if (s->force[i].live) return 1;
  return 0;
}

static void drop_completed(orc_sys* s, double time, double dt) {
  for (int i = 0; i < s->nmot; ++i)
    if (s->mot[i].live && mot_completed(&s->mot[i], time, dt)) s->mot[i].live = 0;
  for (int i = 0; i < s->nforce; ++i)
    if (s->force[i].live && time > s->force[i].t1 + dt * 1.e-5) s->force[i].live = 0;
}

/* ObstacleImposedForce constructor, ConstantTranslation (:52-72) and
   ObstacleKinematicsForce::append (ObstacleKinematicsForce.cpp:28-31) */
int orc_add_wall_force(orc_sys* s, int wall, double t0, double t1, const double* p) {
  if (wall < 0 || wall >= s->nwalls || s->nforce >= ORC_MAX_MOTIONS || !(t1 >= t0))
    return ORC_ERR_ARG;
  force_t* f = &s->force[s->nforce++];
  memset(f, 0, sizeof(*f));
  f->wall = wall;
  f->live = 1;
  f->t0 = t0;
  f->t1 = t1;
  st(f->amp, ld(p));
  st(f->dir, dvd(ld(p), norm(ld(p))));
  f->automass = !(p[3] > 0.);
  f->mass = f->automass ? 0. : p[3];
  f->vmax = p[4];
  return ORC_OK;
}

/* ObstacleImposedForce on a composite (CompositeObstacle::LinkImposedMotion,
   CompositeObstacle.cpp:112-120) */
int orc_add_composite_force(orc_sys* s, int comp, double t0, double t1, const double* p) {
  if (comp < 0 || comp >= s->ncomp || s->nforce >= ORC_MAX_MOTIONS || !(t1 >= t0) ||
      !(p[3] > 0.))
    return ORC_ERR_ARG;
  force_t* f = &s->force[s->nforce++];
  memset(f, 0, sizeof(*f));
  f->wall = -1 - comp;
  f->live = 1;
  f->t0 = t0;
  f->t1 = t1;
  st(f->amp, ld(p));
  st(f->dir, dvd(ld(p), norm(ld(p))));
  f->mass = p[3];
  f->vmax = p[4];
  return ORC_OK;
}

/* obstacle part of AllComponents::Move (Component/src/AllComponents.cpp:
   151-220) at `time`, over [time - dt, time] */
int orc_move_walls(orc_sys* s, double time, double dt) {
  s->time = time;
  if (!any_live(s)) return ORC_OK;
  int anyactive = 0;
  double subint;
  for (int i = 0; i < s->nmot; ++i)
    if (s->mot[i].live && mot_active(&s->mot[i], time - dt, time, dt, &subint)) anyactive = 1;
  for (int i = 0; i < s->nforce; ++i) {
    const force_t* f = &s->force[i];
    if (!f->live) continue;
    subint = time - (time - dt);   /* ObstacleImposedForce::isActif (:160-167) */
    if (time - dt < f->t0) subint -= (f->t0 - (time - dt));
    if (f->t1 < time) subint -= (time - f->t1);
    if (subint > dt * 1.e-5) anyactive = 1;
  }
  if (anyactive || s->active_prev) reset_walls(s);
  s->active_prev = anyactive;
  if (anyactive) move_members(s, -1, time, dt, 0, 0, 1);
  drop_completed(s, time, dt);
  s->comp_loads = 1;
  return ORC_OK;
}

/* AllComponents::setKinematicsObstacleWithoutMoving (AllComponents.cpp:235-251) */
static void init_wall_kinematics(orc_sys* s, double time, double dt) {
  if (!any_live(s)) return;
  reset_walls(s);
  move_members(s, -1, time, dt, 0, 0, 0);
  drop_completed(s, time, dt);
}

/* CompositeObstacle::computeVolumeCenterOfMass (CompositeObstacle.cpp:228-255),
   SimpleObstacle's (SimpleObstacle.cpp:732-735) with Box::getVolume (Box.cpp:
   273-279): volume-weighted centre of the members, in attach order */
static double volume_centre(orc_sys* s, int c, v3* out) {
  comp_t* C = &s->comp[c];
  v3 centre = mk(0., 0., 0.);
  double vol = 0.;
  for (int i = 0; i < C->nmem; ++i) {
    int mbr = C->mem[i];
    v3 pc;
    double pv;
    if (mbr < 0) {
      pv = volume_centre(s, -1 - mbr, &pc);
    } else {
      const wall_t* wl = &s->walls[mbr];
      pc = ld(wl->origin);
      pv = 8.0 * wl->ext[0] * wl->ext[1] * wl->ext[2];
    }
    centre = add(centre, scl(pc, pv));
    vol += pv;
  }
  if (vol > 0.) centre = dvd(centre, vol);
  st(C->centre, centre);
  *out = centre;
  return vol;
}

static void composite_centres(orc_sys* s) {
  v3 c;
  for (int k = 0; k < s->ncomp; ++k)
    if (s->comp[k].parent < 0) volume_centre(s, k, &c);
}

int orc_add_composite(orc_sys* s, int parent, int* id) {
  if (parent < -1 || parent >= s->ncomp || s->ncomp >= ORC_MAX_COMPOSITES) return ORC_ERR_ARG;
  comp_t* C = &s->comp[s->ncomp];
  memset(C, 0, sizeof(*C));
  C->parent = parent;
  C->basis[0] = C->basis[4] = C->basis[8] = 1.;
  C->kin = kin_zero();
  if (parent >= 0) s->comp[parent].mem[s->comp[parent].nmem++] = -1 - s->ncomp;
  *id = s->ncomp++;
  composite_centres(s);
  return ORC_OK;
}

int orc_attach_wall(orc_sys* s, int wall, int c) {
  if (wall < 0 || wall >= s->nwalls || c < 0 || c >= s->ncomp || s->wall_parent[wall] != -1)
    return ORC_ERR_ARG;
  s->wall_parent[wall] = c;
  s->comp[c].mem[s->comp[c].nmem++] = wall;
  composite_centres(s);
  return ORC_OK;
}

int orc_add_composite_motion(orc_sys* s, int c, int type, double t0, double t1, const double* p) {
  if (c < 0 || c >= s->ncomp) return ORC_ERR_ARG;
  return add_motion(s, -1 - c, type, t0, t1, p);
}

int orc_set_motion_restriction(orc_sys* s, int wall, int c, int mask) {
  if (mask < 0 || mask > 7) return ORC_ERR_ARG;
  if (wall >= 0) {
    if (wall >= s->nwalls || c >= 0) return ORC_ERR_ARG;
    s->wall_restrict[wall] = mask;
    return ORC_OK;
  }
  if (c < 0 || c >= s->ncomp) return ORC_ERR_ARG;
  s->comp[c].restrict_mask = mask;
  return ORC_OK;
}

int orc_get_composite(const orc_sys* s, int c, double* centre, double* vel, double* omega) {
  if (c < 0 || c >= s->ncomp) return ORC_ERR_ARG;
  if (centre) memcpy(centre, s->comp[c].centre, sizeof(double) * 3);
  if (vel) memcpy(vel, s->comp[c].vel, sizeof(double) * 3);
  if (omega) memcpy(omega, s->comp[c].omega, sizeof(double) * 3);
  return ORC_OK;
}

/* CylindricalShell with box elements (Component/src/CylindricalShell.cpp:
   13-93): box i of edges (width, ly, height) at radius + width/2, angle
   (i + 1/2) 2 pi / n, rotated by that angle about z, placed by the shell's
   transform (Transform::composeLeftByTransform, Transform.cpp:355-360;
   Matrix products Matrix.cpp:517-549), then translated by zero */
void orc_cylindrical_shell_boxes(const double* sb, const double* sc, double height, double radius,
                                 double width, int n, double* basis, double* centre,
                                 double* half) {
  double ext_rad = radius + width / 2.;
  double delta = 2. * ORC_PI / (double)n;
  double ly = 2. * (radius + width) * (sin(delta) - (1. - cos(delta)) / tan(delta));
  for (int i = 0; i < n; ++i) {
    double angle = (0.5 + (double)i) * delta;
    double ca = cos(angle), sa = sin(angle);
    double cg[3] = {ext_rad * ca, ext_rad * sa, 0.};
    double m[9] = {ca, -sa, 0., sa, ca, 0., 0., 0., 1.};
    for (int r = 0; r < 3; ++r) {
      centre[3 * i + r] = sc[r] + (sb[3 * r] * cg[0] + sb[3 * r + 1] * cg[1] + sb[3 * r + 2] * cg[2]);
      for (int col = 0; col < 3; ++col)
        basis[9 * i + 3 * r + col] =
            sb[3 * r] * m[col] + sb[3 * r + 1] * m[3 + col] + sb[3 * r + 2] * m[6 + col];
      centre[3 * i + r] = centre[3 * i + r] + 0.;
    }
    /* Box(width, ly, height), Box.cpp:27-33 */
    half[3 * i] = fabs(width) / 2.;
    half[3 * i + 1] = fabs(ly) / 2.;
    half[3 * i + 2] = fabs(height) / 2.;
  }
}

int orc_add_cylindrical_shell(orc_sys* s, const double* basis, const double* centre,
                              double height, double radius, double width, int n, double crust,
                              int material, int parent, int* id) {
  if (n < 3 || s->nwalls + n > ORC_MAX_WALLS) return ORC_ERR_ARG;
  double* wb = (double*)malloc(sizeof(double) * 15 * (size_t)n);
  if (!wb) return ORC_ERR_ARG;
  orc_cylindrical_shell_boxes(basis, centre, height, radius, width, n, wb, wb + 9 * n,
                              wb + 12 * n);
  int rc = orc_add_composite(s, parent, id);
  for (int i = 0; i < n && !rc; ++i) {
    rc = orc_add_wall(s, wb + 9 * i, wb + 9 * n + 3 * i, wb + 12 * n + 3 * i, crust, material);
    if (!rc) rc = orc_attach_wall(s, s->nwalls - 1, *id);
  }
  free(wb);
  return rc;
}

/* CylindricalShell with Rectangle elements (CylindricalShell.cpp:55-87,
   m_box false): Rectangle(height, ly) at the inner radius, basis
   R(angle) * M (Matrix::operator*=, Matrix.cpp:447-460) with M the rows
   (0, 0, -1), (0, 1, 0), (1, 0, 0), then the shell's transform */
void orc_cylindrical_shell_rectangles(const double* sb, const double* sc, double height,
                                      double radius, double width, int n, double* basis,
                                      double* centre, double* half) {
  double delta = 2. * ORC_PI / (double)n;
  double ly = 2. * (radius + width) * (sin(delta) - (1. - cos(delta)) / tan(delta));
  static const double M[3][3] = {{0., 0., -1.}, {0., 1., 0.}, {1., 0., 0.}};
  for (int i = 0; i < n; ++i) {
    double angle = (0.5 + (double)i) * delta;
    double ca = cos(angle), sa = sin(angle);
    double cg[3] = {radius * ca, radius * sa, 0.};
    double R[3][3] = {{ca, -sa, 0.}, {sa, ca, 0.}, {0., 0., 1.}}, m[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m[r][c] = R[r][0] * M[0][c] + R[r][1] * M[1][c] + R[r][2] * M[2][c];
    for (int r = 0; r < 3; ++r) {
      double* cr = centre + 3 * i + r;
      *cr = sc[r] + (sb[3 * r] * cg[0] + sb[3 * r + 1] * cg[1] + sb[3 * r + 2] * cg[2]);
      for (int c = 0; c < 3; ++c)
        basis[9 * i + 3 * r + c] = sb[3 * r] * m[0][c] + sb[3 * r + 1] * m[1][c] + sb[3 * r + 2] * m[2][c];
      *cr = *cr + 0.;
    }
    half[3 * i] = height / 2.;
    half[3 * i + 1] = ly / 2.;
    half[3 * i + 2] = 0.;
  }
}

int orc_add_cylindrical_shell_rectangles(orc_sys* s, const double* basis, const double* centre,
                                         double height, double radius, double width, int n,
                                         double crust, int material, int parent, int* id) {
  if (n < 3 || s->nwalls + n > ORC_MAX_WALLS) return ORC_ERR_ARG;
  double* wb = (double*)malloc(sizeof(double) * 15 * (size_t)n);
  if (!wb) return ORC_ERR_ARG;
  orc_cylindrical_shell_rectangles(basis, centre, height, radius, width, n, wb, wb + 9 * n,
                                   wb + 12 * n);
  int rc = orc_add_composite(s, parent, id);
  for (int i = 0; i < n && !rc; ++i) {
    const double* h = wb + 12 * n + 3 * i;
    rc = orc_add_rectangle_wall(s, wb + 9 * i, wb + 9 * n + 3 * i, 2. * h[0], 2. * h[1], crust,
                                material);
    if (!rc) rc = orc_attach_wall(s, s->nwalls - 1, *id);
  }
  free(wb);
  return rc;
}

int orc_set_time(orc_sys* s, double time) {
  s->time = time;
  return ORC_OK;
}

double orc_get_time(const orc_sys* s) { return s->time; }

int orc_get_wall_loads(const orc_sys* s, double* force, double* torque) {
  for (int w = 0; w < s->nwalls; ++w) {
    if (force) memcpy(force + 3 * w, s->walls[w].load_f, sizeof(double) * 3);
    if (torque) memcpy(torque + 3 * w, s->walls[w].load_t, sizeof(double) * 3);
  }
  return ORC_OK;
}

int orc_get_wall(const orc_sys* s, int w, double* centre, double* basis, double* vel,
                 double* omega) {
  if (w < 0 || w >= s->nwalls) return ORC_ERR_ARG;
  const wall_t* wl = &s->walls[w];
  if (centre) memcpy(centre, wl->origin, sizeof(wl->origin));
  if (basis) memcpy(basis, wl->basis, sizeof(wl->basis));
  if (vel) memcpy(vel, wl->vel, sizeof(wl->vel));
  if (omega) memcpy(omega, wl->omega, sizeof(wl->omega));
  return ORC_OK;
}

/* HookeContactForceModel ctor (Component/src/HookeContactForceModel.cpp:11-20),
   HertzContactForceModel ctor (HertzContactForceModel.cpp:12-22);
   ContactBuilderFactory::contactForceModel (ContactBuilderFactory.cpp:22-26)
   keys a model by the OR of two material bits, i.e. an unordered pair. */
int orc_set_contact_model(orc_sys* s, int a, int b, int kind, const double* p) {
  if (a < 0 || b < 0 || a >= ORC_MAX_MATERIALS || b >= ORC_MAX_MATERIALS) return ORC_ERR_ARG;
  int np = kind == ORC_MODEL_HOOKE || kind == ORC_MODEL_HERTZ ? 5
         : kind == ORC_MODEL_HOOKE_MEMORY ? 8 : kind == ORC_MODEL_HERTZ_MEMORY ? 7 : 0;
  if (np == 0) return ORC_ERR_ARG;
  model_t m;
  memset(&m, 0, sizeof(m));
  m.kind = kind;
  memcpy(m.p, p, sizeof(double) * (size_t)np);
  double en = p[1];
  /* HookeMemoryContactForceModel.cpp:28, HertzMemoryContactForceModel.cpp:29-30 */
  m.beta = log(en) / sqrt(ORC_PI * ORC_PI + log(en) * log(en));
  m.m2sqrt56 = -2. * sqrt(5. / 6.);
  s->models[a][b] = m;
  s->models[b][a] = m;
  if (kind >= ORC_MODEL_HOOKE_MEMORY) s->has_memory = 1;
  return ORC_OK;
}

int orc_set_time_step(orc_sys* s, double dt) {
  if (!(dt > 0.)) return ORC_ERR_ARG;
  s->dt_mem = dt;
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Cell::GetCell (Base/src/Cell.cpp:172-184).  Returns the cell number
   j*nk*ni + k*ni + i (LinkedCell.cpp:1165-1168) or -1 outside the grid. */
static int get_cell(const orc_sys* s, const double* p, int* ijk) {
  int id[3];
  for (int d = 0; d < 3; ++d) {
    id[d] = (int)floor((p[d] - s->lc_origin[d]) / s->edge[d]);
    if (s->cfg.periodic[d]) {
      if (id[d] >= s->nc[d]) id[d] = s->nc[d] - 1;
      if (id[d] < 0) id[d] = 0;
    }
    if (id[d] < 0 || id[d] >= s->nc[d]) return -1;
  }
  if (ijk) { ijk[0] = id[0]; ijk[1] = id[1]; ijk[2] = id[2]; }
  return id[1] * s->nc[2] * s->nc[0] + id[2] * s->nc[0] + id[0];
}

static void cell_push_front(orc_sys* s, int c, int p) {
  s->prev[p] = -1;
  s->next[p] = s->head[c];
  if (s->head[c] >= 0) s->prev[s->head[c]] = p;
  s->head[c] = p;
  s->cell[p] = c;
}

static void cell_remove(orc_sys* s, int p) {
  int c = s->cell[p];
  if (s->prev[p] >= 0) s->next[s->prev[p]] = s->next[p];
  else s->head[c] = s->next[p];
  if (s->next[p] >= 0) s->prev[s->next[p]] = s->prev[p];
  s->next[p] = s->prev[p] = -1;
  s->cell[p] = -1;
}

/* LinkedCell::Link(Particle*) (LinkedCell.cpp:1352-1363), called at insertion
   in particle order. */
int orc_link(orc_sys* s) {
  for (int c = 0; c < s->ncell; ++c) s->head[c] = -1;
  for (int p = 0; p < s->n; ++p) {
    int c = get_cell(s, s->x + 3 * p, NULL);
    if (c < 0) return set_error(s, ORC_ERR_SIMULATION, p, -1, 0., 0.);
    cell_push_front(s, c, p);
  }
  return ORC_OK;
}

/* LinkedCell::LinkUpdate / LinkUpdateActiveParticle (LinkedCell.cpp:1452-1547) */
int orc_link_update(orc_sys* s) {
  for (int p = 0; p < s->n; ++p) {
    int c = get_cell(s, s->x + 3 * p, NULL);
    if (c < 0) return set_error(s, ORC_ERR_SIMULATION, p, -1, 0., 0.);
    if (c != s->cell[p]) {
      cell_remove(s, p);
      cell_push_front(s, c, p);
    }
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* ClosestPointSPHERESPHERE (Geometry/src/RigidBodyWithCrust.cpp:596-640) */
int orc_sphere_sphere(const double* a, double ra, double cra, const double* b,
                      double rb, double crb, double* point, double* ovec,
                      double* dist) {
  v3 A = ld(a), B = ld(b);
  v3 AB = sub(B, A);
  double distance = norm(AB) - (ra + rb);
  *dist = distance;
  if (distance > 0.) return 0;
  if (-distance >= cra + crb) return -1;
  v3 c = add(A, dvd(scl(AB, ra + 0.5 * distance), norm(AB)));
  v3 ov = scl(dvd(AB, norm(AB)), distance);
  st(point, c);
  st(ovec, ov);
  return 1;
}

/* Box::ContactCornerSPHERE (Box.cpp:725-742) */
static v3 box_corner(const wall_t* wl, v3 g, double r, int k, double* overlap) {
  v3 cp = mk(0., 0., 0.);
  v3 dist = sub(g, ld(wl->corners[k]));
  double nd = norm(dist);
  if (nd < r) {
    *overlap = nd - r;
    cp = sub(g, scl(dist, 1. + 0.5 * *overlap / nd));
  }
  return cp;
}

/* Box::ContactEdgeSPHERE (Box.cpp:750-769) */
static v3 box_edge(const wall_t* wl, v3 g, double r, int k, int dir, double* overlap) {
  v3 cp = mk(0., 0., 0.);
  double pr[3] = {g.x, g.y, g.z};
  pr[dir] = wl->corners[k][dir];
  v3 dist = sub(ld(pr), ld(wl->corners[k]));
  double nd = norm(dist);
  if (nd < r) {
    *overlap = nd - r;
    cp = sub(g, scl(dist, 1. + 0.5 * *overlap / nd));
  }
  return cp;
}

/* Box::IntersectionPointSPHERE (Box.cpp:466-716), 27-region classification */
static v3 box_sphere(const wall_t* wl, v3 g, double r, double* overlap) {
  const double ex = wl->ext[0], ey = wl->ext[1], ez = wl->ext[2];
  v3 cp = mk(0., 0., 0.);
  double nd;
  *overlap = 1.;
  if (g.x > ex) {
    if (g.y > ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 6, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 5, overlap);
      else cp = box_edge(wl, g, r, 5, 2, overlap);
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 2, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 1, overlap);
      else cp = box_edge(wl, g, r, 1, 2, overlap);
    } else {
      if (g.z > ez) cp = box_edge(wl, g, r, 6, 1, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 1, 1, overlap);
      else {
        nd = g.x - ex;
        if (nd < r) { *overlap = nd - r; cp = mk(ex + 0.5 * *overlap, g.y, g.z); }
      }
    }
  } else if (g.x < -ex) {
    if (g.y > ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 7, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 4, overlap);
      else cp = box_edge(wl, g, r, 4, 2, overlap);
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 3, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 0, overlap);
      else cp = box_edge(wl, g, r, 0, 2, overlap);
    } else {
      if (g.z > ez) cp = box_edge(wl, g, r, 3, 1, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 0, 1, overlap);
      else {
        nd = -ex - g.x;
        if (nd < r) { *overlap = nd - r; cp = mk(-ex - 0.5 * *overlap, g.y, g.z); }
      }
    }
  } else {
    if (g.y > ey) {
      if (g.z > ez) cp = box_edge(wl, g, r, 6, 0, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 4, 0, overlap);
      else {
        nd = g.y - ey;
        if (nd < r) { *overlap = nd - r; cp = mk(g.x, ey + 0.5 * *overlap, g.z); }
      }
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_edge(wl, g, r, 3, 0, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 0, 0, overlap);
      else {
        nd = -ey - g.y;
        if (nd < r) { *overlap = nd - r; cp = mk(g.x, -ey - 0.5 * *overlap, g.z); }
      }
    } else {
      if (g.z > ez) {
        nd = g.z - ez;
        if (nd < r) { *overlap = nd - r; cp = mk(g.x, g.y, ez + 0.5 * *overlap); }
      } else if (g.z < -ez) {
        nd = -ez - g.z;
        if (nd < r) { *overlap = nd - r; cp = mk(g.x, g.y, -ez - 0.5 * *overlap); }
      } else {
        *overlap = -10000.;
      }
    }
  }
  return cp;
}

/* Transform::operator() (Transform.cpp:312-317) */
static v3 xform(const double* basis, const double* origin, v3 p) {
  return mk(dot(mk(basis[0], basis[1], basis[2]), p) + origin[0],
            dot(mk(basis[3], basis[4], basis[5]), p) + origin[1],
            dot(mk(basis[6], basis[7], basis[8]), p) + origin[2]);
}

/* ClosestPointSPHEREBOX, sphere first (RigidBodyWithCrust.cpp:647-683) */
static int sphere_wall(const wall_t* wl, v3 A, double r, double cr, v3* point,
                       v3* ov, double* dist) {
  double overlap = 0.;
  v3 local = xform(wl->inv_basis, wl->inv_origin, A);
  v3 cp = box_sphere(wl, local, r, &overlap);
  *dist = overlap;
  if (overlap < 0.) {
    if (-overlap >= cr + wl->crust) return -1;
    v3 contact = xform(wl->basis, wl->origin, cp);
    v3 AB = sub(contact, A);
    *ov = scl(AB, overlap / norm(AB));
    *point = contact;
    return 1;
  }
  return 0;
}

/* ClosestPointRECTANGLE with the particle first (RigidBodyWithCrust.cpp:
   1107-1226, branch :1172-1218, checkoverlap and checkCGInRec false), after
   ClosestPoint's circumscribed-sphere test (:282-296; m_colDetBoundingVolume
   is 0 by default, GrainsExec.cpp:53).  B is the sphere's accumulated basis
   (its Transform, rotated by composeLeftByRotation at every move).  The
   rectangle normal rNorm = basis * (0, 0, 1) is not normalised
   (Vector3::normalized returns a copy that is discarded). */
static int sphere_rect(const wall_t* wl, v3 A, double r, const double* B, v3* point, v3* ov,
                       double* dist) {
  v3 rp = ld(wl->origin);
  v3 gcagcb = sub(A, rp);
  if (!(norm(gcagcb) < r + wl->circ)) return 0;
  /* Matrix * Vector3 (Matrix.cpp:517-520) with (0, 0, 1) */
  v3 rn = mk(dot(mk(wl->basis[0], wl->basis[1], wl->basis[2]), mk(0., 0., 1.)),
             dot(mk(wl->basis[3], wl->basis[4], wl->basis[5]), mk(0., 0., 1.)),
             dot(mk(wl->basis[6], wl->basis[7], wl->basis[8]), mk(0., 0., 1.)));
  double sg = copysign(1., dot(rn, sub(A, rp)));
  rn = mk(sg * rn.x, sg * rn.y, sg * rn.z);
  /* (-rNorm) * basis: Matrix::tdot (Matrix.cpp:238-241, 527-530) */
  v3 d = neg(rn);
  v3 dl = mk(B[0] * d.x + B[3] * d.y + B[6] * d.z, B[1] * d.x + B[4] * d.y + B[7] * d.z,
             B[2] * d.x + B[5] * d.y + B[8] * d.z);
  /* Sphere::support (Sphere.cpp:99-109) */
  v3 sp = mk(0., 0., 0.);
  double s = norm(dl);
  if (s > ORC_EPSILON) {
    double rr = r / s;
    sp = mk(dl.x * rr, dl.y * rr, dl.z * rr);
  }
  double Ao[3] = {A.x, A.y, A.z};
  v3 pa = xform(B, Ao, sp);
  if (!(dot(rn, sub(pa, rp)) < 0.)) return 0;
  double k = dot(sub(rp, pa), rn);
  v3 pb = add(mk(k * rn.x, k * rn.y, k * rn.z), pa);
  /* Rectangle::isIn (Rectangle.cpp:365-370) in the rectangle's frame */
  v3 l = xform(wl->inv_basis, wl->inv_origin, pb);
  if (!(fabs(l.z) < ORC_EPSILON && fabs(l.y) < wl->ext[1] && fabs(l.x) < wl->ext[0])) return 0;
  *point = add(mk(pa.x / 2.0, pa.y / 2.0, pa.z / 2.0), mk(pb.x / 2.0, pb.y / 2.0, pb.z / 2.0));
  *ov = sub(pb, pa);
  *dist = -norm(*ov);
  return 1;
}

int orc_sphere_rectangle(const double* centre, double r, const double* sphere_basis,
                         const double* basis, const double* origin, double lx, double ly,
                         double* point, double* ovec, double* dist) {
  wall_t wl;
  double half[3] = {lx / 2., ly / 2., 0.};
  build_wall(&wl, basis, origin, half, 0., 0, ORC_WALL_RECTANGLE);
  v3 pt = mk(0., 0., 0.), ov = mk(0., 0., 0.);
  *dist = 0.;
  int rc = sphere_rect(&wl, ld(centre), r, sphere_basis, &pt, &ov, dist);
  st(point, pt);
  st(ovec, ov);
  return rc;
}

int orc_sphere_box(const double* centre, double r, double cr,
                   const double* basis, const double* origin,
                   const double* half_extent, double crb, double* point,
                   double* ovec, double* dist) {
  orc_sys tmp;
  memset(&tmp, 0, sizeof(tmp));
  orc_add_wall(&tmp, basis, origin, half_extent, crb, 0);
  v3 pt = mk(0., 0., 0.), ov = mk(0., 0., 0.);
  int rc = sphere_wall(&tmp.walls[0], ld(centre), r, cr, &pt, &ov, dist);
  st(point, pt);
  st(ovec, ov);
  return rc;
}

/* Velocity at a point: ParticleKinematics::Velocity (ParticleKinematics.cpp:229-232)
   via Particle::getVelocityAtPoint (Particle.cpp:662-666) */
static inline v3 vel_at(v3 x, v3 v, v3 w, v3 pt) { return add(v, cross(w, sub(pt, x))); }

/* HookeContactForceModel::performForcesCalculus (HookeContactForceModel.cpp:44-114)
   and HertzContactForceModel::performForcesCalculus (HertzContactForceModel.cpp:46-114);
   returns delFN + delFT; *delM receives the rolling-resistance moment on body 0
   (HookeContactForceModel.cpp:95-111, HertzContactForceModel.cpp:99-113), zero
   when kr = 0. */
static v3 contact_force(const model_t* m, v3 pt, v3 pen, double odist, v3 x0, v3 v0,
                        v3 w0, double m0, double req0, v3 x1, v3 v1, v3 w1, double m1,
                        double req1, v3* delM) {
  v3 normal = dvd(pen, norm(pen));
  normal = round3(normal);
  v3 tmpV = sub(vel_at(x0, v0, w0, pt), vel_at(x1, v1, w1, pt));
  v3 v_n = scl(normal, dot(tmpV, normal));
  v3 v_t = sub(tmpV, v_n);
  v3 delFN, delFT;
  double normFN, muc;
  if (m->kind == ORC_MODEL_HOOKE) {
    double kn = m->p[0], etat_in = m->p[2];
    muc = m->p[3];
    delFN = scl(pen, kn);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double gamman = -2. * m->beta * sqrt(avmass * kn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double etat = etat_in == -1. ? -m->beta * sqrt(kn / avmass) : etat_in;
    double gammat = 2. * etat * avmass;
    delFT = scl(v_t, -gammat);
  } else {
    double Es = m->p[0], Gs = m->p[2];
    muc = m->p[3];
    double deltan = -odist;
    double Req = 1. / (1. / req0 + 1. / req1);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double sq = sqrt(Req * deltan);
    double Sn = 2. * Es * sq;
    double St = 8. * Gs * sq;
    double kn = (4. / 3.) * Es * sq;
    delFN = scl(pen, kn);
    double gamman = m->m2sqrt56 * m->beta * sqrt(avmass * Sn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double gammat = m->m2sqrt56 * m->beta * sqrt(avmass * St);
    delFT = scl(v_t, -gammat);
  }
  double fn = muc * normFN;
  double ft = norm(delFT);
  if (fn < ft) {
    double nvt = norm(v_t);
    v3 tangent = mk(0., 0., 0.);
    if (nvt > ORC_EPSILON) tangent = dvd(neg(v_t), nvt);
    delFT = scl(tangent, fn);
  }
  *delM = mk(0., 0., 0.);
  if ((m->kind == ORC_MODEL_HOOKE || m->kind == ORC_MODEL_HERTZ) && m->p[4] != 0.) {
    /* Vector3 wrel = w0 - w1; normwtrel = |w0 ^ (x0 - pt) - w1 ^ (x1 - pt)|;
       delM = -(kr Req normFN normwtrel / normwrel) wrel when |wrel| > EPSILON */
    v3 wrel = sub(w0, w1);
    double normwrel = norm(wrel);
    double normwtrel = norm(sub(cross(w0, sub(x0, pt)), cross(w1, sub(x1, pt))));
    if (normwrel > ORC_EPSILON) {
      double Req = 1. / (1. / req0 + 1. / req1);
      *delM = scl(wrel, -(m->p[4] * Req * normFN * normwtrel / normwrel));
    }
  }
  return add(delFN, delFT);
}

void orc_contact_force(int kind, const double* p, const double* point,
                       const double* ovec, double dist, const double* x0,
                       const double* v0, const double* w0, double m0,
                       double req0, const double* x1, const double* v1,
                       const double* w1, double m1, double req1, double* f,
                       double* moment) {
  model_t m;
  memset(&m, 0, sizeof(m));
  m.kind = kind;
  memcpy(m.p, p, sizeof(double) * 5);
  m.beta = log(p[1]) / sqrt(ORC_PI * ORC_PI + log(p[1]) * log(p[1]));
  m.m2sqrt56 = -2. * sqrt(5. / 6.);
  v3 dm;
  v3 r = contact_force(&m, ld(point), ld(ovec), dist, ld(x0), ld(v0), ld(w0), m0, req0,
                       ld(x1), ld(v1), ld(w1), m1, req1, &dm);
  st(f, r);
  if (moment) st(moment, dm);
}

/* ---------------------------------------------------------------------- */
/* Memory models */

/* Quaternion::setRotFromTwoVectors (Geometry/src/Quaternion.cpp:263-286):
   (vect, real_part) scaled by 1 / Norm(qq) (Quaternion.cpp:420-424, 558-562) */
static void rot_two(v3 u, v3 v, v3* qv, double* qw) {
  double norm_u_norm_v = sqrt(dot(u, u) * dot(v, v));
  double real_part = norm_u_norm_v + dot(u, v);
  v3 vect;
  if (real_part < 1.e-6 * norm_u_norm_v) {
    real_part = 0.;
    vect = fabs(u.x) > fabs(u.z) ? mk(-u.y, u.x, 0.) : mk(0., -u.z, u.y);
  } else {
    vect = cross(u, v);
  }
  double nq = sqrt(vect.x * vect.x + vect.y * vect.y + vect.z * vect.z + real_part * real_part);
  double d = 1. / nq;
  *qv = scl(vect, d);
  *qw = d * real_part;
}

/* Quaternion::rotateVector (Quaternion.cpp:293-298) */
static v3 rotate_vec(v3 qv, double qw, v3 v) {
  double a = qw * qw - norm(qv) * norm(qv);
  double b = 2. * dot(v, qv);
  double c = 2. * qw;
  return add(add(scl(v, a), scl(qv, b)), scl(cross(qv, v), c));
}

void orc_rotate_two_vectors(const double* u, const double* v, const double* x, double* out) {
  v3 qv;
  double qw;
  rot_two(ld(u), ld(v), &qv, &qw);
  st(out, rotate_vec(qv, qw, ld(x)));
}

/* computeTangentialVector (HookeMemoryContactForceModel.cpp:54-65) */
static v3 tangential_vector(double n_t, v3 ut, v3 kdelta, double epst) {
  if (norm(ut) > 0.001) return dvd(neg(ut), norm(ut));
  if ((norm(kdelta) > epst || norm(scl(ut, n_t)) > epst) && norm(add(kdelta, scl(ut, n_t))) > epst)
    return dvd(neg(add(kdelta, scl(ut, n_t))), norm(add(kdelta, scl(ut, n_t))));
  return mk(0., 0., 0.);
}

/* HookeMemoryContactForceModel::performForcesCalculus
   (HookeMemoryContactForceModel.cpp:72-193) and the Hertz one
   (HertzMemoryContactForceModel.cpp:73-195).  kd, nrm, spr are the first
   body's map entry (created as zeros when the contact did not exist); returns
   delFN + delFT, the rolling moment in delM. */
static v3 memory_force(const model_t* m, v3 pt, v3 pen, double odist, v3 x0, v3 v0, v3 w0,
                       double m0, double req0, v3 x1, v3 v1, v3 w1, double m1, double req1,
                       double dt, int existed, double* kd, double* nrm, double* spr, v3* delM) {
  int hooke = m->kind == ORC_MODEL_HOOKE_MEMORY;
  v3 normal = dvd(pen, norm(pen));
  normal = round3(normal);
  v3 tmpV = sub(vel_at(x0, v0, w0, pt), vel_at(x1, v1, w1, pt));
  v3 v_n = scl(normal, dot(tmpV, normal));
  v3 v_t = sub(tmpV, v_n);
  double muc = m->p[3], mur, etarpf, epst, kn, kt, gamman, gammat, normFN;
  v3 delFN;
  if (hooke) {
    kn = m->p[0];
    kt = m->p[4];
    mur = m->p[5];
    etarpf = m->p[6];
    epst = m->p[7];
    delFN = scl(pen, kn);
    double avmass = 1. / (1. / m0 + 1. / m1);
    gamman = -2. * m->beta * sqrt(avmass * kn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double etat = m->p[2] == -1. ? -m->beta * sqrt(kn / avmass) : m->p[2];
    gammat = 2. * etat * avmass;
  } else {
    double Es = m->p[0], Gs = m->p[2];
    mur = m->p[4];
    etarpf = m->p[5];
    epst = m->p[6];
    double deltan = -odist;
    double Req = 1. / (1. / req0 + 1. / req1);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double sq = sqrt(Req * deltan);
    double Sn = 2. * Es * sq;
    double St = 8. * Gs * sq;
    (void)St;
    kn = (4. / 3.) * Es * sq;
    delFN = scl(pen, kn);
    gamman = m->m2sqrt56 * m->beta * sqrt(avmass * Sn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    gammat = m->m2sqrt56 * m->beta * sqrt(avmass * St);
    kt = 8 * Gs * sq;
  }
  v3 viscousFT = scl(v_t, -gammat);
  v3 K = ld(kd), qv = mk(0., 0., 0.);
  double qw = 0.;
  if (existed) {
    rot_two(ld(nrm), normal, &qv, &qw);
    K = rotate_vec(qv, qw, K);
  }
  K = add(K, scl(v_t, dt * kt));
  st(nrm, normal);
  v3 tij = tangential_vector(gammat, v_t, K, epst);
  double normFT = norm(add(neg(K), viscousFT));
  v3 delFT;
  if (normFT <= muc * normFN) {
    delFT = scl(tij, norm(add(neg(K), viscousFT)));
  } else {
    if (!hooke || kt != 0.) K = dvd(add(scl(tij, -muc * normFN), viscousFT), kt);
    delFT = scl(tij, muc * normFN);
  }
  st(kd, K);
  if (mur != 0.) {
    double Req = 1. / (1. / req0 + 1. / req1);
    double kr = 3. * kn * mur * mur * Req * Req;
    double etar = 3. * gamman * mur * mur * Req * Req;
    double max_normMk = mur * Req * normFN;
    v3 wrel = sub(w0, w1);
    v3 S = ld(spr);
    if (existed) S = rotate_vec(qv, qw, S);
    S = sub(S, scl(wrel, kr * dt));
    double normMk = norm(S);
    if (normMk > max_normMk) S = scl(S, max_normMk / normMk);
    st(spr, S);
    *delM = sub(S, scl(wrel, etarpf * etar));
  } else {
    *delM = mk(0., 0., 0.);
  }
  return add(delFN, delFT);
}

static int rolling_model(const model_t* m) {
  if (m->kind == ORC_MODEL_HOOKE || m->kind == ORC_MODEL_HERTZ) return m->p[4] != 0.;
  return m->p[m->kind == ORC_MODEL_HOOKE_MEMORY ? 5 : 4] != 0.;
}

/* Component::getContactMemory (Component/src/Component.cpp:634-674) with
   createContact = true: the entry of key in particle p's map (activated),
   created as zeros when missing; *existed tells which. */
static mem_t* get_memory(orc_sys* s, int p, int key, int* existed) {
  mem_t* tab = s->mem + (size_t)p * ORC_MEM_CAP;
  for (int k = 0; k < s->mem_n[p]; ++k)
    if (tab[k].key == key) {
      tab[k].active = 1;
      *existed = 1;
      return tab + k;
    }
  *existed = 0;
  if (s->mem_n[p] >= ORC_MEM_CAP) return NULL;
  mem_t* e = tab + s->mem_n[p]++;
  memset(e, 0, sizeof(*e));
  e->key = key;
  e->active = 1;
  return e;
}

/* Component::addDeplContactInMap (Component.cpp:680-689): sets (creating if
   needed) particle p's entry of key, active. */
static int set_memory(orc_sys* s, int p, int key, v3 kd, v3 nrm, v3 spr) {
  int existed;
  mem_t* e = get_memory(s, p, key, &existed);
  if (!e) return -1;
  st(e->kd, kd);
  st(e->nrm, nrm);
  st(e->spr, spr);
  return 0;
}

int orc_get_memory(const orc_sys* s, int p, int cap, int32_t* key, double* kd, double* nrm,
                   double* spr) {
  if (p < 0 || p >= s->n) return -1;
  int n = s->mem_n[p];
  for (int k = 0; k < n && k < cap; ++k) {
    const mem_t* e = s->mem + (size_t)p * ORC_MEM_CAP + k;
    key[k] = e->key;
    memcpy(kd + 3 * k, e->kd, sizeof(e->kd));
    memcpy(nrm + 3 * k, e->nrm, sizeof(e->nrm));
    memcpy(spr + 3 * k, e->spr, sizeof(e->spr));
  }
  return n;
}

static void log_contact(orc_sys* s, int id0, int id1, v3 pt, v3 ov, double d, v3 f) {
  if (s->log_n < s->log_cap) {
    int k = s->log_n;
    s->log_id0[k] = id0;
    s->log_id1[k] = id1;
    st(s->log_pt + 3 * k, pt);
    st(s->log_ov + 3 * k, ov);
    s->log_d[k] = d;
    st(s->log_f + 3 * k, f);
  }
  s->log_n++;
}

/* Particle::InterAction (Component/src/Particle.cpp:478-546) for two spheres,
   HookeContactForceModel::computeForces (HookeContactForceModel.cpp:121-163)
   and Torsor::addForce(point, f) (Torsor.cpp:184-189) on both bodies.
   shift is the periodic vector applied to the second body (minimum image). */
/* AppCollision::addToContactsFeatures (Base/src/AppCollision.cpp:160-190):
   running mean and maximum of the overlaps, running mean of the reported GJK
   iterations, over every closest-point result of the run. */
static void add_contact_features(orc_sys* s, double overlap, int nbIter) {
  if (overlap < 0.) {
    overlap = fabs(overlap);
    s->ov_mean = (s->ov_mean * s->ov_count + overlap) / (s->ov_count + 1.);
    s->ov_count += 1.;
    if (overlap > s->ov_max) {
      s->ov_max = overlap;
      s->ov_time_max = s->time;
    }
  }
  if (nbIter) {
    s->gjk_mean = (s->gjk_mean * s->gjk_count + (double)nbIter) / (s->gjk_count + 1.);
    s->gjk_count += 1.;
  }
}

/* RigidBodyWithCrust::ClosestPoint_ErreurHandling (Geometry/src/
   RigidBodyWithCrust.cpp:534-588), the retry of a pair whose overlap reaches
   the crust sum (Particle.cpp:508-525): both bodies shrunk by ten crusts
   (getTransformWithCrust(10, 0.5), :770-795), closest points, contact point
   midway, overlap imposed to the crust sum along B -> A; a second-level
   ContactError when the shrunk bodies still touch.  The sphere's rotation is
   not used (the GPU step keeps none; it cancels for a sphere up to
   rounding). */
static double retry_scale(double extent, double crust) {
  double sc = extent < ORC_EPSILON ? 1. : (extent - 10. * crust) / extent;
  return sc < 0.5 ? 0.5 : sc;
}

static int retry_contact(v3 pA, v3 pB, double gjk_dist, double crA, double crB, double* pt,
                         double* ov, double* dist) {
  if (gjk_dist < ORC_EPSILON) return -1;
  v3 ba = sub(pA, pB);
  ba = dvd(ba, norm(ba));
  v3 c = add(dvd(pA, 2.0), dvd(pB, 2.0));
  double imposed = 1. * (crA + crB);
  st(pt, c);
  st(ov, scl(ba, imposed));
  *dist = -imposed;
  return 1;
}

/* Two spheres: Johnson GJK (Convex.cpp:536-627) on the shrunk spheres starts
   from v = A - B and takes the support points along -v and +v.  Apart, their
   difference w points along v and GJK stops at its second iteration with that
   one-point simplex (the new supports repeat the first ones to rounding): its
   witness points are the first iteration's, computed here with the same
   operations; two iterations.  Touching (w not along v), the second
   iteration's supports lie on the far sides, the segment of the two simplex
   points holds the origin and GJK ends at distance 0: the second-level
   ContactError. */
int orc_sphere_sphere_retry(const double* a, double ra, double cra, const double* b, double rb,
                            double crb, double* pt, double* ov, double* dist, int* nbIter) {
  double sa = retry_scale(ra, cra), sb = retry_scale(rb, crb);
  v3 A = ld(a), B = ld(b);
  v3 v = sub(A, B);  /* a2w(support(0)) - b2w(support(0)): the centres */
  /* Sphere::support of -v seen through the scaled basis, back to world */
  v3 ua = mk(sa * -v.x + 0. * -v.y + 0. * -v.z, 0. * -v.x + sa * -v.y + 0. * -v.z,
             0. * -v.x + 0. * -v.y + sa * -v.z);
  v3 ub = mk(sb * v.x + 0. * v.y + 0. * v.z, 0. * v.x + sb * v.y + 0. * v.z,
             0. * v.x + 0. * v.y + sb * v.z);
  double na = norm(ua), nb = norm(ub);
  v3 qa = na > ORC_EPSILON ? scl(ua, ra / na) : mk(0., 0., 0.);
  v3 qb = nb > ORC_EPSILON ? scl(ub, rb / nb) : mk(0., 0., 0.);
  v3 pA = mk(sa * qa.x + 0. * qa.y + 0. * qa.z + A.x, 0. * qa.x + sa * qa.y + 0. * qa.z + A.y,
             0. * qa.x + 0. * qa.y + sa * qa.z + A.z);
  v3 pB = mk(sb * qb.x + 0. * qb.y + 0. * qb.z + B.x, 0. * qb.x + sb * qb.y + 0. * qb.z + B.y,
             0. * qb.x + 0. * qb.y + sb * qb.z + B.z);
  *nbIter = 2;
  v3 w = sub(pA, pB);
  if (!(dot(w, v) > 0.)) return -1;
  return retry_contact(pA, pB, norm(w), cra, crb, pt, ov, dist);
}

/* A sphere against a box wall: the exact closest points of the shrunk box
   (the point nearest to the centre) and of the shrunk sphere, the limit GJK
   converges to (the reference stops at its 1e-10 optimality gap).  Reports
   no GJK iterations. */
static int sphere_wall_retry(const wall_t* wl, v3 A, double r, double cr, double* pt, double* ov,
                             double* dist, int* nbIter) {
  *nbIter = 0;
  v3 l = xform(wl->inv_basis, wl->inv_origin, A);
  double q[3] = {l.x, l.y, l.z};
  int inside = 1;
  for (int d = 0; d < 3; ++d) {
    double e = wl->ext[d] * retry_scale(wl->ext[d], wl->crust);
    if (q[d] > e) { q[d] = e; inside = 0; }
    else if (q[d] < -e) { q[d] = -e; inside = 0; }
  }
  if (inside) return -1;
  v3 pB = xform(wl->basis, wl->origin, mk(q[0], q[1], q[2]));
  v3 u = sub(pB, A);
  double nu = norm(u);
  double ra = r * retry_scale(r, cr);
  v3 pA = add(A, scl(dvd(u, nu), ra));
  return retry_contact(pA, pB, nu - ra, cr, wl->crust, pt, ov, dist);
}

static int interact(orc_sys* s, int p0, int p1, v3 shift) {
  v3 A = ld(s->x + 3 * p0);
  v3 B = add(ld(s->x + 3 * p1), shift);
  double a[3], b[3], pt[3], ov[3], dist;
  st(a, A);
  st(b, B);
  s->ncand++;
  int rc = orc_sphere_sphere(a, s->r[p0], s->crust[p0], b, s->r[p1], s->crust[p1], pt,
                             ov, &dist);
  int nbIter = 1;  /* ClosestPointSPHERESPHERE reports one iteration */
  double over = -dist;
  if (rc < 0)
    rc = orc_sphere_sphere_retry(a, s->r[p0], s->crust[p0], b, s->r[p1], s->crust[p1], pt, ov,
                                 &dist, &nbIter);
  if (rc < 0)
    return set_error(s, ORC_ERR_CONTACT, p0, p1, over, s->crust[p0] + s->crust[p1]);
  if (rc == 0) return ORC_OK;
  add_contact_features(s, dist, nbIter);
  if (!(dist < 0.)) return ORC_OK;
  const model_t* m = &s->models[s->mat[p0]][s->mat[p1]];
  if (m->kind == ORC_MODEL_NONE) return set_error(s, ORC_ERR_ARG, p0, p1, 0., 0.);
  v3 P = ld(pt);
  v3 f, delM = mk(0., 0., 0.);
  int rolling = 0;
  if (m->kind >= ORC_MODEL_HOOKE_MEMORY) {
    /* HookeMemoryContactForceModel::computeForces (:200-249): body 0's map
       entry (created if missing), then body 1's gets the negation */
    int existed;
    mem_t* e = get_memory(s, p0, p1, &existed);
    if (!e) return set_error(s, ORC_ERR_ARG, p0, p1, 0., 0.);
    f = memory_force(m, P, ld(ov), dist, A, ld(s->v + 3 * p0), ld(s->w + 3 * p0), s->mass[p0],
                     s->req[p0], B, ld(s->v + 3 * p1), ld(s->w + 3 * p1), s->mass[p1],
                     s->req[p1], s->dt_mem, existed, e->kd, e->nrm, e->spr, &delM);
    if (set_memory(s, p1, p0, neg(ld(e->kd)), neg(ld(e->nrm)), neg(ld(e->spr))))
      return set_error(s, ORC_ERR_ARG, p1, p0, 0., 0.);
    rolling = rolling_model(m);
  } else {
    f = contact_force(m, P, ld(ov), dist, A, ld(s->v + 3 * p0), ld(s->w + 3 * p0),
                      s->mass[p0], s->req[p0], B, ld(s->v + 3 * p1), ld(s->w + 3 * p1),
                      s->mass[p1], s->req[p1], &delM);
    rolling = rolling_model(m);
  }
  /* coef = 1: body 0 gets coef*(delFN+delFT); body 1 gets coef*(-delFN-delFT),
     which is -(delFN+delFT) bit for bit */
  v3 f0 = scl(f, 1.);
  v3 f1 = neg(f);
  st(s->F + 3 * p0, add(ld(s->F + 3 * p0), f0));
  st(s->T + 3 * p0, add(ld(s->T + 3 * p0), cross(sub(P, A), f0)));
  if (rolling) st(s->T + 3 * p0, add(ld(s->T + 3 * p0), scl(delM, 1.)));
  st(s->F + 3 * p1, add(ld(s->F + 3 * p1), f1));
  st(s->T + 3 * p1, add(ld(s->T + 3 * p1), cross(sub(P, B), f1)));
  if (rolling) st(s->T + 3 * p1, add(ld(s->T + 3 * p1), scl(neg(delM), 1.)));
  log_contact(s, p0, p1, P, ld(ov), dist, f0);
  return ORC_OK;
}

/* SimpleObstacle::InterAction (Component/src/SimpleObstacle.cpp:222-271):
   BBox::InZone pre-test (BBox.cpp:166-171), sphere-box closest point, force
   with the particle as first body and the obstacle (mass 1e20,
   Component.cpp:16, velocity at a point vel + omega x (point - centre),
   Obstacle.cpp:128-133) as second body. */
static int interact_wall(orc_sys* s, int p, int wi) {
  const wall_t* wl = &s->walls[wi];
  const double* xp = s->x + 3 * p;
  double r = s->r[p];
  if (!(fabs(wl->bb_centre[0] - xp[0]) <= wl->bb_ext[0] + r &&
        fabs(wl->bb_centre[1] - xp[1]) <= wl->bb_ext[1] + r &&
        fabs(wl->bb_centre[2] - xp[2]) <= wl->bb_ext[2] + r))
    return ORC_OK;
  v3 A = ld(xp), pt = mk(0., 0., 0.), ov = mk(0., 0., 0.);
  double dist;
  int rc = wl->kind == ORC_WALL_RECTANGLE
               ? sphere_rect(wl, A, r, s->basis + 9 * p, &pt, &ov, &dist)
               : sphere_wall(wl, A, r, s->crust[p], &pt, &ov, &dist);
  int nbIter = wl->kind == ORC_WALL_RECTANGLE ? 0 : 1;  /* ClosestPointRECTANGLE reports none */
  double over = -dist;
  if (rc < 0) {
    double p3[3], o3[3];
    rc = sphere_wall_retry(wl, A, r, s->crust[p], p3, o3, &dist, &nbIter);
    pt = ld(p3);
    ov = ld(o3);
  }
  if (rc < 0) return set_error(s, ORC_ERR_CONTACT, p, -1 - wi, over, s->crust[p] + wl->crust);
  if (rc == 0 || !(dist < 0.)) return ORC_OK;
  add_contact_features(s, dist, nbIter);
  const model_t* m = &s->models[s->mat[p]][wl->material];
  if (m->kind == ORC_MODEL_NONE) return set_error(s, ORC_ERR_ARG, p, -1 - wi, 0., 0.);
  v3 wv = ld(wl->vel), ww = ld(wl->omega);
  v3 f, delM = mk(0., 0., 0.);
  int rolling = 0;
  if (m->kind >= ORC_MODEL_HOOKE_MEMORY) {
    /* the particle is body 0 (SimpleObstacle.cpp:262-264); its map keys the
       wall by -1 - index (the obstacle's own copy is never read back) */
    int existed;
    mem_t* e = get_memory(s, p, -1 - wi, &existed);
    if (!e) return set_error(s, ORC_ERR_ARG, p, -1 - wi, 0., 0.);
    f = memory_force(m, pt, ov, dist, A, ld(s->v + 3 * p), ld(s->w + 3 * p), s->mass[p],
                     s->req[p], ld(wl->origin), wv, ww, 1.e20, wl->req, s->dt_mem, existed,
                     e->kd, e->nrm, e->spr, &delM);
    rolling = rolling_model(m);
  } else {
    f = contact_force(m, pt, ov, dist, A, ld(s->v + 3 * p), ld(s->w + 3 * p), s->mass[p],
                      s->req[p], ld(wl->origin), wv, ww, 1.e20, wl->req, &delM);
    rolling = rolling_model(m);
  }
  st(s->F + 3 * p, add(ld(s->F + 3 * p), f));
  st(s->T + 3 * p, add(ld(s->T + 3 * p), cross(sub(pt, A), f)));
  if (rolling) st(s->T + 3 * p, add(ld(s->T + 3 * p), scl(delM, 1.)));
  /* the reaction on the wall, the second body (HookeContactForceModel.cpp:
     150-153, Obstacle::addForce / addTorque, Obstacle.cpp:575-597, Torsor::
     addForce, Torsor.cpp:184-189) */
  {
    wall_t* ww_ = &s->walls[wi];
    v3 fw = neg(f);
    st(ww_->load_f, add(ld(ww_->load_f), fw));
    st(ww_->load_t, add(ld(ww_->load_t), cross(sub(pt, ld(ww_->origin)), fw)));
    if (rolling) st(ww_->load_t, add(ld(ww_->load_t), neg(delM)));
  }
  log_contact(s, p, -1 - wi, pt, ov, dist, f);
  return ORC_OK;
}

/* Component::setContactMapToFalse / updateContactMap (Component.cpp:567-596),
   over all particles (AllComponents.cpp:1777-1800) */
static void memory_maps_inactive(orc_sys* s) {
  for (int p = 0; p < s->n; ++p)
    for (int k = 0; k < s->mem_n[p]; ++k) s->mem[(size_t)p * ORC_MEM_CAP + k].active = 0;
}

static void memory_maps_prune(orc_sys* s) {
  for (int p = 0; p < s->n; ++p) {
    mem_t* tab = s->mem + (size_t)p * ORC_MEM_CAP;
    int w = 0;
    for (int k = 0; k < s->mem_n[p]; ++k)
      if (tab[k].active) tab[w++] = tab[k];
    s->mem_n[p] = w;
  }
}

static int compute_forces_body(orc_sys* s);

/* Grains::computeParticlesForceAndTorque (Base/src/Grains.cpp:459-482): the
   contact maps are marked inactive, forces computed, inactive entries erased. */
int orc_compute_forces(orc_sys* s) {
  if (s->has_memory && !(s->dt_mem > 0.)) return set_error(s, ORC_ERR_ARG, -1, -1, 0., 0.);
  memory_maps_inactive(s);
  int rc = compute_forces_body(s);
  if (rc) return rc;
  memory_maps_prune(s);
  return ORC_OK;
}

/* Inside Grains::computeParticlesForceAndTorque (Base/src/Grains.cpp:459-482):
   AllComponents::InitializeForces (AllComponents.cpp:259-270) sets each torsor
   to the weight m (1 - rho_f / rho) g (Particle.cpp:414-419, rho_f = 0) at the
   centre; then LinkedCell::ComputeForces (LinkedCell.cpp:1019-1103): for each
   particle, the particles after it in its own cell list, then the particles of
   its contact-neighbour cells in list order (the 13 cells were added with
   push_front, Cell.cpp:100-103, so they are visited in reverse stencil order),
   then the obstacles. */
static int compute_forces_body(orc_sys* s) {
  const double* g = s->cfg.gravity;
  s->log_n = 0;
  s->ncand = 0;
  /* obstacle torsors at their centres, zero (Component::InitializeForce,
     Component.cpp:460-463, from AllComponents::InitializeForces :268-269) */
  for (int w = 0; w < s->nwalls; ++w)
    for (int d = 0; d < 3; ++d) s->walls[w].load_f[d] = s->walls[w].load_t[d] = 0.;
  for (int p = 0; p < s->n; ++p) {
    double m = s->mass[p] * (1. - 0. / 1.);
    s->F[3 * p] = g[0] * m;
    s->F[3 * p + 1] = g[1] * m;
    s->F[3 * p + 2] = g[2] * m;
    s->T[3 * p] = s->T[3 * p + 1] = s->T[3 * p + 2] = 0.;
  }
  const int ni = s->nc[0];
  for (int p = 0; p < s->n; ++p) {
    int ijk[3];
    int c = get_cell(s, s->x + 3 * p, ijk);
    if (c < 0) return set_error(s, ORC_ERR_SIMULATION, p, -1, 0., 0.);
    v3 zero = mk(0., 0., 0.);
    for (int q = s->next[p]; q >= 0; q = s->next[q]) {
      int rc = interact(s, p, q, zero);
      if (rc) return rc;
    }
    for (int jj = 12; jj >= 0; --jj) {
      int t[3];
      double sh[3] = {0., 0., 0.};
      int ok = 1;
      for (int d = 0; d < 3; ++d) {
        t[d] = ijk[d] + s->stencil[jj][d];
        int n = s->nc[d];
        if (t[d] < 0 || t[d] >= n) {
          if (!s->cfg.periodic[d]) { ok = 0; break; }
          if (t[d] < 0) { t[d] += n; sh[d] = -s->cfg.size[d]; }
          else { t[d] -= n; sh[d] = s->cfg.size[d]; }
        }
      }
      if (!ok) continue;
      int cn = t[1] * s->nc[2] * ni + t[2] * ni + t[0];
      for (int q = s->head[cn]; q >= 0; q = s->next[q]) {
        int rc = interact(s, p, q, mk(sh[0], sh[1], sh[2]));
        if (rc) return rc;
      }
    }
    for (int wi = 0; wi < s->nwalls; ++wi) {
      int rc = interact_wall(s, p, wi);
      if (rc) return rc;
    }
  }
  return ORC_OK;
}

/* ParticleKinematicsSphere::computeAngularAccelerationBodyFixed
   (Component/src/ParticleKinematicsSphere.cpp:47-55), coupling factor 1 */
static inline v3 ang_acc(v3 tq, double I) {
  return mk(tq.x / (1. * I), tq.y / (1. * I), tq.z / (1. * I));
}

/* SSPRK3 angular update of SecondOrderLeapFrog (TimeIntegration/src/SecondOrderLeapFrog.cpp:56-66) */
static inline v3 ssprk3(v3 tq, v3 wbf, double I, double dt) {
  /* a sphere's angular acceleration does not depend on w, so K1 = K2 = K3 */
  v3 K1 = scl(ang_acc(tq, I), dt);
  v3 K2 = scl(ang_acc(tq, I), dt);
  v3 K3 = scl(ang_acc(tq, I), dt);
  return add(wbf, dvd(add(add(K1, K2), scl(K3, 4.)), 6.));
}

/* ParticleKinematics::Move (Component/src/ParticleKinematics.cpp:62-104) with
   SecondOrderLeapFrog::Move (SecondOrderLeapFrog.cpp:44-67) and
   Transform::composeLeftByRotation (Geometry/src/Transform.cpp:140-156).
   Returns |dx| (Particle::Move compares it with the crust, Particle.cpp:437-460). */
static double move_one(orc_sys* s, int p, double dtv, double dtd) {
  quat q = qmk(mk(s->q[4 * p], s->q[4 * p + 1], s->q[4 * p + 2]), s->q[4 * p + 3]);
  quat pc = qconj(q);
  v3 tq = qmulv(pc, qleftvec(q, ld(s->T + 3 * p)));
  double mcf = s->mass[p] * 1.;
  v3 v = add(ld(s->v + 3 * p), scl(dvd(ld(s->F + 3 * p), mcf), dtv));
  v3 dx = scl(v, dtd);
  v3 wbf = ssprk3(tq, ld(s->wbf + 3 * p), s->inertia[p], dtv);
  v3 w = qmulv(q, qleftvec(pc, wbf));
  v3 avg = qmulv(q, qleftvec(pc, wbf));
  st(s->v + 3 * p, v);
  st(s->wbf + 3 * p, wbf);
  st(s->w + 3 * p, w);
  /* Component::Translate (Component.cpp:376-379) */
  st(s->x + 3 * p, add(ld(s->x + 3 * p), dx));
  double nOmega = norm(avg);
  quat dq;
  if (nOmega > ORC_EPSILON) {
    double c = cos(nOmega * dtd / 2.);
    double sn = sin(nOmega * dtd / 2.);
    v3 t = scl(avg, sn * 1. / nOmega);
    dq = qmk(t, c);
  } else {
    dq = qmk(mk(0., 0., 0.), 1.);
  }
  quat qn = qmul(dq, q);
  s->q[4 * p] = qn.v.x; s->q[4 * p + 1] = qn.v.y; s->q[4 * p + 2] = qn.v.z; s->q[4 * p + 3] = qn.w;
  compose_left_rotation(s->basis + 9 * p, dq);
  return norm(dx);
}

/* AllComponents::Move (Component/src/AllComponents.cpp:151-166); then the
   periodic wrap that stands in for the clone update of
   LinkedCell::updateDestroyPeriodicClones (LinkedCell.cpp:2008-2126). */
int orc_move(orc_sys* s, double dtv, double dtd) {
  for (int p = 0; p < s->n; ++p) {
    double depl = move_one(s, p, dtv, dtd);
    if (depl > s->crust[p]) return set_error(s, ORC_ERR_MOTION, p, -1, depl, s->crust[p]);
    for (int d = 0; d < 3; ++d) {
      if (!s->cfg.periodic[d]) continue;
      double lo = s->cfg.origin[d], L = s->cfg.size[d];
      double* xd = s->x + 3 * p + d;
      if (*xd >= lo + L) *xd -= L;
      else if (*xd < lo) *xd += L;
    }
  }
  return ORC_OK;
}

/* ParticleKinematics::advanceVelocity (ParticleKinematics.cpp:111-128) with
   SecondOrderLeapFrog::advanceVelocity (SecondOrderLeapFrog.cpp:74-94) */
int orc_advance_velocity(orc_sys* s, double dtv) {
  for (int p = 0; p < s->n; ++p) {
    quat q = qmk(mk(s->q[4 * p], s->q[4 * p + 1], s->q[4 * p + 2]), s->q[4 * p + 3]);
    quat pc = qconj(q);
    v3 tq = qmulv(pc, qleftvec(q, ld(s->T + 3 * p)));
    double mcf = s->mass[p] * 1.;
    v3 v = add(ld(s->v + 3 * p), scl(dvd(ld(s->F + 3 * p), mcf), dtv));
    v3 wbf = ssprk3(tq, ld(s->wbf + 3 * p), s->inertia[p], dtv);
    st(s->v + 3 * p, v);
    st(s->wbf + 3 * p, wbf);
    st(s->w + 3 * p, qmulv(q, qleftvec(pc, wbf)));
  }
  return ORC_OK;
}

/* Grains::do_before_time_stepping (Base/src/Grains.cpp:95-179): particles are
   linked at insertion, then a0 is computed for the leap-frog (:169-174). */
int orc_prime(orc_sys* s) {
  s->err_code = ORC_OK;
  /* initial obstacle kinematics (Grains.cpp:123-124) */
  if (s->dt_mem > 0.) init_wall_kinematics(s, s->time, s->dt_mem);
  int rc = orc_link(s);
  if (rc) return rc;
  return orc_compute_forces(s);
}

/* Grains::Simulation kick-drift-kick body (Base/src/Grains.cpp:290, 344-364):
   time += dt, particles then walls move (AllComponents::Move), forces, kick */
int orc_step(orc_sys* s, int nsteps, double dt) {
  s->dt_mem = dt;
  for (int k = 0; k < nsteps; ++k) {
    s->time += dt;
    int rc = orc_move(s, 0.5 * dt, dt);
    if (rc) return rc;
    rc = orc_move_walls(s, s->time, dt);
    if (rc) return rc;
    rc = orc_link_update(s);
    if (rc) return rc;
    rc = orc_compute_forces(s);
    if (rc) return rc;
    rc = orc_advance_velocity(s, 0.5 * dt);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* All-core variant of orc_step: the timed CPU baseline on every host core.
   It stands in for GrainsMPI's domain decomposition (Base/src/GrainsMPI.cpp)
   on shared memory.  The cell rows along j are split into an even number of
   slabs; slabs of one colour (even, then odd) are swept concurrently, one
   thread each.  The half stencil of a row reaches only rows j and j+1
   (LinkedCell.cpp:965-967), so a slab writes torsors of its own rows and of
   the next slab's first row, which no slab of the same colour touches.  Each
   pair is evaluated once, with the serial step's body order and arithmetic;
   only the order in which a particle's contacts are summed differs.
   Memoryless models, static walls, no contact log, no wall loads. */

/* interact() without the candidate counter, contact log and memory maps */
static int interact_plain(orc_sys* s, int p0, int p1, v3 shift) {
  v3 A = ld(s->x + 3 * p0);
  v3 B = add(ld(s->x + 3 * p1), shift);
  double a[3], b[3], pt[3], ov[3], dist;
  st(a, A);
  st(b, B);
  int rc = orc_sphere_sphere(a, s->r[p0], s->crust[p0], b, s->r[p1], s->crust[p1], pt,
                             ov, &dist);
  int nbIter;
  if (rc < 0)
    rc = orc_sphere_sphere_retry(a, s->r[p0], s->crust[p0], b, s->r[p1], s->crust[p1], pt, ov,
                                 &dist, &nbIter);
  if (rc < 0) return ORC_ERR_CONTACT;
  if (rc == 0 || !(dist < 0.)) return ORC_OK;
  const model_t* m = &s->models[s->mat[p0]][s->mat[p1]];
  if (m->kind == ORC_MODEL_NONE) return ORC_ERR_ARG;
  v3 P = ld(pt);
  v3 delM;
  v3 f = contact_force(m, P, ld(ov), dist, A, ld(s->v + 3 * p0), ld(s->w + 3 * p0),
                       s->mass[p0], s->req[p0], B, ld(s->v + 3 * p1), ld(s->w + 3 * p1),
                       s->mass[p1], s->req[p1], &delM);
  const int rolling = rolling_model(m);
  v3 f0 = scl(f, 1.);
  v3 f1 = neg(f);
  st(s->F + 3 * p0, add(ld(s->F + 3 * p0), f0));
  st(s->T + 3 * p0, add(ld(s->T + 3 * p0), cross(sub(P, A), f0)));
  if (rolling) st(s->T + 3 * p0, add(ld(s->T + 3 * p0), scl(delM, 1.)));
  st(s->F + 3 * p1, add(ld(s->F + 3 * p1), f1));
  st(s->T + 3 * p1, add(ld(s->T + 3 * p1), cross(sub(P, B), f1)));
  if (rolling) st(s->T + 3 * p1, add(ld(s->T + 3 * p1), scl(neg(delM), 1.)));
  return ORC_OK;
}

/* interact_wall() without the wall's torsor */
static int wall_plain(orc_sys* s, int p, int wi) {
  const wall_t* wl = &s->walls[wi];
  const double* xp = s->x + 3 * p;
  double r = s->r[p];
  if (!(fabs(wl->bb_centre[0] - xp[0]) <= wl->bb_ext[0] + r &&
        fabs(wl->bb_centre[1] - xp[1]) <= wl->bb_ext[1] + r &&
        fabs(wl->bb_centre[2] - xp[2]) <= wl->bb_ext[2] + r))
    return ORC_OK;
  v3 A = ld(xp), pt = mk(0., 0., 0.), ov = mk(0., 0., 0.);
  double dist;
  int rc = wl->kind == ORC_WALL_RECTANGLE
               ? sphere_rect(wl, A, r, s->basis + 9 * p, &pt, &ov, &dist)
               : sphere_wall(wl, A, r, s->crust[p], &pt, &ov, &dist);
  if (rc < 0) {
    double p3[3], o3[3];
    int nbIter;
    rc = sphere_wall_retry(wl, A, r, s->crust[p], p3, o3, &dist, &nbIter);
    pt = ld(p3);
    ov = ld(o3);
  }
  if (rc < 0) return ORC_ERR_CONTACT;
  if (rc == 0 || !(dist < 0.)) return ORC_OK;
  const model_t* m = &s->models[s->mat[p]][wl->material];
  if (m->kind == ORC_MODEL_NONE) return ORC_ERR_ARG;
  v3 delM;
  v3 f = contact_force(m, pt, ov, dist, A, ld(s->v + 3 * p), ld(s->w + 3 * p), s->mass[p],
                       s->req[p], ld(wl->origin), ld(wl->vel), ld(wl->omega), 1.e20, wl->req,
                       &delM);
  st(s->F + 3 * p, add(ld(s->F + 3 * p), f));
  st(s->T + 3 * p, add(ld(s->T + 3 * p), cross(sub(pt, A), f)));
  if (rolling_model(m)) st(s->T + 3 * p, add(ld(s->T + 3 * p), scl(delM, 1.)));
  return ORC_OK;
}

/* The pairs particle p detects and its walls (compute_forces_body's body). */
static int sweep_particle(orc_sys* s, int p) {
  const int ni = s->nc[0];
  int ijk[3];
  if (get_cell(s, s->x + 3 * p, ijk) < 0) return ORC_ERR_SIMULATION;
  v3 zero = mk(0., 0., 0.);
  for (int q = s->next[p]; q >= 0; q = s->next[q]) {
    int rc = interact_plain(s, p, q, zero);
    if (rc) return rc;
  }
  for (int jj = 12; jj >= 0; --jj) {
    int t[3];
    double sh[3] = {0., 0., 0.};
    int ok = 1;
    for (int d = 0; d < 3; ++d) {
      t[d] = ijk[d] + s->stencil[jj][d];
      int n = s->nc[d];
      if (t[d] < 0 || t[d] >= n) {
        if (!s->cfg.periodic[d]) { ok = 0; break; }
        if (t[d] < 0) { t[d] += n; sh[d] = -s->cfg.size[d]; }
        else { t[d] -= n; sh[d] = s->cfg.size[d]; }
      }
    }
    if (!ok) continue;
    int cn = t[1] * s->nc[2] * ni + t[2] * ni + t[0];
    for (int q = s->head[cn]; q >= 0; q = s->next[q]) {
      int rc = interact_plain(s, p, q, mk(sh[0], sh[1], sh[2]));
      if (rc) return rc;
    }
  }
  for (int wi = 0; wi < s->nwalls; ++wi) {
    int rc = wall_plain(s, p, wi);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* rows [j0, j1) of the cell grid: every particle of those cells, cell by cell */
static int sweep_rows(orc_sys* s, int j0, int j1) {
  const int ni = s->nc[0], nk = s->nc[2];
  for (int j = j0; j < j1; ++j)
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < ni; ++i)
        for (int p = s->head[j * nk * ni + k * ni + i]; p >= 0; p = s->next[p]) {
          int rc = sweep_particle(s, p);
          if (rc) return rc;
        }
  return ORC_OK;
}

int orc_step_parallel(orc_sys* s, int nsteps, double dt, int nthreads) {
  if (s->has_memory || any_live(s) || nthreads < 1) return ORC_ERR_ARG;
  const int nj = s->nc[1];
  int nslab = 2 * nthreads;
  if (nslab > nj) nslab = nj - (nj & 1);
  if (nslab < 2) nslab = 0; /* one row: serial sweep */
  s->dt_mem = dt;
  for (int k = 0; k < nsteps; ++k) {
    s->time += dt;
    int err = ORC_OK, err_p = -1;
    double err_v = 0.;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int p = 0; p < s->n; ++p) {
      double depl = move_one(s, p, 0.5 * dt, dt);
      if (depl > s->crust[p]) {
#pragma omp critical(orc_err)
        if (!err) { err = ORC_ERR_MOTION; err_p = p; err_v = depl; }
      }
      for (int d = 0; d < 3; ++d) {
        if (!s->cfg.periodic[d]) continue;
        double lo = s->cfg.origin[d], L = s->cfg.size[d];
        double* xd = s->x + 3 * p + d;
        if (*xd >= lo + L) *xd -= L;
        else if (*xd < lo) *xd += L;
      }
    }
    if (err) return set_error(s, err, err_p, -1, err_v, s->crust[err_p]);
    int rc = orc_link_update(s);
    if (rc) return rc;
    const double* g = s->cfg.gravity;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int p = 0; p < s->n; ++p) {
      double m = s->mass[p] * (1. - 0. / 1.);
      s->F[3 * p] = g[0] * m;
      s->F[3 * p + 1] = g[1] * m;
      s->F[3 * p + 2] = g[2] * m;
      s->T[3 * p] = s->T[3 * p + 1] = s->T[3 * p + 2] = 0.;
    }
    if (nslab == 0) {
      err = sweep_rows(s, 0, nj);
    } else {
      for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for (int h = colour; h < nslab; h += 2) {
          int e = sweep_rows(s, (int)((long)h * nj / nslab), (int)((long)(h + 1) * nj / nslab));
          if (e) {
#pragma omp critical(orc_err)
            if (!err) err = e;
          }
        }
      }
    }
    if (err) return set_error(s, err, -1, -1, 0., 0.);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int p = 0; p < s->n; ++p) {
      quat q = qmk(mk(s->q[4 * p], s->q[4 * p + 1], s->q[4 * p + 2]), s->q[4 * p + 3]);
      quat pc = qconj(q);
      v3 tq = qmulv(pc, qleftvec(q, ld(s->T + 3 * p)));
      double mcf = s->mass[p] * 1.;
      v3 v = add(ld(s->v + 3 * p), scl(dvd(ld(s->F + 3 * p), mcf), 0.5 * dt));
      v3 wbf = ssprk3(tq, ld(s->wbf + 3 * p), s->inertia[p], 0.5 * dt);
      st(s->v + 3 * p, v);
      st(s->wbf + 3 * p, wbf);
      st(s->w + 3 * p, qmulv(q, qleftvec(pc, wbf)));
    }
  }
  return ORC_OK;
}

int orc_get_state(const orc_sys* s, double* x, double* v, double* w, double* q,
                  double* F, double* T) {
  size_t n3 = sizeof(double) * 3 * (size_t)s->n;
  if (x) memcpy(x, s->x, n3);
  if (v) memcpy(v, s->v, n3);
  if (w) memcpy(w, s->w, n3);
  if (q) memcpy(q, s->q, sizeof(double) * 4 * (size_t)s->n);
  if (F) memcpy(F, s->F, n3);
  if (T) memcpy(T, s->T, n3);
  return ORC_OK;
}

int orc_set_state(orc_sys* s, const double* x, const double* v, const double* w,
                  const double* q, const double* F, const double* T) {
  size_t n3 = sizeof(double) * 3 * (size_t)s->n;
  if (x) memcpy(s->x, x, n3);
  if (v) memcpy(s->v, v, n3);
  if (q) memcpy(s->q, q, sizeof(double) * 4 * (size_t)s->n);
  if (w) {
    memcpy(s->w, w, n3);
    for (int i = 0; i < s->n; ++i) {
      quat qq = qmk(mk(s->q[4 * i], s->q[4 * i + 1], s->q[4 * i + 2]), s->q[4 * i + 3]);
      st(s->wbf + 3 * i, qmulv(qconj(qq), qleftvec(qq, ld(s->w + 3 * i))));
    }
  }
  if (F) memcpy(s->F, F, n3);
  if (T) memcpy(s->T, T, n3);
  return ORC_OK;
}

int orc_get_grid(const orc_sys* s, int* nc, double* edge, double* origin) {
  for (int d = 0; d < 3; ++d) {
    if (nc) nc[d] = s->nc[d];
    if (edge) edge[d] = s->edge[d];
    if (origin) origin[d] = s->lc_origin[d];
  }
  return ORC_OK;
}

void orc_enable_contact_log(orc_sys* s, int cap) {
  free(s->log_id0); free(s->log_id1); free(s->log_pt); free(s->log_ov);
  free(s->log_d); free(s->log_f);
  s->log_cap = cap;
  s->log_n = 0;
  s->log_id0 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
  s->log_id1 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
  s->log_pt = (double*)malloc(sizeof(double) * 3 * (size_t)(cap > 0 ? cap : 1));
  s->log_ov = (double*)malloc(sizeof(double) * 3 * (size_t)(cap > 0 ? cap : 1));
  s->log_d = (double*)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
  s->log_f = (double*)malloc(sizeof(double) * 3 * (size_t)(cap > 0 ? cap : 1));
}

int orc_get_contacts(const orc_sys* s, int cap, int* n_out, int32_t* id0, int32_t* id1,
                     double* point, double* ovec, double* overlap, double* force) {
  *n_out = s->log_n;
  int n = s->log_n < s->log_cap ? s->log_n : s->log_cap;
  if (n > cap) n = cap;
  for (int k = 0; k < n; ++k) {
    if (id0) id0[k] = s->log_id0[k];
    if (id1) id1[k] = s->log_id1[k];
    if (point) memcpy(point + 3 * k, s->log_pt + 3 * k, sizeof(double) * 3);
    if (ovec) memcpy(ovec + 3 * k, s->log_ov + 3 * k, sizeof(double) * 3);
    if (overlap) overlap[k] = s->log_d[k];
    if (force) memcpy(force + 3 * k, s->log_f + 3 * k, sizeof(double) * 3);
  }
  return s->log_n > s->log_cap ? ORC_ERR_ARG : ORC_OK;
}

int orc_last_error(const orc_sys* s, int* code, int32_t* id0, int32_t* id1,
                   double* value, double* limit) {
  if (code) *code = s->err_code;
  if (id0) *id0 = s->err_id0;
  if (id1) *id1 = s->err_id1;
  if (value) *value = s->err_value;
  if (limit) *limit = s->err_limit;
  return s->err_code;
}

int64_t orc_candidates(const orc_sys* s) { return s->ncand; }

/* AppCollision::getOverlapMax / getOverlapMean / getTimeOverlapMax /
   getNbIterGJKMean (Base/src/AppCollision.cpp:205-240): out[0..3], and the
   number of overlaps and of GJK reports counted, out[4..5]. */
void orc_get_contact_stats(const orc_sys* s, double* out) {
  out[0] = s->ov_max;
  out[1] = s->ov_mean;
  out[2] = s->ov_time_max;
  out[3] = s->gjk_mean;
  out[4] = s->ov_count;
  out[5] = s->gjk_count;
}
