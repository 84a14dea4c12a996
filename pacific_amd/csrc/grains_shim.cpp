// grains_shim.cpp -- host classes of include/gdem_grains.hpp over the C ABI.
#include "../../include/gdem_grains.hpp"

#include <cmath>
#include <sstream>

namespace gdem {
namespace grains {

namespace {
const double kPi = 3.14159265358979323846;

std::string describe(const char* what, int a, int b, double t) {
  std::ostringstream o;
  o << what << " between components " << a << " and " << b << " at t=" << t;
  return o.str();
}

struct Registry {
  std::map<std::string, int> materials;  // material -> index (bit number)
  std::map<std::pair<int, int>, std::unique_ptr<ContactForceModel>> models;
};

Registry& registry() {
  static Registry r;
  return r;
}
}  // namespace

ContactError::ContactError(int a, int b, double t)
    : std::runtime_error(describe("ContactError: overlap exceeds the crusts", a, b, t)),
      id0(a), id1(b), time(t) {}
void ContactError::Message(std::ostream& f) const { f << what() << std::endl; }

MotionError::MotionError(int a, double t)
    : std::runtime_error(describe("MotionError: displacement exceeds the crust", a, a, t)),
      id(a), time(t) {}
void MotionError::Message(std::ostream& f) const { f << what() << std::endl; }

SimulationError::SimulationError(int a, double t)
    : std::runtime_error(describe("SimulationError: particle outside the linked cell", a, a, t)),
      id(a), time(t) {}
void SimulationError::Message(std::ostream& f) const { f << what() << std::endl; }

std::array<double, 3> GrainsExec::m_vgravity = {0., 0., 0.};
std::string GrainsExec::m_TIScheme = "SecondOrderLeapFrog";

// ------------------------------------------------------------ contact models
ContactForceModel::ContactForceModel(std::map<std::string, double>&) {
  for (double& p : m_p) p = 0.;
}
ContactForceModel::~ContactForceModel() = default;

static double need(std::map<std::string, double>& p, const char* key) {
  auto it = p.find(key);
  if (it == p.end()) throw std::invalid_argument(std::string("missing contact parameter ") + key);
  return it->second;
}

HookeContactForceModel::HookeContactForceModel(std::map<std::string, double>& p)
    : ContactForceModel(p) {
  m_p[0] = need(p, "kn");
  m_p[1] = need(p, "en");
  m_p[2] = p.count("etat") ? p["etat"] : -1.;
  m_p[3] = need(p, "muc");
  m_p[4] = p.count("kr") ? p["kr"] : 0.;
}

// HookeContactForceModel::computeAndWriteEstimates (HookeContactForceModel.cpp:208-230)
double HookeContactForceModel::contactTime(double m0, double m1) const {
  const double avmass = 1. / (1. / m0 + 1. / m1);
  const double beta = log(m_p[1]) / sqrt(kPi * kPi + log(m_p[1]) * log(m_p[1]));
  const double etan = -sqrt(m_p[0] / avmass) * beta;
  const double omega0 = sqrt(m_p[0] / avmass);
  const double theta = sqrt(pow(omega0, 2.) - pow(etan, 2.));
  return kPi / theta;
}

// Maximum overlap of the damped oscillator: the first zero of the velocity,
// solved by Newton from tc / 2 as computeDeltaMax does (:238-256).
double HookeContactForceModel::maxOverlap(double m0, double m1, double v0) const {
  const double avmass = 1. / (1. / m0 + 1. / m1);
  const double beta = log(m_p[1]) / sqrt(kPi * kPi + log(m_p[1]) * log(m_p[1]));
  const double eta = -sqrt(m_p[0] / avmass) * beta;
  const double omega0 = sqrt(m_p[0] / avmass);
  const double theta = sqrt(omega0 * omega0 - eta * eta);
  const double tc = kPi / theta;
  double f = 1., t0 = tc / (m_p[1] > 0.2 ? 2. : 100.);
  for (int it = 0; it < 200 && fabs(f) > 1e-10; ++it) {
    f = (v0 / theta) * exp(-eta * t0) * (-eta * sin(theta * t0) + theta * cos(theta * t0));
    const double df = (v0 / theta) * exp(-eta * t0) *
                      (eta * eta * sin(theta * t0) - 2. * eta * theta * cos(theta * t0) -
                       theta * theta * sin(theta * t0));
    t0 -= f / df;
  }
  return (v0 / theta) * exp(-eta * t0) * sin(theta * t0);
}

void HookeContactForceModel::computeAndWriteEstimates(double m0, double m1, double crust0,
                                                      double crust1, double, double, double v0,
                                                      double, std::ostream& OUT) const {
  const double tc = contactTime(m0, m1);
  const double dmax = maxOverlap(m0, m1, v0);
  OUT << "  Maximum overlap allowed by crusts = " << crust0 + crust1 << "\n";
  OUT << "  Collisional relative velocity = " << v0 << "\n";
  OUT << "  Tc = " << tc << "  delta_max = " << dmax << "\n";
  if (dmax > crust0 + crust1) OUT << "  WARNING: delta_max > maximum overlap allowed by crusts\n";
}

HookeMemoryContactForceModel::HookeMemoryContactForceModel(std::map<std::string, double>& p)
    : ContactForceModel(p) {
  m_p[0] = need(p, "kn");
  m_p[1] = need(p, "en");
  m_p[2] = p.count("etat") ? p["etat"] : -1.;
  m_p[3] = need(p, "muc");
  m_p[4] = need(p, "kt");
  m_p[5] = p.count("mur") ? p["mur"] : 0.;
  m_p[6] = p.count("etarpf") ? p["etarpf"] : 0.;
  m_p[7] = p.count("epst") ? p["epst"] : 0.;
}

// HookeMemoryContactForceModel::computeAndWriteEstimates (:319-353) uses the
// normal Hooke law of HookeContactForceModel.
void HookeMemoryContactForceModel::computeAndWriteEstimates(double m0, double m1, double crust0,
                                                            double crust1, double req0,
                                                            double req1, double v0, double dt,
                                                            std::ostream& OUT) const {
  std::map<std::string, double> q{{"kn", m_p[0]}, {"en", m_p[1]}, {"muc", m_p[3]}};
  HookeContactForceModel(q).computeAndWriteEstimates(m0, m1, crust0, crust1, req0, req1, v0, dt,
                                                     OUT);
}

HertzMemoryContactForceModel::HertzMemoryContactForceModel(std::map<std::string, double>& p)
    : ContactForceModel(p) {
  m_p[0] = need(p, "Es");
  m_p[1] = need(p, "en");
  m_p[2] = p.count("Gs") ? p["Gs"] : m_p[0] / 4.;
  m_p[3] = need(p, "muc");
  m_p[4] = p.count("mur") ? p["mur"] : 0.;
  m_p[5] = p.count("etarpf") ? p["etarpf"] : 0.;
  m_p[6] = p.count("epst") ? p["epst"] : 0.;
}

// HertzMemoryContactForceModel::computeAndWriteEstimates (:316-360) uses the
// normal Hertz law of HertzContactForceModel.
void HertzMemoryContactForceModel::computeAndWriteEstimates(double m0, double m1, double crust0,
                                                            double crust1, double req0,
                                                            double req1, double v0, double dt,
                                                            std::ostream& OUT) const {
  std::map<std::string, double> q{{"Es", m_p[0]}, {"en", m_p[1]}, {"Gs", m_p[2]}, {"muc", m_p[3]}};
  HertzContactForceModel(q).computeAndWriteEstimates(m0, m1, crust0, crust1, req0, req1, v0, dt,
                                                     OUT);
}

HertzContactForceModel::HertzContactForceModel(std::map<std::string, double>& p)
    : ContactForceModel(p) {
  m_p[0] = need(p, "Es");
  m_p[1] = need(p, "en");
  m_p[2] = p.count("Gs") ? p["Gs"] : m_p[0] / 4.;
  m_p[3] = need(p, "muc");
  m_p[4] = p.count("kr") ? p["kr"] : 0.;
}

// HertzContactForceModel::computeAndWriteEstimates (HertzContactForceModel.cpp:209-260):
// RK2 integration of the normal motion until the overlap closes.
void HertzContactForceModel::computeAndWriteEstimates(double m0, double m1, double crust0,
                                                      double crust1, double req0, double req1,
                                                      double v0, double dt,
                                                      std::ostream& OUT) const {
  const double Req = 1. / (1. / req0 + 1. / req1);
  const double avmass = 1. / (1. / m0 + 1. / m1);
  const double beta = log(m_p[1]) / sqrt(kPi * kPi + log(m_p[1]) * log(m_p[1]));
  const double m2sqrt56 = -2. * sqrt(5. / 6.);
  auto dvdt = [&](double dn, double v) {
    if (dn <= 0.) return 0.;
    const double sq = sqrt(Req * dn);
    const double kn = (4. / 3.) * m_p[0] * sq;
    const double gn = m2sqrt56 * beta * sqrt(avmass * 2. * m_p[0] * sq);
    return -(kn * dn + gn * v) / avmass;
  };
  double t = 0., dn = 0., v = v0, dmax = -1e20;
  while (dn >= 0. && t < 1.) {
    const double k1d = v * dt, k1v = dvdt(dn, v) * dt;
    const double k2d = (v + k1v / 2.) * dt, k2v = dvdt(dn + k1d / 2., v + k1v / 2.) * dt;
    dn += k2d;
    v += k2v;
    t += dt;
    dmax = std::max(dn, dmax);
  }
  OUT << "  Maximum overlap allowed by crusts = " << crust0 + crust1 << "\n";
  OUT << "  Collisional relative velocity = " << v0 << "\n";
  OUT << "  Tc = " << t << "  delta_max = " << dmax << "\n";
  if (dmax > crust0 + crust1) OUT << "  WARNING: delta_max > maximum overlap allowed by crusts\n";
}

// ------------------------------------------------------------ factory
void ContactBuilderFactory::defineMaterial(std::string const& material, bool) {
  Registry& r = registry();
  if (r.materials.count(material)) return;
  const int idx = (int)r.materials.size();
  if (idx >= GDEM_MAX_MATERIALS) throw std::length_error("too many materials");
  r.materials[material] = idx;
}

void ContactBuilderFactory::define(std::string const& a, std::string const& b,
                                   std::unique_ptr<ContactForceModel> model) {
  defineMaterial(a, false);
  defineMaterial(b, false);
  int ia = materialIndex(a), ib = materialIndex(b);
  registry().models[{std::min(ia, ib), std::max(ia, ib)}] = std::move(model);
}

ContactForceModel* ContactBuilderFactory::contactForceModel(std::string const& a,
                                                            std::string const& b) {
  int ia = materialIndex(a), ib = materialIndex(b);
  auto it = registry().models.find({std::min(ia, ib), std::max(ia, ib)});
  return it == registry().models.end() ? nullptr : it->second.get();
}

int ContactBuilderFactory::materialIndex(std::string const& material) {
  auto it = registry().materials.find(material);
  if (it == registry().materials.end()) throw std::invalid_argument("unknown material " + material);
  return it->second;
}

bool ContactBuilderFactory::checkContactForceModelsExist(std::string& matA, std::string& matB) {
  for (auto const& a : registry().materials)
    for (auto const& b : registry().materials)
      if (!contactForceModel(a.first, b.first)) {
        matA = a.first;
        matB = b.first;
        return false;
      }
  return true;
}

void ContactBuilderFactory::eraseAllContactForceModels() { registry().models.clear(); }

void ContactBuilderFactory::reset() {
  registry().models.clear();
  registry().materials.clear();
}

TimeIntegrator::~TimeIntegrator() = default;

TimeIntegrator* TimeIntegratorBuilderFactory::create(std::string const& scheme) {
  if (scheme == "SecondOrderLeapFrog") return new SecondOrderLeapFrog();
  throw std::invalid_argument("time integration scheme " + scheme +
                              " has no device implementation (SecondOrderLeapFrog only)");
}

// ------------------------------------------------------------ linked cell
LinkedCell::LinkedCell(const double origin[3], const double size[3], const bool periodic[3],
                       int device)
    : m_device(device) {
  for (int d = 0; d < 3; ++d) {
    m_origin[d] = origin[d];
    m_size[d] = size[d];
    m_periodic[d] = periodic[d];
  }
}

LinkedCell::~LinkedCell() {
  if (m_ctx) gdem_destroy(m_ctx);
}

size_t LinkedCell::set(double cellsize, std::string const&) {
  gdem_config cfg{};
  for (int d = 0; d < 3; ++d) {
    cfg.origin[d] = m_origin[d];
    cfg.size[d] = m_size[d];
    cfg.periodic[d] = m_periodic[d] ? 1 : 0;
    cfg.gravity[d] = GrainsExec::m_vgravity[d];
  }
  cfg.cell_size = cellsize;
  cfg.device = m_device;
  if (m_ctx) gdem_destroy(m_ctx);
  m_ctx = nullptr;
  return gdem_create(&cfg, &m_ctx) == GDEM_OK ? 0 : 1;
}

void LinkedCell::push_models() {
  for (auto const& kv : registry().models) {
    const ContactForceModel* m = kv.second.get();
    if (gdem_set_contact_model(m_ctx, kv.first.first, kv.first.second, m->kind(), m->parameters()))
      throw std::invalid_argument("contact model " + m->name() + " rejected");
  }
}

void LinkedCell::Link(std::vector<SphereParticle> const& ps) {
  if (!m_ctx) throw std::logic_error("LinkedCell::set must be called first");
  push_models();
  const size_t n = ps.size();
  std::vector<double> x(3 * n), v(3 * n), w(3 * n), q(4 * n), r(n), rho(n), cr(n);
  std::vector<int32_t> mat(n);
  for (size_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      x[3 * i + d] = ps[i].position[d];
      v[3 * i + d] = ps[i].velocity[d];
      w[3 * i + d] = ps[i].angular_velocity[d];
    }
    for (int d = 0; d < 4; ++d) q[4 * i + d] = ps[i].quaternion[d];
    r[i] = ps[i].radius;
    rho[i] = ps[i].density;
    cr[i] = ps[i].crust;
    ContactBuilderFactory::defineMaterial(ps[i].material, false);
    mat[i] = ContactBuilderFactory::materialIndex(ps[i].material);
  }
  if (gdem_set_particles(m_ctx, (int64_t)n, x.data(), v.data(), w.data(), q.data(), r.data(),
                         rho.data(), cr.data(), mat.data()))
    throw std::invalid_argument("particle set rejected");
  m_n = n;
}

void LinkedCell::Link(BoxObstacle const& ob) {
  if (!m_ctx) throw std::logic_error("LinkedCell::set must be called first");
  ContactBuilderFactory::defineMaterial(ob.material, true);
  const double half[3] = {fabs(ob.edges[0]) / 2., fabs(ob.edges[1]) / 2., fabs(ob.edges[2]) / 2.};
  if (gdem_add_wall(m_ctx, ob.basis, ob.centre, half, ob.crust,
                    ContactBuilderFactory::materialIndex(ob.material)))
    throw std::invalid_argument("obstacle " + ob.name + " rejected");
  m_obstacles.emplace(ob.name, m_nobstacles++);
  push_models();
}

void LinkedCell::Link(RectangleObstacle const& ob) {
  if (!m_ctx) throw std::logic_error("LinkedCell::set must be called first");
  ContactBuilderFactory::defineMaterial(ob.material, true);
  if (gdem_add_rectangle_wall(m_ctx, ob.basis, ob.centre, ob.LX, ob.LY, ob.crust,
                              ContactBuilderFactory::materialIndex(ob.material)))
    throw std::invalid_argument("obstacle " + ob.name + " rejected");
  m_obstacles.emplace(ob.name, m_nobstacles++);
  push_models();
}

int LinkedCell::obstacleIndex(std::string const& name) const {
  auto it = m_obstacles.find(name);
  return it == m_obstacles.end() ? -1 : it->second;
}

void LinkedCell::setRestrictedGeomDirMotion(std::string const& obstacle,
                                            std::list<size_t> const& dirs) {
  int32_t mask = 0;
  for (size_t d : dirs) {
    if (d > 2) throw std::invalid_argument("direction out of range");
    mask |= 1 << d;
  }
  const int w = obstacleIndex(obstacle), c = compositeIndex(obstacle);
  if (w < 0 && c < 0) throw std::invalid_argument("no obstacle named " + obstacle);
  if (gdem_set_motion_restriction(m_ctx, w >= 0 ? w : -1, w >= 0 ? -1 : c, mask))
    throw std::invalid_argument("restriction of " + obstacle + " rejected");
}

int LinkedCell::compositeIndex(std::string const& name) const {
  auto it = m_composites.find(name);
  return it == m_composites.end() ? -1 : it->second;
}

void LinkedCell::LinkComposite(std::string const& name, std::string const& parent) {
  if (!m_ctx) throw std::logic_error("LinkedCell::set must be called first");
  const int p = parent.empty() ? -1 : compositeIndex(parent);
  if (!parent.empty() && p < 0) throw std::invalid_argument("no composite named " + parent);
  int32_t id = -1;
  if (gdem_add_composite(m_ctx, p, &id)) throw std::invalid_argument("composite " + name + " rejected");
  m_composites[name] = id;
}

void LinkedCell::Append(std::string const& obstacle, std::string const& composite) {
  const int w = obstacleIndex(obstacle), c = compositeIndex(composite);
  if (w < 0 || c < 0 || gdem_attach_wall(m_ctx, w, c))
    throw std::invalid_argument("cannot append " + obstacle + " to " + composite);
}

void LinkedCell::Link(CylindricalShell const& sh, std::string const& parent) {
  if (!m_ctx) throw std::logic_error("LinkedCell::set must be called first");
  ContactBuilderFactory::defineMaterial(sh.material, true);
  const int p = parent.empty() ? -1 : compositeIndex(parent);
  if (!parent.empty() && p < 0) throw std::invalid_argument("no composite named " + parent);
  int32_t id = -1;
  const int32_t mat = ContactBuilderFactory::materialIndex(sh.material);
  const int rc = sh.elementary_obstacle == "Box"
                     ? gdem_add_cylindrical_shell(m_ctx, sh.basis, sh.centre, sh.height, sh.radius,
                                                  sh.width, sh.n, sh.crust, mat, p, &id)
                     : gdem_add_cylindrical_shell_rectangles(m_ctx, sh.basis, sh.centre, sh.height,
                                                             sh.radius, sh.width, sh.n, sh.crust,
                                                             mat, p, &id);
  if (rc) throw std::invalid_argument("cylindrical shell " + sh.name + " rejected");
  m_composites[sh.name] = id;
  for (int i = 0; i < sh.n; ++i) m_obstacles.emplace(sh.name + std::to_string(i), m_nobstacles++);
  push_models();
}

void LinkedCell::LinkUpdate(double, double) {
  if (gdem_link_update(m_ctx)) throw std::runtime_error("gdem_link_update failed");
}

void LinkedCell::ComputeForces(double, double dt) {
  if (dt > 0. && gdem_set_time_step(m_ctx, dt))
    throw std::runtime_error("gdem_set_time_step failed");
  if (gdem_compute_forces(m_ctx)) throw std::runtime_error("gdem_compute_forces failed");
}

void LinkedCell::getCells(int32_t nc[3], double edge[3]) const { gdem_get_grid(m_ctx, nc, edge); }

// ------------------------------------------------------------ step driver
Grains::Grains(LinkedCell& lc, TimeIntegrator const& ti) : m_lc(lc) {
  if (ti.name() != "SecondOrderLeapFrog")
    throw std::invalid_argument("only SecondOrderLeapFrog has a device implementation");
}

void Grains::check(int rc) const {
  if (rc == GDEM_OK) return;
  int32_t code = 0, a = -1, b = -1;
  gdem_last_error(m_lc.context(), &code, &a, &b);
  switch (rc) {
    case GDEM_ERR_CONTACT: throw ContactError(a, b, m_time);
    case GDEM_ERR_MOTION: throw MotionError(a, m_time);
    case GDEM_ERR_SIMULATION: throw SimulationError(a, m_time);
    default: throw std::runtime_error("gdem call failed with code " + std::to_string(rc));
  }
}

void Grains::do_before_time_stepping() {
  if (m_dt > 0.) check(gdem_set_time_step(m_lc.context(), m_dt));
  check(gdem_prime(m_lc.context()));
  check(gdem_synchronize(m_lc.context()));
  m_primed = true;
}

void Grains::Simulation(double time_interval, double dt) {
  m_dt = dt;
  if (!m_primed) do_before_time_stepping();
  const int32_t nsteps = (int32_t)llround(time_interval / dt);
  check(gdem_step(m_lc.context(), nsteps, dt));
  check(gdem_synchronize(m_lc.context()));
  check(gdem_get_time(m_lc.context(), &m_time));  // m_time += dt at every step
}

void Grains::setTime(double time) {
  m_time = time;
  check(gdem_set_time(m_lc.context(), time));
}

// The parameters of the reference's XML node, in the order gdem_add_wall_motion
// takes them; a rotation about a given point when CX is present.
void Grains::LinkImposedMotion(ObstacleImposedVelocity const& mo) {
  const int wall = m_lc.obstacleIndex(mo.obstacle);
  const int comp = m_lc.compositeIndex(mo.obstacle);
  if (wall < 0 && comp < 0) throw std::invalid_argument("no obstacle named " + mo.obstacle);
  auto get = [&](const char* k) {
    auto it = mo.parameters.find(k);
    if (it == mo.parameters.end())
      throw std::invalid_argument(mo.type + " of " + mo.obstacle + ": no " + k);
    return it->second;
  };
  std::vector<double> p;
  int32_t type = 0;
  if (mo.type == "ConstantTranslation") {
    type = GDEM_MOTION_CONSTANT_TRANSLATION;
    p = {get("VX"), get("VY"), get("VZ")};
  } else if (mo.type == "SinTranslation") {
    type = GDEM_MOTION_SIN_TRANSLATION;
    p = {get("VX"), get("VY"), get("VZ"), get("Amplitude"), get("Period"), get("PhaseShift")};
  } else if (mo.type == "CrenelTranslation") {
    type = GDEM_MOTION_CRENEL_TRANSLATION;
    p = {get("VX"), get("VY"), get("VZ"), get("Amplitude"), get("Period")};
  } else if (mo.type == "MultiSinTranslation") {
    type = GDEM_MOTION_MULTISIN_TRANSLATION;
    p = {get("VX"), get("VY"), get("VZ"), get("AX"), get("AY"), get("AZ"),
         get("PX"), get("PY"), get("PZ"), get("PhiX"), get("PhiY"), get("PhiZ")};
  } else if (mo.type == "ConstantRotation") {
    type = GDEM_MOTION_CONSTANT_ROTATION;
    const bool centre = mo.parameters.count("CX") != 0;
    p = {get("WX"), get("WY"), get("WZ"), centre ? get("CX") : 0., centre ? get("CY") : 0.,
         centre ? get("CZ") : 0., centre ? 0. : 1.};
  } else {
    throw std::invalid_argument("imposed velocity type " + mo.type + " is not supported");
  }
  const int rc =
      comp >= 0 ? gdem_add_composite_motion(m_lc.context(), comp, type, mo.tstart, mo.tend, p.data())
                : gdem_add_wall_motion(m_lc.context(), wall, type, mo.tstart, mo.tend, p.data());
  if (rc) throw std::invalid_argument(mo.type + " of " + mo.obstacle + " rejected");
}

void Grains::saveParticles2014(std::string const& path, bool binary, int id_offset,
                               int wall_id_offset) {
  check(gdem_write_particles_2014(m_lc.context(), path.c_str(), binary ? 1 : 0, id_offset,
                                  wall_id_offset, nullptr));
}

void Grains::readParticles2014(std::string const& path, bool binary, int id_offset,
                               int wall_id_offset) {
  check(gdem_read_particles_2014(m_lc.context(), path.c_str(), binary ? 1 : 0, id_offset,
                                 wall_id_offset, nullptr));
  m_primed = true;  // the stored F, T are a0
}

void Grains::LinkImposedMotion(ObstacleImposedForce const& mo) {
  // a simple obstacle, or a composite (CompositeObstacle::LinkImposedMotion)
  const int wall = m_lc.obstacleIndex(mo.obstacle);
  const int comp = wall < 0 ? m_lc.compositeIndex(mo.obstacle) : -1;
  if (wall < 0 && comp < 0) throw std::invalid_argument("no obstacle named " + mo.obstacle);
  if (mo.type != "ConstantTranslation")
    throw std::invalid_argument("imposed force type " + mo.type + " is not supported");
  auto get = [&](const char* k, double dflt, bool required) {
    auto it = mo.parameters.find(k);
    if (it == mo.parameters.end()) {
      if (required) throw std::invalid_argument("imposed force on " + mo.obstacle + ": no " + k);
      return dflt;
    }
    return it->second;
  };
  const double p[5] = {get("AX", 0., true), get("AY", 0., true), get("AZ", 0., true),
                       get("Mass", 0., false), get("Vmax", 0., true)};
  const int rc = wall >= 0 ? gdem_add_wall_force(m_lc.context(), wall, mo.tstart, mo.tend, p)
                           : gdem_add_composite_force(m_lc.context(), comp, mo.tstart, mo.tend, p);
  if (rc) throw std::invalid_argument("imposed force on " + mo.obstacle + " rejected");
}

void Grains::trackObstaclesLoad(bool on) { check(gdem_set_wall_loads(m_lc.context(), on ? 1 : 0)); }

void Grains::getObstacleLoad(std::string const& name, double force[3], double torque[3]) const {
  const int wall = m_lc.obstacleIndex(name);
  if (wall < 0) throw std::invalid_argument("no box obstacle named " + name);
  int64_t n = 0;
  int32_t nw = 0;
  check(gdem_info(m_lc.context(), &n, &nw, nullptr, nullptr));
  std::vector<double> f(3 * (size_t)std::max(nw, 1)), t(3 * (size_t)std::max(nw, 1));
  check(gdem_get_wall_loads(m_lc.context(), f.data(), t.data()));
  for (int d = 0; d < 3; ++d) {
    force[d] = f[3 * wall + d];
    torque[d] = t[3 * wall + d];
  }
}

void Grains::moveParticlesAndObstacles(double dtv, double dtd, double dto) {
  check(gdem_move(m_lc.context(), dtv, dtd));
  if (dto > 0.) check(gdem_move_walls(m_lc.context(), m_time, dto));
  m_lc.LinkUpdate(m_time, dtd);
}

void Grains::computeParticlesForceAndTorque() { m_lc.ComputeForces(m_time, m_dt); }

void Grains::advanceParticlesVelocity(double dtv) {
  check(gdem_advance_velocity(m_lc.context(), dtv));
  check(gdem_synchronize(m_lc.context()));
}

void Grains::getState(std::vector<double>& x, std::vector<double>& v, std::vector<double>& w,
                      std::vector<double>& q, std::vector<double>& F,
                      std::vector<double>& T) const {
  const size_t n = m_lc.numberOfParticles();
  x.resize(3 * n);
  v.resize(3 * n);
  w.resize(3 * n);
  q.resize(4 * n);
  F.resize(3 * n);
  T.resize(3 * n);
  check(gdem_get_state(m_lc.context(), x.data(), v.data(), w.data(), q.data(), F.data(), T.data()));
}

}  // namespace grains
}  // namespace gdem
