// grains_run.cpp -- command-line driver of the DEM step through the C++ host
// classes (the role of src/Grains3D/Main/src/main.cpp, with a binary input in
// place of the XML file).
//
//   grains_run <input.bin> <output.bin>
//
// Input (little endian): "GDEMIN01", "GDEMIN02" or "GDEMIN03", int32 n, nwalls,
// nmodels, nsteps, double dt, origin[3], size[3], int32 periodic[3], double
// cell_size, gravity[3]; per model int32 kind, mat_a, mat_b, double p[5]
// (GDEMIN01) or p[8] (GDEMIN02/03: the memory models' parameters,
// include/gdem.h); per wall double basis[9], centre[3], half_extent[3], crust,
// int32 material; GDEMIN03 then: int32 nmotions, per motion int32 wall, type
// (GDEM_MOTION_*), double t_start, t_end, p[12] (gdem_add_wall_motion's
// parameters); GDEMIN04 then also: int32 ncomposites, per composite int32
// parent (-1 root), nmembers, member walls[nmembers]; int32 nshells, per shell
// double basis[9], centre[3], height, radius, width, int32 n, double crust,
// int32 material, parent; int32 ncomposite_motions, per motion int32
// composite, type, double t_start, t_end, p[12] (composites numbered in
// creation order, then the shells; a shell's boxes follow the walls); int32
// nforces, per imposed force int32 wall (-1 - c for composite c), double
// t_start, t_end, ax, ay, az, mass (<= 0: automatic), vmax.  GDEMIN05 is GDEMIN04 with, after the walls,
// int32 nrectangles, per rectangle double basis[9], centre[3], lx, ly,
// crust, int32 material (wall indices after the boxes), and each shell
// record ending with int32 elements (1 Box, 0 Rectangle); then x[3n], v[3n],
// w[3n], q[4n], radius[n], density[n], crust[n], int32 material[n].
// Output: "GDEMOUT1", int32 n, status (0 or the GDEM_ERR_* the step raised),
// id0, id1, then x, v, w, q, F, T of the final state.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

#include "../../include/gdem_grains.hpp"

using namespace gdem::grains;

template <class T>
static bool rd(FILE* f, T* p, size_t n) {
  return fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: grains_run <input.bin> <output.bin>\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  char magic[8];
  int32_t n, nwalls, nmodels, nsteps, periodic[3];
  double dt, origin[3], size[3], cell, grav[3];
  if (!rd(f, magic, 8) || (memcmp(magic, "GDEMIN01", 8) && memcmp(magic, "GDEMIN02", 8) &&
                            memcmp(magic, "GDEMIN03", 8) && memcmp(magic, "GDEMIN04", 8) &&
                            memcmp(magic, "GDEMIN05", 8)))
    return 2;
  const bool with_rect = memcmp(magic, "GDEMIN05", 8) == 0;
  const bool with_tree = with_rect || memcmp(magic, "GDEMIN04", 8) == 0;
  const bool with_motions = with_tree || memcmp(magic, "GDEMIN03", 8) == 0;
  const int np = memcmp(magic, "GDEMIN01", 8) == 0 ? 5 : 8;
  if (!rd(f, &n, 1) || !rd(f, &nwalls, 1) ||
      !rd(f, &nmodels, 1) || !rd(f, &nsteps, 1) || !rd(f, &dt, 1) || !rd(f, origin, 3) ||
      !rd(f, size, 3) || !rd(f, periodic, 3) || !rd(f, &cell, 1) || !rd(f, grav, 3))
    return 2;
  for (int d = 0; d < 3; ++d) GrainsExec::m_vgravity[d] = grav[d];
  ContactBuilderFactory::reset();
  for (int m = 0; m < nmodels; ++m) {
    int32_t kind, ma, mb;
    double p[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    if (!rd(f, &kind, 1) || !rd(f, &ma, 1) || !rd(f, &mb, 1) || !rd(f, p, (size_t)np)) return 2;
    std::map<std::string, double> par;
    std::unique_ptr<ContactForceModel> model;
    if (kind == GDEM_MODEL_HOOKE) {
      par = {{"kn", p[0]}, {"en", p[1]}, {"etat", p[2]}, {"muc", p[3]}, {"kr", p[4]}};
      model.reset(new HookeContactForceModel(par));
    } else if (kind == GDEM_MODEL_HERTZ) {
      par = {{"Es", p[0]}, {"en", p[1]}, {"Gs", p[2]}, {"muc", p[3]}, {"kr", p[4]}};
      model.reset(new HertzContactForceModel(par));
    } else if (kind == GDEM_MODEL_HOOKE_MEMORY) {
      par = {{"kn", p[0]}, {"en", p[1]}, {"etat", p[2]}, {"muc", p[3]}, {"kt", p[4]},
             {"mur", p[5]}, {"etarpf", p[6]}, {"epst", p[7]}};
      model.reset(new HookeMemoryContactForceModel(par));
    } else if (kind == GDEM_MODEL_HERTZ_MEMORY) {
      par = {{"Es", p[0]}, {"en", p[1]}, {"Gs", p[2]}, {"muc", p[3]}, {"mur", p[4]},
             {"etarpf", p[5]}, {"epst", p[6]}};
      model.reset(new HertzMemoryContactForceModel(par));
    } else {
      return 2;
    }
    // materials are named by index so that indices match the input file
    for (int q = 0; q <= std::max(ma, mb); ++q)
      ContactBuilderFactory::defineMaterial("mat" + std::to_string(q), false);
    ContactBuilderFactory::define("mat" + std::to_string(ma), "mat" + std::to_string(mb),
                                  std::move(model));
  }
  std::vector<BoxObstacle> walls((size_t)nwalls);
  for (auto& w : walls) {
    double half[3];
    int32_t mat;
    if (!rd(f, w.basis, 9) || !rd(f, w.centre, 3) || !rd(f, half, 3) || !rd(f, &w.crust, 1) ||
        !rd(f, &mat, 1))
      return 2;
    for (int d = 0; d < 3; ++d) w.edges[d] = 2. * half[d];
    w.material = "mat" + std::to_string(mat);
    w.name = "wall" + std::to_string(&w - walls.data());
  }
  std::vector<RectangleObstacle> rects;
  if (with_rect) {
    int32_t nr = 0;
    if (!rd(f, &nr, 1) || nr < 0) return 2;
    rects.resize((size_t)nr);
    for (auto& rc : rects) {
      int32_t mat;
      if (!rd(f, rc.basis, 9) || !rd(f, rc.centre, 3) || !rd(f, &rc.LX, 1) || !rd(f, &rc.LY, 1) ||
          !rd(f, &rc.crust, 1) || !rd(f, &mat, 1))
        return 2;
      rc.material = "mat" + std::to_string(mat);
      rc.name = "wall" + std::to_string(walls.size() + (size_t)(&rc - rects.data()));
    }
  }
  // imposed velocities as the reference's XML nodes name their parameters
  // imposed velocities, read as (target, record); target = wall index, or
  // -1 - composite index for a composite's motion
  std::vector<std::pair<int32_t, ObstacleImposedVelocity>> motions;
  int32_t nmotions = 0;
  if (with_motions && !rd(f, &nmotions, 1)) return 2;
  static const char* const kNames[6][12] = {
      {},
      {"VX", "VY", "VZ"},
      {"VX", "VY", "VZ", "Amplitude", "Period", "PhaseShift"},
      {"VX", "VY", "VZ", "Amplitude", "Period"},
      {"VX", "VY", "VZ", "AX", "AY", "AZ", "PX", "PY", "PZ", "PhiX", "PhiY", "PhiZ"},
      {"WX", "WY", "WZ", "CX", "CY", "CZ"}};
  static const char* const kTypes[6] = {"", "ConstantTranslation", "SinTranslation",
                                        "CrenelTranslation", "MultiSinTranslation",
                                        "ConstantRotation"};
  auto read_motion = [&](bool composite) -> bool {
    int32_t target, type;
    double t[2], p[12];
    if (!rd(f, &target, 1) || !rd(f, &type, 1) || !rd(f, t, 2) || !rd(f, p, 12)) return false;
    if (type < 1 || type > 5) return false;
    ObstacleImposedVelocity mo;
    mo.type = kTypes[type];
    mo.tstart = t[0];
    mo.tend = t[1];
    // a rotation about the obstacle's centre has no CX, CY, CZ (p[6] != 0)
    static const int kCount[5] = {0, 3, 6, 5, 12};
    const int np_k = type == 5 ? (p[6] != 0. ? 3 : 6) : kCount[type];
    for (int i = 0; i < np_k; ++i) mo.parameters[kNames[type][i]] = p[i];
    motions.emplace_back(composite ? -1 - target : target, mo);
    return true;
  };
  for (int32_t k = 0; k < nmotions; ++k)
    if (!read_motion(false)) return 2;
  // composites, cylindrical shells and their motions
  std::vector<std::pair<int32_t, std::vector<int32_t>>> composites;
  std::vector<std::pair<CylindricalShell, int32_t>> shells;
  std::vector<std::pair<int32_t, ObstacleImposedForce>> forces;
  if (with_tree) {
    int32_t nc = 0, ns = 0, ncm = 0;
    if (!rd(f, &nc, 1)) return 2;
    for (int32_t k = 0; k < nc; ++k) {
      int32_t parent, nm;
      if (!rd(f, &parent, 1) || !rd(f, &nm, 1) || nm < 0) return 2;
      std::vector<int32_t> mem((size_t)nm);
      if (nm && !rd(f, mem.data(), (size_t)nm)) return 2;
      composites.emplace_back(parent, mem);
    }
    if (!rd(f, &ns, 1)) return 2;
    for (int32_t k = 0; k < ns; ++k) {
      CylindricalShell sh;
      int32_t nbox, mat, parent;
      if (!rd(f, sh.basis, 9) || !rd(f, sh.centre, 3) || !rd(f, &sh.height, 1) ||
          !rd(f, &sh.radius, 1) || !rd(f, &sh.width, 1) || !rd(f, &nbox, 1) || !rd(f, &sh.crust, 1) ||
          !rd(f, &mat, 1) || !rd(f, &parent, 1))
        return 2;
      int32_t box = 1;
      if (with_rect && !rd(f, &box, 1)) return 2;
      sh.elementary_obstacle = box ? "Box" : "";
      sh.n = nbox;
      sh.material = "mat" + std::to_string(mat);
      sh.name = "shell" + std::to_string(k) + "_";
      shells.emplace_back(sh, parent);
    }
    if (!rd(f, &ncm, 1)) return 2;
    for (int32_t k = 0; k < ncm; ++k)
      if (!read_motion(true)) return 2;
    int32_t nf = 0;
    if (!rd(f, &nf, 1)) return 2;
    for (int32_t k = 0; k < nf; ++k) {
      int32_t wall;
      double v[7];
      if (!rd(f, &wall, 1) || !rd(f, v, 7)) return 2;
      ObstacleImposedForce fo;
      fo.type = "ConstantTranslation";
      fo.tstart = v[0];
      fo.tend = v[1];
      fo.parameters = {{"AX", v[2]}, {"AY", v[3]}, {"AZ", v[4]}, {"Vmax", v[6]}};
      if (v[5] > 0.) fo.parameters["Mass"] = v[5];
      forces.emplace_back(wall, fo);
    }
  }
  std::vector<double> x(3 * n), v(3 * n), w(3 * n), q(4 * n), r(n), rho(n), cr(n);
  std::vector<int32_t> mat(n);
  if (!rd(f, x.data(), x.size()) || !rd(f, v.data(), v.size()) || !rd(f, w.data(), w.size()) ||
      !rd(f, q.data(), q.size()) || !rd(f, r.data(), r.size()) || !rd(f, rho.data(), rho.size()) ||
      !rd(f, cr.data(), cr.size()) || !rd(f, mat.data(), mat.size()))
    return 2;
  fclose(f);

  std::vector<SphereParticle> ps((size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      ps[i].position[d] = x[3 * i + d];
      ps[i].velocity[d] = v[3 * i + d];
      ps[i].angular_velocity[d] = w[3 * i + d];
    }
    for (int d = 0; d < 4; ++d) ps[i].quaternion[d] = q[4 * i + d];
    ps[i].radius = r[i];
    ps[i].density = rho[i];
    ps[i].crust = cr[i];
    ps[i].material = "mat" + std::to_string(mat[i]);
  }
  const bool per[3] = {periodic[0] != 0, periodic[1] != 0, periodic[2] != 0};
  LinkedCell lc(origin, size, per);
  if (lc.set(cell)) {
    fprintf(stderr, "LinkedCell::set failed\n");
    return 3;
  }
  for (auto const& wl : walls) lc.Link(wl);
  for (auto const& rc : rects) lc.Link(rc);
  // names of composites in creation order, and of every wall by index
  std::vector<std::string> comp_names, wall_names;
  for (auto const& wl : walls) wall_names.push_back(wl.name);
  for (auto const& rc : rects) wall_names.push_back(rc.name);
  for (auto const& c : composites) {
    const std::string name = "comp" + std::to_string(comp_names.size());
    lc.LinkComposite(name, c.first < 0 ? "" : comp_names.at((size_t)c.first));
    comp_names.push_back(name);
    for (int32_t w : c.second) lc.Append(wall_names.at((size_t)w), name);
  }
  for (auto const& sh : shells) {
    lc.Link(sh.first, sh.second < 0 ? "" : comp_names.at((size_t)sh.second));
    comp_names.push_back(sh.first.name);
    for (int i = 0; i < sh.first.n; ++i) wall_names.push_back(sh.first.name + std::to_string(i));
  }
  lc.Link(ps);
  std::unique_ptr<TimeIntegrator> ti(TimeIntegratorBuilderFactory::create());
  Grains app(lc, *ti);
  int32_t status = 0, id0 = -1, id1 = -1;
  for (auto& tm : motions) {
    tm.second.obstacle = tm.first >= 0 ? wall_names.at((size_t)tm.first)
                                       : comp_names.at((size_t)(-1 - tm.first));
    app.LinkImposedMotion(tm.second);
  }
  for (auto& fo : forces) {
    fo.second.obstacle = fo.first >= 0 ? wall_names.at((size_t)fo.first)
                                       : comp_names.at((size_t)(-1 - fo.first));
    app.LinkImposedMotion(fo.second);
  }
  try {
    app.setTimeStep(dt);  // the input's time step, before a0 (memory models use it)
    app.do_before_time_stepping();
    app.Simulation(nsteps * dt, dt);
  } catch (ContactError const& e) {
    e.Message(std::cout);
    status = GDEM_ERR_CONTACT;
    id0 = e.id0;
    id1 = e.id1;
  } catch (MotionError const& e) {
    e.Message(std::cout);
    status = GDEM_ERR_MOTION;
    id0 = e.id;
  } catch (SimulationError const& e) {
    e.Message(std::cout);
    status = GDEM_ERR_SIMULATION;
    id0 = e.id;
  }
  std::vector<double> X, V, W, Q, F, T;
  app.getState(X, V, W, Q, F, T);
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite("GDEMOUT1", 1, 8, o);
  fwrite(&n, sizeof(n), 1, o);
  fwrite(&status, sizeof(status), 1, o);
  fwrite(&id0, sizeof(id0), 1, o);
  fwrite(&id1, sizeof(id1), 1, o);
  for (auto* a : {&X, &V, &W, &Q, &F, &T}) fwrite(a->data(), sizeof(double), a->size(), o);
  fclose(o);
  printf("grains_run: %d particles, %d steps, status %d\n", n, nsteps, status);
  return 0;
}
