// restart.cpp -- checkpoint / restart through writeSerialization and
// readSerialization (CUDAenv.hh:268-285 in the reference): a run of
// 2k steps is compared with a run of k steps, a checkpoint, a fresh object
// restored from it and k more steps.  The restart is exact when the two
// end states agree bit for bit.
//   md  : liquid argon under GaussianThermo (testMD2 set-up, 864 atoms)
//   dem : spheres settling with rotation and tangential history (testDEM5)
//   sph : a small dam break (testSPH5)
//   sphc: the same dam break with the config physics (cubic spline, Tait,
//         artificial viscosity)
//
//   restart <md|dem|sph|sphc> [k=20] [file=restart.ck]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <string>
#include <valarray>
#include <vector>
#include "../include/CUDAenv.hh"
#include "../include/GaussianThermo.hh"
#include "../include/cudaParticleDEM.hh"
#include "../include/cudaParticleSPH_NS.hh"

static double maxdiff(const std::vector<double> &a, const std::vector<double> &b) {
  double m = 0;
  for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::fabs(a[i] - b[i]));
  return m;
}

static std::vector<double> field(pb_sys *s, int f, uint32_t n) {
  std::vector<double> x(3 * (size_t)n);
  pbamd::check(pb_sys_get(s, f, x.data()), "get");
  return x;
}

// ---------------------------------------------------------------- MD
static void md_setup(GaussianThermo &p) {
  const real kB = 8.3145e-7f, Temp = 85.6f;
  const double L = 6 * 5.26;   // 6^3 fcc cells of argon
  real cell[6] = {0, (real)L, 0, (real)L, 0, (real)L};
  std::mt19937 eng(2);
  std::normal_distribution<> g(0.0, std::sqrt(kB * Temp / 39.95));
  std::vector<ParticleBase> G;
  const double b[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      for (int k = 0; k < 6; ++k)
        for (int q = 0; q < 4; ++q) {
          ParticleBase x{};
          x.r[0] = (i + b[q][0] + 0.25) * 5.26;
          x.r[1] = (j + b[q][1] + 0.25) * 5.26;
          x.r[2] = (k + b[q][2] + 0.25) * 5.26;
          x.m = 39.95;
          for (int c = 0; c < 3; ++c) x.v[c] = g(eng);
          G.push_back(x);
        }
  p.setE(1e-10f);
  p.setup((int)G.size());
  p.kB = kB;
  p.setCell(cell);
  p.rmax2 = 10.0f * 10.0f;
  p.import(G);
  p.setupCutoffBlock(10.0f);
  p.calcBlockID();
  p.adjustVelocities(Temp);
  p.clearForce();
}
static void md_params(GaussianThermo &p) { p.setLJparams({3.4f, 120.0f * 8.3145e-7f}, 1); }
static void md_steps(GaussianThermo &p, uint32_t k) {
  for (uint32_t j = 0; j < k; ++j) {
    p.calcBlockID();
    p.calcForce();
    p.TimeEvolution(2.0f);
    p.treatPeriodicCondition();
  }
}

// ---------------------------------------------------------------- DEM
static void dem_setup(cudaParticleDEM &p) {
  const double rmin = 0.9e-3, rmax = 1.1e-3, box[3] = {0.02, 0.02, 0.04}, rho = 7874.0;
  const double pitch = 2.0 * rmax * 1.02;
  std::mt19937_64 eng(4);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<ParticleBase> G;
  std::vector<real> r0;
  for (int k = 0; k < 16; ++k)
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 8; ++j) {
        ParticleBase x{};
        const double jit = pitch - 2.0 * rmax;
        x.r[0] = (i + 0.5) * pitch + (u(eng) - 0.5) * jit;
        x.r[1] = (j + 0.5) * pitch + (u(eng) - 0.5) * jit;
        x.r[2] = (k + 0.5) * pitch + (u(eng) - 0.5) * jit;
        const double rad = rmin + (rmax - rmin) * u(eng);
        x.m = rho * 4.0 / 3.0 * M_PI * rad * rad * rad;
        for (int c = 0; c < 3; ++c) x.v[c] = 0.05 * (u(eng) - 0.5);
        G.push_back(x);
        r0.push_back((real)rad);
      }
  real cell[6] = {0, (real)box[0], 0, (real)box[1], 0, (real)box[2]};
  p.switchBlockAlgorithm(ParticleBlockType::many);
  p.setup((int)G.size());
  p.setCell(cell);
  p.import(G);
  p.autotunetimestep = true;
  p.setDEMProperties(2.11e9f, 0.5f, 0.29f, 0.14f, 0.1f, r0);
  p.setInertia(r0);
  p.setupCutoffBlock((real)(2.0 * rmax), false);
  p.calcBlockID();
  p.calcVinit(2e-6f);
}
static void dem_steps(cudaParticleDEM &p, uint32_t k) {
  for (uint32_t j = 0; j < k; ++j) {
    p.calcBlockID();
    p.calcForce(2e-6f);
    p.calcAcceleration();
    p.addAccelerationZ(-9.8f);
    p.TimeEvolution(2e-6f);
    p.treatRefrectCondition();
  }
}

// ---------------------------------------------------------------- SPH
template <bool config>
static void sph_setup(cudaParticleSPH_NS &p) {
  const double dx = 0.05, h = (config ? 1.3 : 1.7) * dx, c0 = 10.0 * std::sqrt(2.0 * 9.81 * 0.4);
  const double mass = (config ? 1000.0 : 1.0) * dx * dx * dx;
  std::vector<ParticleBase> G;
  auto add = [&](double x, double y, double z, bool fixed) {
    ParticleBase q{};
    q.r[0] = x;
    q.r[1] = y;
    q.r[2] = z;
    q.m = mass;
    q.isFixed = fixed;
    q.type = fixed ? 1 : 0;
    G.push_back(q);
  };
  for (int i = -3; i < 16; ++i)
    for (int j = -3; j < 8; ++j)
      for (int k = -3; k < 8; ++k)
        if (!(i >= 0 && j >= 0 && j < 5 && k >= 0)) {
          if (i < 13 && j < 8) add((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx, true);
        }
  const size_t nwall = G.size();
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 5; ++j)
      for (int k = 0; k < 8; ++k) add((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx, false);
  const real lo = (real)(-3 * dx);
  real cell[6] = {lo, (real)(13 * dx), lo, (real)(8 * dx), lo, (real)(8 * dx)};
  std::valarray<real> mu(1.0e-6f, G.size()), c1((real)c0, G.size());
  mu[std::slice(0, nwall, 1)] = 1.0e-3f;
  p.setup((int)G.size());
  p.setCell(cell);
  p.import(G);
  if (config) {
    p.setWCSPHProperties(h, 1000.0, c0, 7.0, 0.1, 0.0);
    p.setupCutoffBlock((real)(2 * h), false);
  } else {
    p.setSPHProperties(mu, c1, (real)h);
    p.setupCutoffBlock((real)h, false);
  }
  p.calcBlockID();
  p.calcVinit((real)(0.25 * h / c0));
}
template <bool config>
static void sph_steps(cudaParticleSPH_NS &p, uint32_t k) {
  const double dx = 0.05, h = (config ? 1.3 : 1.7) * dx, c0 = 10.0 * std::sqrt(2.0 * 9.81 * 0.4);
  const real dt = (real)(0.25 * h / c0);
  for (uint32_t j = 0; j < k; ++j) {
    p.calcBlockID();
    p.calcDensity(true);
    p.calcAcceleration(true);
    p.addAccelerationZ(-9.81f);
    p.TimeEvolution(dt);
    p.treatAbsoluteCondition();
  }
}

template <typename T, typename Setup, typename Params, typename Steps>
static int check(const char *mode, uint32_t k, const std::string &fn, Setup setup, Params params,
                 Steps steps, const std::vector<int> &fields) {
  std::vector<std::vector<double>> ref, got;
  uint32_t n = 0;
  {
    CUDAenv<T> env;
    env.setup();
    T &p = env[0];
    setup(p);
    params(p);
    steps(p, k);
    env.writeSerialization(fn);
    steps(p, k);
    n = p.size();
    for (int f : fields) ref.push_back(field(p.handle(), f, n));
  }
  {
    CUDAenv<T> env;
    env.setup();
    T &p = env[0];
    env.readSerialization(fn);
    params(p);               // LJ parameters are not part of the archive (as in the reference)
    steps(p, k);
    for (int f : fields) got.push_back(field(p.handle(), f, n));
  }
  double worst = 0;
  for (size_t i = 0; i < fields.size(); ++i) worst = std::max(worst, maxdiff(ref[i], got[i]));
  std::cout << "done mode=" << mode << " N=" << n << " k=" << k << " maxdiff=" << worst << std::endl;
  return 0;
}

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "md";
  const uint32_t k = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 20;
  const std::string fn = argc > 3 ? argv[3] : "restart.ck";
  if (mode == "md")
    return check<GaussianThermo>("md", k, fn, md_setup, md_params, md_steps,
                                 {PB_FIELD_R, PB_FIELD_V, PB_FIELD_A});
  if (mode == "dem")
    return check<cudaParticleDEM>("dem", k, fn, dem_setup, [](cudaParticleDEM &) {}, dem_steps,
                                  {PB_FIELD_R, PB_FIELD_V, PB_FIELD_W, PB_FIELD_L});
  if (mode == "sph")
    return check<cudaParticleSPH_NS>("sph", k, fn, sph_setup<false>, [](cudaParticleSPH_NS &) {},
                                     sph_steps<false>, {PB_FIELD_R, PB_FIELD_V});
  if (mode == "sphc")
    return check<cudaParticleSPH_NS>("sphc", k, fn, sph_setup<true>, [](cudaParticleSPH_NS &) {},
                                     sph_steps<true>, {PB_FIELD_R, PB_FIELD_V});
  std::cerr << "mode: md | dem | sph | sphc" << std::endl;
  return 2;
}
