// sph_dam.cpp -- SPH dam break driver in the call order of the reference's
// testSPH5.cu:193-258 (calcBlockID, calcKernels, calcDensity,
// calcDensityPost, calcForce, calcAcceleration, RestoreAcceleration,
// addAccelerationZ, TimeEvolution, inspectVelocity, AdaptiveTime::Progress,
// treatAbsoluteCondition).  A fluid column of spacing dx in the corner of a
// box lined with three layers of fixed wall particles.
//   reference physics (default): Lucy kernel of support h = 1.7 dx,
//     p = c^2 (rho - 1), relative density units (mass dx^3);
//   config physics ("config"): M4 cubic spline of support 2h, h = 1.3 dx,
//     Tait (rho0 = 1000, gamma = 7, c0 = 10 sqrt(2 g H)), artificial
//     viscosity alpha = 0.1, SI units; the mass is rho0 dx^3 over the
//     kernel's lattice sum, so the initial lattice is at rest density.
//
//   sph_dam [dx=0.04] [steps=200] [reference|config]
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <string>
#include <valarray>
#include <vector>
#include "../include/AdaptiveTime.hh"
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleSPH_NS.hh"

int main(int argc, char **argv) {
  const double dx = argc > 1 ? std::atof(argv[1]) : 0.04;
  const uint32_t steps = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 200;
  const bool config = argc > 3 && std::string(argv[3]) == "config";
  const double box[3] = {1.6, 0.4, 0.8}, col[3] = {0.4, 0.4, 0.8};
  const int layers = 3;
  const double hfac = config ? 1.3 : 1.7;
  const double g = 9.81, h = hfac * dx, sup = config ? 2.0 * h : h, rho0 = 1000.0;
  const double c0 = 10.0 * std::sqrt(2.0 * g * col[2]);
  double mass = dx * dx * dx;
  if (config) {
    // sum of W_M4(|n| dx) dx^3 over the cubic lattice
    double S = 0.0;
    const int k = (int)std::ceil(2.0 * hfac);
    for (int a = -k; a <= k; ++a)
      for (int b = -k; b <= k; ++b)
        for (int c = -k; c <= k; ++c) {
          const double q = std::sqrt((double)(a * a + b * b + c * c)) / hfac;
          const double t1 = std::max(1.0 - q, 0.0), t2 = std::max(2.0 - q, 0.0);
          S += 0.25 * t2 * t2 * t2 - t1 * t1 * t1;
        }
    S /= M_PI * hfac * hfac * hfac;
    mass = rho0 * dx * dx * dx / S;
  }

  std::vector<ParticleBase> G;
  auto add = [&](double x, double y, double z, bool fixed) {
    ParticleBase p{};
    p.r[0] = x;
    p.r[1] = y;
    p.r[2] = z;
    p.m = mass;
    p.isFixed = fixed;
    p.type = fixed ? 1 : 0;
    G.push_back(p);
  };
  const int wx = (int)std::lround(box[0] / dx), wy = (int)std::lround(box[1] / dx),
            wz = (int)std::lround(box[2] / dx);
  for (int i = -layers; i < wx + layers; ++i)
    for (int j = -layers; j < wy + layers; ++j)
      for (int k = -layers; k < wz; ++k)
        if (!(i >= 0 && i < wx && j >= 0 && j < wy && k >= 0))
          add((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx, true);
  const size_t nwall = G.size();
  const int nx = (int)std::lround(col[0] / dx), ny = (int)std::lround(col[1] / dx),
            nz = (int)std::lround(col[2] / dx);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) add((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx, false);
  const uint32_t N = (uint32_t)G.size();

  CUDAenv<cudaParticleSPH_NS> env;
  env.setup();
  cudaParticleSPH_NS &particles = env[0];
  const real lo = (real)(-layers * dx);
  real cell[6] = {lo, (real)(box[0] - lo), lo, (real)(box[1] - lo), lo,
                  (real)(box[2] + layers * dx)};
  std::valarray<real> mu(1.0e-6f, N), c1((real)c0, N);
  mu[std::slice(0, nwall, 1)] = 1.0e-3f;
  particles.setup(N);
  particles.setCell(cell);
  particles.import(G);
  particles.timestep = 0;
  if (config)
    particles.setWCSPHProperties(h, rho0, c0, 7.0, 0.1, 0.0);
  else
    particles.setSPHProperties(mu, c1, (real)h);
  particles.setupCutoffBlock((real)sup, false);
  particles.calcBlockID();
  particles.calcDensity();
  std::cerr << (config ? "config" : "reference") << " physics N= " << N << " walls= " << nwall
            << " h= " << h << " mean num density= "
            << particles.inspectDensity() << std::endl;

  const real initDeltaT = (real)(0.25 * h / c0);       // CFL 0.25
  const real ulim = 0.01f * 4, llim = ulim / 16.0f;
  std::vector<int> res(1);
  AdaptiveTime<CUDAenv<cudaParticleSPH_NS>> thistime;
  thistime.init(initDeltaT);
  particles.calcVinit(initDeltaT);
  uint32_t j = 0, iter = 0;
  while (j < steps && iter < 4 * steps) {
    ++iter;
    particles.calcBlockID();
    particles.calcKernels();
    particles.calcDensity(true);
    particles.calcDensityPost(true);
    particles.selectBlocks();
    particles.calcForce();
    particles.calcAcceleration(true);
    particles.RestoreAcceleration();
    particles.addAccelerationZ((real)-g);
    particles.TimeEvolution(thistime.currentDeltaT());
    real _r = 0;
    uint32_t _r1 = 0;
    res[0] = particles.inspectVelocity((real)(2 * sup) / thistime.currentDeltaT(), ulim, llim, _r,
                                       _r1);
    j = thistime.Progress(env, res, j);
    particles.treatAbsoluteCondition();
  }
  particles.getPosition();
  // front position of the fluid (largest x of a movable particle)
  std::vector<double> r = particles.getState(PB_FIELD_R);
  double front = 0;
  for (size_t i = nwall; i < N; ++i) front = std::max(front, r[3 * i]);
  std::cout << "done N=" << N << " steps=" << j << " t=" << thistime() << " dt="
            << thistime.currentDeltaT() << " front_x=" << front
            << " mean_num=" << particles.inspectDensity() << std::endl;
  return 0;
}
