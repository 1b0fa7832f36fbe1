// sph_dam.cpp -- SPH dam break driver in the call order of the reference's
// testSPH5.cu:193-258 (calcBlockID, calcKernels, calcDensity,
// calcDensityPost, calcForce, calcAcceleration, RestoreAcceleration,
// addAccelerationZ, TimeEvolution, inspectVelocity, AdaptiveTime::Progress,
// treatAbsoluteCondition) with the reference physics (Lucy kernel, support
// h = 1.7 dx, p = c^2 (rho - 1)).  A fluid column of spacing dx in the
// corner of a box lined with three layers of fixed wall particles; relative
// density units (every particle has mass dx^3).
//
//   sph_dam [dx=0.04] [steps=200]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <valarray>
#include <vector>
#include "../include/AdaptiveTime.hh"
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleSPH_NS.hh"

int main(int argc, char **argv) {
  const double dx = argc > 1 ? std::atof(argv[1]) : 0.04;
  const uint32_t steps = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 200;
  const double box[3] = {1.6, 0.4, 0.8}, col[3] = {0.4, 0.4, 0.8};
  const int layers = 3;
  const double g = 9.81, h = 1.7 * dx;
  const double c0 = 10.0 * std::sqrt(2.0 * g * col[2]);

  std::vector<ParticleBase> G;
  auto add = [&](double x, double y, double z, bool fixed) {
    ParticleBase p{};
    p.r[0] = x;
    p.r[1] = y;
    p.r[2] = z;
    p.m = dx * dx * dx;
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
  particles.setSPHProperties(mu, c1, (real)h);
  particles.setupCutoffBlock((real)h, false);
  particles.calcBlockID();
  particles.calcDensity();
  std::cerr << "N= " << N << " walls= " << nwall << " h= " << h << " mean num density= "
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
    res[0] = particles.inspectVelocity((real)(2 * h) / thistime.currentDeltaT(), ulim, llim, _r,
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
