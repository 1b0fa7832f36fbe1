// dem_settle.cpp -- granular settle driver in the call order of the
// reference's testDEM5.cu:181-234 (calcBlockID, selectBlocks, calcForce(dt),
// calcAcceleration, addAccelerationZ, TimeEvolution, inspectVelocity,
// AdaptiveTime::Progress, treatRefrectCondition) with the reference's Hertz
// contact law: n spheres of radius U[0.9, 1.1] mm at non-overlapping
// jittered-lattice sites in a 0.1 x 0.1 x 0.4 m box, steel-like density,
// gravity, reflecting walls, many-particle cells of 2 r_max.
//
//   dem_settle [n=20000] [steps=500]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <random>
#include <vector>
#include "../include/AdaptiveTime.hh"
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleDEM.hh"

int main(int argc, char **argv) {
  const uint32_t N = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000;
  const uint32_t steps = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 500;
  const double rmin = 0.9e-3, rmax = 1.1e-3, box[3] = {0.1, 0.1, 0.4}, rho = 7874.0;
  const double pitch = 2.0 * rmax * 1.05;
  const int nx = (int)(box[0] / pitch), ny = (int)(box[1] / pitch), nz = (int)(box[2] / pitch);
  if ((uint64_t)nx * ny * nz < N) {
    std::cerr << "too many spheres for the box" << std::endl;
    return 1;
  }
  std::mt19937_64 eng(4);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<ParticleBase> G;
  std::vector<real> r0;
  for (int k = 0; k < nz && G.size() < N; ++k)
    for (int i = 0; i < nx && G.size() < N; ++i)
      for (int j = 0; j < ny && G.size() < N; ++j) {
        ParticleBase p{};
        const double jit = pitch - 2.0 * rmax;
        p.r[0] = (i + 0.5) * pitch + (u(eng) - 0.5) * jit;
        p.r[1] = (j + 0.5) * pitch + (u(eng) - 0.5) * jit;
        p.r[2] = (k + 0.5) * pitch + (u(eng) - 0.5) * jit;
        const double rad = rmin + (rmax - rmin) * u(eng);
        p.m = rho * 4.0 / 3.0 * M_PI * rad * rad * rad;
        p.isFixed = false;
        G.push_back(p);
        r0.push_back((real)rad);
      }

  CUDAenv<cudaParticleDEM> env;
  env.setup();
  cudaParticleDEM &particles = env[0];
  real cell[6] = {0, (real)box[0], 0, (real)box[1], 0, (real)box[2]};
  particles.switchBlockAlgorithm(ParticleBlockType::many);
  particles.setup(N);
  particles.setCell(cell);
  particles.import(G);
  particles.autotunetimestep = true;
  const double e = 0.8;
  const double gamma = -std::log(e) / std::sqrt(M_PI * M_PI + std::log(e) * std::log(e)) * 2;
  particles.setDEMProperties(2.11e9f, 0.5f, 0.29f, (real)gamma, 0.1f, r0);
  particles.setInertia(r0);
  particles.setupCutoffBlock((real)(2.0 * rmax), false);
  particles.checkPidRange();       // testDEM.cu:93
  particles.calcBlockID();

  const real initDeltaT = 2e-6f, ulim = 0.01f * 4, llim = ulim / 16.0f;
  std::vector<int> res(1);
  AdaptiveTime<CUDAenv<cudaParticleDEM>> thistime;
  thistime.init(initDeltaT);
  particles.calcVinit(initDeltaT);
  uint32_t j = 0, iter = 0, maxc = 0;
  while (j < steps && iter < 4 * steps) {
    ++iter;
    particles.calcBlockID();
    particles.selectBlocks();
    particles.calcForce(thistime.currentDeltaT());
    maxc = std::max(maxc, particles.numContacts());
    particles.calcAcceleration();
    particles.addAccelerationZ(-9.8f);
    particles.TimeEvolution(thistime.currentDeltaT());
    real _r = 0;
    uint32_t _r1 = 0;
    res[0] = particles.inspectVelocity((real)(2 * rmin) / thistime.currentDeltaT(), ulim, llim,
                                       _r, _r1);
    j = thistime.Progress(env, res, j);
    particles.treatRefrectCondition();
  }
  std::vector<double> r = particles.getState(PB_FIELD_R);
  double zmean = 0;
  for (uint32_t i = 0; i < N; ++i) zmean += r[3 * i + 2];
  std::cout << "done N=" << N << " steps=" << j << " t=" << thistime() << " dt="
            << thistime.currentDeltaT() << " max_contacts=" << maxc << " zmean=" << zmean / N
            << " selected_cells=" << particles.numSelectedBlocks()
            << " cells=" << particles.numBlocks() << std::endl;
  return 0;
}
