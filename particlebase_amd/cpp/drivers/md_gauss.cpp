// md_gauss.cpp -- liquid argon under the Gaussian isokinetic thermostat, in
// the set-up and call order of the reference's testMD2.cu:14-175: units of
// Angstrom, fs and g/NA; 60 A box at 0.0228 atoms/A^3 on a simple lattice,
// LJ sigma = 3.4 A, epsilon = 120 kB, cutoff 15 A, T = 85.6 K, dt = 2 fs;
// adjustVelocities, then per step calcBlockID, calcForce, calcTemp (abort
// above 1.2 T, scaleTemp above 1.1 T), TimeEvolution,
// treatPeriodicCondition.  Prints T every 50 steps and a summary line.
//
//   md_gauss [steps=500]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <random>
#include <vector>
#include "../include/CUDAenv.hh"
#include "../include/GaussianThermo.hh"

int main(int argc, char **argv) {
  const uint32_t steps = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 500;
  const real kB = 8.3145e-7f;   // kB in A, fs, g/NA units
  const real Temp = 85.6f;
  real cell[6] = {0.0f, 60.0f, 0.0f, 60.0f, 0.0f, 60.0f};
  const double rho = 0.0228;
  const double vol = 60.0 * 60.0 * 60.0;
  const uint32_t N = (uint32_t)(vol * rho);
  const double l0 = std::cbrt(vol / N);
  const uint32_t n1 = (uint32_t)(60.0 / l0) + 1;
  const double l1 = 60.0 / n1;

  std::mt19937 engine;
  std::vector<ParticleBase> G;
  for (uint32_t i = 0; i < N; ++i) {
    ParticleBase pb{};
    pb.r[0] = l1 * (i % n1 + 0.5);
    pb.r[1] = l1 * ((i / n1) % n1 + 0.5);
    pb.r[2] = l1 * (i / (n1 * n1) + 0.5);
    pb.m = 39.95;
    pb.type = 0;
    std::normal_distribution<> t1(0.0, std::sqrt(kB * Temp / pb.m));
    for (int c = 0; c < 3; ++c) pb.v[c] = t1(engine);
    G.push_back(pb);
  }

  CUDAenv<GaussianThermo> env;
  env.setup();
  GaussianThermo &particles = env[0];
  particles.setE(1e-10f);
  particles.setup(N);
  particles.kB = kB;
  particles.setCell(cell);
  particles.timestep = 0;
  particles.rmax2 = 15.0f * 15.0f;
  particles.import(G);
  particles.setM();
  particles.setLJparams({3.4f, 120.0f * kB}, 1);

  particles.setupCutoffBlock(15.0f);
  particles.calcBlockID();
  particles.adjustVelocities(Temp, true);

  const real dt = 2.0f;
  double tmin = 1e30, tmax = 0, t0 = 0, t1 = 0;
  uint32_t scaled = 0;
  particles.clearForce();
  // MAIN LOOP (testMD2.cu:149-175)
  for (uint32_t j = 0; j < steps; ++j) {
    particles.calcBlockID();
    particles.calcForce();
    const real t = particles.calcTemp();
    if (t > Temp * 1.2f || std::isnan(t)) {
      std::cerr << "abort T= " << t << std::endl;
      return 1;
    }
    if (t > Temp * 1.1f) {
      std::cerr << "scaling temp by: " << particles.scaleTemp(Temp) << std::endl;
      ++scaled;
    }
    particles.TimeEvolution(dt);
    particles.treatPeriodicCondition();
    if (j == 2) t0 = t;
    if (j >= 2) {
      tmin = std::min(tmin, (double)t);
      tmax = std::max(tmax, (double)t);
    }
    t1 = t;
    if ((j + 1) % 50 == 0) std::cerr << "T= " << t << std::endl;
  }
  std::cout << "done N=" << N << " steps=" << steps << " T0=" << t0 << " T1=" << t1
            << " Tmin=" << tmin << " Tmax=" << tmax << " scaled=" << scaled
            << " xi=" << particles.getXi() << " failed=" << particles.failedSteps() << std::endl;
  return 0;
}
