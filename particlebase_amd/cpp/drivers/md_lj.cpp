// md_lj.cpp -- LJ fluid driver in the call order of the reference's
// testMD1.cu:133-169 (calcBlockID, calcForce, calcTemp, TimeEvolution,
// treatPeriodicCondition; K-P every 25 steps), in reduced units:
// FCC n^3 x 4 atoms at density 0.8442, T = 0.72, rc = 2.5 (shifted
// energies), dt = 0.005.  "fused" runs the same steps as the engine's fused
// loop.  Prints the energies and, at the end, the energy drift.
//
//   md_lj [ncell=10] [steps=1000] [api|fused] [fp64|mixed]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <vector>
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleMD.hh"

static std::vector<ParticleBase> fcc(int nc, double rho, double T, uint64_t seed, double &L) {
  const double a = std::cbrt(4.0 / rho);
  L = a * nc;
  const double b[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  std::mt19937_64 eng(seed);
  std::normal_distribution<double> g(0.0, std::sqrt(T));
  std::vector<ParticleBase> P;
  double p[3] = {0, 0, 0};
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j)
      for (int k = 0; k < nc; ++k)
        for (int q = 0; q < 4; ++q) {
          ParticleBase x{};
          x.r[0] = (i + b[q][0]) * a;
          x.r[1] = (j + b[q][1]) * a;
          x.r[2] = (k + b[q][2]) * a;
          x.m = 1.0;
          for (int c = 0; c < 3; ++c) {
            x.v[c] = g(eng);
            p[c] += x.v[c];
          }
          x.type = 0;
          P.push_back(x);
        }
  // zero momentum, exact T with 3N-3 degrees of freedom
  double mv2 = 0;
  for (auto &x : P)
    for (int c = 0; c < 3; ++c) {
      x.v[c] -= p[c] / P.size();
      mv2 += x.v[c] * x.v[c];
    }
  const double s = std::sqrt(T * (3.0 * P.size() - 3.0) / mv2);
  for (auto &x : P)
    for (int c = 0; c < 3; ++c) x.v[c] *= s;
  return P;
}

int main(int argc, char **argv) {
  const int nc = argc > 1 ? std::atoi(argv[1]) : 10;
  const uint32_t steps = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 1000;
  const bool fused = argc > 3 && std::strcmp(argv[3], "fused") == 0;
  const bool mixed = argc > 4 && std::strcmp(argv[4], "mixed") == 0;
  const real dt = 0.005f;

  CUDAenv<cudaParticleMD> env;
  env.setup();
  cudaParticleMD &particles = env[0];
  double L = 0;
  std::vector<ParticleBase> G = fcc(nc, 0.8442, 0.72, 1, L);
  const uint32_t N = (uint32_t)G.size();
  real cell[6] = {0, (real)L, 0, (real)L, 0, (real)L};
  particles.setup(N);
  particles.kB = 1.0;
  particles.setCell(cell);
  particles.timestep = 0;
  particles.rmax2 = 2.5f * 2.5f;
  particles.shifted = true;
  particles.precision = mixed ? PB_MIXED : PB_FP64;
  particles.import(G);
  particles.setM();
  particles.setLJparams({1.0f, 1.0f}, 1);
  particles.setupCutoffBlock(2.5f);
  particles.calcBlockID();

  std::cerr << "N= " << N << " L= " << L << " " << (fused ? "fused" : "api") << " "
            << (mixed ? "mixed" : "fp64") << std::endl;
  double e0 = 0, e1 = 0;
  particles.clearForce();
  if (fused) {
    std::vector<double> e = particles.runFused(steps, dt, particles.precision, 0.3, 20, 25);
    for (size_t r = 0; r < e.size() / 3; ++r) {
      std::cerr << "K-P\t" << e[3 * r] << "\t" << e[3 * r + 1] << std::endl;
      const double tot = (e[3 * r] + e[3 * r + 1]) / N;
      if (r == 0) e0 = tot;
      e1 = tot;
    }
  } else {
    // MAIN LOOP (testMD1.cu:135-169)
    for (uint32_t j = 0; j < steps; ++j) {
      particles.calcBlockID();
      particles.calcForce();
      const real t = particles.calcTemp();   // of the velocities before the step, as there
      particles.TimeEvolution(dt);
      particles.treatPeriodicCondition();
      if ((j + 1) % 25 == 0) {
        const double ke = particles.calcKineticE(), pe = particles.calcPotentialE();
        std::cerr << "K-P\t" << ke << "\t" << pe << "\tT= " << t << std::endl;
        const double tot = (ke + pe) / N;
        if (j + 1 == 25) e0 = tot;
        e1 = tot;
      }
    }
  }
  // energies per particle at the first and last record
  std::cout << "done N=" << N << " steps=" << steps << " E0=" << e0 << " E1=" << e1
            << " dE=" << (e1 - e0) << std::endl;
  return 0;
}
