// md_gas.cpp -- N2/O2 gas at 353 K in the set-up and call order of the
// reference's testMD1.cu:14-170: units of Angstrom, fs and g/NA; 800 A box
// at 2.074e-5 molecules/A^3 (10,618 molecules, 20 % O2) at uniform random
// positions; two-type LJ, cutoff 40 A, cells of 80 A; adjustVelocities,
// initialAnnealing (soft core to 5 A, then LJ, both under constTemp), then
// per step calcBlockID, calcForce, calcTemp (abort above 1.2 T, scaleTemp
// above 1.1 T), TimeEvolution, treatPeriodicCondition.  "nlist" computes the
// main-loop forces from a neighbour list (useNList / makeNList every
// 10 steps with a 10 A thickness) instead of the cells.
//
//   md_gas [steps=200] [anneal=1000] [cells|nlist]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <random>
#include <vector>
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleMD.hh"

int main(int argc, char **argv) {
  const uint32_t steps = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 200;
  const uint32_t anneal = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 1000;
  const bool nlist = argc > 3 && std::strcmp(argv[3], "nlist") == 0;
  const real kB = 8.3145e-7f;
  const real Temp = 353.0f;
  real cell[6] = {0.0f, 800.0f, 0.0f, 800.0f, 0.0f, 800.0f};
  const double rho = 0.00002074;
  const uint32_t N = (uint32_t)(800.0 * 800.0 * 800.0 * rho);

  std::mt19937 engine;
  std::uniform_real_distribution<> s0(0.0, 1.0), s1(0.0, 800.0), s2(0.0, 800.0), s3(0.0, 800.0);
  std::vector<ParticleBase> G;
  for (uint32_t i = 0; i < N; ++i) {
    ParticleBase pb{};
    pb.r[0] = s1(engine);
    pb.r[1] = s2(engine);
    pb.r[2] = s3(engine);
    if (s0(engine) < 0.2) {
      pb.m = 32.0;   // O2
      pb.type = 0;
    } else {
      pb.m = 28.0;   // N2
      pb.type = 1;
    }
    std::normal_distribution<> t1(0.0, std::sqrt(kB * Temp / pb.m));
    for (int c = 0; c < 3; ++c) pb.v[c] = t1(engine);
    G.push_back(pb);
  }

  CUDAenv<cudaParticleMD> env;
  env.setup();
  cudaParticleMD &particles = env[0];
  particles.setup(N);
  particles.kB = kB;
  particles.setCell(cell);
  particles.timestep = 0;
  particles.rmax2 = 40.0f * 40.0f;
  particles.import(G);
  particles.setM();
  particles.setLJparams({3.1062f, 3.59044e-5f, 3.21365f, 4.82095e-5f,
                         3.21365f, 4.82095e-5f, 3.3211f, 4.19355e-5f}, 2);

  particles.setupCutoffBlock(80.0f);
  particles.calcBlockID();
  particles.adjustVelocities(Temp, true);
  particles.initialAnnealing(anneal, 1.0f, 5.0f, 1e-1f, Temp);
  const double t_anneal = particles.calcTemp();

  if (nlist) {
    particles.thickness = 10.0f;
    particles.useNList();
  }
  const real dt = 5.0f;
  double tmin = 1e30, tmax = 0;
  uint32_t scaled = 0;
  particles.clearForce();
  // MAIN LOOP (testMD1.cu:134-170)
  for (uint32_t j = 0; j < steps; ++j) {
    particles.calcBlockID();
    if (nlist && j % 10 == 0) particles.makeNList();
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
    tmin = std::min(tmin, (double)t);
    tmax = std::max(tmax, (double)t);
    if ((j + 1) % 25 == 0)
      std::cerr << "K-P\t" << particles.calcKineticE() << "\t" << particles.calcPotentialE() << std::endl;
  }
  // minimum pair distance of the final state (the annealing separated the
  // random start)
  particles.getPosition();
  std::vector<double> r(3 * (size_t)N);
  pbamd::check(pb_sys_get(particles.handle(), PB_FIELD_R, r.data()), "get r");
  double dmin = 1e30;
  {
    // cell-list search over 10 A cells for the closest pair
    const int nc = 80;
    std::vector<std::vector<uint32_t>> cl((size_t)nc * nc * nc);
    auto ci = [&](double x) { int c = (int)(x / 10.0); return std::min(std::max(c, 0), nc - 1); };
    for (uint32_t i = 0; i < N; ++i)
      cl[((size_t)ci(r[3 * i]) * nc + ci(r[3 * i + 1])) * nc + ci(r[3 * i + 2])].push_back(i);
    for (uint32_t i = 0; i < N; ++i) {
      const int a = ci(r[3 * i]), b = ci(r[3 * i + 1]), c = ci(r[3 * i + 2]);
      for (int da = -1; da <= 1; ++da)
        for (int db = -1; db <= 1; ++db)
          for (int dc = -1; dc <= 1; ++dc) {
            const size_t k = ((size_t)((a + da + nc) % nc) * nc + (b + db + nc) % nc) * nc +
                             (c + dc + nc) % nc;
            for (uint32_t j : cl[k]) {
              if (j <= i) continue;
              double d2 = 0;
              for (int q = 0; q < 3; ++q) {
                double d = r[3 * i + q] - r[3 * j + q];
                d -= 800.0 * std::nearbyint(d / 800.0);
                d2 += d * d;
              }
              dmin = std::min(dmin, std::sqrt(d2));
            }
          }
    }
  }
  std::cout << "done N=" << N << " steps=" << steps << " Tanneal=" << t_anneal << " Tmin=" << tmin
            << " Tmax=" << tmax << " scaled=" << scaled << " dmin=" << dmin << std::endl;
  return 0;
}
