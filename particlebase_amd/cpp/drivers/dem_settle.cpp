// dem_settle.cpp -- granular settle driver in the call order of the
// reference's testDEM5.cu:181-234 (calcBlockID, selectBlocks, calcForce(dt),
// calcAcceleration, addAccelerationZ, TimeEvolution, inspectVelocity,
// AdaptiveTime::Progress, treatRefrectCondition): spheres of radius
// U[0.9, 1.1] mm in a 0.1 x 0.1 x 0.4 m box, steel-like density, gravity,
// reflecting walls, many-particle cells of 2 r_max.
//
//   dem_settle [n=20000] [steps=500] [law=hertz|linear] [input]
//
// law hertz: the reference's Hertz contact law (setDEMProperties) on
// non-overlapping jittered-lattice sites.  law linear: the granular settle
// workload's linear spring-dashpot law (kn = 1e4 N/m, restitution 0.8,
// tangential spring kt = 2/7 kn with Coulomb mu = 0.5, setLinearContact).
// Under the linear law the adaptive step is also held at a tenth of the
// explicit update's stability bound 2 sqrt(meff / kn) for the lightest
// pair, i.e. at 0.2 sqrt(meff_min / kn) (about a sixteenth of its contact
// time): a step at that cap is not reported as slow, so AdaptiveTime does
// not grow it further.
// input (optional): binary file, uint32 n, then FP64 positions [3n], radii
// [n], masses [n], velocities [3n] (the workload's random non-overlapping
// insertion, written by the caller); n then comes from the file.
// Prints the steps, the timed loop's wall time and particle-steps/s.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>
#include "../include/AdaptiveTime.hh"
#include "../include/CUDAenv.hh"
#include "../include/cudaParticleDEM.hh"

int main(int argc, char **argv) {
  uint32_t N = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000;
  const uint32_t steps = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 500;
  const std::string law = argc > 3 ? argv[3] : "hertz";
  const char *input = argc > 4 ? argv[4] : nullptr;
  if (law != "hertz" && law != "linear") {
    std::cerr << "law must be hertz or linear" << std::endl;
    return 1;
  }
  const double rmin = 0.9e-3, rmax = 1.1e-3, box[3] = {0.1, 0.1, 0.4}, rho = 7874.0;
  std::vector<ParticleBase> G;
  std::vector<real> r0;
  if (input) {
    std::ifstream f(input, std::ios::binary);
    uint32_t n = 0;
    f.read((char *)&n, 4);
    std::vector<double> r(3 * (size_t)n), rad(n), m(n), v(3 * (size_t)n);
    f.read((char *)r.data(), 8 * r.size());
    f.read((char *)rad.data(), 8 * rad.size());
    f.read((char *)m.data(), 8 * m.size());
    f.read((char *)v.data(), 8 * v.size());
    if (!f || n == 0) {
      std::cerr << "cannot read " << input << std::endl;
      return 1;
    }
    N = n;
    for (uint32_t i = 0; i < n; ++i) {
      ParticleBase p{};
      for (int a = 0; a < 3; ++a) {
        p.r[a] = r[3 * i + a];
        p.v[a] = v[3 * i + a];
      }
      p.m = m[i];
      p.isFixed = false;
      G.push_back(p);
      r0.push_back((real)rad[i]);
    }
  } else {
    const double pitch = 2.0 * rmax * 1.05;
    const int nx = (int)(box[0] / pitch), ny = (int)(box[1] / pitch), nz = (int)(box[2] / pitch);
    if ((uint64_t)nx * ny * nz < N) {
      std::cerr << "too many spheres for the box" << std::endl;
      return 1;
    }
    std::mt19937_64 eng(4);
    std::uniform_real_distribution<double> u(0.0, 1.0);
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
  const double e = 0.8, kn = 1.0e4;
  double dtmax = 1e30;
  if (law == "linear") {
    particles.setLinearContact((real)kn, (real)e, (real)(kn * 2.0 / 7.0), 0.5f, r0);
    double mmin = 1e30;
    for (const auto &p : G) mmin = std::min(mmin, p.m);
    dtmax = 0.2 * std::sqrt(0.5 * mmin / kn);
  } else {
    const double gamma = -std::log(e) / std::sqrt(M_PI * M_PI + std::log(e) * std::log(e)) * 2;
    particles.setDEMProperties(2.11e9f, 0.5f, 0.29f, (real)gamma, 0.1f, r0);
  }
  particles.setInertia(r0);
  real rcut = 0;
  for (real x : r0) rcut = std::max(rcut, x);
  particles.setupCutoffBlock(2 * rcut, false);
  particles.checkPidRange();       // testDEM.cu:93
  particles.calcBlockID();

  const real initDeltaT = 2e-6f, ulim = 0.01f * 4, llim = ulim / 16.0f;
  std::vector<int> res(1);
  AdaptiveTime<CUDAenv<cudaParticleDEM>> thistime;
  thistime.init(initDeltaT);
  particles.calcVinit(initDeltaT);
  uint32_t j = 0, iter = 0, maxc = 0, rollbacks = 0;
  double sumc = 0;
  const double growth = std::sqrt(2.0);
  pbamd::check(pb_ctx_sync(pbamd::context()), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  while (j < steps && iter < 4 * steps) {
    ++iter;
    particles.calcBlockID();
    particles.selectBlocks();
    particles.calcForce(thistime.currentDeltaT());
    maxc = std::max(maxc, particles.numContacts());
    sumc += particles.numContacts();
    particles.calcAcceleration();
    particles.addAccelerationZ(-9.8f);
    particles.TimeEvolution(thistime.currentDeltaT());
    real _r = 0;
    uint32_t _r1 = 0;
    res[0] = particles.inspectVelocity((real)(2 * rmin) / thistime.currentDeltaT(), ulim, llim,
                                       _r, _r1);
    if (res[0] == -1 && thistime.currentDeltaT() * growth > dtmax) res[0] = 0;
    if (res[0] == 1) ++rollbacks;
    j = thistime.Progress(env, res, j);
    particles.treatRefrectCondition();
  }
  pbamd::check(pb_ctx_sync(pbamd::context()), "sync");
  const double el =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::vector<double> r = particles.getState(PB_FIELD_R);
  double zmean = 0;
  for (uint32_t i = 0; i < N; ++i) zmean += r[3 * i + 2];
  std::cout << "done N=" << N << " steps=" << j << " iterations=" << iter
            << " rollbacks=" << rollbacks << " t=" << thistime() << " dt="
            << thistime.currentDeltaT() << " max_contacts=" << maxc
            << " mean_contacts=" << sumc / std::max(iter, 1u) << " zmean=" << zmean / N
            << " selected_cells=" << particles.numSelectedBlocks()
            << " cells=" << particles.numBlocks() << " time_s=" << el
            << " particle_steps_per_s=" << (double)N * iter / el << std::endl;
  return 0;
}
