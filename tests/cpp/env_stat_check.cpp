// env_stat_check.cpp -- CUDAenv's device queries and display (CUDAenv.hh:
// 29-34, 177-194), cudaCutoffBlock::calcBlockRange (cudaCutoffBlock.hh:
// 55-70), cudaFrame::nowtime and cudaParticleMD::statMV2
// (cudaParticleMD.cu:157-166).
// statMV2's listing is parsed back and compared with v^2 of the imported
// velocities; prints "done ok=1" with the queried limits.  Built with the
// drivers (make -C particlebase_amd/cpp).
#include <cmath>
#include <cstdint>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include "../../particlebase_amd/cpp/include/CUDAenv.hh"
#include "../../particlebase_amd/cpp/include/cudaParticleMD.hh"

int main() {
  const int n = 500;
  std::mt19937_64 eng(5);
  std::uniform_real_distribution<double> u(0.0, 8.0), uv(-2.0, 2.0);
  std::vector<ParticleBase> G(n);
  for (auto &p : G) {
    for (int c = 0; c < 3; ++c) {
      p.r[c] = u(eng);
      p.v[c] = uv(eng);
      p.a[c] = 0;
    }
    p.m = 3.0;
    p.isFixed = false;
    p.type = 0;
  }
  real cell[6] = {0, 8, 0, 8, 0, 8};
  CUDAenv<cudaParticleMD> env;
  env.setup();
  std::ostringstream disp;
  env.display(disp);
  bool ok = disp.str().find("shared memory size") != std::string::npos;
  ok = ok && env.getShMem() >= 65536 && env.getNumTh() >= 256 && env.getMaxGrid() > 0 &&
       env.getMemSize() > (1ull << 32) && env.getNumCpMP() == 64 && env.getNumMP() > 0;
  env.exchangeForceSelected(ExchangeMode::acceleration);   // one GPU: nothing to move
  // calcBlockRange: 10 cells over 3 devices -> [0,4) [4,7) [7,10)
  const uint32_t want_off[3] = {0, 4, 7}, want_num[3] = {4, 3, 3};
  for (uint32_t id = 0; id < 3; ++id)
    cudaCutoffBlock::calcBlockRange(10, 3, id, [&](uint32_t o, uint32_t k) {
      ok = ok && o == want_off[id] && k == want_num[id];
    });
  std::ostringstream clock;
  ok = ok && env[0].nowtime(clock) > 0 && !clock.str().empty();

  cudaParticleMD &P = env[0];
  P.setup(n);
  P.setCell(cell);
  P.import(G);
  std::ostringstream o;
  P.statMV2(o);
  std::istringstream in(o.str());
  std::string line;
  int blank = 0, rows = 0;
  double worst = 0;
  while (std::getline(in, line)) {
    if (line.empty()) {
      ++blank;
      continue;
    }
    std::istringstream f(line);
    long i = -1;
    double v2 = -1;
    f >> i >> v2;
    if (i != rows) ok = false;
    if (i >= 0 && i < n) {
      const ParticleBase &p = G[i];
      const double want = p.v[0] * p.v[0] + p.v[1] * p.v[1] + p.v[2] * p.v[2];
      worst = std::fmax(worst, std::fabs(v2 - want) / std::fmax(want, 1e-30));
    }
    ++rows;
  }
  // the listing has the stream's default 6 significant digits
  ok = ok && blank == 2 && rows == n && worst < 1e-5;
  std::cout << "done ok=" << (ok ? 1 : 0) << " N=" << n << " rows=" << rows
            << " relerr=" << worst << " lds=" << env.getShMem() << " cp=" << env.getNumCpMP()
            << std::endl;
  return ok ? 0 : 1;
}
