// select_check.cpp -- putTMPselected on an SPH block (putTMPselected.cu,
// cudaParticleSPH_NS.cu:40-69): a small tank of fluid particles (ids
// [0, N1)) over wall particles (ids [N1, N)).  putUnSelected writes the
// walls; getSelectedTypeID + getSelectedPosition + putTMP write the fluid
// with its n, rho, p columns.  Both dumps are parsed back and compared with
// the imported positions and the density pass.  importForceSelected copies
// one block's density (device) and acceleration (TMP) into a second block.
// Prints "done ok=1".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <valarray>
#include <vector>
#include "../../particlebase_amd/cpp/include/CUDAenv.hh"
#include "../../particlebase_amd/cpp/include/cudaParticleSPH_NS.hh"

struct Dump {
  uint32_t n = 0;
  std::string header;
  std::vector<std::vector<double>> rows;
};

static Dump parse(std::istream &in) {
  Dump d;
  std::string line;
  int item = 0;
  while (std::getline(in, line)) {
    if (line.rfind("ITEM:", 0) == 0) {
      ++item;
      if (line.rfind("ITEM: ATOMS", 0) == 0) d.header = line;
      continue;
    }
    if (item == 2) d.n = (uint32_t)std::stoul(line);
    if (item == 4) {
      std::istringstream f(line);
      std::vector<double> v;
      double x;
      while (f >> x) v.push_back(x);
      d.rows.push_back(v);
    }
  }
  return d;
}

int main(int argc, char **argv) {
  const std::string tmpdir = argc > 1 ? argv[1] : ".";
  const double dx = 0.01, h = 1.3 * dx;
  std::vector<ParticleBase> fluid, wall;
  auto mk = [&](int i, int j, int k, bool fixed) {
    ParticleBase p{};
    p.r[0] = (i + 0.5) * dx;
    p.r[1] = (j + 0.5) * dx;
    p.r[2] = (k + 0.5) * dx;
    p.m = 1000.0 * dx * dx * dx;
    p.isFixed = fixed;
    p.type = fixed ? 1 : 0;
    return p;
  };
  for (int i = -3; i < 15; ++i)
    for (int j = -3; j < 15; ++j)
      for (int k = -3; k < 12; ++k) {
        const bool inside = i >= 0 && i < 12 && j >= 0 && j < 12 && k >= 0;
        if (!inside) wall.push_back(mk(i, j, k, true));
        else if (k < 8) fluid.push_back(mk(i, j, k, false));
      }
  std::vector<ParticleBase> G(fluid);
  G.insert(G.end(), wall.begin(), wall.end());
  const uint32_t N1 = (uint32_t)fluid.size(), N = (uint32_t)G.size();

  CUDAenv<cudaParticleSPH_NS> env;
  env.setup();
  cudaParticleSPH_NS &P = env[0];
  real cell[6] = {(real)(-3 * dx), (real)(15 * dx), (real)(-3 * dx), (real)(15 * dx),
                  (real)(-3 * dx), (real)(12 * dx)};
  P.setup(N);
  P.setCell(cell);
  P.import(G);
  P.setWCSPHProperties(h, 1000.0, 10.0, 7.0, 0.1, 0.0);
  P.setupCutoffBlock((real)(2 * h), false);
  P.calcBlockID();
  P.calcDensity();
  const std::vector<double> num = P.getState(PB_FIELD_SPH_NUM);

  bool ok = true;
  double worst = 0;
  P.setupSelectedTMP(0, N1, N1, N - N1);
  const std::string wfile = tmpdir + "/select_check_walls.dump";
  P.putUnSelected(wfile.c_str());
  {
    std::ifstream in(wfile);
    const Dump d = parse(in);
    ok = ok && d.n == N - N1 && d.rows.size() == N - N1 && d.header == "ITEM: ATOMS id type x y z";
    for (size_t i = 0; ok && i < d.rows.size(); ++i) {
      const std::vector<double> &v = d.rows[i];
      ok = v.size() == 5 && v[0] == (double)i && v[1] == 1.0;
      for (int k = 0; ok && k < 3; ++k)
        worst = std::fmax(worst, std::fabs(v[2 + k] - wall[i].r[k]) / 0.15);
    }
  }
  std::remove(wfile.c_str());

  P.getSelectedTypeID();
  P.getSelectedPosition();
  std::ostringstream o;
  P.putTMP(o);
  P.waitPutTMP();
  std::istringstream in(o.str());
  const Dump d = parse(in);
  ok = ok && d.n == N1 && d.rows.size() == N1 && d.header == "ITEM: ATOMS id type x y z n rho p";
  double worst_num = 0;
  for (size_t i = 0; ok && i < d.rows.size(); ++i) {
    const std::vector<double> &v = d.rows[i];
    ok = v.size() == 8 && v[0] == (double)i && v[1] == 0.0;
    for (int k = 0; ok && k < 3; ++k) {
      worst = std::fmax(worst, std::fabs(v[2 + k] - fluid[i].r[k]) / 0.15);
      // 9 significant digits: the FP32 value itself
      const double want = (double)(float)num[3 * i + k];
      worst_num = std::fmax(worst_num, std::fabs(v[5 + k] - want) / std::fmax(std::fabs(want), 1e-30));
    }
  }
  // positions print with the stream's 6 digits; n, rho, p with 9
  ok = ok && worst < 1e-5 && worst_num < 1e-7;

  // back to whole-system dumps: getTypeID restores TMP2 to N particles and
  // the n, rho, p block moves behind all N positions
  P.getTypeID();
  P.getPosition();
  std::ostringstream o2;
  P.putTMP(o2);
  P.waitPutTMP();
  {
    std::istringstream in2(o2.str());
    const Dump d2 = parse(in2);
    ok = ok && d2.n == N && d2.rows.size() == N;
    for (size_t i = 0; ok && i < d2.rows.size(); ++i) {
      const std::vector<double> &v = d2.rows[i];
      ok = v.size() == 8 && v[1] == (i < N1 ? 0.0 : 1.0);
      for (int k = 0; ok && k < 3; ++k) {
        const double want = (double)(float)num[3 * i + k];
        ok = std::fabs(v[5 + k] - want) <= 1e-7 * std::fabs(want);
      }
    }
  }

  // importForceSelected between two blocks (copies, no summation): the
  // density block from the device, then the acceleration from TMP
  CUDAenv<cudaParticleSPH_NS> env2;
  env2.setup();
  cudaParticleSPH_NS &B = env2[0];
  // B's own densities come from a compressed copy of the particles (every
  // coordinate scaled by 0.97), then B takes P's positions back, so only the
  // imported densities give P's acceleration
  std::vector<ParticleBase> Gc(G);
  for (auto &p : Gc)
    for (int c = 0; c < 3; ++c) p.r[c] *= 0.97;
  B.setup(N);
  B.setCell(cell);
  B.import(Gc);
  B.setWCSPHProperties(h, 1000.0, 10.0, 7.0, 0.1, 0.0);
  B.setupCutoffBlock((real)(2 * h), false);
  B.calcBlockID();
  B.calcDensity();
  const std::vector<double> rP = P.getState(PB_FIELD_R);
  pbamd::check(pb_sys_set(B.handle(), PB_FIELD_R, rP.data()), "positions of P");
  B.calcBlockID();
  std::vector<double> zero(3 * (size_t)N, 1.0);
  pbamd::check(pb_sys_set(B.handle(), PB_FIELD_SPH_NUM, zero.data()), "reset num");
  B.importForceSelected(P, ExchangeMode::density, true);
  ok = ok && B.getState(PB_FIELD_SPH_NUM) == num;
  // the sequence of testSPH5.cu:208-220: density import, calcDensityPost,
  // calcAcceleration -- equal to P's acceleration from its own densities
  B.calcDensityPost();
  B.calcAcceleration();
  P.calcAcceleration();
  const bool acc_from_import = B.getState(PB_FIELD_A) == P.getState(PB_FIELD_A);
  ok = ok && acc_from_import;
  P.getForceSelected(ExchangeMode::acceleration);
  B.importForceSelected(P, ExchangeMode::acceleration, false);
  const std::vector<double> pa = P.getState(PB_FIELD_A), ba = B.getState(PB_FIELD_A);
  for (size_t k = 0; ok && k < pa.size(); ++k) ok = ba[k] == (double)(float)pa[k];
  std::cout << "done ok=" << (ok ? 1 : 0) << " N=" << N << " N1=" << N1 << " poserr=" << worst
            << " numerr=" << worst_num << " acc_from_import=" << (acc_from_import ? 1 : 0)
            << std::endl;
  return ok ? 0 : 1;
}
