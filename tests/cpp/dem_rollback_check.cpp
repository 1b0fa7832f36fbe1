// dem_rollback_check.cpp -- cudaParticleDEM under AdaptiveTime with scripted
// inspectVelocity results, so that a rollback (cudaParticleDEM::rollback,
// then rollback2 on the next Progress) happens at a chosen step.  The loop
// is testDEM5.cu:181-234's: calcBlockID, calcForce(dt), calcAcceleration,
// addAccelerationZ, TimeEvolution, Progress, treatRefrectCondition.  After
// every iteration r, v, w, L and the latest contact matrix are appended to
// the output file, for the test to compare with the oracle's sequence.
//
//   dem_rollback_check <input> <output>
//
// input:  u32 n, u32 niter; f64 r[3n], v[3n], m[n]; f32 rad[n]; u8 fixed[n];
//         i32 res[niter]; f32 dt0, g, cell[6], E, mu, sigma, gamma, mu_r, rcell
// output: per iteration f64 r[3n], v[3n], w[3n], L[3n]; u32 nnz
//         (0xFFFFFFFF: no matrix), u32 row[n+1], u32 col[nnz], f64 val[nnz]
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <vector>
#include "../../particlebase_amd/cpp/include/AdaptiveTime.hh"
#include "../../particlebase_amd/cpp/include/CUDAenv.hh"
#include "../../particlebase_amd/cpp/include/cudaParticleDEM.hh"

template <typename T>
static bool rd(FILE *f, T *p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}
template <typename T>
static void wr(FILE *f, const T *p, size_t n) {
  if (n) std::fwrite(p, sizeof(T), n, f);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: dem_rollback_check <input> <output>" << std::endl;
    return 2;
  }
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  uint32_t n = 0, niter = 0;
  if (!rd(in, &n, 1) || !rd(in, &niter, 1)) return 2;
  std::vector<double> r(3 * n), v(3 * n), m(n);
  std::vector<float> rad(n);
  std::vector<uint8_t> fixed(n);
  std::vector<int32_t> script(niter);
  float par[14];
  if (!rd(in, r.data(), r.size()) || !rd(in, v.data(), v.size()) || !rd(in, m.data(), n) ||
      !rd(in, rad.data(), n) || !rd(in, fixed.data(), n) || !rd(in, script.data(), niter) ||
      !rd(in, par, 14))
    return 2;
  std::fclose(in);
  const float dt0 = par[0], g = par[1];
  real cell[6];
  for (int k = 0; k < 6; ++k) cell[k] = par[2 + k];
  const float E = par[8], mu = par[9], sigma = par[10], gamma = par[11], mu_r = par[12],
              rcell = par[13];

  std::vector<ParticleBase> G(n);
  std::vector<real> r0(n);
  for (uint32_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      G[i].r[c] = r[3 * i + c];
      G[i].v[c] = v[3 * i + c];
    }
    G[i].m = m[i];
    G[i].isFixed = fixed[i] != 0;
    r0[i] = rad[i];
  }
  CUDAenv<cudaParticleDEM> env;
  env.setup();
  cudaParticleDEM &P = env[0];
  P.switchBlockAlgorithm(ParticleBlockType::many);
  P.setup((int)n);
  P.setCell(cell);
  P.import(G);
  P.autotunetimestep = true;
  P.setDEMProperties(E, mu, sigma, gamma, mu_r, r0);
  P.setInertia(r0);
  P.setupCutoffBlock(rcell, false);
  P.calcBlockID();

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  AdaptiveTime<CUDAenv<cudaParticleDEM>> thistime;
  thistime.init(dt0);
  P.calcVinit(dt0);
  std::vector<int> res(1);
  uint32_t j = 0, rollbacks = 0;
  std::vector<uint32_t> row(n + 1), col;
  std::vector<double> val;
  for (uint32_t it = 0; it < niter; ++it) {
    P.calcBlockID();
    P.selectBlocks();
    P.calcForce(thistime.currentDeltaT());
    P.calcAcceleration();
    P.addAccelerationZ(-g);
    P.TimeEvolution(thistime.currentDeltaT());
    res[0] = script[it];
    if (res[0] == 1 && !thistime.isRollbacking()) ++rollbacks;
    j = thistime.Progress(env, res, j);
    P.treatRefrectCondition();
    for (int f : {PB_FIELD_R, PB_FIELD_V, PB_FIELD_W, PB_FIELD_L}) {
      const std::vector<double> x = P.getState(f);
      wr(out, x.data(), x.size());
    }
    uint32_t nnz = 0;
    if (pb_dem_get_contacts(P.handle(), row.data(), nullptr, nullptr, 0, &nnz) != PB_OK) {
      const uint32_t none = 0xFFFFFFFFu;
      wr(out, &none, 1);
      continue;
    }
    col.resize(nnz);
    val.resize(nnz);
    pbamd::check(pb_dem_get_contacts(P.handle(), row.data(), col.data(), val.data(), nnz, &nnz),
                 "pb_dem_get_contacts");
    wr(out, &nnz, 1);
    wr(out, row.data(), row.size());
    wr(out, col.data(), nnz);
    wr(out, val.data(), nnz);
  }
  std::fclose(out);
  std::cout << "done N=" << n << " iterations=" << niter << " rollbacks=" << rollbacks
            << " steps=" << j << " dt=" << thistime.currentDeltaT() << std::endl;
  return 0;
}
