// cudaParticleSPH_NS.hh -- SPH Navier-Stokes particles (reference
// cudaParticleSPHBase.hh/.cu + cudaParticleSPH_NS.hh/.cu) over the engine's
// SPH kernels: Lucy kernel of support h, p = c^2 (rho - 1), viscosity mu.
// The sortedOutput flags of the reference exist for its multi-GPU exchange;
// one GPU owns every particle here, so they are accepted and have no effect.
#pragma once
#include <cmath>
#include <valarray>
#include "cudaCutoffBlock.hh"
#include "cudaParticleLF.hh"

class cudaParticleSPH_NS : public cudaParticleLF, virtual public cudaCutoffBlock {
protected:
  real h = 0;
  double meanNum = 0.0;
  std::vector<double> mu0s, cs;    // setSPHProperties arguments, kept for restarts

public:
  void setup(int n) override {
    cudaParticleBase::setup(n);
    TMP.resize(8 * (size_t)n);       // x, y, z, _, num, rho, mu, _ (cudaParticleSPH_NS.cu:13-27)
  }

  // SPH needs no force pass of its own (testSPH5.cu:216 "do nothing")
  void calcForce(void) override {}
  // cudaParticleSPHBase::calcKernels is a no-op with cutoff blocks
  void calcKernels(void) {}

  // cudaParticleSPH_NS::setSPHProperties (cudaParticleSPH_NS.cu:56-66)
  void setSPHProperties(const std::valarray<real> &mu0, std::valarray<real> c1, real _h) {
    std::vector<double> m(N), c(N);
    for (size_t i = 0; i < N; ++i) {
      m[i] = mu0[i];
      c[i] = c1[i];
    }
    h = _h;
    mu0s = m;
    cs = c;
    pbamd::check(pb_sph_set_properties(sys, m.data(), c.data(), _h), "setSPHProperties");
  }

  // cudaParticleSPHBase / cudaParticleSPH_NS fields (cudaParticleSPHBase.hh:
  // 83-106, cudaParticleSPH_NS.hh:129-175): h, w0, viscosity and sound
  // speed (c itself is kept too, so a restart sets c^2 bit for bit)
  void save(pbamd::OArchive &a) const override {
    saveBase(a);
    saveBlock(a);
    a.scalar("cudaParticleSPHBase.h", h);
    a.scalar("cudaParticleSPHBase.w0", 105.0 / (16.0 * M_PI * std::pow((double)h, 7)));
    a.put("cudaParticleSPH_NS.viscosityMu", mu0s);
    std::vector<double> c2(cs.size());
    for (size_t i = 0; i < cs.size(); ++i) c2[i] = cs[i] * cs[i];
    a.put("cudaParticleSPH_NS.soundVelocityC2", c2);
    a.put("cudaParticleSPH_NS.soundVelocityC", cs);
  }
  void load(const pbamd::IArchive &a) override {
    loadBase(a);
    loadBlock(a);
    h = a.scalar<real>("cudaParticleSPHBase.h");
    mu0s = a.get<double>("cudaParticleSPH_NS.viscosityMu");
    cs = a.get<double>("cudaParticleSPH_NS.soundVelocityC");
    if (!mu0s.empty()) {
      if (mu0s.size() != N || cs.size() != N) throw std::runtime_error("readSerialization: SPH arrays");
      pbamd::check(pb_sph_set_properties(sys, mu0s.data(), cs.data(), h), "readSerialization");
    }
  }

  // calcDensity + calcDensityPost (cudaParticleSPH_NS.cu:136-207)
  void calcDensity(bool sortedOutput = false) {
    (void)sortedOutput;
    pbamd::check(pb_sph_density(sys), "calcDensity");
  }
  void calcDensityPost(bool sortedOutput = false) { (void)sortedOutput; }

  // calcAcceleration (cudaParticleSPH_NS.cu:71-134)
  void calcAcceleration(bool sortedOutput = false) {
    (void)sortedOutput;
    pbamd::check(pb_sph_acceleration(sys), "calcAcceleration");
  }
  // accelerations are written in particle order already
  void RestoreAcceleration(void) {}

  // inspectDensity (cudaParticleSPH_NS.cu:209-217)
  real inspectDensity(void) {
    pbamd::check(pb_sph_inspect_density(sys, &meanNum), "inspectDensity");
    return (real)meanNum;
  }

  // getPosition: x, y, z in TMP[0..4N), num, rho, mu in TMP[4N..8N)
  // (cudaParticleSPH_NS.cu:29-38)
  void getPosition(void) {
    fetch(PB_FIELD_R);
    std::vector<double> num(3 * (size_t)N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_SPH_NUM, num.data()), "getPosition");
    std::lock_guard<std::mutex> lk(mutTMP);
    for (size_t i = 0; i < N; ++i)
      for (int k = 0; k < 3; ++k) TMP[4 * (size_t)N + 4 * i + k] = (real)num[3 * i + k];
  }

  // SPHBase::rollback / rollback2 (cudaParticleSPHBase.cu:127-132); the
  // second stage re-initialises v with the (empty) SPH force, as there
  void rollback(real dt) { cudaParticleLF::rollback(dt); }
  void rollback2(real dt) { cudaParticleLF::calcVinit(dt / 2); }

  // multi-GPU exchange ranges: one GPU owns all particles
  void getExchangePidRange1(void) {}
  void getExchangePidRange2(void) {}
  void selectBlocks(void) {}
};
