// cudaParticleSPH_NS.hh -- SPH Navier-Stokes particles (reference
// cudaParticleSPHBase.hh/.cu + cudaParticleSPH_NS.hh/.cu) over the engine's
// SPH kernels: Lucy kernel of support h, p = c^2 (rho - 1), viscosity mu.
// setWCSPHProperties switches the same class to the weakly compressible
// form (cubic spline of support 2h, Tait, artificial viscosity), which the
// reference does not have; setSPHProperties switches back.
// The sortedOutput flags of the reference exist for its multi-GPU exchange;
// one GPU owns every particle here, so they are accepted and have no effect.
#pragma once
#include <cmath>
#include <limits>
#include <sstream>
#include <valarray>
#include "ExchangeMode.hh"
#include "cudaCutoffBlock.hh"
#include "cudaParticleLF.hh"
#include "putTMPselected.hh"

class cudaParticleSPH_NS : public cudaParticleLF,
                           virtual public cudaCutoffBlock,
                           public putTMPselected {
protected:
  static int selectedField(const ExchangeMode typeID, const char *what) {
    switch (typeID) {
      case ExchangeMode::force: return PB_FIELD_F;
      case ExchangeMode::acceleration: return PB_FIELD_A;
      case ExchangeMode::density: return PB_FIELD_SPH_NUM;
      default: throw std::runtime_error(std::string(what) + ": SPH exchanges force, acceleration or density");
    }
  }
  real h = 0;
  // particles held in TMP's first block: the n, rho, p block starts after
  // them (cudaParticleSPH_NS.hh:111)
  uint32_t offset_additional_output = 0;
  double meanNum = 0.0;
  std::vector<double> mu0s, cs;    // setSPHProperties arguments, kept for restarts
  // setWCSPHProperties arguments (h, rho0, c0, gamma, alpha, pmin); empty
  // under the reference physics
  std::vector<double> wcsph;

public:
  // the dump thread calls this class's additionalTag / additionalOutput:
  // join it while the object is still whole (the base destructors would
  // otherwise run while a final dump is being written)
  ~cudaParticleSPH_NS() override { waitPutTMP(); }

  void setup(int n) override {
    cudaParticleBase::setup(n);
    TMP.resize(8 * (size_t)n);       // x, y, z, _, num, rho, mu, _ (cudaParticleSPH_NS.cu:13-27)
    offset_additional_output = (uint32_t)n;
  }

  // dump columns after x y z (cudaParticleSPH_NS.hh:37, cudaParticleSPH_NS.cu:
  // 64-69): the three density-pass values of the particle, FP32 printed
  // with the 9 digits that round-trip
  std::string additionalTag(void) const override { return " n rho p"; }
  std::string additionalOutput(uint32_t i) const override {
    std::ostringstream o;
    o.precision(std::numeric_limits<float>::max_digits10);
    const size_t b = 4 * ((size_t)i + offset_additional_output);
    o << " " << TMP[b] << " " << TMP[b + 1] << " " << TMP[b + 2];
    return o.str();
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
    wcsph.clear();
    pbamd::check(pb_sph_set_properties(sys, m.data(), c.data(), _h), "setSPHProperties");
  }

  // weakly compressible form: M4 cubic spline of support 2h, summation
  // density, p = max(B ((rho/rho0)^gamma - 1), pmin), B = rho0 c0^2/gamma,
  // Monaghan artificial viscosity alpha; the cell length must be >= 2h
  void setWCSPHProperties(double _h, double rho0, double c0, double gamma = 7.0,
                          double alpha = 0.1, double pmin = 0.0) {
    h = (real)_h;
    wcsph = {_h, rho0, c0, gamma, alpha, pmin};
    pbamd::check(pb_sph_set_wcsph(sys, _h, rho0, c0, gamma, alpha, pmin), "setWCSPHProperties");
  }

  // cudaParticleSPHBase / cudaParticleSPH_NS fields (cudaParticleSPHBase.hh:
  // 83-106, cudaParticleSPH_NS.hh:129-175): h, w0, viscosity and sound
  // speed (c itself is kept too, so a restart sets c^2 bit for bit)
  void save(pbamd::OArchive &a) const override {
    saveBase(a);
    saveBlock(a);
    saveSelected(a);
    a.scalar("cudaParticleSPHBase.h", h);
    a.scalar("cudaParticleSPHBase.w0", 105.0 / (16.0 * M_PI * std::pow((double)h, 7)));
    a.put("cudaParticleSPH_NS.viscosityMu", mu0s);
    std::vector<double> c2(cs.size());
    for (size_t i = 0; i < cs.size(); ++i) c2[i] = cs[i] * cs[i];
    a.put("cudaParticleSPH_NS.soundVelocityC2", c2);
    a.put("cudaParticleSPH_NS.soundVelocityC", cs);
    if (!wcsph.empty()) a.put("cudaParticleSPH_NS.wcsph", wcsph);
  }
  void load(const pbamd::IArchive &a) override {
    loadBase(a);
    loadBlock(a);
    loadSelected(a);
    h = a.scalar<real>("cudaParticleSPHBase.h");
    mu0s = a.get<double>("cudaParticleSPH_NS.viscosityMu");
    cs = a.get<double>("cudaParticleSPH_NS.soundVelocityC");
    if (a.has("cudaParticleSPH_NS.wcsph")) {
      const std::vector<double> p = a.get<double>("cudaParticleSPH_NS.wcsph");
      if (p.size() != 6) throw std::runtime_error("readSerialization: SPH wcsph parameters");
      setWCSPHProperties(p[0], p[1], p[2], p[3], p[4], p[5]);
    } else if (!mu0s.empty()) {
      if (mu0s.size() != N || cs.size() != N) throw std::runtime_error("readSerialization: SPH arrays");
      wcsph.clear();
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
    offset_additional_output = N;
    for (size_t i = 0; i < N; ++i)
      for (int k = 0; k < 3; ++k) TMP[4 * (size_t)N + 4 * i + k] = (real)num[3 * i + k];
  }
  // getSelectedPosition (cudaParticleSPH_NS.cu:40-50): the selected range's
  // x, y, z, then its n, rho, p block right after it
  void getSelectedPosition(void) override {
    fetchRange(selected[0], selected[1]);
    std::vector<double> num(3 * (size_t)N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_SPH_NUM, num.data()), "getSelectedPosition");
    std::lock_guard<std::mutex> lk(mutTMP);
    offset_additional_output = selected[1];
    const size_t b = 4 * (size_t)selected[1];
    for (size_t i = 0; i < selected[1]; ++i) {
      for (int k = 0; k < 3; ++k) TMP[b + 4 * i + k] = (real)num[3 * (selected[0] + i) + k];
      TMP[b + 4 * i + 3] = 0;
    }
  }

  // SPHBase::rollback / rollback2 (cudaParticleSPHBase.cu:127-132); the
  // second stage re-initialises v with the (empty) SPH force, as there
  void rollback(real dt) { cudaParticleLF::rollback(dt); }
  void rollback2(real dt) { cudaParticleLF::calcVinit(dt / 2); }

  // getForceSelected / importForceSelected (cudaParticleSPH_NS.hh:96-108,
  // cudaSelectedBlock.cu:46-81).  One GPU owns every moving particle, so
  // the selected range is the whole system.  force / acceleration / density
  // (n, rho, p) are staged in TMP (4 per particle); the import copies the
  // other object's values (no summation) from its device arrays
  // (directAccess) or from its TMP.
  void getForceSelected(const ExchangeMode typeID = ExchangeMode::force) {
    fetch(selectedField(typeID, "getForceSelected"));
  }
  void importForceSelected(const cudaParticleSPH_NS &A, const ExchangeMode typeID = ExchangeMode::force,
                           bool directAccess = false, int idMe = 0, int idPeer = 0) {
    (void)idMe;
    (void)idPeer;
    const int field = selectedField(typeID, "importForceSelected");
    if (A.N != N) throw std::runtime_error("importForceSelected: particle counts differ");
    std::vector<double> other(3 * (size_t)N);
    if (directAccess) {
      pbamd::check(pb_sys_get(A.sys, field, other.data()), "importForceSelected");
    } else {
      if (A.TMP.size() < 4 * (size_t)N) throw std::runtime_error("importForceSelected: TMP");
      for (size_t i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) other[3 * i + c] = (double)A.TMP[4 * i + c];
    }
    pbamd::check(pb_sys_set(sys, field, other.data()), "importForceSelected");
  }

  // multi-GPU exchange ranges: one GPU owns all particles
  void getExchangePidRange1(void) {}
  void getExchangePidRange2(void) {}
  void selectBlocks(void) {}
};
