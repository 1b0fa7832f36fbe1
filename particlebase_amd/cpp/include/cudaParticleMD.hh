// cudaParticleMD.hh -- LJ molecular dynamics (reference cudaParticleMD.hh/.cu).
// calcForce computes F from the current positions with the engine's
// Verlet-list pair kernel (FP64 by default, or FP32 pair math with FP64
// accumulation); calcPotentialE returns the correct 1/2 sum phi in FP64
// (the reference's per-cell assignment bug, kerneltemplate.hh:582, is not
// reproduced).  runFused() is the engine's fused loop (no reference
// counterpart as one call).
#pragma once
#include <cmath>
#include <iostream>
#include <numeric>
#include <vector>
#include "cudaCutoffBlock.hh"
#include "cudaParticleVV.hh"
#include "cudaSortedPositions.hh"

class cudaParticleMD : public cudaParticleVV,
                       virtual public cudaCutoffBlock,
                       virtual public cudaSortedPositions {
protected:
  real m0 = 0;
  real mv2 = 0;
  double lastPE = 0.0, lastVirial = 0.0;
  bool paramsSet = false;
  void syncCutoff() {
    if (rmax2 != cutoffSent || shifted != shiftSent) {
      pbamd::check(pb_lj_set_cutoff(sys, rmax2, shifted ? 1 : 0), "rmax2");
      cutoffSent = rmax2;
      shiftSent = shifted;
    }
  }
  double cutoffSent = -1;
  bool shiftSent = false;

public:
  real kB = 1.0;
  real rmax2 = 0.0;
  real thickness = 0.0;
  bool useNlist = false;
  bool shifted = false;          //!< shift energies so phi(rc) = 0 (forces unchanged)
  int precision = PB_FP64;       //!< PB_FP64 or PB_MIXED pair math

  void setup(int n) override { cudaParticleVV::setup(n); }

  // cudaParticleMD fields (cudaParticleMD.hh:195-240) after particleVV and
  // cutoffBlock; the engine's own settings follow
  void save(pbamd::OArchive &a) const override {
    saveBase(a);
    saveBlock(a);
    a.scalar("cudaParticleMD.BoltzmannConstKb", kB);
    a.scalar("cudaParticleMD.rmax2", rmax2);
    a.scalar("cudaParticleMD.m0", m0);
    a.scalar("cudaParticleMD.thickness", thickness);
    a.scalar("cudaParticleMD.shifted", (uint8_t)shifted);
    a.scalar("cudaParticleMD.precision", (int32_t)precision);
  }
  void load(const pbamd::IArchive &a) override {
    loadBase(a);
    loadBlock(a);
    kB = a.scalar<real>("cudaParticleMD.BoltzmannConstKb");
    rmax2 = a.scalar<real>("cudaParticleMD.rmax2");
    m0 = a.scalar<real>("cudaParticleMD.m0");
    thickness = a.scalar<real>("cudaParticleMD.thickness");
    shifted = a.scalar<uint8_t>("cudaParticleMD.shifted") != 0;
    precision = a.scalar<int32_t>("cudaParticleMD.precision");
  }

  // calcBlockID "steals" both parents (cudaParticleMD.hh:64-67)
  void calcBlockID(void) {
    cudaCutoffBlock::calcBlockID();
    cudaSortedPositions::calcBlockID();
  }
  void setupCutoffBlock(real rmax, bool periodic = true) {
    cudaCutoffBlock::setupCutoffBlock(rmax, periodic);
    cudaSortedPositions::setupSortedPositions();
  }

  // calcForce, cudaParticleMD.cu:24-65: from the neighbour list when
  // useNList() was called (the list of the last makeNList), else the cells
  void calcForce(void) override {
    syncCutoff();
    if (useNlist)
      pbamd::check(pb_lj_force_nlist(sys, &lastPE, &lastVirial), "cudaParticleMD::calcForce");
    else
      pbamd::check(pb_lj_force(sys, precision, &lastPE, &lastVirial), "cudaParticleMD::calcForce");
  }

  // useNList / makeNList, cudaParticleMD.hh:167-175, cudaParticleMD.cu:338-394:
  // CSR list of the pairs within sqrt(rmax2) + thickness
  void useNList(void) { useNlist = true; }
  void makeNList(void) {
    if (!useNlist) return;
    const real rt = std::sqrt(rmax2) + thickness;
    calcBlockID();
    pbamd::check(pb_nlist_build(sys, rt * rt, nullptr), "cudaParticleMD::makeNList");
  }
  // the list as cudaSparseMat's rowPtr / colIdx (original indices)
  void getNList(std::vector<uint32_t> &rowPtr, std::vector<uint32_t> &colIdx) {
    uint32_t nnz = 0;
    rowPtr.resize(N + 1);
    pbamd::check(pb_nlist_get(sys, rowPtr.data(), nullptr, 0, &nnz), "getNList");
    colIdx.resize(nnz);
    pbamd::check(pb_nlist_get(sys, nullptr, colIdx.data(), nnz, &nnz), "getNList");
  }

  // initialAnnealing, cudaParticleMD.cu:205-283: anealstep Euler steps with
  // the soft-core force of radius _rc (i+1)/anealstep and strength _f0
  // under constTemp, scaleTemp(T), then anealstep velocity-Verlet steps
  // with LJ under constTemp, and scaleTemp(T) again
  void initialAnnealing(uint32_t anealstep, real dt, real _rc, real _f0, real T) {
    std::cerr << "performs initial annealing by soft core potential to rmax= " << _rc
              << " with steps " << anealstep << std::endl;
    for (uint32_t i = 0; i < anealstep; ++i) {
      const double rc = (double)_rc * (i + 1) / anealstep;
      calcBlockID();
      pbamd::check(pb_softcore_force(sys, rc, _f0, rmax2), "initialAnnealing soft core");
      constTemp();
      pbamd::check(pb_calc_acceleration(sys), "initialAnnealing calcA");
      pbamd::check(pb_euler_step(sys, dt), "initialAnnealing Euler");
      treatPeriodicCondition();
    }
    std::cerr << "temperature scaling: " << scaleTemp(T) << std::endl;
    std::cerr << "with LJ " << std::endl;
    pbamd::check(pb_clear_force(sys, 1), "initialAnnealing clear Fold");
    syncCutoff();
    for (uint32_t i = 0; i < anealstep; ++i) {
      calcBlockID();
      pbamd::check(pb_lj_force(sys, precision, &lastPE, &lastVirial), "initialAnnealing LJ");
      constTemp();
      pbamd::check(pb_vv_step(sys, dt), "initialAnnealing VV");
      treatPeriodicCondition();
    }
    std::cerr << "temperature scaling: " << scaleTemp(T) << std::endl;
  }

  void setM(void) {}

  // sum m v^2 in FP64 (cudaParticleMD.cu:124-133)
  real calcMV2(void) {
    double out = 0;
    pbamd::check(pb_mv2(sys, &out), "calcMV2");
    return (real)out;
  }
  // statMV2 (cudaParticleMD.cu:157-166): two blank lines, then "i<TAB>v_i^2"
  // per particle; like the reference it lists the squared speed (the mass
  // enters only calcMV2's sum)
  void statMV2(std::ostream &o = std::cout) {
    calcMV2();
    const std::vector<double> v = getState(PB_FIELD_V);
    o << std::endl << std::endl;
    for (uint32_t i = 0; i < N; ++i) {
      const double v2 = v[3 * (size_t)i] * v[3 * (size_t)i] +
                        v[3 * (size_t)i + 1] * v[3 * (size_t)i + 1] +
                        v[3 * (size_t)i + 2] * v[3 * (size_t)i + 2];
      o << i << "\t" << (float)v2 << std::endl;
    }
  }
  real calcTemp(void) {
    mv2 = calcMV2();
    return mv2 / (N * 3 * kB);
  }
  real calcKineticE(void) const { return mv2 / 2.0; }
  real calcPotentialE(void) {
    syncCutoff();
    double pe = 0;
    pbamd::check(pb_lj_force(sys, precision, &pe, &lastVirial), "calcPotentialE");
    lastPE = pe;
    return (real)pe;
  }
  double lastPotentialE(void) const { return lastPE; }
  double lastVirialW(void) const { return lastVirial; }

  // scaleTemp, cudaParticleMD.cu:168-175
  virtual real scaleTemp(real Temp) {
    const real T0 = calcTemp();
    const real s = std::sqrt(static_cast<real>(3 * N - 1) / static_cast<real>(3 * N) * Temp / T0);
    pbamd::check(pb_scale_velocities(sys, s), "scaleTemp");
    return s;
  }

  // adjustVelocities, cudaParticleMD.cu:177-198: zero mean and variance
  // kB T / m0 per axis; debug prints the statistics before the adjustment
  virtual void adjustVelocities(real Temp, bool debug = false) {
    const double v1 = std::sqrt((double)kB * Temp / m0);
    double st[6];
    if (debug)
      std::cerr << std::endl << std::endl << "adjustVelocity currentTemp\t" << calcTemp()
                << "\tTarget Temp " << Temp << std::endl;
    pbamd::check(pb_adjust_velocities(sys, v1, st), "adjustVelocities");
    if (debug) {
      std::cerr << "velocity statistics";
      for (int i = 0; i < 6; ++i) std::cerr << "\t" << st[i];
      std::cerr << std::endl << "after adjusted T=" << calcTemp() << std::endl;
    }
  }

  // constTemp, cudaParticleMD.cu:140-155: F -= lambda m v with
  // lambda = sum F.v / sum m v^2; returns the temperature
  real constTemp(void) {
    double m2 = 0, lam = 0;
    pbamd::check(pb_const_temp(sys, &m2, &lam), "constTemp");
    mv2 = (real)m2;
    return (real)(m2 / (3.0 * N * kB));
  }

  // setLJparams, cudaParticleMD.cu:200-202 (sigma, epsilon per type pair)
  void setLJparams(const std::vector<real> &p, uint32_t elemnum) {
    std::vector<double> d(p.begin(), p.end());
    pbamd::check(pb_lj_set_params(sys, d.data(), (int)elemnum), "setLJparams");
    paramsSet = true;
  }

  void import(const std::vector<ParticleBase> &P) {
    cudaParticleBase::import(P);
    const double s = std::accumulate(P.begin(), P.end(), 0.0,
                                     [](double acc, const ParticleBase &c) { return acc + c.m; });
    m0 = (real)(s / P.size());
  }

  // neighbour counts of makeNList step 1 (cudaParticleMD.cu:343-366)
  std::vector<uint32_t> coordinationNumbers(void) {
    const real rmax2t = (std::sqrt(rmax2) + thickness) * (std::sqrt(rmax2) + thickness);
    std::vector<uint32_t> c(N);
    pbamd::check(pb_coord_counts(sys, rmax2t, c.data()), "coordinationNumbers");
    return c;
  }

  // fused loop: nsteps of [calcForce, TimeEvolution, treatPeriodicCondition]
  std::vector<double> runFused(uint32_t nsteps, real dt, int prec = PB_MIXED, double skin = 0.3,
                               uint32_t rebuild = 20, uint32_t energy_every = 0) {
    syncCutoff();
    pb_run_opts o{};
    o.dt = dt;
    o.nsteps = nsteps;
    o.precision = prec;
    o.skin = skin;
    o.rebuild_every = rebuild;
    o.energy_every = energy_every;
    const uint32_t nrec = energy_every ? (nsteps + energy_every - 1) / energy_every : 0;
    std::vector<double> e(3 * (size_t)std::max<uint32_t>(nrec, 1));
    pbamd::check(pb_lj_run(sys, &o, e.data()), "runFused");
    e.resize(3 * (size_t)nrec);
    return e;
  }
};
