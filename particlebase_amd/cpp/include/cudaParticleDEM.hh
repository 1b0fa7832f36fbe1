// cudaParticleDEM.hh -- DEM spheres with rotation (reference
// cudaParticleDEM.hh/.cu + cudaParticleRotation.hh/.cu) over the engine's
// contact kernels: contact CSR with tangential history, Hertz normal force,
// Coulomb-capped tangential force, rolling friction, Euler rotation.
#pragma once
#include <vector>
#include "cudaCutoffBlock.hh"
#include "cudaParticleLF.hh"
#include "putTMPselected.hh"

class cudaParticleRotation : public virtual cudaParticleBase {
protected:
  std::vector<double> inertiaR;    // radii given to setInertia (restarts rebuild m0inv)
  bool inertiaSet = false;

  // cudaParticleRotation fields (cudaParticleRotation.hh:68-125); m0inv is
  // rebuilt from the radii, exactly as setInertia computed it
  void saveRotation(pbamd::OArchive &a) const {
    a.scalar("cudaParticleRotation.inertiaSet", (uint8_t)inertiaSet);
    if (!inertiaSet) return;
    a.put("cudaParticleRotation.inertiaRadius", inertiaR);
    std::vector<double> x(3 * (size_t)N);
    const std::pair<const char *, int> vecs[3] = {{"angularVelocityW", PB_FIELD_W},
                                                  {"angularMomentumL", PB_FIELD_L},
                                                  {"torqueT", PB_FIELD_T}};
    for (const auto &f : vecs) {
      pbamd::check(pb_sys_get(sys, f.second, x.data()), "writeSerialization");
      a.put(std::string("cudaParticleRotation.") + f.first, x);
    }
  }
  void loadRotation(const pbamd::IArchive &a) {
    if (!a.scalar<uint8_t>("cudaParticleRotation.inertiaSet")) return;
    inertiaR = a.get<double>("cudaParticleRotation.inertiaRadius");
    if (inertiaR.size() != N) throw std::runtime_error("readSerialization: inertia radii");
    pbamd::check(pb_rot_set_inertia(sys, inertiaR.data(), 0.0), "readSerialization");
    inertiaSet = true;
    const char *names[3] = {"angularVelocityW", "angularMomentumL", "torqueT"};
    const int fields[3] = {PB_FIELD_W, PB_FIELD_L, PB_FIELD_T};
    for (int k = 0; k < 3; ++k) {
      const auto x = a.get<double>(std::string("cudaParticleRotation.") + names[k]);
      if (x.size() != 3 * (size_t)N) throw std::runtime_error("readSerialization: rotation state");
      pbamd::check(pb_sys_set(sys, fields[k], x.data()), "readSerialization");
    }
  }

public:
  // cudaParticleRotation::setInertia (cudaParticleRotation.cu:35-61)
  void setInertia(const std::vector<real> &r0) {
    inertiaR.assign(r0.begin(), r0.end());
    pbamd::check(pb_rot_set_inertia(sys, inertiaR.data(), 0.0), "setInertia");
    inertiaSet = true;
  }
  void setInertia(real r0_all) {
    inertiaR.assign(N, (double)r0_all);
    pbamd::check(pb_rot_set_inertia(sys, nullptr, r0_all), "setInertia");
    inertiaSet = true;
  }
  // cudaParticleRotation::TimeEvolution (cudaParticleRotation.cu:29-33)
  void TimeEvolution(real dt) { pbamd::check(pb_rot_step(sys, dt), "Rotation::TimeEvolution"); }
};

class cudaParticleDEM : public cudaParticleLF,
                        public cudaParticleRotation,
                        virtual public cudaCutoffBlock,
                        public putTMPselected {
protected:
  uint32_t lastNNZ = 0;
  long p1 = 0, p2 = 0;          //!< id range of this device (checkPidRange)
  uint32_t p3 = 0, p4 = 0;      //!< moving-particle id range of this device
  // the arguments of setDEMProperties, kept for restarts
  double demE = 0, demMu = 0, demSigma = 0, demGamma = 0, demMuR = 0;
  std::vector<double> demR;
  bool demSet = false;
  // the arguments of setLinearContact (law 1), kept for restarts
  int demLaw = 0;
  double demKn = 0, demRest = 0, demKt = 0;

public:
  bool autotunetimestep = false;   //!< true when AdaptiveTime drives dt (no Cundall-Strack abort)

  cudaParticleDEM() { SingleParticleBlock = true; }   // cudaParticleDEM.hh:37

  void setup(int n) override { cudaParticleBase::setup(n); }

  // the one-argument form is the DEM force (cudaParticleDEM.hh:53-54)
  void calcForce(void) override {}
  void calcForce(real dt) {
    pbamd::check(pb_dem_calc_force(sys, dt, autotunetimestep ? 1 : 0, &lastNNZ),
                 "cudaParticleDEM::calcForce");
  }
  uint32_t numContacts(void) const { return lastNNZ; }

  // cudaParticleDEM::TimeEvolution: translation (LF) then rotation
  void TimeEvolution(real dt) {
    cudaParticleLF::TimeEvolution(dt);
    cudaParticleRotation::TimeEvolution(dt);
  }

  // both overloads of setDEMProperties (cudaParticleDEM.cu:221-252)
  void setDEMProperties(real E, real mu, real sigma, real gamma, real mu_r, real r0_all) {
    setDEMProperties(E, mu, sigma, gamma, mu_r, std::vector<real>(N, r0_all));
  }
  void setDEMProperties(real E, real mu, real sigma, real gamma, real mu_r,
                        const std::vector<real> &r0) {
    demR.assign(r0.begin(), r0.end());
    demE = E;
    demMu = mu;
    demSigma = sigma;
    demGamma = gamma;
    demMuR = mu_r;
    pbamd::check(pb_dem_set_properties(sys, demE, demMu, demSigma, demGamma, demMuR, demR.data(), 0.0),
                 "setDEMProperties");
    demLaw = 0;
    demSet = true;
  }

  // Linear spring-dashpot contacts (normal stiffness kn, restitution e),
  // tangential spring kt with Coulomb friction mu: the granular settle
  // workload's law, set up in place of setDEMProperties' Hertz law.  No
  // counterpart in the reference's class, which has only the Hertz law.
  void setLinearContact(real kn, real e, real kt, real mu, const std::vector<real> &r0) {
    demR.assign(r0.begin(), r0.end());
    demKn = kn;
    demRest = e;
    demKt = kt;
    demMu = mu;
    pbamd::check(pb_dem_set_linear(sys, demKn, demRest, demKt, demMu, demR.data(), 0.0),
                 "setLinearContact");
    demLaw = 1;
    demSet = true;
  }

  // cudaParticleDEM fields (cudaParticleDEM.hh:124-180) after cudaParticleLF,
  // cudaParticleRotation and the cells: the material constants as given,
  // the radii, and the latest contact matrix (cudaSparseMat) whose val
  // holds the tangential history
  void save(pbamd::OArchive &a) const override {
    saveBase(a);
    saveBlock(a);
    saveRotation(a);
    saveSelected(a);
    a.scalar("cudaParticleDEM.demSet", (uint8_t)demSet);
    a.scalar("cudaParticleDEM.autotunetimestep", (uint8_t)autotunetimestep);
    if (!demSet) return;
    a.scalar("cudaParticleDEM.law", (uint32_t)demLaw);
    a.scalar("cudaParticleDEM.kn", demKn);
    a.scalar("cudaParticleDEM.restitution", demRest);
    a.scalar("cudaParticleDEM.kt", demKt);
    a.scalar("cudaParticleDEM.E", demE);
    a.scalar("cudaParticleDEM.mu", demMu);
    a.scalar("cudaParticleDEM.sigma", demSigma);
    a.scalar("cudaParticleDEM.gamma", demGamma);
    a.scalar("cudaParticleDEM.mu_r", demMuR);
    a.put("cudaParticleDEM.radius", demR);
    uint32_t nnz = 0;
    std::vector<uint32_t> row(N + 1), col;
    std::vector<double> val;
    if (pb_dem_get_contacts(sys, row.data(), nullptr, nullptr, 0, &nnz) != PB_OK) {
      a.scalar("cudaSparseMat.NNZ", (uint32_t)0xFFFFFFFFu);   // no matrix yet
      return;
    }
    uint32_t hw = 1;
    pbamd::check(pb_dem_history_width(sys, &hw), "writeSerialization");
    col.resize(nnz);
    val.resize((size_t)nnz * hw);
    pbamd::check(pb_dem_get_contacts(sys, row.data(), col.data(), val.data(), nnz, &nnz),
                 "writeSerialization");
    a.scalar("cudaSparseMat.NNZ", nnz);
    a.put("cudaSparseMat.rowPtr", row);
    a.put("cudaSparseMat.colIdx", col);
    a.put("cudaSparseMat.val", val);
  }
  void load(const pbamd::IArchive &a) override {
    loadBase(a);
    loadBlock(a);
    loadSelected(a);
    autotunetimestep = a.scalar<uint8_t>("cudaParticleDEM.autotunetimestep") != 0;
    if (a.scalar<uint8_t>("cudaParticleDEM.demSet")) {
      demR = a.get<double>("cudaParticleDEM.radius");
      if (demR.size() != N) throw std::runtime_error("readSerialization: DEM radii");
      demE = a.scalar<double>("cudaParticleDEM.E");
      demMu = a.scalar<double>("cudaParticleDEM.mu");
      demSigma = a.scalar<double>("cudaParticleDEM.sigma");
      demGamma = a.scalar<double>("cudaParticleDEM.gamma");
      demMuR = a.scalar<double>("cudaParticleDEM.mu_r");
      demLaw = (int)a.scalar<uint32_t>("cudaParticleDEM.law");
      demKn = a.scalar<double>("cudaParticleDEM.kn");
      demRest = a.scalar<double>("cudaParticleDEM.restitution");
      demKt = a.scalar<double>("cudaParticleDEM.kt");
      if (demLaw == 1)
        pbamd::check(pb_dem_set_linear(sys, demKn, demRest, demKt, demMu, demR.data(), 0.0),
                     "readSerialization");
      else
        pbamd::check(pb_dem_set_properties(sys, demE, demMu, demSigma, demGamma, demMuR,
                                           demR.data(), 0.0),
                     "readSerialization");
      demSet = true;
    }
    loadRotation(a);
    if (demSet) {
      const uint32_t nnz = a.scalar<uint32_t>("cudaSparseMat.NNZ");
      if (nnz != 0xFFFFFFFFu) {
        const auto row = a.get<uint32_t>("cudaSparseMat.rowPtr");
        const auto col = a.get<uint32_t>("cudaSparseMat.colIdx");
        const auto val = a.get<double>("cudaSparseMat.val");
        pbamd::check(pb_dem_set_contacts(sys, row.data(), col.data(), val.data(), nnz),
                     "readSerialization");
      }
    }
  }

  // selectBlocks (cudaSelectedBlock.cu:18-44) marks the cells holding
  // movable particles for the multi-GPU split; the contact search here
  // skips fixed pairs by itself, so the call has nothing to prepare
  void selectBlocks(void) {}
  // numSelectedBlocks (cudaSelectedBlock.hh:47): cells of the last
  // calcBlockID that hold at least one movable particle
  long numSelectedBlocks(void) {
    std::vector<uint32_t> pid(N), bid(N), bindex(totalNumBlock + 1);
    pbamd::check(pb_get_bins(sys, pid.data(), bid.data(), bindex.data()), "numSelectedBlocks");
    std::vector<uint8_t> mv(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_MOVE, mv.data()), "numSelectedBlocks");
    long n = 0;
    uint32_t last = UINT32_MAX;
    for (uint32_t k = 0; k < N; ++k)      // bid is sorted: one count per cell
      if (mv[pid[k]] && bid[k] != last) {
        ++n;
        last = bid[k];
      }
    return n;
  }
  // checkPidRange (cudaSelectedBlock.hh:50-61): the id range [p1, p2) and
  // the moving-particle range [p3, p4) of device devId of ndev
  void checkPidRange(int devId = 0, int ndev = 1, uint32_t Mstart = 0, uint32_t Mend = UINT32_MAX) {
    if (ndev != 1 || devId != 0) {
      std::cerr << "checkPidRange: one GPU owns every particle in this build" << std::endl;
      std::exit(1);
    }
    p1 = 0;
    p2 = N;
    p3 = Mstart;
    p4 = Mend == UINT32_MAX ? N : Mend;
    std::cerr << "prange:" << p1 << ":" << p2 << ", " << p3 << ":" << p4 << std::endl;
  }
  void calcBlockID(void) { cudaCutoffBlock::calcBlockID(); }

  // cudaParticleDEM::rollback / rollback2 (cudaParticleDEM.cu:302-310):
  // both undo the rotation step; rollback also rewinds the contact matrix so
  // the next calcForce carries the history of the step before the discarded one
  void rollback(real dt) {
    cudaParticleLF::rollback(dt);
    cudaParticleRotation::TimeEvolution(-dt);
    pbamd::check(pb_dem_rewind(sys), "cudaParticleDEM::rollback");
  }
  void rollback2(real dt) {
    cudaParticleLF::calcVinit(dt / 2);
    cudaParticleRotation::TimeEvolution(-dt);
  }
};
