// cudaCutoffBlock.hh -- cell grid ("cutoff blocks") of the reference
// (cudaCutoffBlock.hh:241-240, cudaCutoffBlock.cu:75-307).  One GPU holds the
// whole domain, so only the full block range is accepted; the force and
// acceleration imports merge another object's arrays as the reference does
// (its peer copy becomes a read of the other object on the same GPU).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include "cudaParticleBase.hh"

enum class ParticleBlockType { many, single };

class cudaCutoffBlock : public virtual cudaParticleBase {
protected:
  std::vector<real> blocklen = std::vector<real>(6, 0);
  std::vector<uint32_t> blocknum = std::vector<uint32_t>(3, 0);
  uint32_t totalNumBlock = 0;
  bool SingleParticleBlock = false;
  bool uninitialized = true;
  bool periodicCells = true;
  real rmaxCells = 0;

  // cudaCutoffBlock fields (cudaCutoffBlock.hh:135-240); the sort arrays
  // (bid, pid, bindex) and the neighbour table are not stored: the grid is
  // set up again from rmax and the next calcBlockID rebuilds them
  void saveBlock(pbamd::OArchive &a) const {
    a.scalar("cudaCutoffBlock.totalNumBlock", totalNumBlock);
    a.put("cudaCutoffBlock.blocklen", blocklen);
    a.put("cudaCutoffBlock.blocknum", blocknum);
    a.scalar("cudaCutoffBlock.SingleParticleBlock", (uint8_t)SingleParticleBlock);
    a.scalar("cudaCutoffBlock.uninitialized", (uint8_t)uninitialized);
    a.scalar("cudaCutoffBlock.rmax", rmaxCells);
    a.scalar("cudaCutoffBlock.periodic", (uint8_t)periodicCells);
  }
  void loadBlock(const pbamd::IArchive &a) {
    SingleParticleBlock = a.scalar<uint8_t>("cudaCutoffBlock.SingleParticleBlock") != 0;
    if (a.scalar<uint8_t>("cudaCutoffBlock.uninitialized")) return;
    setupCutoffBlock(a.scalar<real>("cudaCutoffBlock.rmax"),
                     a.scalar<uint8_t>("cudaCutoffBlock.periodic") != 0);
    if (a.get<uint32_t>("cudaCutoffBlock.blocknum") != blocknum)
      throw std::runtime_error("readSerialization: cell grid differs");
  }

  void importField(const cudaCutoffBlock &A, int field, bool directAccess, const char *what) {
    if (A.N != N) throw std::runtime_error(std::string(what) + ": particle counts differ");
    std::vector<double> mine(3 * (size_t)N), other(3 * (size_t)N);
    pbamd::check(pb_sys_get(sys, field, mine.data()), what);
    if (directAccess) {
      pbamd::check(pb_sys_get(A.sys, field, other.data()), what);
    } else {
      if (A.TMP.size() < 4 * (size_t)N) throw std::runtime_error(std::string(what) + ": TMP");
      for (size_t i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) other[3 * i + c] = (double)A.TMP[4 * i + c];
    }
    for (size_t k = 0; k < mine.size(); ++k) mine[k] += other[k];
    pbamd::check(pb_sys_set(sys, field, mine.data()), what);
  }

public:
  // cudaCutoffBlock::setupCutoffBlock, cudaCutoffBlock.cu:75-231
  void setupCutoffBlock(real rmax, bool periodic = true) {
    pbamd::check(pb_setup_cells(sys, rmax, periodic ? 1 : 0, SingleParticleBlock ? 1 : 0),
                 "setupCutoffBlock");
    float bl[6];
    pbamd::check(pb_get_cell_info(sys, blocknum.data(), bl), "setupCutoffBlock");
    for (int i = 0; i < 6; ++i) blocklen[i] = bl[i];
    totalNumBlock = blocknum[0] * blocknum[1] * blocknum[2];
    periodicCells = periodic;
    rmaxCells = rmax;
    uninitialized = false;
  }
  // calcBID_F4 + sort + makeBindex (cudaCutoffBlock.cu:233-258)
  void calcBlockID(void) { pbamd::check(pb_bin(sys), "calcBlockID"); }
  uint32_t numBlocks(void) { return totalNumBlock; }
  // the whole domain is on this GPU: only the full range is accepted
  void setBlockRange(uint32_t blockNum, uint32_t n, uint32_t myID) {
    uint32_t off = 0, num = 0;
    calcBlockRange(blockNum, n, myID, [&](uint32_t o, uint32_t k) { off = o; num = k; });
    if (off != 0 || num != totalNumBlock) {
      std::cerr << "setBlockRange: one GPU owns every block in this build" << std::endl;
      std::exit(1);
    }
  }
  void switchBlockAlgorithm(const ParticleBlockType t) {
    if (uninitialized) SingleParticleBlock = (t == ParticleBlockType::single);
    else std::cerr << "cudaCutoffBlock::switchBlockAlgorithm() rejects" << std::endl;
  }
  // calcBlockRange (cudaCutoffBlock.hh:55-70): split blockNum cells over n
  // devices, the first blockNum % n taking one extra; setPara(offset, count)
  // receives device myID's share
  template <typename F>
  static void calcBlockRange(uint32_t blockNum, uint32_t n, uint32_t myID, F setPara) {
    if (n == 0 || myID >= n) {
      std::cerr << "calcBlockRange: device " << myID << " of " << n << std::endl;
      std::exit(1);
    }
    const uint32_t base = blockNum / n, extra = blockNum % n;
    const uint32_t count = base + (myID < extra ? 1u : 0u);
    const uint32_t offset = myID * base + (myID < extra ? myID : extra);
    setPara(offset, count);
  }
  // importAcceleration / importForce (cudaCutoffBlock.cu:260-286): add the
  // other object's a or F (directAccess: its device array; otherwise its
  // TMP host array, float4 per particle) to this object's a or F.
  // idMe / idPeer name GPUs of a multi-GPU run and have no effect here.
  void importAcceleration(const cudaCutoffBlock &A, bool directAccess = false, int idMe = 0,
                          int idPeer = 0) {
    importField(A, PB_FIELD_A, directAccess, "importAcceleration");
    (void)idMe;
    (void)idPeer;
  }
  void importForce(const cudaCutoffBlock &A, bool directAccess = false, int idMe = 0,
                   int idPeer = 0) {
    importField(A, PB_FIELD_F, directAccess, "importForce");
    (void)idMe;
    (void)idPeer;
  }

  // downloads of the sort (pid, sorted bid, bindex with UINT32_MAX = empty)
  void getBlocks(std::vector<uint32_t> &pid, std::vector<uint32_t> &bid,
                 std::vector<uint32_t> &bindex) {
    pid.resize(N);
    bid.resize(N);
    bindex.resize(totalNumBlock + 1);
    pbamd::check(pb_get_bins(sys, pid.data(), bid.data(), bindex.data()), "getBlocks");
  }
};
