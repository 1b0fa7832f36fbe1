// cudaCutoffBlock.hh -- cell grid ("cutoff blocks") of the reference
// (cudaCutoffBlock.hh:241-240, cudaCutoffBlock.cu:75-307).  One GPU holds the
// whole domain, so the multi-GPU block ranges and force imports are refused.
#pragma once
#include <cstdint>
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
    if (n != 1 || myID != 0 || blockNum != totalNumBlock) {
      std::cerr << "setBlockRange: one GPU owns every block in this build" << std::endl;
      std::exit(1);
    }
  }
  void switchBlockAlgorithm(const ParticleBlockType t) {
    if (uninitialized) SingleParticleBlock = (t == ParticleBlockType::single);
    else std::cerr << "cudaCutoffBlock::switchBlockAlgorithm() rejects" << std::endl;
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
