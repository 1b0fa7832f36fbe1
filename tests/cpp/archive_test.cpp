// Round trip of the checkpoint container (particlebase_amd/cpp/include/
// Serialization.hh): typed, counted, named fields written and read back;
// wrong types, missing fields, a foreign file and a truncated file are
// errors.  Host only.
#include <cstdio>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>
#include "../../particlebase_amd/cpp/include/Serialization.hh"

static int fails = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::cout << "FAIL line " << __LINE__ << ": " #c << std::endl; \
      ++fails;                                                       \
    }                                                                \
  } while (0)

template <typename F>
static bool throws(F f) {
  try {
    f();
  } catch (const std::exception &) {
    return true;
  }
  return false;
}

int main(int argc, char **argv) {
  const std::string fn = argc > 1 ? argv[1] : "archive_test.ck";
  const std::vector<double> r = {1.5, -2.25, 3e-300, 1e300, -0.0};
  const std::vector<uint16_t> t = {0, 1, 65535};
  const std::vector<uint8_t> mv = {1, 0, 1, 1};
  const float cell[9] = {0, 1, 0, 2, 0, 3, 0.25f, 1, 2.25f};
  {
    pbamd::OArchive a(fn);
    a.put("cudaFrame.cellsize", cell, 9);
    a.scalar("cudaFrame.timestep", (uint32_t)1234);
    a.put("cudaParticleBase.positionR", r);
    a.put("cudaParticleBase.typeID", t);
    a.put("cudaParticleBase.moveFlag", mv);
    a.put("empty", std::vector<uint32_t>());
    a.scalar("GaussianThermo.xi", -0.125);
  }
  pbamd::IArchive b(fn);
  CHECK(b.get<double>("cudaParticleBase.positionR") == r);
  CHECK(b.get<uint16_t>("cudaParticleBase.typeID") == t);
  CHECK(b.get<uint8_t>("cudaParticleBase.moveFlag") == mv);
  CHECK(b.get<float>("cudaFrame.cellsize") == std::vector<float>(cell, cell + 9));
  CHECK(b.scalar<uint32_t>("cudaFrame.timestep") == 1234u);
  CHECK(b.scalar<double>("GaussianThermo.xi") == -0.125);
  CHECK(b.get<uint32_t>("empty").empty());
  CHECK(b.has("cudaParticleBase.typeID") && !b.has("nope"));
  CHECK(b.names().size() == 7);
  CHECK(throws([&] { b.get<float>("cudaParticleBase.positionR"); }));
  CHECK(throws([&] { b.get<double>("nope"); }));
  CHECK(throws([&] { b.scalar<double>("cudaParticleBase.positionR"); }));
  // a foreign file and a truncated one
  {
    std::ofstream o(fn + ".x", std::ios::binary);
    o << "not a checkpoint at all";
  }
  CHECK(throws([&] { pbamd::IArchive c(fn + ".x"); }));
  {
    std::ifstream in(fn, std::ios::binary);
    std::string all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::ofstream o(fn + ".t", std::ios::binary);
    o.write(all.data(), (std::streamsize)(all.size() - 7));
  }
  CHECK(throws([&] { pbamd::IArchive c(fn + ".t"); }));
  std::remove((fn + ".x").c_str());
  std::remove((fn + ".t").c_str());
  std::cout << (fails ? "FAILED" : "OK") << std::endl;
  return fails ? 1 : 0;
}
