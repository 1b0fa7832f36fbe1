// cudaCalcGR.hh -- radial distribution function (reference cudaCalcGR.hh/.cu).
// The engine bins every unordered pair directly, so makePairInfo's host-side
// tile list is not needed; the integer histogram equals the reference's
// (bin 0 dropped) and igr2gr normalises as kigr2gr does.
#pragma once
#include <cmath>
#include <ostream>
#include <vector>
#include "cudaFrame.hh"
#include "cudaParticleMD.hh"

class cudaCalcGR : public cudaFrame {
public:
  // cudaCalcGR::setup, cudaCalcGR.hh:25-34
  void setup(const real rstep, const real rmax) {
    r2max = rmax * rmax;
    rstepv = rstep;
    rstepinv = (real)(1.0 / (double)rstep);
    rnum = static_cast<uint32_t>(rmax * rstepinv);
    igr.assign(rnum, 0);
    gr.assign(rnum, 0.0f);
  }
  void copyCell(const cudaParticleMD &P) {
    (void)P;
  }
  void makePairInfo(const cudaParticleMD &) {}
  void cleargr(void) { std::fill(igr.begin(), igr.end(), 0); }
  // calcgr: accumulate the pair histogram of the particles' current positions
  void calcgr(cudaParticleMD &part) {
    std::vector<uint64_t> h(rnum);
    pbamd::check(pb_gr_hist(part.handle(), &rstepv, rnum, h.data()), "calcgr");
    for (uint32_t j = 1; j < rnum; ++j) igr[j] += h[j];   // bin 0 is dropped
    np = part.size();
  }
  // igr2gr / kigr2gr, cudaCalcGR.cu:14-28, 115-132
  void igr2gr(cudaParticleMD &part, const real *box) {
    const real rhoinv = (box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]) / part.size();
    const real Vinv = M_1_PI * rstepinv * rstepinv * rstepinv;
    const real Q = 2.0 / part.size() * Vinv * rhoinv;
    const real onethird = 1.0 / 3.0;
    for (uint32_t j = 1; j < rnum; ++j) {
      const real rinv2 = 1.0 / (4 * j * j + onethird);
      gr[j] = static_cast<real>(igr[j]) * Q * rinv2;
    }
    if (rnum) gr[0] = 0;
  }
  void getGr(std::ostream &o = std::cout) {
    for (uint32_t i = 0; i < rnum; ++i) o << i / rstepinv << "  " << gr[i] << std::endl;
  }
  const std::vector<uint64_t> &histogram(void) const { return igr; }
  const std::vector<real> &values(void) const { return gr; }

private:
  real rstepinv = 0, r2max = 0, rstepv = 0;
  uint32_t rnum = 0, np = 0;
  std::vector<uint64_t> igr;
  std::vector<real> gr;
};
