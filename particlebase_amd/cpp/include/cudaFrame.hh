// cudaFrame.hh -- base of the class surface (reference cudaFrame.hh:20-112):
// cell bounds, advisory launch-shape fields, ErrorInfo and the asynchronous
// LAMMPS dump writer.  Launch shapes are chosen inside the engine; MPnum,
// THnum* stay as fields so drivers that set or print them still compile.
#pragma once
#include <cstdint>
#include <ctime>
#include <mutex>
#include <ostream>
#include <string>
#include <thread>
#include <valarray>
#include <vector>
#include "pbamd_env.hh"

class cudaFrame {
protected:
  bool withInfo = true;
  std::valarray<real> TMP;               //!< host staging (4 per particle)
  std::valarray<unsigned short> TMP2;    //!< type ids for the dump
  real cell[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   //!< bounds + (L/2)^2
  std::mutex mutTMP;
  std::thread writer;

public:
  int MPnum = 0, THnum1D = 0, THnum1DX = 0, THnum2D = 0;
  uint32_t threadsMax = 0, maxGrid = 0;
  uint32_t timestep = 0;

  virtual ~cudaFrame() { waitPutTMP(); }

  void ErrorInfo(const std::string &s) const {
    const char *e = pb_ctx_last_error(pbamd::context());
    if (e && *e) {
      std::cerr << s << ": " << e << std::endl;
      std::exit(1);
    }
  }

  // nowtime (cudaFrame.cu:93-97): print the wall-clock date, return it
  time_t nowtime(std::ostream &o = std::cout) {
    const time_t T = time(nullptr);
    o << ctime(&T);
    return T;
  }

  // cudaFrame::setCell, cudaFrame.cu:85-91
  void setCell(const real *_cell) {
    for (int i = 0; i < 6; ++i) cell[i] = _cell[i];
    for (int i = 0; i < 3; ++i) {
      cell[i + 6] = (cell[i * 2 + 1] - cell[i * 2]) / 2;
      cell[i + 6] *= cell[i + 6];
    }
    onCell();
  }

  virtual std::string additionalTag(void) const { return ""; }
  virtual std::string additionalOutput(uint32_t) const { return ""; }

  // putTMP / putTMPinThread (cudaFrame.cu:34-83): LAMMPS dump of TMP/TMP2
  // written by a background thread; waitPutTMP joins it.
  void putTMP(std::ostream &o) {
    waitPutTMP();
    writer = std::thread([this, &o]() {
      std::lock_guard<std::mutex> lk(mutTMP);
      const size_t N = TMP2.size();
      o << "ITEM: TIMESTEP\n" << timestep << "\nITEM: NUMBER OF ATOMS\n" << N
        << "\nITEM: BOX BOUNDS\n" << cell[0] << " " << cell[1] << "\n" << cell[2] << " "
        << cell[3] << "\n" << cell[4] << " " << cell[5] << "\nITEM: ATOMS id type x y z"
        << additionalTag() << "\n";
      for (size_t i = 0; i < N; ++i)
        o << i << "  " << TMP2[i] << " " << TMP[i * 4] << " " << TMP[i * 4 + 1] << " "
          << TMP[i * 4 + 2] << additionalOutput((uint32_t)i) << "\n";
      o.flush();
    });
  }
  void waitPutTMP(void) {
    if (writer.joinable()) writer.join();
  }

protected:
  virtual void onCell() {}
};
