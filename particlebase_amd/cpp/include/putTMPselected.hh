// putTMPselected.hh -- dumps of one contiguous particle-id range (reference
// putTMPselected.hh/.cu): the "selected" range holds the moving particles
// and replaces getPosition / getTypeID for putTMP; the "unselected" range
// (walls, fixed particles) is written once by putUnSelected.  Ranges are
// (first id, count) in the order the particles were imported.
#pragma once
#include <cstdint>
#include <fstream>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include "Serialization.hh"
#include "cudaParticleBase.hh"

class putTMPselected : public virtual cudaParticleBase {
protected:
  std::vector<uint32_t> selected{0, 0};     //!< first id and count of the moving particles
  std::vector<uint32_t> unselected{0, 0};   //!< first id and count of the fixed particles

  void checkRange(const std::vector<uint32_t> &s, const char *what) const {
    if ((uint64_t)s[0] + s[1] > N) throw std::runtime_error(std::string(what) + ": range beyond N");
  }
  // positions of ids [s0, s0+cnt) into TMP[0 .. 4 cnt), FP32 as the reference
  void fetchRange(uint32_t s0, uint32_t cnt) {
    const std::vector<double> r = getState(PB_FIELD_R);
    std::lock_guard<std::mutex> lk(mutTMP);
    for (uint32_t i = 0; i < cnt; ++i) {
      for (int k = 0; k < 3; ++k) TMP[4 * (size_t)i + k] = (real)r[3 * ((size_t)s0 + i) + k];
      TMP[4 * (size_t)i + 3] = 0;
    }
  }
  std::vector<uint16_t> typesRange(uint32_t s0, uint32_t cnt) {
    std::vector<uint16_t> t(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_TYPE, t.data()), "getSelectedTypeID");
    return std::vector<uint16_t>(t.begin() + s0, t.begin() + s0 + cnt);
  }
  void saveSelected(pbamd::OArchive &a) const {
    a.put("putTMPselected.selected", selected);
    a.put("putTMPselected.unselected", unselected);
  }
  void loadSelected(const pbamd::IArchive &a) {
    if (!a.has("putTMPselected.selected")) return;
    selected = a.get<uint32_t>("putTMPselected.selected");
    unselected = a.get<uint32_t>("putTMPselected.unselected");
    if (selected.size() != 2 || unselected.size() != 2)
      throw std::runtime_error("readSerialization: putTMPselected ranges");
  }

public:
  // setupSelectedTMP (putTMPselected.cu:4-10): (start, count) of the
  // selected and of the unselected particles
  void setupSelectedTMP(uint32_t s1, uint32_t s2, uint32_t u1, uint32_t u2) {
    selected = {s1, s2};
    unselected = {u1, u2};
    checkRange(selected, "setupSelectedTMP");
    checkRange(unselected, "setupSelectedTMP");
  }
  // getSelectedTypeID (putTMPselected.cu:12-17): TMP2 holds the selected
  // particles' types, so putTMP writes those particles only
  void getSelectedTypeID(void) {
    const std::vector<uint16_t> t = typesRange(selected[0], selected[1]);
    std::lock_guard<std::mutex> lk(mutTMP);
    TMP2.resize(selected[1]);
    for (uint32_t i = 0; i < selected[1]; ++i) TMP2[i] = t[i];
  }
  // getSelectedPosition (putTMPselected.cu:19-28)
  virtual void getSelectedPosition(void) { fetchRange(selected[0], selected[1]); }
  // putUnSelected (putTMPselected.cu:30-63): LAMMPS dump of the unselected
  // range, timestep 0, ids from 0
  void putUnSelected(const char *filename) {
    waitPutTMP();
    const std::vector<uint16_t> t = typesRange(unselected[0], unselected[1]);
    fetchRange(unselected[0], unselected[1]);
    std::ofstream o(filename);
    if (!o.is_open()) {
      std::cerr << "cannot open " << filename << std::endl;
      return;
    }
    const uint32_t n = unselected[1];
    o << "ITEM: TIMESTEP" << std::endl << 0 << std::endl
      << "ITEM: NUMBER OF ATOMS" << std::endl << n << std::endl
      << "ITEM: BOX BOUNDS" << std::endl
      << cell[0] << " " << cell[1] << std::endl
      << cell[2] << " " << cell[3] << std::endl
      << cell[4] << " " << cell[5] << std::endl
      << "ITEM: ATOMS id type x y z" << std::endl;
    for (uint32_t i = 0; i < n; ++i)
      o << i << "  " << t[i] << " " << TMP[4 * (size_t)i] << " " << TMP[4 * (size_t)i + 1] << " "
        << TMP[4 * (size_t)i + 2] << std::endl;
  }
};
