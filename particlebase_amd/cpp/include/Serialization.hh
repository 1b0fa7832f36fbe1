// Serialization.hh -- checkpoint / restart files of the class surface.
//
// The reference writes Boost binary archives (readSerialization /
// writeSerialization, CUDAenv.hh:268-285; field lists in
// cudaParticleBase.hh:228-354, cudaCutoffBlock.hh:135-240,
// cudaParticleMD.hh:195-240, GaussianThermo.hh:92-120,
// cudaParticleRotation.hh:68-125, cudaParticleDEM.hh:124-180,
// cudaParticleSPHBase.hh:83-106, cudaParticleSPH_NS.hh:129-175,
// cudaSparseMat.hh:70-122).  Boost is not part of this build, so the same
// field set is kept in a self-describing binary: the magic "PBAMDAR1", then
// records of
//   u16 name length, name ("<class>.<field>", the reference's nvp names),
//   u8 element type, u64 element count, the elements (little endian),
// ended by a record with an empty name.  A reader looks fields up by name
// and checks their type, so a file written by one class hierarchy can be
// inspected by any tool.
#pragma once
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pbamd {

template <typename T> struct ArType;
template <> struct ArType<double>   { static constexpr uint8_t code = 1; };
template <> struct ArType<float>    { static constexpr uint8_t code = 2; };
template <> struct ArType<uint32_t> { static constexpr uint8_t code = 3; };
template <> struct ArType<uint16_t> { static constexpr uint8_t code = 4; };
template <> struct ArType<uint8_t>  { static constexpr uint8_t code = 5; };
template <> struct ArType<int32_t>  { static constexpr uint8_t code = 6; };
template <> struct ArType<uint64_t> { static constexpr uint8_t code = 7; };

static constexpr char kArMagic[8] = {'P', 'B', 'A', 'M', 'D', 'A', 'R', '1'};

class OArchive {
  std::ofstream o;

public:
  explicit OArchive(const std::string &fn) : o(fn, std::ios::binary) {
    if (!o) throw std::runtime_error("cannot open " + fn + " for writing");
    o.write(kArMagic, 8);
  }
  ~OArchive() {
    const uint16_t z = 0;
    o.write((const char *)&z, 2);
  }
  template <typename T>
  void put(const std::string &name, const T *p, size_t n) {
    if (name.empty() || name.size() > 65535) throw std::runtime_error("bad field name");
    const uint16_t l = (uint16_t)name.size();
    const uint8_t t = ArType<T>::code;
    const uint64_t c = n;
    o.write((const char *)&l, 2);
    o.write(name.data(), l);
    o.write((const char *)&t, 1);
    o.write((const char *)&c, 8);
    if (n) o.write((const char *)p, (std::streamsize)(sizeof(T) * n));
    if (!o) throw std::runtime_error("write failed: " + name);
  }
  template <typename T>
  void put(const std::string &name, const std::vector<T> &v) { put(name, v.data(), v.size()); }
  template <typename T>
  void scalar(const std::string &name, T x) { put(name, &x, 1); }
};

class IArchive {
  struct Field {
    uint8_t type = 0;
    uint64_t count = 0;
    std::vector<char> bytes;
  };
  std::map<std::string, Field> f;
  std::string fn;

public:
  explicit IArchive(const std::string &name) : fn(name) {
    std::ifstream in(name, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + name);
    char m[8];
    in.read(m, 8);
    if (!in || std::memcmp(m, kArMagic, 8) != 0) throw std::runtime_error(name + ": not a checkpoint");
    for (;;) {
      uint16_t l = 0;
      in.read((char *)&l, 2);
      if (!in) throw std::runtime_error(name + ": truncated");
      if (l == 0) break;
      std::string key(l, '\0');
      in.read(&key[0], l);
      Field x;
      in.read((char *)&x.type, 1);
      in.read((char *)&x.count, 8);
      static const size_t sz[8] = {0, 8, 4, 4, 2, 1, 4, 8};
      if (!in || x.type == 0 || x.type > 7) throw std::runtime_error(name + ": bad record " + key);
      x.bytes.resize((size_t)(sz[x.type] * x.count));
      if (x.count) in.read(x.bytes.data(), (std::streamsize)x.bytes.size());
      if (!in) throw std::runtime_error(name + ": truncated in " + key);
      f[key] = std::move(x);
    }
  }
  bool has(const std::string &name) const { return f.count(name) != 0; }
  std::vector<std::string> names(void) const {
    std::vector<std::string> k;
    for (const auto &e : f) k.push_back(e.first);
    return k;
  }
  template <typename T>
  std::vector<T> get(const std::string &name) const {
    auto it = f.find(name);
    if (it == f.end()) throw std::runtime_error(fn + ": no field " + name);
    if (it->second.type != ArType<T>::code) throw std::runtime_error(fn + ": wrong type of " + name);
    std::vector<T> v(it->second.count);
    if (!v.empty()) std::memcpy(v.data(), it->second.bytes.data(), sizeof(T) * v.size());
    return v;
  }
  template <typename T>
  T scalar(const std::string &name) const {
    const std::vector<T> v = get<T>(name);
    if (v.size() != 1) throw std::runtime_error(fn + ": " + name + " is not a scalar");
    return v[0];
  }
};

}  // namespace pbamd
