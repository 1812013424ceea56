// Minimal .npz (zip of .npy) reader for TensorMD model files.
// Handles STORED and DEFLATE members (zlib) and zip64 size fields, which is
// what numpy.savez / savez_compressed and the reference's cnpy writer
// (src/ML-TENSORMD/cnpy.cpp) produce.  The reference reader is cnpy::npz_load
// (src/ML-TENSORMD/cnpy.h:27-72); this is an independent implementation.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace tmd {

struct NpyArray {
  std::vector<size_t> shape;
  char kind = 0;    // 'f', 'i', 'u', 'b'
  int word = 0;     // bytes per element
  bool fortran = false;
  std::vector<unsigned char> bytes;
  size_t size() const
  {
    size_t n = 1;
    for (size_t s : shape) n *= s;
    return n;
  }
  double get_double(size_t i) const;
  long long get_int(size_t i) const;
  std::vector<double> as_doubles() const;
  std::vector<int> as_ints() const;
};

using Npz = std::map<std::string, NpyArray>;

// Returns true on success; on failure `err` holds the reason.
bool read_npz(const std::string &path, Npz &out, std::string &err);

}  // namespace tmd
