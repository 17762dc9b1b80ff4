// Binary RDM files (restating read_rdms_binary / write_rdms_binary,
// src/macis/fcidump.cxx:209-282): an int32 norb, then the 1-RDM (norb^2)
// and the 2-RDM (norb^4) as column-major doubles. The in-memory arrays
// have leading dimensions LDD1 / LDD2; reading zeroes them first.
#pragma once
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace macis {

inline void read_rdms_binary(std::string fname, size_t norb, double* ORDM, size_t LDD1,
                             double* TRDM, size_t LDD2) {
  for(size_t j = 0; j < norb; ++j)
    for(size_t i = 0; i < norb; ++i) ORDM[i + j * LDD1] = 0;
  for(size_t l = 0; l < norb; ++l)
    for(size_t k = 0; k < norb; ++k)
      for(size_t j = 0; j < norb; ++j)
        for(size_t i = 0; i < norb; ++i) TRDM[i + LDD2 * (j + LDD2 * (k + LDD2 * l))] = 0;
  std::ifstream in(fname, std::ios::binary);
  if(!in) throw std::runtime_error(fname + " not available");
  int32_t nread = 0;
  in.read(reinterpret_cast<char*>(&nread), sizeof(int32_t));
  if(nread != (int32_t)norb)
    throw std::runtime_error("NORB in RDM file doesn't match " + std::to_string(norb) + " " +
                             std::to_string(nread));
  const size_t n2 = norb * norb;
  std::vector<double> raw(n2 * n2);
  in.read(reinterpret_cast<char*>(raw.data()), n2 * sizeof(double));
  for(size_t j = 0; j < norb; ++j)
    for(size_t i = 0; i < norb; ++i) ORDM[i + j * LDD1] = raw[i + j * norb];
  in.read(reinterpret_cast<char*>(raw.data()), n2 * n2 * sizeof(double));
  for(size_t l = 0; l < norb; ++l)
    for(size_t k = 0; k < norb; ++k)
      for(size_t j = 0; j < norb; ++j)
        for(size_t i = 0; i < norb; ++i)
          TRDM[i + LDD2 * (j + LDD2 * (k + LDD2 * l))] = raw[i + norb * (j + norb * (k + norb * l))];
}

inline void write_rdms_binary(std::string fname, size_t norb, const double* ORDM, size_t LDD1,
                              const double* TRDM, size_t LDD2) {
  std::ofstream out(fname, std::ios::binary);
  if(!out) throw std::runtime_error(fname + " not available");
  const int32_t nw = (int32_t)norb;
  out.write(reinterpret_cast<const char*>(&nw), sizeof(int32_t));
  const size_t n2 = norb * norb;
  std::vector<double> raw(n2 * n2);
  for(size_t j = 0; j < norb; ++j)
    for(size_t i = 0; i < norb; ++i) raw[i + j * norb] = ORDM[i + j * LDD1];
  out.write(reinterpret_cast<const char*>(raw.data()), n2 * sizeof(double));
  for(size_t l = 0; l < norb; ++l)
    for(size_t k = 0; k < norb; ++k)
      for(size_t j = 0; j < norb; ++j)
        for(size_t i = 0; i < norb; ++i)
          raw[i + norb * (j + norb * (k + norb * l))] = TRDM[i + LDD2 * (j + LDD2 * (k + LDD2 * l))];
  out.write(reinterpret_cast<const char*>(raw.data()), n2 * n2 * sizeof(double));
}

}  // namespace macis
