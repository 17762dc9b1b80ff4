// ASCII wavefunction files (restating include/macis/wavefunction_io.hpp):
// a header line "nstates norb nalpha nbeta", then one line per determinant,
// "coefficient canonical-string", the coefficient in scientific notation
// with 16 digits after the point, right-aligned in 30 columns.
#pragma once
#include <bitset>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <macis/sd_operations.hpp>

namespace macis {

template <size_t N>
void read_wavefunction(std::string fname, std::vector<std::bitset<N>>& states,
                       std::vector<double>& coeffs) {
  states.clear();
  coeffs.clear();
  std::ifstream file(fname);
  std::string line;
  std::getline(file, line);
  std::istringstream head(line);
  std::string nstate, norb, nalpha, nbeta;
  head >> nstate >> norb >> nalpha >> nbeta;
  const size_t n = std::stoi(nstate);  // the header is parsed as the reference does
  (void)std::stoi(norb);
  (void)std::stoi(nalpha);
  (void)std::stoi(nbeta);
  states.reserve(n);
  coeffs.reserve(n);
  while(std::getline(file, line)) {
    std::istringstream ss(line);
    std::string c, d;
    ss >> c >> d;
    states.emplace_back(from_canonical_string<N>(d));
    coeffs.emplace_back(std::stod(c));
  }
}

template <size_t N>
void write_wavefunction(std::string fname, size_t norb, const std::vector<std::bitset<N>>& states,
                        const std::vector<double>& coeffs) {
  if(states.size() != coeffs.size()) throw std::runtime_error("Invalid Wave Function Dimensions");
  if(!states.size()) return;
  const auto nalpha = (states[0] << N / 2).count();
  const auto nbeta = (states[0] >> N / 2).count();
  std::ofstream file(fname);
  file << states.size() << " " << norb << " " << nalpha << " " << nbeta << std::endl;
  file << std::scientific << std::setprecision(16);
  for(size_t i = 0; i < states.size(); ++i)
    file << std::setw(30) << coeffs[i] << " " << to_canonical_string(states[i]) << " "
         << std::endl;
}

}  // namespace macis
