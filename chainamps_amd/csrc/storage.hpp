// Reference binary formats (.MPX_matrix / .MPS_matrix, basis, .UNITCELL).
#pragma once
#include <string>
#include <vector>

#include "bstensor.hpp"

namespace cam {

// Basis: charge rules, the charges of the d physical states, their energies.
struct StoredBasis {
  Rules rules;
  std::vector<Charge> charges;
  std::vector<double> energies;
};

void store_tensor(const std::string& path, const BSTensor& t, hipStream_t s);
// physical indices take the charges `phys` (the reference reads them from the basis)
TensorPtr load_tensor(const std::string& path, const Rules& rules, const std::vector<Charge>& phys, hipStream_t s);
void store_unitcell(const std::string& path, const StoredBasis& basis, const std::vector<TensorPtr>& matrices,
                    const std::vector<std::vector<double>>& lambdas, hipStream_t s);
void load_unitcell(const std::string& path, StoredBasis& basis, std::vector<TensorPtr>& matrices,
                   std::vector<std::vector<double>>& lambdas, hipStream_t s);

}  // namespace cam
