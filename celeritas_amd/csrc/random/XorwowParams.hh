//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file random/XorwowParams.hh
//---------------------------------------------------------------------------//
#pragma once

#include <vector>

#include "XorwowEngine.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Shared XORWOW data: the seed and the jump ladders.
 *
 * The jump polynomials are derived at construction: the characteristic
 * polynomial p(z) of the 160-bit xorshift transform is found with
 * Berlekamp-Massey over F2, then z^(4^i) and z^(2^67 * 4^i) are reduced
 * modulo p(z) by repeated squaring. Same role as XorwowRngParams
 * (celeritas/random/XorwowRngParams.cc:30-125), which stores them as a table.
 */
class XorwowParams
{
  public:
    explicit XorwowParams(std::uint32_t seed);

    std::uint32_t seed() const { return seed_; }
    XorwowJumpTables const& tables() const { return tables_; }

    //! Characteristic polynomial of the transform (bits 0..160)
    std::vector<std::uint32_t> const& char_poly() const { return charpoly_; }

  private:
    std::uint32_t seed_;
    XorwowJumpTables tables_;
    std::vector<std::uint32_t> charpoly_;
};

//---------------------------------------------------------------------------//
/*!
 * Fill states with well-distributed data from a 32-bit Mersenne twister.
 *
 * Same stream as initialize_xorwow (celeritas/random/XorwowRngData.cc:28-60):
 * std::seed_seq over {seed} (plus the stream index for streams other than
 * 0), std::mt19937, and full-range 32-bit uniform draws, in the order
 * x,y,z,w,v,d for each slot. \c out receives 6 words per slot in that order.
 */
void initialize_xorwow_states(std::uint32_t seed,
                              std::uint32_t stream,
                              size_type num_slots,
                              std::uint32_t* out);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
