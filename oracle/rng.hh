//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/rng.hh
//! XORWOW restated from celeritas/random/XorwowRngEngine.hh:54-328,
//! celeritas/random/XorwowRngData.cc:28-60 (state seeding) and
//! celeritas/random/detail/GenerateCanonical32.hh:60-84.
//!
//! Skip-ahead here uses powers of the 160x160 F2 transition matrix (square
//! and multiply), a different route to T^d than the product's jump
//! polynomials; both must agree bit for bit.
//!
//! Pinned by: XorwowRngEngine.test.cc:138-187 (seeded state words),
//! :36-99 (canonical bit patterns), :208-260 (jump == repeated next),
//! RngReseed.test.cc:33-90 (reseed equivalence and golden draws).
//---------------------------------------------------------------------------//
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace orc
{
//---------------------------------------------------------------------------//
struct XorwowState
{
    std::uint32_t x[5];
    std::uint32_t d;
};

//! 160x160 matrix over F2, rows as 5 words
struct Mat160
{
    std::array<std::array<std::uint32_t, 5>, 160> row;
};

//! Matrices T^(4^i) and T^(2^67 * 4^i) for i < 32
struct JumpMatrices
{
    std::vector<Mat160> step;
    std::vector<Mat160> subsequence;
    static JumpMatrices const& instance();
};

std::uint32_t xorwow_next(XorwowState& s);
void xorwow_advance_raw(XorwowState& s);  // transform without Weyl
void xorwow_discard(XorwowState& s, unsigned long long count);
void xorwow_discard_subsequence(XorwowState& s, unsigned long long count);
void xorwow_initialize(XorwowState& s,
                       std::uint32_t seed,
                       unsigned long long subsequence,
                       unsigned long long offset);
double canonical_double(XorwowState& s);
float canonical_float(XorwowState& s);

//! Seeded states (std::seed_seq + std::mt19937), words x0..x4, d per slot
std::vector<XorwowState>
seeded_states(std::uint32_t seed, std::uint32_t stream, std::size_t n);

//---------------------------------------------------------------------------//
}  // namespace orc
