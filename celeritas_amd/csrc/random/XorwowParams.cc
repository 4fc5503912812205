//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file random/XorwowParams.cc
//---------------------------------------------------------------------------//
#include "XorwowParams.hh"

#include <array>
#include <random>
#include <stdexcept>

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
// F2 polynomials of degree < 321 stored as little-endian 64-bit limbs
constexpr int kDeg = 160;
constexpr int kLimbs = 6;  // 384 bits: room for products before reduction
using Poly = std::array<std::uint64_t, kLimbs>;

bool get_bit(Poly const& p, int i)
{
    return (p[i / 64] >> (i % 64)) & 1u;
}
void flip_bit(Poly* p, int i)
{
    (*p)[i / 64] ^= (std::uint64_t(1) << (i % 64));
}
int degree(Poly const& p)
{
    for (int i = kLimbs * 64 - 1; i >= 0; --i)
    {
        if (get_bit(p, i))
            return i;
    }
    return -1;
}

//! Reduce modulo the monic degree-160 polynomial m
Poly reduce(Poly a, Poly const& m)
{
    for (int i = kLimbs * 64 - 1; i >= kDeg; --i)
    {
        if (get_bit(a, i))
        {
            int const shift = i - kDeg;
            for (int j = 0; j <= kDeg; ++j)
            {
                if (get_bit(m, j))
                    flip_bit(&a, j + shift);
            }
        }
    }
    return a;
}

//! Multiply modulo m (operands already reduced)
Poly mulmod(Poly const& a, Poly const& b, Poly const& m)
{
    Poly r{};
    for (int i = 0; i < kDeg; ++i)
    {
        if (!get_bit(a, i))
            continue;
        for (int j = 0; j < kDeg; ++j)
        {
            if (get_bit(b, j))
                flip_bit(&r, i + j);
        }
    }
    return reduce(r, m);
}

//! Pack the low 160 bits into five 32-bit words (low word first)
void pack(Poly const& p, std::uint32_t* out)
{
    for (int w = 0; w < 5; ++w)
    {
        out[w] = static_cast<std::uint32_t>(p[(32 * w) / 64]
                                            >> ((32 * w) % 64));
    }
}

//---------------------------------------------------------------------------//
/*!
 * Characteristic polynomial of the xorshift transform via Berlekamp-Massey.
 *
 * The output bit sequence of any nonzero state satisfies the minimal
 * polynomial of the transform; for the full-period xorwow transform it is
 * the degree-160 characteristic polynomial.
 */
Poly find_char_poly()
{
    XorwowEngine e;
    e.v[0] = 123456789u;
    e.v[1] = 362436069u;
    e.v[2] = 521288629u;
    e.v[3] = 88675123u;
    e.v[4] = 5783321u;
    e.d = 0;
    constexpr int n = 2 * kDeg + 16;
    std::vector<int> s(n);
    for (int i = 0; i < n; ++i)
    {
        s[i] = e.v[0] & 1u;
        e.next();
    }
    // Berlekamp-Massey: connection polynomial C(x) = 1 + c1 x + ... + cL x^L
    std::vector<int> c(n + 1, 0), b(n + 1, 0), t;
    c[0] = b[0] = 1;
    int L = 0, m = 1;
    for (int i = 0; i < n; ++i)
    {
        int disc = s[i];
        for (int j = 1; j <= L; ++j)
            disc ^= c[j] & s[i - j];
        if (disc == 0)
        {
            ++m;
        }
        else if (2 * L <= i)
        {
            t = c;
            for (int j = 0; j + m <= n; ++j)
                c[j + m] ^= b[j];
            L = i + 1 - L;
            b = t;
            m = 1;
        }
        else
        {
            for (int j = 0; j + m <= n; ++j)
                c[j + m] ^= b[j];
            ++m;
        }
    }
    if (L != kDeg)
        throw std::runtime_error("xorwow characteristic polynomial degree");
    // p(z) = z^L C(1/z): reverse the connection polynomial
    Poly p{};
    for (int j = 0; j <= L; ++j)
    {
        if (c[j])
            flip_bit(&p, L - j);
    }
    return p;
}
//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
XorwowParams::XorwowParams(std::uint32_t seed) : seed_(seed)
{
    Poly const p = find_char_poly();
    charpoly_.resize(6);
    for (int w = 0; w < 6; ++w)
        charpoly_[w] = static_cast<std::uint32_t>(p[(32 * w) / 64]
                                                  >> ((32 * w) % 64));

    // Jump by 4^i steps: start from z, square twice per ladder rung
    Poly g{};
    flip_bit(&g, 1);
    for (int i = 0; i < 32; ++i)
    {
        pack(g, tables_.jump[i]);
        g = mulmod(g, g, p);
        g = mulmod(g, g, p);
    }
    // Jump by 2^67 * 4^i steps: z^(2^67) by 67 squarings
    Poly h{};
    flip_bit(&h, 1);
    for (int i = 0; i < 67; ++i)
        h = mulmod(h, h, p);
    for (int i = 0; i < 32; ++i)
    {
        pack(h, tables_.jump_subsequence[i]);
        h = mulmod(h, h, p);
        h = mulmod(h, h, p);
    }
}

//---------------------------------------------------------------------------//
void initialize_xorwow_states(std::uint32_t seed,
                              std::uint32_t stream,
                              size_type num_slots,
                              std::uint32_t* out)
{
    std::vector<std::seed_seq::result_type> host_seeds{seed};
    if (stream != 0)
        host_seeds.push_back(stream);
    std::seed_seq seq(host_seeds.begin(), host_seeds.end());
    std::mt19937 rng(seq);
    std::uniform_int_distribution<std::uint32_t> sample;
    for (size_type i = 0; i < num_slots * 6; ++i)
        out[i] = sample(rng);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
