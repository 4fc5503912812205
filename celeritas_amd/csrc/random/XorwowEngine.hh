//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file random/XorwowEngine.hh
//! Register-resident XORWOW generator over a structure-of-arrays state.
//!
//! Algorithm: Marsaglia (2003) "xorwow": a 160-bit xorshift combined with a
//! Weyl sequence (step 362437), as used by the reference engine
//! (celeritas/random/XorwowRngEngine.hh:191-236). Skip-ahead evaluates the
//! F2 jump polynomial g(z) = z^d mod p(z) with the Horner-like scheme of
//! Haramoto et al. (2008) (XorwowRngEngine.hh:246-314). Seeding runs
//! SplitMix64 (Steele, Lea, Flood 2014; XorwowRngEngine.hh:139-166,322-328).
//!
//! The per-slot state is six 32-bit words stored as six separate arrays so a
//! wave64 load of one word is a single 256-byte coalesced access. Kernels
//! load the state into registers once, draw, and store it back once.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Types.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Jump polynomials: index i skips 4^i draws (or 4^i * 2^67 for subsequences)
struct XorwowJumpTables
{
    std::uint32_t jump[32][5];
    std::uint32_t jump_subsequence[32][5];
};

//! Structure-of-arrays view of the per-slot generator state
struct RngStateRef
{
    std::uint32_t* xs[5];  //!< xorshift words x, y, z, w, v
    std::uint32_t* weyl;  //!< Weyl sequence word d
    size_type size;
};

//! Initializer following the reference's XorwowRngInitializer
struct XorwowInitializer
{
    std::uint32_t seed{0};
    ull_int subsequence{0};
    ull_int offset{0};
};

//---------------------------------------------------------------------------//
/*!
 * XORWOW engine with its state held in registers.
 */
class XorwowEngine
{
  public:
    std::uint32_t v[5];
    std::uint32_t d;
    //! Set by a sampling loop that ran out of its iteration budget (not part
    //! of the stored state)
    bool exhausted{false};

    //! Load from SoA state
    CELER_FORCEINLINE_FUNCTION static XorwowEngine
    load(RngStateRef const& s, size_type slot)
    {
        XorwowEngine e;
        for (int i = 0; i < 5; ++i)
            e.v[i] = s.xs[i][slot];
        e.d = s.weyl[slot];
        return e;
    }

    //! Store to SoA state
    CELER_FORCEINLINE_FUNCTION void
    store(RngStateRef const& s, size_type slot) const
    {
        for (int i = 0; i < 5; ++i)
            s.xs[i][slot] = v[i];
        s.weyl[slot] = d;
    }

    //! Apply the xorshift transform (does not touch the Weyl word)
    CELER_FORCEINLINE_FUNCTION void next()
    {
        std::uint32_t const t = v[0] ^ (v[0] >> 2u);
        v[0] = v[1];
        v[1] = v[2];
        v[2] = v[3];
        v[3] = v[4];
        v[4] = (v[4] ^ (v[4] << 4u)) ^ (t ^ (t << 1u));
    }

    //! Draw a 32-bit word
    CELER_FORCEINLINE_FUNCTION std::uint32_t operator()()
    {
        this->next();
        d += 362437u;
        return d + v[4];
    }

    //! Apply one jump polynomial: state <- g(T) state
    CELER_FUNCTION void apply_poly(std::uint32_t const* poly)
    {
        std::uint32_t acc[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < 5; ++i)
        {
            std::uint32_t const word = poly[i];
            for (int j = 0; j < 32; ++j)
            {
                // Branch-free conditional accumulate keeps the wave coherent
                std::uint32_t const mask = 0u - ((word >> j) & 1u);
                for (int k = 0; k < 5; ++k)
                    acc[k] ^= v[k] & mask;
                this->next();
            }
        }
        for (int k = 0; k < 5; ++k)
            v[k] = acc[k];
    }

    //! Skip ahead \c count steps of the given jump ladder (base-4 digits)
    CELER_FUNCTION void jump(ull_int count, std::uint32_t const (*table)[5])
    {
        int idx = 0;
        while (count > 0)
        {
            unsigned int const reps = static_cast<unsigned int>(count) & 3u;
            for (unsigned int r = 0; r < reps; ++r)
                this->apply_poly(table[idx]);
            ++idx;
            count >>= 2;
        }
    }

    //! Advance \c count draws (also advances the Weyl sequence)
    CELER_FUNCTION void discard(ull_int count, XorwowJumpTables const& t)
    {
        this->jump(count, t.jump);
        d += static_cast<unsigned int>(count) * 362437u;
    }

    //! Advance \c count subsequences of 2^67 draws (Weyl period divides it)
    CELER_FUNCTION void
    discard_subsequence(ull_int count, XorwowJumpTables const& t)
    {
        this->jump(count, t.jump_subsequence);
    }

    //! Seed with SplitMix64, then skip to subsequence and offset
    CELER_FUNCTION void
    initialize(XorwowInitializer const& init, XorwowJumpTables const& t)
    {
        std::uint64_t sm = init.seed;
        auto splitmix = [&sm]() -> std::uint64_t {
            std::uint64_t z = (sm += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        std::uint64_t w = splitmix();
        v[0] = static_cast<std::uint32_t>(w);
        v[1] = static_cast<std::uint32_t>(w >> 32);
        w = splitmix();
        v[2] = static_cast<std::uint32_t>(w);
        v[3] = static_cast<std::uint32_t>(w >> 32);
        w = splitmix();
        v[4] = static_cast<std::uint32_t>(w);
        d = static_cast<std::uint32_t>(w >> 32);
        this->discard_subsequence(init.subsequence, t);
        this->discard(init.offset, t);
    }
};

//---------------------------------------------------------------------------//
/*!
 * Uniform real on [0, 1) from 32-bit draws.
 *
 * Double: 53 bits from (upper << 21) ^ lower, times 2^-53
 * (random/detail/GenerateCanonical32.hh:60-84). Float: one draw times 2^-32
 * in single precision (the product may round up to 1 for the largest words,
 * exactly as the reference does).
 */
template<class T>
struct GenerateCanonical;

template<>
struct GenerateCanonical<double>
{
    template<class Engine>
    CELER_FORCEINLINE_FUNCTION double operator()(Engine& rng) const
    {
        std::uint32_t const upper = rng();
        std::uint32_t const lower = rng();
        constexpr double norm = 1.1102230246251565e-16;  // 2^-53
        return norm
               * static_cast<double>((static_cast<ull_int>(upper) << 21)
                                     ^ static_cast<ull_int>(lower));
    }
};

template<>
struct GenerateCanonical<float>
{
    template<class Engine>
    CELER_FORCEINLINE_FUNCTION float operator()(Engine& rng) const
    {
        constexpr float norm = 2.32830643654e-10f;  // 2^-32
        return norm * static_cast<float>(rng());
    }
};

template<class T, class Engine>
CELER_FORCEINLINE_FUNCTION T generate_canonical(Engine& rng)
{
    return GenerateCanonical<T>{}(rng);
}

//---------------------------------------------------------------------------//
/*!
 * Engine that replays precomputed uniform doubles.
 *
 * Used by the interactor verification kernels so that a device interactor
 * consumes exactly the stream given to a CPU implementation.
 */
struct ReplayEngine
{
    double const* u;
    std::uint32_t i;
    bool exhausted{false};
    CELER_FORCEINLINE_FUNCTION double next() { return u[i++]; }
};

template<>
CELER_FORCEINLINE_FUNCTION double generate_canonical<double, ReplayEngine>(ReplayEngine& rng)
{
    return rng.next();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
