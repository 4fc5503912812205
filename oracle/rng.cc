//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see rng.hh).
//---------------------------------------------------------------------------//
#include "rng.hh"

#include <random>

namespace orc
{
namespace
{
//---------------------------------------------------------------------------//
using Vec160 = std::array<std::uint32_t, 5>;

//! Bit b of a 160-bit vector
inline int bit(Vec160 const& v, int b)
{
    return (v[b / 32] >> (b % 32)) & 1u;
}

//! Apply the xorshift transform to a bare 160-bit vector
Vec160 apply_t(Vec160 v)
{
    XorwowState s;
    for (int i = 0; i < 5; ++i)
        s.x[i] = v[i];
    s.d = 0;
    xorwow_advance_raw(s);
    Vec160 r;
    for (int i = 0; i < 5; ++i)
        r[i] = s.x[i];
    return r;
}

//! y = M x where M is stored by rows
Vec160 matvec(Mat160 const& m, Vec160 const& x)
{
    Vec160 y{0, 0, 0, 0, 0};
    for (int r = 0; r < 160; ++r)
    {
        unsigned int acc = 0;
        for (int w = 0; w < 5; ++w)
            acc ^= __builtin_popcount(m.row[r][w] & x[w]) & 1u;
        y[r / 32] |= (acc & 1u) << (r % 32);
    }
    return y;
}

//! C = A B
Mat160 matmul(Mat160 const& a, Mat160 const& b)
{
    Mat160 c;
    for (int r = 0; r < 160; ++r)
    {
        Vec160 acc{0, 0, 0, 0, 0};
        for (int k = 0; k < 160; ++k)
        {
            if (bit(a.row[r], k))
            {
                for (int w = 0; w < 5; ++w)
                    acc[w] ^= b.row[k][w];
            }
        }
        c.row[r] = acc;
    }
    return c;
}

//! Transition matrix: column j is T applied to unit vector e_j
Mat160 transition()
{
    Mat160 m;
    for (auto& r : m.row)
        r = Vec160{0, 0, 0, 0, 0};
    for (int j = 0; j < 160; ++j)
    {
        Vec160 e{0, 0, 0, 0, 0};
        e[j / 32] = 1u << (j % 32);
        Vec160 col = apply_t(e);
        for (int r = 0; r < 160; ++r)
        {
            if (bit(col, r))
                m.row[r][j / 32] |= 1u << (j % 32);
        }
    }
    return m;
}

void apply_ladder(XorwowState& s, std::vector<Mat160> const& ladder,
                  unsigned long long count)
{
    int idx = 0;
    while (count > 0)
    {
        unsigned int digit = static_cast<unsigned int>(count & 3u);
        for (unsigned int r = 0; r < digit; ++r)
        {
            Vec160 v;
            for (int i = 0; i < 5; ++i)
                v[i] = s.x[i];
            v = matvec(ladder[idx], v);
            for (int i = 0; i < 5; ++i)
                s.x[i] = v[i];
        }
        ++idx;
        count >>= 2;
    }
}
//---------------------------------------------------------------------------//
}  // namespace

JumpMatrices const& JumpMatrices::instance()
{
    static JumpMatrices const jm = [] {
        JumpMatrices r;
        Mat160 m = transition();
        for (int i = 0; i < 32; ++i)
        {
            r.step.push_back(m);
            m = matmul(m, m);
            m = matmul(m, m);
        }
        Mat160 s = transition();
        for (int i = 0; i < 67; ++i)
            s = matmul(s, s);
        for (int i = 0; i < 32; ++i)
        {
            r.subsequence.push_back(s);
            s = matmul(s, s);
            s = matmul(s, s);
        }
        return r;
    }();
    return jm;
}

void xorwow_advance_raw(XorwowState& s)
{
    std::uint32_t t = s.x[0] ^ (s.x[0] >> 2);
    s.x[0] = s.x[1];
    s.x[1] = s.x[2];
    s.x[2] = s.x[3];
    s.x[3] = s.x[4];
    s.x[4] = (s.x[4] ^ (s.x[4] << 4)) ^ (t ^ (t << 1));
}

std::uint32_t xorwow_next(XorwowState& s)
{
    xorwow_advance_raw(s);
    s.d += 362437u;
    return s.d + s.x[4];
}

void xorwow_discard(XorwowState& s, unsigned long long count)
{
    apply_ladder(s, JumpMatrices::instance().step, count);
    s.d += static_cast<unsigned int>(count) * 362437u;
}

void xorwow_discard_subsequence(XorwowState& s, unsigned long long count)
{
    apply_ladder(s, JumpMatrices::instance().subsequence, count);
}

void xorwow_initialize(XorwowState& s,
                       std::uint32_t seed,
                       unsigned long long subsequence,
                       unsigned long long offset)
{
    // SplitMix64 (Steele, Lea & Flood 2014)
    std::uint64_t state = seed;
    auto next64 = [&state] {
        std::uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    for (int i = 0; i < 3; ++i)
    {
        std::uint64_t v = next64();
        if (i < 2)
        {
            s.x[2 * i] = static_cast<std::uint32_t>(v);
            s.x[2 * i + 1] = static_cast<std::uint32_t>(v >> 32);
        }
        else
        {
            s.x[4] = static_cast<std::uint32_t>(v);
            s.d = static_cast<std::uint32_t>(v >> 32);
        }
    }
    xorwow_discard_subsequence(s, subsequence);
    xorwow_discard(s, offset);
}

double canonical_double(XorwowState& s)
{
    unsigned long long upper = xorwow_next(s);
    unsigned long long lower = xorwow_next(s);
    return static_cast<double>((upper << 21) ^ lower) * 0x1p-53;
}

float canonical_float(XorwowState& s)
{
    return 0x1p-32f * static_cast<float>(xorwow_next(s));
}

std::vector<XorwowState>
seeded_states(std::uint32_t seed, std::uint32_t stream, std::size_t n)
{
    std::vector<unsigned int> seeds{seed};
    if (stream != 0)
        seeds.push_back(stream);
    std::seed_seq sseq(seeds.begin(), seeds.end());
    std::mt19937 mt(sseq);
    std::uniform_int_distribution<std::uint32_t> u32;
    std::vector<XorwowState> result(n);
    for (auto& st : result)
    {
        for (auto& w : st.x)
            w = u32(mt);
        st.d = u32(mt);
    }
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace orc
