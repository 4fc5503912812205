//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file io/G4EmlowReaders.cc
//---------------------------------------------------------------------------//
#include "G4EmlowReaders.hh"

#include <cstdlib>
#include <fstream>
#include <sstream>

#include "corecel/DeviceRuntime.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
std::string data_dir(char const* subdir, char const* what)
{
    char const* env = std::getenv("G4LEDATA");
    CELER_VALIDATE(env && *env, << "environment variable G4LEDATA is not defined (needed "
                                   "to locate " << what << " data)");
    return std::string(env) + subdir;
}

std::string strip_slash(std::string p)
{
    CELER_VALIDATE(!p.empty(), << "empty G4EMLOW data directory");
    if (p.size() > 1 && p.back() == '/')
        p.pop_back();
    return p;
}

void check_z(int z)
{
    CELER_VALIDATE(z >= 1 && z <= 100, << "G4EMLOW data covers 1 <= Z <= 100 (got " << z << ")");
}

//! Whitespace-separated numbers of one file, with its name for errors
class Tokens
{
  public:
    Tokens(std::string const& file, char const* what) : file_(file)
    {
        std::ifstream in(file);
        CELER_VALIDATE(in, << "failed to open '" << file << "' (should contain " << what
                           << ")");
        std::ostringstream ss;
        ss << in.rdbuf();
        text_ = ss.str();
        is_.str(text_);
    }

    bool empty() const { return text_.find_first_not_of(" \t\r\n") == std::string::npos; }

    double real()
    {
        double v = 0;
        is_ >> v;
        CELER_VALIDATE(is_, << "truncated or malformed data in '" << file_ << "'");
        return v;
    }

    long long integer()
    {
        double const v = real();
        CELER_VALIDATE(v >= 0 && v == static_cast<double>(static_cast<long long>(v)),
                       << "expected a count in '" << file_ << "' (got " << v << ")");
        return static_cast<long long>(v);
    }

  private:
    std::string file_;
    std::string text_;
    std::istringstream is_;
};

//! "emin emax n n" header then n (x, y) pairs: a free physics vector
JsonValue read_table(Tokens& t, std::size_t n)
{
    std::vector<double> x(n), y(n);
    for (std::size_t i = 0; i < n; ++i)
    {
        x[i] = t.real();
        y[i] = t.real();
    }
    JsonValue r = JsonValue::object();
    r.obj["x"] = JsonValue::array(x);
    r.obj["y"] = JsonValue::array(y);
    return r;
}

constexpr int num_fit_params = 6;

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
SeltzerBergerReader::SeltzerBergerReader()
    : path_(data_dir("/brem_SB", "Seltzer-Berger"))
{
}

SeltzerBergerReader::SeltzerBergerReader(std::string path)
    : path_(strip_slash(std::move(path)))
{
}

SBTable SeltzerBergerReader::operator()(int z) const
{
    check_z(z);
    Tokens t(path_ + "/br" + std::to_string(z), "SB cross section data");
    SBTable r;
    t.integer();  // unused leading field
    auto const ny = static_cast<std::size_t>(t.integer());
    auto const nx = static_cast<std::size_t>(t.integer());
    CELER_VALIDATE(nx >= 2 && ny >= 2, << "SB table of Z=" << z << " has " << nx << " x "
                                       << ny << " points");
    r.y.resize(ny);
    r.x.resize(nx);
    for (auto& v : r.y)
        v = t.real();
    for (auto& v : r.x)
        v = t.real();
    r.values.resize(nx * ny);
    for (auto& v : r.values)
        v = t.real();
    return r;
}

//---------------------------------------------------------------------------//
LivermorePEReader::LivermorePEReader()
    : path_(data_dir("/livermore/phot_epics2014", "Livermore"))
{
}

LivermorePEReader::LivermorePEReader(std::string path)
    : path_(strip_slash(std::move(path)))
{
}

JsonValue LivermorePEReader::operator()(int z) const
{
    check_z(z);
    std::string const zs = std::to_string(z) + ".dat";
    JsonValue r = JsonValue::object();

    // Total cross section above the K edge, below the fit threshold
    {
        Tokens t(path_ + "/pe-cs-" + zs, "cross section data");
        t.real();
        t.real();
        t.integer();
        r.obj["xs_hi"] = read_table(t, static_cast<std::size_t>(t.integer()));
    }
    // Total below the K edge: absent for one-shell elements
    {
        Tokens t(path_ + "/pe-le-cs-" + zs, "cross section data");
        if (!t.empty())
        {
            t.real();
            t.real();
            t.integer();
            r.obj["xs_lo"] = read_table(t, static_cast<std::size_t>(t.integer()));
        }
    }
    // Per-shell binding energies and fit parameters, low then high interval
    JsonValue shells = JsonValue::array();
    std::size_t num_shells = 0;
    {
        Tokens t(path_ + "/pe-low-" + zs, "subshell fit parameters");
        t.integer();
        num_shells = static_cast<std::size_t>(t.integer());
        r.obj["thresh_lo"] = JsonValue::number(t.real());
        for (std::size_t s = 0; s < num_shells; ++s)
        {
            JsonValue sh = JsonValue::object();
            sh.obj["binding_energy"] = JsonValue::number(t.real());
            std::vector<double> p(num_fit_params);
            for (auto& v : p)
                v = t.real();
            sh.obj["param_lo"] = JsonValue::array(p);
            shells.arr.push_back(std::move(sh));
        }
    }
    {
        Tokens t(path_ + "/pe-high-" + zs, "subshell fit parameters");
        t.integer();
        auto const n = static_cast<std::size_t>(t.integer());
        CELER_VALIDATE(n == num_shells, << "pe-high-" << zs << " lists " << n
                                        << " shells, pe-low " << num_shells);
        r.obj["thresh_hi"] = JsonValue::number(t.real());
        for (auto& sh : shells.arr)
        {
            double const binding = t.real();
            CELER_VALIDATE(binding == sh.obj["binding_energy"].num,
                           << "binding energies differ between pe-low and pe-high of Z="
                           << z);
            std::vector<double> p(num_fit_params);
            for (auto& v : p)
                v = t.real();
            sh.obj["param_hi"] = JsonValue::array(p);
        }
    }
    // Tabulated subshell cross sections: "emin emax n shell_id" then n pairs
    {
        Tokens t(path_ + "/pe-ss-cs-" + zs, "subshell cross sections");
        for (auto& sh : shells.arr)
        {
            t.real();
            t.real();
            auto const n = static_cast<std::size_t>(t.integer());
            t.integer();
            JsonValue tab = read_table(t, n);
            sh.obj["x"] = tab.obj["x"];
            sh.obj["y"] = tab.obj["y"];
        }
    }
    r.obj["shells"] = std::move(shells);
    return r;
}

//---------------------------------------------------------------------------//
JsonValue to_json(SBTable const& t)
{
    JsonValue r = JsonValue::object();
    r.obj["x"] = JsonValue::array(t.x);
    r.obj["y"] = JsonValue::array(t.y);
    r.obj["values"] = JsonValue::array(t.values);
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
