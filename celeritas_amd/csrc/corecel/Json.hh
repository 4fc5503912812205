//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/Json.hh
//! Minimal JSON document model and recursive-descent parser (RFC 8259
//! subset: no \u escapes beyond ASCII) for problem and geometry input.
//---------------------------------------------------------------------------//
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
class JsonValue
{
  public:
    enum class Type
    {
        null,
        boolean,
        number,
        string,
        array,
        object
    };

    Type type{Type::null};
    bool b{false};
    double num{0};
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool is_null() const { return type == Type::null; }
    bool has(std::string const& key) const
    {
        return type == Type::object && obj.count(key);
    }
    JsonValue const& at(std::string const& key) const;
    JsonValue const& operator[](std::size_t i) const;
    std::size_t size() const
    {
        return type == Type::array ? arr.size() : obj.size();
    }

    double as_double() const;
    long long as_int() const;
    bool as_bool() const;
    std::string const& as_string() const;

    //! Get with a default if absent
    double get(std::string const& key, double dflt) const;
    long long get_int(std::string const& key, long long dflt) const;
    bool get_bool(std::string const& key, bool dflt) const;
    std::string get_str(std::string const& key, std::string const& dflt) const;

    std::vector<double> as_doubles() const;
    std::vector<long long> as_ints() const;

    //!@{
    //! Construction helpers
    static JsonValue number(double x);
    static JsonValue string(std::string s);
    static JsonValue boolean(bool x);
    static JsonValue array(std::vector<double> const& xs);
    static JsonValue array();
    static JsonValue object();
    //!@}
};

//! Parse a JSON document; throws RuntimeError with a position on error
JsonValue parse_json(std::string const& text);

//! Serialize a JSON value (numbers with 17 significant digits)
std::string dump_json(JsonValue const& v);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
