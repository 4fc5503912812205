//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/Json.cc
//---------------------------------------------------------------------------//
#include "Json.hh"

#include <cctype>
#include <cstdio>
#include <cstdlib>

#include "DeviceRuntime.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
class Parser
{
  public:
    explicit Parser(std::string const& t) : text_(t) {}

    JsonValue parse_document()
    {
        JsonValue v = this->parse_value();
        this->skip_ws();
        if (pos_ != text_.size())
            this->fail("trailing characters");
        return v;
    }

  private:
    std::string const& text_;
    std::size_t pos_{0};

    [[noreturn]] void fail(char const* what)
    {
        throw RuntimeError(std::string("JSON parse error: ") + what
                           + " at offset " + std::to_string(pos_));
    }
    void skip_ws()
    {
        while (pos_ < text_.size()
               && std::isspace(static_cast<unsigned char>(text_[pos_])))
            ++pos_;
    }
    char peek()
    {
        this->skip_ws();
        if (pos_ >= text_.size())
            this->fail("unexpected end");
        return text_[pos_];
    }
    void expect(char c)
    {
        if (this->peek() != c)
            this->fail("unexpected character");
        ++pos_;
    }

    JsonValue parse_value()
    {
        char c = this->peek();
        JsonValue v;
        if (c == '{')
        {
            v.type = JsonValue::Type::object;
            ++pos_;
            if (this->peek() == '}')
            {
                ++pos_;
                return v;
            }
            while (true)
            {
                std::string key = this->parse_string();
                this->expect(':');
                v.obj[key] = this->parse_value();
                char d = this->peek();
                ++pos_;
                if (d == '}')
                    break;
                if (d != ',')
                    this->fail("expected ',' or '}'");
            }
        }
        else if (c == '[')
        {
            v.type = JsonValue::Type::array;
            ++pos_;
            if (this->peek() == ']')
            {
                ++pos_;
                return v;
            }
            while (true)
            {
                v.arr.push_back(this->parse_value());
                char d = this->peek();
                ++pos_;
                if (d == ']')
                    break;
                if (d != ',')
                    this->fail("expected ',' or ']'");
            }
        }
        else if (c == '"')
        {
            v.type = JsonValue::Type::string;
            v.str = this->parse_string();
        }
        else if (text_.compare(pos_, 4, "true") == 0)
        {
            v.type = JsonValue::Type::boolean;
            v.b = true;
            pos_ += 4;
        }
        else if (text_.compare(pos_, 5, "false") == 0)
        {
            v.type = JsonValue::Type::boolean;
            pos_ += 5;
        }
        else if (text_.compare(pos_, 4, "null") == 0)
        {
            pos_ += 4;
        }
        else
        {
            char const* start = text_.c_str() + pos_;
            char* end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start)
                this->fail("invalid value");
            v.type = JsonValue::Type::number;
            pos_ += static_cast<std::size_t>(end - start);
        }
        return v;
    }

    std::string parse_string()
    {
        if (this->peek() != '"')
            this->fail("expected string");
        ++pos_;
        std::string result;
        while (pos_ < text_.size() && text_[pos_] != '"')
        {
            char c = text_[pos_++];
            if (c == '\\')
            {
                if (pos_ >= text_.size())
                    this->fail("bad escape");
                char e = text_[pos_++];
                switch (e)
                {
                    case 'n':
                        result += '\n';
                        break;
                    case 't':
                        result += '\t';
                        break;
                    case 'r':
                        result += '\r';
                        break;
                    case 'b':
                        result += '\b';
                        break;
                    case 'f':
                        result += '\f';
                        break;
                    case 'u':
                        if (pos_ + 4 > text_.size())
                            this->fail("bad unicode escape");
                        result += static_cast<char>(
                            std::strtol(text_.substr(pos_, 4).c_str(), nullptr, 16)
                            & 0x7f);
                        pos_ += 4;
                        break;
                    default:
                        result += e;
                }
            }
            else
            {
                result += c;
            }
        }
        if (pos_ >= text_.size())
            this->fail("unterminated string");
        ++pos_;
        return result;
    }
};
//---------------------------------------------------------------------------//
}  // namespace

JsonValue parse_json(std::string const& text)
{
    return Parser(text).parse_document();
}

JsonValue const& JsonValue::at(std::string const& key) const
{
    CELER_VALIDATE(type == Type::object, << "JSON: expected object for key '"
                                         << key << "'");
    auto iter = obj.find(key);
    CELER_VALIDATE(iter != obj.end(), << "JSON: missing key '" << key << "'");
    return iter->second;
}

JsonValue const& JsonValue::operator[](std::size_t i) const
{
    CELER_VALIDATE(type == Type::array && i < arr.size(),
                   << "JSON: bad array index " << i);
    return arr[i];
}

double JsonValue::as_double() const
{
    CELER_VALIDATE(type == Type::number, << "JSON: expected number");
    return num;
}
long long JsonValue::as_int() const
{
    CELER_VALIDATE(type == Type::number, << "JSON: expected integer");
    return static_cast<long long>(num);
}
bool JsonValue::as_bool() const
{
    CELER_VALIDATE(type == Type::boolean, << "JSON: expected boolean");
    return b;
}
std::string const& JsonValue::as_string() const
{
    CELER_VALIDATE(type == Type::string, << "JSON: expected string");
    return str;
}
double JsonValue::get(std::string const& key, double dflt) const
{
    return this->has(key) && !obj.at(key).is_null() ? obj.at(key).as_double()
                                                     : dflt;
}
long long JsonValue::get_int(std::string const& key, long long dflt) const
{
    return this->has(key) && !obj.at(key).is_null() ? obj.at(key).as_int()
                                                     : dflt;
}
bool JsonValue::get_bool(std::string const& key, bool dflt) const
{
    return this->has(key) && !obj.at(key).is_null() ? obj.at(key).as_bool()
                                                     : dflt;
}
std::string
JsonValue::get_str(std::string const& key, std::string const& dflt) const
{
    return this->has(key) && !obj.at(key).is_null() ? obj.at(key).as_string()
                                                     : dflt;
}
std::vector<double> JsonValue::as_doubles() const
{
    CELER_VALIDATE(type == Type::array, << "JSON: expected array");
    std::vector<double> result;
    for (auto const& v : arr)
        result.push_back(v.as_double());
    return result;
}
std::vector<long long> JsonValue::as_ints() const
{
    CELER_VALIDATE(type == Type::array, << "JSON: expected array");
    std::vector<long long> result;
    for (auto const& v : arr)
        result.push_back(v.as_int());
    return result;
}

//---------------------------------------------------------------------------//
namespace
{
void dump_to(JsonValue const& v, std::string& out)
{
    switch (v.type)
    {
        case JsonValue::Type::null:
            out += "null";
            break;
        case JsonValue::Type::boolean:
            out += v.b ? "true" : "false";
            break;
        case JsonValue::Type::number: {
            char buf[32];
            std::snprintf(buf, sizeof(buf), "%.17g", v.num);
            out += buf;
            break;
        }
        case JsonValue::Type::string:
            out += '"';
            for (char c : v.str)
            {
                if (c == '"' || c == '\\')
                    out += '\\';
                out += c;
            }
            out += '"';
            break;
        case JsonValue::Type::array:
            out += '[';
            for (std::size_t i = 0; i < v.arr.size(); ++i)
            {
                if (i)
                    out += ',';
                dump_to(v.arr[i], out);
            }
            out += ']';
            break;
        case JsonValue::Type::object: {
            out += '{';
            bool first = true;
            for (auto const& kv : v.obj)
            {
                if (!first)
                    out += ',';
                first = false;
                out += '"';
                out += kv.first;
                out += "\":";
                dump_to(kv.second, out);
            }
            out += '}';
            break;
        }
    }
}
}  // namespace

std::string dump_json(JsonValue const& v)
{
    std::string out;
    dump_to(v, out);
    return out;
}

JsonValue JsonValue::number(double x)
{
    JsonValue v;
    v.type = Type::number;
    v.num = x;
    return v;
}

JsonValue JsonValue::string(std::string s)
{
    JsonValue v;
    v.type = Type::string;
    v.str = std::move(s);
    return v;
}

JsonValue JsonValue::array(std::vector<double> const& xs)
{
    JsonValue v;
    v.type = Type::array;
    for (double x : xs)
        v.arr.push_back(number(x));
    return v;
}

JsonValue JsonValue::array()
{
    JsonValue v;
    v.type = Type::array;
    return v;
}

JsonValue JsonValue::object()
{
    JsonValue v;
    v.type = Type::object;
    return v;
}

JsonValue JsonValue::boolean(bool x)
{
    JsonValue v;
    v.type = Type::boolean;
    v.b = x;
    return v;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
