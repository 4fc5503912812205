//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/json.hh
//! Small JSON reader for problem descriptions given to the oracle.
//---------------------------------------------------------------------------//
#pragma once

#include <cstdlib>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc
{
//---------------------------------------------------------------------------//
struct Json
{
    enum Kind { null_k, bool_k, num_k, str_k, arr_k, obj_k } kind = null_k;
    double n = 0;
    bool b = false;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    Json const* find(std::string const& k) const
    {
        for (auto const& kv : o)
            if (kv.first == k)
                return &kv.second;
        return nullptr;
    }
    Json const& operator()(std::string const& k) const
    {
        auto p = find(k);
        if (!p)
            throw std::runtime_error("oracle json: missing key " + k);
        return *p;
    }
    double num(std::string const& k, double d) const
    {
        auto p = find(k);
        return (p && p->kind == num_k) ? p->n : d;
    }
    std::string str(std::string const& k, std::string const& d) const
    {
        auto p = find(k);
        return (p && p->kind == str_k) ? p->s : d;
    }
    bool flag(std::string const& k, bool d) const
    {
        auto p = find(k);
        return (p && p->kind == bool_k) ? p->b : d;
    }
    std::vector<double> nums() const
    {
        std::vector<double> r;
        for (auto const& v : a)
            r.push_back(v.n);
        return r;
    }
};

//! Parse JSON text (throws std::runtime_error)
inline Json parse(std::string const& t)
{
    std::size_t i = 0;
    auto ws = [&] {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\t' || t[i] == '\r'))
            ++i;
    };
    auto err = [&](char const* m) -> Json {
        throw std::runtime_error(std::string("oracle json: ") + m + " @"
                                 + std::to_string(i));
    };
    std::function<Json()> value;
    auto string_lit = [&]() {
        std::string r;
        ++i;  // opening quote
        while (i < t.size() && t[i] != '"')
        {
            if (t[i] == '\\' && i + 1 < t.size())
            {
                char e = t[i + 1];
                r += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                i += 2;
                if (e == 'u')
                    i += 4;
                continue;
            }
            r += t[i++];
        }
        ++i;
        return r;
    };
    value = [&]() -> Json {
        ws();
        Json j;
        if (i >= t.size())
            return err("eof");
        char c = t[i];
        if (c == '{')
        {
            j.kind = Json::obj_k;
            ++i;
            ws();
            if (t[i] == '}')
            {
                ++i;
                return j;
            }
            for (;;)
            {
                ws();
                std::string k = string_lit();
                ws();
                ++i;  // colon
                j.o.emplace_back(k, value());
                ws();
                if (t[i++] == '}')
                    break;
            }
        }
        else if (c == '[')
        {
            j.kind = Json::arr_k;
            ++i;
            ws();
            if (t[i] == ']')
            {
                ++i;
                return j;
            }
            for (;;)
            {
                j.a.push_back(value());
                ws();
                if (t[i++] == ']')
                    break;
            }
        }
        else if (c == '"')
        {
            j.kind = Json::str_k;
            j.s = string_lit();
        }
        else if (c == 't' || c == 'f')
        {
            j.kind = Json::bool_k;
            j.b = (c == 't');
            i += j.b ? 4 : 5;
        }
        else if (c == 'n')
        {
            i += 4;
        }
        else
        {
            char* end = nullptr;
            j.kind = Json::num_k;
            j.n = std::strtod(t.c_str() + i, &end);
            i = static_cast<std::size_t>(end - t.c_str());
        }
        return j;
    };
    return value();
}

//---------------------------------------------------------------------------//
}  // namespace orc
