// qk_conf.cpp -- .conf reader (grammar of parse.py:8-66, storage of
// var::read, src/low-level.cpp:376-444).
#include "qk_conf.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace qkd {

namespace {

// character classes of the grammar
bool is_word(unsigned char c) { return std::isalnum(c) || c == '_' || c >= 0x80; }
bool is_text(unsigned char c) { return is_word(c) || (c >= '&' && c <= '.') || c == '/'; }

struct Cursor {
    const std::string &s;
    size_t i;
    bool end() const { return i >= s.size(); }
    char peek() const { return end() ? '\0' : s[i]; }
};

bool digits(Cursor &c)
{
    const size_t b = c.i;
    while (!c.end() && std::isdigit((unsigned char)c.peek())) ++c.i;
    return c.i > b;
}

bool number(Cursor &c)
{
    const size_t b = c.i;
    if (c.peek() == '+' || c.peek() == '-') ++c.i;
    if (std::isdigit((unsigned char)c.peek())) {
        digits(c);
        if (c.peek() == '.') {
            ++c.i;
            digits(c);
        }
    } else if (c.peek() == '.') {
        ++c.i;
        if (!digits(c)) { c.i = b; return false; }
    } else {
        c.i = b;
        return false;
    }
    if (c.peek() == 'e' || c.peek() == 'E') {
        const size_t e = c.i;
        ++c.i;
        if (c.peek() == '+' || c.peek() == '-') ++c.i;
        if (!digits(c)) c.i = e;   // "1e" is the number 1 followed by text
    }
    return true;
}

bool array(Cursor &c)
{
    const size_t b = c.i;
    if (c.peek() != '[') return false;
    ++c.i;
    if (!number(c)) { c.i = b; return false; }
    while (c.peek() == ';') {
        ++c.i;
        if (!number(c)) { c.i = b; return false; }
    }
    if (c.peek() != ']') { c.i = b; return false; }
    ++c.i;
    return true;
}

bool unit(Cursor &c)
{
    const size_t b = c.i;
    while (!c.end() && is_word((unsigned char)c.peek())) ++c.i;
    if (c.i == b) return false;
    if (c.peek() == '^') {
        const size_t h = c.i;
        ++c.i;
        if (c.peek() == '{') {
            ++c.i;
            if (c.peek() == '+' || c.peek() == '-') ++c.i;
            if (!digits(c) || c.peek() != '}') { c.i = h; return true; }
            ++c.i;
        } else if (!digits(c)) {
            c.i = h;
        }
    }
    return true;
}

bool dimension(Cursor &c)
{
    if (c.peek() == '%') {
        ++c.i;
        return true;
    }
    if (!unit(c)) return false;
    for (;;) {
        const size_t b = c.i;
        if (c.peek() != '/' && c.peek() != '*') break;
        ++c.i;
        if (!unit(c)) { c.i = b; break; }
    }
    return true;
}

// whole-string matches
bool whole(const std::string &v, bool (*f)(Cursor &))
{
    Cursor c{v, 0};
    return f(c) && c.end();
}
bool numeric_value(Cursor &c) { return array(c) || number(c); }

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

void store_numeric(const std::string &tok, ConfVar &v)
{
    if (tok[0] == '[') {
        // entries in double (atof), value stays 0
        std::string body = tok.substr(1, tok.size() - 2);
        std::stringstream ss(body);
        std::string item;
        while (std::getline(ss, item, ';')) v.array.push_back(std::atof(item.c_str()));
    } else {
        v.value = (double)std::strtof(tok.c_str(), nullptr);   // stof: float precision
    }
}

}  // namespace

bool read_conf_text(const std::string &text, std::vector<ConfVar> &out, std::string &err)
{
    std::stringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        line = trim(line);
        if (line.empty()) continue;
        if (line.find("eval{") != std::string::npos || line.find("exec{") != std::string::npos) {
            err = "line " + std::to_string(lineno) + ": eval{} / exec{} run Python and are not supported";
            return false;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos || line.find('=', eq + 1) != std::string::npos) {
            err = "line " + std::to_string(lineno) + ": [" + line + "] is not key = value";
            return false;
        }
        ConfVar v;
        v.name = trim(line.substr(0, eq));
        const std::string val = trim(line.substr(eq + 1));
        if (v.name.empty() || v.name.find_first_of(" \t") != std::string::npos) {
            err = "line " + std::to_string(lineno) + ": bad key [" + v.name + "]";
            return false;
        }
        if (!val.empty() && whole(val, numeric_value)) {
            store_numeric(val, v);
            v.units = "#";
        } else if (!val.empty() && [&] {
                       for (unsigned char ch : val)
                           if (!is_text(ch)) return false;
                       return true;
                   }()) {
            v.units = val;
        } else {
            Cursor c{val, 0};
            bool ok = numeric_value(c);
            const size_t vend = c.i;
            size_t ws = 0;
            while (!c.end() && std::isspace((unsigned char)c.peek())) { ++c.i; ++ws; }
            const size_t dstart = c.i;
            ok = ok && ws > 0 && dimension(c) && c.end();
            if (!ok) {
                err = "line " + std::to_string(lineno) + ": syntax of value [" + val + "] of key [" + v.name +
                      "] is incorrect";
                return false;
            }
            store_numeric(val.substr(0, vend), v);
            v.units = val.substr(dstart);
        }
        out.push_back(v);
    }
    return true;
}

bool read_conf_file(const std::string &path, std::vector<ConfVar> &out, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "config file [" + path + "] not found";
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return read_conf_text(ss.str(), out, err);
}

size_t find_index(const std::vector<ConfVar> &vars, const std::string &name, size_t from)
{
    for (size_t n = from; n < vars.size(); ++n)
        if (vars[n].name == name) return n;
    return vars.size();
}

const ConfVar &find(const std::vector<ConfVar> &vars, const std::string &name, size_t from)
{
    static const ConfVar empty;
    const size_t n = find_index(vars, name, from);
    return n < vars.size() ? vars[n] : empty;
}

}  // namespace qkd
