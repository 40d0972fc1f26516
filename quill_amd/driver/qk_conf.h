// qk_conf.h -- reader of QUILL .conf files for the quillk-run driver.
//
// The reference pipes each config through parse.py (parse.py:8-66), which
// emits key / value / units records, and reads them back with var::read
// (src/low-level.cpp:376-444).  This reader accepts the same grammar
// directly:
//
//   line      := key '=' value            ('#' starts a comment)
//   value     := number | array | text | (number | array) ws+ dimension
//   number    := [+-]? (digits ('.' digits?)? | '.' digits) ([eE] [+-]? digits)?
//   array     := '[' number (';' number)* ']'
//   text      := one or more of word characters and & ' ( ) * + , - . /
//   dimension := '%' | unit (('/' | '*') unit)*,  unit := word ('^' ('{' [+-]? digits '}' | digits))?
//
// and stores them as the reference does: a plain number gets units "#"
// (parse.py prints '#' in the units slot), a text value gets value 0 and
// the text as units, numbers are rounded through float as var::read's stof
// does, array entries are read in double as its atof does.  Lines with the
// Python hooks eval{...} and exec{...} are rejected: they would run code.
#pragma once

#include <string>
#include <vector>

namespace qkd {

struct ConfVar {
    std::string name;
    double value = 0;
    std::string units;
    std::vector<double> array;
};

// Parse a config file.  Returns false and sets err (with the line number)
// on the first line that does not follow the grammar.
bool read_conf_file(const std::string &path, std::vector<ConfVar> &out, std::string &err);
bool read_conf_text(const std::string &text, std::vector<ConfVar> &out, std::string &err);

// The first variable named `name` at or after index `from`, or the empty
// variable (value 0, units "") when there is none, as find()
// (src/low-level.cpp:446-461) returns.
const ConfVar &find(const std::vector<ConfVar> &vars, const std::string &name, size_t from = 0);
// index of that variable, or vars.size() when absent
size_t find_index(const std::vector<ConfVar> &vars, const std::string &name, size_t from = 0);

}  // namespace qkd
