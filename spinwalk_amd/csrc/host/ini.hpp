// ini.hpp -- INI reader with the semantics of the parser the reference reads
// its configs with (mINI 0.9.17, include/ini.h, built with MINI_CASE_SENSITIVE,
// CMakeLists.txt:36):
//   - a UTF-8 byte-order mark is skipped; lines split on '\n', '\r' and '\0' dropped;
//   - every line is trimmed of " \t\n\r\f\v";
//   - a line starting with ';' is a comment (there are no inline value comments:
//     "TR = 10e3 ; x" keeps "10e3 ; x", and the numeric conversion stops at the space);
//   - "[name]" (text after ';' on a section line ignored) opens a section;
//   - "key = value" splits at the first '=' not escaped as "\=", key and value
//     trimmed, "\=" in the key becomes '=';
//   - key/value lines before the first section are ignored; a repeated key keeps
//     the last value; names are case sensitive.
#pragma once
#include <map>
#include <string>

namespace swhost {

class Ini {
public:
    bool read(const std::string &path);
    bool has(const std::string &sec, const std::string &key) const;
    // the value, or "" when absent (mINI's operator[] yields an empty value)
    std::string get(const std::string &sec, const std::string &key) const;

private:
    std::map<std::string, std::map<std::string, std::string>> data_;
};

}  // namespace swhost
