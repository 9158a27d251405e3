#include "ini.hpp"

#include <fstream>
#include <sstream>
#include <vector>

namespace swhost {
namespace {

const char *kWs = " \t\n\r\f\v";

void trim(std::string &s) {
    s.erase(s.find_last_not_of(kWs) + 1);
    s.erase(0, s.find_first_not_of(kWs));
}

void replace_all(std::string &s, const std::string &a, const std::string &b) {
    if (a.empty()) return;
    for (size_t pos = 0; (pos = s.find(a, pos)) != std::string::npos; pos += b.size()) s.replace(pos, a.size(), b);
}

}  // namespace

bool Ini::read(const std::string &path) {
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string txt = ss.str();
    if (txt.size() >= 3 && (unsigned char)txt[0] == 0xEF && (unsigned char)txt[1] == 0xBB &&
        (unsigned char)txt[2] == 0xBF)
        txt.erase(0, 3);
    std::vector<std::string> lines;
    std::string buf;
    for (char c : txt) {
        if (c == '\n') { lines.push_back(buf); buf.clear(); continue; }
        if (c != '\0' && c != '\r') buf += c;
    }
    lines.push_back(buf);
    std::string section;
    bool in_section = false;
    for (std::string line : lines) {
        trim(line);
        if (line.empty() || line[0] == ';') continue;
        if (line[0] == '[') {
            std::string l = line;
            const size_t sc = l.find_first_of(';');
            if (sc != std::string::npos) l = l.substr(0, sc);
            const size_t cb = l.find_last_of(']');
            if (cb != std::string::npos) {
                section = l.substr(1, cb - 1);
                trim(section);
                data_[section];
                in_section = true;
                continue;
            }
        }
        std::string norm = line;
        replace_all(norm, "\\=", "  ");
        const size_t eq = norm.find_first_of('=');
        if (eq == std::string::npos || !in_section) continue;
        std::string key = line.substr(0, eq), value = line.substr(eq + 1);
        trim(key);
        replace_all(key, "\\=", "=");
        trim(value);
        data_[section][key] = value;
    }
    return true;
}

bool Ini::has(const std::string &sec, const std::string &key) const {
    auto s = data_.find(sec);
    return s != data_.end() && s->second.count(key) > 0;
}

std::string Ini::get(const std::string &sec, const std::string &key) const {
    auto s = data_.find(sec);
    if (s == data_.end()) return "";
    auto k = s->second.find(key);
    return k == s->second.end() ? "" : k->second;
}

}  // namespace swhost
