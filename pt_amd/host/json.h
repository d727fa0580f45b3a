// Minimal JSON document model and parser, enough for glTF 2.0 files.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace pth {

struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool has(const std::string& k) const { return kind == OBJ && obj.count(k) != 0; }
    const Json& operator[](const std::string& k) const;
    const Json& operator[](size_t i) const;
    size_t size() const { return kind == ARR ? arr.size() : (kind == OBJ ? obj.size() : 0); }
    double number(double fallback) const { return kind == NUM ? num : fallback; }
};

// Throws std::runtime_error on malformed input.
Json parse_json(const std::string& text);

}  // namespace pth
