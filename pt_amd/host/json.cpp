#include "json.h"

#include <cstdlib>
#include <stdexcept>

namespace pth {

static const Json kNull;

const Json& Json::operator[](const std::string& k) const
{
    if (kind != OBJ)
        return kNull;
    auto it = obj.find(k);
    return it == obj.end() ? kNull : it->second;
}

const Json& Json::operator[](size_t i) const
{
    if (kind != ARR || i >= arr.size())
        return kNull;
    return arr[i];
}

namespace {

struct Parser {
    const std::string& s;
    size_t p = 0;

    explicit Parser(const std::string& text) : s(text) {}

    [[noreturn]] void fail(const char* what)
    {
        throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset " + std::to_string(p));
    }

    void ws()
    {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\r' || s[p] == '\t'))
            p++;
    }

    bool lit(const char* w)
    {
        size_t n = 0;
        while (w[n])
            n++;
        if (s.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }

    static void put_utf8(std::string& out, unsigned cp)
    {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }

    unsigned hex4()
    {
        if (p + 4 > s.size())
            fail("short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; i++) {
            char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= unsigned(c - '0');
            else if (c >= 'a' && c <= 'f') v |= unsigned(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= unsigned(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }

    std::string string_body()
    {
        std::string out;
        p++;  // opening quote
        while (true) {
            if (p >= s.size())
                fail("unterminated string");
            char c = s[p++];
            if (c == '"')
                break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size())
                fail("unterminated escape");
            char e = s[p++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && p + 1 < s.size() && s[p] == '\\' && s[p + 1] == 'u') {
                    p += 2;
                    unsigned lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }

    Json value()
    {
        ws();
        if (p >= s.size())
            fail("unexpected end");
        Json j;
        char c = s[p];
        if (c == '{') {
            j.kind = Json::OBJ;
            p++;
            ws();
            if (p < s.size() && s[p] == '}') {
                p++;
                return j;
            }
            while (true) {
                ws();
                if (p >= s.size() || s[p] != '"')
                    fail("expected key");
                std::string k = string_body();
                ws();
                if (p >= s.size() || s[p] != ':')
                    fail("expected ':'");
                p++;
                j.obj[k] = value();
                ws();
                if (p < s.size() && s[p] == ',') {
                    p++;
                    continue;
                }
                if (p < s.size() && s[p] == '}') {
                    p++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            j.kind = Json::ARR;
            p++;
            ws();
            if (p < s.size() && s[p] == ']') {
                p++;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (p < s.size() && s[p] == ',') {
                    p++;
                    continue;
                }
                if (p < s.size() && s[p] == ']') {
                    p++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            j.kind = Json::STR;
            j.str = string_body();
        } else if (lit("true")) {
            j.kind = Json::BOOL;
            j.b = true;
        } else if (lit("false")) {
            j.kind = Json::BOOL;
        } else if (lit("null")) {
            j.kind = Json::NUL;
        } else {
            const char* begin = s.c_str() + p;
            char* end = nullptr;
            j.num = std::strtod(begin, &end);
            if (end == begin)
                fail("bad value");
            j.kind = Json::NUM;
            p += size_t(end - begin);
        }
        return j;
    }
};

}  // namespace

Json parse_json(const std::string& text)
{
    Parser ps(text);
    Json j = ps.value();
    ps.ws();
    if (ps.p != text.size())
        ps.fail("trailing characters");
    return j;
}

}  // namespace pth
