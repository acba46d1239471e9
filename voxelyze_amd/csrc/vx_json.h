/*
 * vx_json.h -- a small JSON reader and writer for the .vxl.json model format
 * (CVoxelyze::loadJSON/saveJSON, src/Voxelyze.cpp:61-241; CVX_Material::readJSON/writeJSON,
 * src/VX_Material.cpp:75-163).  Host code only.
 *
 * The reader takes RFC 8259 JSON (numbers, strings with the standard escapes, true/false/null,
 * arrays, objects) and also accepts a document whose outer braces or brackets are left open at
 * the end of the text: the reference's writer closes the top-level object only when the model
 * has externals (src/Voxelyze.cpp:237), so its files without externals end early.
 */
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace vx {

struct JValue {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    bool integral = false; /* written without a fraction or exponent */
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;

    const JValue* get(const char* key) const {
        if (type != Object) return nullptr;
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool isNum() const { return type == Number; }
    bool isArr(size_t n = (size_t)-1) const { return type == Array && (n == (size_t)-1 || arr.size() == n); }
};

class JReader {
public:
    explicit JReader(const std::string& t) : s(t) {}
    bool parse(JValue* out, std::string* err) {
        pos = 0;
        if (!value(out, 0)) {
            if (err) *err = "JSON parse error at offset " + std::to_string(pos) + ": " + why;
            return false;
        }
        ws();
        if (pos != s.size()) {
            if (err) *err = "unexpected text after the JSON value at offset " + std::to_string(pos);
            return false;
        }
        return true;
    }

private:
    const std::string& s;
    size_t pos = 0;
    std::string why;

    void ws() {
        while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r')) pos++;
    }
    bool fail(const char* w) {
        why = w;
        return false;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) n++;
        if (s.compare(pos, n, w) != 0) return false;
        pos += n;
        return true;
    }
    bool value(JValue* v, int depth) {
        if (depth > 256) return fail("nesting too deep");
        ws();
        if (pos >= s.size()) return fail("unexpected end of text");
        const char c = s[pos];
        if (c == '{') return object(v, depth);
        if (c == '[') return array(v, depth);
        if (c == '"') {
            v->type = JValue::String;
            return string(&v->str);
        }
        if (lit("true")) { v->type = JValue::Bool; v->b = true; return true; }
        if (lit("false")) { v->type = JValue::Bool; v->b = false; return true; }
        if (lit("null")) { v->type = JValue::Null; return true; }
        return number(v);
    }
    bool number(JValue* v) {
        const size_t start = pos;
        if (pos < s.size() && s[pos] == '-') pos++;
        bool digits = false, integral = true;
        while (pos < s.size() && s[pos] >= '0' && s[pos] <= '9') { pos++; digits = true; }
        if (pos < s.size() && s[pos] == '.') {
            integral = false;
            pos++;
            while (pos < s.size() && s[pos] >= '0' && s[pos] <= '9') { pos++; digits = true; }
        }
        if (!digits) return fail("invalid value");
        if (pos < s.size() && (s[pos] == 'e' || s[pos] == 'E')) {
            integral = false;
            pos++;
            if (pos < s.size() && (s[pos] == '+' || s[pos] == '-')) pos++;
            if (pos >= s.size() || s[pos] < '0' || s[pos] > '9') return fail("invalid exponent");
            while (pos < s.size() && s[pos] >= '0' && s[pos] <= '9') pos++;
        }
        v->type = JValue::Number;
        v->integral = integral;
        v->num = std::strtod(s.substr(start, pos - start).c_str(), nullptr);
        return true;
    }
    static void utf8(std::string* out, unsigned cp) {
        if (cp < 0x80) {
            out->push_back((char)cp);
        } else if (cp < 0x800) {
            out->push_back((char)(0xC0 | (cp >> 6)));
            out->push_back((char)(0x80 | (cp & 0x3F)));
        } else {
            out->push_back((char)(0xE0 | (cp >> 12)));
            out->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
            out->push_back((char)(0x80 | (cp & 0x3F)));
        }
    }
    bool string(std::string* out) {
        pos++; /* opening quote */
        while (pos < s.size()) {
            const char c = s[pos++];
            if (c == '"') return true;
            if (c != '\\') {
                out->push_back(c);
                continue;
            }
            if (pos >= s.size()) break;
            const char e = s[pos++];
            switch (e) {
                case '"': out->push_back('"'); break;
                case '\\': out->push_back('\\'); break;
                case '/': out->push_back('/'); break;
                case 'b': out->push_back('\b'); break;
                case 'f': out->push_back('\f'); break;
                case 'n': out->push_back('\n'); break;
                case 'r': out->push_back('\r'); break;
                case 't': out->push_back('\t'); break;
                case 'u': {
                    if (pos + 4 > s.size()) return fail("short \\u escape");
                    const unsigned cp = (unsigned)std::strtoul(s.substr(pos, 4).c_str(), nullptr, 16);
                    pos += 4;
                    utf8(out, cp);
                    break;
                }
                default: return fail("invalid escape");
            }
        }
        return fail("unterminated string");
    }
    bool array(JValue* v, int depth) {
        v->type = JValue::Array;
        pos++;
        ws();
        if (pos >= s.size()) return true; /* left open at the end of the text */
        if (s[pos] == ']') { pos++; return true; }
        for (;;) {
            JValue e;
            if (!value(&e, depth + 1)) return false;
            v->arr.push_back(std::move(e));
            ws();
            if (pos >= s.size()) return true;
            if (s[pos] == ',') { pos++; continue; }
            if (s[pos] == ']') { pos++; return true; }
            return fail("expected ',' or ']'");
        }
    }
    bool object(JValue* v, int depth) {
        v->type = JValue::Object;
        pos++;
        ws();
        if (pos >= s.size()) return true;
        if (s[pos] == '}') { pos++; return true; }
        for (;;) {
            ws();
            if (pos >= s.size() || s[pos] != '"') return fail("expected a key");
            std::string key;
            if (!string(&key)) return false;
            ws();
            if (pos >= s.size() || s[pos] != ':') return fail("expected ':'");
            pos++;
            JValue e;
            if (!value(&e, depth + 1)) return false;
            v->obj.emplace_back(std::move(key), std::move(e));
            ws();
            if (pos >= s.size()) return true; /* left open at the end of the text */
            if (s[pos] == ',') { pos++; continue; }
            if (s[pos] == '}') { pos++; return true; }
            return fail("expected ',' or '}'");
        }
    }
};

/* pretty-printing writer (four-space indent); doubles in the shortest form that reads back
   to the same value */
class JWriter {
public:
    std::string out;
    void beginObject() { open('{'); }
    void endObject() { close('}'); }
    void beginArray() { open('['); }
    void endArray() { close(']'); }
    void key(const char* k) {
        item();
        quoted(k);
        out += ": ";
        afterKey = true;
    }
    void number(double d) {
        item();
        char buf[40];
        for (int prec = 15; prec <= 17; prec++) {
            std::snprintf(buf, sizeof buf, "%.*g", prec, d);
            if (std::strtod(buf, nullptr) == d) break;
        }
        std::string t(buf);
        if (t.find_first_of(".eEn") == std::string::npos) t += ".0"; /* keep it a double */
        out += t;
    }
    void integer(long long i) {
        item();
        out += std::to_string(i);
    }
    void boolean(bool b) {
        item();
        out += b ? "true" : "false";
    }
    void string(const std::string& str) {
        item();
        quoted(str.c_str());
    }

private:
    std::vector<int> counts; /* items written at each open level */
    bool afterKey = false;
    void newline() {
        out += '\n';
        out.append(4 * counts.size(), ' ');
    }
    void item() {
        if (afterKey) {
            afterKey = false;
            return;
        }
        if (!counts.empty()) {
            if (counts.back()++ > 0) out += ',';
            newline();
        }
    }
    void open(char c) {
        item();
        out += c;
        counts.push_back(0);
    }
    void close(char c) {
        const int n = counts.back();
        counts.pop_back();
        if (n > 0) newline();
        out += c;
    }
    void quoted(const char* k) {
        out += '"';
        for (const char* p = k; *p; p++) {
            const unsigned char ch = (unsigned char)*p;
            if (ch == '"' || ch == '\\') { out += '\\'; out += (char)ch; }
            else if (ch == '\n') out += "\\n";
            else if (ch == '\t') out += "\\t";
            else if (ch < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", ch); out += b; }
            else out += (char)ch;
        }
        out += '"';
    }
};

}  // namespace vx
