// Reads a JSON document with vx_json.h, writes it back with the same header's writer.
// Usage: json_roundtrip <in> <out>; exit status 1 and the reason on stderr on a parse error.
#include <fstream>
#include <iostream>
#include <sstream>

#include "vx_json.h"

static void emit(vx::JWriter& w, const vx::JValue& v) {
    switch (v.type) {
        case vx::JValue::Null: w.out += "null"; break;  // not produced by the model format
        case vx::JValue::Bool: w.boolean(v.b); break;
        case vx::JValue::Number:
            if (v.integral) w.integer((long long)v.num);
            else w.number(v.num);
            break;
        case vx::JValue::String: w.string(v.str); break;
        case vx::JValue::Array:
            w.beginArray();
            for (auto& e : v.arr) emit(w, e);
            w.endArray();
            break;
        case vx::JValue::Object:
            w.beginObject();
            for (auto& kv : v.obj) {
                w.key(kv.first.c_str());
                emit(w, kv.second);
            }
            w.endObject();
            break;
    }
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    std::ifstream f(argv[1]);
    std::stringstream b;
    b << f.rdbuf();
    const std::string text = b.str();
    vx::JValue doc;
    std::string err;
    if (!vx::JReader(text).parse(&doc, &err)) {
        std::cerr << err << "\n";
        return 1;
    }
    vx::JWriter w;
    emit(w, doc);
    std::ofstream o(argv[2]);
    o << w.out << "\n";
    return 0;
}
