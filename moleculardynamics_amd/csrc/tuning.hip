// tuning.hip -- launch shapes of the LDS-staged kernels from the tuning table.
//
// The table is a JSON file written by the project's tuner
// (python -m moleculardynamics_amd.tune): {"format": "mdx-tuning-1",
// "entries": [{"arch": "gfx950", "kind": "lj", "atoms_min": 0, "atoms_max":
// 65536, "block": 512, "fill": 0.85, ...}, ...]}.  An entry applies to a
// system of its kind ("lj": the periodic Lennard-Jones box; "replicas": a
// batch, by its total atom count) on a device whose architecture name starts
// with its arch, when atoms_min <= atoms < atoms_max; the first such entry
// wins.  Keys other than the shape's are carried for the record (the tuner's
// timings) and ignored here.  The file is MDX_TUNING_FILE when that is set (an
// empty value: no table), else data/tuning.json beside libmdx.so; without a
// table, or without an entry, the built-in shape applies.
#include <dlfcn.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mdx_internal.h"

namespace mdx {

namespace {

// A small strict JSON reader: objects, arrays, strings (escapes \" \\ \/ \n
// \t \r \b \f; \u is rejected), numbers, true, false, null.
struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0;
    std::string str;
    std::vector<Value> items;
    std::vector<std::pair<std::string, Value>> members;
    const Value* get(const char* key) const {
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct Parser {
    const std::string& t;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string& text) : t(text) {}
    void ws() {
        while (i < t.size() && std::isspace((unsigned char)t[i])) i++;
    }
    bool fail(const char* what) {
        if (err.empty()) err = std::string(what) + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (t.compare(i, n, w) != 0) return fail("unexpected token");
        i += n;
        return true;
    }
    bool string(std::string& out) {
        if (i >= t.size() || t[i] != '"') return fail("expected a string");
        i++;
        out.clear();
        while (i < t.size() && t[i] != '"') {
            char ch = t[i++];
            if (ch == '\\') {
                if (i >= t.size()) return fail("bad escape");
                const char e = t[i++];
                switch (e) {
                    case '"': ch = '"'; break;
                    case '\\': ch = '\\'; break;
                    case '/': ch = '/'; break;
                    case 'n': ch = '\n'; break;
                    case 't': ch = '\t'; break;
                    case 'r': ch = '\r'; break;
                    case 'b': ch = '\b'; break;
                    case 'f': ch = '\f'; break;
                    default: return fail("unsupported escape");
                }
            }
            out.push_back(ch);
        }
        if (i >= t.size()) return fail("unterminated string");
        i++;
        return true;
    }
    bool value(Value& v, int depth) {
        if (depth > 32) return fail("nesting too deep");
        ws();
        if (i >= t.size()) return fail("unexpected end");
        const char ch = t[i];
        if (ch == '{') {
            v.kind = Value::Object;
            i++;
            ws();
            if (i < t.size() && t[i] == '}') { i++; return true; }
            for (;;) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (i >= t.size() || t[i] != ':') return fail("expected ':'");
                i++;
                Value m;
                if (!value(m, depth + 1)) return false;
                v.members.emplace_back(std::move(key), std::move(m));
                ws();
                if (i < t.size() && t[i] == ',') { i++; continue; }
                if (i < t.size() && t[i] == '}') { i++; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (ch == '[') {
            v.kind = Value::Array;
            i++;
            ws();
            if (i < t.size() && t[i] == ']') { i++; return true; }
            for (;;) {
                Value m;
                if (!value(m, depth + 1)) return false;
                v.items.push_back(std::move(m));
                ws();
                if (i < t.size() && t[i] == ',') { i++; continue; }
                if (i < t.size() && t[i] == ']') { i++; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (ch == '"') {
            v.kind = Value::String;
            return string(v.str);
        }
        if (ch == 't') { v.kind = Value::Bool; v.num = 1; return lit("true"); }
        if (ch == 'f') { v.kind = Value::Bool; v.num = 0; return lit("false"); }
        if (ch == 'n') { v.kind = Value::Null; return lit("null"); }
        // number
        const char* begin = t.c_str() + i;
        char* end = nullptr;
        v.num = std::strtod(begin, &end);
        if (end == begin) return fail("expected a value");
        v.kind = Value::Number;
        i += (size_t)(end - begin);
        return true;
    }
};

bool number(const Value& obj, const char* key, double& out) {
    const Value* v = obj.get(key);
    if (!v || v->kind != Value::Number) return false;
    out = v->num;
    return true;
}

// the table's entry for (arch, kind, atoms): 1 found, 0 none, -1 a bad file
int lookup(const std::string& text, const char* arch, const char* kind, int64_t atoms, mdx_shape* out,
           std::string& err) {
    Value root;
    Parser p(text);
    if (!p.value(root, 0)) {
        err = "tuning table: " + p.err;
        return -1;
    }
    p.ws();
    if (p.i != text.size()) {
        err = "tuning table: trailing characters";
        return -1;
    }
    const Value* fmt = root.get("format");
    if (root.kind != Value::Object || !fmt || fmt->kind != Value::String || fmt->str != "mdx-tuning-1") {
        err = "tuning table: not an mdx-tuning-1 object";
        return -1;
    }
    const Value* entries = root.get("entries");
    if (!entries || entries->kind != Value::Array) {
        err = "tuning table: no entries array";
        return -1;
    }
    for (const Value& e : entries->items) {
        if (e.kind != Value::Object) continue;
        const Value* a = e.get("arch");
        const Value* k = e.get("kind");
        double lo = 0, hi = 0;
        if (!a || a->kind != Value::String || !k || k->kind != Value::String) continue;
        if (std::strncmp(arch, a->str.c_str(), a->str.size()) != 0 || a->str.empty()) continue;
        if (k->str != kind) continue;
        if (!number(e, "atoms_min", lo) || !number(e, "atoms_max", hi)) continue;
        if (!((double)atoms >= lo && (double)atoms < hi)) continue;
        double block, fill, aim;
        if (number(e, "block", block)) {
            if (block != 512 && block != 1024) {
                err = "tuning table: block must be 512 or 1024";
                return -1;
            }
            out->block = (int32_t)block;
        }
        if (number(e, "fill", fill)) {
            if (!(fill >= 0.25 && fill <= 1.0)) {
                err = "tuning table: fill must lie in [0.25, 1]";
                return -1;
            }
            out->fill = (float)fill;
        }
        if (number(e, "stage_aim", aim)) {
            if (!(aim >= 256 && aim <= 2560)) {
                err = "tuning table: stage_aim must lie in [256, 2560]";
                return -1;
            }
            out->stage_aim = (float)aim;
        }
        out->source = MDX_SHAPE_TUNED;
        return 1;
    }
    return 0;
}

bool read_file(const std::string& path, std::string& text) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    text = ss.str();
    return true;
}

std::string table_path() {
    if (const char* env = std::getenv("MDX_TUNING_FILE")) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&mdx_tuning_lookup), &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/data/tuning.json";
    }
    return "";
}

}  // namespace

// The built-in shape: 512-thread workgroups; segments of about 0.85 of one
// pass of home atoms (a few cells' worth of slack keeps most segments to one
// pass); a replica segment's staging aims at 1250 atoms (the step kernel's
// FP64 staging at 1.2 x the largest segment then stays within a quarter of
// a CU's 160 KiB of LDS).
mdx_shape default_shape() {
    mdx_shape s{};
    s.block = 512;
    s.fill = 0.85f;
    s.stage_aim = 1250.0f;
    s.source = MDX_SHAPE_DEFAULT;
    return s;
}

int tuned_shape(const char* kind, int64_t atoms, mdx_shape* out, std::string& err) {
    *out = default_shape();
    const std::string path = table_path();
    if (path.empty()) return MDX_OK;
    std::string text;
    if (!read_file(path, text)) return MDX_OK;   // no table: the built-in shape
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        err = "tuning: cannot read the device's architecture";
        return MDX_ERR_HIP;
    }
    mdx_shape s = *out;
    const int r = lookup(text, prop.gcnArchName, kind, atoms, &s, err);
    if (r < 0) {
        err += " (" + path + ")";
        return MDX_ERR_ARG;
    }
    if (r > 0) *out = s;
    return MDX_OK;
}

}  // namespace mdx

namespace {
thread_local std::string g_tuning_err;
}

extern "C" int mdx_tuning_lookup(const char* path, const char* arch, const char* kind, int64_t atoms,
                                 mdx_shape* out) {
    if (!path || !arch || !kind || !out) {
        g_tuning_err = "null argument";
        return MDX_ERR_ARG;
    }
    *out = mdx::default_shape();
    std::string text, err;
    if (!mdx::read_file(path, text)) {
        g_tuning_err = std::string("cannot read ") + path;
        return MDX_ERR_ARG;
    }
    mdx_shape s = *out;
    const int r = mdx::lookup(text, arch, kind, atoms, &s, err);
    if (r < 0) {
        g_tuning_err = err;
        return MDX_ERR_ARG;
    }
    if (r > 0) *out = s;
    return MDX_OK;
}

extern "C" const char* mdx_tuning_last_error(void) { return g_tuning_err.c_str(); }
