// rc_scenejson.cpp — scene files: SceneSerializer.load
// (src/main/java/com/retrocam/scene/SceneSerializer.java:72-212) and
// SceneEditor.buildScene (SceneEditor.java:101-107, SceneObject.toPrimitive
// SceneObject.java:80-88) restated for C++ callers.
//
// The reader is not a general JSON parser: it reproduces the Java reader's
// substring scans, so that any file gives the values the Java host reads from
// it (first occurrence of a quoted key anywhere in an object's text, bracket and
// brace matching that ignores strings, numbers taken from the characters
// [0-9.+-eE] after the key, Float.parseFloat's grammar for the parts of an
// array, and 0 for anything it rejects).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rc_host.h"

namespace {

struct JsonMaterial {
    std::string name;
    float albedo[3] = {0.8f, 0.8f, 0.8f};  // Material field initialisers, Material.java:15-22
    float metallic = 0.0f;
    float emission[3] = {0.0f, 0.0f, 0.0f};
    float emissionStrength = 0.0f;
    float roughness = 0.5f;
    float ior = 1.5f;
    float transmission = 0.0f;
};

struct JsonObject {
    std::string name;
    rcs_json_object o;
};

bool is_ws(char c) { return (unsigned char)c <= ' '; }  // String.trim()

// Float.parseFloat(String) (FloatingDecimal): optional whitespace, sign, then
// "NaN", "Infinity", a hexadecimal significand with a binary exponent, or a
// decimal significand with an optional exponent, then an optional [fFdD].
// Returns false where Java throws NumberFormatException.
bool java_parse_float(const std::string &in, float *out) {
    size_t a = 0, b = in.size();
    while (a < b && is_ws(in[a])) ++a;
    while (b > a && is_ws(in[b - 1])) --b;
    if (a == b) return false;
    std::string s = in.substr(a, b - a);
    size_t i = 0;
    bool neg = false;
    if (s[i] == '+' || s[i] == '-') neg = s[i++] == '-';
    const std::string rest = s.substr(i);
    if (rest == "NaN") {
        *out = std::nanf("");
        return true;
    }
    if (rest == "Infinity") {
        *out = neg ? -INFINITY : INFINITY;
        return true;
    }
    std::string body = rest;
    if (!body.empty() && std::strchr("fFdD", body.back())) body.pop_back();
    if (body.empty()) return false;
    auto digits = [&](size_t &k, bool hex) {
        size_t n = 0;
        while (k < body.size() && (hex ? std::isxdigit((unsigned char)body[k]) : std::isdigit((unsigned char)body[k]))) {
            ++k;
            ++n;
        }
        return n;
    };
    size_t k = 0;
    if (body.size() > 1 && body[0] == '0' && (body[1] == 'x' || body[1] == 'X')) {
        k = 2;
        size_t n = digits(k, true);
        if (k < body.size() && body[k] == '.') {
            ++k;
            n += digits(k, true);
        }
        if (n == 0 || k >= body.size() || (body[k] != 'p' && body[k] != 'P')) return false;  // 'p' is required
        ++k;
        if (k < body.size() && (body[k] == '+' || body[k] == '-')) ++k;
        if (digits(k, false) == 0 || k != body.size()) return false;
    } else {
        size_t n = digits(k, false);
        if (k < body.size() && body[k] == '.') {
            ++k;
            n += digits(k, false);
        }
        if (n == 0) return false;
        if (k < body.size() && (body[k] == 'e' || body[k] == 'E')) {
            ++k;
            if (k < body.size() && (body[k] == '+' || body[k] == '-')) ++k;
            if (digits(k, false) == 0) return false;
        }
        if (k != body.size()) return false;
    }
    // the grammar is checked; strtof rounds the value correctly (to nearest,
    // ties to even), as Float.parseFloat does
    const std::string text = (neg ? "-" : "") + body;
    errno = 0;
    *out = std::strtof(text.c_str(), nullptr);
    return true;
}

std::string quoted(const char *key) { return std::string("\"") + key + "\""; }

// arrayBody (:131-144): text between the '[' after the first "key" and its
// matching ']' (no awareness of strings); empty optional when absent.
bool array_body(const std::string &json, const char *key, std::string *out) {
    const std::string tok = quoted(key);
    size_t ki = json.find(tok);
    if (ki == std::string::npos) return false;
    size_t start = json.find('[', ki + tok.size());
    if (start == std::string::npos) return false;
    int depth = 0;
    for (size_t i = start; i < json.size(); ++i) {
        const char c = json[i];
        if (c == '[') {
            ++depth;
        } else if (c == ']') {
            if (--depth == 0) {
                *out = json.substr(start + 1, i - start - 1);
                return true;
            }
        }
    }
    return false;
}

// splitObjects (:147-156): every top-level {...} of an array body.
std::vector<std::string> split_objects(const std::string &body) {
    std::vector<std::string> out;
    int depth = 0;
    long start = -1;
    for (size_t i = 0; i < body.size(); ++i) {
        const char c = body[i];
        if (c == '{') {
            if (depth++ == 0) start = (long)i;
        } else if (c == '}') {
            if (--depth == 0 && start >= 0) {
                out.push_back(body.substr((size_t)start, i + 1 - (size_t)start));
                start = -1;
            }
        }
    }
    return out;
}

void replace_all(std::string &s, const std::string &from, const std::string &to) {
    size_t p = 0;
    while ((p = s.find(from, p)) != std::string::npos) {
        s.replace(p, from.size(), to);
        p += to.size();
    }
}

// strField (:159-176): the text between the first '"' after the key and the
// next unescaped '"', with \" and then \\ unescaped. Returns 1, 0 when
// absent, or -1 where Java's String.substring would throw (an escape at the
// text's last character; objects end in '}', so it does not occur for them).
int str_field(const std::string &obj, const char *key, std::string *out) {
    const std::string tok = quoted(key);
    size_t ki = obj.find(tok);
    if (ki == std::string::npos) return 0;
    size_t scan = ki + tok.size();
    while (scan < obj.size() && obj[scan] != '"') ++scan;
    if (scan >= obj.size()) return 0;
    size_t close = scan + 1;
    while (close < obj.size()) {
        const char c = obj[close];
        if (c == '\\') {
            close += 2;
            continue;
        }
        if (c == '"') break;
        ++close;
    }
    if (close > obj.size()) return -1;
    std::string v = obj.substr(scan + 1, close - scan - 1);
    replace_all(v, "\\\"", "\"");
    replace_all(v, "\\\\", "\\");
    *out = v;
    return 1;
}

// floatField (:179-195): the characters [0-9.+-eE] after the key and any
// ':', ' ', '\t', '\n', '\r'; 0 when absent or rejected.
float float_field(const std::string &obj, const char *key) {
    const std::string tok = quoted(key);
    size_t ki = obj.find(tok);
    if (ki == std::string::npos) return 0.0f;
    size_t i = ki + tok.size();
    while (i < obj.size() && (obj[i] == ':' || obj[i] == ' ' || obj[i] == '\t' || obj[i] == '\n' || obj[i] == '\r')) ++i;
    if (i >= obj.size()) return 0.0f;
    size_t end = i;
    while (end < obj.size()) {
        const char c = obj[end];
        if (std::isdigit((unsigned char)c) || c == '.' || c == '-' || c == '+' || c == 'e' || c == 'E') ++end;
        else break;
    }
    float v = 0.0f;
    if (end > i && java_parse_float(obj.substr(i, end - i), &v)) return v;
    return 0.0f;
}

// floatArr (:198-212): the parts between the '[' after the key and the next
// ']', split at ',' as String.split does (trailing empty parts dropped, a
// lone empty part kept); a part Java rejects reads 0.
bool float_arr(const std::string &obj, const char *key, std::vector<float> *out) {
    const std::string tok = quoted(key);
    size_t ki = obj.find(tok);
    if (ki == std::string::npos) return false;
    const size_t start = obj.find('[', ki + tok.size());
    const size_t end = obj.find(']', start != std::string::npos && start > 0 ? start : 0);
    if (start == std::string::npos || end == std::string::npos) return false;
    if (end < start + 1) return false;  // Java: substring(start + 1, end) throws when end <= start
    const std::string inner = obj.substr(start + 1, end - start - 1);
    std::vector<std::string> parts;
    size_t p = 0;
    for (;;) {
        size_t q = inner.find(',', p);
        parts.push_back(inner.substr(p, q == std::string::npos ? std::string::npos : q - p));
        if (q == std::string::npos) break;
        p = q + 1;
    }
    if (!inner.empty())
        while (!parts.empty() && parts.back().empty()) parts.pop_back();
    out->clear();
    for (const std::string &s : parts) {
        float v = 0.0f;
        out->push_back(java_parse_float(s, &v) ? v : 0.0f);
    }
    return true;
}

// (int) f in Java: NaN -> 0, truncation toward zero, saturated to the int range
int32_t java_f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)-2147483647 - 1;
    return (int32_t)f;
}

bool iequals(const std::string &a, const char *b) {
    const size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::toupper((unsigned char)a[i]) != std::toupper((unsigned char)b[i])) return false;
    return true;
}

int32_t copy_name(const std::string &s, char *dst, int32_t cap) {
    if (dst && cap > 0) {
        const size_t n = s.size() < (size_t)(cap - 1) ? s.size() : (size_t)(cap - 1);
        std::memcpy(dst, s.data(), n);
        dst[n] = '\0';
    }
    return (int32_t)s.size();
}

}  // namespace

struct rcs_doc {
    std::vector<JsonMaterial> mats;
    std::vector<JsonObject> objs;
};

extern "C" {

rcs_doc *rcs_json_parse(const char *text, size_t len, int32_t *status) {  // SceneSerializer.load, :72-124
    int32_t dummy;
    int32_t &st = status ? *status : dummy;
    st = RCS_JSON_OK;
    if (!text && len) {
        st = RCS_JSON_ERR_ARG;
        return nullptr;
    }
    rcs_doc *d = new (std::nothrow) rcs_doc();
    if (!d) {
        st = RCS_JSON_ERR_OOM;
        return nullptr;
    }
    auto fail = [&]() -> rcs_doc * {
        delete d;
        st = RCS_JSON_ERR_STRING;
        return nullptr;
    };
    const std::string json(text ? text : "", len);
    std::string body;
    if (array_body(json, "materials", &body)) {
        for (const std::string &obj : split_objects(body)) {
            JsonMaterial m;
            std::string name;
            const int hasName = str_field(obj, "name", &name);
            if (hasName < 0) return fail();
            std::vector<float> alb, em;
            if (float_arr(obj, "albedo", &alb) && alb.size() == 3) std::memcpy(m.albedo, alb.data(), 12);
            if (float_arr(obj, "emission", &em) && em.size() == 3) std::memcpy(m.emission, em.data(), 12);
            m.metallic = float_field(obj, "metallic");
            m.emissionStrength = float_field(obj, "emissionStrength");
            m.roughness = float_field(obj, "roughness");
            m.ior = float_field(obj, "ior");
            if (m.ior < 0.01f) m.ior = 1.5f;  // missing field guard (:90); NaN is kept
            m.transmission = float_field(obj, "transmission");
            m.name = hasName > 0 ? name : "Material";
            d->mats.push_back(m);
        }
    }
    if (array_body(json, "objects", &body)) {
        for (const std::string &obj : split_objects(body)) {
            JsonObject jo;
            std::string name, type;
            const int hasName = str_field(obj, "name", &name);
            const int hasType = hasName < 0 ? 0 : str_field(obj, "type", &type);
            if (hasName < 0 || hasType < 0) return fail();
            rcs_json_object &o = jo.o;
            o.pos[0] = float_field(obj, "px");
            o.pos[1] = float_field(obj, "py");
            o.pos[2] = float_field(obj, "pz");
            o.scale[0] = float_field(obj, "sx");
            o.scale[1] = float_field(obj, "sy");
            o.scale[2] = float_field(obj, "sz");
            o.rot[0] = float_field(obj, "rx");
            o.rot[1] = float_field(obj, "ry");
            o.rot[2] = float_field(obj, "rz");
            const int32_t stacks = java_f2i(float_field(obj, "stacks"));
            const int32_t slices = java_f2i(float_field(obj, "slices"));
            const int32_t matIdx = java_f2i(float_field(obj, "materialIndex"));
            o.type = (hasType > 0 && iequals(type, "SPHERE")) ? RCS_OBJ_SPHERE : RCS_OBJ_BOX;
            o.materialIndex = matIdx > 0 ? matIdx : 0;
            o.stacks = stacks > 2 ? stacks : 24;  // SceneObject defaults, SceneObject.java:28
            o.slices = slices > 2 ? slices : 24;
            jo.name = hasName > 0 ? name : "Object";
            d->objs.push_back(jo);
        }
    }
    return d;
}

void rcs_json_free(rcs_doc *d) { delete d; }

int32_t rcs_json_material_count(const rcs_doc *d) { return d ? (int32_t)d->mats.size() : -1; }

int32_t rcs_json_object_count(const rcs_doc *d) { return d ? (int32_t)d->objs.size() : -1; }

int32_t rcs_json_get_material(const rcs_doc *d, int32_t i, float *mat16, char *name, int32_t nameCap) {
    if (!d || i < 0 || i >= (int32_t)d->mats.size()) return -1;
    const JsonMaterial &m = d->mats[i];
    if (mat16) {  // Material.write layout (Material.java:61-66), albedoTex = -1
        float w[16] = {m.albedo[0], m.albedo[1], m.albedo[2], m.metallic,
                       m.emission[0], m.emission[1], m.emission[2], m.emissionStrength,
                       m.roughness, m.ior, m.transmission, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const int32_t noTex = -1;
        std::memcpy(&w[11], &noTex, 4);
        std::memcpy(mat16, w, sizeof w);
    }
    return copy_name(m.name, name, nameCap);
}

int32_t rcs_json_get_object(const rcs_doc *d, int32_t i, rcs_json_object *o, char *name, int32_t nameCap) {
    if (!d || i < 0 || i >= (int32_t)d->objs.size()) return -1;
    if (o) *o = d->objs[i].o;
    return copy_name(d->objs[i].name, name, nameCap);
}

int32_t rcs_add_object(rcs_scene *s, const rcs_json_object *o) {  // SceneObject.toPrimitive, :80-88
    if (!s || !o) return -1;
    const int32_t first = rcs_triangle_count(s);
    int32_t n;
    if (o->type == RCS_OBJ_SPHERE)
        n = rcs_add_sphere(s, o->pos[0], o->pos[1], o->pos[2], o->scale[0], o->materialIndex, o->stacks, o->slices);
    else
        n = rcs_add_box(s, o->pos[0], o->pos[1], o->pos[2], o->scale[0], o->scale[1], o->scale[2], o->materialIndex);
    if (n < 0) return -1;
    if (o->rot[0] != 0.0f || o->rot[1] != 0.0f || o->rot[2] != 0.0f)
        if (rcs_rotate_since(s, first, o->pos[0], o->pos[1], o->pos[2], o->rot[0], o->rot[1], o->rot[2]) < 0) return -1;
    return n;
}

int32_t rcs_json_build(const rcs_doc *d, rcs_scene *s) {  // SceneEditor.buildScene, :101-107 (before Scene.build)
    if (!d || !s) return -1;
    for (const JsonMaterial &m : d->mats) {
        float w[16];
        rcs_json_get_material(d, (int32_t)(&m - d->mats.data()), w, nullptr, 0);
        if (rcs_add_material(s, w) < 0) return -1;
    }
    int32_t tris = 0;
    for (const JsonObject &jo : d->objs) {
        const int32_t n = rcs_add_object(s, &jo.o);
        if (n < 0) return -1;
        tris += n;
    }
    return tris;
}

}  // extern "C"
