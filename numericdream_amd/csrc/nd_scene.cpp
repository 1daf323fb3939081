// nd_scene.cpp - .dream scene files (JSON) -> the export's Uniforms block (host side).
//
// Follows parseScene / applyScene (src/app/scene_file.zig:532-535, 447-755) for what a file
// holds and which fields it may omit, then buildUniforms (src/app/export_image.zig:97-167)
// with the export's render settings (src/ui/main_panels.zig:222-245, src/app/bench.zig:
// 132-160), and the renderer's own stamps: formula codes (fractal_renderer.zig:2882-2894),
// approximations (:3033-3042) and the sky (:3107-3109). The photon map and the empty-space
// accelerator that an export builds are reported in nd_dream_info, not stamped: the first is
// not part of this library's path trace, the second is the caller's choice (nd_accel_setup).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "nd_config.h"

namespace {

// ------------------------------------------------------------------------------- JSON
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal* get(const char* key) const {
        if (kind != OBJ) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JParser {
   public:
    JParser(const char* p, size_t n) : p_(p), b_(p), e_(p + n) {}
    bool parse(JVal& out, std::string& err) {
        ws();
        if (!value(out, 0)) { err = err_; return false; }
        ws();
        if (p_ != e_) { err = at("trailing characters after the scene object"); return false; }
        return true;
    }

   private:
    const char* p_;
    const char* b_;
    const char* e_;
    std::string err_;

    std::string at(const char* what) const {
        char buf[160];
        std::snprintf(buf, sizeof buf, "invalid JSON at byte %zu: %s", (size_t)(p_ - b_), what);
        return buf;
    }
    bool fail(const char* what) { if (err_.empty()) err_ = at(what); return false; }
    void ws() { while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_; }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if ((size_t)(e_ - p_) < n || std::memcmp(p_, w, n) != 0) return false;
        p_ += n;
        return true;
    }
    static void utf8(std::string& s, unsigned cp) {
        if (cp < 0x80) s += (char)cp;
        else if (cp < 0x800) { s += (char)(0xC0 | (cp >> 6)); s += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) { s += (char)(0xE0 | (cp >> 12)); s += (char)(0x80 | ((cp >> 6) & 0x3F)); s += (char)(0x80 | (cp & 0x3F)); }
        else { s += (char)(0xF0 | (cp >> 18)); s += (char)(0x80 | ((cp >> 12) & 0x3F)); s += (char)(0x80 | ((cp >> 6) & 0x3F)); s += (char)(0x80 | (cp & 0x3F)); }
    }
    bool hex4(unsigned& v) {
        if (e_ - p_ < 4) return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool str(std::string& s) {
        ++p_;  // opening quote
        while (p_ < e_ && *p_ != '"') {
            unsigned char c = (unsigned char)*p_;
            if (c < 0x20) return fail("control character in a string");
            if (c != '\\') { s += (char)c; ++p_; continue; }
            if (++p_ >= e_) return fail("unterminated escape");
            char x = *p_++;
            switch (x) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    unsigned cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        unsigned lo;
                        if (!lit("\\u") || !hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return fail("unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(s, cp);
                    break;
                }
                default: return fail("unknown escape");
            }
        }
        if (p_ >= e_) return fail("unterminated string");
        ++p_;
        return true;
    }
    bool number(double& v) {
        const char* s = p_;
        if (p_ < e_ && *p_ == '-') ++p_;
        if (p_ >= e_ || !(*p_ >= '0' && *p_ <= '9')) return fail("bad number");
        while (p_ < e_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '+' || *p_ == '-')) ++p_;
        std::string t(s, p_);
        char* end = nullptr;
        v = std::strtod(t.c_str(), &end);
        if (!end || *end) return fail("bad number");
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        if (p_ >= e_) return fail("unexpected end");
        char c = *p_;
        if (c == '{') {
            v.kind = JVal::OBJ;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == '}') { ++p_; return true; }
            for (;;) {
                ws();
                if (p_ >= e_ || *p_ != '"') return fail("expected a key");
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p_ >= e_ || *p_ != ':') return fail("expected ':'");
                ++p_;
                ws();
                v.o.emplace_back(std::move(k), JVal());
                if (!value(v.o.back().second, depth + 1)) return false;
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == '}') { ++p_; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::ARR;
            ++p_;
            ws();
            if (p_ < e_ && *p_ == ']') { ++p_; return true; }
            for (;;) {
                ws();
                v.a.emplace_back();
                if (!value(v.a.back(), depth + 1)) return false;
                ws();
                if (p_ < e_ && *p_ == ',') { ++p_; continue; }
                if (p_ < e_ && *p_ == ']') { ++p_; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.kind = JVal::STR; return str(v.s); }
        if (lit("true")) { v.kind = JVal::BOOL; v.b = true; return true; }
        if (lit("false")) { v.kind = JVal::BOOL; v.b = false; return true; }
        if (lit("null")) { v.kind = JVal::NUL; return true; }
        v.kind = JVal::NUM;
        return number(v.n);
    }
};

// -------------------------------------------------------------------- typed field reads
// The file's fields map onto the Zig structs the app parses it into: a field without a
// default in the struct is required, an unknown field is ignored (ignore_unknown_fields).
struct Reader {
    std::string err;
    bool ok() const { return err.empty(); }
    void fail(const std::string& path, const char* what) {
        if (err.empty()) err = path + ": " + what;
    }
    const JVal* field(const JVal& obj, const std::string& path, const char* key, bool required) {
        const JVal* v = obj.get(key);
        if (!v && required) fail(path + "." + key, "missing");
        return v;
    }
    float f32(const JVal& obj, const std::string& path, const char* key, bool required, float def) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return def;
        if (v->kind != JVal::NUM) { fail(path + "." + key, "not a number"); return def; }
        return (float)v->n;
    }
    long long integer(const JVal& obj, const std::string& path, const char* key, bool required, long long def) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return def;
        if (v->kind != JVal::NUM || v->n != std::floor(v->n)) { fail(path + "." + key, "not an integer"); return def; }
        return (long long)v->n;
    }
    bool boolean(const JVal& obj, const std::string& path, const char* key, bool required, bool def) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return def;
        if (v->kind != JVal::BOOL) { fail(path + "." + key, "not a boolean"); return def; }
        return v->b;
    }
    std::string string(const JVal& obj, const std::string& path, const char* key, bool required, const char* def) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return def;
        if (v->kind != JVal::STR) { fail(path + "." + key, "not a string"); return def; }
        return v->s;
    }
    // a Vec3 object {x, y, z}
    void vec3(const JVal& obj, const std::string& path, const char* key, bool required, float out[3]) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return;
        if (v->kind != JVal::OBJ) { fail(path + "." + key, "not an object"); return; }
        const std::string p = path + "." + key;
        out[0] = f32(*v, p, "x", true, out[0]);
        out[1] = f32(*v, p, "y", true, out[1]);
        out[2] = f32(*v, p, "z", true, out[2]);
    }
    // a [3]f32 array
    void arr3(const JVal& obj, const std::string& path, const char* key, bool required, float out[3]) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return;
        if (v->kind != JVal::ARR || v->a.size() != 3) { fail(path + "." + key, "not an array of 3 numbers"); return; }
        for (int i = 0; i < 3; ++i) {
            if (v->a[i].kind != JVal::NUM) { fail(path + "." + key, "not an array of 3 numbers"); return; }
            out[i] = (float)v->a[i].n;
        }
    }
    // a SliderRange {min, max}: UI state, read only to hold the file to the app's format
    void range(const JVal& obj, const std::string& path, const char* key, bool required) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return;
        if (v->kind != JVal::OBJ) { fail(path + "." + key, "not an object"); return; }
        (void)f32(*v, path + "." + key, "min", true, 0);
        (void)f32(*v, path + "." + key, "max", true, 0);
    }
    // an enum by its tag name
    int tag(const JVal& obj, const std::string& path, const char* key, bool required, int def,
            std::initializer_list<const char*> names) {
        const JVal* v = field(obj, path, key, required);
        if (!v) return def;
        if (v->kind != JVal::STR) { fail(path + "." + key, "not an enum name"); return def; }
        int i = 0;
        for (const char* n : names) {
            if (v->s == n) return i;
            ++i;
        }
        fail(path + "." + key, "unknown enum value");
        return def;
    }
    const JVal* array(const JVal& obj, const std::string& path, const char* key, bool required) {
        const JVal* v = field(obj, path, key, required);
        if (v && v->kind != JVal::ARR) { fail(path + "." + key, "not an array"); return nullptr; }
        return v;
    }
    const JVal* object(const JVal& obj, const std::string& path, const char* key, bool required) {
        const JVal* v = field(obj, path, key, required);
        if (v && v->kind != JVal::OBJ) { fail(path + "." + key, "not an object"); return nullptr; }
        return v;
    }
};

std::string idx(const std::string& path, const char* key, size_t i) {
    char b[32];
    std::snprintf(b, sizeof b, "[%zu]", i);
    return path + "." + key + b;
}

// ------------------------------------------------------------------------ formula table
struct FormulaRow {
    int id;
    const char* stem;
    int n;
    const char* names[16];
    float defaults[16];
};
const FormulaRow kFormulas[] = {
#include "nd_formula_table.inc"
};

// A formula path as a scene names it: "" is the built-in Mandelbulb (mandelbrot.wgsl);
// otherwise the file's stem (directory and .wgsl dropped) selects the bound formula.
const FormulaRow* formula_of(const std::string& path) {
    std::string stem = path;
    size_t slash = stem.find_last_of("/\\");
    if (slash != std::string::npos) stem = stem.substr(slash + 1);
    if (stem.size() > 5 && stem.compare(stem.size() - 5, 5, ".wgsl") == 0) stem.resize(stem.size() - 5);
    if (path.empty()) stem = "mandelbrot";
    for (const FormulaRow& r : kFormulas)
        if (stem == r.stem) return &r;
    return nullptr;
}

// The formula's parameters at their @param defaults, then each saved (name, value) applied to
// the first parameter of that label (applyParamDtosTo, scene_file.zig:586-601; labels hold at
// most 23 bytes, formula.zig:19-32). Packed in declaration order (packParams, formula.zig:104-108).
void formula_params(Reader& R, const FormulaRow& f, const JVal* saved, const std::string& path, float out[16]) {
    for (int i = 0; i < 16; ++i) out[i] = i < f.n ? f.defaults[i] : 0.0f;
    if (!saved) return;
    for (size_t k = 0; k < saved->a.size() && R.ok(); ++k) {
        const JVal& p = saved->a[k];
        const std::string pp = idx(path, "params", k);
        if (p.kind != JVal::OBJ) { R.fail(pp, "not an object"); return; }
        std::string name = R.string(p, pp, "name", true, "");
        const float value = R.f32(p, pp, "value", true, 0.0f);
        R.range(p, pp, "range", true);
        if (name.size() > 23) name.resize(23);
        for (int i = 0; i < f.n; ++i) {
            std::string label = f.names[i];
            if (label.size() > 23) label.resize(23);
            if (label == name) { out[i] = value; break; }
        }
    }
}

void set3(float* d, float a, float b, float c) { d[0] = a; d[1] = b; d[2] = c; }

float radians(float deg) { return deg * (float)(M_PI / 180.0); }

// eulerFrame (camera.zig:41-60): the columns of Rz(rz) Ry(ry) Rx(rx), each padded to 4 floats
void euler_frame(float rx, float ry, float rz, float frame[3][4]) {
    const float cx = std::cos(rx), sx = std::sin(rx), cy = std::cos(ry), sy = std::sin(ry);
    const float cz = std::cos(rz), sz = std::sin(rz);
    const float cols[3][3] = {{cy * cz, cy * sz, -sy},
                              {sx * sy * cz - cx * sz, sx * sy * sz + cx * cz, sx * cy},
                              {cx * sy * cz + sx * sz, cx * sy * sz - sx * cz, cx * cy}};
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) frame[i][k] = cols[i][k];
        frame[i][3] = 0.0f;
    }
}

// The code table of the scene: code c -> bound formula, in order of first use (the shader's
// CodeTable assigns each distinct formula of the scene a code).
struct Codes {
    int builtin[ND_MAX_CODES];
    uint32_t n = 0;
    int code_of(int id, Reader& R, const std::string& path) {
        for (uint32_t c = 0; c < n; ++c)
            if (builtin[c] == id) return (int)c;
        if (n == ND_MAX_CODES) { R.fail(path, "more distinct formulas than codes"); return 0; }
        builtin[n] = id;
        return (int)n++;
    }
};

// ColorStopState (scene_state.zig:33-66) -> ColorStop (toGpu, :67-87)
void color_stop(Reader& R, const JVal& c, const std::string& p, nd_color_stop* s) {
    std::memset(s, 0, sizeof(*s));
    s->position = R.f32(c, p, "position", true, 0);
    R.arr3(c, p, "color", true, s->color);
    s->glossiness = R.f32(c, p, "glossiness", false, 0.3f);
    s->transparency = R.f32(c, p, "transparency", false, 0.0f);
    s->reflectiveness = R.f32(c, p, "reflectiveness", false, 0.1f);
    s->ior = R.f32(c, p, "ior", false, 1.5f);
    s->subsurface = R.f32(c, p, "subsurface", false, 0.0f);
    s->subsurface_radius = R.f32(c, p, "subsurface_radius", false, 0.1f);
    s->abbe = R.f32(c, p, "abbe", false, 0.0f);
    s->inner_max_steps = R.f32(c, p, "inner_max_steps", false, 64.0f);
    s->roughness = R.f32(c, p, "roughness", false, 0.0f);
    s->film_thickness = R.f32(c, p, "film_thickness", false, 400.0f);
    s->film_ior = R.f32(c, p, "film_ior", false, 1.8f);
    s->film_strength = R.f32(c, p, "film_strength", false, 0.0f);
    s->film_angle_scale = R.f32(c, p, "film_angle_scale", false, 1.0f);
    s->film_perturb = R.f32(c, p, "film_perturb", false, 0.0f);
    s->film_perturb_scale = R.f32(c, p, "film_perturb_scale", false, 0.2f);
}

// InstanceDto (scene_file.zig:151-193) -> FractalInstanceState -> FractalInstance (toGpu,
// scene_state.zig:278-327), formula codes stamped as the renderer does.
// images[0] receives the path of the instance formula's image (its "image" field), images[1 + j]
// mix-in j's: the host loads them into image slots instance * 4 and instance * 4 + 1 + j.
void instance(Reader& R, const JVal& in, const std::string& p, Codes& codes, nd_fractal_instance* g,
              std::string images[1 + ND_MAX_MIXINS]) {
    std::memset(g, 0, sizeof(*g));
    float offset[3] = {0, 0, 0}, scale[3] = {1, 1, 1}, rot[3] = {0, 0, 0};
    R.vec3(in, p, "offset", true, offset);
    R.range(in, p, "offset_range_x", true); R.range(in, p, "offset_range_y", true); R.range(in, p, "offset_range_z", true);
    R.vec3(in, p, "scale", true, scale);
    R.range(in, p, "scale_range_x", true); R.range(in, p, "scale_range_y", true); R.range(in, p, "scale_range_z", true);
    const float su = R.f32(in, p, "scale_uniform", false, 1.0f);
    R.range(in, p, "scale_uniform_range", false);
    R.vec3(in, p, "rotation", true, rot);
    R.range(in, p, "rotation_range_x", true); R.range(in, p, "rotation_range_y", true); R.range(in, p, "rotation_range_z", true);
    g->step_safety = R.f32(in, p, "step_safety", true, 1.0f);
    R.range(in, p, "step_safety_range", true);
    g->combine_mode = (float)R.tag(in, p, "combine_mode", true, 0,
                                   {"hard_union", "de_combinate", "smooth_max", "smooth_inv_max", "smooth_min_lin",
                                    "smooth_min_nlin", "smooth_mix"});
    g->blend_k = R.f32(in, p, "blend_k", true, 0.5f);
    R.range(in, p, "blend_k_range", true);
    g->hybrid_base_iters = R.f32(in, p, "hybrid_base_iters", true, 2.0f);
    R.range(in, p, "hybrid_base_iters_range", true);
    g->hybrid_total_iters = R.f32(in, p, "hybrid_total_iters", true, 12.0f);
    R.range(in, p, "hybrid_total_iters_range", true);
    g->hidden = R.boolean(in, p, "visible", false, true) ? 0.0f : 1.0f;
    if (R.boolean(in, p, "ft_view", false, false)) R.fail(p + ".ft_view", "the FFT view is not implemented");
    for (int a = 0; a < 3; ++a) {
        g->offset[a] = offset[a];
        g->scale[a] = scale[a] * su;
        g->rotation[a] = radians(rot[a]);
    }
    euler_frame(g->rotation[0], g->rotation[1], g->rotation[2], g->frame);

    // formula and mix-ins
    const JVal* fm = R.object(in, p, "formula", true);
    if (fm) {
        const std::string fp = p + ".formula";
        const std::string path = R.string(*fm, fp, "path", true, "");
        const JVal* params = R.array(*fm, fp, "params", true);
        images[0] = R.string(*fm, fp, "image", false, "");
        const FormulaRow* f = formula_of(path);
        if (!f) { R.fail(fp + ".path", ("no bound formula for '" + path + "'").c_str()); return; }
        formula_params(R, *f, params, fp, g->params);
        const JVal* mixins = R.array(in, p, "mixins", true);
        const size_t nm = mixins ? std::min(mixins->a.size(), (size_t)ND_MAX_MIXINS) : 0;
        g->mixin_count = (float)nm;
        // a formula with mix-ins runs as the base of a hybrid cycle (codeOf(base, standalone))
        g->code = (float)codes.code_of(f->id, R, fp);
        for (int j = 0; j < ND_MAX_MIXINS; ++j) {
            g->mixins[j].iterations = 1.0f;
            g->mixins[j].code = (float)ND_NO_CODE;
        }
        for (size_t j = 0; j < nm && R.ok(); ++j) {
            const JVal& m = mixins->a[j];
            const std::string mp = idx(p, "mixins", j);
            if (m.kind != JVal::OBJ) { R.fail(mp, "not an object"); return; }
            g->mixins[j].iterations = R.f32(m, mp, "iterations", true, 1.0f);
            R.range(m, mp, "iterations_range", true);
            const JVal* mf = R.object(m, mp, "formula", true);
            if (!mf) return;
            const std::string mfp = mp + ".formula";
            const std::string mpath = R.string(*mf, mfp, "path", true, "");
            const JVal* mparams = R.array(*mf, mfp, "params", true);
            images[1 + j] = R.string(*mf, mfp, "image", false, "");
            const FormulaRow* mfr = formula_of(mpath);
            if (!mfr) { R.fail(mfp + ".path", ("no bound formula for '" + mpath + "'").c_str()); return; }
            formula_params(R, *mfr, mparams, mfp, g->mixins[j].params);
            g->mixins[j].code = (float)codes.code_of(mfr->id, R, mfp);
        }
    }

    // colour stops, sorted by position (sortedColorOrder, scene_state.zig:97-108: a stable
    // insertion sort); unused slots zero
    const JVal* colors = R.array(in, p, "colors", true);
    if (colors) {
        const size_t nc = std::min(colors->a.size(), (size_t)ND_MAX_COLOR_STOPS);
        nd_color_stop st[ND_MAX_COLOR_STOPS];
        for (size_t k = 0; k < nc && R.ok(); ++k) {
            if (colors->a[k].kind != JVal::OBJ) { R.fail(idx(p, "colors", k), "not an object"); return; }
            color_stop(R, colors->a[k], idx(p, "colors", k), &st[k]);
        }
        size_t order[ND_MAX_COLOR_STOPS];
        for (size_t k = 0; k < nc; ++k) order[k] = k;
        for (size_t k = 1; k < nc; ++k)
            for (size_t j = k; j > 0 && st[order[j - 1]].position > st[order[j]].position; --j) std::swap(order[j], order[j - 1]);
        for (size_t k = 0; k < nc; ++k) g->colors[k] = st[order[k]];
        g->color_count = (float)nc;
    }

    // orbit trap (OrbitTrapState, scene_state.zig:190-211)
    g->trap_box[0] = g->trap_box[1] = g->trap_box[2] = 1.0f;
    g->trap_radius = 1.0f;
    g->trap_span = 1.5f;
    if (const JVal* t = R.object(in, p, "trap", false)) {
        const std::string tp = p + ".trap";
        g->trap_shape = (float)R.tag(*t, tp, "shape", false, 0, {"point", "sphere", "box", "cross", "torus"});
        g->trap_mode = (float)R.tag(*t, tp, "mode", false, 0, {"min", "max", "average", "last", "iteration"});
        g->trap_repeat = (float)R.tag(*t, tp, "repeat", false, 0, {"clamp", "repeat", "mirror"});
        R.vec3(*t, tp, "center", false, g->trap_center);
        g->trap_radius = R.f32(*t, tp, "radius", false, 1.0f);
        g->trap_tube = R.f32(*t, tp, "tube", false, 0.0f);
        R.vec3(*t, tp, "box", false, g->trap_box);
        g->trap_span = R.f32(*t, tp, "span", false, 1.5f);
        g->trap_offset = R.f32(*t, tp, "offset", false, 0.0f);
    }
}

// LightState (scene_state.zig:572-623): toGpu
void light(Reader& R, const JVal& l, const std::string& p, nd_light* g, bool* visible) {
    std::memset(g, 0, sizeof(*g));
    const int kind = R.tag(l, p, "kind", true, 1, {"point", "global", "ray"});
    R.arr3(l, p, "color", true, g->color);
    g->brightness = R.f32(l, p, "brightness", true, 1.0f);
    R.range(l, p, "brightness_range", true);
    float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 0};
    R.vec3(l, p, "position", true, pos);
    R.range(l, p, "position_range_x", true); R.range(l, p, "position_range_y", true); R.range(l, p, "position_range_z", true);
    R.vec3(l, p, "direction", true, dir);
    R.range(l, p, "direction_range_x", true); R.range(l, p, "direction_range_y", true); R.range(l, p, "direction_range_z", true);
    const float spread = R.f32(l, p, "spread", false, 25.0f);
    R.range(l, p, "spread_range", false);
    const bool cast = R.boolean(l, p, "cast_shadows", true, true);
    const bool hard = R.boolean(l, p, "hard_shadows", true, false);
    g->shadow_softness = R.f32(l, p, "shadow_softness", true, 8.0f);
    R.range(l, p, "shadow_softness_range", true);
    (void)R.boolean(l, p, "window_open", true, false);
    *visible = R.boolean(l, p, "visible", false, true);
    const float* pd = kind == 1 ? dir : pos;
    set3(g->position_or_direction, pd[0], pd[1], pd[2]);
    g->light_type = (float)kind;
    g->cast_shadows = cast ? 1.0f : 0.0f;
    g->hard_shadows = hard ? 1.0f : 0.0f;
    g->spread = radians(spread);
    set3(g->ray_direction, dir[0], dir[1], dir[2]);
    g->waist = 0.05f;  // beam_waist
}

// FogEmitterState (scene_state.zig:625-660): toGpu
void fog(Reader& R, const JVal& f, const std::string& p, nd_fog_emitter* g, bool* visible) {
    std::memset(g, 0, sizeof(*g));
    R.vec3(f, p, "position", true, g->position);
    R.range(f, p, "position_range_x", true); R.range(f, p, "position_range_y", true); R.range(f, p, "position_range_z", true);
    g->radius = R.f32(f, p, "radius", true, 3.0f);
    R.range(f, p, "radius_range", true);
    g->density = R.f32(f, p, "density", true, 1.0f);
    R.range(f, p, "density_range", true);
    g->softness = R.f32(f, p, "softness", true, 0.6f);
    R.range(f, p, "softness_range", true);
    R.arr3(f, p, "color", true, g->color);
    g->anisotropy = R.f32(f, p, "anisotropy", true, 0.3f);
    R.range(f, p, "anisotropy_range", true);
    (void)R.boolean(f, p, "window_open", true, false);
    *visible = R.boolean(f, p, "visible", false, true);
}

// MarchPrecision (render_precision.zig:3-19)
struct Precision {
    float eps_c = 0.0005f, eps_f = 0.00005f, budget = 2.0f, refine_fast = 4, refine_hq = 20;
    bool verify = false;
};
Precision precision(Reader& R, const JVal* v, const std::string& p) {
    Precision q;
    if (!v) return q;
    q.eps_c = R.f32(*v, p, "epsilon_coefficient", false, q.eps_c);
    q.eps_f = R.f32(*v, p, "epsilon_floor", false, q.eps_f);
    q.budget = R.f32(*v, p, "hq_footprint_budget_px", false, q.budget);
    q.refine_fast = R.f32(*v, p, "refine_fast", false, q.refine_fast);
    q.refine_hq = R.f32(*v, p, "refine_hq", false, q.refine_hq);
    q.verify = R.boolean(*v, p, "verify_each_step", false, q.verify);
    return q;
}

void cpy_err(char* err, size_t cap, const std::string& s) {
    if (!err || cap == 0) return;
    std::snprintf(err, cap, "%s", s.c_str());
}

}  // namespace

extern "C" nd_status nd_scene_from_dream(const char* text, size_t len, uint32_t width, uint32_t height,
                                         nd_uniforms* u, nd_dream_info* info, char* err, size_t err_cap) {
    if (err && err_cap) err[0] = 0;
    if (!text || !u) { cpy_err(err, err_cap, "null scene text or output"); return ND_E_INVALID; }
    JVal root;
    std::string perr;
    if (!JParser(text, len).parse(root, perr)) { cpy_err(err, err_cap, "invalid scene file: " + perr); return ND_E_INVALID; }
    if (root.kind != JVal::OBJ) { cpy_err(err, err_cap, "invalid scene file: not an object"); return ND_E_INVALID; }
    Reader R;
    const std::string S = "scene";
    nd_dream_info di;
    std::memset(&di, 0, sizeof(di));
    const long long version = R.integer(root, S, "version", false, 12);
    if (R.ok() && (version < 0 || version > 12)) {
        cpy_err(err, err_cap, "Scene file is from a newer version of the app.");  // scene_file.zig:825-827
        return ND_E_UNSUPPORTED;
    }
    di.version = (uint32_t)version;

    // ---- camera (FreeCamera, camera.zig:64-83)
    float cpos[3] = {0, 0, 0}, cfwd[3] = {0, 0, 1}, cup[3] = {0, 1, 0};
    bool dof = false, mode_2d = false;
    float focus_distance = 3.2f, aperture = 0.05f, focus_range = 0.5f, zoom_2d = 2.0f;
    // FlatView (flat_view.zig:39-101)
    int flat_coloring = 0, flat_interior = 0;
    float fb = 16, fx = 0, fbd = 0.25f, fbo = 0, fh = 0.35f, frel = 0, fla = 135, fc = 0, fcd = 2, fss = 1, fsd = 0.5f,
          fso = 0.35f, fsa = 8;
    if (const JVal* cam = R.object(root, S, "camera", true)) {
        const std::string P = S + ".camera";
        R.vec3(*cam, P, "position", true, cpos);
        R.vec3(*cam, P, "forward", true, cfwd);
        R.vec3(*cam, P, "up", true, cup);
        dof = R.boolean(*cam, P, "dof_enabled", false, false);
        focus_distance = R.f32(*cam, P, "focus_distance", false, focus_distance);
        aperture = R.f32(*cam, P, "aperture", false, aperture);
        focus_range = R.f32(*cam, P, "focus_range", false, focus_range);
        mode_2d = R.boolean(*cam, P, "mode_2d", false, false);
        zoom_2d = R.f32(*cam, P, "zoom_2d", false, zoom_2d);
        if (const JVal* fl = R.object(*cam, P, "flat", false)) {
            const std::string F = P + ".flat";
            flat_coloring = R.tag(*fl, F, "coloring", false, 0, {"section", "escape", "distance", "orbit"});
            flat_interior = R.tag(*fl, F, "interior", false, 0, {"trap", "orbit_radius", "dark"});
            fb = R.f32(*fl, F, "bailout", false, fb);
            fx = R.f32(*fl, F, "extra_iterations", false, fx);
            fbd = R.f32(*fl, F, "band_density", false, fbd);
            fbo = R.f32(*fl, F, "band_offset", false, fbo);
            fh = R.f32(*fl, F, "halo", false, fh);
            frel = R.f32(*fl, F, "relief", false, frel);
            fla = R.f32(*fl, F, "light_angle", false, fla);
            fc = R.f32(*fl, F, "contours", false, fc);
            fcd = R.f32(*fl, F, "contour_density", false, fcd);
            fss = R.f32(*fl, F, "slab_slices", false, fss);
            fsd = R.f32(*fl, F, "slab_depth", false, fsd);
            fso = R.f32(*fl, F, "slab_opacity", false, fso);
            fsa = R.f32(*fl, F, "samples", false, fsa);
        }
    }
    zoom_2d = std::min(std::max(zoom_2d, 1e-6f), 1e4f);  // applyScene, scene_file.zig:735

    // ---- top-level settings
    (void)R.f32(root, S, "max_steps", true, 0);
    R.range(root, S, "max_steps_range", true);
    (void)R.f32(root, S, "max_dist", true, 0);
    R.range(root, S, "max_dist_range", true);
    (void)R.f32(root, S, "preview_quality", true, 0);
    R.range(root, S, "preview_quality_range", true);
    (void)R.f32(root, S, "max_reflection_bounces", true, 0);
    R.range(root, S, "max_reflection_bounces_range", true);
    di.export_width = (int32_t)R.integer(root, S, "export_width", true, 0);
    di.export_height = (int32_t)R.integer(root, S, "export_height", true, 0);
    // the export's render settings (RenderSettingsState, render_precision.zig:21-35)
    float rs_steps = 256, rs_dist = 24, rs_bounces = 1, rs_fog = 8;
    Precision prec;
    if (const JVal* rs = R.object(root, S, "render_settings", false)) {
        const std::string P = S + ".render_settings";
        rs_steps = R.f32(*rs, P, "max_steps", false, rs_steps);
        rs_dist = R.f32(*rs, P, "max_dist", false, rs_dist);
        rs_bounces = R.f32(*rs, P, "max_reflection_bounces", false, rs_bounces);
        rs_fog = R.f32(*rs, P, "fog_samples", false, rs_fog);
        prec = precision(R, R.object(*rs, P, "precision", false), P + ".precision");
    }
    // PhotonSettings (photon_state.zig:12-60): bounces reach the block, the map is the renderer's
    float photon_bounces = 2;
    bool photon_enabled = true;
    if (const JVal* ph = R.object(root, S, "photon", false)) {
        photon_bounces = R.f32(*ph, S + ".photon", "bounces", false, photon_bounces);
        photon_enabled = R.boolean(*ph, S + ".photon", "enabled", false, photon_enabled);
    }
    // McRenderState (mc_render.zig:3-28)
    bool mc_on = false, mc_adaptive = false, mc_light_pick = true, mc_fog_tracking = true, mc_cache = true;
    float mc_thr = 0.02f, mc_min = 16, mc_inner = 1;
    di.mc_max_samples = 256;
    di.mc_export_samples = 512;
    if (const JVal* mc = R.object(root, S, "mc", false)) {
        const std::string P = S + ".mc";
        mc_on = R.boolean(*mc, P, "enabled", false, mc_on);
        di.mc_max_samples = R.f32(*mc, P, "max_samples", false, di.mc_max_samples);
        di.mc_export_samples = R.f32(*mc, P, "export_samples", false, di.mc_export_samples);
        mc_adaptive = R.boolean(*mc, P, "adaptive", false, mc_adaptive);
        mc_thr = R.f32(*mc, P, "adaptive_threshold", false, mc_thr);
        mc_min = R.f32(*mc, P, "adaptive_min_samples", false, mc_min);
        mc_inner = R.f32(*mc, P, "inner_samples", false, mc_inner);
        mc_light_pick = R.boolean(*mc, P, "light_pick", false, mc_light_pick);
        mc_fog_tracking = R.boolean(*mc, P, "fog_tracking", false, mc_fog_tracking);
        mc_cache = R.boolean(*mc, P, "primary_cache", false, mc_cache);
    }
    // Approximations (approximations.zig:45-101): mask bits in declaration order
    uint32_t approx_mask = 0;
    float ap_cs = 1.0f, ap_csc = 0.3f, ap_b = 0.35f, ap_t = 1.0f, ap_ls = 2.0f, ap_lst = 1.5f;
    bool ap_fog = false, ap_caustics = false, ap_bounce = false;
    if (const JVal* ap = R.object(root, S, "approximations", false)) {
        const std::string P = S + ".approximations";
        const char* bits[] = {"simple_fog", "fake_caustics", "fake_bounce", "sky_ambient", "thin_glass", "tint",
                              "cheap_dispersion", "coarse_secondary", "distance_lod", "sky_reflections"};
        for (int b = 0; b < 10; ++b)
            if (R.boolean(*ap, P, bits[b], false, false)) approx_mask |= 1u << b;
        ap_fog = approx_mask & 1u;
        ap_caustics = approx_mask & 2u;
        ap_bounce = approx_mask & 4u;
        ap_cs = R.f32(*ap, P, "caustic_strength", false, ap_cs);
        ap_csc = R.f32(*ap, P, "caustic_scale", false, ap_csc);
        ap_b = R.f32(*ap, P, "bounce_strength", false, ap_b);
        ap_t = R.f32(*ap, P, "tint_density", false, ap_t);
        ap_ls = R.f32(*ap, P, "lod_start", false, ap_ls);
        ap_lst = R.f32(*ap, P, "lod_strength", false, ap_lst);
    }
    // SkyDto (scene_file.zig:66-93) -> SkyState.toGpu(no image loaded) (sky.zig:76-104)
    int sky_mode = 0;
    float zen[3] = {0.02f, 0.02f, 0.05f}, hor[3] = {0.04f, 0.035f, 0.07f}, gnd[3] = {0.06f, 0.05f, 0.09f};
    float falloff = 1.0f, sun_size = 3.0f, sun_int = 20.0f, sky_int = 1.0f, yaw = 0.0f;
    float sun_dir[3] = {0.4f, 0.55f, 0.3f}, sun_col[3] = {1.0f, 0.95f, 0.85f};
    bool sun_on = false, sky_photons = false;
    if (const JVal* sk = R.object(root, S, "sky", false)) {
        const std::string P = S + ".sky";
        sky_mode = R.tag(*sk, P, "mode", false, 0, {"procedural", "image"});
        R.arr3(*sk, P, "zenith", false, zen);
        R.arr3(*sk, P, "horizon", false, hor);
        R.arr3(*sk, P, "ground", false, gnd);
        falloff = R.f32(*sk, P, "falloff", false, falloff);
        sun_on = R.boolean(*sk, P, "sun_enabled", false, sun_on);
        R.vec3(*sk, P, "sun_direction", false, sun_dir);
        sun_size = R.f32(*sk, P, "sun_size", false, sun_size);
        R.arr3(*sk, P, "sun_color", false, sun_col);
        sun_int = R.f32(*sk, P, "sun_intensity", false, sun_int);
        sky_int = R.f32(*sk, P, "intensity", false, sky_int);
        yaw = R.f32(*sk, P, "yaw", false, yaw);
        sky_photons = R.boolean(*sk, P, "photons", false, sky_photons);
        di.sky_image = sky_mode == 1 ? 1u : 0u;
    }
    if (const JVal* st = R.object(root, S, "stereo", false)) di.stereo = R.boolean(*st, S + ".stereo", "enabled", false, false) ? 1u : 0u;
    if (const JVal* ss = R.array(root, S, "screen_shaders", false)) di.screen_shaders = (uint32_t)ss->a.size();
    if (const JVal* kf = R.array(root, S, "keyframes", false)) di.keyframes = (uint32_t)kf->a.size();
    if (const JVal* w = R.array(root, S, "warps", false))
        if (!w->a.empty()) R.fail(S + ".warps", "domain warps are not implemented");
    if (const JVal* ps = R.array(root, S, "particle_systems", false))
        if (!ps->a.empty()) R.fail(S + ".particle_systems", "particle systems are not implemented");

    if (!R.ok()) { cpy_err(err, err_cap, "invalid scene file: " + R.err); return ND_E_INVALID; }

    // ---- the block: reference defaults, then buildUniforms (export_image.zig:97-167)
    nd_uniforms_defaults(u);
    const uint32_t W = width ? width : (uint32_t)std::max(di.export_width, 1);
    const uint32_t H = height ? height : (uint32_t)std::max(di.export_height, 1);
    // eye basis: right = normalize(forward x up), up and forward as saved (camera.zig:91-94)
    double r[3] = {(double)cfwd[1] * cup[2] - (double)cfwd[2] * cup[1], (double)cfwd[2] * cup[0] - (double)cfwd[0] * cup[2],
                   (double)cfwd[0] * cup[1] - (double)cfwd[1] * cup[0]};
    float rf[3] = {(float)r[0], (float)r[1], (float)r[2]};
    const float rl = std::sqrt(rf[0] * rf[0] + rf[1] * rf[1] + rf[2] * rf[2]);
    for (float& x : rf) x = rl < 1e-6f ? 0.0f : x * (1.0f / rl);
    for (int a = 0; a < 3; ++a) {
        u->camera_pos[a] = cpos[a];
        u->camera_forward[a] = cfwd[a];
        u->camera_up[a] = cup[a];
        u->camera_right[a] = rf[a];
    }
    u->time = 0.0f;
    u->max_steps = rs_steps;
    u->max_dist = rs_dist;
    u->resolution[0] = (float)W;
    u->resolution[1] = (float)H;
    u->max_reflection_bounces = rs_bounces;
    u->light_bounces = photon_bounces;
    u->high_quality = 1.0f;
    u->epsilon_coefficient = prec.eps_c;
    u->epsilon_floor = prec.eps_f;
    u->hq_footprint_budget_px = prec.budget;
    u->refine_fast = prec.refine_fast;
    u->refine_hq = prec.refine_hq;
    u->march_verify = prec.verify ? 1.0f : 0.0f;
    u->dof_enabled = dof ? 1.0f : 0.0f;
    u->focus_distance = focus_distance;
    u->aperture = aperture;
    u->focus_range = focus_range;
    u->fog_samples = std::max(rs_fog, 1.0f);
    u->mode_2d = mode_2d ? 1.0f : 0.0f;
    u->slice_zoom = zoom_2d;
    u->mc_enabled = mc_on ? 1.0f : 0.0f;
    u->adapt_enabled = (mc_on && mc_adaptive) ? 1.0f : 0.0f;
    u->adapt_threshold = mc_thr;
    u->adapt_min_samples = mc_min;
    u->mc_inner_samples = std::max(mc_inner, 1.0f);
    u->mc_light_pick = mc_light_pick ? 1.0f : 0.0f;
    u->fog_tracking = mc_fog_tracking ? 1.0f : 0.0f;
    u->primary_cache_key = (mc_on && mc_cache) ? 1.0f : 0.0f;
    // FlatView.stamp (flat_view.zig:78-94)
    u->flat_coloring = (float)flat_coloring;
    u->flat_interior = (float)flat_interior;
    u->flat_bailout = fb;
    u->flat_extra_iters = std::round(std::max(fx, 0.0f));
    u->flat_band_density = fbd;
    u->flat_band_offset = fbo;
    u->flat_relief = std::max(frel, 0.0f);
    u->flat_light_angle = fla;
    u->flat_contours = std::min(std::max(fc, 0.0f), 1.0f);
    u->flat_contour_density = std::max(fcd, 0.01f);
    u->flat_slab_slices = std::round(std::min(std::max(fss, 1.0f), 32.0f));
    u->flat_slab_depth = std::max(fsd, 0.0f);
    u->flat_slab_opacity = std::min(std::max(fso, 0.01f), 1.0f);
    u->flat_samples = std::min(std::max(std::round(fsa), 1.0f), 1024.0f);
    u->flat_halo = std::max(fh, 0.0f);

    // instances
    Codes codes;
    const JVal* insts = R.array(root, S, "instances", true);
    const size_t ni = insts ? std::min(insts->a.size(), (size_t)ND_MAX_INSTANCES) : 0;
    bool transparent = false;
    for (size_t i = 0; insts && i < insts->a.size() && R.ok(); ++i) {
        const JVal& in = insts->a[i];
        if (in.kind != JVal::OBJ) { R.fail(idx(S, "instances", i), "not an object"); break; }
        if (i >= ni) {  // parsed like the others, then dropped (applyScene keeps max_instances)
            Codes spare;
            nd_fractal_instance g;
            std::string spare_images[1 + ND_MAX_MIXINS];
            instance(R, in, idx(S, "instances", i), spare, &g, spare_images);
            continue;
        }
        std::string images[1 + ND_MAX_MIXINS];
        instance(R, in, idx(S, "instances", i), codes, &u->instances[i], images);
        for (int k = 0; k < 1 + ND_MAX_MIXINS; ++k) {
            if (images[k].size() >= ND_DREAM_PATH_MAX)
                R.fail(idx(S, "instances", i) + ".formula.image", "image path longer than ND_DREAM_PATH_MAX - 1 bytes");
            else
                std::memcpy(di.image_path[4 * i + k], images[k].c_str(), images[k].size() + 1);
        }
        const nd_fractal_instance& g = u->instances[i];
        if (g.hidden <= 0.5f)
            for (int k = 0; k < (int)g.color_count; ++k)
                if (g.colors[k].transparency > 0.001f) transparent = true;  // anyTransparent, export_image.zig:169-186
    }
    u->instance_count = (float)ni;
    u->has_transparency = transparent ? 1.0f : 0.0f;
    // lights and fog emitters: the visible ones, packed (packVisible, export_image.zig:188-197)
    int nl = 0, nf = 0;
    if (const JVal* ls = R.array(root, S, "lights", true)) {
        for (size_t i = 0; i < ls->a.size() && R.ok(); ++i) {
            if (ls->a[i].kind != JVal::OBJ) { R.fail(idx(S, "lights", i), "not an object"); break; }
            nd_light g;
            bool vis = true;
            light(R, ls->a[i], idx(S, "lights", i), &g, &vis);
            if (i < (size_t)ND_MAX_LIGHTS && vis && nl < ND_MAX_LIGHTS) u->lights[nl++] = g;
        }
    }
    for (int i = nl; i < ND_MAX_LIGHTS; ++i) std::memset(&u->lights[i], 0, sizeof(nd_light));
    u->light_count = (float)nl;
    if (const JVal* fs = R.array(root, S, "fog_emitters", true)) {
        for (size_t i = 0; i < fs->a.size() && R.ok(); ++i) {
            if (fs->a[i].kind != JVal::OBJ) { R.fail(idx(S, "fog_emitters", i), "not an object"); break; }
            nd_fog_emitter g;
            bool vis = true;
            fog(R, fs->a[i], idx(S, "fog_emitters", i), &g, &vis);
            if (i < (size_t)ND_MAX_FOG_EMITTERS && vis && nf < ND_MAX_FOG_EMITTERS) u->fog_emitters[nf++] = g;
        }
    }
    for (int i = nf; i < ND_MAX_FOG_EMITTERS; ++i) std::memset(&u->fog_emitters[i], 0, sizeof(nd_fog_emitter));
    u->fog_count = (float)nf;
    if (!R.ok()) { cpy_err(err, err_cap, "invalid scene file: " + R.err); return ND_E_INVALID; }

    // renderer stamps: approximations, sky
    u->approx_flags = (float)approx_mask;
    u->approx_caustic_strength = ap_cs;
    u->approx_caustic_scale = ap_csc;
    u->approx_bounce = ap_b;
    u->approx_tint = ap_t;
    u->approx_lod_start = ap_ls;
    u->approx_lod_strength = ap_lst;
    {
        const float dl = std::sqrt(sun_dir[0] * sun_dir[0] + sun_dir[1] * sun_dir[1] + sun_dir[2] * sun_dir[2]);
        const float inv = dl > 1e-6f ? 1.0f / dl : 0.0f;
        const bool sun = sun_on && dl > 1e-6f;  // no sky image is loaded: the procedural sky
        nd_sky& k = u->sky;
        k.mode = 0.0f;
        k.intensity = std::max(sky_int, 0.0f);
        k.falloff = std::max(falloff, 0.01f);
        k.photons = sky_photons ? 1.0f : 0.0f;
        set3(k.zenith, zen[0], zen[1], zen[2]);
        k.sun_cos = sun ? std::cos(radians(std::min(std::max(sun_size, 0.01f), 89.0f))) : -2.0f;
        set3(k.horizon, hor[0], hor[1], hor[2]);
        k.sun_intensity = std::max(sun_int, 0.0f);
        set3(k.ground, gnd[0], gnd[1], gnd[2]);
        k.yaw = radians(yaw);
        set3(k.sun_color, sun_col[0], sun_col[1], sun_col[2]);
        set3(k.sun_dir, sun_dir[0] * inv, sun_dir[1] * inv, sun_dir[2] * inv);
    }

    if (info) {
        di.width = W;
        di.height = H;
        di.mc_enabled = mc_on ? 1u : 0u;
        // an export traces a photon map when photons are on and not replaced, in 3D
        di.photons_wanted = (photon_enabled && !(ap_fog && ap_caustics && ap_bounce) && !mode_2d) ? 1u : 0u;
        di.n_codes = codes.n;
        for (uint32_t c = 0; c < ND_MAX_CODES; ++c) di.code_to_builtin[c] = c < codes.n ? codes.builtin[c] : -1;
        *info = di;
    }
    return ND_OK;
}
