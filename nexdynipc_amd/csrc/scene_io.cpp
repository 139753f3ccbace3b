// Scene and state I/O of the C ABI (include/ndipc_io.h): the reference's JSON scene
// format and OBJ meshes (nexdynipc/src/App/SceneLoader.cpp), the settings
// Simulation::run applies (nexdynipc/src/App/Simulation.cpp:14-56) and StateExporter's
// CSV (nexdynipc/src/App/StateExporter.cpp). Host code only.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <variant>
#include <vector>

#include "../../include/ndipc_io.h"

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return -1;
}

// ------------------------------------------------------------------ JSON
// A small recursive-descent reader of RFC 8259 JSON (objects keep their key order;
// lookups take the last duplicate key, as nlohmann::json's object does).
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0.0;
    bool is_int = false;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const std::string& k) const {
        if (kind != Obj) return nullptr;
        const JVal* r = nullptr;
        for (auto& kv : obj)
            if (kv.first == k) r = &kv.second;
        return r;
    }
    bool has(const std::string& k) const { return get(k) != nullptr; }
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    explicit JParser(const std::string& src) : s(src) {}

    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool bad(const std::string& m) {
        if (err.empty()) err = m + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return bad(std::string("expected ") + w);
        i += n;
        return true;
    }
    static void utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) {
            o += char(cp);
        } else if (cp < 0x800) {
            o += char(0xC0 | (cp >> 6));
            o += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += char(0xE0 | (cp >> 12));
            o += char(0x80 | ((cp >> 6) & 0x3F));
            o += char(0x80 | (cp & 0x3F));
        } else {
            o += char(0xF0 | (cp >> 18));
            o += char(0x80 | ((cp >> 12) & 0x3F));
            o += char(0x80 | ((cp >> 6) & 0x3F));
            o += char(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(unsigned& v) {
        if (i + 4 > s.size()) return bad("short \\u escape");
        v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= unsigned(c - '0');
            else if (c >= 'a' && c <= 'f') v |= unsigned(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= unsigned(c - 'A' + 10);
            else return bad("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& o) {
        if (i >= s.size() || s[i] != '"') return bad("expected string");
        ++i;
        o.clear();
        while (i < s.size()) {
            const char c = s[i++];
            if (c == '"') return true;
            if (static_cast<unsigned char>(c) < 0x20) return bad("control character in string");
            if (c != '\\') {
                o += c;
                continue;
            }
            if (i >= s.size()) break;
            const char e = s[i++];
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    unsigned cp = 0;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                        i += 2;
                        unsigned lo = 0;
                        if (!hex4(lo)) return false;
                        if (lo < 0xDC00 || lo > 0xDFFF) return bad("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(o, cp);
                    break;
                }
                default: return bad("bad escape");
            }
        }
        return bad("unterminated string");
    }
    bool number(JVal& v) {
        const size_t b = i;
        if (i < s.size() && s[i] == '-') ++i;
        if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return bad("bad number");
        if (s[i] == '0') ++i;
        else while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
        bool integral = true;
        if (i < s.size() && s[i] == '.') {
            integral = false;
            ++i;
            if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return bad("bad fraction");
            while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            integral = false;
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return bad("bad exponent");
            while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
        }
        const std::string t = s.substr(b, i - b);
        errno = 0;
        v.kind = JVal::Num;
        v.num = std::strtod(t.c_str(), nullptr);  // correctly rounded
        v.is_int = integral;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 256) return bad("nesting too deep");
        ws();
        if (i >= s.size()) return bad("unexpected end");
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = JVal::Obj;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return bad("expected ':'");
                ++i;
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.obj.emplace_back(std::move(k), std::move(x));
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; return true; }
                return bad("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i;
            v.kind = JVal::Arr;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return true; }
            for (;;) {
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.arr.push_back(std::move(x));
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; return true; }
                return bad("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return string(v.str);
        }
        if (c == 't') { v.kind = JVal::Bool; v.b = true; return lit("true"); }
        if (c == 'f') { v.kind = JVal::Bool; v.b = false; return lit("false"); }
        if (c == 'n') { v.kind = JVal::Null; return lit("null"); }
        return number(v);
    }
    bool document(JVal& v) {
        if (!value(v, 0)) return false;
        ws();
        if (i != s.size()) return bad("trailing characters");
        return true;
    }
};

// Canonical text of a value (keys sorted, numbers %.17g): equal values, equal text.
void dump(const JVal& v, std::string& o) {
    char buf[40];
    switch (v.kind) {
        case JVal::Null: o += "null"; break;
        case JVal::Bool: o += v.b ? "true" : "false"; break;
        case JVal::Num: std::snprintf(buf, sizeof buf, "%.17g", v.num); o += buf; break;
        case JVal::Str: o += '"'; o += v.str; o += '"'; break;
        case JVal::Arr:
            o += '[';
            for (size_t k = 0; k < v.arr.size(); ++k) {
                if (k) o += ',';
                dump(v.arr[k], o);
            }
            o += ']';
            break;
        case JVal::Obj: {
            std::map<std::string, const JVal*> m;
            for (auto& kv : v.obj) m[kv.first] = &kv.second;
            o += '{';
            bool first = true;
            for (auto& kv : m) {
                if (!first) o += ',';
                first = false;
                o += '"';
                o += kv.first;
                o += "\":";
                dump(*kv.second, o);
            }
            o += '}';
            break;
        }
    }
}

// typed reads with the reference's defaults (nlohmann::json::value)
double num_or(const JVal& o, const char* k, double d) {
    const JVal* v = o.get(k);
    return (v && v->kind == JVal::Num) ? v->num : d;
}
int int_or(const JVal& o, const char* k, int d) {
    const JVal* v = o.get(k);
    return (v && v->kind == JVal::Num) ? int(v->num) : d;
}
bool bool_or(const JVal& o, const char* k, bool d) {
    const JVal* v = o.get(k);
    return (v && v->kind == JVal::Bool) ? v->b : d;
}
std::string str_or(const JVal& o, const char* k, const std::string& d) {
    const JVal* v = o.get(k);
    return (v && v->kind == JVal::Str) ? v->str : d;
}
bool vec3(const JVal* v, double out[3]) {
    if (!v || v->kind != JVal::Arr || v->arr.size() < 3) return false;
    for (int k = 0; k < 3; ++k) {
        if (v->arr[k].kind != JVal::Num) return false;
        out[k] = v->arr[k].num;
    }
    return true;
}

// ------------------------------------------------------------------ OBJ
// SceneLoader.cpp:40-109: vertices "v x y z"; faces "f" with 1-based indices ("a/b/c"
// tokens read up to the first '/'), fan-triangulated; a line whose tokens do not all
// parse as integers is dropped, and so is a triangle with a negative index.
bool read_obj(const std::string& path, std::vector<double>& V, std::vector<int>& F) {
    std::ifstream in(path);
    if (!in.is_open()) return false;
    V.clear();
    F.clear();
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string tag;
        iss >> tag;
        if (tag == "v") {
            double x = 0.0, y = 0.0, z = 0.0;
            iss >> x >> y >> z;
            V.push_back(x);
            V.push_back(y);
            V.push_back(z);
        } else if (tag == "f") {
            std::vector<int> idx;
            std::string tok;
            bool ok = true;
            while (iss >> tok) {
                const size_t sl = tok.find('/');
                const std::string t = sl == std::string::npos ? tok : tok.substr(0, sl);
                char* end = nullptr;
                errno = 0;
                const long v = t.empty() ? 0 : std::strtol(t.c_str(), &end, 10);
                if (t.empty()) {
                    idx.push_back(-1);
                    continue;
                }
                if (end == t.c_str() || errno) {
                    ok = false;
                    break;
                }
                idx.push_back(int(v) - 1);
            }
            if (!ok || idx.size() < 3) continue;
            for (size_t k = 1; k + 1 < idx.size(); ++k) {
                const int a = idx[0], b = idx[k], c = idx[k + 1];
                if (a < 0 || b < 0 || c < 0) continue;
                F.push_back(a);
                F.push_back(b);
                F.push_back(c);
            }
        }
    }
    if (V.empty() || F.empty()) return false;
    const int nv = int(V.size() / 3);
    for (int f : F)
        if (f >= nv) return false;  // the reference would index out of bounds
    return true;
}

}  // namespace

struct ndipc_scene_file {
    ndipc_sim_settings set{};
    std::vector<int32_t> body_off{0, 0}, body_id, vert_off, n_verts, face_off, n_faces;
    std::vector<double> verts, mass, inertia;
    std::vector<int32_t> faces;
    std::vector<int32_t> joint_off{0, 0}, joint_type, joint_body;
    std::vector<double> joint_data, joint_mu;
    std::vector<int32_t> allowed_off{0, 0}, allowed_pairs;
    std::vector<double> allowed_mu;
    std::vector<double> p;
    std::vector<std::string> names;
    std::vector<std::string> unsupported;
    ndipc_params params{};
};

struct ndipc_exporter {
    std::string csv, kpi;
};

namespace {

void copy_str(char* dst, size_t cap, const std::string& s) {
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

void default_settings(ndipc_sim_settings& s) {
    // SimulationConfig defaults (Simulation.h:19-53)
    s = ndipc_sim_settings{};
    s.dt = 0.01;
    s.max_time = 1.0;
    s.newmark_beta = 0.25;
    s.newmark_gamma = 0.5;
    s.joint_stiffness = 1000.0;
    s.alm_max_iters = 6;
    s.alm_constraint_tolerance = 1e-4;
    s.alm_dual_tolerance = 1e-3;
    s.alm_hardening_trigger = 2.0;
    s.alm_hardening_ratio = 2.0;
    s.kpi_gate_enabled = 0;
    s.kpi_v_w_max = 0.2;
    s.kpi_t_sat_max = 0.01;
    s.kpi_r_dual_max = 1e-3;
    s.contact_enabled = 1;
    s.contact_thickness = 1e-3;
    s.contact_stiffness = -1.0;
    s.contact_friction = 0.0;
    copy_str(s.integrator_type, sizeof s.integrator_type, "ImplicitNewmark");
    copy_str(s.output_dir, sizeof s.output_dir, "output");
    copy_str(s.output_name, sizeof s.output_name, "simulation_results");
}

struct JointEntry {
    int type;  // 0 fixed, 1 hinge
    int body_a, body_b;  // body ids
    double data[16];
    double stiffness;
};

// add_joint_from_entry (SceneLoader.cpp:144-457): fixed and hinge joints are kept; other
// types are recorded as entries the device path does not run.
int read_joint(const JVal& j, std::unordered_set<std::string>& sigs, std::vector<JointEntry>& out,
               std::vector<std::string>& unsupported) {
    if (j.kind != JVal::Obj || !j.has("type")) return 0;
    // make_joint_signature: the entry with "revolute" as "hinge" and without "stiffness"
    JVal norm = j;
    std::string type = str_or(j, "type", "");
    if (type == "revolute") type = "hinge";
    for (auto& kv : norm.obj)
        if (kv.first == "type" && kv.second.kind == JVal::Str && kv.second.str == "revolute") kv.second.str = "hinge";
    norm.obj.erase(std::remove_if(norm.obj.begin(), norm.obj.end(),
                                  [](const std::pair<std::string, JVal>& kv) { return kv.first == "stiffness"; }),
                   norm.obj.end());
    std::string sig;
    dump(norm, sig);
    if (sigs.count(sig)) return 0;  // duplicated entry, skipped
    JointEntry e{};
    e.stiffness = num_or(j, "stiffness", 1e8);  // Joint.h:44 default
    if (type == "hinge") {
        double aa[3], ab[3], xa[3], xb[3];
        if (!vec3(j.get("anchor_a"), aa) || !vec3(j.get("anchor_b"), ab) || !vec3(j.get("axis_a"), xa) ||
            !vec3(j.get("axis_b"), xb) || !j.has("body_a") || !j.has("body_b"))
            return fail("hinge joint: missing body_a / body_b / anchors / axes");
        e.type = 1;
        e.body_a = int_or(j, "body_a", -1);
        e.body_b = int_or(j, "body_b", -1);
        for (int k = 0; k < 3; ++k) {
            e.data[k] = aa[k];
            e.data[3 + k] = ab[k];
            e.data[6 + k] = xa[k];
            e.data[9 + k] = xb[k];
        }
    } else if (type == "fixed") {
        double pos[3];
        if (!j.has("body") || !vec3(j.get("position"), pos)) return fail("fixed joint: missing body / position");
        e.type = 0;
        e.body_a = int_or(j, "body", -1);
        e.body_b = -1;
        for (int k = 0; k < 3; ++k) e.data[k] = pos[k];
        e.data[3] = 1.0;  // identity (w, x, y, z)
        if (const JVal* o = j.get("orientation")) {
            if (o->kind != JVal::Arr || o->arr.size() < 4) return fail("fixed joint: orientation needs 4 numbers");
            // [x, y, z, w] in the file, Eigen::Quaterniond(w, x, y, z) (SceneLoader.cpp:261-266)
            e.data[3] = o->arr[3].num;
            e.data[4] = o->arr[0].num;
            e.data[5] = o->arr[1].num;
            e.data[6] = o->arr[2].num;
        }
    } else {
        unsupported.push_back("joint type '" + type + "'");
        sigs.insert(sig);
        return 0;
    }
    sigs.insert(sig);
    out.push_back(e);
    return 0;
}

}  // namespace

extern "C" {

const char* ndipc_io_last_error(void) { return g_err.c_str(); }

int32_t ndipc_obj_load(const char* path, int32_t* nv, double* verts, int32_t v_cap, int32_t* nf, int32_t* faces,
                       int32_t f_cap) {
    if (!path || !nv || !nf) return fail("obj_load: null argument");
    std::vector<double> V;
    std::vector<int> F;
    if (!read_obj(path, V, F)) return fail(std::string("obj_load: no valid mesh in ") + path);
    *nv = int32_t(V.size() / 3);
    *nf = int32_t(F.size() / 3);
    if (verts && v_cap >= *nv) std::copy(V.begin(), V.end(), verts);
    if (faces && f_cap >= *nf) std::copy(F.begin(), F.end(), faces);
    return 0;
}

int32_t ndipc_scene_load(const char* json_path, ndipc_scene_file** out) {
    namespace fs = std::filesystem;
    if (!json_path || !out) return fail("scene_load: null argument");
    *out = nullptr;
    const std::string fn = json_path;
    if (fn.size() < 5 || fn.substr(fn.size() - 5) != ".json")
        return fail("scene_load: only .json scenes are read (the reference's built-in scenes are not data)");
    std::ifstream f(fn, std::ios::binary);
    if (!f.is_open()) return fail("scene_load: cannot open " + fn);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    JVal scene;
    JParser P(text);
    if (!P.document(scene)) return fail("scene_load: " + fn + ": " + P.err);
    if (scene.kind != JVal::Obj) return fail("scene_load: the scene is not a JSON object");

    auto sf = std::make_unique<ndipc_scene_file>();
    default_settings(sf->set);
    const fs::path scene_dir = fs::path(fn).has_parent_path() ? fs::path(fn).parent_path() : fs::current_path();

    // bodies (SceneLoader.cpp:472-514)
    if (const JVal* bodies = scene.get("bodies")) {
        if (bodies->kind != JVal::Arr) return fail("scene_load: 'bodies' is not an array");
        for (const JVal& b : bodies->arr) {
            if (b.kind != JVal::Obj) return fail("scene_load: a body is not an object");
            sf->body_id.push_back(int_or(b, "id", -1));
            sf->names.push_back(str_or(b, "name", "Body"));
            sf->mass.push_back(num_or(b, "mass", 1.0));
            double pos[3] = {0.0, 0.0, 0.0};
            if (b.has("position") && !vec3(b.get("position"), pos)) return fail("scene_load: bad body position");
            sf->p.insert(sf->p.end(), pos, pos + 3);
            double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            if (const JVal* in = b.get("inertia")) {
                if (in->kind != JVal::Arr || in->arr.size() < 3) return fail("scene_load: inertia needs 3 rows");
                for (int r = 0; r < 3; ++r) {
                    double row[3];
                    if (!vec3(&in->arr[r], row)) return fail("scene_load: inertia rows need 3 numbers");
                    for (int c = 0; c < 3; ++c) I[3 * r + c] = row[c];
                }
            }
            sf->inertia.insert(sf->inertia.end(), I, I + 9);
            sf->vert_off.push_back(int32_t(sf->verts.size() / 3));
            sf->face_off.push_back(int32_t(sf->faces.size() / 3));
            int32_t nv = 0, nfc = 0;
            if (const JVal* m = b.get("mesh"); m && m->kind == JVal::Str) {
                const fs::path rel = m->str;
                fs::path mp = rel;
                if (rel.is_relative()) {
                    const fs::path from_scene = scene_dir / rel;
                    mp = fs::exists(from_scene) ? from_scene : fs::current_path() / rel;
                }
                std::vector<double> V;
                std::vector<int> F;
                if (read_obj(mp.string(), V, F)) {
                    nv = int32_t(V.size() / 3);
                    nfc = int32_t(F.size() / 3);
                    sf->verts.insert(sf->verts.end(), V.begin(), V.end());
                    sf->faces.insert(sf->faces.end(), F.begin(), F.end());
                } else {
                    sf->unsupported.push_back("body " + std::to_string(sf->body_id.back()) + ": mesh '" + mp.string() +
                                              "' could not be read (the body has no MeshShape)");
                }
            }
            sf->n_verts.push_back(nv);
            sf->n_faces.push_back(nfc);
        }
    }
    const int N = int(sf->body_id.size());
    sf->body_off = {0, N};

    // joints, then constraints (SceneLoader.cpp:516-528)
    std::unordered_set<std::string> sigs;
    std::vector<JointEntry> joints;
    for (const char* key : {"joints", "constraints"}) {
        if (const JVal* arr = scene.get(key)) {
            if (arr->kind != JVal::Arr) return fail(std::string("scene_load: '") + key + "' is not an array");
            for (const JVal& j : arr->arr)
                if (read_joint(j, sigs, joints, sf->unsupported) < 0) return -1;
        }
    }

    // settings (SceneLoader.cpp:622-644)
    ndipc_sim_settings& st = sf->set;
    const std::string stem = fs::path(fn).stem().string();
    const std::string model_name = str_or(scene, "case_name", stem);
    if (const JVal* s = scene.get("settings"); s && s->kind == JVal::Obj) {
        st.dt = num_or(*s, "dt", st.dt);
        st.max_time = num_or(*s, "max_time", st.max_time);
        st.newmark_gamma = num_or(*s, "newmark_gamma", st.newmark_gamma);
        st.newmark_beta = num_or(*s, "newmark_beta", st.newmark_beta);
        st.joint_stiffness = num_or(*s, "joint_stiffness", st.joint_stiffness);
        if (s->has("integrator_type"))
            copy_str(st.integrator_type, sizeof st.integrator_type, str_or(*s, "integrator_type", ""));
        if (s->has("output_name")) copy_str(st.output_name, sizeof st.output_name, str_or(*s, "output_name", ""));
        if (s->has("output_dir")) copy_str(st.output_dir, sizeof st.output_dir, str_or(*s, "output_dir", ""));
        st.alm_max_iters = int_or(*s, "alm_max_iters", st.alm_max_iters);
        st.alm_constraint_tolerance = num_or(*s, "alm_constraint_tolerance", st.alm_constraint_tolerance);
        st.alm_dual_tolerance = num_or(*s, "alm_dual_tolerance", st.alm_dual_tolerance);
        st.alm_hardening_trigger = num_or(*s, "alm_hardening_trigger", st.alm_hardening_trigger);
        st.alm_hardening_ratio = num_or(*s, "alm_hardening_ratio", st.alm_hardening_ratio);
        st.kpi_gate_enabled = bool_or(*s, "kpi_gate_enabled", st.kpi_gate_enabled != 0) ? 1 : 0;
        st.kpi_v_w_max = num_or(*s, "kpi_v_w_max", st.kpi_v_w_max);
        st.kpi_t_sat_max = num_or(*s, "kpi_t_sat_max", st.kpi_t_sat_max);
        st.kpi_r_dual_max = num_or(*s, "kpi_r_dual_max", st.kpi_r_dual_max);
    }

    // contact interactions (SceneLoader.cpp:646-698)
    std::vector<int32_t> pairs;
    std::vector<double> pair_mu;
    if (const JVal* it = scene.get("interactions"); it && it->kind == JVal::Obj) {
        if (const JVal* c = it->get("contact"); c && c->kind == JVal::Obj) {
            st.contact_enabled = 1;
            if (const JVal* gp = c->get("global_params"); gp && gp->kind == JVal::Obj) {
                st.contact_thickness = num_or(*gp, "default_contact_thickness", st.contact_thickness);
                st.contact_stiffness = num_or(*gp, "min_contact_stiffness", st.contact_stiffness);
            }
            std::unordered_map<std::string, int> name_to_id;
            for (int b = 0; b < N; ++b) name_to_id[sf->names[b]] = sf->body_id[b];
            auto resolve = [&](const JVal& v) -> int {  // resolve_body_id (SceneLoader.cpp:111-124)
                if (v.kind == JVal::Num && v.is_int) return int(v.num);
                if (v.kind == JVal::Str) {
                    auto f2 = name_to_id.find(v.str);
                    if (f2 != name_to_id.end()) return f2->second;
                }
                return -1;
            };
            double sum_mu = 0.0;
            int count_mu = 0;
            if (const JVal* ps = c->get("pairs"); ps && ps->kind == JVal::Arr) {
                for (const JVal& pr : ps->arr) {
                    if (pr.kind != JVal::Obj || !pr.has("object1") || !pr.has("object2")) continue;
                    const int a = resolve(*pr.get("object1")), b = resolve(*pr.get("object2"));
                    const bool enabled = bool_or(pr, "enabled", true);
                    const double fr = num_or(pr, "friction", -1.0);
                    if (a < 0 || b < 0) continue;
                    ++st.n_contact_pairs;
                    if (enabled && fr >= 0.0) {
                        sum_mu += fr;
                        ++count_mu;
                    }
                    if (!enabled) continue;  // Simulation.cpp:48-54 adds the enabled pairs only
                    pairs.push_back(a);
                    pairs.push_back(b);
                    pair_mu.push_back(fr);
                }
            }
            if (count_mu > 0) st.contact_friction = sum_mu / double(count_mu);
        }
    }
    if (std::string(st.output_name) == "simulation_results") copy_str(st.output_name, sizeof st.output_name, model_name);
    if (std::string(st.output_dir) == "output")
        copy_str(st.output_dir, sizeof st.output_dir, "output/" + model_name);

    // joints as upload arrays: bodies by id -> world index (the last body with the id, an
    // unknown id -> 0: the reference's unordered_map lookups, IPCSolver.cpp:966-1066);
    // stiffness left at the default 1e8 -> settings.joint_stiffness (Simulation.cpp:20-24)
    std::unordered_map<int, int> id_to_idx;
    for (int b = 0; b < N; ++b) id_to_idx[sf->body_id[b]] = b;
    auto idx = [&](int id) {
        auto it2 = id_to_idx.find(id);
        return it2 == id_to_idx.end() ? 0 : it2->second;
    };
    for (const JointEntry& e : joints) {
        if (N == 0) return fail("scene_load: joints in a scene without bodies");
        sf->joint_type.push_back(e.type);
        sf->joint_body.push_back(idx(e.body_a));
        sf->joint_body.push_back(e.type == 1 ? idx(e.body_b) : -1);
        sf->joint_data.insert(sf->joint_data.end(), e.data, e.data + 16);
        sf->joint_mu.push_back(e.stiffness == 1e8 ? st.joint_stiffness : e.stiffness);
    }
    sf->joint_off = {0, int32_t(sf->joint_type.size())};
    sf->allowed_pairs = pairs;
    sf->allowed_mu = pair_mu;
    sf->allowed_off = {0, int32_t(pair_mu.size())};

    // IPCSolver settings of Simulation::run (Simulation.cpp:25-56)
    ndipc_params& q = sf->params;
    const bool fixed_k = st.contact_stiffness > 0.0;
    q.dhat = std::max(1e-9, st.contact_thickness);
    q.kappa = fixed_k ? st.contact_stiffness : 0.0;
    q.friction_mu = st.contact_friction;
    q.enable_friction = st.contact_friction > 0.0 ? 1 : 0;
    q.enable_ccd = 1;
    q.enable_adaptive_barrier = (st.contact_enabled && !fixed_k) ? 1 : 0;
    q.alm_max_iters = std::max(1, st.alm_max_iters);
    q.alm_constraint_tol = std::max(1e-12, st.alm_constraint_tolerance);
    q.alm_dual_tol = std::max(1e-12, st.alm_dual_tolerance);
    q.alm_contact_tol = 1e-3;
    q.alm_hardening_trigger = std::max(1.0, st.alm_hardening_trigger);
    q.alm_hardening_ratio = std::max(1.0, st.alm_hardening_ratio);
    q.ccd_safety_factor = 0.8;
    q.newton_tol = 1e-6;
    q.newton_max_iters = 20;
    *out = sf.release();
    return 0;
}

void ndipc_scene_free(ndipc_scene_file* f) { delete f; }

int32_t ndipc_scene_settings(const ndipc_scene_file* f, ndipc_sim_settings* out) {
    if (!f || !out) return fail("scene_settings: null argument");
    *out = f->set;
    return 0;
}

int32_t ndipc_scene_describe(const ndipc_scene_file* f, ndipc_scene_desc* d) {
    if (!f || !d) return fail("scene_describe: null argument");
    *d = ndipc_scene_desc{};
    d->n_scenes = 1;
    d->n_bodies = int32_t(f->body_id.size());
    d->scene_body_off = f->body_off.data();
    d->body_id = f->body_id.data();
    d->vert_off = f->vert_off.data();
    d->n_verts = f->n_verts.data();
    d->face_off = f->face_off.data();
    d->n_faces = f->n_faces.data();
    d->verts = f->verts.data();
    d->faces = f->faces.data();
    d->mass = f->mass.data();
    d->inertia = f->inertia.data();
    d->n_joints = int32_t(f->joint_type.size());
    d->scene_joint_off = f->joint_off.data();
    d->joint_type = f->joint_type.data();
    d->joint_body = f->joint_body.data();
    d->joint_data = f->joint_data.data();
    d->joint_mu = f->joint_mu.data();
    d->n_allowed = int32_t(f->allowed_mu.size());
    d->scene_allowed_off = f->allowed_off.data();
    d->allowed_pairs = f->allowed_pairs.data();
    d->allowed_mu = f->allowed_mu.data();
    return 0;
}

int32_t ndipc_scene_params(const ndipc_scene_file* f, ndipc_params* out) {
    if (!f || !out) return fail("scene_params: null argument");
    *out = f->params;
    return 0;
}

int32_t ndipc_scene_state(const ndipc_scene_file* f, double* p, double* q, double* v, double* w) {
    if (!f) return fail("scene_state: null argument");
    const size_t n = f->body_id.size();
    for (size_t b = 0; b < n; ++b) {
        if (p) for (int k = 0; k < 3; ++k) p[3 * b + k] = f->p[3 * b + k];
        if (q) { q[4 * b] = 1.0; q[4 * b + 1] = q[4 * b + 2] = q[4 * b + 3] = 0.0; }
        if (v) for (int k = 0; k < 3; ++k) v[3 * b + k] = 0.0;
        if (w) for (int k = 0; k < 3; ++k) w[3 * b + k] = 0.0;
    }
    return 0;
}

int32_t ndipc_scene_body_name(const ndipc_scene_file* f, int32_t body, char* out, int32_t cap) {
    if (!f || !out || cap <= 0) return fail("scene_body_name: null argument");
    if (body < 0 || body >= int32_t(f->names.size())) return fail("scene_body_name: no such body");
    copy_str(out, size_t(cap), f->names[body]);
    return 0;
}

int32_t ndipc_scene_unsupported(const ndipc_scene_file* f, int32_t k, char* out, int32_t cap) {
    if (!f) return fail("scene_unsupported: null argument");
    if (out && cap > 0) {
        if (k < 0 || k >= int32_t(f->unsupported.size())) return fail("scene_unsupported: no such entry");
        copy_str(out, size_t(cap), f->unsupported[k]);
    }
    return int32_t(f->unsupported.size());
}

int32_t ndipc_export_open(const char* dir, const char* name, ndipc_exporter** out) {
    if (!dir || !name || !out) return fail("export_open: null argument");
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) return fail(std::string("export_open: cannot create ") + dir + ": " + ec.message());
    auto* x = new ndipc_exporter;
    x->csv = std::string(dir) + "/" + name + ".csv";
    x->kpi = std::string(dir) + "/" + name + "_kpi.csv";
    *out = x;
    return 0;
}

int32_t ndipc_export_frame(ndipc_exporter* x, const ndipc_scene_file* f, int32_t frame, double time, const double* p,
                           const double* q, const double* v, const double* w, const double* al, const double* aa) {
    if (!x || !f || !p || !q || !v || !w || !al || !aa) return fail("export_frame: null argument");
    std::FILE* fp = std::fopen(x->csv.c_str(), frame == 0 ? "w" : "a");
    if (!fp) return fail("export_frame: cannot open " + x->csv);
    if (frame == 0)
        std::fputs("frame_idx,time,body_id,body_name,pos_x,pos_y,pos_z,qx,qy,qz,qw,vx,vy,vz,ax,ay,az,wx,wy,wz,"
                   "alphax,alphay,alphaz\n", fp);
    for (size_t b = 0; b < f->body_id.size(); ++b) {
        const double* P = p + 3 * b;
        const double* Q = q + 4 * b;  // w, x, y, z
        const double* V = v + 3 * b;
        const double* A = al + 3 * b;
        const double* W = w + 3 * b;
        const double* B = aa + 3 * b;
        std::fprintf(fp,
                     "%d,%.17g,%d,%s,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,"
                     "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n",
                     frame, time, f->body_id[b], f->names[b].c_str(), P[0], P[1], P[2], Q[1], Q[2], Q[3], Q[0], V[0],
                     V[1], V[2], A[0], A[1], A[2], W[0], W[1], W[2], B[0], B[1], B[2]);
    }
    const bool ok = std::fclose(fp) == 0;
    return ok ? 0 : fail("export_frame: write failed");
}

int32_t ndipc_export_kpi(ndipc_exporter* x, int32_t frame, double time, double v_w, double t_sat, double r_dual,
                         double max_constraint_violation) {
    if (!x) return fail("export_kpi: null argument");
    std::FILE* fp = std::fopen(x->kpi.c_str(), frame == 0 ? "w" : "a");
    if (!fp) return fail("export_kpi: cannot open " + x->kpi);
    if (frame == 0) std::fputs("frame_idx,time,V_w,T_sat,R_dual,max_constraint_violation\n", fp);
    std::fprintf(fp, "%d,%.17g,%.17g,%.17g,%.17g,%.17g\n", frame, time, v_w, t_sat, r_dual, max_constraint_violation);
    const bool ok = std::fclose(fp) == 0;
    return ok ? 0 : fail("export_kpi: write failed");
}

void ndipc_export_close(ndipc_exporter* x) { delete x; }

}  // extern "C"
