// hr_scene.cpp -- scene.cfg, OBJ / MTL, camera.cfg and light.cfg readers.
//
// Formats and defaults follow the reference:
//   loadSceneConfiguration        src/Application/SceneConfigurationReader.cpp:29-190
//   AnimationManager::read        src/Application/AnimationLoader.cpp:50-101
//   WFObject::read / loadWFObj    src/Application/WFObjectReader.cpp:261-636
//   readMtlLib                    src/Application/WFObjectReader.cpp:135-256
//   Material::setupPhongCoefficients src/Application/WFObjectReader.cpp:49-57
//   CameraManager::read / rotate  src/Application/CameraManager.cpp:33-66, :96-168
//   LightSourceLoader::loadFromFile src/Application/LightSourceLoader.cpp:29-100
//   SceneLoader::loadScene, insertLightSourceGeometry, createLightSources
//                                 src/Application/SceneLoader.cpp:41-332
// Differences: leading / trailing blanks around scene.cfg keys and values are
// trimmed (the reference keeps them and then fails to match or open); OBJ
// lines and points are skipped (only triangles are rendered); a material
// named by `usemtl` starts with zero emission.
#include "hr_scene.h"
#include "hr_error.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

namespace hr {
namespace scene {

namespace {

const float kPiF = 3.14159265358979323846f;

std::string trim(const std::string& s)
{
    const size_t b = s.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_last_not_of(" \t\r\n");
    return s.substr(b, e - b + 1);
}

std::string cut_comment(const std::string& s)
{
    const size_t p = s.find('#');
    return p == std::string::npos ? s : s.substr(0, p);
}

std::string dir_name(const std::string& path)
{
    const size_t p = path.find_last_of("/\\");
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

std::string resolve(const std::string& base_dir, const std::string& name)
{
    if (name.empty() || name[0] == '/' || base_dir.empty()) return name;
    std::ifstream probe(name);
    if (probe) return name;              // relative to the working directory (the reference)
    return base_dir + name;              // else relative to the scene.cfg
}

// float3 helpers with the reference's host evaluation order
void sub(const float* a, const float* b, float* r) { r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2]; }
void add(const float* a, const float* b, float* r) { r[0] = a[0] + b[0]; r[1] = a[1] + b[1]; r[2] = a[2] + b[2]; }
void cross(const float* a, const float* b, float* r)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
float dot(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void normalize(const float* a, float* r)
{
    const float s = 1.f / sqrtf(dot(a, a));
    r[0] = a[0] * s; r[1] = a[1] * s; r[2] = a[2] * s;
}
float length(const float* a) { return sqrtf(dot(a, a)); }

void skip_ws(const char*& c)
{
    while (*c && std::isspace((unsigned char)*c)) ++c;
}

// one face corner v[/t][/n]; the normal index is taken when a number follows
// the second '/' (WFObjectReader.cpp:403-476)
bool parse_corner(const char*& c, long& v, long& n)
{
    char* e = nullptr;
    v = strtol(c, &e, 10);
    if (e == c) return false;
    c = e;
    n = 0;
    skip_ws(c);
    if (*c != '/') return true;
    ++c;
    skip_ws(c);
    if (*c != '/') {
        strtol(c, &e, 10);                 // texture index (not used)
        if (e == c) return false;
        c = e;
        skip_ws(c);
    }
    if (*c != '/') return true;
    ++c;
    skip_ws(c);
    const long k = strtol(c, &e, 10);
    if (e != c) {
        n = k;
        c = e;
    }
    return true;
}

bool parse_floats(const char*& c, float* out, int n)
{
    for (int i = 0; i < n; ++i) {
        while (std::isspace((unsigned char)*c)) ++c;
        char* e = nullptr;
        const double v = strtod(c, &e);
        if (e == c) return false;
        out[i] = (float)v;
        c = e;
    }
    return true;
}

bool read_mtl(const std::string& path, Obj& o, const std::map<std::string, size_t>& map)
{
    std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
    if (!in) {
        o.error = "cannot open material library " + path;
        return false;
    }
    std::string line;
    long cur = -1;
    while (std::getline(in, line)) {
        line = trim(cut_comment(line));
        if (line.empty()) continue;
        const char* c = line.c_str();
        if (strncasecmp(c, "newmtl", 6) == 0) {
            // a material the OBJ never selects is not in the map: the line is
            // an error and the previous material stays current, as in the
            // reference (WFObjectReader.cpp:155-163)
            const std::string name = trim(std::string(c + 6));
            auto it = map.find(name);
            if (it != map.end()) cur = (long)it->second;
        } else if (cur >= 0 && (strncasecmp(c, "Kd", 2) == 0 || strncasecmp(c, "Ks", 2) == 0 ||
                                strncasecmp(c, "Ka", 2) == 0 || strncasecmp(c, "Ke", 2) == 0)) {
            const char t = c[1];           // the case of the second letter matters (:188-198)
            c += 2;
            float col[3];
            if (!parse_floats(c, col, 3)) continue;
            Mat& m = o.mats[(size_t)cur];
            float* dst = t == 'd' ? m.kd : t == 's' ? m.ks : t == 'a' ? m.ka : t == 'e' ? m.ke : nullptr;
            if (dst) std::memcpy(dst, col, sizeof col);
        } else if (cur >= 0 && strncasecmp(c, "Ns", 2) == 0) {
            c += 2;
            float v;
            if (parse_floats(c, &v, 1)) o.mats[(size_t)cur].ns = v;
        } else if (cur >= 0 && strncasecmp(c, "Ni", 2) == 0) {
            c += 2;
            float v;
            if (parse_floats(c, &v, 1)) o.mats[(size_t)cur].ni = v;
        } else if (cur >= 0 && strncasecmp(c, "illum", 5) == 0) {
            const long v = strtol(c + 5, nullptr, 10);
            if (v > 3) o.mats[(size_t)cur].refractive = true;
        }
    }
    return true;
}

// AreaLightSource::create (src/RT/Primitive/LightSource.hpp:83-95)
Light make_light(const float* v0, const float* v1, const float* v2, const float* v3, const float* inten,
                 const float* normal)
{
    (void)v2;
    Light l;
    std::memcpy(l.position, v0, 3 * sizeof(float));
    std::memcpy(l.normal, normal, 3 * sizeof(float));
    sub(v1, v0, l.edge1);
    sub(v3, v0, l.edge2);
    std::memcpy(l.intensity, inten, 3 * sizeof(float));
    return l;
}

float light_weight(const Light& l)
{
    float c[3];
    cross(l.edge1, l.edge2, c);
    return length(c) * length(l.intensity);
}

// SceneLoader::insertLightSourceGeometry (SceneLoader.cpp:41-92)
void insert_light_geometry(const Light& l, Obj& o)
{
    for (const Mat& m : o.mats)
        if (m.ke[0] + m.ke[1] + m.ke[2] > 0.f) return;
    Mat m;
    m.name = "lightsource";
    std::memcpy(m.ke, l.intensity, sizeof m.ke);
    o.mats.push_back(m);
    const uint32_t mi = (uint32_t)o.mats.size() - 1u;
    const uint32_t ni = (uint32_t)(o.normals.size() / 3);
    o.normals.insert(o.normals.end(), l.normal, l.normal + 3);
    float p[4][3];
    std::memcpy(p[0], l.position, sizeof p[0]);
    add(p[0], l.edge1, p[1]);
    add(p[0], l.edge2, p[2]);
    add(p[1], l.edge2, p[3]);
    const uint32_t vb = (uint32_t)(o.verts.size() / 3);
    const int order[4] = { 0, 1, 3, 2 };      // pt1, pt2, pt4, pt3
    for (int k : order) o.verts.insert(o.verts.end(), p[k], p[k] + 3);
    Face f1{ { vb, vb + 1, vb + 2 }, { ni, ni, ni }, mi };
    Face f2{ { vb, vb + 2, vb + 3 }, { ni, ni, ni }, mi };
    o.faces.push_back(f1);
    o.faces.push_back(f2);
}

// SceneLoader::createLightSources (SceneLoader.cpp:94-210)
void create_lights(Obj& o, std::vector<Light>& lights)
{
    const uint32_t nv = (uint32_t)(o.verts.size() / 3);
    const uint32_t nm = (uint32_t)o.mats.size();
    uint32_t ids[4] = { nv, nv, nv, nv };
    uint32_t material = nm;
    float emission[3] = { 0.f, 0.f, 0.f };
    auto vtx = [&](uint32_t i) { return &o.verts[3 * (size_t)i]; };
    for (const Face& f : o.faces) {
        const Mat& m = o.mats[f.material];
        if (m.ke[0] + m.ke[1] + m.ke[2] > 0 && ids[0] == nv) {
            ids[0] = f.v[0]; ids[1] = f.v[1]; ids[2] = f.v[2];
            material = f.material;
            std::memcpy(emission, m.ke, sizeof emission);
        } else if (f.material == material && ids[3] == nv) {
            uint32_t matches = 0, first[4] = { 4, 4, 4, 4 }, self[4] = { 4, 4, 4, 4 };
            for (uint32_t i = 0; i < 3; ++i) {
                for (uint32_t k = 0; k < 3; ++k) {
                    if (f.v[k] == ids[i]) {
                        self[matches] = k;
                        first[matches] = i;
                        ++matches;
                        break;
                    }
                }
            }
            if (matches != 2) continue;
            const uint32_t non_matched = 3u - first[0] - first[1];
            const uint32_t last = 3u - self[0] - self[1];
            if (last > 2) {             // no valid fourth vertex: drop the pair
                ids[0] = ids[1] = ids[2] = ids[3] = nv;
                material = nm;
                emission[0] = emission[1] = emission[2] = 0.f;
                continue;
            }
            ids[3] = f.v[last];
            const float* v1 = vtx(ids[non_matched]);
            const float* v2 = vtx(ids[(non_matched + 2) % 3]);
            const float* v3 = vtx(ids[3]);
            const float* v4 = vtx(ids[(non_matched + 1) % 3]);
            const float* normal = &o.normals[3 * (size_t)f.n[0]];
            lights.push_back(make_light(v1, v2, v3, v4, emission, normal));
            ids[0] = ids[1] = ids[2] = ids[3] = nv;
            material = nm;
            emission[0] = emission[1] = emission[2] = 0.f;
        }
    }
    if (!lights.empty() || nv == 0) return;
    // default light above the scene
    float lo[3] = { 3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f };
    float hi[3] = { -3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f };
    for (uint32_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], vtx(i)[k]);
            hi[k] = std::max(hi[k], vtx(i)[k]);
        }
    float d[3];
    sub(hi, lo, d);
    const float md = std::max(d[0], std::max(d[1], d[2]));
    for (int k = 0; k < 3; ++k) {
        lo[k] += md * 0.025f;
        hi[k] -= md * 0.025f;
    }
    const float top = std::max(hi[0], std::max(hi[1], hi[2]));
    const float normal[3] = { 0.f, 0.f, -1.f };
    const float one[3] = { 1.f, 1.f, 1.f };
    const float v1[3] = { lo[0], lo[1], top }, v2[3] = { lo[0], hi[1], top };
    const float v3[3] = { hi[0], lo[1], top }, v4[3] = { hi[0], hi[1], top };
    const Light l = make_light(v1, v2, v3, v4, one, normal);
    lights.push_back(l);
    insert_light_geometry(l, o);
}

}  // namespace

std::string frame_file_name(const std::string& prefix, const std::string& suffix, uint32_t num_frames,
                            uint32_t i)
{
    int digits = 0;
    for (int n = (int)num_frames - 1; n > 0; n /= 10) ++digits;
    int cur = 1;
    for (int k = (int)i; k > 9; k /= 10) ++cur;
    std::string id;
    for (; cur < digits; ++cur) id += "0";
    id += std::to_string(i);
    return prefix + id + suffix;
}

bool read_config(const std::string& path, Config& cfg, std::string& err)
{
    std::ifstream in(path.c_str());
    if (!in) {
        err = "cannot open scene configuration " + path;
        return false;
    }
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t cr = line.find('\r');
        if (cr != std::string::npos) line = line.substr(0, cr);
        line = cut_comment(line);
        if (trim(line).empty()) continue;
        const size_t eq = line.find('=');
        const std::string key = trim(line.substr(0, eq));
        const std::string val = eq == std::string::npos ? std::string() : trim(line.substr(eq + 1));
        if (key == "obj" || key == "scene") cfg.obj_files.push_back(val);
        else if (key == "camera" || key == "view" || key == "cam") cfg.camera_file = val;
        else if (key == "camera path" || key == "camerapath" || key == "camera_path") cfg.camera_path_file = val;
        else if (key == "frame prefix" || key == "frameprefix" || key == "framename") cfg.frame_prefix = val;
        else if (key == "frame suffix" || key == "framesuffix") cfg.frame_suffix = val;
        else if (key == "frame number" || key == "number" || key == "framenumber" || key == "number of frames") {
            char* e = nullptr;
            const long v = strtol(val.c_str(), &e, 10);
            if (e == val.c_str() || v < 0) {
                err = path + ":" + std::to_string(lineno) + ": invalid number of frames";
                return false;
            }
            cfg.num_frames = (uint32_t)v;
        } else if (key == "frame step" || key == "framestep" || key == "step size") {
            char* e = nullptr;
            const double v = strtod(val.c_str(), &e);
            if (e == val.c_str()) {
                err = path + ":" + std::to_string(lineno) + ": invalid frame step";
                return false;
            }
            cfg.frame_step = (float)v;
        } else if (key == "light sources" || key == "light source" || key == "light" || key == "lights")
            cfg.lights_file = val;
        else if (key == "instances") cfg.instances_file = val;
        else cfg.warnings.push_back("skipping unrecognized token '" + key + "' at line " + std::to_string(lineno));
    }
    return true;
}

bool read_obj(const std::string& path, Obj& o)
{
    std::ifstream in(path.c_str(), std::ios::in);
    if (!in) {
        o.error = "cannot open " + path;
        return false;
    }
    o = Obj();
    Mat def;
    def.name = "Default";
    def.kd[0] = def.kd[1] = def.kd[2] = .6774f;
    def.ka[0] = def.ka[1] = def.ka[2] = 0.2f;
    o.mats.push_back(def);
    std::map<std::string, size_t> map;
    map["Default"] = 0;
    std::vector<std::string> libs;
    uint32_t cur_mat = 0;
    const uint32_t kNone = 0xFFFFFFFFu;
    std::string line;
    size_t lineno = 0;
    std::vector<uint32_t> cv, cn;
    size_t obj_begin = 0;
    o.obj_first.push_back(0);
    while (std::getline(in, line)) {
        ++lineno;
        line = cut_comment(line);
        const char* c = line.c_str();
        while (std::isspace((unsigned char)*c)) ++c;
        if (!*c) continue;
        const char c0 = (char)std::tolower((unsigned char)c[0]);
        if (c0 == 'v') {
            // v x y z | vn x y z | vt ...: exactly three coordinates, or the
            // line is skipped as malformed (WFObjectReader.cpp:339-380)
            const char c1 = (char)std::tolower((unsigned char)c[1]);
            if (c1 == 't') continue;
            const bool normal = c1 == 'n';
            if (!normal && !std::isspace((unsigned char)c1)) {
                ++o.skipped_lines;
                continue;
            }
            c += normal ? 2 : 1;
            float v[4];
            int k = 0;
            bool bad = false;
            for (;;) {
                skip_ws(c);
                if (!*c) break;
                char* e = nullptr;
                const double x = strtod(c, &e);
                if (e == c || k == 4) { bad = true; break; }
                v[k++] = (float)x;
                c = e;
            }
            if (bad || k != 3) {
                ++o.skipped_lines;
                continue;
            }
            (normal ? o.normals : o.verts).insert((normal ? o.normals : o.verts).end(), v, v + 3);
        } else if (c0 == 'f' || c0 == 'l') {
            // corners v, v/t, v//n, v/t/n; polygons become triangle fans
            // (c0, ck, ck+1); one or two corners (points, lines) are not
            // rendered (WFObjectReader.cpp:382-505)
            ++c;
            if (std::tolower((unsigned char)*c) == 'o') ++c;
            cv.clear();
            cn.clear();
            bool bad = false;
            for (;;) {
                skip_ws(c);
                if (!*c) break;
                long vi = 0, ni = 0;
                if (!parse_corner(c, vi, ni)) { bad = true; break; }
                if (vi <= 0) {
                    o.error = path + ":" + std::to_string(lineno) + ": relative or zero vertex index";
                    return false;
                }
                cv.push_back((uint32_t)(vi - 1));
                cn.push_back(ni > 0 ? (uint32_t)(ni - 1) : kNone);
            }
            if (bad || cv.empty()) {
                ++o.skipped_lines;
                continue;
            }
            if (cv.size() < 3) continue;
            for (size_t k = 1; k + 1 < cv.size(); ++k) {
                Face f{ { cv[0], cv[k], cv[k + 1] }, { cn[0], cn[k], cn[k + 1] }, cur_mat };
                o.faces.push_back(f);
            }
        } else if (c0 == 'o' && (c[1] == 0 || std::isspace((unsigned char)c[1]))) {
            // a new object starts; an empty one is not recorded (WFObjectReader.cpp:516-530)
            if (obj_begin < o.faces.size()) o.obj_first.push_back((uint32_t)o.faces.size());
            obj_begin = o.faces.size();
        } else if (strncasecmp(c, "usemtl", 6) == 0) {
            const std::string name = trim(std::string(c + 6));
            if (name.empty()) {
                ++o.skipped_lines;
                continue;
            }
            auto it = map.find(name);
            if (it == map.end()) {
                Mat m;
                m.name = name;
                o.mats.push_back(m);
                it = map.emplace(name, o.mats.size() - 1).first;
            }
            cur_mat = (uint32_t)it->second;
        } else if (strncasecmp(c, "mtllib", 6) == 0) {
            const std::string name = trim(std::string(c + 6));
            if (!name.empty()) libs.push_back(name);
        }
        // g, s and unknown entities carry nothing the renderer uses
    }
    // the last object always closes (WFObjectReader.cpp:583)
    o.obj_first.push_back((uint32_t)o.faces.size());
    const uint32_t nv = (uint32_t)(o.verts.size() / 3);
    for (const Face& f : o.faces)
        for (int k = 0; k < 3; ++k)
            if (f.v[k] >= nv) {
                o.error = path + ": face vertex index out of range";
                return false;
            }
    const std::string dir = dir_name(path);
    for (const std::string& lib : libs)
        if (!read_mtl(dir + lib, o, map)) return false;
    // missing normals -> face normal (WFObjectReader.cpp:597-614)
    const uint32_t nn = (uint32_t)(o.normals.size() / 3);
    for (Face& f : o.faces) {
        if (nn == 0) f.n[0] = f.n[1] = f.n[2] = kNone;
        for (int k = 0; k < 3; ++k)
            if (f.n[k] != kNone && f.n[k] >= nn) f.n[k] = kNone;
        if (f.n[0] == kNone || f.n[1] == kNone || f.n[2] == kNone) {
            float e1[3], e2[3], c3[3], n[3];
            sub(&o.verts[3 * (size_t)f.v[1]], &o.verts[3 * (size_t)f.v[0]], e1);
            sub(&o.verts[3 * (size_t)f.v[2]], &o.verts[3 * (size_t)f.v[0]], e2);
            cross(e1, e2, c3);
            normalize(c3, n);
            const uint32_t id = (uint32_t)(o.normals.size() / 3);
            for (int k = 0; k < 3; ++k)
                if (f.n[k] == kNone) f.n[k] = id;
            o.normals.insert(o.normals.end(), n, n + 3);
        }
    }
    // setupPhongCoefficients (WFObjectReader.cpp:49-57)
    for (Mat& m : o.mats) {
        for (int k = 0; k < 3; ++k) {
            const float kd = m.kd[k];
            const float ks = std::min(m.ks[k], 0.99f - std::min(0.99f, kd));
            m.ks[k] = ks * (m.ns + 2) * 0.5f / kPiF;
            m.kd[k] = std::min(kd, 0.99f) / kPiF;
        }
    }
    return true;
}

std::vector<float> phong_table(const Obj& o)
{
    std::vector<float> t;
    t.reserve(o.mats.size() * 12);
    for (const Mat& m : o.mats) {
        const float row[12] = { m.kd[0], m.kd[1], m.kd[2], m.ni + (m.refractive ? 10.f : 0.f),
                                m.ks[0], m.ks[1], m.ks[2], m.ns, m.ke[0], m.ke[1], m.ke[2], 0.f };
        t.insert(t.end(), row, row + 12);
    }
    return t;
}

bool read_camera(const std::string& path, Camera& cam, std::string& err)
{
    std::ifstream in(path.c_str());
    if (!in) {
        err = "cannot open camera file " + path;
        return false;
    }
    // CameraManager defaults (src/Application/CameraManager.hpp:54-75)
    cam = Camera();
    cam.dir[2] = 1.f;
    cam.up[1] = -1.f;
    cam.fov = 66.f;
    cam.res_x = cam.res_y = 512;
    std::string line, key;
    while (std::getline(in, line)) {
        if (line.find('#') != std::string::npos) continue;
        std::stringstream ss(line);
        key.clear();
        ss >> key;
        if (key == "position") ss >> cam.pos[0] >> cam.pos[1] >> cam.pos[2];
        else if (key == "orientation") ss >> cam.dir[0] >> cam.dir[1] >> cam.dir[2];
        else if (key == "up") ss >> cam.up[0] >> cam.up[1] >> cam.up[2];
        else if (key == "rotation") ss >> cam.rotation[0] >> cam.rotation[1] >> cam.rotation[2];
        else if (key == "resX") ss >> cam.res_x;
        else if (key == "resY") ss >> cam.res_y;
        else if (key == "FOV") ss >> cam.fov;
    }
    // CameraManager::rotate (CameraManager.cpp:33-66), written as the reference has it
    const float ax = cam.rotation[0] * kPiF / 180.f, ay = cam.rotation[1] * kPiF / 180.f,
                az = cam.rotation[2] * kPiF / 180.f;
    const float sx = sinf(ax), cx = cosf(ax), sy = sinf(ay), cy = cosf(ay), sz = sinf(az), cz = cosf(az);
    const float c1[3] = { cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz };
    const float c2[3] = { cy * sz, sx * sy * sz + cx * cz, sx * sy * sz - sx * cz };
    const float c3[3] = { -sy, sx * cy, cx * cy };
    float d[3], u[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = c1[k] * cam.dir[0] + c2[k] * cam.dir[1] + c3[k] * cam.dir[2];
        u[k] = c1[k] * cam.up[0] + c2[k] * cam.up[1] + c3[k] * cam.up[2];
    }
    std::memcpy(cam.dir, d, sizeof d);
    std::memcpy(cam.up, u, sizeof u);
    if (cam.res_x <= 0 || cam.res_y <= 0) {
        err = path + ": invalid resolution";
        return false;
    }
    cam.present = true;
    return true;
}

bool read_lights(const std::string& path, std::vector<Light>& out, std::string& err)
{
    std::ifstream in(path.c_str());
    if (!in) {
        err = "cannot open light file " + path;
        return false;
    }
    Light cur;
    std::memset(&cur, 0, sizeof cur);
    bool hp = false, hn = false, hi = false, h1 = false, h2 = false;
    std::string line, key;
    // the key of a line without one stays that of the line before, as in the
    // reference (a blank line repeats the previous key's flag, not its values)
    while (std::getline(in, line)) {
        line = cut_comment(line);
        std::stringstream ss(line);
        ss >> key;
        if (key == "position") { ss >> cur.position[0] >> cur.position[1] >> cur.position[2]; hp = true; }
        else if (key == "normal") { ss >> cur.normal[0] >> cur.normal[1] >> cur.normal[2]; hn = true; }
        else if (key == "intensity") { ss >> cur.intensity[0] >> cur.intensity[1] >> cur.intensity[2]; hi = true; }
        else if (key == "edge1") { ss >> cur.edge1[0] >> cur.edge1[1] >> cur.edge1[2]; h1 = true; }
        else if (key == "edge2") { ss >> cur.edge2[0] >> cur.edge2[1] >> cur.edge2[2]; h2 = true; }
        if (hp && hi && h1 && h2) {
            if (!hn) {
                float c[3];
                cross(cur.edge1, cur.edge2, c);
                normalize(c, cur.normal);
            }
            out.push_back(cur);
            hp = hn = hi = h1 = h2 = false;
        }
    }
    return true;
}

bool read_instances(const std::string& path, std::vector<Instance>& out, std::string& err)
{
    std::ifstream in(path.c_str());
    if (!in) {
        err = "cannot open instances file " + path;
        return false;
    }
    out.clear();
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t cr = line.find('\r');
        if (cr != std::string::npos) line = line.substr(0, cr);
        line = cut_comment(line);
        std::stringstream ss(line);
        std::string key;
        if (!(ss >> key)) continue;
        const std::string where = path + ":" + std::to_string(lineno);
        if (key == "obj_id") {
            // a new instance: identity transform, box unbounded until AABB lines
            long long id = -1;
            if (!(ss >> id) || id < 0 || id > 0xFFFFFFFFll) {
                err = where + ": invalid object id";
                return false;
            }
            Instance n;
            n.object = (uint32_t)id;
            out.push_back(n);
            continue;
        }
        int row = -1;
        if (key == "m_row_0") row = 0;
        else if (key == "m_row_1") row = 1;
        else if (key == "m_row_2") row = 2;
        const bool box_lo = key == "AABB_min", box_hi = key == "AABB_max";
        if (row < 0 && !box_lo && !box_hi) continue;   // unknown keys are ignored
        if (out.empty()) {
            err = where + ": " + key + " before the first obj_id";
            return false;
        }
        Instance& I = out.back();
        float v[4];
        const int nval = row >= 0 ? 4 : 3;
        for (int k = 0; k < nval; ++k)
            if (!(ss >> v[k])) {
                err = where + ": " + key + " needs " + std::to_string(nval) + " numbers";
                return false;
            }
        if (row >= 0)
            for (int k = 0; k < 4; ++k) I.m[row][k] = v[k];
        else
            for (int k = 0; k < 3; ++k) (box_lo ? I.lo : I.hi)[k] = v[k];
    }
    return true;
}

bool load_scene(const std::string& cfg_path, Scene& sc)
{
    Config cfg;
    if (!read_config(cfg_path, cfg, sc.error)) return false;
    const std::string base = dir_name(cfg_path);
    const bool anim = !cfg.frame_prefix.empty() && !cfg.frame_suffix.empty() && cfg.num_frames != 0;
    if (anim) {
        for (uint32_t i = 0; i < cfg.num_frames; ++i) {
            Obj o;
            const std::string f = resolve(base, frame_file_name(cfg.frame_prefix, cfg.frame_suffix,
                                                                cfg.num_frames, i));
            if (!read_obj(f, o)) {
                sc.error = o.error;
                return false;
            }
            sc.frames.push_back(std::move(o));
            sc.interpolatable.push_back(1);
        }
        sc.frame_step = cfg.frame_step;
    }
    for (const std::string& f : cfg.obj_files) {
        Obj o;
        if (!read_obj(resolve(base, f), o)) {
            sc.error = o.error;
            return false;
        }
        sc.frames.push_back(std::move(o));
        sc.interpolatable.push_back(0);
    }
    if (sc.frames.empty()) {
        sc.error = "no scene specified in " + cfg_path;
        return false;
    }
    for (size_t i = 1; i < sc.frames.size(); ++i)
        if (sc.frames[i].verts.size() != sc.frames[0].verts.size() ||
            sc.frames[i].faces.size() != sc.frames[0].faces.size() ||
            sc.frames[i].normals.size() != sc.frames[0].normals.size()) {
            sc.error = "key frames differ in vertex, normal or triangle count";
            return false;
        }
    if (!cfg.camera_file.empty() && !read_camera(resolve(base, cfg.camera_file), sc.camera, sc.error))
        return false;
    if (!cfg.lights_file.empty()) {
        if (!read_lights(resolve(base, cfg.lights_file), sc.lights, sc.error)) return false;
        for (const Light& l : sc.lights)
            for (Obj& o : sc.frames) insert_light_geometry(l, o);
    } else {
        create_lights(sc.frames[0], sc.lights);
        // the default light's quad is added to frame 0 only by the reference;
        // keep every key frame the same size
        for (size_t i = 1; i < sc.frames.size(); ++i)
            if (sc.frames[i].faces.size() < sc.frames[0].faces.size())
                for (const Light& l : sc.lights) insert_light_geometry(l, sc.frames[i]);
    }
    for (const Light& l : sc.lights) sc.light_weights.push_back(light_weight(l));
    if (!cfg.instances_file.empty()) {
        if (!read_instances(resolve(base, cfg.instances_file), sc.instances, sc.error)) return false;
        const size_t nobj = sc.frames[0].obj_first.size() - 1;
        for (const Instance& I : sc.instances)
            if (I.object >= nobj) {
                sc.error = "instance of object " + std::to_string(I.object) + " but the scene has " +
                           std::to_string(nobj) + " objects";
                return false;
            }
    }
    return true;
}

}  // namespace scene
}  // namespace hr

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

struct hr_scene {
    hr::scene::Scene s;
};

namespace {

using hr::set_error;

hr_status finish_load(hr_scene* sc, hr_scene** out)
{
    const hr::scene::Scene& s = sc->s;
    const size_t nv = s.frames[0].verts.size() / 3, nt = s.frames[0].faces.size();
    if (nv > 0xFFFFFFFFull || nt > 0x3FFFFFFFull) {
        delete sc;
        return set_error(HR_ERR_CAPACITY, "scene too large");
    }
    *out = sc;
    return HR_OK;
}

hr_status soa(const std::vector<float>& aos, float* xyz)
{
    const size_t n = aos.size() / 3;
    for (size_t i = 0; i < n; ++i) {
        xyz[i] = aos[3 * i];
        xyz[n + i] = aos[3 * i + 1];
        xyz[2 * n + i] = aos[3 * i + 2];
    }
    return HR_OK;
}

hr_status check_scene(const hr_scene* sc)
{
    if (!sc) return set_error(HR_ERR_ARG, "null scene");
    return HR_OK;
}

}  // namespace

extern "C" {

hr_status hr_scene_load(const char* cfg_path, hr_scene** out)
{
    if (!cfg_path || !out) return set_error(HR_ERR_ARG, "null argument");
    *out = nullptr;
    hr_scene* sc = new hr_scene();
    if (!hr::scene::load_scene(cfg_path, sc->s)) {
        const std::string e = sc->s.error;
        delete sc;
        return set_error(e.rfind("cannot open", 0) == 0 ? HR_ERR_IO : HR_ERR_PARSE, e);
    }
    return finish_load(sc, out);
}

hr_status hr_scene_load_obj(const char* obj_path, hr_scene** out)
{
    if (!obj_path || !out) return set_error(HR_ERR_ARG, "null argument");
    *out = nullptr;
    hr_scene* sc = new hr_scene();
    hr::scene::Obj o;
    if (!hr::scene::read_obj(obj_path, o)) {
        const std::string e = o.error;
        delete sc;
        return set_error(e.rfind("cannot open", 0) == 0 ? HR_ERR_IO : HR_ERR_PARSE, e);
    }
    sc->s.frames.push_back(std::move(o));
    sc->s.interpolatable.push_back(0);
    hr::scene::create_lights(sc->s.frames[0], sc->s.lights);
    for (const hr::scene::Light& l : sc->s.lights) sc->s.light_weights.push_back(hr::scene::light_weight(l));
    return finish_load(sc, out);
}

hr_status hr_scene_free(hr_scene* sc)
{
    delete sc;
    return HR_OK;
}

hr_status hr_scene_get_info(const hr_scene* sc, hr_scene_info* out)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!out) return set_error(HR_ERR_ARG, "null output pointer");
    const hr::scene::Scene& s = sc->s;
    const hr::scene::Obj& f = s.frames[0];
    std::memset(out, 0, sizeof *out);
    out->num_frames = (uint32_t)s.frames.size();
    out->num_vertices = (uint32_t)(f.verts.size() / 3);
    out->num_normals = (uint32_t)(f.normals.size() / 3);
    out->num_triangles = (uint32_t)f.faces.size();
    out->num_materials = (uint32_t)f.mats.size();
    out->num_lights = (uint32_t)s.lights.size();
    out->frame_step = s.frame_step;
    out->has_camera = s.camera.present ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
        out->cam_pos[k] = s.camera.pos[k];
        out->cam_dir[k] = s.camera.dir[k];
        out->cam_up[k] = s.camera.up[k];
    }
    out->cam_fov = s.camera.fov;
    out->cam_res_x = s.camera.res_x;
    out->cam_res_y = s.camera.res_y;
    out->num_objects = (uint32_t)(f.obj_first.size() - 1);
    out->num_instances = (uint32_t)s.instances.size();
    return HR_OK;
}

hr_status hr_scene_get_instances(const hr_scene* sc, uint32_t* obj_first, hr_instance* inst)
{
    if (hr_status st = check_scene(sc)) return st;
    const hr::scene::Scene& s = sc->s;
    if (obj_first)
        std::copy(s.frames[0].obj_first.begin(), s.frames[0].obj_first.end(), obj_first);
    if (inst)
        for (size_t i = 0; i < s.instances.size(); ++i) {
            const hr::scene::Instance& I = s.instances[i];
            inst[i].object = I.object;
            for (int r = 0; r < 3; ++r) {
                for (int k = 0; k < 4; ++k) inst[i].m[r][k] = I.m[r][k];
                inst[i].aabb_min[r] = I.lo[r];
                inst[i].aabb_max[r] = I.hi[r];
            }
        }
    return HR_OK;
}

hr_status hr_scene_get_vertices(const hr_scene* sc, uint32_t frame, float* xyz)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!xyz) return set_error(HR_ERR_ARG, "null output pointer");
    if (frame >= sc->s.frames.size()) return set_error(HR_ERR_ARG, "key frame index out of range");
    return soa(sc->s.frames[frame].verts, xyz);
}

hr_status hr_scene_get_normals(const hr_scene* sc, uint32_t frame, float* xyz)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!xyz) return set_error(HR_ERR_ARG, "null output pointer");
    if (frame >= sc->s.frames.size()) return set_error(HR_ERR_ARG, "key frame index out of range");
    return soa(sc->s.frames[frame].normals, xyz);
}

hr_status hr_scene_get_triangles(const hr_scene* sc, uint32_t* idx, uint32_t* nidx, uint32_t* mat_idx)
{
    if (hr_status st = check_scene(sc)) return st;
    const std::vector<hr::scene::Face>& faces = sc->s.frames[0].faces;
    for (size_t i = 0; i < faces.size(); ++i) {
        for (int k = 0; k < 3; ++k) {
            if (idx) idx[3 * i + k] = faces[i].v[k];
            if (nidx) nidx[3 * i + k] = faces[i].n[k];
        }
        if (mat_idx) mat_idx[i] = faces[i].material;
    }
    return HR_OK;
}

hr_status hr_scene_get_materials(const hr_scene* sc, float* mats)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!mats) return set_error(HR_ERR_ARG, "null output pointer");
    const std::vector<float> t = hr::scene::phong_table(sc->s.frames[0]);
    std::memcpy(mats, t.data(), t.size() * sizeof(float));
    return HR_OK;
}

hr_status hr_scene_get_lights(const hr_scene* sc, hr_area_light* lights, float* weights)
{
    if (hr_status st = check_scene(sc)) return st;
    for (size_t i = 0; i < sc->s.lights.size(); ++i) {
        const hr::scene::Light& l = sc->s.lights[i];
        if (lights) {
            std::memcpy(lights[i].position, l.position, sizeof l.position);
            std::memcpy(lights[i].normal, l.normal, sizeof l.normal);
            std::memcpy(lights[i].intensity, l.intensity, sizeof l.intensity);
            std::memcpy(lights[i].edge1, l.edge1, sizeof l.edge1);
            std::memcpy(lights[i].edge2, l.edge2, sizeof l.edge2);
        }
        if (weights) weights[i] = sc->s.light_weights[i];
    }
    return HR_OK;
}

hr_status hr_scene_upload(hr_ctx* ctx, const hr_scene* sc)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!ctx) return set_error(HR_ERR_ARG, "null context");
    const hr::scene::Scene& s = sc->s;
    const uint32_t nf = (uint32_t)s.frames.size();
    const uint32_t nv = (uint32_t)(s.frames[0].verts.size() / 3);
    const uint32_t nn = (uint32_t)(s.frames[0].normals.size() / 3);
    const uint32_t nt = (uint32_t)s.frames[0].faces.size();
    if (nv == 0) return set_error(HR_ERR_ARG, "scene has no vertices");
    std::vector<float> keys((size_t)nf * 3 * nv);
    for (uint32_t f = 0; f < nf; ++f) soa(s.frames[f].verts, keys.data() + (size_t)f * 3 * nv);
    std::vector<uint32_t> idx(3 * (size_t)nt), nidx(3 * (size_t)nt), mat_idx(nt);
    hr_scene_get_triangles(sc, idx.data(), nidx.data(), mat_idx.data());
    if (hr_status st = hr_set_keyframes(ctx, keys.data(), nf, nv, idx.data(), nt)) return st;
    if (hr_status st = hr_set_frame(ctx, 0, 0, 0.f)) return st;
    if (nn > 0) {
        std::vector<float> nrm((size_t)nf * 3 * nn);
        for (uint32_t f = 0; f < nf; ++f) soa(s.frames[f].normals, nrm.data() + (size_t)f * 3 * nn);
        if (hr_status st = hr_set_normals(ctx, nrm.data(), nf, nn, nidx.data(), nt)) return st;
    }
    const std::vector<float> mats = hr::scene::phong_table(s.frames[0]);
    if (hr_status st = hr_upload_materials(ctx, mats.data(), (uint32_t)(mats.size() / 12), mat_idx.data(), nt))
        return st;
    std::vector<hr_area_light> lights(s.lights.size());
    hr_scene_get_lights(sc, lights.data(), nullptr);
    if (hr_status st = hr_set_lights(ctx, lights.data(), s.light_weights.data(), (int)lights.size())) return st;
    if (s.camera.present) {
        const hr::scene::Camera& c = s.camera;
        if (hr_status st = hr_set_camera(ctx, c.pos, c.dir, c.up, c.fov, c.res_x, c.res_y)) return st;
    }
    if (!s.instances.empty()) {
        std::vector<hr_instance> inst(s.instances.size());
        hr_scene_get_instances(sc, nullptr, inst.data());
        const std::vector<uint32_t>& of = s.frames[0].obj_first;
        if (hr_status st = hr_set_instances(ctx, of.data(), (uint32_t)(of.size() - 1), inst.data(),
                                            (uint32_t)inst.size()))
            return st;
    }
    return HR_OK;
}

hr_status hr_scene_frame(const hr_scene* sc, float current, uint32_t* k1, uint32_t* k2, float* coeff)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!k1 || !k2 || !coeff) return set_error(HR_ERR_ARG, "null output pointer");
    const size_t n = sc->s.frames.size();
    if (!(current >= 0.f) || (size_t)current >= n) return set_error(HR_ERR_ARG, "animation time out of range");
    const size_t f = (size_t)current;
    const size_t next = (f + 1u + (size_t)sc->s.frame_step) % n;
    *k1 = (uint32_t)f;
    *k2 = (uint32_t)(sc->s.interpolatable[next] && sc->s.interpolatable[f] ? next : f);
    *coeff = current - (float)((size_t)current);
    return HR_OK;
}

hr_status hr_scene_next_frame(const hr_scene* sc, float* current)
{
    if (hr_status st = check_scene(sc)) return st;
    if (!current) return set_error(HR_ERR_ARG, "null argument");
    const size_t n = sc->s.frames.size();
    if (!(*current >= 0.f) || (size_t)*current >= n) return set_error(HR_ERR_ARG, "animation time out of range");
    *current += sc->s.interpolatable[(size_t)*current] ? sc->s.frame_step : 1.f;
    if ((size_t)*current >= n) *current -= (float)n;
    return HR_OK;
}

}  // extern "C"
