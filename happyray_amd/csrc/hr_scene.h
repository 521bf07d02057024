// hr_scene.h -- host-side scene files: scene.cfg, Wavefront OBJ / MTL key
// frames, camera.cfg and light.cfg, in the formats the reference reads.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hr {
namespace scene {

struct Mat {                  // WFObject::Material (src/Application/WFObject.hpp:52-90)
    std::string name;
    float kd[3] = { 0.f, 0.f, 0.f };
    float ks[3] = { 0.f, 0.f, 0.f };
    float ka[3] = { 0.f, 0.f, 0.f };
    float ke[3] = { 0.f, 0.f, 0.f };
    float ns = 1.f;
    float ni = 1.f;
    bool refractive = false;
};

struct Face {                 // WFObject::Face: vertex and normal index per corner
    uint32_t v[3];
    uint32_t n[3];
    uint32_t material;
};

struct Obj {                  // WFObject (triangles only)
    std::vector<float> verts;      // xyz per vertex
    std::vector<float> normals;    // xyz per normal
    std::vector<Face> faces;
    std::vector<Mat> mats;
    std::vector<uint32_t> obj_first;   // object k: faces [obj_first[k], obj_first[k+1]) ('o' lines)
    size_t skipped_lines = 0;      // malformed lines, skipped as the reference does
    std::string error;
};

struct Light {                // AreaLightSource
    float position[3], normal[3], intensity[3], edge1[3], edge2[3];
};

struct Instance {             // WFObject::Instance (src/Application/WFObject.hpp:136-146)
    uint32_t object = 0;
    float m[3][4] = { { 1.f, 0.f, 0.f, 0.f }, { 0.f, 1.f, 0.f, 0.f }, { 0.f, 0.f, 1.f, 0.f } };
    float lo[3] = { -3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f };
    float hi[3] = { 3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f };
};

struct Camera {               // CameraManager (src/Application/CameraManager.cpp:96-168)
    float pos[3] = { 0.f, 0.f, 0.f };
    float dir[3] = { 0.f, 0.f, -1.f };
    float up[3] = { 0.f, 1.f, 0.f };
    float rotation[3] = { 0.f, 0.f, 0.f };
    float fov = 60.f;
    int res_x = 720, res_y = 480;
    bool present = false;
};

struct Config {               // SceneConfiguration (src/Application/SceneConfiguration.hpp:51-64)
    std::vector<std::string> obj_files;
    std::string camera_file, camera_path_file, lights_file;
    std::string frame_prefix, frame_suffix, instances_file;
    uint32_t num_frames = 0;
    float frame_step = 0.8f;
    std::vector<std::string> warnings;
};

struct Scene {
    std::vector<Obj> frames;       // key frames (same topology)
    std::vector<uint8_t> interpolatable;   // per key frame (AnimationManager::mInterpolatable)
    float frame_step = 0.8f;
    Camera camera;
    std::vector<Light> lights;
    std::vector<float> light_weights;
    std::vector<Instance> instances;  // the scene.cfg's instances file
    std::string error;
};

bool read_config(const std::string& path, Config& out, std::string& err);
bool read_obj(const std::string& path, Obj& out);
bool read_camera(const std::string& path, Camera& out, std::string& err);
bool read_lights(const std::string& path, std::vector<Light>& out, std::string& err);
// WFObject::loadInstances (src/Application/WFObjectReader.cpp:762-824)
bool read_instances(const std::string& path, std::vector<Instance>& out, std::string& err);
// SceneLoader::loadScene (src/Application/SceneLoader.cpp:212-332)
bool load_scene(const std::string& cfg_path, Scene& out);
// PhongMaterial rows (12 floats) after setupPhongCoefficients
std::vector<float> phong_table(const Obj& o);
// the file name of key frame i (src/Application/AnimationLoader.cpp:49-92)
std::string frame_file_name(const std::string& prefix, const std::string& suffix, uint32_t num_frames,
                            uint32_t i);

}  // namespace scene
}  // namespace hr
