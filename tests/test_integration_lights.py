"""The reference-side lights binding of the integration notes compiles against
include/happyray.h and turns an AreaLightSourceCollection into hr_area_light
records and weights as SceneLoader::createLightSources fills it
(src/Application/SceneLoader.cpp:94-210: weight = area x |intensity|,
src/RT/Primitive/LightSource.hpp:38-99 field order)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# float3 and AreaLightSource as the reference declares them
# (src/RT/Primitive/LightSource.hpp:38-99), with the two helpers getArea uses
PRELUDE = r"""
#include <cmath>
#include <cstdio>
#include <vector>
#include "happyray.h"
struct float3 { float x, y, z; };
static float3 make_float3(float x, float y, float z) { float3 r = { x, y, z }; return r; }
static float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
static float3 operator%(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static float len(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
struct AreaLightSource {
    float3 position, normal, intensity, edge1, edge2;
    int e1MaxDimension, e2MaxDimension;
    float getArea() const { return len(edge1 % edge2); }
    void create(const float3& v0, const float3& v1, const float3& v2, const float3& v3,
                const float3& inten, const float3& n)
    {
        (void)v2;
        position = v0; normal = n; edge1 = v1 - v0; edge2 = v3 - v0; intensity = inten;
    }
};
"""

MAIN = r"""
int main()
{
    AreaLightSourceCollection L;
    AreaLightSource a, b;
    a.create(make_float3(0, 4, 0), make_float3(2, 4, 0), make_float3(2, 4, 1), make_float3(0, 4, 1),
             make_float3(5, 5, 5), make_float3(0, -1, 0));
    b.create(make_float3(-1, 0, 3), make_float3(-1, 3, 3), make_float3(0, 3, 3), make_float3(0, 0, 3),
             make_float3(1, 2, 2), make_float3(0, 0, -1));
    L.upload(a.getArea() * len(a.intensity), a);     // SceneLoader.cpp:181
    L.upload(b.getArea() * len(b.intensity), b);
    L.normalizeALSIntensities();
    const std::vector<hr_area_light> v = toHrLights(L);
    for (size_t i = 0; i < v.size(); ++i) {
        const float* f = &v[i].position[0];
        for (int k = 0; k < 15; ++k) printf("%.9g ", f[k]);
        printf("%.9g\n", L.weights()[i]);
    }
    // the records bind to the library's entry point (no context: an argument error)
    return hr_set_lights(nullptr, v.data(), L.weights(), (int)v.size()) == HR_OK ? 1 : 0;
}
"""


def binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"// \[lights-binding\].*?\n(.*?)// \[/lights-binding\]", text, re.S)
    assert m, "the integration notes have no lights-binding block"
    return m.group(1)


def test_lights_binding_compiles_and_converts(tmp_path, hr):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = tmp_path / "lights.cpp"
    src.write_text(PRELUDE + binding_source() + MAIN)
    exe = tmp_path / "lights"
    libdir = os.path.dirname(hr.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe), "-L", libdir, "-lhappyray", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    rows = np.array([[float(x) for x in line.split()] for line in out.stdout.strip().splitlines()])
    assert rows.shape == (2, 16)
    # light a: position (0,4,0), normal (0,-1,0), intensity 5, edges (2,0,0), (0,0,1)
    np.testing.assert_array_equal(rows[0, :15], [0, 4, 0, 0, -1, 0, 5, 5, 5, 2, 0, 0, 0, 0, 1])
    # light b: edges (0,3,0) and (1,0,0)
    np.testing.assert_array_equal(rows[1, :15], [-1, 0, 3, 0, 0, -1, 1, 2, 2, 0, 3, 0, 1, 0, 0])
    # weights: area x |intensity| = 2 * 5 sqrt(3) and 3 * 3
    np.testing.assert_allclose(rows[:, 15], [2 * 5 * np.sqrt(3.0), 9.0], rtol=1e-6)
