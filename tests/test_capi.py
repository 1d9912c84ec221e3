"""C ABI checks that need no GPU: the library loads, exports every entry point include/*.h declares,
the ctypes structures match the C layouts, and the host-built M4 table equals the oracle's."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

from opensph_amd import gpu, gravity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "osph_gpu.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                 if f.endswith(".h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names |= set(re.findall(r"^\s*(?:int|void|const char\*)\s+(osph_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = gpu.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(gpu.EXPORTS + gravity.EXPORTS) == names


def test_gravity_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "osph_gravity.h"
int main(void) {
    printf("%zu %zu %zu %zu\n", sizeof(osph_gravity_config), sizeof(osph_gravity_stats),
           offsetof(osph_gravity_config, leaf_size), offsetof(osph_gravity_config, G));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(gravity.GravityConfig), C.sizeof(gravity.GravityStats), gravity.GravityConfig.leaf_size.offset,
            gravity.GravityConfig.G.offset]
    assert got == want


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "osph_gpu.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu\n", sizeof(osph_material), sizeof(osph_config), sizeof(osph_stats),
           offsetof(osph_material, young_modulus), offsetof(osph_config, precision));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(gpu.Material), C.sizeof(gpu.Config), C.sizeof(gpu.Stats),
            gpu.Material.young_modulus.offset, gpu.Config.precision.offset]
    assert got == want


def test_device_lut_equals_oracle_lut(oracle):
    v1, g1 = oracle.lut()
    v2, g2 = gpu.kernel_lut()
    assert np.array_equal(v1, v2)
    assert np.array_equal(g1, g2)


def test_quantity_ids_follow_header():
    text = open(HEADER).read()
    body = text[text.index("OSPH_Q_POSITION"):text.index("OSPH_Q_COUNT")]
    names = re.findall(r"OSPH_Q_(\w+)", body)
    assert {k: i for i, k in enumerate(names)} == gpu.Q


def test_setup_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "osph_gpu.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(osph_sphere_body), offsetof(osph_sphere_body, velocity),
           offsetof(osph_sphere_body, material), offsetof(osph_sphere_body, flag), sizeof(osph_collision),
           sizeof(osph_body_info), offsetof(osph_body_info, volume), offsetof(osph_body_info, mass));
    return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [C.sizeof(gpu.SphereBody), gpu.SphereBody.velocity.offset, gpu.SphereBody.material.offset,
            gpu.SphereBody.flag.offset, C.sizeof(gpu.Collision), C.sizeof(gpu.BodyInfo), gpu.BodyInfo.volume.offset,
            gpu.BodyInfo.mass.offset]
    assert got == want
