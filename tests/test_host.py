"""Host-side checks that need no GPU: the C ABI library loads and exports every entry point
include/prt_mi355x.h declares, the HitReg layouts match the reference headers for all 32
field masks, the no-device behaviour of the API, the deterministic workload generators, and
the C++ API headers compiling (and, without a device, taking the no-backend path)."""
import ctypes
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import portablert_amd as prt
from portablert_amd import scenes
from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "prt_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(prt_mi355x_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 14
    lib = ctypes.CDLL(prt.LIB)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_layouts_match_reference_headers():
    ref = np.load(os.path.join(GOLDEN, "hitreg_layouts.npy"))
    for mask in range(32):
        size, off = prt.layout(mask)
        assert [size] + off == list(ref[mask]), (mask, size, off, ref[mask])


def test_layout_sizes_known_values():
    assert prt.layout(31)[0] == 32
    assert prt.layout(prt.T | prt.PRIMITIVE_ID) == (16, [-1, -1, 4, 8, -1, -1, -1, -1])
    assert prt.layout(prt.T)[0] == 16
    assert prt.layout(prt.VALID)[0] == 8


def _gpu_present():
    return prt.probe()[0] > 0


@pytest.mark.skipif(_gpu_present(), reason="checks the no-device behaviour")
def test_no_device_behaviour():
    n, name = prt.probe()
    assert n == 0 and name == ""
    with pytest.raises(prt.PrtError) as e:
        prt.Context(0)
    assert e.value.code == prt.ERR_NODEV
    b = prt.MI355XBackend()
    assert not b.is_available()


def test_null_and_bad_arguments_need_no_device():
    lib = prt.lib()
    assert lib.prt_mi355x_probe(None, None, 0) == prt.ERR_ARG
    out = np.zeros(9, np.int32)
    assert lib.prt_mi355x_layout(32, out.ctypes.data_as(ctypes.c_void_p)) == prt.ERR_ARG
    assert lib.prt_mi355x_set_tris(None, None, 0) == prt.ERR_ARG
    assert lib.prt_mi355x_refit(None, None, 0) == prt.ERR_ARG
    assert lib.prt_mi355x_refit_device(None, None, 0) == prt.ERR_ARG
    assert lib.prt_mi355x_nearest_hits(None, None, 0, 0, None, 0, None) == prt.ERR_ARG
    assert lib.prt_mi355x_last_error(None) == b"null context"


def test_generators_are_deterministic_and_sized():
    for w, n in ((0, 16384), (1, 81920), (2, 262144)):
        a, b = scenes.scene(w), scenes.scene(w)
        assert a.shape == (n, 9) and np.array_equal(a, b)
        assert np.isfinite(a).all()
    assert scenes.n_rays(2) == 8388608 and scenes.n_rays(0) == 1 << 20
    r = scenes.rays(2, first=1000, count=50)
    assert np.array_equal(r, scenes.ray_sample(2, np.arange(1000, 1050)))
    assert scenes.n_scenes(3) == 64
    counts = [scenes.lib().prt_scene_tris(3, i) for i in range(64)]
    assert min(counts) >= 1000 and max(counts) <= 200000
    rc = [scenes.n_rays(3, i) for i in range(64)]
    assert min(rc) >= 10000 and max(rc) <= 2000000


def test_room_scene_shape():
    t = scenes.scene(2).reshape(-1, 3, 3)
    assert (t.min(axis=(0, 1)) >= -1e-6).all() and (t.max(axis=(0, 1)) <= np.array([30, 12, 18]) + 1e-5).all()
    a = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    med = float(np.median(np.log(a[a > 0])))
    assert -9 < med < -4  # log-normal areas around mu = -6 (caps and clamps move it a little)


def test_pinhole_directions_unit_and_centered():
    r = scenes.rays(0)
    d = r[:, 3:]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert (r[:, :3] == np.array([0, 0, -3], np.float32)).all()
    mid = d[512 * 1024 + 512]
    assert abs(mid[0]) < 2e-3 and abs(mid[1]) < 2e-3 and mid[2] > 0.99


def test_bounce_rays_leave_the_surface():
    import oracle
    tris = scenes.scene(1)
    prim = scenes.rays(1, first=1920 * 540, count=1920)
    hits = oracle.trace(tris, prim)
    p = np.stack([hits["px"], hits["py"], hits["pz"]], 1)
    b = scenes.bounce_rays(tris, prim, hits["t"], hits["primitive_id"], p)
    nhit = int((hits["t"] < np.inf).sum())
    assert len(b) == 4 * nhit and nhit > 0
    assert np.allclose(np.linalg.norm(b[:, 3:], axis=1), 1.0, atol=1e-5)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_cpp_api_compiles_and_runs(tmp_path):
    exe = tmp_path / "api_test"
    libdir = os.path.dirname(prt.LIB)
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "api_test.cpp"), "-L", libdir, "-lprt_mi355x",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    if not _gpu_present():
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "no-backend path" in r.stdout


REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference sources not present")
@pytest.mark.parametrize("name", ["triangle", "simple", "backends"])
def test_reference_examples_compile_unchanged(name, tmp_path):
    """The reference's own example programs build against these headers without edits."""
    libdir = os.path.dirname(prt.LIB)
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(REF_EXAMPLES, name, "main.cpp"), "-L", libdir, "-lprt_mi355x",
                    "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0


def test_batch_tiles_cover_every_ray_once():
    """The work tiles of the batch traversal: each inside one scene, every ray exactly once,
    tile sizes per size bucket, scenes with more triangles first (ties by index)."""
    rng = np.random.default_rng(5)
    tri = rng.integers(0, 300000, 40)
    tri[3], tri[4] = 0, 1
    ray = rng.integers(0, 5000, 40)
    ray[7] = 0
    chunk = (7, 64, 1, 500)
    t = prt.batch_tiles(tri, ray, chunk)
    ray_off = np.concatenate([[0], np.cumsum(ray)])
    first = t[:, 0].astype(np.int64) | (t[:, 1].astype(np.int64) << 32)
    cnt, scene = t[:, 2].astype(np.int64), t[:, 3]
    assert (cnt >= 1).all()
    assert (first >= ray_off[scene]).all() and (first + cnt <= ray_off[scene + 1]).all()
    cover = np.zeros(ray_off[-1], np.int32)
    for f, c in zip(first, cnt):
        cover[f:f + c] += 1
    assert (cover == 1).all()
    for s in range(40):
        c = cnt[scene == s]
        want = chunk[prt.batch_bucket(int(tri[s]))]
        assert (c[:-1] == want).all() and (len(c) == 0 or c[-1] <= want)
    order = [int(s) for k, s in enumerate(scene) if k == 0 or scene[k - 1] != s]
    assert order == sorted(order, key=lambda s: (-int(tri[s]), s))
    assert len(prt.batch_tiles([], [], chunk)) == 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says why, before any GPU work."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr


def test_bench_cpu_baseline_leg():
    """The bench's CPU leg on a small scene: the compiled reference where it was built (else the
    restatement), a positive rate, the restatement's rate alongside, and the host recorded."""
    sys.path.insert(0, ROOT)
    import bench
    import oracle
    tris = scenes.scene(0)
    rays = scenes.rays(0)[:20000]
    out = bench.cpu_baseline(tris, rays, target_s=0.3)
    assert out["unit"] == "Mrays/s" and out["value"] > 0 and out["cores"] >= 1
    assert out["kind"] == ("reference" if oracle.have_ref() else "port")
    assert out["host"]["affinity"] == bench.host_threads()
    if out["kind"] == "reference":
        assert out["port_value"] > 0


REF_TREE = "/root/reference"


def _integration_snippets():
    """The fenced code blocks of INTEGRATION.md section 2, in order, without the list indent."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text.split("## 2. Into the reference tree")[1].split("\n## 3.")[0]
    blocks, cur = [], None
    for line in sec.splitlines():
        body = line[3:] if line.startswith("   ") else line
        if body.startswith("```"):
            if cur is None:
                cur = []
            else:
                blocks.append("\n".join(cur) + "\n")
                cur = None
        elif cur is not None:
            cur.append(body)
    return blocks


def _insert(path, anchor, snippet, after=False):
    src = open(path).read()
    lines = src.split("\n")
    idx = [i for i, ln in enumerate(lines) if ln == anchor]
    assert len(idx) == 1, (path, anchor, len(idx))
    at = idx[0] + (1 if after else 0)
    lines[at:at] = snippet.rstrip("\n").split("\n")
    open(path, "w").write("\n".join(lines))


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_TREE, "src")), reason="reference sources not present")
def test_reference_tree_integration(tmp_path):
    """INTEGRATION.md section 2 applied verbatim to a copy of the reference tree: the reference's
    CPU backend and the MI355X backend are both compiled in (-DUSE_MI355X), the reference's
    examples build and run, CPU stays the default selection, MI355X is listed among the compiled
    backends (and among the available ones only where a gfx950 device is visible)."""
    import re
    import shutil
    tree = tmp_path / "ref"
    for d in ("include", "src"):
        shutil.copytree(os.path.join(REF_TREE, d), tree / d)
    for ex in ("triangle", "simple", "backends"):
        shutil.copytree(os.path.join(REF_TREE, "examples", ex), tree / "examples" / ex)
    # version.hpp: what the reference's configure_file(version.hpp.in ... @ONLY) generates
    ver = re.search(r"project\(portableRT VERSION (\d+)\.(\d+)\.(\d+)\)",
                    open(os.path.join(REF_TREE, "CMakeLists.txt")).read()).groups()
    tpl = open(os.path.join(REF_TREE, "version.hpp.in")).read()
    tpl = (tpl.replace("@PROJECT_VERSION@", ".".join(ver)).replace("@PROJECT_VERSION_MAJOR@", ver[0])
           .replace("@PROJECT_VERSION_MINOR@", ver[1]).replace("@PROJECT_VERSION_PATCH@", ver[2]))
    (tree / "include" / "portableRT" / "version.hpp").write_text(tpl)

    blocks = _integration_snippets()
    assert len(blocks) == 6, blocks  # cmake, forward declaration, variant, include, to_variant, include
    inc = tree / "include" / "portableRT"
    shutil.copy(os.path.join(ROOT, "include", "prt_mi355x.h"), tree / "include")
    shutil.copy(os.path.join(ROOT, "include", "portableRT", "intersect_mi355x.hpp"), inc)
    _insert(inc / "backend.hpp", "class EmbreeCPUBackend;", blocks[1], after=True)
    _insert(inc / "backend.hpp", "    CPUBackend *>;", blocks[2])
    _insert(tree / "src" / "backend.cpp", "namespace portableRT {", blocks[3])
    _insert(tree / "src" / "backend.cpp", "\tif (auto *q = dynamic_cast<CPUBackend *>(backend))", blocks[4])
    _insert(inc / "portableRT.hpp", '#include "nearest_hits_impl.hpp"', blocks[5])
    _insert(inc / "nearest_hits_impl.hpp", '#include "backend.hpp"', blocks[5])

    libdir = os.path.dirname(prt.LIB)
    srcs = [str(tree / "src" / f) for f in ("portableRT.cpp", "intersect_cpu.cpp", "backend.cpp")]
    outs = {}
    for ex in ("triangle", "simple", "backends"):
        exe = tmp_path / ex
        subprocess.run(["g++", "-std=c++17", "-O1", "-DUSE_MI355X", "-I", str(tree / "include")] + srcs +
                       [str(tree / "examples" / ex / "main.cpp"), "-L", libdir, "-lprt_mi355x", "-lpthread",
                        "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[ex] = r.stdout
    compiled = outs["backends"].split("compiled backends:")[1].split("available backends:")[0]
    assert "CPU" in compiled and "MI355X" in compiled, outs["backends"]
    available = outs["backends"].split("available backends:")[1]
    assert "CPU" in available
    assert ("MI355X" in available) == _gpu_present()
    # the examples test every available backend; CPU (the reference's) gives the known answer
    assert "Testing CPU" in outs["triangle"] and "Ray t: 1" in outs["triangle"], outs["triangle"]
    assert "Testing CPU" in outs["simple"] and "Ray 1: 1\nRay 2: 0" in outs["simple"], outs["simple"]


HELPERS_CPP = r"""
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>
#include <portableRT/core.hpp>
int main() {
    portableRT::clear_affinity();
    for (const auto &s : portableRT::all_tags_combinations()) std::cout << s << "\n";
    std::cout << "--\n";
    portableRT::Tri tri;
    portableRT::Ray ray;
    while (true) {
        for (float &x : tri) if (std::scanf("%a", &x) != 1) return 0;
        for (float &x : ray.origin) if (std::scanf("%a", &x) != 1) return 1;
        for (float &x : ray.direction) if (std::scanf("%a", &x) != 1) return 1;
        float t = 0, u = 0, v = 0;
        const bool hit = portableRT::intersect_tri(tri, ray, t, u, v);
        std::printf("%d %a %a %a\n", hit ? 1 : 0, hit ? t : 0.0f, hit ? u : 0.0f, hit ? v : 0.0f);
    }
}
"""


def _run_helpers(include_dir, tmp_path, tris, rays):
    src = tmp_path / "helpers.cpp"
    src.write_text(HELPERS_CPP)
    exe = tmp_path / ("helpers_" + os.path.basename(os.path.dirname(include_dir)))
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", include_dir, str(src), "-o", str(exe)],
                   check=True)
    inp = "\n".join(" ".join(float(x).hex() for x in np.concatenate([t, r])) for t, r in zip(tris, rays))
    r = subprocess.run([str(exe)], input=inp + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    names, rows = r.stdout.split("--\n")
    return names.split(), [ln.split() for ln in rows.strip().splitlines()]


def test_core_helpers(tmp_path):
    """intersect_tri, clear_affinity and all_tags_combinations of the restated core.hpp: the
    leaf test equals the oracle's restatement of core.hpp:27-65 bit for bit (hits, misses,
    det == 0, negative t), and the 31 names are TAG_COMBOS' order; where the reference headers
    are present, the same program built against them prints the same."""
    import oracle
    rng = np.random.default_rng(7)
    n = 3000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    rays = np.concatenate([rng.uniform(-2, 2, (n, 3)), rng.normal(size=(n, 3))], 1).astype(np.float32)
    tris[:5] = [-1, -1, 0, 1, -1, 0, 0, 1, 0]
    rays[:5] = [[0.1, 0, -1, 0, 0, 1], [0.1, 0, 1, 0, 0, 1], [0.1, 0, -1, 1, 0, 0], [-2, 0, -1, 0, 0, 1],
                [0.3, 0.2, 1, 0, 0, -1]]
    names, rows = _run_helpers(os.path.join(ROOT, "include"), tmp_path, tris, rays)
    assert len(names) == 31 and names[0] == "uv" and names[5] == "uv_t" and names[-1] == "uv_t_primitive_id_p_valid"
    hits = 0
    for tri, ray, row in zip(tris, rays, rows):
        hit, t, u, v = oracle.mt(tri, ray)
        assert int(row[0]) == int(hit)
        if hit:
            hits += 1
            got = [np.float32(float.fromhex(x)) for x in row[1:]]
            assert [g.tobytes() for g in got] == [np.float32(x).tobytes() for x in (t, u, v)]
    assert hits > 50
    if os.path.isdir(os.path.join(REF_TREE, "include")):
        ref_names, ref_rows = _run_helpers(os.path.join(REF_TREE, "include"), tmp_path, tris, rays)
        assert ref_names == names
        assert ref_rows == rows
