"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU): the C ABI (table builders for robots, scenes, cull tables, kept-pose
slots, cloud tile sizing, trajectory / SQP offsets, graphs, the no-allocation
entry point), the host mirror of the reference classes and its URDF reader,
built with g++ against host stand-ins of the HIP runtime
(tests/cpp/hip_host_stub.cpp) and driven by tests/cpp/host_sanitize.cpp over
the benchmark robots and the reference-shaped ones.  Kernels do not run on
the CPU; the run must finish without a sanitizer report, and the host
mirror's sphere reorder must match libstdc++'s unordered_set order."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from plainmp_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("host_sanitize")
    exe = str(out / "host_sanitize")
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    srcs = [os.path.join(CPP, "host_sanitize.cpp"), os.path.join(CPP, "hip_host_stub.cpp"),
            os.path.join(ROOT, "plainmp_amd", "csrc", "capi.cpp")]
    subprocess.run([cxx, *flags, *srcs, "-o", exe], check=True)
    return exe, out


def _prim_line(tag, p):
    R = " ".join(f"{v:.17g}" for v in np.asarray(p.get("rotation", np.eye(3)), float).ravel())
    pos = " ".join(f"{v:.17g}" for v in np.asarray(p.get("position", np.zeros(3)), float))
    t = p["type"]
    if t == "box":
        return f"{tag} box " + " ".join(f"{v:.17g}" for v in p["width"]) + f" {pos} {R}"
    if t == "sphere":
        return f"{tag} sphere {p['radius']:.17g} {pos} {R}"
    if t == "cylinder":
        return f"{tag} cylinder {p['radius']:.17g} {p['height']:.17g} {pos} {R}"
    return f"{tag} ground {p['height']:.17g}"


def _case(out, name, robot, prims, base=0, extra_cols=0, fixed=(), cloud=None, B=48, seed=0):
    rng = np.random.default_rng(seed)
    urdf = out / f"{name}.urdf"
    urdf.write_text(robot.urdf)
    Q = rng.uniform(robot.lower, robot.upper, size=(B, robot.dof))
    if extra_cols:
        Q = np.concatenate([Q, rng.normal(0, 0.1, size=(B, extra_cols))], axis=1)
    lines = [f"urdf_file {urdf}", f"joints {len(robot.joint_names)} " + " ".join(robot.joint_names), f"base {base}"]
    for ln, pos, r, only in robot.sphere_specs:
        vals = " ".join(f"{x:.17g} {y:.17g} {z:.17g} {rr:.17g}" for (x, y, z), rr in zip(np.asarray(pos), r))
        lines.append(f"spec {ln} {int(only)} {len(r)} {vals}")
    lines += [f"pair {a} {b}" for a, b in robot.self_pairs]
    lines += [_prim_line("prim", p) for p in prims]
    lines += [_prim_line("fixed", p) for p in fixed]
    if cloud is not None:
        lines.append(f"cloud {len(cloud)} 0.002 " + " ".join(f"{v:.9g}" for v in cloud.ravel()))
    lines.append(f"q {B} " + " ".join(f"{v:.17g}" for v in Q.ravel()))
    lines.append("width " + " ".join(["0.05"] * Q.shape[1]))
    path = out / f"{name}.case"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


CASES = {
    "panda_two_bars": lambda: (synth.panda_like(0), synth.two_bars(0), {}),
    "dual_arm_self": lambda: (synth.dual_arm(0), synth.mixed_primitives(24, seed=5), {}),
    "fetch_like": lambda: (synth.fetch_like(0), synth.mixed_primitives(16, seed=9), {"fixed": [
        {"type": "cylinder", "radius": 0.29, "height": 0.32, "position": [0.005, 0.0, 0.2], "rotation": np.eye(3)}]}),
    "humanoid31_floating": lambda: (synth.humanoid31(0), synth.mixed_primitives(8, seed=10), {"base": 2,
                                                                                               "extra_cols": 6}),
    "panda_planar_ground": lambda: (synth.panda_like(0), synth.mixed_primitives(6, seed=3) + [
        {"type": "ground", "height": 0.0}], {"base": 1, "extra_cols": 3}),
    "torso_cloud": lambda: (synth.mobile_torso(0), [], {"cloud": synth.shelf_cloud(2000, seed=2)}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_code_under_sanitizers(harness, name):
    exe, out = harness
    robot, prims, kw = CASES[name]()
    case = _case(out, name, robot, prims, **kw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, case], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout + r.stderr
    assert "runtime error" not in r.stderr, r.stderr
