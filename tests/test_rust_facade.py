"""The Rust drop-in (rust/sfm-ba) against the C ABI and the build recipe.

No Rust toolchain exists here, so the crate is not compiled; these checks
pin what can be checked without one: every #[repr(C)] mirror in ffi.rs has
the fields of its struct in include/sfmba.h, in order, with the
corresponding types; every bound entry point is declared in the header with
the same parameter count; build.rs compiles the translation units and flags
that sfmtory_amd/build.py does; and the facade exports every public item of
the reference crate (crates/sfm-ba/src/lib.rs:49, :57-247, :708, :747;
intrinsics.rs:49, :66, :116).
"""
from __future__ import annotations

import re
from pathlib import Path

from sfmtory_amd import build as engine_build

ROOT = Path(__file__).resolve().parents[1]
HEADER = (ROOT / "include" / "sfmba.h").read_text()
CRATE = ROOT / "rust" / "sfm-ba"
FFI = (CRATE / "src" / "ffi.rs").read_text()
LIB = (CRATE / "src" / "lib.rs").read_text()
INTR = (CRATE / "src" / "intrinsics.rs").read_text()
BUILD_RS = (CRATE / "build.rs").read_text()

C_SCALAR = {"int32_t": "i32", "uint64_t": "u64", "size_t": "usize", "double": "f64", "uint8_t": "u8"}


def c_structs() -> dict[str, list[tuple[str, str]]]:
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", HEADER, re.S):
        body = re.sub(r"/\*.*?\*/", "", m.group(2), flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            const = decl.startswith("const ")
            decl = decl[6:] if const else decl
            base, names = decl.split(None, 1)
            base_ptr = base.endswith("*")
            base = base.rstrip("*")
            for name in names.split(","):
                name = name.strip()
                ptr = base_ptr or name.startswith("*")
                name = name.lstrip("*").strip()
                arr = re.match(r"(\w+)\[(\d+)\]", name)
                t = C_SCALAR.get(base, base)
                if arr:
                    name, t = arr.group(1), f"[{t}; {arr.group(2)}]"
                if ptr:
                    t = ("*const " if const else "*mut ") + t
                fields.append((name, t))
        out[m.group(1)] = fields
    return out


def rust_structs() -> dict[str, list[tuple[str, str]]]:
    out = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^)]*\)\]\s*)?pub struct (\w+) \{(.*?)\n\}", FFI, re.S):
        out[m.group(1)] = re.findall(r"pub (\w+): ([^,\n]+),", m.group(2))
    return out


def test_repr_c_mirrors_match_the_header():
    cs, rs = c_structs(), rust_structs()
    assert set(rs) == {"sfmba_camera", "sfmba_pose", "sfmba_observation", "sfmba_input", "sfmba_params",
                       "sfmba_output", "sfmba_intrinsics_params"}
    for name, fields in rs.items():
        assert fields == cs[name], name


def test_bound_entry_points_are_declared_with_the_same_arity():
    block = re.search(r'extern "C" \{(.*?)\n\}', FFI, re.S).group(1)
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)", block, re.S):
        name, params = m.group(1), m.group(2).strip()
        rust_n = len([p for p in params.split(",") if p.strip()])
        decl = re.search(rf"^(?:int|void|const char\*) {name}\((.*?)\);", HEADER, re.S | re.M)
        assert decl, name
        c_params = decl.group(1).strip()
        c_n = 0 if c_params in ("", "void") else len(c_params.split(","))
        assert rust_n == c_n, name


def test_build_rs_compiles_what_the_python_build_compiles():
    def listed(const):
        block = re.search(rf"const {const}: &\[&str\] = &\[(.*?)\];", BUILD_RS, re.S).group(1)
        return re.findall(r'"([^"]+)"', block)
    assert listed("HIP_SOURCES") == engine_build.HIP_SOURCES
    assert listed("HOST_SOURCES") == engine_build.HOST_SOURCES
    flags = listed("HIPCC_FLAGS")
    for f in ("--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-packed-fp32-ops"):
        assert f in flags and f in engine_build.HIPCC_FLAGS


def test_facade_exports_the_reference_public_api():
    for item in ("pub struct Observation", "pub struct BaInput", "pub enum RobustLoss", "pub struct BaParams",
                 "impl Default for BaParams", "pub struct BaOutput", "pub fn bundle_adjust(input: BaInput, params: &BaParams) -> BaOutput",
                 "pub fn mean_reprojection_error(input: &BaInput) -> f64",
                 "pub use intrinsics::{default_fixed_params_mask, refine_intrinsics, IntrinsicsRefineParams};"):
        assert item in LIB, item
    for item in ("pub struct IntrinsicsRefineParams", "impl Default for IntrinsicsRefineParams",
                 "pub fn default_fixed_params_mask(cameras: &[CameraModel]) -> Vec<Vec<bool>>",
                 "pub fn refine_intrinsics("):
        assert item in INTR, item
    # the defaults the ABI's *_default functions fill in
    assert "max_iterations: 50, initial_lambda: 1e-3, robust_loss: RobustLoss::Huber, loss_scale_px: 2.0" in LIB
    assert re.search(r"max_iterations: 30,\s*initial_lambda: 1e-3,\s*robust_loss: RobustLoss::Huber,\s*loss_scale_px: 2.0",
                     INTR)
