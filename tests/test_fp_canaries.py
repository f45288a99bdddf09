"""Level-0 compiler canaries for the exactness contract.

The reference's CUDA backend probes its compiled module before use:
(1+2^-29)(1-2^-29)-1 must be +0 (no fused multiply-add), and binary64
multiply-then-add, division and f64 -> f32 narrowing must equal the host on
100k samples (docs/cuda-decision-record.md:23-40); its CPU build has the
same canary (tests/test_fast_cpu_bitexact.py:347-375).  Here:

* on the GPU, the same operations run in dice_canary.hip's kernels over
  operands read from device memory, and every result is compared bit for
  bit with numpy (which rounds each operation once), binary32 denormals
  included;
* on the CPU, the canary's multiply-add kernels are compiled to gfx950
  assembly: with this build's flags, and with HIP's default (which honours
  the file's ``#pragma clang fp contract(off)``), they hold no fused
  instruction; built with -ffp-contract=fast (which disregards pragmas), or
  without the pragma, they do, so the GPU comparison would fail on such a
  build.
"""

from __future__ import annotations

import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import hip_device_or_skip
from dice_amd._abi import DiceFpCanary, ptr
from dice_amd.hip_backend import engine as hip_engine

ROOT = Path(__file__).resolve().parents[1]
CANARY_SOURCE = ROOT / "digital-fauxice_amd" / "csrc" / "dice_canary.hip"
MAKEFILE = ROOT / "digital-fauxice_amd" / "Makefile"
HIPCC = Path("/opt/rocm/bin/hipcc")
EXACT_FLAGS = ["-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
               "-fhip-fp32-correctly-rounded-divide-sqrt"]
FUSED = re.compile(r"\b(v_fma_f64|v_fmac_f64|v_fma_f32|v_fmac_f32|v_mad_f32|v_mac_f32|"
                   r"v_fma_mix\w*|v_pk_fma_f32|v_fmaak_f32|v_fmamk_f32|v_dot\w*)"
                   r"(_e32|_e64|_dpp|_sdwa)?\b")


def canary_inputs(n: int = 100_000, seed: int = 20260716):
    rng = np.random.default_rng(seed)

    def wide(size, low, high):
        signs = rng.choice([-1.0, 1.0], size)
        return signs * np.exp2(rng.uniform(low, high, size)) * rng.uniform(1.0, 2.0, size)

    a = wide(n, -60, 60)
    b = wide(n, -60, 60)
    c = wide(n, -60, 60)
    # half the samples cancel: c = -(a*b) rounded, where a fused multiply-add
    # leaves the product's rounding error instead of +0
    cancel = rng.random(n) < 0.5
    c[cancel] = -(a[cancel] * b[cancel])
    # the reference's probe, and binary64 subnormal operands
    a[0], b[0], c[0] = 1.0 + 2.0**-29, 1.0 - 2.0**-29, -1.0
    a[1:5] = [5e-324, -2.5e-310, 1e-308, 3e-40]
    b[1:5] = [1.0, 0.5, 3.0, 1.0]
    fa = wide(n, -40, 40).astype(np.float32)
    fb = wide(n, -40, 40).astype(np.float32)
    fc = wide(n, -40, 40).astype(np.float32)
    fcancel = rng.random(n) < 0.5
    fc[fcancel] = -(fa[fcancel] * fb[fcancel])
    fa[0], fb[0], fc[0] = np.float32(1 + 2.0**-13), np.float32(1 - 2.0**-13), np.float32(-1)
    # binary32 subnormal operands and results
    tiny = np.float32(2.0**-70)
    fa[1:6] = [np.float32(1e-40), tiny, np.float32(-3e-39), np.float32(2.0**-149), tiny]
    fb[1:6] = [np.float32(1.0), np.float32(2.0**-60), np.float32(0.25), np.float32(1.0),
               np.float32(2.0**60)]
    # narrowing into binary32's subnormal range
    a[5:8] = [1e-40, -7.1e-42, 2.0**-140 * 1.5]
    return a, b, c, fa, fb, fc


@pytest.mark.gpu
def test_device_arithmetic_is_ieee_exact():
    hip_device_or_skip()
    a, b, c, fa, fb, fc = canary_inputs()
    n = a.size
    out = {name: np.empty(n, np.float64) for name in ("mul_add", "quotient")}
    out.update({name: np.empty(n, np.float32) for name in ("narrowed", "f32_mul_add",
                                                           "f32_quotient", "f32_product")})
    io = DiceFpCanary(n, ptr(a, C.c_double), ptr(b, C.c_double), ptr(c, C.c_double),
                      ptr(fa, C.c_float), ptr(fb, C.c_float), ptr(fc, C.c_float),
                      ptr(out["mul_add"], C.c_double), ptr(out["quotient"], C.c_double),
                      ptr(out["narrowed"], C.c_float), ptr(out["f32_mul_add"], C.c_float),
                      ptr(out["f32_quotient"], C.c_float), ptr(out["f32_product"], C.c_float))
    code = hip_engine.library().dice_t_fp_canaries(hip_engine.context(), C.byref(io))
    assert code == 0, hip_engine.last_error()
    with np.errstate(all="ignore"):
        want = {
            "mul_add": a * b + c,
            "quotient": a / b,
            "narrowed": a.astype(np.float32),
            "f32_mul_add": fa * fb + fc,
            "f32_quotient": fa / fb,
            "f32_product": fa * fb,
        }
    # the probes first: +0, not the product's rounding error
    assert out["mul_add"][0] == 0.0 and not np.signbit(out["mul_add"][0])
    assert out["f32_mul_add"][0] == 0.0 and not np.signbit(out["f32_mul_add"][0])
    for name, expected in want.items():
        got_bits = out[name].view(np.uint64 if out[name].dtype == np.float64 else np.uint32)
        want_bits = expected.view(got_bits.dtype)
        bad = np.flatnonzero(got_bits != want_bits)
        assert bad.size == 0, (name, bad[:5], out[name][bad[:5]], expected[bad[:5]])
    # subnormals survived: they are in the operands and in the results
    assert np.count_nonzero((np.abs(out["f32_product"]) < np.finfo(np.float32).tiny)
                            & (out["f32_product"] != 0)) >= 3
    assert np.count_nonzero((np.abs(out["narrowed"]) < np.finfo(np.float32).tiny)
                            & (out["narrowed"] != 0)) >= 3
    assert (np.count_nonzero(out["mul_add"][1:] == 0.0)
            > 0.4 * n)  # the cancelling half stays exactly zero


def device_assembly(source: Path, flags: list[str], tmp_path: Path) -> str:
    out = tmp_path / (source.stem + "-" + str(abs(hash(tuple(flags)))) + ".s")
    subprocess.run([str(HIPCC), "-O3", "-std=c++17", "--offload-arch=gfx950", "-S",
                    "--cuda-device-only", *flags, "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-o", str(out), str(source)],
                   check=True, capture_output=True)
    return out.read_text()


def kernel_body(assembly: str, kernel: str) -> str:
    names = [m for m in re.findall(r"^(_Z\w+):", assembly, re.M) if kernel in m]
    assert len(names) == 1, (kernel, names)
    start = assembly.index(names[0] + ":")
    end = assembly.index("s_endpgm", start)
    return assembly[start:end]


@pytest.mark.skipif(not HIPCC.exists(), reason="hipcc is not installed")
def test_canary_kernels_hold_fused_instructions_only_when_contraction_is_allowed(tmp_path):
    makefile = MAKEFILE.read_text()
    for flag in EXACT_FLAGS:
        assert flag in makefile, flag
    exact = device_assembly(CANARY_SOURCE, EXACT_FLAGS, tmp_path)
    for kernel in ("k_canary_mul_add_f64", "k_canary_mul_add_f32"):
        body = kernel_body(exact, kernel)
        assert "v_mul_f" in body and "v_add_f" in body, kernel
        assert not FUSED.search(body), (kernel, FUSED.search(body))
    # without the flag, HIP's default contraction honours the file's pragma
    default = device_assembly(CANARY_SOURCE, [], tmp_path)
    for kernel in ("k_canary_mul_add_f64", "k_canary_mul_add_f32"):
        assert not FUSED.search(kernel_body(default, kernel)), kernel
    # -ffp-contract=fast disregards pragmas: both kernels fuse, so a library
    # built that way fails test_device_arithmetic_is_ieee_exact
    fused = device_assembly(CANARY_SOURCE, ["-ffp-contract=fast"], tmp_path)
    for kernel in ("k_canary_mul_add_f64", "k_canary_mul_add_f32"):
        assert FUSED.search(kernel_body(fused, kernel)), kernel
    # and so does a source without the pragma under the default
    stripped = tmp_path / "canary_without_pragma.hip"
    text = CANARY_SOURCE.read_text()
    assert "#pragma clang fp contract(off)" in text
    stripped.write_text(text.replace("#pragma clang fp contract(off)", ""))
    unprotected = device_assembly(stripped, [], tmp_path)
    for kernel in ("k_canary_mul_add_f64", "k_canary_mul_add_f32"):
        assert FUSED.search(kernel_body(unprotected, kernel)), kernel


def test_every_translation_unit_keeps_contraction_off():
    """Every source the HIP compiler builds carries the pragma; a host-only
    source the host compiler builds (no device code) gets the flag."""

    sources = sorted((ROOT / "digital-fauxice_amd" / "csrc").glob("*.hip")) + sorted(
        (ROOT / "digital-fauxice_amd" / "csrc").glob("dice_*.cpp"))
    assert sources
    makefile = (ROOT / "digital-fauxice_amd" / "Makefile").read_text()
    host_rules = re.findall(r"\$\(SRC\)/(\w+\.cpp)[^\n]*\n(?:\t[^\n]*\n)*?\t\$\(CXX\)([^\n]*)",
                            makefile)
    host_only = {name: flags for name, flags in host_rules}
    assert set(host_only) == {"dice_sha.cpp"}
    for source in sources:
        if source.name in host_only:
            assert "-ffp-contract=off" in host_only[source.name], source.name
        else:
            assert "#pragma clang fp contract(off)" in source.read_text(), source.name
    assert shutil.which("make")
