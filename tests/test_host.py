"""CPU tests of the host logic and the C ABI (no GPU): reference test/parser.cpp,
test/codegen.cpp and test/error.cpp restated, plus the library's own host pieces."""
import ctypes
import json
import os
import re
import struct
import tempfile

import numpy as np
import pytest

import bbfft_amd as bb
from bbfft_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "bbfft_hip.h")).read()
    declared = set(re.findall(r"BBFFT_EXPORT_C[^;(]*?\b(bbfft_\w+)\s*\(", header))
    assert declared == set(_capi.EXPORTED)
    lib = _capi.load()
    for name in declared:
        assert hasattr(lib, name), name


def _shape(*v):
    return list(v) + [0] * (5 - len(v))


# reference test/parser.cpp:33-125: (descriptor, dim, shape, fp, dir, type, inplace-default,
# explicit istride, explicit ostride)
PARSER_CASES = [
    ("srfi5", 1, _shape(1, 5, 1), 4, -1, 1, True, None, None),
    ("dcbi4*5", 1, _shape(1, 4, 5), 8, 1, 0, True, None, None),
    ("dcbi4.5", 1, _shape(4, 5, 1), 8, 1, 0, True, None, None),
    ("drbo4.5*6", 1, _shape(4, 5, 6), 8, 1, 2, False, None, None),
    ("drfo5x6x7", 3, _shape(1, 5, 6, 7, 1), 8, -1, 1, False, None, None),
    ("srbo4.5x6*7", 2, _shape(4, 5, 6, 7), 4, 1, 2, False, None, None),
    ("scfo16*32i1,1,20", 1, _shape(1, 16, 32), 4, -1, 0, False, _shape(1, 1, 20), None),
    ("scfo16*32i1,1,20o1,2,32", 1, _shape(1, 16, 32), 4, -1, 0, False, _shape(1, 1, 20),
     _shape(1, 2, 32)),
    ("scfi16*32i1,1,20o1,1,20", 1, _shape(1, 16, 32), 4, -1, 0, False, _shape(1, 1, 20),
     _shape(1, 1, 20)),
]


@pytest.mark.parametrize("case", PARSER_CASES, ids=[c[0] for c in PARSER_CASES])
def test_fft_descriptor_parser(case):
    desc, dim, shape, fp, d, t, inplace, ist, ost = case
    cfg = bb.parse_fft_descriptor(desc)
    assert cfg.dim == dim
    assert list(cfg.shape) == shape
    assert int(cfg.fp) == fp and int(cfg.dir) == d and int(cfg.type) == t
    assert list(cfg.istride) == (ist or bb.default_istride(dim, shape, t, inplace))
    assert list(cfg.ostride) == (ost or bb.default_ostride(dim, shape, t, inplace))
    assert cfg.to_string() == desc


@pytest.mark.parametrize("bad", ["srb4x5*6x7", "", "x", "sxfo4", "scxo4", "scfx4", "scfo",
                                 "scfo4.", "scfo4x5x6x7", "scfo4*5.6", "scfo16*32i1,1",
                                 "scfo16*32q"])
def test_malformed_descriptors_throw(bad):
    with pytest.raises(bb.DescriptorError) as e:
        bb.parse_fft_descriptor(bad)
    assert "is malformed" in str(e.value)


@pytest.mark.parametrize("bad", ["scfo2147483648*1", "scfo8.256*99999999999999999999",
                                 "scfo18446744073709551617*2", "scfi16*32i1,4294967296"])
def test_descriptor_numbers_out_of_int_range_throw(bad):
    """Descriptor numbers are ints, as the reference reads them (std::stoi throws beyond
    2^31 - 1); nothing wraps around."""
    with pytest.raises(bb.DescriptorError, match="out of range"):
        bb.parse_fft_descriptor(bad)
    assert bb.parse_fft_descriptor("scfo2147483647*1").shape[1] == 2147483647


def test_malformed_descriptor_caret_points_at_error():
    with pytest.raises(bb.DescriptorError) as e:
        bb.parse_fft_descriptor("srb4x5*6x7")
    lines = str(e.value).splitlines()
    # "==> " prefix is 4 characters; the fourth descriptor character ('4') is offending
    assert lines[1].index("^") == 4 + 3


def test_device_info_roundtrip():
    # reference test/parser.cpp:25-31
    for s in ["{1024, {16, 32}, 131072, gpu}", "{8192, {4, 8, 16, 32, 64}, 32768, cpu}",
              "{1024, {64}, 163840, gpu}"]:
        assert bb.normalize_device_info(s) == s
    assert bb.normalize_device_info("{ 1024 ,{ 64 } , 163840 ,gpu }") == "{1024, {64}, 163840, gpu}"
    with pytest.raises(bb.DescriptorError):
        bb.normalize_device_info("{1024, {64}, 163840, tpu}")


@pytest.mark.parametrize("dim,shape,type_,inplace,expect", [
    (1, [8, 256, 32768], 0, False, [1, 8, 2048, 0, 0]),
    (1, [1, 64, 16384], 0, True, [1, 1, 64, 0, 0]),
    (1, [3, 10, 5], 1, True, [1, 3, 36, 0, 0]),        # in-place r2c pads 2 (N/2 + 1)
    (1, [3, 10, 5], 1, False, [1, 3, 30, 0, 0]),
    (1, [3, 10, 5], 2, False, [1, 3, 18, 0, 0]),       # c2r input has N/2 + 1 complex
    (3, [1, 128, 128, 128, 32], 1, False, [1, 1, 128, 16384, 2097152]),
])
def test_default_istride_closed_form(dim, shape, type_, inplace, expect):
    assert bb.default_istride(dim, shape, type_, inplace) == expect


def test_default_ostride_mirrors_istride():
    shape = [1, 128, 128, 128, 32]
    assert bb.default_ostride(3, shape, 1, False) == [1, 1, 65, 8320, 1064960]
    assert bb.default_ostride(3, shape, 2, True) == bb.default_istride(3, shape, 1, True)
    assert bb.default_ostride(3, shape, 0, True) == bb.default_istride(3, shape, 0, True)


def test_to_string_rejects_wrong_real_direction():
    cfg = bb.configuration(1, [1, 8, 1], bb.precision.f32, bb.direction.backward,
                           bb.transform_type.r2c)
    with pytest.raises(bb.DescriptorError):
        cfg.to_string()


# reference test/codegen.cpp:71-125
FACTOR_TABLE = [
    (4096, 0, []), (4096, 1, [4096]), (4096, 2, [64, 64]), (4096, 3, [16, 16, 16]),
    (2080, 2, [40, 52]), (2080, 3, [10, 13, 16]), (2080, 4, [4, 5, 8, 13]),
    (3465, 2, [55, 63]), (3465, 3, [11, 15, 21]), (3465, 4, [5, 7, 9, 11]),
    (9216, 2, [96, 96]), (9216, 3, [16, 24, 24]), (9216, 4, [8, 8, 12, 12]),
    (64536, 2, [24, 2689]), (64536, 3, [1, 24, 2689]), (64536, 4, [1, 2, 12, 2689]),
    (65536, 2, [256, 256]), (65536, 3, [32, 32, 64]), (65536, 4, [16, 16, 16, 16]),
]


@pytest.mark.parametrize("n,index,expect", FACTOR_TABLE)
def test_factor_reference_table(n, index, expect):
    assert bb.factor(n, index) == expect


def test_iroot_is_prime_reference_values():
    L = _capi.load()
    # reference test/codegen.cpp:61-69,97-103
    for (r, i, e) in [(0, 5, 0), (3, 1, 3), (25, 2, 5), (1331, 3, 11), (1337, 3, 11),
                      (4096, 3, 16), (64539, 2, 254), (64539, 3, 40), (64539, 4, 15)]:
        assert L.bbfft_iroot(r, i) == e
    for n, p in [(1, 0), (2, 1), (13, 1), (25, 0), (2689, 1)]:
        assert L.bbfft_is_prime(n) == p


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference sources not built")
def test_factor_matches_reference_build():
    """The library's factor() against the reference's own prime_factorization.cpp."""
    ref = ctypes.CDLL(REF_LIB)
    out = (ctypes.c_uint * 4)()
    for index in (1, 2, 3, 4):
        for n in list(range(1, 1200)) + [2080, 3465, 4096, 9216, 64536, 65536]:
            ref.ref_factor(n, index, out)
            assert bb.factor(n, index) == list(out)[:index], (n, index)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference sources not built")
def test_iroot_is_prime_match_reference_build():
    """Compared where the reference is exact: its bisection multiplies in `int` and overflows
    once a trial root exceeds ~46341 (e.g. iroot(2^20, 2) returns 983040 there); the library's
    iroot stays exact (checked separately below)."""
    ref = ctypes.CDLL(REF_LIB)
    L = _capi.load()
    assert L.bbfft_iroot(1 << 20, 2) == 1024 and L.bbfft_iroot((1 << 32) - 1, 2) == 65535
    for r in list(range(0, 3000, 7)) + [40000, 64539]:
        for i in (1, 2, 3, 4):
            assert L.bbfft_iroot(r, i) == ref.ref_iroot(r, i)
    for n in range(0, 5000):
        assert L.bbfft_is_prime(n) == ref.ref_is_prime(n)


def test_roots_match_reference_exact_tables():
    """Twiddle constants vs the reference's correctly rounded tables (fixture extracted from
    src/base/root_of_unity.hpp by tests/golden/make_w_lut.py)."""
    data = json.load(open(os.path.join(ROOT, "tests", "golden", "w_lut.json")))["values"]
    n_checked = 0
    for n, vals in data.items():
        for i, (re_, im_) in enumerate(vals):
            w = bb.root_of_unity(i, int(n), 8)
            assert (w.real, w.imag) == (float.fromhex(re_), float.fromhex(im_)), (n, i)
            # conjugate symmetry is exact
            wc = bb.root_of_unity(-i, int(n), 8)
            assert (wc.real, wc.imag) == (w.real, -w.imag)
            n_checked += 1
    assert n_checked > 150


def test_roots_are_correctly_rounded():
    """Roots are the exact value rounded once to float / double: checked against 40-digit
    mpmath values (round-to-nearest of the exact value), including exact zeros on the axes."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    for n in (3, 7, 12, 60, 100, 256, 500, 512):
        for k in range(n):
            c = mpmath.cos(2 * mpmath.pi * k / n)
            s_ = mpmath.sin(2 * mpmath.pi * k / n)
            w4 = bb.root_of_unity(k, n, 4)
            w8 = bb.root_of_unity(k, n, 8)
            for got, ex in ((w4.real, c), (w4.imag, s_)):
                assert got == float(np.float32(float(mpmath.nstr(ex, 30))) if abs(ex) > 1e-30
                                    else 0.0) or abs(got - float(ex)) <= float(
                    np.spacing(np.float32(abs(float(ex))))) / 2, (n, k)
            for got, ex in ((w8.real, c), (w8.imag, s_)):
                if abs(ex) < 1e-30:
                    assert got == 0.0, (n, k)
                else:
                    assert abs(mpmath.mpf(got) - ex) <= mpmath.mpf(np.spacing(abs(got))) / 2, (n, k)
    assert bb.root_of_unity(1, 4, 4) == 1j
    assert bb.root_of_unity(1, 8, 4).real == float(np.float32(np.sqrt(0.5)))


def test_unsupported_fft_dim():
    """reference test/error.cpp:13-21: dim 0 and dim 4 -> bad_configuration (no GPU needed)."""
    for dim in (0, 4):
        cfg = bb.configuration(dim, [1, 8, 8, 8, 1], istride=[1, 1, 8, 64, 512],
                               ostride=[1, 1, 8, 64, 512])
        with pytest.raises(bb.BadConfiguration):
            bb.make_plan(cfg, stream=0)


def test_nd_callbacks_placed_and_custom_strides_rejected():
    """The reference refuses callbacks on 2-D/3-D plans; here (an extension) the load callback
    goes to the first pass and the store callback to the last. Non-default nd strides stay
    refused, as in the reference."""
    cfg = bb.configuration(2, [1, 8, 8, 1], callbacks=bb.user_module(
        "__device__ float2 ld(float2* p, size_t o) { return p[o]; }", "ld"))
    assert "_ld_cb" in bb.kernel_configuration(cfg, 0)["name"]
    assert "_ld_cb" not in bb.kernel_configuration(cfg, 1)["name"]
    cfg = bb.configuration(2, [1, 8, 8, 1], istride=[1, 1, 9, 72, 0], ostride=[1, 1, 9, 72, 0])
    with pytest.raises(bb.BadConfiguration):
        bb.make_plan(cfg, stream=0)


def test_unit_stride_m_required():
    cfg = bb.configuration(1, [2, 8, 3], istride=[2, 4, 32], ostride=[2, 4, 32])
    with pytest.raises(bb.BadConfiguration):
        bb.make_plan(cfg, stream=0)


@pytest.mark.parametrize("desc,fac,Nt,Mb,Kb", [
    ("scfo3.256*32768", [16, 16], 16, 3, 5),
    ("scfo3.16*32768", [16], 1, 3, 85),
    ("scfo3.512*32768", [8, 8, 8], 64, 3, 1),
    ("dcfo3.512*32768", [8, 8, 8], 64, 3, 1),
    ("scfo3.60*32768", [10, 6], 6, 3, 14),
    ("scfo2.64*16384", [8, 8], 8, 2, 16),
])
def test_kernel_configuration_heuristic_tiling(desc, fac, Nt, Mb, Kb):
    """Heuristics for shapes without a tuned entry (M = 2, 3 are not in the table)."""
    kc = bb.kernel_configuration(bb.parse_fft_descriptor(desc))
    assert (kc["factorization"], kc["Nt"], kc["Mb"], kc["Kb"]) == (fac, Nt, Mb, Kb)
    assert kc["work_group_size"] == Mb * Nt * Kb <= 1024
    assert kc["lds_bytes"] <= 65536
    assert kc["name"].startswith("bbfft_")


def _table_entries():
    path = os.path.join(ROOT, "double-batched-fft-library_amd", "tuning", "gfx950.inc")
    out = []
    for line in open(path):
        line = line.split("#")[0]
        if ":" not in line:
            continue
        lhs, rhs = line.split(":")
        prec, typ, N, M = lhs.split()
        fac, Nt, Mb, Kb, pad, flags, *occupancy = rhs.split()  # [waves_per_eu [groups]]
        assert len(occupancy) <= 2, line
        out.append((prec, typ, int(N), int(M), [int(f) for f in fac.split("x")], int(Nt),
                    int(Mb), int(Kb), int(pad), int(flags)))
    return out


def test_tuning_table_is_applied():
    """Without explicit tuning, configure_fft1d uses the persisted autotuning table."""
    entries = _table_entries()
    assert len(entries) >= 20
    for prec, typ, N, M, fac, Nt, Mb, Kb, pad, flags in entries:
        if typ == "c2c-ragged":  # ragged variants: tests/test_autotune_host.py checks them
            continue
        if "-" in typ:  # half-input kernels: c2c-{f16,bf16}[-mfma]
            _, storage, *mfma = typ.split("-")
            kc = bb.half_input_kernel_configuration(f"scfo{M}.{N}*65536", storage, bool(mfma))
            assert (kc["factorization"], kc["Nt"], kc["Mb"], kc["lds_pad"]) == \
                (fac, Nt, Mb, pad), typ
            assert kc["Kb"] >= Kb and kc["work_group_size"] % 64 == 0
            continue
        d = {"c2c": "c", "r2c": "r", "c2r": "r"}[typ]
        dirn = "b" if typ == "c2r" else "f"
        desc = f"{prec[0].replace('f', 's') if prec == 'f32' else 'd'}{d}{dirn}o{M}.{N}*65536"
        kc = bb.kernel_configuration(bb.parse_fft_descriptor(desc))
        assert (kc["factorization"], kc["Nt"], kc["Mb"], kc["Kb"], kc["lds_pad"]) == \
            (fac, Nt, Mb, Kb, pad), desc
        if flags:
            assert f"_g{flags}" in kc["name"], desc


@pytest.mark.parametrize("desc", ["scfo8.256*32768", "srfo8.256*32768", "scfo8.16*32768",
                                  "scfo8.512*32768", "dcfo8.60*32768"])
def test_callback_plans_take_table_tiling_without_prefetch(desc):
    """Tuned shapes whose table entry carries the load-first tile loop (flag 64) still accept
    callbacks: the plan keeps the entry's tiling and drops the flag."""
    cfg = bb.parse_fft_descriptor(desc)
    plain = bb.kernel_configuration(cfg)
    fp = "float" if cfg.fp == bb.precision.f32 else "double"
    t = fp if cfg.type == bb.transform_type.r2c else fp + "2"
    cfg.callbacks = bb.user_module(
        f"__device__ {t} ld({t}* p, size_t o) {{ return p[o]; }}", "ld")
    kc = bb.kernel_configuration(cfg)
    assert (kc["factorization"], kc["Nt"], kc["Mb"], kc["Kb"]) == \
        (plain["factorization"], plain["Nt"], plain["Mb"], plain["Kb"])
    flags = int(kc["name"].split("_g")[1].split("_")[0]) if "_g" in kc["name"] else 0
    assert not flags & 64
    assert "_cb" in kc["name"] and "_cb" not in plain["name"]


def test_callback_source_and_language_enter_the_identifier():
    """Same callback name, different code or language: different kernel identifiers (the
    identifier is the jit_cache key and the symbol name)."""
    def name(src, lang):
        cfg = bb.parse_fft_descriptor("srfo1.64*16")
        cfg.callbacks = bb.user_module(src, None, "store", lang)
        return bb.kernel_configuration(cfg)["name"]
    a = "__device__ void store(float2* o, size_t i, float2 v) { o[i] = v; }"
    b = "__device__ void store(float2* o, size_t i, float2 v) { v.x *= 2.0f; o[i] = v; }"
    names = {name(a, bb.kernel_language.hip), name(b, bb.kernel_language.hip),
             name(a, bb.kernel_language.opencl_c)}
    assert len(names) == 3
    assert name(a, bb.kernel_language.hip) == name(a, bb.kernel_language.hip)


def test_load_callback_kernels_stream_their_stores(monkeypatch):
    """The tuning table's entry for the 32 x 128^3 plans' 128-point pass (M = 8320) stores
    plainly; a kernel of that shape with a load callback streams its stores (flag 2 in the
    identifier's _g field), so the callback's side data is not evicted from the last-level
    cache by the output. A store callback does its own stores: no flag is added.
    BBFFT_CALLBACK_NT=off keeps the entry's stores."""
    def name(load, store):
        cfg = bb.parse_fft_descriptor("scbo8320.128*32")
        src = ("__device__ float2 ld(float2* in, size_t i) { return in[i]; }\n"
               "__device__ void st(float2* o, size_t i, float2 v) { o[i] = v; }")
        if load or store:
            cfg.callbacks = bb.user_module(src, "ld" if load else None, "st" if store else None,
                                           bb.kernel_language.hip)
        return bb.kernel_configuration(cfg)["name"]
    def flags(kernel):  # the access-hint field after the padding fields ("" when none)
        m = re.search(r"_in0(?:_g(\d+))?", kernel)
        return m.group(1) or ""
    assert flags(name(False, False)) == ""
    assert flags(name(True, False)) == "2"
    assert flags(name(False, True)) == ""
    monkeypatch.setenv("BBFFT_CALLBACK_NT", "off")
    assert flags(name(True, False)) == ""


def test_nd_decomposition_passes():
    """2-D/3-D plans run one double-batched pass per mode (reference nd_fft.hpp:61-99)."""
    cfg = bb.parse_fft_descriptor("srfo128x128x128*32")
    assert bb.num_passes(cfg) == 3
    ks = [bb.kernel_configuration(cfg, p) for p in range(3)]
    assert [(k["M"], k["N"], k["K"]) for k in ks] == [(1, 128, 128 * 128 * 32),
                                                      (65, 128, 128 * 32), (65 * 128, 128, 32)]
    icfg = bb.parse_fft_descriptor("srbo128x128x128*32")
    ks = [bb.kernel_configuration(icfg, p) for p in range(3)]
    assert [(k["M"], k["N"], k["K"]) for k in ks] == [(65 * 128, 128, 32), (65, 128, 128 * 32),
                                                      (1, 128, 128 * 128 * 32)]


def test_generated_sources_compile_with_hiprtc():
    for desc in ["scfo8.256*32768", "drbo3.27*5", "srfi1.64*7"]:
        assert bb.compile_kernel(bb.parse_fft_descriptor(desc)) > 1000


def test_aot_archive_format(tmp_path):
    out = str(tmp_path / "k.bin")
    n = bb.aot_generate_shard(["scfo8.64*100", "dcbo3.12*5"], out)
    assert n == 2
    data = open(out, "rb").read()
    assert data[:8] == b"BBFFTAOT"
    version, count = struct.unpack_from("<II", data, 8)
    assert version == 1 and count == 1
    p = 16
    (alen,) = struct.unpack_from("<I", data, p)
    assert data[p + 4:p + 4 + alen] == b"gfx950"
    p += 4 + alen
    (nn,) = struct.unpack_from("<I", data, p)
    p += 4
    names = []
    for _ in range(nn):
        (ln,) = struct.unpack_from("<I", data, p)
        names.append(data[p + 4:p + 4 + ln].decode())
        p += 4 + ln
    (size,) = struct.unpack_from("<Q", data, p)
    assert data[p + 8:p + 12] == b"\x7fELF" and p + 8 + size == len(data)
    expect = {bb.kernel_configuration(bb.parse_fft_descriptor(d))["name"]
              for d in ["scfo8.64*100", "dcbo3.12*5"]}
    assert set(names) == expect


def test_c_abi_rejects_invalid_arguments():
    L = _capi.load()
    assert L.bbfft_parse_fft_descriptor(None, None) == _capi.ERROR_INVALID_ARGUMENT
    assert b"null" in L.bbfft_last_error()
    c = _capi.CConfiguration()
    c.dim = 1
    c.precision = 3
    assert L.bbfft_configuration_to_string(ctypes.byref(c), None, 0) == \
        _capi.ERROR_INVALID_ARGUMENT
    assert L.bbfft_plan_execute(None, None, None, None, 0, None) == _capi.ERROR_INVALID_ARGUMENT


def test_plan_execute_checks_tensor_arguments():
    """Plan.execute rejects host tensors and device buffers smaller than the configured tensor
    before anything reaches the library (a smaller buffer would be addressed out of bounds);
    the byte counts follow the configuration's extents: scfo8.256*4 spans 8 x 256 x 4 complex
    elements, srfo256*4 reads 256 x 4 reals and writes 129 x 4 complex elements."""
    import ctypes
    import torch
    cfg = bb.parse_fft_descriptor("scfo8.256*4")
    assert bb._io_bytes(cfg) == (8 * 256 * 4 * 8, 8 * 256 * 4 * 8)
    r2c = bb.parse_fft_descriptor("srfo256*4")
    assert bb._io_bytes(r2c) == (256 * 4 * 4, 129 * 4 * 8)
    c2r = bb.parse_fft_descriptor("drbo256*4")
    assert bb._io_bytes(c2r) == (129 * 4 * 16, 256 * 4 * 8)
    assert bb._io_bytes(cfg, in_elem_bytes=4)[0] == 8 * 256 * 4 * 4
    plan = bb.Plan(ctypes.c_void_p(), cfg, 0)
    x = torch.empty(8 * 256 * 4, dtype=torch.complex64)
    with pytest.raises(ValueError, match="GPU tensor"):
        plan.execute(x, x)
    with pytest.raises(ValueError, match="GPU tensor"):
        plan.execute(x)


def test_ragged_plan_checks_tensor_arguments():
    """Ragged plans check tensor arguments against the jobs' extents like plain plans."""
    import ctypes
    import torch
    from bbfft_amd import ragged
    jobs = [ragged.Job(M=2, N=60, K=10, in_offset=0, out_offset=5000),
            ragged.Job(M=1, N=12, K=100, in_offset=1200, out_offset=0)]
    assert ragged.io_bytes(jobs, 4) == ((1200 + 1200) * 8, (5000 + 1200) * 8)
    assert ragged.io_bytes(jobs, 8) == ((1200 + 1200) * 16, (5000 + 1200) * 16)
    plan = bb.Plan(ctypes.c_void_p(), None, 0, io_bytes=ragged.io_bytes(jobs, 4))
    x = torch.empty(6200, dtype=torch.complex64)
    with pytest.raises(ValueError, match="GPU tensor"):
        plan.execute(x, x)
    plan._h = None
