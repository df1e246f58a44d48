"""Python mirror of the bbfft C++ API over the gfx950 C ABI (include/bbfft_hip.h).

Names follow the reference's C++ interface (include/bbfft/configuration.hpp,
include/bbfft/parser.hpp, include/bbfft/plan.hpp) so tests read like the reference's tests:

    cfg = parse_fft_descriptor("scfo8.256*32768")
    plan = make_plan(cfg)                # current torch stream, current device
    plan.execute(x, y)                   # device pointers or torch tensors; asynchronous

PyTorch (when present) is only used for device pointers and the current stream.
"""
import ctypes
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

from . import _capi

__all__ = [
    "precision", "direction", "transform_type", "kernel_language", "configuration",
    "user_module", "parse_fft_descriptor", "default_istride", "default_ostride", "make_plan",
    "Plan", "Cache", "BadConfiguration", "BbfftError", "DescriptorError", "kernel_configuration",
    "generate_source", "aot_generate", "aot_generate_shard", "device_info_query", "MAX_FFT_DIM",
    "MAX_TENSOR_DIM", "default_cache", "root_of_unity", "factor", "num_passes", "compile_kernel",
    "normalize_device_info", "GFX950_INFO", "adjoint_configuration", "make_adjoint_plan",
]

MAX_FFT_DIM = 3
MAX_TENSOR_DIM = MAX_FFT_DIM + 2


class precision(enum.IntEnum):
    f32 = 4
    f64 = 8


class direction(enum.IntEnum):
    forward = -1
    backward = 1


class transform_type(enum.IntEnum):
    c2c = 0
    r2c = 1
    c2r = 2


class kernel_language(enum.IntEnum):
    opencl_c = 0
    hip = 1


class BbfftError(RuntimeError):
    """HIP runtime / hiprtc failure or internal error."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class BadConfiguration(BbfftError):
    """Unsupported or faulty configuration (bbfft::bad_configuration)."""


class DescriptorError(BbfftError):
    """Malformed descriptor (std::runtime_error from the parser)."""


def _check(status):
    if status != _capi.SUCCESS:
        msg = _capi.load().bbfft_last_error().decode(errors="replace")
        if status == _capi.ERROR_BAD_CONFIGURATION:
            raise BadConfiguration(status, msg)
        if status == _capi.ERROR_PARSE:
            raise DescriptorError(status, msg)
        if status == _capi.ERROR_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise BbfftError(status, msg)


@dataclass
class user_module:
    """Load/store callbacks (include/bbfft/user_module.hpp of the reference)."""
    data: Optional[str] = None
    load_function: Optional[str] = None
    store_function: Optional[str] = None
    language: kernel_language = kernel_language.hip

    def __bool__(self):
        return bool(self.data) and (self.load_function is not None
                                    or self.store_function is not None)


def _arr5(v):
    v = list(v) + [0] * (5 - len(v))
    return (ctypes.c_size_t * 5)(*v[:5])


def default_istride(dim, shape, type_, inplace):
    out = (ctypes.c_size_t * 5)()
    _check(_capi.load().bbfft_default_istride(dim, _arr5(shape), int(type_), int(inplace), out))
    return list(out)


def default_ostride(dim, shape, type_, inplace):
    out = (ctypes.c_size_t * 5)()
    _check(_capi.load().bbfft_default_ostride(dim, _arr5(shape), int(type_), int(inplace), out))
    return list(out)


@dataclass
class configuration:
    """Mirror of bbfft::configuration; strides default to the in-place layout like the
    reference's default member initialisers (include/bbfft/configuration.hpp:168-179)."""
    dim: int
    shape: Sequence[int]
    fp: precision = precision.f32
    dir: direction = direction.forward
    type: transform_type = transform_type.c2c
    istride: Optional[Sequence[int]] = None
    ostride: Optional[Sequence[int]] = None
    callbacks: user_module = field(default_factory=user_module)

    def __post_init__(self):
        self.shape = list(self.shape) + [0] * (5 - len(self.shape))
        if self.istride is None:
            self.istride = default_istride(self.dim, self.shape, self.type, True)
        if self.ostride is None:
            self.ostride = default_ostride(self.dim, self.shape, self.type, True)
        self.istride = list(self.istride) + [0] * (5 - len(self.istride))
        self.ostride = list(self.ostride) + [0] * (5 - len(self.ostride))

    def set_strides_default(self, inplace):
        self.istride = default_istride(self.dim, self.shape, self.type, inplace)
        self.ostride = default_ostride(self.dim, self.shape, self.type, inplace)

    def _to_c(self):
        c = _capi.CConfiguration()
        c.dim = self.dim
        c.shape = _arr5(self.shape)
        c.precision = int(self.fp)
        c.direction = int(self.dir)
        c.type = int(self.type)
        c.istride = _arr5(self.istride)
        c.ostride = _arr5(self.ostride)
        cb = self.callbacks
        if cb:
            src = cb.data.encode()
            self._keep = src  # keep the buffer alive while the C struct is in use
            c.callback_source = src
            c.callback_length = len(src)
            c.load_function = cb.load_function.encode() if cb.load_function else None
            c.store_function = cb.store_function.encode() if cb.store_function else None
            c.callback_language = int(cb.language)
        return c

    def to_string(self):
        buf = ctypes.create_string_buffer(256)
        _check(_capi.load().bbfft_configuration_to_string(ctypes.byref(self._to_c()), buf, 256))
        return buf.value.decode()

    def __str__(self):
        return self.to_string()


def parse_fft_descriptor(desc: str) -> configuration:
    c = _capi.CConfiguration()
    _check(_capi.load().bbfft_parse_fft_descriptor(desc.encode(), ctypes.byref(c)))
    return configuration(dim=c.dim, shape=list(c.shape), fp=precision(c.precision),
                         dir=direction(c.direction), type=transform_type(c.type),
                         istride=list(c.istride), ostride=list(c.ostride))


def normalize_device_info(desc: str) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(_capi.load().bbfft_device_info_normalize(desc.encode(), buf, 256))
    return buf.value.decode()


# What hipDeviceProp_t reports for MI355X (maxThreadsPerBlock, warpSize, sharedMemPerBlock).
GFX950_INFO = "{1024, {64}, 163840, gpu}"


def root_of_unity(k, n, precision_=8) -> complex:
    """exp(2 pi i k / n) correctly rounded to the given precision (as emitted in kernels)."""
    re, im = ctypes.c_double(), ctypes.c_double()
    _check(_capi.load().bbfft_root_of_unity(k, n, int(precision_), ctypes.byref(re),
                                            ctypes.byref(im)))
    return complex(re.value, im.value)


def factor(n, index):
    """factor(n, index) of the library's tiling logic."""
    out = (ctypes.c_uint * max(1, index))()
    _check(_capi.load().bbfft_factor(n, index, out))
    return list(out)[:index]


def num_passes(cfg: configuration) -> int:
    n = ctypes.c_uint()
    _check(_capi.load().bbfft_num_passes(ctypes.byref(cfg._to_c()), ctypes.byref(n)))
    return n.value


def tuning_struct(t):
    """dict(factorization=[...], Nt=, Mb=, Kb=, lds_pad=, no_pad=, flags=) -> CTuning."""
    c = _capi.CTuning()
    fac = list(t.get("factorization", []))
    c.num_factors = len(fac)
    for i, f in enumerate(fac):
        c.factorization[i] = f
    for k in ("Nt", "Mb", "Kb", "lds_pad"):
        setattr(c, k, int(t.get(k, 0)))
    c.no_pad = int(bool(t.get("no_pad", False)))
    c.waves_per_eu = int(t.get("waves_per_eu", 0))
    c.groups_per_cu = int(t.get("groups_per_cu", 0))
    if "flags" in t:
        c.flags = int(t["flags"])
        c.has_flags = 1
    return c


def kernel_configuration(cfg: configuration, pass_index=0, device_info=GFX950_INFO,
                         tuning=None):
    """Tiling of the kernel of 1-D pass `pass_index` (dict) plus its identifier."""
    kc = _capi.CKernelConfig()
    name = ctypes.create_string_buffer(512)
    if tuning is not None:
        _check(_capi.load().bbfft_kernel_configuration_tuned(
            ctypes.byref(cfg._to_c()), ctypes.byref(tuning_struct(tuning)),
            device_info.encode(), ctypes.byref(kc), name, 512))
    else:
        _check(_capi.load().bbfft_kernel_configuration(ctypes.byref(cfg._to_c()), pass_index,
                                                       device_info.encode(), ctypes.byref(kc),
                                                       name, 512))
    d = {f: getattr(kc, f) for f, _ in _capi.CKernelConfig._fields_ if f != "factorization"}
    d["factorization"] = list(kc.factorization)[:kc.num_factors]
    d["name"] = name.value.decode()
    return d


def generate_source(cfg: configuration, pass_index=0, device_info=GFX950_INFO,
                    tuning=None) -> str:
    L = _capi.load()
    c = cfg._to_c()
    n = ctypes.c_size_t()
    if tuning is not None:
        t = tuning_struct(tuning)
        _check(L.bbfft_generate_source_tuned(ctypes.byref(c), ctypes.byref(t),
                                             device_info.encode(), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(L.bbfft_generate_source_tuned(ctypes.byref(c), ctypes.byref(t),
                                             device_info.encode(), buf, n.value, None))
        return buf.value.decode()
    _check(L.bbfft_generate_source(ctypes.byref(c), pass_index, device_info.encode(), None, 0,
                                   ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    _check(L.bbfft_generate_source(ctypes.byref(c), pass_index, device_info.encode(), buf,
                                   n.value, None))
    return buf.value.decode()


def compile_kernel(cfg: configuration, pass_index=0, device_info=GFX950_INFO,
                   arch="gfx950") -> int:
    """Compile the kernel of one pass (callbacks included) with hiprtc; returns its size."""
    n = ctypes.c_size_t()
    _check(_capi.load().bbfft_compile_kernel(ctypes.byref(cfg._to_c()), pass_index,
                                             device_info.encode(), arch.encode(),
                                             ctypes.byref(n)))
    return n.value


def aot_generate_shard(descriptors: List[str], out_path: str, shard=0, num_shards=1,
                       device_info=GFX950_INFO, arch="gfx950") -> int:
    """Compile shard `shard` of the kernels of `descriptors` into one archive (no GPU)."""
    arr = (ctypes.c_char_p * len(descriptors))(*[d.encode() for d in descriptors])
    n = ctypes.c_size_t()
    _check(_capi.load().bbfft_aot_generate(arr, len(descriptors), device_info.encode(),
                                           arch.encode(), out_path.encode(), shard, num_shards,
                                           ctypes.byref(n)))
    return n.value


def aot_generate(descriptors: List[str], out_dir: str, device_info=GFX950_INFO, arch="gfx950",
                 jobs=8) -> List[str]:
    """bbfft-aot-generate: archives out_dir/kernels_<i>.bin built by `jobs` processes."""
    from .aot import generate
    return generate(descriptors, out_dir, device_info=device_info, arch=arch, jobs=jobs)


def device_info_query(device=0) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(_capi.load().bbfft_device_info_query(device, buf, 256))
    return buf.value.decode()


class Cache:
    """jit_cache_all layered over an aot_cache (see bbfft_cache_create)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        _check(_capi.load().bbfft_cache_create(ctypes.byref(self._h)))

    def load_aot(self, path, device=0):
        """Register one archive file, or every kernels_*.bin archive of a directory."""
        import glob
        import os
        paths = (sorted(glob.glob(os.path.join(path, "kernels_*.bin"))) if os.path.isdir(path)
                 else [path])
        total = 0
        for p in paths:
            n = ctypes.c_size_t()
            _check(_capi.load().bbfft_cache_load_aot(self._h, p.encode(), device,
                                                     ctypes.byref(n)))
            total += n.value
        return total

    def __len__(self):
        n = ctypes.c_size_t()
        _check(_capi.load().bbfft_cache_size(self._h, ctypes.byref(n)))
        return n.value

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            lib = _capi._lib
        except (AttributeError, NameError):  # interpreter shutdown cleared the module
            return
        if lib is not None and getattr(self, "_h", None) and self._h.value:
            lib.bbfft_cache_destroy(self._h)
            self._h = ctypes.c_void_p()


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError("expected a device pointer (int) or a tensor with data_ptr()")


def _current_stream():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.current_stream().cuda_stream
    except ImportError:
        pass
    return 0


def _io_bytes(cfg, in_elem_bytes=None):
    """Bytes from the first element to the last one a configuration addresses, (input,
    output)."""
    rb = int(cfg.fp)
    t = int(cfg.type)
    ishape, oshape = list(cfg.shape), list(cfg.shape)
    if t == 2:
        ishape[1] = cfg.shape[1] // 2 + 1
    if t == 1:
        oshape[1] = cfg.shape[1] // 2 + 1

    def extent(shape, stride):
        return sum((shape[i] - 1) * stride[i] for i in range(cfg.dim + 2)) + 1

    ieb = in_elem_bytes or (rb if t == 1 else 2 * rb)
    oeb = rb if t == 2 else 2 * rb
    return extent(ishape, cfg.istride) * ieb, extent(oshape, cfg.ostride) * oeb


def _check_buffer(x, need, what):
    """A tensor argument must live on the GPU and hold `need` bytes from its data pointer on:
    the kernels address the whole configured tensor, and a smaller buffer would be read or
    written out of bounds. Plain integer pointers are passed through unchecked."""
    if not hasattr(x, "untyped_storage"):
        return
    if not x.is_cuda:
        raise ValueError(f"{what} must be a GPU tensor (got device {x.device})")
    avail = x.untyped_storage().nbytes() - x.storage_offset() * x.element_size()
    if avail < need:
        raise ValueError(f"{what} holds {avail} bytes from its data pointer, the configuration "
                         f"addresses {need}")


class Plan:
    """Mirror of plan_unmanaged_event<hipEvent_t> (include/bbfft/plan.hpp:144-184)."""

    def __init__(self, handle, cfg, stream, in_elem_bytes=None, io_bytes=None):
        self._h = handle
        self.cfg = cfg
        # bytes the input and output span (None: not checked, callbacks interpret offsets);
        # io_bytes gives them directly for plans without a configuration (ragged batches)
        self._io = _io_bytes(cfg, in_elem_bytes) if cfg is not None and not cfg.callbacks \
            else io_bytes
        self.stream = stream

    def execute(self, inp, out=None, signal=None, wait: Sequence = (), static_scratch=False):
        """Out-of-place if out is given, else in-place. Events are hipEvent_t handles (int).
        Tensors are checked for device and size (see _check_buffer); integer pointers are not.
        static_scratch=True runs bbfft_plan_execute_static_scratch (allocates nothing; plans
        needing a scratch buffer use the library's 64 MiB per-GPU array)."""
        if self._io is not None:
            ib, ob = self._io
            if out is None:
                _check_buffer(inp, max(ib, ob), "the in-place buffer")
            else:
                _check_buffer(inp, ib, "the input")
                _check_buffer(out, ob, "the output")
        src = _ptr(inp)
        dst = src if out is None else _ptr(out)
        nw = len(wait)
        warr = (ctypes.c_void_p * max(1, nw))(*[int(w) for w in wait]) if nw else None
        L = _capi.load()
        fn = L.bbfft_plan_execute_static_scratch if static_scratch else L.bbfft_plan_execute
        _check(fn(self._h, src, dst, signal, nw, warr))

    def launches(self):
        L = _capi.load()
        n = ctypes.c_uint()
        _check(L.bbfft_plan_num_launches(self._h, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            name = ctypes.create_string_buffer(512)
            g, b, lds = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_size_t()
            _check(L.bbfft_plan_launch_info(self._h, i, name, 512, ctypes.byref(g),
                                            ctypes.byref(b), ctypes.byref(lds)))
            out.append({"kernel": name.value.decode(), "grid": g.value, "block": b.value,
                        "lds_bytes": lds.value})
        return out

    def __del__(self):
        try:
            lib = _capi._lib
        except (AttributeError, NameError):  # interpreter shutdown cleared the module
            return
        if lib is not None and getattr(self, "_h", None):
            lib.bbfft_plan_destroy(self._h)
            self._h = None


def adjoint_configuration(cfg: configuration) -> configuration:
    """Configuration of the adjoint transform of cfg (see bbfft_adjoint_configuration)."""
    c = _capi.CConfiguration()
    _check(_capi.load().bbfft_adjoint_configuration(ctypes.byref(cfg._to_c()), ctypes.byref(c)))
    return configuration(dim=c.dim, shape=list(c.shape), fp=precision(c.precision),
                         dir=direction(c.direction), type=transform_type(c.type),
                         istride=list(c.istride), ostride=list(c.ostride))


def make_adjoint_plan(cfg, stream=None, cache: Optional["Cache"] = None) -> "Plan":
    """Plan computing the adjoint of cfg's transform (gradient of a real loss)."""
    if isinstance(cfg, str):
        cfg = parse_fft_descriptor(cfg)
    if stream is None:
        stream = _current_stream()
    h = ctypes.c_void_p()
    _check(_capi.load().bbfft_plan_create_adjoint(ctypes.byref(cfg._to_c()), stream,
                                                  cache.handle if cache is not None else None,
                                                  ctypes.byref(h)))
    return Plan(h, adjoint_configuration(cfg), stream)


def fused_configuration(cfg: configuration, device_info=GFX950_INFO):
    """The fused pair of passes of a 2-D/3-D configuration on a device (None if not fused)."""
    fused = ctypes.c_int()
    info = _capi.CFusedInfo()
    name = ctypes.create_string_buffer(256)
    _check(_capi.load().bbfft_fused_configuration(ctypes.byref(cfg._to_c()),
                                                  device_info.encode(), ctypes.byref(fused),
                                                  ctypes.byref(info), name, 256))
    if not fused.value:
        return None
    out = {k: getattr(info, k) for k, _ in _capi.CFusedInfo._fields_}
    out["name"] = name.value.decode()
    return out


def generate_fused_source(cfg: configuration, device_info=GFX950_INFO) -> str:
    L = _capi.load()
    n = ctypes.c_size_t()
    c = cfg._to_c()
    _check(L.bbfft_generate_fused_source(ctypes.byref(c), device_info.encode(), None, 0,
                                         ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    _check(L.bbfft_generate_fused_source(ctypes.byref(c), device_info.encode(), buf, n.value,
                                         ctypes.byref(n)))
    return buf.value.decode()


INPUT_F16 = _capi.INPUT_F16
INPUT_BF16 = _capi.INPUT_BF16
MFMA = _capi.MFMA
PREFETCH = _capi.PREFETCH  # software-pipelined tile loop (fft1d flag)


def half_input_flags(storage="f16", mfma=True) -> int:
    """Kernel-variant flags of a c2c fp32 plan reading "f16" or "bf16" (re, im) pairs."""
    if storage not in ("f16", "bf16"):
        raise ValueError("storage must be 'f16' or 'bf16'")
    return (INPUT_F16 if storage == "f16" else INPUT_BF16) | (MFMA if mfma else 0)


def half_input_kernel_configuration(cfg, storage="f16", mfma=True, device_info=GFX950_INFO):
    """Kernel tiling of a half-input plan (tuning table applied), like kernel_configuration."""
    if isinstance(cfg, str):
        cfg = parse_fft_descriptor(cfg)
    kc = _capi.CKernelConfig()
    name = ctypes.create_string_buffer(512)
    _check(_capi.load().bbfft_half_input_kernel_configuration(
        ctypes.byref(cfg._to_c()), half_input_flags(storage, mfma), device_info.encode(),
        ctypes.byref(kc), name, 512))
    return {"M": kc.M, "Mb": kc.Mb, "N": kc.N, "K": kc.K, "Nt": kc.Nt, "Kb": kc.Kb,
            "lds_pad": kc.lds_pad, "lds_bytes": kc.lds_bytes,
            "work_group_size": kc.work_group_size,
            "factorization": [kc.factorization[i] for i in range(kc.num_factors)],
            "name": name.value.decode()}


def make_half_input_plan(cfg, storage="f16", mfma=True, stream=None,
                         cache: Optional["Cache"] = None) -> "Plan":
    """1-D c2c fp32 plan whose input is stored as half-precision (re, im) pairs (4 bytes per
    complex element, strides as in cfg) and whose output is fp32 complex; with mfma the radix-16
    first stage runs on the matrix cores (split DFT matrix, fp32 accumulation)."""
    if isinstance(cfg, str):
        cfg = parse_fft_descriptor(cfg)
    if stream is None:
        stream = _current_stream()
    h = ctypes.c_void_p()
    _check(_capi.load().bbfft_plan_create_half_input(
        ctypes.byref(cfg._to_c()), half_input_flags(storage, mfma), stream,
        cache.handle if cache is not None else None, ctypes.byref(h)))
    return Plan(h, cfg, stream, in_elem_bytes=4)


DEFAULT_AOT_DIR = os.path.join(os.path.dirname(_capi.DEFAULT_LIB), "aot")
_default_cache = None


def default_cache(device=0) -> "Cache":
    """Process-wide cache preloaded with the archives in lib/aot (if they exist)."""
    global _default_cache
    if _default_cache is None:
        c = Cache()
        if os.path.isdir(DEFAULT_AOT_DIR):
            c.load_aot(DEFAULT_AOT_DIR, device)
        _default_cache = c
    return _default_cache


def make_plan(cfg, stream=None, cache: Optional[Cache] = None, tuning=None) -> Plan:
    """make_plan(cfg, queue, cache) for the current HIP device; cfg may be a descriptor.
    `tuning` (dict, 1-D only) overrides the tiling heuristics."""
    if isinstance(cfg, str):
        cfg = parse_fft_descriptor(cfg)
    if stream is None:
        stream = _current_stream()
    h = ctypes.c_void_p()
    ch = cache.handle if cache is not None else None
    if tuning is not None:
        _check(_capi.load().bbfft_plan_create_tuned(ctypes.byref(cfg._to_c()),
                                                    ctypes.byref(tuning_struct(tuning)), stream,
                                                    ch, ctypes.byref(h)))
    else:
        _check(_capi.load().bbfft_plan_create(ctypes.byref(cfg._to_c()), stream, ch,
                                              ctypes.byref(h)))
    # a tuning with half-input bits reads (re, im) half pairs: 4 bytes per input element
    half = tuning is not None and int(tuning.get("flags", 0)) & (INPUT_F16 | INPUT_BF16)
    return Plan(h, cfg, stream, in_elem_bytes=4 if half else None)
