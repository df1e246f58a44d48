// hiprtc compilation, code-object loading and the AOT archive.
// Plays the role of the reference's src/ze/online_compiler.cpp:18-228 and
// src/cl/online_compiler.cpp:30-123 (build from source / native binary, build log on failure).
#include "bbfft/hip/online_compiler.hpp"
#include "aot_archive.hpp"
#include "bbfft/hip/device.hpp"
#include "bbfft/hip/error.hpp"

#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <unistd.h>

namespace bbfft::hip {

namespace {

std::string program_log(hiprtcProgram prog) {
    std::size_t n = 0;
    if (hiprtcGetProgramLogSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
        return {};
    }
    std::string log(n, '\0');
    hiprtcGetProgramLog(prog, log.data());
    return log;
}

std::uint64_t fnv1a(std::string const &s, std::uint64_t h = 0xcbf29ce484222325ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 0x100000001b3ull;
    }
    return h;
}

/// Directory of the on-disk code-object cache (BBFFT_CACHE_DIR), or empty when disabled.
std::string cache_dir() {
    char const *d = std::getenv("BBFFT_CACHE_DIR");
    return d && *d ? std::string(d) : std::string();
}

std::vector<std::uint8_t> read_file(std::string const &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        return {};
    }
    return std::vector<std::uint8_t>(std::istreambuf_iterator<char>(f), {});
}

void write_file_atomic(std::string const &path, std::vector<std::uint8_t> const &data) {
    std::string tmp = path + ".tmp" + std::to_string(::getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) {
            return; // cache is best effort
        }
        f.write(reinterpret_cast<char const *>(data.data()),
                static_cast<std::streamsize>(data.size()));
    }
    std::rename(tmp.c_str(), path.c_str());
}

void release_module(module_handle_t m) { (void)hipModuleUnload(reinterpret_cast<hipModule_t>(m)); }

} // namespace

/// Options of every kernel compile. Multiply-adds are contracted within a source expression
/// only, so a kernel's arithmetic depends on its generated expressions and not on the code it
/// is inlined into (the ragged kernel then computes each job exactly like the plain kernel).
/// Packed fp32 instructions are disabled.
static std::vector<std::string> base_options(std::string const &arch) {
    return {"--offload-arch=" + arch, "-O3", "-std=c++17", "-ffp-contract=on", "-Xclang",
            "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

std::vector<std::uint8_t> compile_to_native(std::string const &source, std::string const &arch,
                                            std::vector<std::string> const &options) {
    std::vector<std::string> opts = base_options(arch);
    opts.insert(opts.end(), options.begin(), options.end());
    std::vector<char const *> copts;
    for (auto const &o : opts) {
        copts.push_back(o.c_str());
    }
    hiprtcProgram prog;
    check(hiprtcCreateProgram(&prog, source.c_str(), "bbfft_kernel.hip", 0, nullptr, nullptr),
          "hiprtcCreateProgram");
    auto st = hiprtcCompileProgram(prog, static_cast<int>(copts.size()), copts.data());
    if (st != HIPRTC_SUCCESS) {
        auto log = program_log(prog);
        hiprtcDestroyProgram(&prog);
        check(st, "hiprtcCompileProgram", log);
    }
    std::size_t size = 0;
    check(hiprtcGetCodeSize(prog, &size), "hiprtcGetCodeSize");
    std::vector<std::uint8_t> code(size);
    check(hiprtcGetCode(prog, reinterpret_cast<char *>(code.data())), "hiprtcGetCode");
    hiprtcDestroyProgram(&prog);
    return code;
}

hipModule_t build_kernel_bundle(std::uint8_t const *binary, std::size_t binary_size,
                                module_format format, int device) {
    if (format != module_format::native) {
        throw error("Only native code objects can be loaded on HIP devices", 0);
    }
    (void)binary_size;
    int cur = 0;
    BBFFT_HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) {
        BBFFT_HIP_CHECK(hipSetDevice(device));
    }
    hipModule_t mod = nullptr;
    auto st = hipModuleLoadData(&mod, binary);
    if (cur != device) {
        (void)hipSetDevice(cur);
    }
    check(st, "hipModuleLoadData");
    return mod;
}

hipModule_t build_kernel_bundle(std::string const &source, int device,
                                std::vector<std::string> const &options) {
    std::string const arch = get_device_arch(device);
    std::string const dir = cache_dir();
    std::string path;
    if (!dir.empty()) {
        int major = 0, minor = 0;
        hiprtcVersion(&major, &minor);
        std::string key = arch + '\n' + std::to_string(major) + '.' + std::to_string(minor);
        for (auto const &o : base_options(arch)) {
            key += '\n' + o;
        }
        for (auto const &o : options) {
            key += '\n' + o;
        }
        char name[64];
        std::snprintf(name, sizeof(name), "/%016llx.co",
                      static_cast<unsigned long long>(fnv1a(source, fnv1a(key))));
        path = dir + name;
        auto cached = read_file(path);
        if (!cached.empty()) {
            return build_kernel_bundle(cached.data(), cached.size(), module_format::native,
                                       device);
        }
    }
    auto code = compile_to_native(source, arch, options);
    if (!path.empty()) {
        write_file_atomic(path, code);
    }
    return build_kernel_bundle(code.data(), code.size(), module_format::native, device);
}

hipFunction_t create_kernel(hipModule_t mod, std::string const &name) {
    hipFunction_t fn = nullptr;
    auto st = hipModuleGetFunction(&fn, mod, name.c_str());
    if (st != hipSuccess) {
        check(st, ("hipModuleGetFunction(" + name + ")").c_str());
    }
    return fn;
}

aot_module create_aot_module(std::uint8_t const *binary, std::size_t binary_size,
                             module_format format, std::vector<std::string> const &kernel_names,
                             int device) {
    auto mod = build_kernel_bundle(binary, binary_size, format, device);
    aot_module m;
    m.mod = shared_handle<module_handle_t>(reinterpret_cast<module_handle_t>(mod),
                                           &release_module);
    m.kernel_names.insert(kernel_names.begin(), kernel_names.end());
    m.device_id = get_device_id(device);
    return m;
}

namespace {
std::size_t register_entries(aot_cache &cache, std::vector<archive_entry> const &entries,
                             std::string const &path, int device) {
    std::string const arch = get_device_arch(device);
    std::size_t n = 0;
    for (auto const &e : entries) {
        if (e.arch != arch) {
            throw error("AOT archive " + path + " was built for " + e.arch + ", device is " + arch,
                        0);
        }
        cache.register_module(create_aot_module(e.code.data(), e.code.size(),
                                                module_format::native, e.kernel_names, device));
        n += e.kernel_names.size();
    }
    return n;
}
} // namespace

std::size_t register_aot_archive(aot_cache &cache, std::string const &path, int device) {
    return register_entries(cache, read_archive(path), path, device);
}

std::size_t register_aot_archive(aot_cache &cache, std::uint8_t const *data, std::size_t size,
                                 int device) {
    return register_entries(cache, parse_archive(data, size, "(memory)"), "(memory)", device);
}

// ---- archive ------------------------------------------------------------------------------

namespace {
void put_u32(std::ostream &os, std::uint32_t v) { os.write(reinterpret_cast<char *>(&v), 4); }
void put_u64(std::ostream &os, std::uint64_t v) { os.write(reinterpret_cast<char *>(&v), 8); }
void put_str(std::ostream &os, std::string const &s) {
    put_u32(os, static_cast<std::uint32_t>(s.size()));
    os.write(s.data(), static_cast<std::streamsize>(s.size()));
}
struct reader {
    std::uint8_t const *d;
    std::size_t size;
    std::size_t p = 0;
    std::string const &path;
    void need(std::size_t n) {
        if (n > size || p > size - n) {
            throw error("Truncated AOT archive " + path, 0);
        }
    }
    std::uint32_t u32() {
        need(4);
        std::uint32_t v;
        std::memcpy(&v, d + p, 4);
        p += 4;
        return v;
    }
    std::uint64_t u64() {
        need(8);
        std::uint64_t v;
        std::memcpy(&v, d + p, 8);
        p += 8;
        return v;
    }
    std::string str() {
        auto n = u32();
        need(n);
        std::string s(reinterpret_cast<char const *>(d + p), n);
        p += n;
        return s;
    }
};
} // namespace

void write_archive(std::string const &path, std::vector<archive_entry> const &entries) {
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        throw error("Cannot write AOT archive " + path, 0);
    }
    f.write("BBFFTAOT", 8);
    put_u32(f, 1);
    put_u32(f, static_cast<std::uint32_t>(entries.size()));
    for (auto const &e : entries) {
        put_str(f, e.arch);
        put_u32(f, static_cast<std::uint32_t>(e.kernel_names.size()));
        for (auto const &n : e.kernel_names) {
            put_str(f, n);
        }
        put_u64(f, e.code.size());
        f.write(reinterpret_cast<char const *>(e.code.data()),
                static_cast<std::streamsize>(e.code.size()));
    }
}

std::vector<archive_entry> read_archive(std::string const &path) {
    auto data = read_file(path);
    return parse_archive(data.data(), data.size(), path);
}

std::vector<archive_entry> parse_archive(std::uint8_t const *data, std::size_t size,
                                         std::string const &what) {
    if (!data || size < 16 || std::memcmp(data, "BBFFTAOT", 8) != 0) {
        throw error("Not an AOT archive: " + what, 0);
    }
    reader r{data, size, 8, what};
    if (r.u32() != 1) {
        throw error("Unsupported AOT archive version: " + what, 0);
    }
    auto count = r.u32();
    std::vector<archive_entry> out;
    for (std::uint32_t c = 0; c < count; ++c) {
        archive_entry e;
        e.arch = r.str();
        auto nn = r.u32();
        for (std::uint32_t i = 0; i < nn; ++i) {
            e.kernel_names.push_back(r.str());
        }
        auto sz = r.u64();
        r.need(sz);
        e.code.assign(data + r.p, data + r.p + sz);
        r.p += sz;
        out.push_back(std::move(e));
    }
    return out;
}

} // namespace bbfft::hip
