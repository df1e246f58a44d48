"""bbfft-aot-generate: compile the kernels of FFT descriptors ahead of time (no GPU needed).

    python -m bbfft_amd.aot -o OUT_DIR [-d DEVICE_INFO] [-a ARCH] [-j JOBS] DESC... | @FILE

Reference: tools/aot/main.cpp:22-60 and tools/aot/args.cpp:17-102. hiprtc compiles one program
at a time per process, so shards of the kernel list are built by separate processes, each
writing OUT_DIR/kernels_<shard>.bin; a Cache loads the whole directory.
"""
import argparse
import os
import subprocess
import sys

_PKG_PARENT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def generate(descriptors, out_dir, device_info=None, arch="gfx950", jobs=8,
             ragged_serving=False):
    """Archives of every kernel the descriptors need; with ragged_serving also the persistent
    kernel of the seeded mixed serving workload (bbfft_amd.ragged.synthetic_serving_jobs) and
    its per-variant plain kernels, in OUT_DIR/kernels_ragged_serving.bin."""
    from . import GFX950_INFO
    device_info = device_info or GFX950_INFO
    os.makedirs(out_dir, exist_ok=True)
    for f in os.listdir(out_dir):
        if f.startswith("kernels_") and f.endswith(".bin"):
            os.remove(os.path.join(out_dir, f))
    listing = os.path.join(out_dir, "descriptors.txt")
    with open(listing, "w") as f:
        f.write("\n".join(descriptors) + "\n")
    procs = []
    env = dict(os.environ)
    env["PYTHONPATH"] = _PKG_PARENT + os.pathsep + env.get("PYTHONPATH", "")
    for shard in range(jobs):
        out = os.path.join(out_dir, f"kernels_{shard}.bin")
        cmd = [sys.executable, "-m", "bbfft_amd.aot", "--shard", str(shard), "--num-shards",
               str(jobs), "-d", device_info, "-a", arch, "--out-file", out, "@" + listing]
        procs.append(subprocess.Popen(cmd, env=env))
    if ragged_serving:
        out = os.path.join(out_dir, "kernels_ragged_serving.bin")
        cmd = [sys.executable, "-m", "bbfft_amd.aot", "--ragged-serving", "-d", device_info,
               "-a", arch, "--out-file", out]
        procs.append(subprocess.Popen(cmd, env=env))
    rc = [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError(f"AOT generation failed (exit codes {rc})")
    return sorted(os.path.join(out_dir, f) for f in os.listdir(out_dir)
                  if f.startswith("kernels_"))


def embed(archives, out_cpp, symbol="bbfft_aot_archive"):
    """Write a C++ source holding the archives (concatenated entries) as a byte array and a
    function `std::size_t <symbol>_register(bbfft::aot_cache &, int device)` registering them:
    link it into an executable to ship its kernels without archive files."""
    import re
    import struct
    if not re.fullmatch(r"[A-Za-z_][A-Za-z0-9_]*", symbol):
        raise ValueError("symbol must be a C identifier")
    entries, count = b"", 0
    for path in archives:
        with open(path, "rb") as f:
            data = f.read()
        if data[:8] != b"BBFFTAOT" or struct.unpack_from("<I", data, 8)[0] != 1:
            raise ValueError(f"not a version-1 AOT archive: {path}")
        count += struct.unpack_from("<I", data, 12)[0]
        entries += data[16:]
    blob = b"BBFFTAOT" + struct.pack("<II", 1, count) + entries
    lines = []
    for i in range(0, len(blob), 24):
        lines.append("    " + ", ".join(str(b) for b in blob[i:i + 24]) + ",")
    with open(out_cpp, "w") as f:
        f.write(f"""// Ahead-of-time FFT kernels for gfx950 linked into the executable (written by
// `python -m bbfft_amd.aot --embed`). Register them with {symbol}_register(cache, device).
#include "bbfft/aot_cache.hpp"
#include "bbfft/hip/online_compiler.hpp"

#include <cstddef>
#include <cstdint>

alignas(16) extern const std::uint8_t {symbol}[] = {{
""" + "\n".join(lines) + f"""
}};
extern const std::size_t {symbol}_size = {len(blob)};

std::size_t {symbol}_register(bbfft::aot_cache &cache, int device) {{
    return bbfft::hip::register_aot_archive(cache, {symbol}, {symbol}_size, device);
}}
""")
    return len(blob)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-o", "--out-dir")
    ap.add_argument("--out-file")
    ap.add_argument("-d", "--device-info", default=None)
    ap.add_argument("-a", "--arch", default="gfx950")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    ap.add_argument("--shard", type=int, default=None)
    ap.add_argument("--num-shards", type=int, default=1)
    ap.add_argument("--embed", metavar="OUT_CPP",
                    help="write the archives given as arguments (or OUT_DIR/kernels_*.bin) "
                         "into a C++ source to link into an executable")
    ap.add_argument("--symbol", default="bbfft_aot_archive", help="array name for --embed")
    ap.add_argument("--ragged-serving", action="store_true",
                    help="compile the ragged kernel of the seeded serving workload to --out-file")
    ap.add_argument("descriptors", nargs="*")
    a = ap.parse_args(argv)
    if a.embed:
        files = a.descriptors or sorted(
            os.path.join(a.out_dir, f) for f in os.listdir(a.out_dir)
            if f.startswith("kernels_") and f.endswith(".bin"))
        n = embed(files, a.embed, a.symbol)
        print(f"{len(files)} archives ({n} bytes) -> {a.embed}", flush=True)
        return 0
    if a.ragged_serving:
        from . import GFX950_INFO, ragged
        ragged.aot_generate(ragged.synthetic_serving_jobs(), a.out_file,
                            device_info=a.device_info or GFX950_INFO, arch=a.arch)
        print(f"ragged serving kernel -> {a.out_file}", flush=True)
        return 0
    descs = []
    for d in a.descriptors:
        if d.startswith("@"):
            with open(d[1:]) as f:
                descs += [line.strip() for line in f if line.strip()]
        else:
            descs.append(d)
    from . import GFX950_INFO, aot_generate_shard
    info = a.device_info or GFX950_INFO
    if a.shard is not None:
        n = aot_generate_shard(descs, a.out_file, a.shard, a.num_shards, info, a.arch)
        print(f"shard {a.shard}/{a.num_shards}: {n} kernels -> {a.out_file}", flush=True)
        return 0
    files = generate(descs, a.out_dir, info, a.arch, a.jobs)
    print(f"{len(files)} archives in {a.out_dir}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
