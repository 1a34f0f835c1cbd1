// Runs one transform of the engine on the CPU emulation of HIP (hip/hip_runtime.h beside this file):
//   emu_main dft  f32|f64 n batch inverse in.bin out.bin
//   emu_main real f32|f64 n fmt batch inverse in.bin out.bin
//   emu_main dct  f32|f64 n batch inverse in.bin out.bin
//   emu_main prog f32|f64 n want inverse in.bin out.bin   (progressive units of one transform)
//   emu_main ws   f32|f64 n batch inverse mb in.bin out.bin (execute with an mb-MiB caller workspace)
//   emu_main half bf16|f16 n batch inverse in.bin out.bin   (half-precision storage, 4-byte values)
// The input is copied into a buffer of exactly its size and the output buffer holds exactly the
// output, so under AddressSanitizer every access beyond either is reported.
#include "../../kfr_amd/csrc/fft_engine.hpp"
#include "../../kfr_amd/csrc/fft_pow2.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace kfr_amd;

static std::vector<char> read_file(const char* path)
{
    std::vector<char> v;
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return v;
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0)
        v.insert(v.end(), buf, buf + r);
    std::fclose(f);
    return v;
}

template <typename T>
static int run(const std::string& kind, int argc, char** argv)
{
    const size_t n = std::strtoull(argv[3], nullptr, 10);
    int a          = 4;
    if (kind == "ws")
    {
        const long long batch = std::atoll(argv[4]);
        const bool inverse    = std::atoi(argv[5]) != 0;
        const size_t wbytes   = (size_t)std::atoll(argv[6]) << 20;
        const std::vector<char> in = read_file(argv[7]);
        const size_t bytes         = (size_t)batch * n * 2 * sizeof(T);
        if (in.size() != bytes)
            return 2;
        char* din  = static_cast<char*>(std::malloc(bytes));
        char* dout = static_cast<char*>(std::malloc(bytes));
        void* ws   = std::malloc(wbytes);
        std::memcpy(din, in.data(), bytes);
        try
        {
            gpu_dft<T>(n).execute_workspace(dout, din, batch, inverse, nullptr, workspace{ ws, wbytes });
        }
        catch (const std::exception& e)
        {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 3;
        }
        FILE* f = std::fopen(argv[8], "wb");
        if (!f || std::fwrite(dout, 1, bytes, f) != bytes)
            return 4;
        std::fclose(f);
        std::free(din);
        std::free(dout);
        std::free(ws);
        return 0;
    }
    if (kind == "prog")
    {
        const int want     = std::atoi(argv[4]);
        const bool inverse = std::atoi(argv[5]) != 0;
        const std::vector<char> in = read_file(argv[6]);
        const size_t bytes         = n * 2 * sizeof(T);
        if (in.size() != bytes)
            return 2;
        char* d = static_cast<char*>(std::malloc(bytes));
        std::memcpy(d, in.data(), bytes);
        try
        {
            gpu_dft<T> f(n);
            const int units = f.progressive_prepare(want);
            void* mid       = f.progressive_mid_bytes() ? std::malloc(f.progressive_mid_bytes()) : nullptr;
            for (int u = 0; u < units; ++u)
                f.progressive_unit(u, d, mid, inverse, nullptr);
            std::free(mid);
            std::fprintf(stderr, "units=%d\n", units);
        }
        catch (const std::exception& e)
        {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 3;
        }
        FILE* f = std::fopen(argv[7], "wb");
        if (!f || std::fwrite(d, 1, bytes, f) != bytes)
            return 4;
        std::fclose(f);
        std::free(d);
        return 0;
    }
    int fmt        = 1;
    if (kind == "real")
        fmt = std::atoi(argv[a++]);
    const long long batch = std::atoll(argv[a++]);
    const bool inverse    = std::atoi(argv[a++]) != 0;
    if (argc != a + 2)
        return 2;
    const std::vector<char> in = read_file(argv[a]);
    size_t in_bytes = 0, out_bytes = 0;
    const size_t cs = kind == "real" ? (fmt == 1 ? n / 2 + 1 : n / 2) : 0;
    if (kind == "dft")
        in_bytes = out_bytes = (size_t)batch * n * 2 * sizeof(T);
    else if (kind == "dct")
        in_bytes = out_bytes = (size_t)batch * n * sizeof(T);
    else
    {
        in_bytes  = (size_t)batch * (inverse ? cs * 2 : n) * sizeof(T);
        out_bytes = (size_t)batch * (inverse ? n : cs * 2) * sizeof(T);
    }
    if (in.size() != in_bytes)
    {
        std::fprintf(stderr, "input holds %zu bytes, expected %zu\n", in.size(), in_bytes);
        return 2;
    }
    char* din  = static_cast<char*>(std::malloc(in_bytes));
    char* dout = static_cast<char*>(std::malloc(out_bytes));
    std::memcpy(din, in.data(), in_bytes);
    std::memset(dout, 0x7f, out_bytes);
    try
    {
        if (kind == "dft")
            gpu_dft<T>(n).execute(dout, din, batch, inverse, nullptr);
        else if (kind == "dct")
            gpu_dct<T>(n).execute(dout, din, batch, inverse, nullptr);
        else if (inverse)
            gpu_dft_real<T>(n, fmt).inverse(dout, din, batch, nullptr);
        else
            gpu_dft_real<T>(n, fmt).forward(dout, din, batch, nullptr);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    FILE* f = std::fopen(argv[a + 1], "wb");
    if (!f || std::fwrite(dout, 1, out_bytes, f) != out_bytes)
        return 4;
    std::fclose(f);
    std::free(din);
    std::free(dout);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 4)
    {
        std::fprintf(stderr, "usage: emu_main dft|real|dct f32|f64 n [fmt] batch inverse in out\n");
        return 2;
    }
    const std::string kind = argv[1], t = argv[2];
    if (kind == "half")
    {
        if (argc != 8)
            return 2;
        const size_t n        = std::strtoull(argv[3], nullptr, 10);
        const long long batch = std::atoll(argv[4]);
        const bool inverse    = std::atoi(argv[5]) != 0;
        const std::vector<char> in = read_file(argv[6]);
        const size_t bytes         = (size_t)batch * n * 4;
        if (in.size() != bytes)
            return 2;
        char* din  = static_cast<char*>(std::malloc(bytes));
        char* dout = static_cast<char*>(std::malloc(bytes));
        std::memcpy(din, in.data(), bytes);
        try
        {
            pow2_fft<float>(n, nullptr, t == "bf16" ? 1 : 2).run(inverse, din, dout, batch, nullptr);
        }
        catch (const std::exception& e)
        {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 3;
        }
        FILE* f = std::fopen(argv[7], "wb");
        if (!f || std::fwrite(dout, 1, bytes, f) != bytes)
            return 4;
        std::fclose(f);
        std::free(din);
        std::free(dout);
        return 0;
    }
    if (kind != "dft" && kind != "real" && kind != "dct" && kind != "prog" && kind != "ws")
        return 2;
    return t == "f64" ? run<double>(kind, argc, argv) : run<float>(kind, argc, argv);
}
