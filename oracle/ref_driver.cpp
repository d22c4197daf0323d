// Test infrastructure: C entry points over the reference's own host FL loops
// run_afl_compress_cpu<T,1> / run_afl_decompress_cpu<T,1>
// (reference src/compression/afl/afl_gpu.cu:6-65), compiled from the
// reference sources by oracle/ref_build.py into oracle/_ref/libddj_ref.so.
#include <cstddef>
#include <cstdint>

#include "compression/afl/afl_gpu.cuh"

extern "C" {

int ref_fl_pack(int width_bytes, int bits, const void* in, unsigned long n, void* out, unsigned long words) {
    switch (width_bytes) {
    case 4: run_afl_compress_cpu<int, 1>(bits, (int*)in, (int*)out, n, words); return 0;
    case 8: run_afl_compress_cpu<long, 1>(bits, (long*)in, (long*)out, n, words); return 0;
    case 2: run_afl_compress_cpu<short, 1>(bits, (short*)in, (short*)out, n, words); return 0;
    case 1: run_afl_compress_cpu<char, 1>(bits, (char*)in, (char*)out, n, words); return 0;
    default: return -1;
    }
}

int ref_fl_unpack(int width_bytes, int bits, const void* in, unsigned long n, void* out) {
    switch (width_bytes) {
    case 4: run_afl_decompress_cpu<int, 1>(bits, (int*)in, (int*)out, n); return 0;
    case 8: run_afl_decompress_cpu<long, 1>(bits, (long*)in, (long*)out, n); return 0;
    case 2: run_afl_decompress_cpu<short, 1>(bits, (short*)in, (short*)out, n); return 0;
    case 1: run_afl_decompress_cpu<char, 1>(bits, (char*)in, (char*)out, n); return 0;
    default: return -1;
    }
}
}
