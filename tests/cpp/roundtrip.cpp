// C++ use of the engine through include/ddj.hpp: encoding, tree, optimizer
// and column API on synthetic columns; exits non-zero on any mismatch.
#include <cstdio>
#include <cstring>
#include <vector>

#include "ddj.hpp"

static int fails = 0;
#define EXPECT(c)                                                   \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "FAILED %s:%d %s\n", __FILE__, __LINE__, #c); \
            fails++;                                                \
        }                                                           \
    } while (0)

int main(int argc, char** argv) {
    const char* tmp = argc > 1 ? argv[1] : "roundtrip.ddj";
    ddj::Context ctx(0);
    const size_t n = 1 << 20;
    std::vector<int> ts(n);
    unsigned s = 1;
    int t = 1400000000;
    for (size_t i = 0; i < n; i++) {
        s = s * 1103515245u + 12345u;
        t += 1 + (int)((s >> 16) % 7);
        ts[i] = t;
    }
    ddj::DeviceBuffer d = ctx.upload(ts.data(), n * 4);
    // Encoding
    ddj::Encoding delta(ctx, DDJ_ENC_DELTA);
    auto outs = delta.Encode(d, DDJ_D_INT);
    EXPECT(outs.size() == 2);
    auto back = delta.Decode(outs, DDJ_D_INT).copyToHost();
    EXPECT(back.size() == n * 4 && std::memcmp(back.data(), ts.data(), n * 4) == 0);
    // CompressionTree
    ddj::CompressionTree tree(ctx, {DDJ_ENC_DELTA, DDJ_ENC_AFL, DDJ_ENC_NONE}, DDJ_D_INT);
    auto blob = tree.Compress(d);
    EXPECT(blob.size() == tree.GetPredictedSizeAfterCompression(d));
    EXPECT(tree.GetCompressionRatio() > 4.0);
    auto dec = ddj::CompressionTree(ctx).Decompress(blob).copyToHost();
    EXPECT(dec.size() == n * 4 && std::memcmp(dec.data(), ts.data(), n * 4) == 0);
    // CompressionOptimizer
    ddj::CompressionOptimizer opt(ctx);
    auto b2 = opt.CompressData(d, DDJ_D_INT);
    EXPECT(!opt.GetOptimalTree().empty());
    auto dec2 = ddj::CompressionTree(ctx).Decompress(b2).copyToHost();
    EXPECT(std::memcmp(dec2.data(), ts.data(), n * 4) == 0);
    // ParallelTSCompressor
    std::vector<double> px(n);
    for (size_t i = 0; i < n; i++) px[i] = (10000 + (int)(i % 97) - 48) / 100.0;
    ddj::ParallelTSCompressor pts(ctx, {DDJ_D_INT, DDJ_D_DOUBLE}, 300000);
    pts.Compress({ts.data(), px.data()}, n, tmp);
    size_t row = 0;
    bool same = true;
    pts.Decompress(tmp, [&](const std::vector<void*>& cols, size_t rows) {
        same = same && std::memcmp(cols[0], ts.data() + row, rows * 4) == 0;
        same = same && std::memcmp(cols[1], px.data() + row, rows * 8) == 0;
        row += rows;
    });
    EXPECT(same && row == n);
    if (fails) return 1;
    std::printf("cpp roundtrip ok\n");
    return 0;
}
