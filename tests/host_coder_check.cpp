// Host coder self-checks, built and run by test_host_coder.py (CPU only): every fast or parallel
// form of the host's rANS and MV-context coding gives the bytes of its plain serial form.
//  1. code_symbol_stream: groups of 16 and of 8 blocks, by tasks, against one serial writer;
//  2. RansWriter::flush_parallel (one task per block) against flush();
//  3. MvCoder::contexts (one table word per decision) against RansWriter::bit on Ctx states,
//     the per-decision rule of the reference's BinProb (rans.rs:30-84).
#include "../riv2_amd/csrc/entropy.hpp"
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

using namespace rivh;

static int failures = 0;
static void expect(bool ok, const char* what, size_t n) {
    if (!ok) { std::printf("FAIL %s (n=%zu)\n", what, n); failures++; }
}

// Tasks run one after another: the parallel forms must not depend on how the tasks are scheduled.
static void serial_for(int n, const std::function<void(int)>& f) {
    for (int i = n - 1; i >= 0; i--) f(i);   // reverse order on purpose
}

// A symbol as the coders buffer it: a binary decision at P(0) = p/256, or a bypass bit.
static uint16_t symbol(std::mt19937& g) {
    const uint32_t p = 1 + g() % 255, bit = g() & 1u;
    if (g() % 8 == 0) return bit ? 0x8080u : 0x80u;
    return (uint16_t)(bit ? ((p << 8) | (256u - p)) : p);
}

int main() {
    std::mt19937 g(12345);
    const size_t B = 16386;
    // 1 and 2: streams of 0 to 40 blocks, full and partial
    const size_t sizes[] = {0, 1, 1000, B - 1, B, B + 1, 2 * B, 3 * B + 17, 4 * B, 7 * B + 5, 8 * B, 9 * B + 3,
                            16 * B, 17 * B + 1, 24 * B, 33 * B + 99, 40 * B};
    for (size_t n : sizes) {
        std::vector<uint16_t> sym(n);
        for (auto& s : sym) s = symbol(g);
        const std::vector<uint8_t> plain = code_symbol_stream(sym.data(), n, ParallelFor());
        expect(code_symbol_stream(sym.data(), n, serial_for, 16) == plain, "code_symbol_stream, groups of 16", n);
        expect(code_symbol_stream(sym.data(), n, serial_for, 8) == plain, "code_symbol_stream, groups of 8", n);
        if (n <= 8 * B) {
            std::vector<uint8_t> a, b;
            {
                RansWriter w(a);
                uint16_t* o = w.reserve(n);
                for (size_t i = 0; i < n; i++) o[i] = sym[i];
                w.commit(o + n);
                w.flush();
            }
            {
                RansWriter w(b);
                uint16_t* o = w.reserve(n);
                for (size_t i = 0; i < n; i++) o[i] = sym[i];
                w.commit(o + n);
                w.flush_parallel(serial_for);
            }
            expect(a == b, "flush_parallel", n);
            // (an empty symbol sequence codes to no stream at all, while a writer emits its
            // two initial states: residual.rs:670-674 against rans.rs:231-269)
            if (n) expect(a == plain, "RansWriter against code_symbol_stream", n);
        }
    }
    // 3: random decisions over the 80 MV contexts, skewed so that counters saturate, over several
    // frames (the contexts carry from frame to frame)
    MvCoder mv;
    Ctx ref[256];
    for (int frame = 0; frame < 6; frame++) {
        const size_t n = 20000 + 9000 * (size_t)frame;
        std::vector<uint16_t> ev(n);
        for (auto& e : ev) {
            const uint32_t id = g() % 80;
            const uint32_t skew = id % 4 == 0 ? 250u : (id % 4 == 1 ? 6u : 128u);   // P(bit = 1) * 256
            const uint32_t bit = (g() & 255u) < skew ? 1u : 0u;
            e = (uint16_t)(id | (bit << 8));
        }
        std::vector<uint8_t> fast, slow;
        {
            RansWriter w(fast);
            mv.contexts(ev.data(), n, w);
            w.flush();
        }
        {
            RansWriter w(slow);
            for (uint16_t e : ev) w.bit((e >> 8) != 0, ref[e & 0xffu]);
            w.flush();
        }
        expect(fast == slow, "MvCoder::contexts against RansWriter::bit", n);
    }
    std::printf("%s\n", failures ? "host coder checks FAILED" : "host coder checks passed");
    return failures ? 1 : 0;
}
