// Runs one generated expectation kernel (compiled from its text with the
// stand-in header, LQ_KERNEL_FILE) over a state on the CPU and prints the
// sum of its block partials.  argv: n dtype(0: complex128, 1: complex64)
// blocks threads state_file coef_file
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>  // the stand-in (tests/emul/hip)

#include LQ_KERNEL_FILE

int main(int argc, char** argv) {
    if (argc != 7) return 2;
    const int n = std::atoi(argv[1]);
    const int dtype = std::atoi(argv[2]);
    const unsigned blocks = (unsigned)std::atoi(argv[3]);
    const unsigned threads = (unsigned)std::atoi(argv[4]);
    const size_t N = 1ull << n;
    std::vector<double> amps(2 * N);
    FILE* f = std::fopen(argv[5], "rb");
    if (!f || std::fread(amps.data(), sizeof(double), 2 * N, f) != 2 * N) return 3;
    std::fclose(f);
    std::vector<double> coef(4096);
    f = std::fopen(argv[6], "rb");
    if (!f) return 4;
    const size_t nc = std::fread(coef.data(), sizeof(double), coef.size(), f);
    std::fclose(f);
    (void)nc;
    std::vector<double2> s128(N);
    std::vector<float2> s64(N);
    for (size_t i = 0; i < N; ++i) {
        s128[i] = {amps[2 * i], amps[2 * i + 1]};
        s64[i] = {(float)amps[2 * i], (float)amps[2 * i + 1]};
    }
    std::vector<double> partial(blocks, 0.0);
    gridDim = {blocks, 1, 1};
    blockDim = {threads, 1, 1};
    for (unsigned b = 0; b < blocks; ++b) {
        std::barrier<> bar((std::ptrdiff_t)threads);
        emu_barrier = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < threads; ++t)
            th.emplace_back([&, t, b] {
                threadIdx = {t, 0, 0};
                blockIdx = {b, 0, 0};
                if (dtype == 0)
                    lq_expect((const LQ_V*)s128.data(), coef.data(), partial.data(), 0, 0);
                else
                    lq_expect((const LQ_V*)s64.data(), coef.data(), partial.data(), 0, 0);
            });
        for (auto& x : th) x.join();
    }
    double sum = 0.0;
    for (double p : partial) sum += p;
    std::printf("%.17g\n", sum);
    return 0;
}
