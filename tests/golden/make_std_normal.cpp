// Known answers of the standard-library generators the reference's low-rank sketch draws from
// (std::mt19937_64(seed) with std::normal_distribution<double>(0, 1), src/lmm_engine.cpp:916-921):
// the first values for seeds 42 and 7, printed with 17 significant digits.
// Build and run:  g++ -O2 -std=c++17 make_std_normal.cpp -o /tmp/mk && /tmp/mk > std_normal_mt64.txt
#include <cstdio>
#include <random>

int main() {
    for (unsigned long long seed : {42ULL, 7ULL}) {
        std::mt19937_64 gen(seed);
        std::normal_distribution<double> normal(0.0, 1.0);
        std::printf("seed %llu\n", seed);
        for (int i = 0; i < 257; ++i) std::printf("%.17g\n", normal(gen));
    }
    return 0;
}
