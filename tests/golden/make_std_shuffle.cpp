// Known answers of the standard-library shuffle the matrix-free GWAS driver draws its calibration
// SNPs with (std::shuffle over std::mt19937_64(seed), the seed of --lmm-low-rank-seed): the
// shuffled order of 0..n-1 for a few lengths and seeds.
// Build and run:  g++ -O2 -std=c++17 make_std_shuffle.cpp -o /tmp/mk && /tmp/mk > std_shuffle_mt64.txt
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

int main() {
    for (unsigned long long seed : {42ULL, 7ULL}) {
        for (int n : {1, 2, 5, 16, 33, 200, 1001}) {
            std::vector<int> v((size_t)n);
            std::iota(v.begin(), v.end(), 0);
            std::mt19937_64 gen(seed);
            std::shuffle(v.begin(), v.end(), gen);
            std::printf("seed %llu n %d\n", seed, n);
            for (int i = 0; i < n; ++i) std::printf("%d%c", v[(size_t)i], i + 1 == n ? '\n' : ' ');
        }
    }
    return 0;
}
