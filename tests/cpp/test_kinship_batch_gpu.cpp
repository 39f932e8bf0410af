// C++ parity test of the MI355X kinship operator through the C++ adapter, restating the checks
// of the reference's tests/test_kinship_batch.cpp: the same fixture (n = 37 with a partial last
// byte, m = 1031, std::mt19937_64(8128), j%31 monomorphic, j%37 all missing), the six sample /
// variant selections, widths 1 / 7 / 41, trace against the dense oracle, apply vs dense z z^T v/M,
// applyTranspose == apply, aliasing, bitwise repeat, empty RHS and the dimension error.
// The dense oracle is built from an independent decode of the same bytes.
//
// Tolerance: the int8-MFMA operator quantises each right-hand-side column to 23-bit fixed point,
// so apply is checked at max|Y - K X| <= 2e-6 max|K X| per column (the reference's FP64 CPU
// operator uses 1e-10); trace is FP64 and checked at 1e-10.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <vector>

#include "cosmicscanner_amd_kinship.hpp"

static void require(bool v, const char* msg) {
    if (!v) throw std::runtime_error(msg);
}

int main() {
    try {
        const int n = 37, m = 1031, B = (n + 3) / 4;
        std::mt19937_64 rng(8128);
        std::vector<uint8_t> bed((size_t)m * B);
        for (int j = 0; j < m; ++j)
            for (int b = 0; b < B; ++b) {
                unsigned char byte = 0;
                for (int k = 0; k < 4; ++k) {
                    unsigned char code = j % 31 == 0 ? 3 : j % 37 == 0 ? 1 : rng() % 4;
                    byte |= code << (2 * k);
                }
                bed[(size_t)j * B + b] = byte;
            }
        // independent decode + all-sample stats (plink_reader.cpp:187-239)
        auto dosage = [&](int j, int i) -> double {
            const int c = (bed[(size_t)j * B + (i >> 2)] >> (2 * (i & 3))) & 3;
            return c == 0 ? 2.0 : c == 2 ? 1.0 : c == 3 ? 0.0 : NAN;
        };
        std::vector<double> mean(m), inv(m);
        for (int j = 0; j < m; ++j) {
            double s = 0;
            int cnt = 0;
            for (int i = 0; i < n; ++i) {
                const double d = dosage(j, i);
                if (!std::isnan(d)) { s += d; ++cnt; }
            }
            mean[j] = cnt ? s / cnt : 0.0;
            const double p = mean[j] / 2.0, var = 2.0 * p * (1.0 - p);
            inv[j] = (cnt == 0 || var < 1e-9) ? 0.0 : 1.0 / std::sqrt(var);
        }
        cosmic_amd::GpuPlinkGenotypes geno(bed.data(), n, m);

        std::vector<int> subset{35, 1, 20, 3, 4, 31, 7, 8, 10, 12, 14, 16, 18, 22, 24, 26, 28, 30, 32, 34};
        std::vector<int> identity(n), permuted(n), repeated(n), variants;
        for (int i = 0; i < n; ++i) { identity[i] = i; permuted[i] = n - 1 - i; repeated[i] = i / 2; }
        for (int j = 0; j < m; ++j) if (j % 3 != 1) variants.push_back(j);
        std::normal_distribution<double> normal;
        for (int sel = 0; sel < 6; ++sel) {
            const std::vector<int> keep = sel == 1 ? subset : sel == 3 ? identity :
                sel == 4 ? permuted : sel == 5 ? repeated : std::vector<int>{};
            const std::vector<int> snps = sel == 0 ? std::vector<int>{} : variants;
            cosmic_amd::GpuKinshipOperator op(&geno, keep, snps);
            const int r = op.rows(), M = op.numVariants();
            std::vector<double> z((size_t)r * M);
            for (int t = 0; t < M; ++t) {
                const int j = snps.empty() ? t : snps[t];
                for (int i = 0; i < r; ++i) {
                    const double d = dosage(j, keep.empty() ? i : keep[i]);
                    z[(size_t)t * r + i] = std::isnan(d) ? 0.0 : (d - mean[j]) * inv[j];
                }
            }
            double tr = 0;
            for (double v : z) tr += v * v;
            tr /= M;
            require(std::fabs(op.trace() - tr) < 1e-10 * std::fabs(tr), "trace mismatch");
            for (int width : {1, 7, 41}) {
                std::vector<double> X((size_t)r * width);
                for (double& v : X) v = normal(rng);
                // dense oracle Z (Z^T X) / M
                std::vector<double> ref((size_t)r * width, 0.0);
                for (int c = 0; c < width; ++c)
                    for (int t = 0; t < M; ++t) {
                        double u = 0;
                        for (int i = 0; i < r; ++i) u += z[(size_t)t * r + i] * X[(size_t)c * r + i];
                        for (int i = 0; i < r; ++i) ref[(size_t)c * r + i] += z[(size_t)t * r + i] * u / M;
                    }
                std::vector<double> Y;
                op.applyBlock(X.data(), r, width, Y);
                for (int c = 0; c < width; ++c) {
                    double mx = 0, err = 0;
                    for (int i = 0; i < r; ++i) {
                        mx = std::fmax(mx, std::fabs(ref[(size_t)c * r + i]));
                        err = std::fmax(err, std::fabs(Y[(size_t)c * r + i] - ref[(size_t)c * r + i]));
                    }
                    require(err <= 2e-6 * mx, "apply mismatch vs dense oracle");
                }
                std::vector<double> again;
                op.applyBlock(X.data(), r, width, again);
                require(again == Y, "repeat is not bitwise identical");
                std::vector<double> v(X.begin(), X.begin() + r), a1, a2;
                op.apply(v.data(), r, a1);
                op.applyTranspose(v.data(), r, a2);
                require(a1 == a2, "applyTranspose != apply");
                op.apply(v.data(), r, v);          // aliasing out == in
                require(v == a1, "aliased apply differs");
            }
            std::vector<double> empty_out;
            op.applyBlock(nullptr, r, 0, empty_out);
            require(empty_out.empty(), "empty RHS shape mismatch");
            bool rejected = false;
            try {
                std::vector<double> w((size_t)r + 1, 1.0), o;
                op.apply(w.data(), r + 1, o);
            } catch (const std::runtime_error&) {
                rejected = true;
            }
            require(rejected, "bad dimensions not rejected");
            bool bad_threads = false;
            try {
                std::vector<double> o;
                op.applyBlock(z.data(), r, 1, o, 0);
            } catch (const std::invalid_argument&) {
                bad_threads = true;
            }
            require(bad_threads, "non-positive threads not rejected");
        }
        std::printf("BATCH_KOMEGA_PASS\n");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
