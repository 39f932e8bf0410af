// Exact Hardy-Weinberg test (Wigginton, Cutler and Abecasis, Am. J. Hum. Genet. 76:887-893,
// 2005), as the GRM SNP filter applies it (calculateHWE, libCosmicCore/src/qc_utils.cpp:9-77;
// grm_builder.cpp:220-223): the heterozygote-count distribution given the allele counts by its
// two-sided recurrence from the mode, then the mass of the outcomes no more likely than the one
// observed.  The recurrences are evaluated in the published order so that a p-value at a filter
// threshold comes out the same as the reference's.
#pragma once
#include <algorithm>
#include <vector>

namespace csamd {

inline double hwe_exact_p(long long hom1, long long hets, long long hom2) {
    const long long homr = std::min(hom1, hom2);
    const long long rare = 2 * homr + hets;
    const long long n = hom1 + hets + hom2;
    if (n == 0) return 1.0;
    long long mid = rare * (2 * n - rare) / (2 * n);
    if ((mid % 2) != (rare % 2)) ++mid;
    std::vector<double> prob((size_t)rare + 1, 0.0);
    prob[(size_t)mid] = 1.0;
    double total = 1.0;
    {
        long long r = (rare - mid) / 2, c = n - mid - r;
        for (long long h = mid; h >= 2; h -= 2) {
            prob[(size_t)(h - 2)] = prob[(size_t)h] * h * (h - 1.0) / (4.0 * (r + 1.0) * (c + 1.0));
            total += prob[(size_t)(h - 2)];
            ++r;
            ++c;
        }
    }
    {
        long long r = (rare - mid) / 2, c = n - mid - r;
        for (long long h = mid; h <= rare - 2; h += 2) {
            prob[(size_t)(h + 2)] = prob[(size_t)h] * 4.0 * r * c / ((h + 2.0) * (h + 1.0));
            total += prob[(size_t)(h + 2)];
            --r;
            --c;
        }
    }
    const double target = prob[(size_t)hets];
    double p = 0.0;
    for (long long h = rare % 2; h <= rare; h += 2)
        if (prob[(size_t)h] <= target) p += prob[(size_t)h];
    return std::min(1.0, p / total);
}

}  // namespace csamd
