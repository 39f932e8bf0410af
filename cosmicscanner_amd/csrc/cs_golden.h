// Golden-section minimisation shared by the REML searches of the dense null model
// (solveDiagonal, libCosmicCore/src/vce.cpp:3116-3140: log10 delta in [-5, 5]) and of the
// low-rank null model (src/lmm_engine.cpp:1040-1070: log eta around the best point of a grid).
// The two searches differ only in how the probe offset is formed from the bracket width (divided
// by the golden ratio, or multiplied by its inverse), in the stopping test (probe gap or bracket
// width) and in an iteration cap; those choices decide the optimum to the last bit, so they are
// parameters here.  Callable on the host and, with a block-uniform objective, on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace csamd {

struct GoldenSpec {
    double ratio;            // (sqrt 5 + 1) / 2 with divide, (sqrt 5 - 1) / 2 without
    bool divide;             // probe offset = width / ratio, else width * ratio
    double tol;
    int max_it;              // < 0: no cap
    bool stop_on_probe_gap;  // stop when |c - d| <= tol, else when hi - lo <= tol
};

// Minimises f on [lo, hi]; returns the midpoint of the final bracket.
template <class F>
__host__ __device__ inline double golden_section_min(F&& f, double lo, double hi, const GoldenSpec& g) {
    auto off = [&](double w) { return g.divide ? w / g.ratio : w * g.ratio; };
    double c = hi - off(hi - lo);
    double d = lo + off(hi - lo);
    double fc = f(c), fd = f(d);
    for (int it = 0; g.max_it < 0 || it < g.max_it; ++it) {
        if (g.stop_on_probe_gap ? fabs(c - d) <= g.tol : hi - lo <= g.tol) break;
        if (fc < fd) {
            hi = d;
            d = c;
            fd = fc;
            c = hi - off(hi - lo);
            fc = f(c);
        } else {
            lo = c;
            c = d;
            fc = fd;
            d = lo + off(hi - lo);
            fd = f(d);
        }
    }
    return 0.5 * (lo + hi);
}

}  // namespace csamd
