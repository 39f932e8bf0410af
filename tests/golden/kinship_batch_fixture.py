"""Regenerates the BED fixture of the reference's kinship batch test.

The reference test (tests/test_kinship_batch.cpp:21-44) writes n = 37 samples (partial last
byte) x m = 1031 SNPs with ``std::mt19937_64 rng(8128)``: SNPs with j % 31 == 0 are all code 3
(monomorphic), SNPs with j % 37 == 0 are all code 1 (all missing), every other SNP draws
``rng() % 4`` per 2-bit slot, padding slots of the last byte included.  std::mt19937_64 is fully
specified by the C++ standard, so the bytes are reproducible here; MT19937-64 below follows the
published algorithm (Matsumoto & Nishimura; Nishimura 2000).

The selections of the same test (:47-59) are exported too.
"""
import numpy as np

N_SAMPLES = 37
N_SNPS = 1031
SEED = 8128


from oracle import MT19937_64  # noqa: E402  (std::mt19937_64, restated in the oracle)


def make_bed():
    """Returns the BED body (no magic), SNP-major, ceil(37/4) = 10 bytes per SNP."""
    rng = MT19937_64(SEED)
    B = (N_SAMPLES + 3) // 4
    out = np.zeros(N_SNPS * B, dtype=np.uint8)
    for j in range(N_SNPS):
        for b in range(B):
            byte = 0
            for k in range(4):
                if j % 31 == 0:
                    code = 3
                elif j % 37 == 0:
                    code = 1
                else:
                    code = rng() % 4
                byte |= code << (2 * k)
            out[j * B + b] = byte
    return out


def selections():
    """The six (keep, variants) selections of test_kinship_batch.cpp:47-59 (None = all)."""
    n, m = N_SAMPLES, N_SNPS
    subset = [35, 1, 20, 3, 4, 31, 7, 8, 10, 12, 14, 16, 18, 22, 24, 26, 28, 30, 32, 34]
    identity = list(range(n))
    permuted = [n - 1 - i for i in range(n)]
    repeated = [i // 2 for i in range(n)]
    variants = [j for j in range(m) if j % 3 != 1]
    out = []
    for sel in range(6):
        keep = {1: subset, 3: identity, 4: permuted, 5: repeated}.get(sel)
        snps = None if sel == 0 else variants
        out.append((None if keep is None else np.array(keep, dtype=np.int32),
                    None if snps is None else np.array(snps, dtype=np.int32)))
    return out


if __name__ == "__main__":
    bed = make_bed()
    print(bed.size, int(bed.sum()))
