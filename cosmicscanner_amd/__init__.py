"""MI355X-native genotype hot path of CosmicScanner: host-side mirror over the HIP C-ABI.

The compute lives in ``libcsamd.so`` (hand-written HIP for gfx950, ``csrc/``); this package only
binds its C entry points (``include/cosmicscanner_amd.h``) with ctypes and mirrors the
reference's interfaces for the path:

* :class:`PlinkReader`  -- ``libCosmicCore/src/plink_reader.h`` (load, stats, class counts), with
  the packed matrix resident in HBM;
* :class:`PlinkKinshipOperator` -- ``src/kinship_operator.h:19-93`` (apply / applyTranspose /
  trace / getDiagonal / rows / cols / numVariants) and :func:`createKinshipOperator`;
* :class:`GrmBuilder` -- ``libCosmicCore/src/grm_builder.h:15-80`` (compute / save / getGrm /
  getSum2pq / getTotalSnps) for the default VanRaden algorithm;
* :func:`scan_sums` / :func:`lmm_scan` -- the per-SNP association scan.

Importing fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsamd.so")

CS_OK = 0
CS_ERR_INVALID_ARGUMENT = -1
CS_ERR_RUNTIME = -2
CS_PREC_INT8_L3 = 3
CS_PREC_INT8_L2 = 2
CS_PREC_FP32 = 32
CS_PREC_INT8_L4 = 4

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_f32p = _fp
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"cosmicscanner_amd: HIP library not built ({LIB_PATH}); run "
                          "`python -m cosmicscanner_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "cs_last_error": ([], ctypes.c_char_p),
        "cs_version": ([], ctypes.c_int),
        "cs_dev_alloc": ([_i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_dev_free": ([_vp], ctypes.c_int),
        "cs_dev_release_cache": ([], ctypes.c_int),
        "cs_dev_upload": ([_vp, _vp, _i64], ctypes.c_int),
        "cs_dev_download": ([_vp, _vp, _i64], ctypes.c_int),
        "cs_dev_synchronize": ([], ctypes.c_int),
        "cs_dev_copy": ([_vp, _vp, _i64], ctypes.c_int),
        "cs_bed_create": ([_u8p, _i64, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_bed_create_from_file_image": ([_u8p, _i64, _i64, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_bed_create_synthetic": ([ctypes.c_uint64, _i64, _i64, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_bed_destroy": ([_vp], ctypes.c_int),
        "cs_bed_dims": ([_vp, _i64p, _i64p], ctypes.c_int),
        "cs_bed_download": ([_vp, _u8p], ctypes.c_int),
        "cs_bed_download_snps": ([_vp, _i32p, _i64, _u8p], ctypes.c_int),
        "cs_bed_build_sample_major": ([_vp], ctypes.c_int),
        "cs_bed_device_bytes": ([_vp, _i64p], ctypes.c_int),
        "cs_bed_compute_stats": ([_vp], ctypes.c_int),
        "cs_bed_get_stats": ([_vp, _dp, _dp], ctypes.c_int),
        "cs_bed_class_counts": ([_vp, _i32p, _i64p], ctypes.c_int),
        "cs_kinship_create": ([_vp, _i32p, _i64, _i32p, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_kinship_destroy": ([_vp], ctypes.c_int),
        "cs_kinship_rows": ([_vp, _i64p], ctypes.c_int),
        "cs_kinship_num_variants": ([_vp, _i64p], ctypes.c_int),
        "cs_kinship_set_precision": ([_vp, ctypes.c_int], ctypes.c_int),
        "cs_kinship_set_missing_mode": ([_vp, ctypes.c_int], ctypes.c_int),
        "cs_kinship_compact_rows": ([_vp], ctypes.c_int),
        "cs_kinship_device_bytes": ([_vp, _i64p], ctypes.c_int),
        "cs_kinship_apply": ([_vp, _dp, _i64, _i64, _dp, _i64], ctypes.c_int),
        "cs_kinship_apply_device": ([_vp, _vp, _i64, _i64, _vp, _i64], ctypes.c_int),
        "cs_kinship_apply_device_noalloc": ([_vp, _vp, _i64, _i64, _vp, _i64], ctypes.c_int),
        "cs_validate_variance_components": ([_dp, _i64, ctypes.c_char_p], ctypes.c_int),
        "cs_kinship_multiply_zt": ([_vp, _dp, _i64, _i64, _dp, _i64], ctypes.c_int),
        "cs_kinship_multiply_z": ([_vp, _dp, _i64, _i64, _dp, _i64], ctypes.c_int),
        "cs_kinship_multiply_zt_device": ([_vp, _vp, _i64, _i64, _vp, _i64], ctypes.c_int),
        "cs_kinship_multiply_z_device": ([_vp, _vp, _i64, _i64, _vp, _i64], ctypes.c_int),
        "cs_kinship_trace": ([_vp, _dp], ctypes.c_int),
        "cs_kinship_diag": ([_vp, _dp], ctypes.c_int),
        "cs_kinship_last_timing": ([_vp, _dp], ctypes.c_int),
        "cs_scan_sums": ([_vp, _i32p, _i64, _dp, _i64, _i64, ctypes.c_double, _i32p, _i64,
                          ctypes.c_int, _dp, _i64p, _i32p, _dp, _dp], ctypes.c_int),
        "cs_lmm_scan": ([_vp, _i32p, _i64, _dp, _i64, _i64, ctypes.c_double, _i32p, _i64,
                         ctypes.c_int, ctypes.c_void_p, _dp, _i64p, _i32p, _dp, _dp, _dp, _dp],
                        ctypes.c_int),
        "cs_lmm_scan_device": ([_vp, _i32p, _i64, _vp, _i64, _i64, ctypes.c_double, _i32p, _i64,
                                ctypes.c_int, ctypes.c_void_p, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                               ctypes.c_int),
        "cs_scan_last_timing": ([_dp, _dp, _dp, _i64p], ctypes.c_int),
        "cs_sym_eigen_device": ([_vp, _i64, _i64, _vp, _i64], ctypes.c_int),
        "cs_grm_build": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _fp, _dp, _dp, _i64p], ctypes.c_int),
        "cs_grm_build_device": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _vp, _dp, _i64p],
                                ctypes.c_int),
        "cs_grm_last_timing": ([_dp, _dp, ctypes.POINTER(ctypes.c_int), _dp], ctypes.c_int),
        "cs_grm_build_raw": ([_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _i32p, _i64, _vp, ctypes.c_int,
                              _dp, _dp, _i64p], ctypes.c_int),
        "cs_grm_loco_subset": ([_vp, _vp, _i64, _i32p, _i64, ctypes.c_double, ctypes.c_double, _vp],
                               ctypes.c_int),
        "cs_lmm_create": ([_dp, _f32p, _i64, _i64, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_lmm_destroy": ([_vp], ctypes.c_int),
        "cs_lmm_prepare_dense": ([_vp, _vp, _i64, ctypes.c_int], ctypes.c_int),
        "cs_lmm_null_brent": ([_vp, _dp, _dp, _dp], ctypes.c_int),
        "cs_lmm_set_variance": ([_vp, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "cs_lmm_get_eigen": ([_vp, _dp, _dp, _dp, _i64], ctypes.c_int),
        "cs_lmm_scan_rotated": ([_vp, _vp, _i32p, _i64, _i32p, _i64, ctypes.c_double, ctypes.c_int,
                                 _dp, _i64p, _i32p, _dp, _dp, _dp, _dp], ctypes.c_int),
        "cs_pcg_solve": ([_vp, ctypes.c_double, ctypes.c_double, _vp, _i64, _i64, _vp, _i64,
                          ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int), _dp], ctypes.c_int),
        "cs_pcg_solve_cols": ([_vp, _dp, _dp, _vp, _i64, _i64, _vp, _i64, ctypes.c_double,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                               _dp], ctypes.c_int),
        "cs_markers_device": ([_vp, _i32p, _i64, _i32p, _i64, ctypes.c_double, _vp, _i64, _i32p, _dp],
                              ctypes.c_int),
        "cs_pcg_projected_scan": ([_vp, _vp, _i32p, _i64, _dp, _i64, _i64, _f32p, _i64, _i64,
                                   ctypes.c_double, ctypes.c_double, _i32p, _i64, _i32p, _i64,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp, _i64p, _i32p, _dp, _dp, _dp, _dp, ctypes.c_void_p],
                                  ctypes.c_int),
        "cs_lowrank_prepare": ([_vp, _dp, _f32p, _i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp),
                                ctypes.c_void_p], ctypes.c_int),
        "cs_lowrank_destroy": ([_vp], ctypes.c_int),
        "cs_lowrank_get": ([_vp, _i64p, _dp, _dp, _i64, _dp, _dp, _dp], ctypes.c_int),
        "cs_lowrank_scan": ([_vp, _vp, _i32p, _i64, _i32p, _i64, ctypes.c_double, ctypes.c_int, _dp, _i64p,
                             _i32p, _dp, _dp, _dp, _dp], ctypes.c_int),
        "cs_gwas_matrix_free_run": ([ctypes.c_char_p, _dp, _dp, _i64, _i64, ctypes.c_void_p,
                                     ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
        "cs_server_create": ([_vp, _i32p, _i64, _i64, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_server_destroy": ([_vp], ctypes.c_int),
        "cs_server_submit": ([_vp, ctypes.c_void_p, _i64p], ctypes.c_int),
        "cs_server_step": ([_vp, _i32p, _i32p, _i32p], ctypes.c_int),
        "cs_server_result": ([_vp, _i64, _i32p, _dp, _i64p, _i32p, _dp, _dp, _dp, _dp, ctypes.c_void_p],
                             ctypes.c_int),
        "cs_server_release": ([_vp, _i64], ctypes.c_int),
        "cs_gwas_dense_run": ([ctypes.c_char_p, _dp, _dp, _i64, _i64, ctypes.c_void_p,
                               ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
        "cs_mvlmm_create": ([_dp, _i64, _i64, _i64, _f32p, _i64, _i64, ctypes.POINTER(_vp)], ctypes.c_int),
        "cs_mvlmm_destroy": ([_vp], ctypes.c_int),
        "cs_mvlmm_prepare_dense": ([_vp, _vp, _i64, ctypes.c_int], ctypes.c_int),
        "cs_mvlmm_null": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                           ctypes.c_double, _i32p, _i32p, _i32p], ctypes.c_int),
        "cs_mvlmm_get": ([_vp, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
        "cs_mvlmm_set_variance": ([_vp, _dp, _dp], ctypes.c_int),
        "cs_mvlmm_get_eigen": ([_vp, _dp, _dp, _dp, _i64], ctypes.c_int),
        "cs_mvlmm_scan": ([_vp, _vp, _i32p, _i64, _i32p, _i64, _dp, _dp, _dp, _dp, _dp, _i32p], ctypes.c_int),
        "cs_gwas_mv_run": ([ctypes.c_char_p, _dp, _i64, _i64, _dp, _i64, _i64, ctypes.c_void_p,
                            ctypes.c_char_p, ctypes.c_void_p], ctypes.c_int),
        "cs_grm_loco_subset_scaled": ([_vp, _vp, _i64, _i32p, _i64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _vp], ctypes.c_int),
    }
    for name, (argt, rest) in sig.items():
        f = getattr(L, name)
        f.argtypes = argt
        f.restype = rest
    _lib = L
    return L


def lib():
    return _load()


class CsError(RuntimeError):
    pass


def _check(rc):
    if rc == CS_OK:
        return
    msg = lib().cs_last_error().decode()
    if rc == CS_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class DeviceArray:
    """A float64 column-major (n, k) buffer in HBM, owned by this object."""

    def __init__(self, n, k=1):
        self.n, self.k = int(n), int(k)
        self.nbytes = self.n * self.k * 8
        p = _vp()
        _check(lib().cs_dev_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.asarray(a, dtype=np.float64)
        a2 = a[:, None] if a.ndim == 1 else a
        d = cls(a2.shape[0], a2.shape[1])
        d.upload(a2)
        return d

    def upload(self, a):
        f = np.asfortranarray(np.asarray(a, dtype=np.float64).reshape(self.n, self.k, order="A"))
        _check(lib().cs_dev_upload(self.ptr, f.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        out = np.empty((self.n, self.k), order="F")
        _check(lib().cs_dev_download(out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return np.ascontiguousarray(out)

    def close(self):
        if getattr(self, "ptr", None):
            _check(lib().cs_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    _check(lib().cs_dev_synchronize())


def read_plink_text(prefix):
    with open(prefix + ".fam") as f:
        fam = [ln.split() for ln in f if ln.strip()]
    with open(prefix + ".bim") as f:
        bim = [ln.split() for ln in f if ln.strip()]
    return bim, fam


class PlinkReader:
    """PLINK BED genotypes resident in HBM (PlinkReader, libCosmicCore/src/plink_reader.h:30-127).

    Construct with :meth:`load` (prefix.{bed,bim,fam}), :meth:`from_body` (the bytes after the
    magic) or :meth:`synthetic` (generated on the device)."""

    def __init__(self, handle, n_samples, n_snps, bim=None, fam=None):
        self._h = handle
        self.num_samples = int(n_samples)
        self.num_snps = int(n_snps)
        self.bim = bim
        self.fam = fam

    # construction ------------------------------------------------------------------
    @classmethod
    def load(cls, prefix):
        bim, fam = read_plink_text(prefix)
        image = np.fromfile(prefix + ".bed", dtype=np.uint8)
        h = _vp()
        _check(lib().cs_bed_create_from_file_image(_p(image, _u8p), image.size, len(fam), len(bim),
                                                   ctypes.byref(h)))
        return cls(h, len(fam), len(bim), bim, fam)

    @classmethod
    def from_body(cls, body, n_samples, n_snps):
        body = np.ascontiguousarray(body, dtype=np.uint8)
        if body.size < n_snps * ((n_samples + 3) // 4):
            raise RuntimeError("genotype buffer shorter than n_snps * ceil(n_samples/4)")
        h = _vp()
        _check(lib().cs_bed_create(_p(body, _u8p), n_samples, n_snps, ctypes.byref(h)))
        return cls(h, n_samples, n_snps)

    @classmethod
    def synthetic(cls, seed, n_samples, n_snps, pmin, pmax, miss):
        h = _vp()
        _check(lib().cs_bed_create_synthetic(seed, n_samples, n_snps, pmin, pmax, miss,
                                             ctypes.byref(h)))
        return cls(h, n_samples, n_snps)

    def close(self):
        if self._h:
            _check(lib().cs_bed_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference interface ----------------------------------------------------------
    def getNumSamples(self):
        return self.num_samples

    def getNumSnps(self):
        return self.num_snps

    def getBytesPerSnp(self):
        return (self.num_samples + 3) // 4

    def computeStats(self, n_threads=1):
        _check(lib().cs_bed_compute_stats(self._h))

    def getMeans(self):
        return self._stats()[0]

    def getInvStds(self):
        return self._stats()[1]

    def _stats(self):
        m = np.empty(self.num_snps)
        s = np.empty(self.num_snps)
        _check(lib().cs_bed_get_stats(self._h, _p(m, _dp), _p(s, _dp)))
        return m, s

    def class_counts(self, weights=None):
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int32)
        if w is not None and w.size != self.num_samples:
            raise ValueError("weights must have one entry per sample")
        out = np.empty((self.num_snps, 4), dtype=np.int64)
        _check(lib().cs_bed_class_counts(self._h, _p(w, _i32p), _p(out, _i64p)))
        return out

    def download(self):
        out = np.empty(self.num_snps * self.getBytesPerSnp(), dtype=np.uint8)
        _check(lib().cs_bed_download(self._h, _p(out, _u8p)))
        return out

    def download_snps(self, snps):
        """Packed body of the listed SNPs (row i = SNP snps[i])."""
        snps = np.ascontiguousarray(snps, dtype=np.int32)
        out = np.empty(snps.size * self.getBytesPerSnp(), dtype=np.uint8)
        _check(lib().cs_bed_download_snps(self._h, _p(snps, _i32p), snps.size, _p(out, _u8p)))
        return out

    def device_bytes(self):
        v = ctypes.c_int64()
        _check(lib().cs_bed_device_bytes(self._h, ctypes.byref(v)))
        return v.value


class PlinkKinshipOperator:
    """Matrix-free K = Z Z^T / M over selected samples and variants (src/kinship_operator.h:19-93).

    ``keep`` / ``variants`` default to all samples / all SNPs (empty list semantics of the
    reference).  ``apply`` accepts a vector (n,) or a matrix (n, k)."""

    def __init__(self, reader: PlinkReader, keep_indices=None, variant_indices=None):
        self.reader = reader   # the reader must outlive the operator (kinship_operator.h:83)
        keep = None if keep_indices is None or len(keep_indices) == 0 else \
            np.ascontiguousarray(keep_indices, dtype=np.int32)
        var = None if variant_indices is None or len(variant_indices) == 0 else \
            np.ascontiguousarray(variant_indices, dtype=np.int32)
        h = _vp()
        _check(lib().cs_kinship_create(reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                                       _p(var, _i32p), 0 if var is None else var.size,
                                       ctypes.byref(h)))
        self._h = h
        r = ctypes.c_int64()
        _check(lib().cs_kinship_rows(h, ctypes.byref(r)))
        self._rows = r.value
        _check(lib().cs_kinship_num_variants(h, ctypes.byref(r)))
        self._m = r.value

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().cs_kinship_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rows(self):
        return self._rows

    def cols(self):
        return self._rows

    def numVariants(self):
        return self._m

    def isDense(self):
        return False

    def set_precision(self, mode):
        _check(lib().cs_kinship_set_precision(self._h, int(mode)))

    def set_missing_mode(self, mode):
        """0 auto, 1 dense observed-indicator products, 2 sparse missing lists (same results)."""
        _check(lib().cs_kinship_set_missing_mode(self._h, int(mode)))

    def compact_rows(self):
        """Run the genotype products on a compact copy of the cohort's rows (same results; the
        copy costs 2 * roundup(n, 256) * roundup(M, 256) / 4 bytes of HBM)."""
        _check(lib().cs_kinship_compact_rows(self._h))

    def device_bytes(self):
        b = ctypes.c_int64()
        _check(lib().cs_kinship_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def apply(self, X, n_threads=1):
        X = np.asarray(X, dtype=np.float64)
        vec = X.ndim == 1
        Xm = X[:, None] if vec else X
        if Xm.shape[0] != self._rows:
            raise RuntimeError("PlinkKinshipOperator::apply: dimension mismatch")
        k = Xm.shape[1]
        Xf = np.asfortranarray(Xm)
        Y = np.empty((self._rows, k), order="F")
        if k > 0:
            _check(lib().cs_kinship_apply(self._h, _p(Xf, _dp), self._rows, k, _p(Y, _dp),
                                          self._rows))
        return Y[:, 0].copy() if vec else np.ascontiguousarray(Y)

    apply_block = apply

    def applyTranspose(self, v, n_threads=1):
        return self.apply(v, n_threads)

    def apply_device(self, dX_ptr, ldx, k, dY_ptr, ldy):
        """Y = K X on device buffers (FP64, column-major) -- no host transfer."""
        _check(lib().cs_kinship_apply_device(self._h, dX_ptr, ldx, k, dY_ptr, ldy))

    def multiply_Zt_v(self, V):
        """U = Z^T V over all SNPs (PlinkReader::multiply_Zt_v, keep-list form): V n x k or n."""
        V = np.asarray(V, dtype=np.float64)
        vec = V.ndim == 1
        Vf = np.asfortranarray(V[:, None] if vec else V)
        M = self.reader.num_snps
        U = np.empty((M, Vf.shape[1]), order="F")
        _check(lib().cs_kinship_multiply_zt(self._h, _p(Vf, _dp), Vf.shape[0], Vf.shape[1], _p(U, _dp), M))
        return U[:, 0] if vec else np.ascontiguousarray(U)

    def multiply_Z_v(self, U):
        """Y = Z U over all SNPs (PlinkReader::multiply_Z_v, keep-list form): U M x k or M."""
        U = np.asarray(U, dtype=np.float64)
        vec = U.ndim == 1
        Uf = np.asfortranarray(U[:, None] if vec else U)
        Y = np.empty((self._rows, Uf.shape[1]), order="F")
        _check(lib().cs_kinship_multiply_z(self._h, _p(Uf, _dp), Uf.shape[0], Uf.shape[1], _p(Y, _dp), self._rows))
        return Y[:, 0] if vec else np.ascontiguousarray(Y)

    def apply_device_noalloc(self, dX_ptr, ldx, k, dY_ptr, ldy):
        """Y = K X on device buffers through the entry point that allocates nothing (fixed
        per-GPU workspace; see cs_kinship_apply_device_noalloc)."""
        _check(lib().cs_kinship_apply_device_noalloc(self._h, dX_ptr, ldx, k, dY_ptr, ldy))

    def trace(self):
        t = ctypes.c_double()
        _check(lib().cs_kinship_trace(self._h, ctypes.byref(t)))
        return t.value

    def getDiagonal(self):
        d = np.empty(self._rows)
        _check(lib().cs_kinship_diag(self._h, _p(d, _dp)))
        return d

    def last_timing_ms(self):
        t = np.empty(5)
        _check(lib().cs_kinship_last_timing(self._h, _p(t, _dp)))
        return t


def pcg_solve(op, sigma_g, sigma_e, dB, dX, rel_tol=1e-6, max_iter=200, precond=1, warm_start=False):
    """(sigma_g K + sigma_e I) X = B on device buffers (DeviceArray, n x k); returns
    (iterations, true relative residual per column)."""
    k = dB.k
    it = ctypes.c_int()
    res = np.empty(k)
    _check(lib().cs_pcg_solve(op._h, float(sigma_g), float(sigma_e), dB.ptr, dB.n, k, dX.ptr, dX.n,
                              float(rel_tol), int(max_iter), int(precond), int(bool(warm_start)),
                              ctypes.byref(it), _p(res, _dp)))
    return it.value, res


def createKinshipOperator(reader, keep_indices=None, variant_indices=None):
    """Factory (src/kinship_operator.h:98-100, kinship_operator.cpp:370-374)."""
    return PlinkKinshipOperator(reader, keep_indices, variant_indices)


# ----------------------------------------------------------------------------- scan (kernel c)
class ScanModel(ctypes.Structure):
    """Variance model of the fused scan (cs_scan_model in the C header)."""
    _fields_ = [("n_score", ctypes.c_int64), ("alpha", ctypes.c_double),
                ("quad_begin", ctypes.c_int64), ("quad_end", ctypes.c_int64),
                ("quad_w", _dp), ("cov_begin", ctypes.c_int64), ("cov_end", ctypes.c_int64),
                ("cov_C", _dp)]


def _opt_i32(a):
    return None if a is None or len(a) == 0 else np.ascontiguousarray(a, dtype=np.int32)


def scan_sums(reader, keep, R, min_maf=0.01, snps=None, prec=CS_PREC_INT8_L3):
    """Prepared-marker sums per SNP: (freq, n_obs, pass, dot (n_snps, nr), gtg)."""
    keep = _opt_i32(keep)
    snps = _opt_i32(snps)
    n = reader.num_samples if keep is None else keep.size
    R = np.asfortranarray(np.asarray(R, dtype=np.float64).reshape(n, -1))
    nr = R.shape[1]
    ns = reader.num_snps if snps is None else snps.size
    freq = np.empty(ns)
    nobs = np.empty(ns, dtype=np.int64)
    ps = np.empty(ns, dtype=np.int32)
    dot = np.empty((ns, nr))
    gtg = np.empty(ns)
    _check(lib().cs_scan_sums(reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                              _p(R, _dp), n, nr, float(min_maf), _p(snps, _i32p),
                              0 if snps is None else snps.size, int(prec), _p(freq, _dp),
                              _p(nobs, _i64p), _p(ps, _i32p), _p(dot, _dp), _p(gtg, _dp)))
    return freq, nobs, ps.astype(bool), dot, gtg


def lmm_scan(reader, keep, R, n_score=1, alpha=0.0, quad=None, quad_w=None, cov=None, cov_C=None,
             min_maf=0.01, snps=None, prec=CS_PREC_INT8_L3):
    """Fused per-SNP scan; returns dict of freq, n_obs, pass, beta, se, p, log10p."""
    keep = _opt_i32(keep)
    snps = _opt_i32(snps)
    n = reader.num_samples if keep is None else keep.size
    R = np.asfortranarray(np.asarray(R, dtype=np.float64).reshape(n, -1))
    nr = R.shape[1]
    ns = reader.num_snps if snps is None else snps.size
    qb, qe = quad if quad is not None else (0, 0)
    cb, ce = cov if cov is not None else (0, 0)
    qw = np.ascontiguousarray(quad_w if quad_w is not None else np.zeros(0), dtype=np.float64)
    cC = np.ascontiguousarray(cov_C if cov_C is not None else np.zeros(0), dtype=np.float64)
    m = ScanModel(n_score, float(alpha), qb, qe, _p(qw, _dp), cb, ce, _p(cC, _dp))
    out = {k: np.empty((ns, n_score)) for k in ("beta", "se", "p", "log10p")}
    freq = np.empty(ns)
    nobs = np.empty(ns, dtype=np.int64)
    ps = np.empty(ns, dtype=np.int32)
    _check(lib().cs_lmm_scan(reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                             _p(R, _dp), n, nr, float(min_maf), _p(snps, _i32p),
                             0 if snps is None else snps.size, int(prec), ctypes.byref(m),
                             _p(freq, _dp), _p(nobs, _i64p), _p(ps, _i32p), _p(out["beta"], _dp),
                             _p(out["se"], _dp), _p(out["p"], _dp), _p(out["log10p"], _dp)))
    out.update(freq=freq, n_obs=nobs, passes=ps.astype(bool))
    return out


def lmm_scan_device(reader, keep, dR, n_score=1, alpha=0.0, quad=None, quad_w=None, cov=None, cov_C=None,
                    min_maf=0.01, snps=None, prec=CS_PREC_INT8_L3, out=None):
    """Fused per-SNP scan with R (a DeviceArray, n x nr) and the statistics in device memory.
    out: dict of DeviceArrays to fill (freq, n_obs, pass, beta, se, p, log10p; made when None).
    Returns (out, timing) with timing = (ms counts, ms products, ms epilogues, packed bytes)."""
    keep = _opt_i32(keep)
    snps = _opt_i32(snps)
    ns = reader.num_snps if snps is None else snps.size
    qb, qe = quad if quad is not None else (0, 0)
    cb, ce = cov if cov is not None else (0, 0)
    qw = np.ascontiguousarray(quad_w if quad_w is not None else np.zeros(0), dtype=np.float64)
    cC = np.ascontiguousarray(cov_C if cov_C is not None else np.zeros(0), dtype=np.float64)
    m = ScanModel(n_score, float(alpha), qb, qe, _p(qw, _dp), cb, ce, _p(cC, _dp))
    if out is None:
        out = {k: DeviceArray(ns, n_score) for k in ("beta", "se", "p", "log10p")}
        out.update(freq=DeviceArray(ns, 1), n_obs=DeviceArray(ns, 1), passes=DeviceArray((ns + 1) // 2, 1))
    _check(lib().cs_lmm_scan_device(reader._h, _p(keep, _i32p), 0 if keep is None else keep.size, dR.ptr, dR.n,
                                    dR.k, float(min_maf), _p(snps, _i32p), 0 if snps is None else snps.size,
                                    int(prec), ctypes.byref(m), out["freq"].ptr, out["n_obs"].ptr,
                                    out["passes"].ptr, out["beta"].ptr, out["se"].ptr, out["p"].ptr,
                                    out["log10p"].ptr))
    a, b, c, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    _check(lib().cs_scan_last_timing(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
    return out, (a.value, b.value, c.value, d.value)


# ----------------------------------------------------------------------------- GRM (kernel b)
class GrmBuilder:
    """Dense GRM build (GrmBuilder, libCosmicCore/src/grm_builder.h:15-80; compute at
    grm_builder.cpp:63-399, save at :1977-2026).  Algorithm 1 (VanRaden, the default) is computed
    exactly on int8 MFMA: the lower triangle is the correctly rounded float32 of Z Z^T / sum_2pq,
    getGrm() the FP64 matrix (correctly rounded raw Z Z^T divided by sum_2pq).  Algorithms 0 and 3
    (standardised, G = Z Z^T / M) run as FP16 hi/lo split products on MFMA with FP32 accumulation
    flushed to FP64 every 16384 SNPs."""

    class Options:
        def __init__(self, min_maf=0.01, algorithm=1, thread_num=1, block_size=2048, min_hwe=0.0):
            self.min_maf = min_maf
            self.min_hwe = min_hwe
            self.algorithm = algorithm
            self.thread_num = thread_num
            self.block_size = block_size

    def __init__(self, opts=None):
        self.options = opts if opts is not None else GrmBuilder.Options()
        self.tri = None
        self.raw = None
        self.total_sum_2pq = 0.0
        self.total_valid_snps = 0
        self._fam = None

    def compute(self, reader, keep_indices=None, fp64=True):
        """Build the GRM of all samples.  fp64=False skips the n x n FP64 copy (the float32
        lower triangle for save() is always produced)."""
        if keep_indices is not None and len(keep_indices) > 0:
            raise RuntimeError("PlinkReader GRM calculation with subsetting not yet implemented. "
                               "Use PGEN/BGEN or full dataset.")
        if reader.num_samples == 0 or reader.num_snps == 0:
            raise RuntimeError("Empty dataset provided to GrmBuilder")
        n = reader.num_samples
        tri = np.empty(n * (n + 1) // 2, dtype=np.float32)
        raw = np.empty((n, n)) if fp64 else None
        s = ctypes.c_double()
        nv = ctypes.c_int64()
        _check(lib().cs_grm_build(reader._h, int(self.options.algorithm), float(self.options.min_maf),
                                  float(self.options.min_hwe), _p(tri, _fp), _p(raw, _dp), ctypes.byref(s),
                                  ctypes.byref(nv)))
        self.tri, self.raw = tri, raw
        self.total_sum_2pq, self.total_valid_snps = s.value, nv.value
        self._fam = reader.fam
        self._n = n

    def getGrm(self):
        if self.raw is None:
            raise RuntimeError("GRM not computed with fp64=True")
        if self.options.algorithm in (0, 3):
            return self.raw / self.total_valid_snps if self.total_valid_snps > 0 else self.raw.copy()
        return self.raw / self.total_sum_2pq if self.total_sum_2pq > 0 else self.raw.copy()

    def getRaw(self):
        return self.raw

    def getLowerTriangle(self):
        return self.tri

    def getSum2pq(self):
        return self.total_sum_2pq

    def getTotalSnps(self):
        return self.total_valid_snps

    def last_timing_ms(self):
        """(ms of the int8 products, ms of the combine, product launches, int8 MFMA ops issued)."""
        a, b, c, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        _check(lib().cs_grm_last_timing(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                        ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def save(self, out_prefix):
        """GCTA files (grm_builder.cpp:1977-2026): .grm.id (FID<TAB>IID), .grm.bin (float32 lower
        triangle, row-major, j <= i), .grm.N.bin (float32 SNP count of every pair)."""
        if self.tri is None:
            raise RuntimeError("GRM not computed")
        fam = self._fam if self._fam is not None else [[str(i), str(i)] for i in range(self._n)]
        with open(out_prefix + ".grm.id", "w") as f:
            for row in fam:
                f.write(f"{row[0]}\t{row[1]}\n")
        self.tri.astype("<f4").tofile(out_prefix + ".grm.bin")
        np.full(self.tri.size, np.float32(self.total_valid_snps), dtype="<f4").tofile(
            out_prefix + ".grm.N.bin")


def grm_build_device(reader, d_tri_ptr, min_maf=0.01, alg=1, min_hwe=0.0):
    """GRM lower triangle into a device buffer of n(n+1)/2 floats; returns (sum_2pq, n_valid)."""
    s = ctypes.c_double()
    nv = ctypes.c_int64()
    _check(lib().cs_grm_build_device(reader._h, int(alg), float(min_maf), float(min_hwe), d_tri_ptr,
                                     ctypes.byref(s), ctypes.byref(nv)))
    return s.value, nv.value


def sym_eigen_device(dA, dD, leaf=32766):
    """Eigendecomposition of the symmetric DeviceArray dA (n x n, lower triangle read) in place:
    dA <- eigenvectors, dD (n x 1 DeviceArray) <- eigenvalues ascending -- the
    SelfAdjointEigenSolver step of EigenVce::prepare (libCosmicCore/src/eigen_vce.cpp:34-38).
    leaf = 0: rocSOLVER dsyevd (n <= 32766); leaf > 0: the 64-bit tridiagonal divide-and-conquer
    path with blocks of at most `leaf` rows."""
    if dA.k != dA.n or dD.n * dD.k < dA.n:
        raise ValueError("sym_eigen_device: dimension mismatch")
    _check(lib().cs_sym_eigen_device(dA.ptr, dA.n, dA.n, dD.ptr, int(leaf)))


# ----------------------------------------------------------------------- dense null model + scan
class DenseLmm:
    """EigenVce (libCosmicCore/src/eigen_vce.cpp:19-113) + LmmEngine::scan
    (src/lmm_engine.cpp:1363-1525) on the device, with covariate projection.

    y: kept-sample phenotype (n,); X: (n, c) design including the intercept (cast to float32 as
    the reference's Eigen::MatrixXf)."""

    def __init__(self, y, X):
        y = np.ascontiguousarray(y, dtype=np.float64)
        X = np.asfortranarray(np.asarray(X, dtype=np.float32).reshape(y.size, -1))
        self.n, self.c = y.size, X.shape[1]
        h = _vp()
        _check(lib().cs_lmm_create(_p(y, _dp), _p(X, _f32p), self.n, self.n, self.c, ctypes.byref(h)))
        self._h = h
        self.sigma2_g = self.sigma2_e = self.delta = None

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().cs_lmm_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, G):
        """EigenVce::prepare with an n x n GRM (numpy array or DeviceArray)."""
        if isinstance(G, DeviceArray):
            _check(lib().cs_lmm_prepare_dense(self._h, G.ptr, G.n, 1))
        else:
            Gf = np.asfortranarray(np.asarray(G, dtype=np.float64))
            if Gf.shape != (self.n, self.n):
                raise RuntimeError("EigenVce::prepare: Dimension mismatch")
            _check(lib().cs_lmm_prepare_dense(self._h, Gf.ctypes.data_as(_vp), self.n, 0))

    def runNullModel(self):
        a, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cs_lmm_null_brent(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(d)))
        self.sigma2_g, self.sigma2_e, self.delta = a.value, b.value, d.value
        return True

    def set_variance(self, sigma2_g, sigma2_e):
        _check(lib().cs_lmm_set_variance(self._h, float(sigma2_g), float(sigma2_e)))
        self.sigma2_g, self.sigma2_e = sigma2_g, sigma2_e

    def eigen(self, with_w=False):
        ne = self.n - self.c
        D = np.empty(ne)
        Uty = np.empty(ne)
        W = np.empty((self.n, ne), order="F") if with_w else None
        _check(lib().cs_lmm_get_eigen(self._h, _p(D, _dp), _p(Uty, _dp), _p(W, _dp), self.n))
        return (D, Uty, W) if with_w else (D, Uty)

    def scan(self, reader, keep, min_maf=0.01, snps=None, prec=CS_PREC_INT8_L3):
        keep = _opt_i32(keep)
        snps = _opt_i32(snps)
        ns = reader.num_snps if snps is None else snps.size
        out = {k: np.empty(ns) for k in ("freq", "beta", "se", "p", "log10p")}
        nobs = np.empty(ns, dtype=np.int64)
        ps = np.empty(ns, dtype=np.int32)
        _check(lib().cs_lmm_scan_rotated(self._h, reader._h, _p(keep, _i32p),
                                         0 if keep is None else keep.size, _p(snps, _i32p),
                                         0 if snps is None else snps.size, float(min_maf), int(prec),
                                         _p(out["freq"], _dp), _p(nobs, _i64p), _p(ps, _i32p),
                                         _p(out["beta"], _dp), _p(out["se"], _dp), _p(out["p"], _dp),
                                         _p(out["log10p"], _dp)))
        out.update(n_obs=nobs, passes=ps.astype(bool))
        return out


def grm_build_raw(reader, d_raw, min_maf=0.01, snps=None, alg=1, min_hwe=0.0):
    """Raw Z Z^T over a SNP subset into an N x N DeviceArray; returns (S_block, S_seq, n_valid)."""
    snps = _opt_i32(snps)
    a, b, nv = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    _check(lib().cs_grm_build_raw(reader._h, int(alg), float(min_maf), float(min_hwe), _p(snps, _i32p),
                                  0 if snps is None else snps.size, d_raw.ptr, 1, ctypes.byref(a),
                                  ctypes.byref(b), ctypes.byref(nv)))
    return a.value, b.value, nv.value


def grm_loco_subset(d_raw, d_raw_chr, N, keep, S_global, S_chr, d_out):
    keep = _opt_i32(keep)
    _check(lib().cs_grm_loco_subset(d_raw.ptr, None if d_raw_chr is None else d_raw_chr.ptr, N,
                                    _p(keep, _i32p), 0 if keep is None else keep.size,
                                    float(S_global), float(S_chr), d_out.ptr))


# ------------------------------------------------------------------------------- GWAS driver
class GwasOptions(ctypes.Structure):
    _fields_ = [("min_maf", ctypes.c_double), ("no_loco", ctypes.c_int),
                ("scan_chr", ctypes.c_char_p), ("prec", ctypes.c_int), ("grm_alg", ctypes.c_int)]


class GwasSummary(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int64), ("n_snps_grm", ctypes.c_int64),
                ("n_tested", ctypes.c_int64), ("sum_2pq", ctypes.c_double),
                ("loco", ctypes.c_int32), ("n_models", ctypes.c_int32),
                ("model_label", (ctypes.c_char * 16) * 64),
                ("sigma2_g", ctypes.c_double * 64), ("sigma2_e", ctypes.c_double * 64),
                ("ms_grm", ctypes.c_double), ("ms_null", ctypes.c_double),
                ("ms_scan", ctypes.c_double)]


def read_pheno(path, fam, mpheno=1, missing=-9.0):
    """Phenotype column `mpheno` (1-based after FID IID) aligned to the .fam rows
    (PhenoReader::load / align, libCosmicCore/src/pheno_reader.cpp:13-63,139-160): NA / na / NaN /
    nan / NAN, unparsable values and the missing code -9 become NaN; absent samples are NaN."""
    vals = {}
    with open(path) as f:
        for ln in f:
            t = ln.split()
            if len(t) < 2 + mpheno:
                continue
            s = t[1 + mpheno]
            if s in ("NA", "na", "NaN", "nan", "NAN"):
                v = float("nan")
            else:
                try:
                    v = float(s)
                except ValueError:
                    v = float("nan")
            if v == v and abs(v - missing) < 1e-6:
                v = float("nan")
            vals[(t[0], t[1])] = v
    return np.array([vals.get((r[0], r[1]), float("nan")) for r in fam])


class GwasMfOptions(ctypes.Structure):
    _fields_ = [("min_maf", ctypes.c_double), ("no_loco", ctypes.c_int), ("scan_chr", ctypes.c_char_p),
                ("prec", ctypes.c_int), ("mode", ctypes.c_int), ("max_rank", ctypes.c_int),
                ("oversample", ctypes.c_int), ("seed", ctypes.c_int64), ("trace_threshold", ctypes.c_double),
                ("min_rank", ctypes.c_int), ("no_fallback", ctypes.c_int), ("error_tol", ctypes.c_double),
                ("sigma2_g", ctypes.c_double), ("sigma2_e", ctypes.c_double), ("pcg_tol", ctypes.c_double),
                ("pcg_max_iter", ctypes.c_int), ("n_calibration", ctypes.c_int), ("exact_log10p", ctypes.c_double)]


def _summary_dict(sm):
    return dict(n_samples=sm.n_samples, n_snps_grm=sm.n_snps_grm, n_tested=sm.n_tested,
                sum_2pq=sm.sum_2pq, loco=bool(sm.loco),
                models=[(sm.model_label[i].value.decode(), sm.sigma2_g[i], sm.sigma2_e[i])
                        for i in range(sm.n_models)],
                ms_grm=sm.ms_grm, ms_null=sm.ms_null, ms_scan=sm.ms_scan)


def gwas_matrix_free_run(bfile, y, out_prefix, cov=None, min_maf=0.01, no_loco=False, scan_chr=None,
                         mode="lowrank", max_rank=256, oversample=10, seed=42, trace_threshold=0.95,
                         min_rank=1, no_fallback=False, error_tol=1.0, sigma2_g=0.0, sigma2_e=0.0,
                         pcg_tol=1e-6, pcg_max_iter=200, n_calibration=32, prec=CS_PREC_INT8_L3,
                         exact_log10p=0.0):
    """The matrix-free GWAS (--fastlmm --fastlmm-matrix-free) on the GPU; mode "lowrank" (the
    reference path) or "pcg" (PCG-projected scan)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    N = y.size
    if cov is None:
        covf, nc = None, 0
    else:
        covf = np.asfortranarray(np.asarray(cov, dtype=np.float64).reshape(N, -1))
        nc = covf.shape[1]
    o = GwasMfOptions(float(min_maf), int(bool(no_loco)), None if scan_chr is None else str(scan_chr).encode(),
                      int(prec), {"lowrank": 0, "pcg": 1}[mode], int(max_rank), int(oversample), int(seed),
                      float(trace_threshold), int(min_rank), int(bool(no_fallback)), float(error_tol),
                      float(sigma2_g), float(sigma2_e), float(pcg_tol), int(pcg_max_iter), int(n_calibration),
                      float(exact_log10p))
    sm = GwasSummary()
    _check(lib().cs_gwas_matrix_free_run(str(bfile).encode(), _p(y, _dp), _p(covf, _dp), N, nc, ctypes.byref(o),
                                         str(out_prefix).encode(), ctypes.byref(sm)))
    return _summary_dict(sm)


def gwas_dense_run(bfile, y, out_prefix, cov=None, min_maf=0.01, no_loco=False, scan_chr=None,
                   prec=CS_PREC_INT8_L3, grm_alg=1):
    """The default dense PLINK GWAS of GwasScanner::runLocoScan end to end on the GPU; writes
    out_prefix.assoc.txt / .scan.meta.txt / .cosmic.sumstats and returns the run summary.
    grm_alg is --make-grm-alg: 1 (VanRaden, exact) or 0 / 3 (standardised, FP16-split MFMA)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    N = y.size
    if cov is None:
        covf, nc = None, 0
    else:
        covf = np.asfortranarray(np.asarray(cov, dtype=np.float64).reshape(N, -1))
        nc = covf.shape[1]
    o = GwasOptions(float(min_maf), int(bool(no_loco)),
                    None if scan_chr is None else str(scan_chr).encode(), int(prec), int(grm_alg))
    sm = GwasSummary()
    _check(lib().cs_gwas_dense_run(str(bfile).encode(), _p(y, _dp), _p(covf, _dp), N, nc,
                                   ctypes.byref(o), str(out_prefix).encode(), ctypes.byref(sm)))
    k = sm.n_models
    return dict(n_samples=sm.n_samples, n_snps_grm=sm.n_snps_grm, n_tested=sm.n_tested,
                sum_2pq=sm.sum_2pq, loco=bool(sm.loco),
                models=[(sm.model_label[i].value.decode(), sm.sigma2_g[i], sm.sigma2_e[i])
                        for i in range(k)],
                ms_grm=sm.ms_grm, ms_null=sm.ms_null, ms_scan=sm.ms_scan)


# --------------------------------------------------------------------- PCG-projected scan
class PscanInfo(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("n_calibration", ctypes.c_int64),
                ("pcg_iterations", ctypes.c_int), ("max_rel_residual", ctypes.c_double),
                ("n_exact", ctypes.c_int64), ("exact_pcg_iterations", ctypes.c_int),
                ("calibration_rel_sd", ctypes.c_double), ("exact_log10p", ctypes.c_double),
                ("dlog10p_bound_below", ctypes.c_double)]


def markers_device(reader, keep, snps, d_G, min_maf=0.01):
    """Prepared markers of `snps` over `keep` into d_G (DeviceArray n x len(snps)); returns (pass, freq)."""
    keep = _opt_i32(keep)
    snps = np.ascontiguousarray(snps, dtype=np.int32)
    ps = np.empty(snps.size, dtype=np.int32)
    fr = np.empty(snps.size)
    _check(lib().cs_markers_device(reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                                   _p(snps, _i32p), snps.size, float(min_maf), d_G.ptr, d_G.n,
                                   _p(ps, _i32p), _p(fr, _dp)))
    return ps.astype(bool), fr


def pcg_projected_scan(op, reader, keep, Y, X, sigma_g, sigma_e, cal_snps, snps=None, min_maf=0.01,
                       rel_tol=1e-6, max_iter=200, prec=CS_PREC_INT8_L3, exact_log10p=0.0):
    """PCG-projected GLS scan with kinship operator `op` (see cs_pcg_projected_scan).

    exact_log10p: level of the exact stage, on the smallest P over the traits Bonferroni-adjusted
    for their number (0 = the default 5, negative = every tested SNP, ``inf`` = off)."""
    keep = _opt_i32(keep)
    snps = _opt_i32(snps)
    n = op.rows()
    Yf = np.asfortranarray(np.asarray(Y, dtype=np.float64).reshape(n, -1))
    Xf = np.asfortranarray(np.asarray(X, dtype=np.float32).reshape(n, -1))
    T, c = Yf.shape[1], Xf.shape[1]
    cal = np.ascontiguousarray(cal_snps, dtype=np.int32)
    ns = reader.num_snps if snps is None else snps.size
    out = {k: np.empty((ns, T)) for k in ("beta", "se", "p", "log10p")}
    freq = np.empty(ns)
    nobs = np.empty(ns, dtype=np.int64)
    ps = np.empty(ns, dtype=np.int32)
    info = PscanInfo()
    _check(lib().cs_pcg_projected_scan(op._h, reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                                       _p(Yf, _dp), n, T, _p(Xf, _f32p), n, c, float(sigma_g),
                                       float(sigma_e), _p(cal, _i32p), cal.size, _p(snps, _i32p),
                                       0 if snps is None else snps.size, float(min_maf), float(rel_tol),
                                       int(max_iter), int(prec), float(exact_log10p), _p(freq, _dp), _p(nobs, _i64p),
                                       _p(ps, _i32p), _p(out["beta"], _dp), _p(out["se"], _dp),
                                       _p(out["p"], _dp), _p(out["log10p"], _dp), ctypes.byref(info)))
    out.update(freq=freq, n_obs=nobs, passes=ps.astype(bool), gamma=info.gamma,
               n_calibration=info.n_calibration, pcg_iterations=info.pcg_iterations,
               max_rel_residual=info.max_rel_residual, n_exact=info.n_exact,
               exact_pcg_iterations=info.exact_pcg_iterations, calibration_rel_sd=info.calibration_rel_sd,
               exact_log10p=info.exact_log10p, dlog10p_bound_below=info.dlog10p_bound_below)
    return out


# ------------------------------------------------------------------- low-rank matrix-free path
class LowRankDiag(ctypes.Structure):
    _fields_ = [("requested_rank", ctypes.c_int), ("effective_rank", ctypes.c_int),
                ("sketch_size", ctypes.c_int), ("total_rank", ctypes.c_int),
                ("residual_eigenvalue", ctypes.c_double), ("trace", ctypes.c_double),
                ("trace_fraction", ctypes.c_double), ("relative_frobenius_error", ctypes.c_double),
                ("trace_threshold_met", ctypes.c_int), ("fell_back", ctypes.c_int),
                ("sigma2_g", ctypes.c_double), ("sigma2_e", ctypes.c_double), ("delta", ctypes.c_double),
                ("decomposition_seconds", ctypes.c_double), ("estimated_peak_bytes", ctypes.c_int64)]


class LowRankLmm:
    """Randomized low-rank null model on a kinship operator + scanLowRank
    (src/lmm_engine.cpp:897-1095, 1269-1361; src/matrix_free_vce.cpp:32-235)."""

    def __init__(self, op, y, X, rank=256, oversample=10, seed=42, trace_threshold=0.95, min_rank=1,
                 no_fallback=False):
        n = op.rows()
        y = np.ascontiguousarray(y, dtype=np.float64)
        Xf = np.asfortranarray(np.asarray(X, dtype=np.float32).reshape(n, -1))
        self.n, self.c = n, Xf.shape[1]
        h = _vp()
        d = LowRankDiag()
        _check(lib().cs_lowrank_prepare(op._h, _p(y, _dp), _p(Xf, _f32p), n, self.c, int(rank), int(oversample),
                                        int(seed), float(trace_threshold), int(min_rank), int(bool(no_fallback)),
                                        ctypes.byref(h), ctypes.byref(d)))
        self._h = h
        self.diag = {f[0]: getattr(d, f[0]) for f in LowRankDiag._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().cs_lowrank_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, with_u=False):
        r = ctypes.c_int64()
        _check(lib().cs_lowrank_get(self._h, ctypes.byref(r), None, None, 0, None, None, None))
        lam = np.empty(r.value)
        U = np.empty((self.n, r.value), order="F") if with_u else None
        r0, sg, se = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cs_lowrank_get(self._h, None, _p(lam, _dp), _p(U, _dp), self.n, ctypes.byref(r0),
                                    ctypes.byref(sg), ctypes.byref(se)))
        return dict(rank=r.value, lam=lam, U=U, r0=r0.value, sg=sg.value, se=se.value)

    def scan(self, reader, keep, snps=None, min_maf=0.01, prec=CS_PREC_INT8_L3):
        keep = _opt_i32(keep)
        snps = _opt_i32(snps)
        ns = reader.num_snps if snps is None else snps.size
        out = {k: np.empty(ns) for k in ("freq", "beta", "se", "p", "log10p")}
        nobs = np.empty(ns, dtype=np.int64)
        ps = np.empty(ns, dtype=np.int32)
        _check(lib().cs_lowrank_scan(self._h, reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                                     _p(snps, _i32p), 0 if snps is None else snps.size, float(min_maf), int(prec),
                                     _p(out["freq"], _dp), _p(nobs, _i64p), _p(ps, _i32p), _p(out["beta"], _dp),
                                     _p(out["se"], _dp), _p(out["p"], _dp), _p(out["log10p"], _dp)))
        out.update(n_obs=nobs, passes=ps.astype(bool))
        return out


# ------------------------------------------------------------------------------ serving loop
class _Request(ctypes.Structure):
    _fields_ = [("keep", _i32p), ("n_keep", ctypes.c_int64), ("Y", _dp), ("ldy", ctypes.c_int64),
                ("n_traits", ctypes.c_int64), ("X", _f32p), ("ldx", ctypes.c_int64), ("n_cov", ctypes.c_int64),
                ("sigma2_g", ctypes.c_double), ("sigma2_e", ctypes.c_double), ("cal_snps", _i32p),
                ("n_cal", ctypes.c_int64), ("snps", _i32p), ("n_snps", ctypes.c_int64),
                ("min_maf", ctypes.c_double), ("rel_tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("prec", ctypes.c_int), ("exact_log10p", ctypes.c_double)]


class _ReqStats(ctypes.Structure):
    _fields_ = [("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("iterations", ctypes.c_int), ("gamma", ctypes.c_double), ("n_calibration", ctypes.c_int64),
                ("n_columns", ctypes.c_int64), ("n_exact", ctypes.c_int64), ("exact_iterations", ctypes.c_int),
                ("calibration_rel_sd", ctypes.c_double), ("dlog10p_bound_below", ctypes.c_double),
                ("admit_bytes", ctypes.c_int64), ("compact_rows", ctypes.c_int)]


class ScanServer:
    """Continuously batched scan requests over one resident genotype matrix (see cs_server_*)."""

    def __init__(self, reader, variants=None, max_columns=256, hbm_budget_bytes=0):
        self.reader = reader
        var = _opt_i32(variants)
        h = _vp()
        _check(lib().cs_server_create(reader._h, _p(var, _i32p), 0 if var is None else var.size, int(max_columns),
                                      int(hbm_budget_bytes), ctypes.byref(h)))
        self._h = h
        self._meta = {}

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().cs_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, Y, X, sigma2_g, sigma2_e, cal_snps, keep=None, snps=None, min_maf=0.01, rel_tol=1e-6,
               max_iter=200, prec=CS_PREC_INT8_L3, exact_log10p=0.0):
        keep_a = _opt_i32(keep)
        n = self.reader.num_samples if keep_a is None else keep_a.size
        Yf = np.asfortranarray(np.asarray(Y, dtype=np.float64).reshape(n, -1))
        Xf = np.asfortranarray(np.asarray(X, dtype=np.float32).reshape(n, -1))
        cal = np.ascontiguousarray(cal_snps, dtype=np.int32)
        sn = _opt_i32(snps)
        r = _Request(_p(keep_a, _i32p), 0 if keep_a is None else keep_a.size, _p(Yf, _dp), n, Yf.shape[1],
                     _p(Xf, _f32p), n, Xf.shape[1], float(sigma2_g), float(sigma2_e), _p(cal, _i32p), cal.size,
                     _p(sn, _i32p), 0 if sn is None else sn.size, float(min_maf), float(rel_tol), int(max_iter),
                     int(prec), float(exact_log10p))
        rid = ctypes.c_int64()
        _check(lib().cs_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._meta[rid.value] = (Yf.shape[1], self.reader.num_snps if sn is None else sn.size)
        return rid.value

    def step(self):
        a, p, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().cs_server_step(self._h, ctypes.byref(a), ctypes.byref(p), ctypes.byref(c)))
        return a.value, p.value, c.value

    def run_until_idle(self, max_steps=100000):
        for _ in range(max_steps):
            a, p, _ = self.step()
            if a == 0 and p == 0:
                return
        raise RuntimeError("server did not drain")

    def status(self, rid):
        st = ctypes.c_int32()
        s = _ReqStats()
        _check(lib().cs_server_result(self._h, rid, ctypes.byref(st), None, None, None, None, None, None, None,
                                      ctypes.byref(s)))
        return st.value, {f[0]: getattr(s, f[0]) for f in _ReqStats._fields_}

    def result(self, rid):
        T, ns = self._meta[rid]
        out = {k: np.empty((ns, T)) for k in ("beta", "se", "p", "log10p")}
        freq = np.empty(ns)
        nobs = np.empty(ns, dtype=np.int64)
        ps = np.empty(ns, dtype=np.int32)
        st = ctypes.c_int32()
        s = _ReqStats()
        _check(lib().cs_server_result(self._h, rid, ctypes.byref(st), _p(freq, _dp), _p(nobs, _i64p), _p(ps, _i32p),
                                      _p(out["beta"], _dp), _p(out["se"], _dp), _p(out["p"], _dp),
                                      _p(out["log10p"], _dp), ctypes.byref(s)))
        if st.value != 2:
            raise RuntimeError(f"request {rid} not finished (status {st.value})")
        out.update(freq=freq, n_obs=nobs, passes=ps.astype(bool),
                   stats={f[0]: getattr(s, f[0]) for f in _ReqStats._fields_})
        return out

    def release(self, rid):
        _check(lib().cs_server_release(self._h, rid))
        self._meta.pop(rid, None)


# ----------------------------------------------------------------------- multivariate LMM
class MvLmm:
    """MvEigenVce (libCosmicCore/src/eigen_vce.cpp:496-910) + MvLmmEngine::scan
    (src/mvlmm_engine.cpp:210-394, 756-931) on the device, with covariate projection.

    Y: (n, d) kept-sample phenotypes (1 <= d <= 6); X: (n, c) design including the intercept
    (float32, as the reference's covariate matrix)."""

    def __init__(self, Y, X):
        Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
        if Y.ndim == 1:
            Y = Y.reshape(-1, 1, order="F")
        self.n, self.d = Y.shape
        X = np.asfortranarray(np.asarray(X, dtype=np.float32).reshape(self.n, -1))
        self.c = X.shape[1]
        h = _vp()
        _check(lib().cs_mvlmm_create(_p(Y, _dp), self.n, self.n, self.d, _p(X, _f32p), self.n, self.c,
                                     ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().cs_mvlmm_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, G):
        """MvEigenVce::prepare with an n x n GRM (numpy array or DeviceArray)."""
        if isinstance(G, DeviceArray):
            _check(lib().cs_mvlmm_prepare_dense(self._h, G.ptr, G.n, 1))
        else:
            Gf = np.asfortranarray(np.asarray(G, dtype=np.float64))
            if Gf.shape != (self.n, self.n):
                raise RuntimeError("MvEigenVce::prepare: Dimension mismatch")
            _check(lib().cs_mvlmm_prepare_dense(self._h, Gf.ctypes.data_as(_vp), self.n, 0))

    def runNullModel(self, solver="em", max_iter=1000, tol=1e-8, em_min_iter=10, ai_max_iter=50,
                     ai_tol=1e-8):
        """EM (default) or "aireml"; returns (converged, em_iterations, ai_iterations)."""
        conv, em_it, ai_it = (np.zeros(1, dtype=np.int32) for _ in range(3))
        _check(lib().cs_mvlmm_null(self._h, 1 if solver == "aireml" else 0, int(max_iter), float(tol),
                                   int(em_min_iter), int(ai_max_iter), float(ai_tol), _p(conv, _i32p),
                                   _p(em_it, _i32p), _p(ai_it, _i32p)))
        return bool(conv[0]), int(em_it[0]), int(ai_it[0])

    def variance(self):
        """(Vg, Ve, Vg_se, Ve_se, logl); the standard errors are NaN when not computed."""
        d = self.d
        out = [np.empty((d, d)) for _ in range(4)]
        ll = np.empty(1)
        _check(lib().cs_mvlmm_get(self._h, *[_p(a, _dp) for a in out], _p(ll, _dp)))
        return (*out, float(ll[0]))

    def set_variance(self, Vg, Ve):
        Vg = np.ascontiguousarray(Vg, dtype=np.float64)
        Ve = np.ascontiguousarray(Ve, dtype=np.float64)
        _check(lib().cs_mvlmm_set_variance(self._h, _p(Vg, _dp), _p(Ve, _dp)))

    def eigen(self, with_w=False):
        ne = self.n - self.c
        D = np.empty(ne)
        Uty = np.empty((ne, self.d), order="F")
        W = np.empty((self.n, ne), order="F") if with_w else None
        _check(lib().cs_mvlmm_get_eigen(self._h, _p(D, _dp), _p(Uty, _dp), _p(W, _dp), self.n))
        return (D, Uty, W) if with_w else (D, Uty)

    def scan(self, reader, keep=None, snps=None):
        keep = _opt_i32(keep)
        snps = _opt_i32(snps)
        ns = reader.num_snps if snps is None else snps.size
        d = self.d
        freq, wald, p = np.empty(ns), np.empty(ns), np.empty(ns)
        beta = np.empty((ns, d))
        vbeta = np.empty((ns, d, d))
        status = np.empty(ns, dtype=np.int32)
        _check(lib().cs_mvlmm_scan(self._h, reader._h, _p(keep, _i32p), 0 if keep is None else keep.size,
                                   _p(snps, _i32p), 0 if snps is None else snps.size, _p(freq, _dp),
                                   _p(beta, _dp), _p(vbeta, _dp), _p(wald, _dp), _p(p, _dp),
                                   _p(status, _i32p)))
        return dict(freq=freq, beta=beta, vbeta=vbeta, wald=wald, p=p, status=status)


def grm_loco_subset_scaled(d_raw, d_raw_chr, N, keep, norm, norm_chr, S, S_chr, d_out):
    """LOCO GRM of the multivariate driver (multivariate_model.cpp:1040-1063) into d_out."""
    keep = _opt_i32(keep)
    _check(lib().cs_grm_loco_subset_scaled(d_raw.ptr, None if d_raw_chr is None else d_raw_chr.ptr, N,
                                           _p(keep, _i32p), 0 if keep is None else keep.size,
                                           float(norm), float(norm_chr), float(S), float(S_chr), d_out.ptr))


class GwasMvOptions(ctypes.Structure):
    _fields_ = [("min_maf", ctypes.c_double), ("no_loco", ctypes.c_int), ("scan_chr", ctypes.c_char_p),
                ("grm_alg", ctypes.c_int), ("solver", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("tol", ctypes.c_double), ("em_min_iter", ctypes.c_int), ("ai_max_iter", ctypes.c_int),
                ("ai_tol", ctypes.c_double)]


class GwasMvSummary(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int64), ("n_traits", ctypes.c_int64), ("n_tested", ctypes.c_int64),
                ("loco", ctypes.c_int32), ("n_models", ctypes.c_int32),
                ("model_label", (ctypes.c_char * 16) * 64),
                ("vg", (ctypes.c_double * 36) * 64), ("ve", (ctypes.c_double * 36) * 64),
                ("ms_grm", ctypes.c_double), ("ms_null", ctypes.c_double), ("ms_scan", ctypes.c_double)]


def gwas_mv_run(bfile, Y, out_prefix, cov=None, min_maf=0.01, no_loco=False, scan_chr=None, grm_alg=1,
                solver="em", max_iter=1000, tol=1e-8, em_min_iter=10, ai_max_iter=50, ai_tol=1e-8):
    """MultivariateModel::run's fixed-null Wald GWAS end to end on the GPU (Y: N x d over the .fam
    rows, NaN = missing); writes out_prefix.assoc.mv.txt and returns the run summary."""
    Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
    if Y.ndim == 1:
        Y = Y.reshape(-1, 1, order="F")
    N, d = Y.shape
    if cov is None:
        covf, nc = None, 0
    else:
        covf = np.asfortranarray(np.asarray(cov, dtype=np.float64).reshape(N, -1))
        nc = covf.shape[1]
    o = GwasMvOptions(float(min_maf), int(bool(no_loco)), None if scan_chr is None else str(scan_chr).encode(),
                      int(grm_alg), 1 if solver == "aireml" else 0, int(max_iter), float(tol), int(em_min_iter),
                      int(ai_max_iter), float(ai_tol))
    sm = GwasMvSummary()
    _check(lib().cs_gwas_mv_run(str(bfile).encode(), _p(Y, _dp), N, d, _p(covf, _dp), N, nc, ctypes.byref(o),
                                str(out_prefix).encode(), ctypes.byref(sm)))
    k = sm.n_models
    return dict(n_samples=sm.n_samples, n_traits=sm.n_traits, n_tested=sm.n_tested, loco=bool(sm.loco),
                models=[(sm.model_label[i].value.decode(),
                         np.array(sm.vg[i][:d * d]).reshape(d, d), np.array(sm.ve[i][:d * d]).reshape(d, d))
                        for i in range(k)],
                ms_grm=sm.ms_grm, ms_null=sm.ms_null, ms_scan=sm.ms_scan)
