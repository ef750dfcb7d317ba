// Device task layouts and kernel launchers (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tns {

constexpr int DEV_MAXR = 8;

// One strided block copy: dst[i] = (neg ? -1 : 1) * src[i] over a rank-r box,
// i enumerated row-major over dims (destination order).  mul / sh: the
// round-up division magic of each dimension (perm_magic), so that the kernel
// decodes a flat index with multiply-shifts instead of integer divisions.
struct PermDev {
    const void* src;
    void* dst;
    int64_t n;
    int32_t rank, neg;
    int32_t dims[DEV_MAXR];
    int32_t ss[DEV_MAXR];
    int32_t ds[DEV_MAXR];
    uint32_t mul[DEV_MAXR];
    int32_t sh[DEV_MAXR];
};

// Granlund-Montgomery round-up magic for 31-bit dividends: with
// l = ceil(log2 d), mul = floor(2^(31+l) / d) + 1 (< 2^32 for every d >= 1)
// and sh = 31 + l, floor(x / d) = (x * mul) >> sh for all 0 <= x < 2^31.
inline void perm_magic(uint32_t d, uint32_t* mul, int32_t* sh) {
    int l = 0;
    while ((uint64_t(1) << l) < d) l++;
    *mul = uint32_t((uint64_t(1) << (31 + l)) / d + 1);
    *sh = 31 + l;
}

struct GemmDev {
    const void* a;
    const void* b;
    void* c;
    int32_t m, n, k, lda, ldb, ldc, ta, tb;
    double alpha;
};

struct TileDev {
    int32_t task, tm, tn, pad;
};

struct FillDev {
    void* p;
    int64_t n;
};

struct QrDev {
    const void* in;
    void* f1;
    void* f2;
    double* work;  // global scratch (mx*nx + k doubles) when not LDS-resident
    int32_t m, n, lq, pad;
};

struct SvdDev {
    const void* in;
    void* u;
    double* s;  // singular values are always stored in double
    void* v;
    double* work;  // global scratch (mx*kx + kx*kx + kx doubles) when not LDS-resident
    int32_t m, n;
    int32_t max_sweeps;  // Jacobi sweep limit; a sector not converged by then gets NaN singular values
    int32_t pad1;
};
// default Jacobi sweep limit (the host can lower it to force the
// non-convergence path in tests)
constexpr int SVD_MAX_SWEEPS = 60;

// dtype: 0 = float64, 1 = float32
// Element chunk per workgroup of the permute kernel.
constexpr int PERM_CHUNK = 2048;
void launch_permute(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg, int dtype,
                    hipStream_t stream);
// same copy with a float64 source (singular values into S)
void launch_permute_from_f64(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg,
                             int dtype, hipStream_t stream);
void launch_fill(const FillDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg, int dtype,
                 hipStream_t stream);
void launch_permute_to_f64(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg,
                           hipStream_t stream);
// dtype: 0 float64, 1 float32, 2 float32 operands as split-bf16 x3 (hi*hi +
// hi*lo + lo*hi), 3 split-bf16 x6 (three-term split, all products of order
// <= 2); float32 accumulation for 1..3
// variant: tile 0 = 32x32, 1 = 64x64, 2 = 16x16 (split-bf16: 32x32 only)
void launch_gemm(const GemmDev* tasks, const TileDev* tiles, int64_t num_tiles, int dtype, int variant,
                 hipStream_t stream);
constexpr int GEMM_VARIANTS = 3;
constexpr int GEMM_TILE_OF[GEMM_VARIANTS] = {32, 64, 16};
// QR / SVD: one workgroup per block.  lds_bytes > 0 selects the LDS-resident
// kernel with that much dynamic LDS (every task of the launch must fit);
// 0 selects the kernel working in each task's global scratch.
constexpr int QR_WG = 512;
constexpr int SVD_WG = 256;
// largest LDS working set per workgroup (gfx950: 160 KiB per CU)
// LDS size classes: a launch reserves the largest working set among its
// tasks, so tasks are grouped by class (3, 2 or 1 workgroups per CU)
constexpr int64_t LDS_CLASS[3] = {48 * 1024, 76 * 1024, 152 * 1024};
constexpr int64_t QR_LDS_MAX = LDS_CLASS[2];
constexpr int64_t SVD_LDS_MAX = LDS_CLASS[2];
// scratch of a QR / LQ task beyond LDS: the block and its taus
inline int64_t qr_doubles(int64_t m, int64_t n) { return m * n + (m < n ? m : n); }
inline int64_t svd_doubles(int64_t m, int64_t n) {
    const int64_t mx = m > n ? m : n, kx = m < n ? m : n;
    return mx * kx + kx * kx + kx;
}
// variant 1: blocked Householder with the panel in LDS (global scratch
// blocks with at most QB_MAX_ROWS rows); otherwise LDS (lds_bytes > 0) or
// unblocked global scratch
void launch_qr(const QrDev* tasks, int64_t num, int64_t lds_bytes, int variant, int dtype, hipStream_t stream);
constexpr int64_t QB_LDS = 156 * 1024;
// rows for which the blocked kernel's panel keeps at least 4 columns
constexpr int64_t QB_MAX_ROWS = (QB_LDS - 8 * (16 * 16 * 2 + 16 * 128 + 16)) / (8 * 4);
// rpt > 0: register-resident kernel for blocks with at most 8 * rpt rows
// (rpt in {8, 16, 24}), V recovered from U; otherwise LDS (lds_bytes > 0)
// or global scratch.
void launch_svd(const SvdDev* tasks, int64_t num, int64_t lds_bytes, int rpt, int dtype, hipStream_t stream);

// Copy n_i doubles from src_i to dst + off_i for every task, one workgroup
// per task: gathers the singular values of a batch of decompositions into
// one (host-mapped) buffer with a single launch.
struct GatherDev {
    const double* src;
    int64_t n;
    int64_t off;
};
void launch_gather_f64(const GatherDev* tasks, int64_t num, double* dst, hipStream_t stream);
constexpr int SVD_REG_MAX_ROWS = 192;

}  // namespace tns
