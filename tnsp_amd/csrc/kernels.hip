// CDNA4 (gfx950) kernels for the block-sparse tensor path.
//
//  permute  : batched strided block copy with the fermion sign folded in; one
//             flat element space over all tasks, destination-ordered so that
//             stores coalesce.
//  gemm     : grouped GEMM over heterogeneous sector blocks, one workgroup per
//             32x32 output tile, 4 waves x MFMA 16x16x4 (fp64 / fp32).
//  qr       : batched Householder QR / LQ, one workgroup per sector block.
//  svd      : batched one-sided Jacobi SVD (always accumulated in fp64), one
//             workgroup per sector block, singular values sorted descending.
#include "kernels.hpp"

#include <stdexcept>
#include <string>
#include <type_traits>

namespace tns {

namespace {

constexpr int WG = 256;
constexpr int WAVE = 64;

// Device pointers read from task records are generic to the compiler, so
// accesses through them compile to flat instructions, and a flat load is
// counted by the LDS counter as well as the vector-memory one: a wait for an
// LDS read then also waits for every global load in flight (the prefetch of
// a pipelined loop included).  Every task buffer is global memory; accesses
// through these pointers compile to global_load / global_store.
template<typename T>
using gptr = __attribute__((address_space(1))) T*;
template<typename T>
__device__ inline gptr<T> to_global(T* p) {
    return (gptr<T>)(p);
}

template<typename T>
__device__ inline T neg_if(T v, int neg) {
    return neg ? -v : v;
}

// ---------------------------------------------------------------- permute
// floor(x / d) for 0 <= x < 2^31 from d's round-up magic (perm_magic)
__device__ inline uint32_t magic_div(uint32_t x, uint32_t mul, int sh) {
    return uint32_t((uint64_t(x) * mul) >> sh);
}

// source and destination offsets of element `local` of a task, decoded
// row-major over dims (destination order, last dimension fastest)
__device__ inline void perm_decode(uint32_t local, int rank, const int32_t* dims, const int32_t* ss,
                                   const int32_t* ds, const uint32_t* mul, const int32_t* sh, int64_t& a,
                                   int64_t& c) {
    a = 0;
    c = 0;
#pragma unroll
    for (int r = DEV_MAXR - 1; r >= 0; r--) {
        if (r < rank) {
            const uint32_t q = magic_div(local, mul[r], sh[r]);
            const uint32_t i = local - q * uint32_t(dims[r]);
            a += int64_t(i) * ss[r];
            c += int64_t(i) * ds[r];
            local = q;
        }
    }
}

// Each workgroup copies PERM_CHUNK consecutive elements of the flat element
// space of all tasks; each thread handles PERM_CHUNK / WG of them, WG apart,
// all loads issued before the stores.  A chunk inside one task (the common
// case: most elements belong to tasks larger than a chunk) reads that task's
// descriptor once for the whole workgroup; a chunk spanning several tasks
// locates each element's task through the prefix sums.
template<typename TS, typename T>
__global__ __launch_bounds__(WG) void permute_kernel(const PermDev* __restrict__ tasks, const int64_t* __restrict__ prefix,
                                                     const int32_t* __restrict__ wg_first) {
    constexpr int PER = PERM_CHUNK / WG;
    const int64_t base = int64_t(blockIdx.x) * PERM_CHUNK;
    const int t0 = wg_first[blockIdx.x];
    const int t_end = wg_first[gridDim.x];
    const int64_t p0 = prefix[t0], p1 = prefix[t0 + 1];
    if (base + PERM_CHUNK <= p1 || t0 + 1 == t_end) {
        const PermDev& tk = tasks[t0];
        const int rank = tk.rank;
        int32_t dims[DEV_MAXR], ss[DEV_MAXR], ds[DEV_MAXR], sh[DEV_MAXR];
        uint32_t mul[DEV_MAXR];
#pragma unroll
        for (int r = 0; r < DEV_MAXR; r++) {
            dims[r] = tk.dims[r];
            ss[r] = tk.ss[r];
            ds[r] = tk.ds[r];
            mul[r] = tk.mul[r];
            sh[r] = tk.sh[r];
        }
        const gptr<const TS> src = to_global(static_cast<const TS*>(tk.src));
        const gptr<T> dst = to_global(static_cast<T*>(tk.dst));
        const bool neg = tk.neg != 0;
        const int64_t left = p1 - base;  // elements of task t0 from base on
        const uint32_t off = uint32_t(base - p0);
        int64_t so[PER], dof[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int e = threadIdx.x + j * WG;
            so[j] = dof[j] = 0;
            if (e < left) perm_decode(off + uint32_t(e), rank, dims, ss, ds, mul, sh, so[j], dof[j]);
        }
        T v[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) v[j] = (threadIdx.x + j * WG < left) ? T(src[so[j]]) : T(0);
#pragma unroll
        for (int j = 0; j < PER; j++)
            if (threadIdx.x + j * WG < left) dst[dof[j]] = neg_if(v[j], neg ? 1 : 0);
        return;
    }
    int t = t0;
    gptr<const TS> src[PER];
    gptr<T> dst[PER];
    int64_t so[PER], dof[PER];
    bool neg[PER], ok[PER];
    int64_t end = p1;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int64_t e = base + threadIdx.x + int64_t(j) * WG;
        // advance to the task containing e (tasks are contiguous in e)
        while (end <= e && t < t_end) {
            t++;
            if (t < t_end) end = prefix[t + 1];
        }
        ok[j] = t < t_end;
        so[j] = dof[j] = 0;
        src[j] = nullptr;
        dst[j] = nullptr;
        neg[j] = false;
        if (ok[j]) {
            const PermDev& tk = tasks[t];
            perm_decode(uint32_t(e - prefix[t]), tk.rank, tk.dims, tk.ss, tk.ds, tk.mul, tk.sh, so[j], dof[j]);
            src[j] = to_global(static_cast<const TS*>(tk.src));
            dst[j] = to_global(static_cast<T*>(tk.dst));
            neg[j] = tk.neg != 0;
        }
    }
    T v[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) v[j] = ok[j] ? T(src[j][so[j]]) : T(0);
#pragma unroll
    for (int j = 0; j < PER; j++)
        if (ok[j]) dst[j][dof[j]] = neg_if(v[j], neg[j] ? 1 : 0);
}

template<typename T>
__global__ __launch_bounds__(WG) void fill_kernel(const FillDev* __restrict__ tasks, const int64_t* __restrict__ prefix,
                                                  const int32_t* __restrict__ wg_first) {
    const int64_t base = int64_t(blockIdx.x) * PERM_CHUNK;
    int t = wg_first[blockIdx.x];
    for (int j = 0; j < PERM_CHUNK / WG; j++) {
        const int64_t e = base + threadIdx.x + int64_t(j) * WG;
        int64_t end = prefix[t + 1];
        while (end <= e) {
            t++;
            if (t >= wg_first[gridDim.x]) return;
            end = prefix[t + 1];
        }
        to_global(static_cast<T*>(tasks[t].p))[e - prefix[t]] = T(0);
    }
}

// ------------------------------------------------------------------- gemm
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template<typename T>
struct Mfma;
// Accumulator rows: lane l, register i of the 16x16 output tile holds
// D(row(l, i), l % 16).  For the fp64 MFMA the four registers of a lane are
// rows l/16 + 4 i; for the fp32 MFMA they are rows 4 (l/16) + i.
template<>
struct Mfma<double> {
    using acc_t = d4;
    __device__ static int row(int lane, int i) { return lane / 16 + 4 * i; }
    __device__ static acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    __device__ static acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template<>
struct Mfma<float> {
    using acc_t = f4;
    __device__ static int row(int lane, int i) { return 4 * (lane / 16) + i; }
    __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    __device__ static acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

constexpr int KB = 16;

// split-bf16 products on the bf16 MFMA (16x16x16, fp32 accumulation).  An
// fp32 value is split into NTERM bf16 terms (round to nearest even, each
// term taken from the remainder of the previous ones); all term products of
// order i + j < NTERM are accumulated, smallest first.
typedef short s4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_rne(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

__device__ inline float bf16_value(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

template<int NTERM>
__device__ inline void bf16_split(float x, s4* terms, int j) {
#pragma unroll
    for (int t = 0; t < NTERM; t++) {
        const unsigned short h = bf16_rne(x);
        terms[t][j] = static_cast<short>(h);
        x -= bf16_value(h);
    }
}

// C tile (16 WR FM) x (16 WC FN) per workgroup of WR x WC waves; wave
// (wr, wc) owns FM x FN fragments of 16x16.  MFMA 16x16x4 operand layout:
// lane l holds A(l%16, l/16) and B(l/16, l%16); accumulator rows as given by
// Mfma<T>::row.  Variants: 32x32 (2x2 waves, 1x1 fragments), 64x64 (2x2
// waves, 2x2 fragments: each operand fragment feeds two MFMAs), 16x16 (one
// wave, for sectors of at most 16 rows and columns).
// The K loop is software-pipelined: the global loads of k-block kb + 1 are
// issued into registers before the MFMAs of block kb run from one LDS
// buffer, then land in the other buffer; one barrier per k-block.
template<typename T, int WR, int WC, int FM, int FN>
__global__ __launch_bounds__(64 * WR * WC) void gemm_kernel(const GemmDev* __restrict__ tasks,
                                                           const TileDev* __restrict__ tiles) {
    constexpr int NT = 64 * WR * WC, TM = 16 * WR * FM, TN = 16 * WC * FN;
    constexpr int LA = (TM * KB + NT - 1) / NT, LB = (KB * TN + NT - 1) / NT;
    __shared__ T As[2][TM][KB + 1];
    __shared__ T Bs[2][KB][TN + 1];
    const TileDev tile = tiles[blockIdx.x];
    const GemmDev g = tasks[tile.task];
    const gptr<const T> A = to_global(static_cast<const T*>(g.a));
    const gptr<const T> B = to_global(static_cast<const T*>(g.b));
    const int row0 = tile.tm * TM, col0 = tile.tn * TN;
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const int wr = (wave / WC) * 16 * FM, wc = (wave % WC) * 16 * FN;
    typename Mfma<T>::acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::zero();
    // element u of this thread's share of a k-block: A tile (i, p) and B
    // tile (p, j), indexed so that consecutive threads walk the contiguous
    // direction of the stored operand
    auto a_pos = [&](int e, int& i, int& p) {
        if (g.ta) {
            i = e % TM;
            p = e / TM;
        } else {
            p = e % KB;
            i = e / KB;
        }
    };
    auto b_pos = [&](int e, int& p, int& j) {
        if (g.tb) {
            p = e % KB;
            j = e / KB;
        } else {
            j = e % TN;
            p = e / TN;
        }
    };
    T ra[LA], rb[LB];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < LA; u++) {
            const int e = tid + u * NT;
            int i, p;
            a_pos(e, i, p);
            const int gi = row0 + i, gp = k0 + p;
            ra[u] = (e < TM * KB && gi < g.m && gp < g.k)
                        ? (g.ta ? A[int64_t(gp) * g.lda + gi] : A[int64_t(gi) * g.lda + gp])
                        : T(0);
        }
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int e = tid + u * NT;
            int p, j;
            b_pos(e, p, j);
            const int gp = k0 + p, gj = col0 + j;
            rb[u] = (e < KB * TN && gp < g.k && gj < g.n)
                        ? (g.tb ? B[int64_t(gj) * g.ldb + gp] : B[int64_t(gp) * g.ldb + gj])
                        : T(0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < LA; u++) {
            const int e = tid + u * NT;
            int i, p;
            a_pos(e, i, p);
            if (e < TM * KB) As[buf][i][p] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int e = tid + u * NT;
            int p, j;
            b_pos(e, p, j);
            if (e < KB * TN) Bs[buf][p][j] = rb[u];
        }
    };
    const int nk = (g.k + KB - 1) / KB;
    load(0);
    store(0);
    __syncthreads();
    for (int kb = 0; kb < nk; kb++) {
        const int cur = kb & 1;
        if (kb + 1 < nk) load((kb + 1) * KB);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            T a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; i++) a[i] = As[cur][wr + 16 * i + lane % 16][kk + lane / 16];
#pragma unroll
            for (int j = 0; j < FN; j++) b[j] = Bs[cur][kk + lane / 16][wc + 16 * j + lane % 16];
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::step(a[i], b[j], acc[i][j]);
        }
        // buffer cur ^ 1 was last read in iteration kb - 1, before that
        // iteration's barrier
        if (kb + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    const gptr<T> C = to_global(static_cast<T*>(g.c));
    const T alpha = T(g.alpha);
#pragma unroll
    for (int fi = 0; fi < FM; fi++)
#pragma unroll
        for (int fj = 0; fj < FN; fj++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = row0 + wr + 16 * fi + Mfma<T>::row(lane, i);
                const int c = col0 + wc + 16 * fj + lane % 16;
                if (r < g.m && c < g.n) C[int64_t(r) * g.ldc + c] = alpha * acc[fi][fj][i];
            }
}

// Same tiling as the 32x32 gemm_kernel<float>, with each 16-deep k step done as
// split-bf16 MFMA products.  Lane l holds A(l % 16, 4 (l / 16) + j) and
// B(4 (l / 16) + j, l % 16), j = 0..3; accumulator rows 4 (l / 16) + i.
template<int NTERM>
__global__ __launch_bounds__(WG) void gemm_kernel_bf16(const GemmDev* __restrict__ tasks,
                                                       const TileDev* __restrict__ tiles) {
    constexpr int BT = 32;
    __shared__ float As[BT][KB + 1];
    __shared__ float Bs[KB][BT + 1];
    const TileDev tile = tiles[blockIdx.x];
    const GemmDev g = tasks[tile.task];
    const gptr<const float> A = to_global(static_cast<const float*>(g.a));
    const gptr<const float> B = to_global(static_cast<const float*>(g.b));
    const int row0 = tile.tm * BT, col0 = tile.tn * BT;
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const int wr = (wave / 2) * 16, wc = (wave % 2) * 16;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.k; k0 += KB) {
        for (int e = tid; e < BT * KB; e += WG) {
            int i, p;
            if (g.ta) {
                i = e % BT;
                p = e / BT;
            } else {
                p = e % KB;
                i = e / KB;
            }
            const int gi = row0 + i, gp = k0 + p;
            float v = 0.f;
            if (gi < g.m && gp < g.k) v = g.ta ? A[int64_t(gp) * g.lda + gi] : A[int64_t(gi) * g.lda + gp];
            As[i][p] = v;
        }
        for (int e = tid; e < BT * KB; e += WG) {
            int p, j;
            if (g.tb) {
                p = e % KB;
                j = e / KB;
            } else {
                j = e % BT;
                p = e / BT;
            }
            const int gp = k0 + p, gj = col0 + j;
            float v = 0.f;
            if (gp < g.k && gj < g.n) v = g.tb ? B[int64_t(gj) * g.ldb + gp] : B[int64_t(gp) * g.ldb + gj];
            Bs[p][j] = v;
        }
        __syncthreads();
        s4 at[NTERM], bt[NTERM];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bf16_split<NTERM>(As[wr + lane % 16][4 * (lane / 16) + j], at, j);
            bf16_split<NTERM>(Bs[4 * (lane / 16) + j][wc + lane % 16], bt, j);
        }
#pragma unroll
        for (int order = NTERM - 1; order >= 0; order--) {
#pragma unroll
            for (int i = 0; i <= order; i++)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(at[i], bt[order - i], acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const gptr<float> C = to_global(static_cast<float*>(g.c));
    const float alpha = float(g.alpha);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int r = row0 + wr + 4 * (lane / 16) + i;
        const int c = col0 + wc + lane % 16;
        if (r < g.m && c < g.n) C[int64_t(r) * g.ldc + c] = alpha * acc[i];
    }
}

// --------------------------------------------------------- reductions
constexpr int DPP_ROW_ROR1 = 0x121, DPP_ROW_ROR2 = 0x122, DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;

// a double moved across lanes by a DPP row shift (two 32-bit halves)
template<int CTRL>
__device__ inline double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(u)), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(u >> 32)), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo));
}

// v of lane l plus v of lane l ^ 16 (SWAP = 16) or l ^ 32 (SWAP = 32), by the
// gfx950 row / half swaps: permlane16_swap(x, x) leaves each lane one of
// {x_l, x_(l^16)} in each of its two results (even rows keep their own value
// in the first and receive the partner's in the second, odd rows the other
// way round), so the two results add up to the pair sum in every lane; the
// same holds for permlane32_swap with halves.  VALU only, no LDS.
template<int SWAP>
__device__ inline double swap_sum(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(u), hi = unsigned(u >> 32);
    unsigned lo0, lo1, hi0, hi1;
    if constexpr (SWAP == 16) {
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        lo0 = l[0];
        lo1 = l[1];
        hi0 = h[0];
        hi1 = h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        lo0 = l[0];
        lo1 = l[1];
        hi0 = h[0];
        hi1 = h[1];
    }
    const double a = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi0) << 32) | lo0);
    const double b = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi1) << 32) | lo1);
    return a + b;
}

// butterfly sum over the 8 lanes {l ^ 8k}: l ^ 8 by a DPP rotate by 8 within
// the 16-lane row, l ^ 16 and l ^ 32 by the row and half swaps; every lane
// ends with the same value (the additions commute)
__device__ inline double sum8(double v) {
    v += dpp_f64<DPP_ROW_ROR8>(v);
    v = swap_sum<16>(v);
    return swap_sum<32>(v);
}

// sum over the 16 lanes of a DPP row by rotations 8, 4, 2, 1: after the
// step with rotation r every lane holds the sum of its 16/r-lane orbit, and
// the two lanes of each addition hold the same two values, so every lane of
// the row ends with bitwise the same sum (IEEE addition commutes)
__device__ inline double row_sum16(double v) {
    v += dpp_f64<DPP_ROW_ROR8>(v);
    v += dpp_f64<DPP_ROW_ROR4>(v);
    v += dpp_f64<DPP_ROW_ROR2>(v);
    v += dpp_f64<DPP_ROW_ROR1>(v);
    return v;
}

// sum within aligned groups of G lanes (G = 16, 32 or 64), VALU only: every
// lane of the group ends with bitwise the same value
__device__ inline double group_sum(double v, int G) {
    v = row_sum16(v);
    if (G >= 32) v = swap_sum<16>(v);
    if (G >= 64) v = swap_sum<32>(v);
    return v;
}

__device__ inline double wave_sum(double v) { return group_sum(v, WAVE); }

// maximum over the wave (rotations within rows, then the row / half swaps)
__device__ inline double wave_max(double v) {
    v = fmax(v, dpp_f64<DPP_ROW_ROR8>(v));
    v = fmax(v, dpp_f64<DPP_ROW_ROR4>(v));
    v = fmax(v, dpp_f64<DPP_ROW_ROR2>(v));
    v = fmax(v, dpp_f64<DPP_ROW_ROR1>(v));
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(u), hi = unsigned(u >> 32);
    {
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        const double a = __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(h[0])) << 32) | unsigned(l[0]));
        const double b = __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(h[1])) << 32) | unsigned(l[1]));
        v = fmax(a, b);
    }
    const unsigned long long u2 = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo2 = unsigned(u2), hi2 = unsigned(u2 >> 32);
    const auto l = __builtin_amdgcn_permlane32_swap(lo2, lo2, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi2, hi2, false, false);
    const double a = __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(h[0])) << 32) | unsigned(l[0]));
    const double b = __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(h[1])) << 32) | unsigned(l[1]));
    return fmax(a, b);
}

// Scale exponent of a block for the Jacobi SVDs: the block is multiplied by
// 2^-e (exact) so that its largest entry lies in [1, 2); squared column norms
// then neither overflow nor underflow for every column that matters.  0 for
// an all-zero or non-finite block.
__device__ inline int jacobi_scale_exponent(double amax) {
    return (amax > 0.0 && amax <= 1.79e308) ? ilogb(amax) : 0;
}

// Squared column norm (after the scaling) at or below which a column counts
// as zero: it takes part in no rotation and gets singular value 0.  Its true
// singular value is below 1e-100 of the largest entry.  A column of an
// exactly rank-deficient block shrinks by a rounding factor with every
// rotation against the columns spanning its space, and would otherwise keep
// rotating down into the underflow range, where its relative cosines stay
// of order one and the sweeps never end.
constexpr double JACOBI_QUIET = 1e-200;

// A column whose squared norm falls to JACOBI_DEAD2 times the block's squared
// Frobenius norm (1e-15 of the norm, about 4.5 ulp) is rounding noise: it
// takes part in no further rotation.  The columns of an exactly
// rank-deficient block that shrink into the null space are noise inside the
// span of the others, so their cosines with those stay of order one; they
// shrink only slowly as they mix among themselves, and would otherwise keep
// the sweeps going for dozens of sweeps or forever.  Their singular values
// come out at that noise level, as LAPACK's do.
constexpr double JACOBI_DEAD2 = 1e-30;

template<int NT>
__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < NT / WAVE; w++) t += red[w];
    return t;
}

// lanes per group for work along a column of length len
__device__ inline int group_lanes(int len) { return len <= 48 ? 16 : (len <= 96 ? 32 : 64); }

// Jacobi rotation (cs, sn) zeroing the off-diagonal of [[saa, sab], [sab,
// sbb]] (Rutishauser's formulas, sab != 0).  The two reciprocals after the
// exact zeta use the hardware reciprocal / reciprocal square root refined by
// two Newton steps each: their arguments lie in [1, inf), and
// cs^2 + sn^2 = cs^2 (1 + t^2) stays 1 to rounding, which is what keeps the
// accumulated transformation orthogonal.
__device__ inline double refined_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ inline double refined_rsqrt(double y) {
    double r = __builtin_amdgcn_rsq(y);
    r = r * fma(-0.5 * y * r, r, 1.5);
    return r * fma(-0.5 * y * r, r, 1.5);
}

// t = tan of the angle: after a' = cs a - sn b, b' = sn a + cs b the squared
// norms are exactly saa - t sab and sbb + t sab (t solves t^2 + 2 zeta t = 1)
__device__ inline void jacobi_rotation(double saa, double sbb, double sab, double& cs, double& sn, double& t) {
    const double zeta = (sbb - saa) / (2.0 * sab);
    // for |zeta| beyond 1e150, zeta^2 would overflow: t = 1 / (2 zeta) there
    t = fabs(zeta) > 1e150 ? 0.5 / zeta : copysign(refined_rcp(fabs(zeta) + sqrt(fma(zeta, zeta, 1.0))), zeta);
    cs = refined_rsqrt(fma(t, t, 1.0));
    sn = cs * t;
}

__device__ inline void jacobi_rotation(double saa, double sbb, double sab, double& cs, double& sn) {
    double t;
    jacobi_rotation(saa, sbb, sab, cs, sn, t);
}

// Squared column norm after a rotation, tracked instead of recomputed: the
// update keeps its rounding relative to the old norm, so a column that lost
// all but 1e-3 of it is marked unknown (-1: no rotation until the next
// exact refresh) rather than carrying a large relative error.
__device__ inline double jacobi_tracked(double old_n, double new_n) {
    return new_n < 1e-3 * old_n ? -1.0 : new_n;
}

// --------------------------------------------------------------------- qr
// Householder QR of X (mx x nx, column-major), X = M (qr) or X = M^T (lq);
// R is written out first, then Q is formed in place from the reflectors
// (backward accumulation, as LAPACK's orgqr does).  X lives in LDS when it
// fits, otherwise in global scratch.
// Householder reflector of a column with diagonal entry alpha and squared
// norm sigma below it: H = I - tau v v^T, v = (1, x / (alpha - beta)),
// H (alpha, x) = (beta, 0).  Every thread computes the same values.
__device__ inline void householder(double alpha, double sigma, double& tau, double& beta) {
    if (sigma == 0.0) {
        tau = 0.0;
        beta = alpha;
    } else {
        beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
        tau = (beta - alpha) / beta;
    }
}

// Apply the reflector of column j (unit diagonal implied, entries below j in
// cj) to columns j + 1 .. nc - 1 of X (column stride ld), one group of G
// lanes per column.  The group that updates column j + 1 also sums the
// squares of its new entries below row j + 1 (the next reflector's sigma)
// into *next (written when j + 1 < nc).
__device__ inline void qr_update_next(const double* cj, double* X, int ld, int j, int nc, double tj, int g, int gl,
                                      int G, int NG, double* next) {
    const int rows = ld;
    if (tj != 0.0) {
        for (int c = j + 1 + g; c < nc; c += NG) {
            double* cc = X + int64_t(c) * ld;
            double d = 0;
            for (int r = j + 1 + gl; r < rows; r += G) d += cj[r] * cc[r];
            d = group_sum(d, G) + cc[j];
            const double f = tj * d;
            if (gl == 0) cc[j] -= f;
            double sq = 0;
            for (int r = j + 1 + gl; r < rows; r += G) {
                const double v = cc[r] - f * cj[r];
                cc[r] = v;
                if (r > j + 1) sq += v * v;
            }
            if (c == j + 1) {
                sq = group_sum(sq, G);
                if (gl == 0) *next = sq;
            }
        }
    } else if (g == 0 && j + 1 < nc) {
        const double* cc = X + int64_t(j + 1) * ld;
        double sq = 0;
        for (int r = j + 2 + gl; r < rows; r += G) sq += cc[r] * cc[r];
        sq = group_sum(sq, G);
        if (gl == 0) *next = sq;
    }
}

// Lanes per column group of the Householder updates: the widest group (of
// 16, 32, 64, no wider than the column needs) that still gives every
// column to be updated a group of its own, so that a reflector's update is
// one round of groups instead of several in a row.
__device__ inline int qr_group_lanes(int rows, int cols) {
    int G = group_lanes(rows);
    while (G > 16 && QR_WG / G < cols) G /= 2;
    return G;
}

template<typename T>
__device__ __forceinline__ void qr_body(const QrDev& tk, double* X, double* tau, double* red, double* bc) {
    const int m = tk.m, n = tk.n, lq = tk.lq;
    const int mx = lq ? n : m, nx = lq ? m : n, k = mx < nx ? mx : nx;
    const T* in = static_cast<const T*>(tk.in);
    const int tid = threadIdx.x;
    const int G = qr_group_lanes(mx, nx), NG = QR_WG / G, g = tid / G, gl = tid % G;
    for (int64_t e = tid; e < int64_t(mx) * nx; e += QR_WG) {
        const int r = int(e % mx), c = int(e / mx);
        X[e] = double(lq ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]);
    }
    __syncthreads();
    // squared norm below the diagonal of the column to reduce: a block sum
    // for column 0, then from the group that updates the next column
    double sigma;
    {
        double s = 0;
        for (int r = 1 + tid; r < mx; r += QR_WG) s += X[r] * X[r];
        sigma = block_sum<QR_WG>(s, red);
    }
    for (int j = 0; j < k; j++) {
        double* cj = X + int64_t(j) * mx;
        double tj, beta;
        householder(cj[j], sigma, tj, beta);
        if (tj != 0.0) {
            const double scale = 1.0 / (cj[j] - beta);
            for (int r = j + 1 + tid; r < mx; r += QR_WG) cj[r] *= scale;
        }
        // every thread has read cj[j]; the reflector below it is scaled
        __syncthreads();
        if (tid == 0) {
            cj[j] = beta;
            tau[j] = tj;
        }
        qr_update_next(cj, X, mx, j, nx, tj, g, gl, G, NG, bc);
        __syncthreads();
        sigma = bc[0];
    }
    T* f1 = static_cast<T*>(tk.f1);
    T* f2 = static_cast<T*>(tk.f2);
    // R (k x nx, upper trapezoid)
    if (!lq) {
        for (int64_t e = tid; e < int64_t(k) * n; e += QR_WG) {
            const int r = int(e / n), c = int(e % n);
            f2[e] = T(r <= c ? X[int64_t(c) * mx + r] : 0.0);
        }
    } else {
        for (int64_t e = tid; e < int64_t(m) * k; e += QR_WG) {
            const int i = int(e / k), r = int(e % k);
            f1[e] = T(r <= i ? X[int64_t(i) * mx + r] : 0.0);
        }
    }
    __syncthreads();
    for (int j = k - 1; j >= 0; j--) {
        const double tj = tau[j];
        const double* v = X + int64_t(j) * mx;
        if (tj != 0.0) {
            for (int c = j + 1 + g; c < k; c += NG) {
                double* qc = X + int64_t(c) * mx;
                double d = 0;
                for (int r = j + 1 + gl; r < mx; r += G) d += v[r] * qc[r];
                d = group_sum(d, G) + qc[j];
                const double f = tj * d;
                if (gl == 0) qc[j] -= f;
                for (int r = j + 1 + gl; r < mx; r += G) qc[r] -= f * v[r];
            }
        }
        __syncthreads();
        double* cj = X + int64_t(j) * mx;
        for (int r = tid; r < mx; r += QR_WG) {
            if (r > j)
                cj[r] *= -tj;
            else if (r == j)
                cj[r] = 1.0 - tj;
            else
                cj[r] = 0.0;
        }
        __syncthreads();
    }
    if (!lq) {
        for (int64_t e = tid; e < int64_t(m) * k; e += QR_WG) {
            const int r = int(e / k), c = int(e % k);
            f1[e] = T(X[int64_t(c) * mx + r]);
        }
    } else {
        for (int64_t e = tid; e < int64_t(k) * n; e += QR_WG) {
            const int r = int(e / n), i = int(e % n);
            f2[e] = T(X[int64_t(r) * mx + i]);
        }
    }
}

template<typename T>
__global__ __launch_bounds__(QR_WG) void qr_kernel_lds(const QrDev* __restrict__ tasks) {
    extern __shared__ double smem[];
    __shared__ double red[QR_WG / WAVE];
    __shared__ double bc[1];
    const QrDev tk = tasks[blockIdx.x];
    const int mx = tk.lq ? tk.n : tk.m, nx = tk.lq ? tk.m : tk.n;
    qr_body<T>(tk, smem, smem + int64_t(mx) * nx, red, bc);
}

template<typename T>
__global__ __launch_bounds__(QR_WG) void qr_kernel_glb(const QrDev* __restrict__ tasks) {
    __shared__ double red[QR_WG / WAVE];
    __shared__ double bc[1];
    const QrDev tk = tasks[blockIdx.x];
    const int mx = tk.lq ? tk.n : tk.m, nx = tk.lq ? tk.m : tk.n;
    qr_body<T>(tk, tk.work, tk.work + int64_t(mx) * nx, red, bc);
}

// Blocked Householder QR for blocks that do not fit LDS (X in global
// scratch, column-major, as in qr_body).  Panels of nb <= 16 columns are
// factored in LDS (same reflector convention as qr_body), then the panel's
// block reflector H = I - V T V^T (T from the LAPACK forward / columnwise
// recurrence) is applied to the trailing columns on the fp64 MFMA as
// W = V^T A, W = T^T W, A -= V W: two passes over the trailing matrix per
// panel instead of two per column.  Q is formed backwards block by block
// (A -= V (T (V^T A)) on identity-initialised columns).
constexpr int QB_NB = 16, QB_WC = 128;
constexpr int64_t QB_EXTRA = (QB_NB * QB_NB * 2 + QB_NB * QB_WC + QB_NB) * 8;

__device__ inline int qb_panel_width(int mx) {
    const int64_t vb = QB_LDS - QB_EXTRA;
    const int64_t w = vb / (8 * int64_t(mx));
    return int(w < QB_NB ? w : QB_NB);
}

// T (nb x nb, upper) of H = I - V T V^T from the unit-lower V (rows x nb,
// column-major) and taus; Gm scratch nb x nb.  One wave owns row j of T.
__device__ inline void qb_build_t(const double* V, int rows, int nb, const double* taus, double* Tm, double* Gm) {
    const int tid = threadIdx.x;
    // Gram entries Gm[l][i] = v_l . v_i for l < i, 4 lanes per pair
    const int npair = nb * (nb - 1) / 2;
    for (int base = 0; base < npair * 4; base += QR_WG) {
        const int e = base + tid;
        double acc = 0.0;
        int l = -1, i = -1;
        if (e < npair * 4) {
            int q = e / 4;
            i = 1;
            while (q >= i) {
                q -= i;
                i++;
            }
            l = q;
            for (int r = i + (e % 4); r < rows; r += 4) acc += V[int64_t(l) * rows + r] * V[int64_t(i) * rows + r];
        }
        acc += __shfl_xor(acc, 1, WAVE);
        acc += __shfl_xor(acc, 2, WAVE);
        if (e < npair * 4 && e % 4 == 0) Gm[l * QB_NB + i] = acc;
    }
    __syncthreads();
    if (tid < nb) {
        const int j = tid;
        for (int i = 0; i < nb; i++) {
            double v;
            if (i < j) {
                v = 0.0;
            } else if (i == j) {
                v = taus[i];
            } else {
                double sacc = 0.0;
                for (int l = j; l < i; l++) sacc += Tm[j * QB_NB + l] * Gm[l * QB_NB + i];
                v = -taus[i] * sacc;
            }
            Tm[j * QB_NB + i] = v;
        }
    }
    __syncthreads();
}

// A (rows p0.., columns [c_begin, c_end) of X) -= V op(T) V^T A with
// op(T) = T^T when trans (applying H^T) and T otherwise (applying H), in
// chunks of QB_WC columns on the fp64 MFMA (16x16x4): W = V^T A with rows as
// the reduction dimension (one wave per 16-column tile), op(T) W per column,
// then A -= V W computed as (W^T V^T) tiles of 16 columns x 16 rows so that
// the stores run down the columns.  A lives in global memory and each wave
// has few companions on its SIMD, so both passes keep many independent
// loads in flight: 64 rows of A per step in the first pass, the A tiles of
// four row tiles loaded before their MFMAs in the second.
__device__ inline void qb_apply(const double* V, int rows, int nb, const double* Tm, double* W, gptr<double> X,
                                int mx, int p0, int c_begin, int c_end, bool trans) {
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    constexpr int NW = QR_WG / WAVE;
    constexpr int RS = 16;  // 4-row MFMA steps per iteration of the first pass
    constexpr int RG = 4;   // 16-row tiles per task of the second pass
    const int rows_full = rows - rows % (4 * RS);
    for (int cb = c_begin; cb < c_end; cb += QB_WC) {
        const int ncol = min(QB_WC, c_end - cb), ntile = (ncol + 15) / 16;
        for (int t = wave; t < ntile; t += NW) {
            // lanes of columns beyond ncol (or reflectors beyond nb) read
            // column 0 (reflector 0) instead: their results land in entries
            // of W that are never stored
            const int c = t * 16 + lane % 16, j = lane % 16;
            const gptr<const double> xc = X + int64_t(cb + (c < ncol ? c : 0)) * mx + p0;
            const double* vj = V + int64_t(j < nb ? j : 0) * rows;
            const int rl = lane / 16;
            // four independent accumulator chains keep the MFMA pipeline fed
            // (one chain would wait out each MFMA's latency)
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            d4 acc4[4];
#pragma unroll
            for (int q = 0; q < 4; q++) acc4[q] = acc;
            for (int r0 = 0; r0 < rows_full; r0 += 4 * RS) {
                double av[RS], bv[RS];
#pragma unroll
                for (int u = 0; u < RS; u++) bv[u] = xc[r0 + 4 * u + rl];
#pragma unroll
                for (int u = 0; u < RS; u++) av[u] = vj[r0 + 4 * u + rl];
#pragma unroll
                for (int u = 0; u < RS; u++)
                    acc4[u % 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc4[u % 4], 0, 0, 0);
            }
            acc = (acc4[0] + acc4[1]) + (acc4[2] + acc4[3]);
            for (int r0 = rows_full; r0 < rows; r0 += 16) {
                double av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int r = r0 + 4 * u + rl;
                    av[u] = r < rows ? vj[r] : 0.0;
                    bv[u] = r < rows ? xc[r] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int jj = rl + 4 * i;
                if (jj < nb && c < ncol) W[jj * QB_WC + c] = acc[i];
            }
        }
        __syncthreads();
        for (int c = tid; c < ncol; c += QR_WG) {
            double w[QB_NB];
#pragma unroll
            for (int j = 0; j < QB_NB; j++) w[j] = j < nb ? W[j * QB_WC + c] : 0.0;
            for (int j = 0; j < nb; j++) {
                double sacc = 0.0;
                if (trans) {
                    for (int i = 0; i <= j; i++) sacc += Tm[i * QB_NB + j] * w[i];
                } else {
                    for (int i = j; i < nb; i++) sacc += Tm[j * QB_NB + i] * w[i];
                }
                W[j * QB_WC + c] = sacc;
            }
        }
        __syncthreads();
        const int nrt = (rows + 15) / 16, nrg = (nrt + RG - 1) / RG;
        for (int q = wave; q < nrg * ntile; q += NW) {
            const int rg = q / ntile, ct = q % ntile;
            const int ca = ct * 16 + lane % 16;
            double xv[RG][4];
#pragma unroll
            for (int g = 0; g < RG; g++) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = ct * 16 + lane / 16 + 4 * i, r = (rg * RG + g) * 16 + lane % 16;
                    xv[g][i] = (c < ncol && r < rows) ? X[int64_t(cb + c) * mx + p0 + r] : 0.0;
                }
            }
            double wa[QB_NB / 4];
#pragma unroll
            for (int s4 = 0; s4 < QB_NB / 4; s4++) {
                const int jk = 4 * s4 + lane / 16;
                wa[s4] = (jk < nb && ca < ncol) ? W[jk * QB_WC + ca] : 0.0;
            }
#pragma unroll
            for (int g = 0; g < RG; g++) {
                const int rb = (rg * RG + g) * 16 + lane % 16;
                d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s4 = 0; s4 < QB_NB / 4; s4++) {
                    if (4 * s4 < nb) {
                        const int jk = 4 * s4 + lane / 16;
                        const double b = (jk < nb && rb < rows) ? V[int64_t(jk) * rows + rb] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[s4], b, acc, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = ct * 16 + lane / 16 + 4 * i, r = rb;
                    if (c < ncol && r < rows) X[int64_t(cb + c) * mx + p0 + r] = xv[g][i] - acc[i];
                }
            }
        }
        __syncthreads();
    }
}

template<typename T>
__global__ __launch_bounds__(QR_WG) void qr_kernel_blocked(const QrDev* __restrict__ tasks) {
    extern __shared__ double sm[];
    __shared__ double red[QR_WG / WAVE];
    __shared__ double bc[1];
    const QrDev tk = tasks[blockIdx.x];
    const int m = tk.m, n = tk.n, lq = tk.lq;
    const int mx = lq ? n : m, nx = lq ? m : n, k = mx < nx ? mx : nx;
    const int nbmax = qb_panel_width(mx);
    if (nbmax < 1) return;  // the host routes only blocks with nbmax >= 4 here
    double* Tm = sm;
    double* Gm = Tm + QB_NB * QB_NB;
    double* W = Gm + QB_NB * QB_NB;
    double* taus = W + QB_NB * QB_WC;
    double* V = taus + QB_NB;
    const gptr<double> X = to_global(tk.work);
    const gptr<double> tau = X + int64_t(mx) * nx;
    const gptr<const T> in = to_global(static_cast<const T*>(tk.in));
    const int tid = threadIdx.x;
    const int G = 32, NG = QR_WG / G, g = tid / G, gl = tid % G;
    for (int64_t e = tid; e < int64_t(mx) * nx; e += QR_WG) {
        const int r = int(e % mx), c = int(e / mx);
        X[e] = double(lq ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]);
    }
    __syncthreads();
    for (int p0 = 0; p0 < k; p0 += nbmax) {
        const int nb = min(nbmax, k - p0), rows = mx - p0;
        for (int64_t e = tid; e < int64_t(rows) * nb; e += QR_WG) {
            const int c = int(e / rows), r = int(e % rows);
            V[e] = X[int64_t(p0 + c) * mx + p0 + r];
        }
        __syncthreads();
        double sigma;
        {
            double sq = 0;
            for (int r = 1 + tid; r < rows; r += QR_WG) sq += V[r] * V[r];
            sigma = block_sum<QR_WG>(sq, red);
        }
        for (int jj = 0; jj < nb; jj++) {
            double* cj = V + int64_t(jj) * rows;
            double tj, beta;
            householder(cj[jj], sigma, tj, beta);
            if (tj != 0.0) {
                const double scale = 1.0 / (cj[jj] - beta);
                for (int r = jj + 1 + tid; r < rows; r += QR_WG) cj[r] *= scale;
            }
            __syncthreads();
            if (tid == 0) {
                cj[jj] = beta;
                taus[jj] = tj;
                tau[p0 + jj] = tj;
            }
            qr_update_next(cj, V, rows, jj, nb, tj, g, gl, G, NG, bc);
            __syncthreads();
            sigma = bc[0];
        }
        // panel back to X (R on and above the diagonal, reflectors below),
        // then V as unit lower for the block reflector
        for (int64_t e = tid; e < int64_t(rows) * nb; e += QR_WG) {
            const int c = int(e / rows), r = int(e % rows);
            const double v = V[e];
            X[int64_t(p0 + c) * mx + p0 + r] = v;
            V[e] = r < c ? 0.0 : (r == c ? 1.0 : v);
        }
        __syncthreads();
        if (p0 + nb < nx) {
            qb_build_t(V, rows, nb, taus, Tm, Gm);
            qb_apply(V, rows, nb, Tm, W, X, mx, p0, p0 + nb, nx, true);
        }
        __syncthreads();
    }
    const gptr<T> f1 = to_global(static_cast<T*>(tk.f1));
    const gptr<T> f2 = to_global(static_cast<T*>(tk.f2));
    if (!lq) {
        for (int64_t e = tid; e < int64_t(k) * n; e += QR_WG) {
            const int r = int(e / n), c = int(e % n);
            f2[e] = T(r <= c ? X[int64_t(c) * mx + r] : 0.0);
        }
    } else {
        for (int64_t e = tid; e < int64_t(m) * k; e += QR_WG) {
            const int i = int(e / k), r = int(e % k);
            f1[e] = T(r <= i ? X[int64_t(i) * mx + r] : 0.0);
        }
    }
    __syncthreads();
    // Q = H_1 ... H_k I[:, :k], last panel first
    for (int p0 = ((k - 1) / nbmax) * nbmax; p0 >= 0; p0 -= nbmax) {
        const int nb = min(nbmax, k - p0), rows = mx - p0;
        for (int64_t e = tid; e < int64_t(rows) * nb; e += QR_WG) {
            const int c = int(e / rows), r = int(e % rows);
            V[e] = r < c ? 0.0 : (r == c ? 1.0 : X[int64_t(p0 + c) * mx + p0 + r]);
        }
        if (tid < nb) taus[tid] = tau[p0 + tid];
        __syncthreads();
        for (int64_t e = tid; e < int64_t(mx) * nb; e += QR_WG) {
            const int c = int(e / mx), r = int(e % mx);
            X[int64_t(p0 + c) * mx + r] = (r == p0 + c) ? 1.0 : 0.0;
        }
        __syncthreads();
        qb_build_t(V, rows, nb, taus, Tm, Gm);
        qb_apply(V, rows, nb, Tm, W, X, mx, p0, p0, k, false);
        __syncthreads();
    }
    if (!lq) {
        for (int64_t e = tid; e < int64_t(m) * k; e += QR_WG) {
            const int r = int(e / k), c = int(e % k);
            f1[e] = T(X[int64_t(c) * mx + r]);
        }
    } else {
        for (int64_t e = tid; e < int64_t(k) * n; e += QR_WG) {
            const int r = int(e / n), i = int(e % n);
            f2[e] = T(X[int64_t(r) * mx + i]);
        }
    }
}

// -------------------------------------------------------------------- svd
// Ranks the singular values (descending, ties by column) and writes S, U
// and V^T of the block from X (normalised columns), V and sig.
template<typename T>
__device__ __forceinline__ void svd_write(const SvdDev& tk, const double* X, const double* V, const double* sig,
                                          int g, int gl, int G, int NG) {
    const int m = tk.m, n = tk.n;
    const bool trans = m < n;
    const int mx = trans ? n : m, kx = trans ? m : n;
    T* U = static_cast<T*>(tk.u);
    T* Vo = static_cast<T*>(tk.v);
    for (int c = g; c < kx; c += NG) {
        const double sc = sig[c];
        int rank = 0;
        for (int d = 0; d < kx; d++) {
            const double sd = sig[d];
            rank += (sd > sc) || (sd == sc && d < c);
        }
        if (gl == 0) tk.s[rank] = sc;
        const double* xc = X + int64_t(c) * mx;
        const double* vc = V + int64_t(c) * kx;
        if (!trans) {
            // M = X diag(sig) V^T with X orthonormal: U (m x k), Vo = V^T
            for (int r = gl; r < m; r += G) U[int64_t(r) * kx + rank] = T(xc[r]);
            for (int i = gl; i < n; i += G) Vo[int64_t(rank) * n + i] = T(vc[i]);
        } else {
            // M^T = X diag(sig) V^T  =>  M = V diag(sig) X^T
            for (int i = gl; i < m; i += G) U[int64_t(i) * kx + rank] = T(vc[i]);
            for (int r = gl; r < n; r += G) Vo[int64_t(rank) * n + r] = T(xc[r]);
        }
    }
}

// One-sided Jacobi (Hestenes) on the kx = min(m, n) columns of X (X = M when
// m >= n, X = M^T otherwise), round-robin pair ordering; each pair is handled
// by one group of G lanes.  X, V (kx x kx) and sigma live in LDS when they
// fit, otherwise in global scratch.  Singular values sorted descending.
template<typename T>
__device__ __forceinline__ void svd_body(const SvdDev& tk, double* X, double* V, double* sig, int* flag, bool lds) {
    const int m = tk.m, n = tk.n;
    const bool trans = m < n;
    const int mx = trans ? n : m, kx = trans ? m : n;
    const T* in = static_cast<const T*>(tk.in);
    const int tid = threadIdx.x;
    // lanes per pair group: as for the rows, but in LDS narrower while the
    // groups would not cover a round's kx / 2 pairs at once (each lane then
    // keeps more rows of the pair, up to RPT in registers); the long columns
    // of the global-scratch kernel keep the row-sized groups
    int G0 = group_lanes(mx);
    while (lds && G0 > 16 && SVD_WG / G0 < (kx + 1) / 2) G0 /= 2;
    const int G = G0, NG = SVD_WG / G, g = tid / G, gl = tid % G;
    __shared__ double wmax[SVD_WG / WAVE];
    double amax = 0.0;
    for (int64_t e = tid; e < int64_t(mx) * kx; e += SVD_WG) {
        const int r = int(e % mx), c = int(e / mx);
        const double x = double(trans ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]);
        X[e] = x;
        amax = fmax(amax, fabs(x));
    }
    for (int64_t e = tid; e < int64_t(kx) * kx; e += SVD_WG) {
        const int r = int(e % kx), c = int(e / kx);
        V[e] = (r == c) ? 1.0 : 0.0;
    }
    amax = wave_max(amax);
    if (tid % WAVE == 0) wmax[tid / WAVE] = amax;
    __syncthreads();
    for (int w = 0; w < SVD_WG / WAVE; w++) amax = fmax(amax, wmax[w]);
    const int ex = jacobi_scale_exponent(amax);
    const double scl = ldexp(1.0, -ex);
    for (int64_t e = tid; e < int64_t(mx) * kx; e += SVD_WG) X[e] *= scl;
    __syncthreads();
    double fro = 0;
    for (int c = g; c < kx; c += NG) {
        double* xc = X + int64_t(c) * mx;
        double s = 0;
        for (int r = gl; r < mx; r += G) s += xc[r] * xc[r];
        s = group_sum(s, G);
        if (s < JACOBI_QUIET)
            for (int r = gl; r < mx; r += G) xc[r] = 0.0;
        else if (gl == 0)
            fro += s;
    }
    // squared norm at or below which a column takes part in no rotation
    const double live = fmax(JACOBI_QUIET, JACOBI_DEAD2 * block_sum<SVD_WG>(fro, wmax));
    __syncthreads();
    const int P = kx + (kx & 1);
    const double tol = fmax(1e-15, 2.220446049250313e-16 * mx);
    constexpr int RPT = 12;
    const bool cached = mx <= RPT * G;
    bool converged = kx <= 1;
    for (int sweep = 0; sweep < tk.max_sweeps && kx > 1; sweep++) {
        if (tid == 0) *flag = 0;
        // exact squared column norms in sig[], carried through the sweep's
        // rotations (one dot product and one reduction per pair instead of
        // three); a sweep without rotations tested every pair on these
        // exact values
        for (int c = g; c < kx; c += NG) {
            const double* xc = X + int64_t(c) * mx;
            double sq = 0;
            for (int r = gl; r < mx; r += G) sq += xc[r] * xc[r];
            sq = group_sum(sq, G);
            if (gl == 0) sig[c] = sq;
        }
        __syncthreads();
        for (int round = 0; round < P - 1; round++) {
            for (int pi = g; pi < P / 2; pi += NG) {
                int a = (pi == 0) ? 0 : 1 + (round + pi - 1) % (P - 1);
                int b = 1 + (round + P - 2 - pi) % (P - 1);
                if (a >= kx || b >= kx) continue;
                if (a > b) {
                    const int t = a;
                    a = b;
                    b = t;
                }
                double* xa = X + int64_t(a) * mx;
                double* xb = X + int64_t(b) * mx;
                // the pair's columns stay in registers between the dot
                // products and the rotation (one read, one write per element)
                double ra[RPT], rb[RPT];
                const double saa = sig[a], sbb = sig[b];
                double sab = 0;
                if (cached) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) {
                        const int r = gl + i * G;
                        const double u = r < mx ? xa[r] : 0.0, w = r < mx ? xb[r] : 0.0;
                        ra[i] = u;
                        rb[i] = w;
                        sab += u * w;
                    }
                } else {
                    for (int r = gl; r < mx; r += G) sab += xa[r] * xb[r];
                }
                sab = group_sum(sab, G);
                if (saa > live && sbb > live && sab != 0.0 && fabs(sab) > tol * sqrt(saa) * sqrt(sbb)) {
                    double cs, sn, t;
                    jacobi_rotation(saa, sbb, sab, cs, sn, t);
                    if (cached) {
#pragma unroll
                        for (int i = 0; i < RPT; i++) {
                            const int r = gl + i * G;
                            if (r < mx) {
                                xa[r] = cs * ra[i] - sn * rb[i];
                                xb[r] = sn * ra[i] + cs * rb[i];
                            }
                        }
                    } else {
                        for (int r = gl; r < mx; r += G) {
                            const double u = xa[r], w = xb[r];
                            xa[r] = cs * u - sn * w;
                            xb[r] = sn * u + cs * w;
                        }
                    }
                    double* va = V + int64_t(a) * kx;
                    double* vb = V + int64_t(b) * kx;
                    for (int r = gl; r < kx; r += G) {
                        const double u = va[r], w = vb[r];
                        va[r] = cs * u - sn * w;
                        vb[r] = sn * u + cs * w;
                    }
                    if (gl == 0) {
                        sig[a] = jacobi_tracked(saa, saa - t * sab);
                        sig[b] = jacobi_tracked(sbb, sbb + t * sab);
                        *flag = 1;
                    }
                }
            }
            __syncthreads();
        }
        const int any = *flag;
        __syncthreads();
        if (!any) {
            converged = true;
            break;
        }
    }
    if (!converged) {
        // not converged within the sweep limit: NaN singular values tell the
        // host (the truncation raises on them)
        for (int c = tid; c < kx; c += SVD_WG) tk.s[c] = __builtin_nan("");
        return;
    }
    for (int c = g; c < kx; c += NG) {
        double* xc = X + int64_t(c) * mx;
        double s = 0;
        for (int r = gl; r < mx; r += G) s += xc[r] * xc[r];
        s = group_sum(s, G);
        // a quiet or dead column (squared norm at most the live threshold)
        // keeps its tiny singular value, marked negative until its direction
        // is replaced below; an exactly zero column has singular value 0
        const bool quiet = s <= live;
        const double sc = sqrt(s);
        const double inv = (quiet || !(sc > 0)) ? 0.0 : 1.0 / sc;
        for (int r = gl; r < mx; r += G) xc[r] *= inv;
        if (gl == 0) sig[c] = quiet ? -ldexp(sc, ex) : ldexp(sc, ex);
    }
    __syncthreads();
    // rank-deficient blocks: complete the zero and quiet columns to an
    // orthonormal set (LAPACK's gesvd returns orthonormal U and V^T too)
    if (g == 0) {
        for (int c = 0; c < kx; c++) {
            if (sig[c] > 0.0) continue;
            double* xc = X + int64_t(c) * mx;
            for (int j = 0; j < mx; j++) {
                for (int r = gl; r < mx; r += G) xc[r] = (r == j) ? 1.0 : 0.0;
                for (int pass = 0; pass < 2; pass++) {
                    for (int d = 0; d < kx; d++) {
                        if (d == c || (sig[d] <= 0.0 && d > c)) continue;
                        const double* xd = X + int64_t(d) * mx;
                        double dot = 0;
                        for (int r = gl; r < mx; r += G) dot += xc[r] * xd[r];
                        dot = group_sum(dot, G);
                        for (int r = gl; r < mx; r += G) xc[r] -= dot * xd[r];
                    }
                }
                double nn = 0;
                for (int r = gl; r < mx; r += G) nn += xc[r] * xc[r];
                nn = group_sum(nn, G);
                if (nn > 0.25) {
                    const double inv = 1.0 / sqrt(nn);
                    for (int r = gl; r < mx; r += G) xc[r] *= inv;
                    break;
                }
            }
            // the completed column orthogonalises the later ones
            if (gl == 0) sig[c] = -sig[c];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    svd_write<T>(tk, X, V, sig, g, gl, G, NG);
}

template<typename T>
__global__ __launch_bounds__(SVD_WG) void svd_kernel_lds(const SvdDev* __restrict__ tasks) {
    extern __shared__ double smem[];
    __shared__ int flag;
    const SvdDev tk = tasks[blockIdx.x];
    const int mx = tk.m < tk.n ? tk.n : tk.m, kx = tk.m < tk.n ? tk.m : tk.n;
    svd_body<T>(tk, smem, smem + int64_t(mx) * kx, smem + int64_t(mx) * kx + int64_t(kx) * kx, &flag, true);
}

template<typename T>
__global__ __launch_bounds__(SVD_WG) void svd_kernel_glb(const SvdDev* __restrict__ tasks) {
    __shared__ int flag;
    const SvdDev tk = tasks[blockIdx.x];
    const int mx = tk.m < tk.n ? tk.n : tk.m, kx = tk.m < tk.n ? tk.m : tk.n;
    double* X = tk.work;
    svd_body<T>(tk, X, X + int64_t(mx) * kx, X + int64_t(mx) * kx + int64_t(kx) * kx, &flag, false);
}


// Register-resident one-sided Jacobi for the large blocks of the batched
// amplitude engine (rows mx <= 8 * RPT, so columns kx <= 8 * RPT), 32 * RPT
// threads.  Pair slot p (4 * RPT slots, 8 per wave) is held by 8 lanes of one
// wave (lane = 8 * row_group + slot % 8), each keeping RPT rows of the slot's
// top and bottom column in registers.  The 16 NWV columns form 2 NWV
// half-blocks of 8 (a wave's tops, a wave's bottoms).  A sweep is a
// round-robin over the half-blocks: in block round 0 every wave pairs the
// columns inside its two half-blocks (slot p with slot p ^ d, d = 1..7, both
// lanes of a pair computing the same rotation) and then its tops with its
// bottoms (8 cross rounds, the bottoms permuted between rounds by DPP xor
// 1, 3, 1, 7, 1, 3, 1 within each 8-lane group); in block rounds 1 .. 2 NWV - 2
// the half-blocks move one step along the round-robin ring through LDS and
// each wave again pairs its tops with its bottoms.  Every pair of columns
// meets once per sweep, and rounds inside a block round need neither a
// barrier nor LDS: the waves run them out of step, which hides the latency
// of each round's reduction and rotation chain behind the other waves.
// V is not accumulated: after convergence it is recovered as
// V = M^T U diag(1/sigma) (an LDS-tiled product), which is exact up to
// eps * sigma_max / sigma_j; columns with sigma_j <= 1e-13 * sigma_max get
// V = 0 (their products U S V^T are below that level).  The engine only
// consumes U S and V^T of kept values, so no orthonormal completion is done.
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_QUAD_XOR3 = 0x1B, DPP_ROW_HALF_MIRROR = 0x141;

template<int CTRL>
__device__ inline int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}

// value held by the lane whose slot (lane % 8) is this lane's slot ^ D, in
// the same 8-lane group: xor 1, 2, 3 by quad permutes, xor 7 by the half-row
// mirror, xor 4, 5, 6 by the mirror followed by a quad permute
template<int D>
__device__ inline double xor_slot(double v) {
    if constexpr (D == 1) return dpp_f64<DPP_QUAD_XOR1>(v);
    if constexpr (D == 2) return dpp_f64<DPP_QUAD_XOR2>(v);
    if constexpr (D == 3) return dpp_f64<DPP_QUAD_XOR3>(v);
    if constexpr (D == 7) return dpp_f64<DPP_ROW_HALF_MIRROR>(v);
    if constexpr (D == 4) return dpp_f64<DPP_QUAD_XOR3>(dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    if constexpr (D == 5) return dpp_f64<DPP_QUAD_XOR2>(dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    if constexpr (D == 6) return dpp_f64<DPP_QUAD_XOR1>(dpp_f64<DPP_ROW_HALF_MIRROR>(v));
    return v;
}

template<int D>
__device__ inline int xor_slot_int(int v) {
    if constexpr (D == 1) return dpp_i32<DPP_QUAD_XOR1>(v);
    if constexpr (D == 2) return dpp_i32<DPP_QUAD_XOR2>(v);
    if constexpr (D == 3) return dpp_i32<DPP_QUAD_XOR3>(v);
    if constexpr (D == 7) return dpp_i32<DPP_ROW_HALF_MIRROR>(v);
    if constexpr (D == 4) return dpp_i32<DPP_QUAD_XOR3>(dpp_i32<DPP_ROW_HALF_MIRROR>(v));
    if constexpr (D == 5) return dpp_i32<DPP_QUAD_XOR2>(dpp_i32<DPP_ROW_HALF_MIRROR>(v));
    if constexpr (D == 6) return dpp_i32<DPP_QUAD_XOR1>(dpp_i32<DPP_ROW_HALF_MIRROR>(v));
    return v;
}

template<int D, int N>
__device__ inline void move_xor(double (&x)[N], int& id) {
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = xor_slot<D>(x[i]);
    id = xor_slot_int<D>(id);
}

template<typename T, int RPT>
__global__ __launch_bounds__(32 * RPT) void svd_kernel_blk(const SvdDev* __restrict__ tasks) {
    constexpr int NT = 32 * RPT, NWV = NT / WAVE;
    // block exchange: (NWV + 1) half-blocks of 64 lanes x RPT doubles, i-major
    // (lanes contiguous); reused for the recovery tiles
    extern __shared__ double xbuf[];
    __shared__ int ibuf[NWV + 1][8];
    __shared__ double rmax[NWV], rfro[NWV];
    __shared__ double smax;
    const SvdDev tk = tasks[blockIdx.x];
    const int m = tk.m, n = tk.n;
    const bool trans = m < n;
    const int mx = trans ? n : m, kx = trans ? m : n;
    const T* in = static_cast<const T*>(tk.in);
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    if (mx > 8 * RPT || kx > 16 * NWV) {
        // the host routes only fitting blocks here; flag anything else
        for (int c = tid; c < kx; c += NT) tk.s[c] = __builtin_nan("");
        return;
    }
    const int gl = lane / 8, pl = lane % 8;
    // wave w starts with the half-blocks of columns 16 w + pl (top, xa) and
    // 16 w + 8 + pl (bottom, xb); lane 8 gl + pl holds rows gl + 8 i
    int ca = 16 * w + pl, cb = 16 * w + 8 + pl;
    double xa[RPT], xb[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        const int r = gl + 8 * i;
        xa[i] = (r < mx && ca < kx) ? double(trans ? in[int64_t(ca) * n + r] : in[int64_t(r) * n + ca]) : 0.0;
        xb[i] = (r < mx && cb < kx) ? double(trans ? in[int64_t(cb) * n + r] : in[int64_t(r) * n + cb]) : 0.0;
    }
    const double tol = fmax(1e-15, 2.220446049250313e-16 * mx);
    // exact power-of-two scaling, quiet and dead columns, as in svd_body
    int ex;
    double live;
    {
        double amax = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) amax = fmax(amax, fmax(fabs(xa[i]), fabs(xb[i])));
        amax = wave_max(amax);
        if (lane == 0) rmax[w] = amax;
        __syncthreads();
        for (int v = 0; v < NWV; v++) amax = fmax(amax, rmax[v]);
        ex = jacobi_scale_exponent(amax);
        const double scl = ldexp(1.0, -ex);
        double sa = 0, sb = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            xa[i] *= scl;
            xb[i] *= scl;
            sa += xa[i] * xa[i];
            sb += xb[i] * xb[i];
        }
        sa = sum8(sa);
        sb = sum8(sb);
        if (sa < JACOBI_QUIET)
#pragma unroll
            for (int i = 0; i < RPT; i++) xa[i] = 0.0;
        if (sb < JACOBI_QUIET)
#pragma unroll
            for (int i = 0; i < RPT; i++) xb[i] = 0.0;
        // squared Frobenius norm: one lane per slot adds its two columns
        double fr = gl == 0 ? (sa < JACOBI_QUIET ? 0.0 : sa) + (sb < JACOBI_QUIET ? 0.0 : sb) : 0.0;
        fr = wave_sum(fr);
        if (lane == 0) rfro[w] = fr;
        __syncthreads();
        fr = 0.0;
        for (int v = 0; v < NWV; v++) fr += rfro[v];
        live = fmax(JACOBI_QUIET, JACOBI_DEAD2 * fr);
    }
    bool converged = kx <= 1;
    int rot = 0;
    // rotation of the pair (x of column cx, y of column cy) as seen from the
    // lane holding x: both lanes of a duplicated pair order it by column
    // index, so they compute the same rotation and each keeps its own column
    // squared norms of the lane's two columns: exact after refresh(), then
    // carried through the rotations (and the bottoms' moves) of one block
    // round, so that a pair costs one dot product instead of three
    double na = 0.0, nb = 0.0;
    auto refresh = [&]() __attribute__((always_inline)) {
        double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
#pragma unroll
        for (int i = 0; i < RPT; i += 2) {
            a0 += xa[i] * xa[i];
            b0 += xb[i] * xb[i];
            a1 += xa[i + 1] * xa[i + 1];
            b1 += xb[i + 1] * xb[i + 1];
        }
        na = sum8(a0 + a1);
        nb = sum8(b0 + b1);
    };
    auto cross_round = [&]() __attribute__((always_inline)) {
        double sab = 0, sab1 = 0;
#pragma unroll
        for (int i = 0; i < RPT; i += 2) {
            sab += xa[i] * xb[i];
            sab1 += xa[i + 1] * xb[i + 1];
        }
        sab = sum8(sab + sab1);
        const double saa = na, sbb = nb;
        if (ca < kx && cb < kx && saa > live && sbb > live && sab != 0.0 && fabs(sab) > tol * sqrt(saa) * sqrt(sbb)) {
            double cs, sn, t;
            jacobi_rotation(saa, sbb, sab, cs, sn, t);
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const double u = xa[i], v = xb[i];
                xa[i] = cs * u - sn * v;
                xb[i] = sn * u + cs * v;
            }
            na = jacobi_tracked(saa, saa - t * sab);
            nb = jacobi_tracked(sbb, sbb + t * sab);
            rot = 1;
        }
    };
    auto move_b = [&](auto dtag) __attribute__((always_inline)) {
        constexpr int D = decltype(dtag)::value;
        move_xor<D>(xb, cb);
        nb = xor_slot<D>(nb);
    };
    // 8 cross rounds: top slot p meets bottom slot p ^ r in round r (the
    // bottoms move by a DPP lane permutation within each 8-lane group between
    // rounds: xor 1 and 3 by quad permutes, xor 7 by the half-row mirror), so
    // each top meets each bottom of the wave once
    auto cross8 = [&]() __attribute__((always_inline)) {
        using X1 = std::integral_constant<int, 1>;
        using X3 = std::integral_constant<int, 3>;
        using X7 = std::integral_constant<int, 7>;
        cross_round();
        move_b(X1());
        cross_round();
        move_b(X3());
        cross_round();
        move_b(X1());
        cross_round();
        move_b(X7());
        cross_round();
        move_b(X1());
        cross_round();
        move_b(X3());
        cross_round();
        move_b(X1());
        cross_round();
    };
    // intra round: slot p meets slot p ^ D among the tops and among the
    // bottoms; both slots of a pair hold it (each updates its own column)
    auto intra = [&](auto dtag) __attribute__((always_inline)) {
        constexpr int D = decltype(dtag)::value;
        // the partner's rows are fetched by DPP twice (for the sums, then for
        // the update) instead of being held: no extra registers
        const int pa = xor_slot_int<D>(ca), pb = xor_slot_int<D>(cb);
        // own and partner squared norms (the partner's tracked value moves
        // by the same permutation as its rows, so both lanes see one pair)
        const double s0 = na, s1 = xor_slot<D>(na), t0 = nb, t1 = xor_slot<D>(nb);
        double s2 = 0, t2 = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const double ya = xor_slot<D>(xa[i]), yb = xor_slot<D>(xb[i]);
            s2 += xa[i] * ya;
            t2 += xb[i] * yb;
        }
        s2 = sum8(s2);
        t2 = sum8(t2);
        // tops: (lower index, higher index) order fixes the rotation, the
        // lane keeps its own new column
        double ca_c = 1.0, ca_s = 0.0, cb_c = 1.0, cb_s = 0.0;
        bool ra = false, rb = false;
        {
            const bool low = ca < pa;
            const double saa = low ? s0 : s1, sbb = low ? s1 : s0;
            if (ca < kx && pa < kx && saa > live && sbb > live && s2 != 0.0 && fabs(s2) > tol * sqrt(saa) * sqrt(sbb)) {
                double cs, sn, t;
                jacobi_rotation(saa, sbb, s2, cs, sn, t);
                // own' = cs own - sn partner (low) or sn partner + cs own
                ca_c = cs;
                ca_s = low ? -sn : sn;
                na = low ? jacobi_tracked(saa, saa - t * s2) : jacobi_tracked(sbb, sbb + t * s2);
                ra = true;
            }
        }
        {
            const bool low = cb < pb;
            const double saa = low ? t0 : t1, sbb = low ? t1 : t0;
            if (cb < kx && pb < kx && saa > live && sbb > live && t2 != 0.0 && fabs(t2) > tol * sqrt(saa) * sqrt(sbb)) {
                double cs, sn, t;
                jacobi_rotation(saa, sbb, t2, cs, sn, t);
                cb_c = cs;
                cb_s = low ? -sn : sn;
                nb = low ? jacobi_tracked(saa, saa - t * t2) : jacobi_tracked(sbb, sbb + t * t2);
                rb = true;
            }
        }
        if (ra || rb) rot = 1;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const double ya = xor_slot<D>(xa[i]), yb = xor_slot<D>(xb[i]);
            xa[i] = ca_c * xa[i] + ca_s * ya;
            xb[i] = cb_c * xb[i] + cb_s * yb;
        }
    };
    for (int sweep = 0; sweep < tk.max_sweeps && kx > 1; sweep++) {
        rot = 0;
        // block round 0: the pairs inside each half-block, then tops with
        // bottoms
        refresh();
        intra(std::integral_constant<int, 1>());
        intra(std::integral_constant<int, 2>());
        intra(std::integral_constant<int, 3>());
        intra(std::integral_constant<int, 4>());
        intra(std::integral_constant<int, 5>());
        intra(std::integral_constant<int, 6>());
        intra(std::integral_constant<int, 7>());
        cross8();
        // block rounds 1 .. 2 NWV - 2: round-robin of the 2 NWV half-blocks
        // (wave 0's top fixed; tops move to the next wave, bottoms to the
        // previous one, wave 0's bottom becomes wave 1's top and the last
        // wave's top its own bottom), each followed by the 8 cross rounds
        for (int b = 1; b < 2 * NWV - 1; b++) {
            // the lane's LDS column base, made opaque each round: the row
            // offsets i * WAVE then fold into the LDS instructions' immediate
            // offsets instead of being hoisted out of the loops as RPT live
            // addresses (which spilled to scratch at RPT = 24)
            int lb = lane;
            asm volatile("" : "+v"(lb));
            double* const own = xbuf + w * RPT * WAVE + lb;
            double* const spare = xbuf + NWV * RPT * WAVE + lb;
            // bottoms out (wave 0's also into the spare slot NWV)
#pragma unroll
            for (int i = 0; i < RPT; i++) own[i * WAVE] = xb[i];
            if (gl == 0) ibuf[w][pl] = cb;
            if (w == 0) {
#pragma unroll
                for (int i = 0; i < RPT; i++) spare[i * WAVE] = xb[i];
                if (gl == 0) ibuf[NWV][pl] = cb;
            }
            __syncthreads();
            if (w + 1 < NWV) {
                const double* next = xbuf + (w + 1) * RPT * WAVE + lb;
#pragma unroll
                for (int i = 0; i < RPT; i++) xb[i] = next[i * WAVE];
                cb = ibuf[w + 1][pl];
            } else {
#pragma unroll
                for (int i = 0; i < RPT; i++) xb[i] = xa[i];
                cb = ca;
            }
            __syncthreads();
            // tops out
#pragma unroll
            for (int i = 0; i < RPT; i++) own[i * WAVE] = xa[i];
            if (gl == 0) ibuf[w][pl] = ca;
            __syncthreads();
            if (w >= 2) {
                const double* prev = xbuf + (w - 1) * RPT * WAVE + lb;
#pragma unroll
                for (int i = 0; i < RPT; i++) xa[i] = prev[i * WAVE];
                ca = ibuf[w - 1][pl];
            } else if (w == 1) {
#pragma unroll
                for (int i = 0; i < RPT; i++) xa[i] = spare[i * WAVE];
                ca = ibuf[NWV][pl];
            }
            __syncthreads();
            refresh();
            cross8();
        }
        if (!__syncthreads_or(rot)) {
            converged = true;
            break;
        }
    }
    if (!converged) {
        for (int c = tid; c < kx; c += NT) tk.s[c] = __builtin_nan("");
        return;
    }
    // sigma, normalised U columns into scratch
    double* X = tk.work;
    double* V = X + int64_t(mx) * kx;
    double* sig = V + int64_t(kx) * kx;
    if (tid == 0) smax = 0.0;
    __syncthreads();
    {
        double sa = 0, sb = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            sa += xa[i] * xa[i];
            sb += xb[i] * xb[i];
        }
        // quiet and dead columns keep their tiny singular value (no
        // orthonormal completion here: the engine reads U S and V^T of kept
        // values only)
        sa = sqrt(sum8(sa));
        sb = sqrt(sum8(sb));
        const double ia = sa > 0 ? 1.0 / sa : 0.0, ib = sb > 0 ? 1.0 / sb : 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int r = gl + 8 * i;
            if (r < mx) {
                if (ca < kx) X[int64_t(ca) * mx + r] = xa[i] * ia;
                if (cb < kx) X[int64_t(cb) * mx + r] = xb[i] * ib;
            }
        }
        if (gl == 0) {
            if (ca < kx) sig[ca] = ldexp(sa, ex);
            if (cb < kx) sig[cb] = ldexp(sb, ex);
        }
    }
    __syncthreads();
    if (tid == 0) {
        double mxs = 0;
        for (int c = 0; c < kx; c++) mxs = fmax(mxs, sig[c]);
        smax = mxs;
    }
    __syncthreads();
    // V(c, j) = sum_r M0(r, c) U(r, j) / sig_j, with M0 = M (mx = m) or M^T
    constexpr int TB = 64, KB = 32;
    double* As = xbuf;            // KB x TB: M0 rows r0.., columns c0..
    double* Bs = xbuf + KB * TB;  // KB x TB: U rows r0.., columns j0..
    const double vthr = 1e-13 * smax;
    for (int c0 = 0; c0 < kx; c0 += TB) {
        for (int j0 = 0; j0 < kx; j0 += TB) {
            // thread: column c0 + ci, rows j0 + jg + q * JS of the V tile
            constexpr int JS = NT / TB, JN = (TB + JS - 1) / JS;
            const int ci = tid % TB, jg = tid / TB;
            double acc[JN];
#pragma unroll
            for (int q = 0; q < JN; q++) acc[q] = 0.0;
            for (int r0 = 0; r0 < mx; r0 += KB) {
                for (int e = tid; e < KB * TB; e += NT) {
                    const int rr = e / TB, cc = e % TB, r = r0 + rr;
                    const int c = c0 + cc, j = j0 + cc;
                    As[e] = (r < mx && c < kx) ? double(trans ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]) : 0.0;
                    Bs[e] = (r < mx && j < kx) ? X[int64_t(j) * mx + r] : 0.0;
                }
                __syncthreads();
#pragma unroll 8
                for (int rr = 0; rr < KB; rr++) {
                    const double a = As[rr * TB + ci];
#pragma unroll
                    for (int q = 0; q < JN; q++)
                        if (jg + q * JS < TB) acc[q] += a * Bs[rr * TB + jg + q * JS];
                }
                __syncthreads();
            }
            const int c = c0 + ci;
            if (c < kx) {
#pragma unroll
                for (int q = 0; q < JN; q++) {
                    const int j = j0 + jg + q * JS;
                    if (jg + q * JS < TB && j < kx) V[int64_t(j) * kx + c] = sig[j] > vthr ? acc[q] / sig[j] : 0.0;
                }
            }
        }
    }
    __syncthreads();
    const int G = 8, NG = NT / G;
    svd_write<T>(tk, X, V, sig, tid / G, tid % G, G, NG);
}

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

void launch_permute(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg, int dtype,
                    hipStream_t stream) {
    if (num_wg <= 0) return;
    if (dtype == 0)
        hipLaunchKernelGGL((permute_kernel<double, double>), dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix,
                           wg_first);
    else
        hipLaunchKernelGGL((permute_kernel<float, float>), dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix,
                           wg_first);
    check(hipGetLastError(), "permute launch");
}

void launch_permute_from_f64(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg,
                             int dtype, hipStream_t stream) {
    if (num_wg <= 0) return;
    if (dtype == 0)
        hipLaunchKernelGGL((permute_kernel<double, double>), dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix,
                           wg_first);
    else
        hipLaunchKernelGGL((permute_kernel<double, float>), dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix,
                           wg_first);
    check(hipGetLastError(), "permute launch");
}

void launch_fill(const FillDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg, int dtype,
                 hipStream_t stream) {
    if (num_wg <= 0) return;
    if (dtype == 0)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix, wg_first);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix, wg_first);
    check(hipGetLastError(), "fill launch");
}

namespace {
__global__ void __launch_bounds__(256) gather_kernel(const GatherDev* __restrict__ tasks, double* __restrict__ dst) {
    const GatherDev t = tasks[blockIdx.x];
    for (int64_t i = threadIdx.x; i < t.n; i += blockDim.x) dst[t.off + i] = t.src[i];
}
}  // namespace

void launch_gather_f64(const GatherDev* tasks, int64_t num, double* dst, hipStream_t stream) {
    if (num <= 0) return;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned(num)), dim3(256), 0, stream, tasks, dst);
    check(hipGetLastError(), "gather launch");
}

void launch_permute_to_f64(const PermDev* tasks, const int64_t* prefix, const int32_t* wg_first, int64_t num_wg,
                           hipStream_t stream) {
    if (num_wg <= 0) return;
    hipLaunchKernelGGL((permute_kernel<float, double>), dim3(unsigned(num_wg)), dim3(WG), 0, stream, tasks, prefix,
                       wg_first);
    check(hipGetLastError(), "permute launch");
}

void launch_gemm(const GemmDev* tasks, const TileDev* tiles, int64_t num_tiles, int dtype, int variant,
                 hipStream_t stream) {
    if (num_tiles <= 0) return;
    const dim3 grid = dim3(static_cast<unsigned>(num_tiles));
    if (dtype >= 2 && variant != 0) throw std::runtime_error("gemm: split-bf16 uses the 32x32 tile");
    if (dtype == 0) {
        if (variant == 0)
            hipLaunchKernelGGL((gemm_kernel<double, 2, 2, 1, 1>), grid, dim3(256), 0, stream, tasks, tiles);
        else if (variant == 1)
            hipLaunchKernelGGL((gemm_kernel<double, 2, 2, 2, 2>), grid, dim3(256), 0, stream, tasks, tiles);
        else
            hipLaunchKernelGGL((gemm_kernel<double, 1, 1, 1, 1>), grid, dim3(64), 0, stream, tasks, tiles);
    } else if (dtype == 1) {
        if (variant == 0)
            hipLaunchKernelGGL((gemm_kernel<float, 2, 2, 1, 1>), grid, dim3(256), 0, stream, tasks, tiles);
        else if (variant == 1)
            hipLaunchKernelGGL((gemm_kernel<float, 2, 2, 2, 2>), grid, dim3(256), 0, stream, tasks, tiles);
        else
            hipLaunchKernelGGL((gemm_kernel<float, 1, 1, 1, 1>), grid, dim3(64), 0, stream, tasks, tiles);
    } else if (dtype == 2)
        hipLaunchKernelGGL(gemm_kernel_bf16<2>, grid, dim3(WG), 0, stream, tasks, tiles);
    else if (dtype == 3)
        hipLaunchKernelGGL(gemm_kernel_bf16<3>, grid, dim3(WG), 0, stream, tasks, tiles);
    else
        throw std::runtime_error("gemm: unsupported precision");
    check(hipGetLastError(), "gemm launch");
}

namespace {
// Raise a kernel's dynamic-LDS limit to the most any of its launches uses,
// once per kernel and process: the static's initialisation is thread-safe,
// so lanes launching from several host threads never call into the runtime
// for it concurrently or per launch.
template<auto Kernel>
void allow_lds(int64_t max_bytes) {
    static const bool done = [max_bytes] {
        if (max_bytes > 64 * 1024)
            check(hipFuncSetAttribute(reinterpret_cast<const void*>(Kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int(max_bytes)),
                  "dynamic LDS attribute");
        return true;
    }();
    (void)done;
}
}  // namespace

void launch_qr(const QrDev* tasks, int64_t num, int64_t lds_bytes, int variant, int dtype, hipStream_t stream) {
    if (num <= 0) return;
    if (variant == 1) {
        if (dtype == 0) {
            allow_lds<qr_kernel_blocked<double>>(QB_LDS);
            hipLaunchKernelGGL(qr_kernel_blocked<double>, dim3(unsigned(num)), dim3(QR_WG), size_t(QB_LDS), stream,
                               tasks);
        } else {
            allow_lds<qr_kernel_blocked<float>>(QB_LDS);
            hipLaunchKernelGGL(qr_kernel_blocked<float>, dim3(unsigned(num)), dim3(QR_WG), size_t(QB_LDS), stream,
                               tasks);
        }
    } else if (lds_bytes > 0) {
        if (dtype == 0) {
            allow_lds<qr_kernel_lds<double>>(QR_LDS_MAX);
            hipLaunchKernelGGL(qr_kernel_lds<double>, dim3(unsigned(num)), dim3(QR_WG), size_t(lds_bytes), stream, tasks);
        } else {
            allow_lds<qr_kernel_lds<float>>(QR_LDS_MAX);
            hipLaunchKernelGGL(qr_kernel_lds<float>, dim3(unsigned(num)), dim3(QR_WG), size_t(lds_bytes), stream, tasks);
        }
    } else {
        if (dtype == 0)
            hipLaunchKernelGGL(qr_kernel_glb<double>, dim3(unsigned(num)), dim3(QR_WG), 0, stream, tasks);
        else
            hipLaunchKernelGGL(qr_kernel_glb<float>, dim3(unsigned(num)), dim3(QR_WG), 0, stream, tasks);
    }
    check(hipGetLastError(), "qr launch");
}

namespace {
template<typename T, int RPT>
void launch_svd_blk(const SvdDev* tasks, int64_t num, hipStream_t stream) {
    // block exchange ((waves + 1) x 64 x RPT), at least the 32 KiB of the
    // recovery tiles
    const size_t lds = std::max<size_t>(size_t((RPT / 2 + 1) * 64 * RPT), 2 * 32 * 64) * sizeof(double);
    allow_lds<svd_kernel_blk<T, RPT>>(int64_t(lds));
    hipLaunchKernelGGL((svd_kernel_blk<T, RPT>), dim3(unsigned(num)), dim3(32 * RPT), lds, stream, tasks);
}

}  // namespace

void launch_svd(const SvdDev* tasks, int64_t num, int64_t lds_bytes, int rpt, int dtype, hipStream_t stream) {
    if (num <= 0) return;
    if (rpt > 0) {
        if (rpt != 8 && rpt != 16 && rpt != 24) throw std::runtime_error("svd: unsupported register tile");
        if (dtype == 0) {
            if (rpt == 8) launch_svd_blk<double, 8>(tasks, num, stream);
            if (rpt == 16) launch_svd_blk<double, 16>(tasks, num, stream);
            if (rpt == 24) launch_svd_blk<double, 24>(tasks, num, stream);
        } else {
            if (rpt == 8) launch_svd_blk<float, 8>(tasks, num, stream);
            if (rpt == 16) launch_svd_blk<float, 16>(tasks, num, stream);
            if (rpt == 24) launch_svd_blk<float, 24>(tasks, num, stream);
        }
    } else if (lds_bytes > 0) {
        if (dtype == 0) {
            allow_lds<svd_kernel_lds<double>>(SVD_LDS_MAX);
            hipLaunchKernelGGL(svd_kernel_lds<double>, dim3(unsigned(num)), dim3(SVD_WG), size_t(lds_bytes), stream,
                               tasks);
        } else {
            allow_lds<svd_kernel_lds<float>>(SVD_LDS_MAX);
            hipLaunchKernelGGL(svd_kernel_lds<float>, dim3(unsigned(num)), dim3(SVD_WG), size_t(lds_bytes), stream,
                               tasks);
        }
    } else {
        if (dtype == 0)
            hipLaunchKernelGGL(svd_kernel_glb<double>, dim3(unsigned(num)), dim3(SVD_WG), 0, stream, tasks);
        else
            hipLaunchKernelGGL(svd_kernel_glb<float>, dim3(unsigned(num)), dim3(SVD_WG), 0, stream, tasks);
    }
    check(hipGetLastError(), "svd launch");
}

}  // namespace tns
