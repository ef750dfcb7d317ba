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

namespace tns {

namespace {

constexpr int WG = 256;
constexpr int WAVE = 64;

template<typename T>
__device__ inline T neg_if(T v, int neg) {
    return neg ? -v : v;
}

// ---------------------------------------------------------------- permute
template<typename TS, typename T>
__global__ __launch_bounds__(WG) void permute_kernel(const PermDev* __restrict__ tasks, const int64_t* __restrict__ prefix,
                                                     const int32_t* __restrict__ wg_first) {
    const int64_t base = int64_t(blockIdx.x) * PERM_CHUNK;
    int t = wg_first[blockIdx.x];
    for (int j = 0; j < PERM_CHUNK / WG; j++) {
        const int64_t e = base + threadIdx.x + int64_t(j) * WG;
        // advance to the task containing e (tasks are contiguous in e)
        int64_t end = prefix[t + 1];
        while (end <= e) {
            t++;
            if (t >= wg_first[gridDim.x]) return;
            end = prefix[t + 1];
        }
        const PermDev& tk = tasks[t];
        uint32_t local = uint32_t(e - prefix[t]);
        int64_t so = 0, dof = 0;
        for (int r = tk.rank - 1; r >= 0; r--) {
            const uint32_t d = uint32_t(tk.dims[r]);
            const uint32_t i = local % d;
            local /= d;
            so += int64_t(i) * tk.ss[r];
            dof += int64_t(i) * tk.ds[r];
        }
        const T v = T(static_cast<const TS*>(tk.src)[so]);
        static_cast<T*>(tk.dst)[dof] = neg_if(v, tk.neg);
    }
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
        static_cast<T*>(tasks[t].p)[e - prefix[t]] = T(0);
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

// C tile 32x32 per workgroup; wave w owns the 16x16 sub-tile (w/2, w%2).
// MFMA 16x16x4 operand layout: lane l holds A(l%16, l/16) and B(l/16, l%16);
// accumulator rows as given by Mfma<T>::row.
template<typename T>
__global__ __launch_bounds__(WG) void gemm_kernel(const GemmDev* __restrict__ tasks, const TileDev* __restrict__ tiles) {
    __shared__ T As[GEMM_TILE][KB + 1];
    __shared__ T Bs[KB][GEMM_TILE + 1];
    const TileDev tile = tiles[blockIdx.x];
    const GemmDev g = tasks[tile.task];
    const T* A = static_cast<const T*>(g.a);
    const T* B = static_cast<const T*>(g.b);
    const int row0 = tile.tm * GEMM_TILE, col0 = tile.tn * GEMM_TILE;
    const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    const int wr = (wave / 2) * 16, wc = (wave % 2) * 16;
    auto acc = Mfma<T>::zero();
    for (int k0 = 0; k0 < g.k; k0 += KB) {
        for (int e = tid; e < GEMM_TILE * KB; e += WG) {
            // A tile: rows row0.., cols k0..; index so that consecutive threads
            // walk the contiguous direction of the stored operand
            int i, p;
            if (g.ta) {
                i = e % GEMM_TILE;
                p = e / GEMM_TILE;
            } else {
                p = e % KB;
                i = e / KB;
            }
            const int gi = row0 + i, gp = k0 + p;
            T v = T(0);
            if (gi < g.m && gp < g.k) v = g.ta ? A[int64_t(gp) * g.lda + gi] : A[int64_t(gi) * g.lda + gp];
            As[i][p] = v;
        }
        for (int e = tid; e < GEMM_TILE * KB; e += WG) {
            int p, j;
            if (g.tb) {
                p = e % KB;
                j = e / KB;
            } else {
                j = e % GEMM_TILE;
                p = e / GEMM_TILE;
            }
            const int gp = k0 + p, gj = col0 + j;
            T v = T(0);
            if (gp < g.k && gj < g.n) v = g.tb ? B[int64_t(gj) * g.ldb + gp] : B[int64_t(gp) * g.ldb + gj];
            Bs[p][j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            const T a = As[wr + lane % 16][kk + lane / 16];
            const T b = Bs[kk + lane / 16][wc + lane % 16];
            acc = Mfma<T>::step(a, b, acc);
        }
        __syncthreads();
    }
    T* C = static_cast<T*>(g.c);
    const T alpha = T(g.alpha);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int r = row0 + wr + Mfma<T>::row(lane, i);
        const int c = col0 + wc + lane % 16;
        if (r < g.m && c < g.n) C[int64_t(r) * g.ldc + c] = alpha * acc[i];
    }
}

// --------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// butterfly sum within aligned groups of G lanes (G = 16, 32 or 64): every
// lane of the group ends with the same, order-independent value
__device__ inline double group_sum(double v, int G) {
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

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

// --------------------------------------------------------------------- qr
// Householder QR of X (mx x nx, column-major), X = M (qr) or X = M^T (lq);
// R is written out first, then Q is formed in place from the reflectors
// (backward accumulation, as LAPACK's orgqr does).  X lives in LDS when it
// fits, otherwise in global scratch.
template<typename T>
__device__ __forceinline__ void qr_body(const QrDev& tk, double* X, double* tau, double* red, double* bc) {
    const int m = tk.m, n = tk.n, lq = tk.lq;
    const int mx = lq ? n : m, nx = lq ? m : n, k = mx < nx ? mx : nx;
    const T* in = static_cast<const T*>(tk.in);
    const int tid = threadIdx.x;
    const int G = group_lanes(mx), NG = QR_WG / G, g = tid / G, gl = tid % G;
    for (int64_t e = tid; e < int64_t(mx) * nx; e += QR_WG) {
        const int r = int(e % mx), c = int(e / mx);
        X[e] = double(lq ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]);
    }
    __syncthreads();
    for (int j = 0; j < k; j++) {
        double* cj = X + int64_t(j) * mx;
        double s = 0;
        for (int r = j + 1 + tid; r < mx; r += QR_WG) s += cj[r] * cj[r];
        const double sigma = block_sum<QR_WG>(s, red);
        if (tid == 0) {
            const double alpha = cj[j];
            if (sigma == 0.0) {
                bc[0] = 0.0;
                bc[1] = alpha;
            } else {
                const double beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
                bc[0] = (beta - alpha) / beta;
                bc[1] = beta;
            }
        }
        __syncthreads();
        const double tj = bc[0], beta = bc[1];
        if (tj != 0.0) {
            const double scale = 1.0 / (cj[j] - beta);
            for (int r = j + 1 + tid; r < mx; r += QR_WG) cj[r] *= scale;
        }
        __syncthreads();
        if (tid == 0) {
            cj[j] = beta;
            tau[j] = tj;
        }
        if (tj != 0.0) {
            for (int c = j + 1 + g; c < nx; c += NG) {
                double* cc = X + int64_t(c) * mx;
                double d = 0;
                for (int r = j + 1 + gl; r < mx; r += G) d += cj[r] * cc[r];
                d = group_sum(d, G) + cc[j];
                const double f = tj * d;
                if (gl == 0) cc[j] -= f;
                for (int r = j + 1 + gl; r < mx; r += G) cc[r] -= f * cj[r];
            }
        }
        __syncthreads();
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
    __shared__ double bc[2];
    const QrDev tk = tasks[blockIdx.x];
    const int mx = tk.lq ? tk.n : tk.m, nx = tk.lq ? tk.m : tk.n;
    qr_body<T>(tk, smem, smem + int64_t(mx) * nx, red, bc);
}

template<typename T>
__global__ __launch_bounds__(QR_WG) void qr_kernel_glb(const QrDev* __restrict__ tasks) {
    __shared__ double red[QR_WG / WAVE];
    __shared__ double bc[2];
    const QrDev tk = tasks[blockIdx.x];
    const int mx = tk.lq ? tk.n : tk.m, nx = tk.lq ? tk.m : tk.n;
    qr_body<T>(tk, tk.work, tk.work + int64_t(mx) * nx, red, bc);
}

// -------------------------------------------------------------------- svd
// One-sided Jacobi (Hestenes) on the kx = min(m, n) columns of X (X = M when
// m >= n, X = M^T otherwise), round-robin pair ordering; each pair is handled
// by one group of G lanes.  X, V (kx x kx) and sigma live in LDS when they
// fit, otherwise in global scratch.  Singular values sorted descending.
template<typename T>
__device__ __forceinline__ void svd_body(const SvdDev& tk, double* X, double* V, double* sig, int* flag) {
    const int m = tk.m, n = tk.n;
    const bool trans = m < n;
    const int mx = trans ? n : m, kx = trans ? m : n;
    const T* in = static_cast<const T*>(tk.in);
    const int tid = threadIdx.x;
    const int G = group_lanes(mx), NG = SVD_WG / G, g = tid / G, gl = tid % G;
    for (int64_t e = tid; e < int64_t(mx) * kx; e += SVD_WG) {
        const int r = int(e % mx), c = int(e / mx);
        X[e] = double(trans ? in[int64_t(c) * n + r] : in[int64_t(r) * n + c]);
    }
    for (int64_t e = tid; e < int64_t(kx) * kx; e += SVD_WG) {
        const int r = int(e % kx), c = int(e / kx);
        V[e] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int P = kx + (kx & 1);
    const double tol = fmax(1e-15, 2.220446049250313e-16 * mx);
    constexpr int RPT = 4;
    const bool cached = mx <= RPT * G;
    for (int sweep = 0; sweep < 60 && kx > 1; sweep++) {
        if (tid == 0) *flag = 0;
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
                double saa = 0, sbb = 0, sab = 0;
                if (cached) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) {
                        const int r = gl + i * G;
                        const double u = r < mx ? xa[r] : 0.0, w = r < mx ? xb[r] : 0.0;
                        ra[i] = u;
                        rb[i] = w;
                        saa += u * u;
                        sbb += w * w;
                        sab += u * w;
                    }
                } else {
                    for (int r = gl; r < mx; r += G) {
                        const double u = xa[r], w = xb[r];
                        saa += u * u;
                        sbb += w * w;
                        sab += u * w;
                    }
                }
                saa = group_sum(saa, G);
                sbb = group_sum(sbb, G);
                sab = group_sum(sab, G);
                if (fabs(sab) > tol * sqrt(saa * sbb) && sab != 0.0) {
                    const double zeta = (sbb - saa) / (2.0 * sab);
                    const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t);
                    const double sn = cs * t;
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
                    if (gl == 0) *flag = 1;
                }
            }
            __syncthreads();
        }
        const int any = *flag;
        __syncthreads();
        if (!any) break;
    }
    for (int c = g; c < kx; c += NG) {
        double* xc = X + int64_t(c) * mx;
        double s = 0;
        for (int r = gl; r < mx; r += G) s += xc[r] * xc[r];
        s = group_sum(s, G);
        const double sc = sqrt(s);
        const double inv = sc > 0 ? 1.0 / sc : 0.0;
        for (int r = gl; r < mx; r += G) xc[r] *= inv;
        if (gl == 0) sig[c] = sc;
    }
    __syncthreads();
    // exactly rank-deficient blocks: complete the zero columns to an
    // orthonormal set (LAPACK's gesvd returns orthonormal U and V^T too)
    if (g == 0) {
        for (int c = 0; c < kx; c++) {
            if (sig[c] > 0.0) continue;
            double* xc = X + int64_t(c) * mx;
            for (int j = 0; j < mx; j++) {
                for (int r = gl; r < mx; r += G) xc[r] = (r == j) ? 1.0 : 0.0;
                for (int pass = 0; pass < 2; pass++) {
                    for (int d = 0; d < kx; d++) {
                        if (d == c || (sig[d] == 0.0 && d > c)) continue;
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
        }
    }
    __syncthreads();
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

template<typename T>
__global__ __launch_bounds__(SVD_WG) void svd_kernel_lds(const SvdDev* __restrict__ tasks) {
    extern __shared__ double smem[];
    __shared__ int flag;
    const SvdDev tk = tasks[blockIdx.x];
    const int mx = tk.m < tk.n ? tk.n : tk.m, kx = tk.m < tk.n ? tk.m : tk.n;
    svd_body<T>(tk, smem, smem + int64_t(mx) * kx, smem + int64_t(mx) * kx + int64_t(kx) * kx, &flag);
}

template<typename T>
__global__ __launch_bounds__(SVD_WG) void svd_kernel_glb(const SvdDev* __restrict__ tasks) {
    __shared__ int flag;
    const SvdDev tk = tasks[blockIdx.x];
    const int mx = tk.m < tk.n ? tk.n : tk.m, kx = tk.m < tk.n ? tk.m : tk.n;
    double* X = tk.work;
    svd_body<T>(tk, X, X + int64_t(mx) * kx, X + int64_t(mx) * kx + int64_t(kx) * kx, &flag);
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

void launch_gemm(const GemmDev* tasks, const TileDev* tiles, int64_t num_tiles, int dtype, hipStream_t stream) {
    if (num_tiles <= 0) return;
    if (dtype == 0)
        hipLaunchKernelGGL(gemm_kernel<double>, dim3(unsigned(num_tiles)), dim3(WG), 0, stream, tasks, tiles);
    else
        hipLaunchKernelGGL(gemm_kernel<float>, dim3(unsigned(num_tiles)), dim3(WG), 0, stream, tasks, tiles);
    check(hipGetLastError(), "gemm launch");
}

namespace {
template<typename K>
void allow_lds(K kernel, int64_t bytes) {
    static int64_t granted[4] = {0, 0, 0, 0};
    (void)granted;
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(bytes));
}
}  // namespace

void launch_qr(const QrDev* tasks, int64_t num, int64_t lds_bytes, int dtype, hipStream_t stream) {
    if (num <= 0) return;
    if (lds_bytes > 0) {
        if (dtype == 0) {
            allow_lds(qr_kernel_lds<double>, lds_bytes);
            hipLaunchKernelGGL(qr_kernel_lds<double>, dim3(unsigned(num)), dim3(QR_WG), size_t(lds_bytes), stream, tasks);
        } else {
            allow_lds(qr_kernel_lds<float>, lds_bytes);
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

void launch_svd(const SvdDev* tasks, int64_t num, int64_t lds_bytes, int dtype, hipStream_t stream) {
    if (num <= 0) return;
    if (lds_bytes > 0) {
        if (dtype == 0) {
            allow_lds(svd_kernel_lds<double>, lds_bytes);
            hipLaunchKernelGGL(svd_kernel_lds<double>, dim3(unsigned(num)), dim3(SVD_WG), size_t(lds_bytes), stream,
                               tasks);
        } else {
            allow_lds(svd_kernel_lds<float>, lds_bytes);
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
