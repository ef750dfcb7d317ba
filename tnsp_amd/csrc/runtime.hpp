// Launch batching: task lists from any number of plans (and, in the batched
// amplitude engine, from any number of configurations) are bound to device
// pointers, packed into one host blob, uploaded with one copy and executed as
// a short sequence of kernel launches on one stream.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "core.hpp"
#include "kernels.hpp"

namespace tns {

void check_hip(hipError_t e, const char* what);

// Device/pinned staging for task arrays.  Not thread-safe: one user (one
// stream) at a time; the per-op API shares the process-wide instance, each
// engine lane owns its own.
// Task records of a batch travel host -> pinned buffer -> device buffer on
// the launch stream.  One instance per host thread (get()) or per engine
// lane.  The device buffer is reused by the next upload, possibly on another
// stream: release(st), recorded after the launches that read it, makes that
// upload wait for them on the device.
class Staging {
   public:
    static Staging& get();
    Staging() = default;
    // fixed mode: task records go to the given device buffer of cap bytes,
    // copied straight from the host blob; no memory or event is ever
    // allocated, and an upload larger than cap throws.  Uploads on one stream
    // are ordered; uploads on different streams into the same buffer race.
    Staging(char* fixed_dev, size_t cap) : dev_(fixed_dev), dev_cap_(cap), fixed_(true) {}
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging();
    // copies blob to device memory on stream st and returns the device base
    char* upload(const std::vector<char>& blob, hipStream_t st);
    // the kernels reading the last upload are queued on st
    void release(hipStream_t st);

   private:
    bool fixed_ = false;
    char* dev_ = nullptr;
    size_t dev_cap_ = 0;
    char* host_ = nullptr;
    size_t host_cap_ = 0;
    hipEvent_t ev_ = nullptr;    // copy out of host_ done
    bool pending_ = false;
    hipEvent_t done_ = nullptr;  // kernels reading dev_ done
    bool in_use_ = false;
    hipStream_t last_ = nullptr;
};

// Per-kind device timing of batched launches with HIP events recorded on the
// launch stream, plus the algorithmic work of each launch (flops for gemm /
// qr / svd, bytes moved for permute / fill).
struct KernelTimer {
    // one entry per kernel (variant) so that launch averages compare with a
    // kernel trace: QR / SVD are the LDS and unblocked-global kernels
    enum { PERM, FILL, GEMM, QR, SVD, DIAG, QR_BLOCKED, SVD_REG8, SVD_REG16, SVD_REG24, NKIND };
    struct Acc {
        double ms = 0, units = 0, launches = 0;
    } acc[NKIND];
    struct Pending {
        int kind;
        hipEvent_t a, b;
        double units;
    };
    std::vector<hipEvent_t> pool;
    size_t next = 0;
    std::vector<Pending> pending;
    hipEvent_t get();
    void add(int kind, hipEvent_t a, hipEvent_t b, double units) { pending.push_back({kind, a, b, units}); }
    // call after the stream has been synchronised
    void collect();
    void reset() {
        for (auto& a : acc) a = Acc();
    }
    ~KernelTimer();
};

// Jacobi sweep limit written into every SVD task (SVD_MAX_SWEEPS unless
// lowered, e.g. by a test of the non-convergence path)
int svd_max_sweeps();
void set_svd_max_sweeps(int n);

class Batch {
   public:
    explicit Batch(int dtype) : dtype_(dtype) {}
    void perm(const std::vector<PermTask>& ts, void* const* bufs, int64_t es);
    void perm_one(const PermTask& t, void* const* bufs, int64_t es);
    void flush_perm();
    void fill(const std::vector<FillTask>& ts, void* const* bufs, int64_t es);
    void fill_raw(void* p, int64_t n);
    void flush_fill();
    void gemm(const std::vector<GemmTask>& ts, void* const* bufs, int64_t es);
    void flush_gemm();
    // GEMM tile variant per sector size class (max(m, n) <= 16, <= 48,
    // larger); set by the engine's autotuner
    void set_gemm_table(const int* table) {
        for (int c = 0; c < 3; c++) gemm_table_[c] = table[c];
    }
    static int gemm_class(int64_t m, int64_t n) {
        const int64_t x = m > n ? m : n;
        return x <= 16 ? 0 : (x <= 48 ? 1 : 2);
    }
    // reduced-precision contraction pieces (float64 storage, float32 GEMM
    // operands): converting permutes, float32 fills and a GEMM launch in
    // precision `prec` (1 fp32, 2 split-bf16 x3, 3 split-bf16 x6)
    void perm_cvt(const std::vector<PermTask>& ts, void* const* bufs, int64_t es_src, int64_t es_dst);
    void flush_perm_cvt(int conv);  // 1: float64 -> float32, 2: float32 -> float64
    void flush_fill_as(int dt);
    void flush_gemm_as(int prec);
    void qr(const std::vector<QrTask>& ts, void* const* bufs, int64_t es, double* scratch);
    void flush_qr();
    void svd(const std::vector<SvdTask>& ts, void* const* bufs, int64_t es, double* scratch);
    void flush_svd();
    // copies of singular values (double, buffer 4) into the diagonal of S
    void diag(const std::vector<PermTask>& ts, void* const* bufs, int64_t es);
    void flush_diag();
    void svals_to_diag(const std::vector<PermTask>& ts, void* const* bufs, int64_t es) {
        diag(ts, bufs, es);
        flush_diag();
    }
    void run(hipStream_t st);
    bool empty() const { return cmds_.empty(); }
    int64_t perm_elements() const { return perm_elems_; }
    void reset();
    void set_timer(KernelTimer* t) { timer_ = t; }
    void set_staging(Staging* s) { staging_ = s; }
    // large SVD blocks may use the register-resident kernel, whose V is
    // recovered from U (no orthonormal completion for zero singular values)
    void set_fast_svd(bool on) { fast_svd_ = on; }

   private:
    enum Kind { PERM, FILL, GEMM, QR, SVD, DIAG };
    struct Cmd {
        Kind kind;
        size_t tasks, prefix, wg;  // blob offsets
        int64_t count, nwg;
        double units = 0;  // algorithmic flops or bytes of this launch
        int64_t lds = 0;   // dynamic LDS bytes (decompositions)
        int rpt = 0;       // register tile of the large-block SVD kernel
        int dt = -1;       // element type / GEMM precision override (-1: batch dtype)
        int conv = 0;      // permute conversion (1: f64 -> f32, 2: f32 -> f64)
    };
    template<typename D>
    void flush_decomp(Kind kind, std::vector<D>& tasks, std::vector<double>& cost, int64_t lds_max,
                      int64_t (*doubles)(int64_t, int64_t));
    KernelTimer* timer_ = nullptr;
    Staging* staging_ = nullptr;
    double pend_units_[6] = {0, 0, 0, 0, 0, 0};
    template<typename T>
    size_t put(const std::vector<T>& v);
    void flush_flat(Kind kind, std::vector<PermDev>& pd, std::vector<FillDev>& fd);

    int dtype_;
    std::vector<char> blob_;
    std::vector<Cmd> cmds_;
    std::vector<PermDev> pperm_, pdiag_;
    std::vector<FillDev> pfill_;
    std::vector<GemmDev> pgemm_;
    std::vector<TileDev> ptile_[GEMM_VARIANTS];
    double pgemm_units_[GEMM_VARIANTS] = {0, 0, 0};
    int gemm_table_[3] = {2, 0, 1};
    std::vector<QrDev> pqr_;
    std::vector<SvdDev> psvd_;
    std::vector<double> pqr_cost_, psvd_cost_;
    int64_t perm_elems_ = 0;
    double* scratch_cursor_ = nullptr;
    bool fast_svd_ = false;
};

}  // namespace tns
