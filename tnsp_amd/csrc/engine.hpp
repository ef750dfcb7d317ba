// Batched amplitude engine: evaluates psi(s) for many configurations of one
// PEPS by running tetragono's single-layer boundary-MPS op sequence
// (tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498 per
// row, :254-312 / 710-720 for the final row) for all configurations in
// lockstep.  Every op is planned per configuration on host threads (block
// structures differ per configuration) and executed as ONE batched launch
// per kernel kind over all configurations.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "core.hpp"
#include "runtime.hpp"

namespace tns {

class Pool {
   public:
    explicit Pool(int n);
    ~Pool();
    int size() const { return int(workers_.size()) + 1; }
    // fn(begin, end) over [0, n) in contiguous chunks, caller participates
    void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn);

   private:
    void work(int id);
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int64_t, int64_t)>* fn_ = nullptr;
    int64_t n_ = 0;
    int generation_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

struct DevArena {
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    size_t peak = 0;
    size_t alloc(size_t bytes) {
        const size_t off = (used + 255) / 256 * 256;
        used = off + bytes;
        if (used > peak) peak = used;
        return off;
    }
    void reset() { used = 0; }
    // grow to hold `used` bytes, keeping the first `keep` bytes
    void ensure(size_t keep, hipStream_t st);
    void release();
};

struct Slot {
    Shape sh;
    int arena = -1;  // index into Engine::arenas_
    int64_t off = 0; // byte offset
};

struct EngineStats {
    // plan_s: per-configuration planning on the thread pool; bind_s: the
    // lane thread's serial arena binding and task packing; exec_s: upload and
    // launch calls; sync_s: waiting for singular values before truncation
    double plan_s = 0, exec_s = 0, total_s = 0, bind_s = 0, sync_s = 0;
    double gemm_flops = 0, qr_flops = 0, svd_flops = 0, perm_elems = 0;
    double launches = 0, peak_bytes = 0, configs = 0;
};

class Engine {
   public:
    // name ids: P0 T U D L R L1 L2 L3 R1 R2 P Q, then P_{l1}_{l2}_0 row-major
    Engine(int kind, int L1, int L2, const std::vector<Shape>& sites, const std::vector<const void*>& site_data,
           int dtype, const std::vector<int32_t>& names, int threads);
    ~Engine();
    void evaluate(int64_t nconf, const int32_t* occupations, int64_t remain_cut, double relative_cut, double* psi,
                  hipStream_t st);
    void set_chunk(int64_t c) { chunk_ = c > 0 ? c : 1; }
    // contraction GEMM precision with float64 storage: 0 fp64, 1 fp32,
    // 2 split-bf16 x3, 3 split-bf16 x6 (QR, SVD and every stored tensor stay
    // float64)
    void set_precision(int p) { gemm_prec_ = dtype_ == 0 ? p : 0; }
    // GEMM tile autotuner: until finish_gemm_tuning, every contraction times
    // each tile variant on its own sector GEMMs, per size class; the sums
    // over all contractions decide the variant of each class
    void request_gemm_tuning() {
        tune_pending_ = true;
        for (auto& x : tune_ms_) x = 0;
    }
    void finish_gemm_tuning() {
        tune_pending_ = false;
        for (int c = 0; c < 3; c++) {
            int best = -1;
            for (int v = 0; v < 3; v++)
                if (tune_ms_[3 * c + v] > 0 && (best < 0 || tune_ms_[3 * c + v] < tune_ms_[3 * c + best])) best = v;
            if (best >= 0) gemm_table_[c] = best;
        }
    }
    const int* gemm_table() const { return gemm_table_; }
    void set_gemm_table(const int* t) {
        for (int c = 0; c < 3; c++) gemm_table_[c] = t[c];
    }
    // kernel ms of every (class, variant) pair measured by the last tuning
    const double* gemm_tuning_ms() const { return tune_ms_; }
    const EngineStats& stats() const { return stats_; }
    void reset_stats() {
        stats_ = EngineStats();
        timer_.reset();
    }
    void set_profile(bool on) { profile_ = on; }
    // the raw site buffers changed in place: project again on the next call
    void refresh_sites() { proj_.clear(); }
    // frees the working arenas (every one but the projected sites)
    void trim();
    const KernelTimer& timer() const { return timer_; }

   private:
    enum { A_PERSIST = 0, A_ROW0 = 1, A_ROW1 = 2, A_TEMP = 3, A_SVALS = 4, NARENA = 5 };
    void evaluate_chunk(int64_t nconf, const int32_t* occ, double* psi, hipStream_t st);
    void project_sites(hipStream_t st);
    void* ptr(const Slot& s) const { return arenas_[s.arena].base + s.off; }
    // batched ops over configurations; out slots receive results in arena `oa`
    void contract(std::vector<Slot*>& a, std::vector<Slot*>& b, const std::vector<std::pair<int32_t, int32_t>>& pairs,
                  std::vector<Slot*>& out, int oa, hipStream_t st);
    void qr(std::vector<Slot*>& t, const std::vector<int32_t>& free_names, int32_t cq, int32_t cr,
            std::vector<Slot*>& q, std::vector<Slot*>& r, int oa, hipStream_t st);
    void svd(std::vector<Slot*>& t, const std::vector<int32_t>& free_u, int32_t cu, int32_t cv, int32_t su,
             int32_t sv, std::vector<Slot*>& u, std::vector<Slot*>& s, std::vector<Slot*>& v, int oa,
             hipStream_t st);
    void rename(std::vector<Slot*>& t, const std::vector<std::pair<int32_t, int32_t>>& m);
    void grow_all(hipStream_t st);
    void run_batch(Batch& b, hipStream_t st);

    int kind_, L1_, L2_, dtype_;
    int64_t es_;
    std::vector<Shape> sites_;
    std::vector<const void*> site_data_;
    std::vector<int32_t> names_;
    int nphys_ = 0;
    std::vector<Slot> proj_;  // [(l1*L2+l2)*nphys + p]
    Slot one_;
    DevArena arenas_[NARENA];
    size_t keep_[NARENA] = {0, 0, 0, 0, 0};
    Pool pool_;
    int64_t chunk_ = 1024;
    int gemm_prec_ = 0;
    int gemm_table_[3] = {2, 0, 1};  // GEMM tile variant per sector size class
    bool tune_pending_ = false;
    double tune_ms_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int64_t remain_cut_ = -1;
    double relative_cut_ = 0.0;
    EngineStats stats_;
    Staging staging_;
    KernelTimer timer_;
    bool profile_ = false;
    std::vector<double> svals_host_;
};

}  // namespace tns
