// Launch batching and staging of task arrays (see runtime.hpp).
#include "runtime.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <stdexcept>

namespace tns {

namespace {
std::atomic<int> g_svd_max_sweeps{SVD_MAX_SWEEPS};
}

int svd_max_sweeps() { return g_svd_max_sweeps.load(std::memory_order_relaxed); }

void set_svd_max_sweeps(int n) {
    if (n < 1 || n > 1000) throw Error("svd sweep limit must be in [1, 1000]");
    g_svd_max_sweeps.store(n, std::memory_order_relaxed);
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

Staging& Staging::get() {
    // per host thread: callers on different threads (and streams) never share
    // the buffers
    thread_local Staging s;
    return s;
}

Staging::~Staging() {
    if (fixed_) return;
    if (pending_ && ev_) (void)hipEventSynchronize(ev_);
    if (in_use_ && done_) (void)hipEventSynchronize(done_);
    if (ev_) (void)hipEventDestroy(ev_);
    if (done_) (void)hipEventDestroy(done_);
    if (host_) (void)hipHostFree(host_);
    if (dev_) (void)hipFree(dev_);
}

char* Staging::upload(const std::vector<char>& blob, hipStream_t st) {
    const size_t n = blob.size();
    if (n == 0) return dev_;
    if (fixed_) {
        if (n > dev_cap_)
            throw Error("task records (" + std::to_string(n) + " bytes) exceed the fixed staging buffer (" +
                        std::to_string(dev_cap_) + " bytes)");
        // pageable source: the copy has consumed blob when the call returns
        check_hip(hipMemcpyAsync(dev_, blob.data(), n, hipMemcpyHostToDevice, st), "task upload");
        return dev_;
    }
    if (!ev_) check_hip(hipEventCreateWithFlags(&ev_, hipEventDisableTiming), "event create");
    if (!done_) check_hip(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "event create");
    // the pinned staging buffer may still feed the previous copy
    if (pending_) {
        check_hip(hipEventSynchronize(ev_), "staging wait");
        pending_ = false;
    }
    if (n > host_cap_) {
        if (host_) check_hip(hipHostFree(host_), "host free");
        host_cap_ = std::max(n, host_cap_ * 2);
        check_hip(hipHostMalloc(reinterpret_cast<void**>(&host_), host_cap_, hipHostMallocDefault), "host alloc");
    }
    if (n > dev_cap_) {
        // kernels queued earlier may still read the old task buffer
        if (in_use_) check_hip(hipEventSynchronize(done_), "staging synchronize");
        in_use_ = false;
        if (dev_) check_hip(hipFree(dev_), "device free");
        dev_cap_ = std::max(n, dev_cap_ * 2);
        check_hip(hipMalloc(reinterpret_cast<void**>(&dev_), dev_cap_), "device alloc");
    }
    // kernels of the previous upload, if queued on another stream, finish
    // before this copy overwrites their records (same stream: in order)
    if (in_use_ && last_ != st) check_hip(hipStreamWaitEvent(st, done_, 0), "staging stream wait");
    std::memcpy(host_, blob.data(), n);
    check_hip(hipMemcpyAsync(dev_, host_, n, hipMemcpyHostToDevice, st), "task upload");
    check_hip(hipEventRecord(ev_, st), "event record");
    pending_ = true;
    last_ = st;
    return dev_;
}

void Staging::release(hipStream_t st) {
    if (fixed_ || !done_) return;
    check_hip(hipEventRecord(done_, st), "staging release");
    in_use_ = true;
    last_ = st;
}

hipEvent_t KernelTimer::get() {
    if (next == pool.size()) {
        hipEvent_t e;
        check_hip(hipEventCreate(&e), "timer event");
        pool.push_back(e);
    }
    return pool[next++];
}

void KernelTimer::collect() {
    for (auto& p : pending) {
        float ms = 0;
        check_hip(hipEventElapsedTime(&ms, p.a, p.b), "timer elapsed");
        acc[p.kind].ms += ms;
        acc[p.kind].units += p.units;
        acc[p.kind].launches += 1;
    }
    pending.clear();
    next = 0;
}

KernelTimer::~KernelTimer() {
    for (auto e : pool) hipEventDestroy(e);
}

template<typename T>
size_t Batch::put(const std::vector<T>& v) {
    const size_t off = (blob_.size() + 255) / 256 * 256;
    blob_.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob_.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

namespace {
PermDev bind_perm(const PermTask& t, void* const* bufs, int64_t es_src, int64_t es_dst) {
    PermDev d{};
    d.src = static_cast<const char*>(bufs[t.src_buf]) + t.src * es_src;
    d.dst = static_cast<char*>(bufs[t.dst_buf]) + t.dst * es_dst;
    d.n = t.n;
    d.rank = t.rank;
    d.neg = t.neg;
    if (t.n >= (int64_t(1) << 31)) throw Error("permute task too large");
    for (int i = 0; i < MAXR; i++) {
        d.dims[i] = t.dims[i];
        d.ss[i] = t.ss[i];
        d.ds[i] = t.ds[i];
        perm_magic(uint32_t(t.dims[i] > 0 ? t.dims[i] : 1), &d.mul[i], &d.sh[i]);
    }
    return d;
}
}  // namespace

void Batch::perm(const std::vector<PermTask>& ts, void* const* bufs, int64_t es) {
    for (auto& t : ts) perm_one(t, bufs, es);
}

void Batch::perm_one(const PermTask& t, void* const* bufs, int64_t es) {
    if (t.n <= 0) return;
    pperm_.push_back(bind_perm(t, bufs, es, es));
    perm_elems_ += t.n;
    pend_units_[PERM] += double(2 * t.n * es);
}

void Batch::diag(const std::vector<PermTask>& ts, void* const* bufs, int64_t es) {
    for (auto& t : ts)
        if (t.n > 0) {
            pdiag_.push_back(bind_perm(t, bufs, 8, es));
            pend_units_[DIAG] += double(t.n * (8 + es));
        }
}

void Batch::flush_diag() {
    std::vector<FillDev> none;
    flush_flat(DIAG, pdiag_, none);
}

void Batch::fill(const std::vector<FillTask>& ts, void* const* bufs, int64_t es) {
    for (auto& t : ts)
        if (t.n > 0) {
            pfill_.push_back(FillDev{static_cast<char*>(bufs[t.buf]) + t.off * es, t.n});
            pend_units_[FILL] += double(t.n * es);
        }
}

void Batch::fill_raw(void* p, int64_t n) {
    if (n > 0) {
        pfill_.push_back(FillDev{p, n});
        pend_units_[FILL] += double(n * (dtype_ == 0 ? 8 : 4));
    }
}

void Batch::flush_flat(Kind kind, std::vector<PermDev>& pd, std::vector<FillDev>& fd) {
    const bool isfill = kind == FILL;
    const size_t T = isfill ? fd.size() : pd.size();
    if (T == 0) return;
    std::vector<int64_t> prefix(T + 1, 0);
    for (size_t i = 0; i < T; i++) prefix[i + 1] = prefix[i] + (isfill ? fd[i].n : pd[i].n);
    const int64_t total = prefix[T];
    const int64_t nwg = (total + PERM_CHUNK - 1) / PERM_CHUNK;
    if (nwg >= (int64_t(1) << 31)) throw Error("permute launch too large");
    std::vector<int32_t> wg(nwg + 1);
    size_t t = 0;
    for (int64_t w = 0; w < nwg; w++) {
        const int64_t e = w * PERM_CHUNK;
        while (t + 1 < T && prefix[t + 1] <= e) t++;
        wg[w] = int32_t(t);
    }
    wg[nwg] = int32_t(T);
    Cmd c;
    c.kind = kind;
    c.tasks = isfill ? put(fd) : put(pd);
    c.prefix = put(prefix);
    c.wg = put(wg);
    c.count = int64_t(T);
    c.nwg = nwg;
    c.units = pend_units_[kind];
    pend_units_[kind] = 0;
    cmds_.push_back(c);
    pd.clear();
    fd.clear();
}

void Batch::flush_perm() {
    std::vector<FillDev> none;
    flush_flat(PERM, pperm_, none);
}

void Batch::perm_cvt(const std::vector<PermTask>& ts, void* const* bufs, int64_t es_src, int64_t es_dst) {
    for (auto& t : ts)
        if (t.n > 0) {
            pperm_.push_back(bind_perm(t, bufs, es_src, es_dst));
            perm_elems_ += t.n;
            pend_units_[PERM] += double(t.n * (es_src + es_dst));
        }
}

void Batch::flush_perm_cvt(int conv) {
    const size_t before = cmds_.size();
    flush_perm();
    if (cmds_.size() > before) cmds_.back().conv = conv;
}

void Batch::flush_fill_as(int dt) {
    const size_t before = cmds_.size();
    flush_fill();
    if (cmds_.size() > before) cmds_.back().dt = dt;
}

void Batch::flush_gemm_as(int prec) {
    const size_t before = cmds_.size();
    flush_gemm();
    for (size_t i = before; i < cmds_.size(); i++) {
        if (prec >= 2 && cmds_[i].rpt != 0) throw Error("split-bf16 GEMMs need the 32x32 tile for every class");
        cmds_[i].dt = prec;
    }
}

void Batch::flush_fill() {
    std::vector<PermDev> none;
    flush_flat(FILL, none, pfill_);
}

void Batch::gemm(const std::vector<GemmTask>& ts, void* const* bufs, int64_t es) {
    for (auto& t : ts) {
        if (t.m <= 0 || t.n <= 0 || t.k <= 0) continue;
        GemmDev g{};
        g.a = static_cast<const char*>(bufs[t.a_buf]) + t.a * es;
        g.b = static_cast<const char*>(bufs[t.b_buf]) + t.b * es;
        g.c = static_cast<char*>(bufs[t.c_buf]) + t.c * es;
        g.m = t.m;
        g.n = t.n;
        g.k = t.k;
        g.lda = t.lda;
        g.ldb = t.ldb;
        g.ldc = t.ldc;
        g.ta = t.ta;
        g.tb = t.tb;
        g.alpha = t.alpha;
        const double fl = 2.0 * double(t.m) * double(t.n) * double(t.k);
        pend_units_[GEMM] += fl;
        const int32_t id = int32_t(pgemm_.size());
        pgemm_.push_back(g);
        const int v = gemm_table_[gemm_class(t.m, t.n)];
        const int T = GEMM_TILE_OF[v];
        pgemm_units_[v] += fl;
        const int tm = (t.m + T - 1) / T, tn = (t.n + T - 1) / T;
        for (int i = 0; i < tm; i++)
            for (int j = 0; j < tn; j++) ptile_[v].push_back(TileDev{id, i, j, 0});
    }
}

void Batch::flush_gemm() {
    bool any = false;
    for (auto& pt : ptile_) any = any || !pt.empty();
    if (!any) {
        pgemm_.clear();
        return;
    }
    const size_t tasks = put(pgemm_);
    for (int v = 0; v < GEMM_VARIANTS; v++) {
        if (ptile_[v].empty()) continue;
        Cmd c;
        c.kind = GEMM;
        c.tasks = tasks;
        c.prefix = put(ptile_[v]);
        c.wg = 0;
        c.count = int64_t(pgemm_.size());
        c.nwg = int64_t(ptile_[v].size());
        c.units = pgemm_units_[v];
        c.rpt = v;
        cmds_.push_back(c);
        ptile_[v].clear();
        pgemm_units_[v] = 0;
    }
    pend_units_[GEMM] = 0;
    pgemm_.clear();
}

void Batch::qr(const std::vector<QrTask>& ts, void* const* bufs, int64_t es, double* scratch) {
    double* cur = scratch;
    for (auto& t : ts) {
        QrDev q{};
        q.in = static_cast<const char*>(bufs[t.in_buf]) + t.in * es;
        q.f1 = static_cast<char*>(bufs[t.f1_buf]) + t.f1 * es;
        q.f2 = static_cast<char*>(bufs[t.f2_buf]) + t.f2 * es;
        q.work = cur;
        q.m = t.m;
        q.n = t.n;
        q.lq = t.lq;
        cur += qr_doubles(t.m, t.n);
        const double M = double(std::max(t.m, t.n)), N = double(std::min(t.m, t.n));
        const double fl = 4.0 * M * N * N - 4.0 * N * N * N / 3.0;
        pend_units_[QR] += fl;
        pqr_.push_back(q);
        pqr_cost_.push_back(fl);
    }
}

// Split a decomposition launch into an LDS-resident part (working set fits
// the per-workgroup LDS budget) and a global-scratch part; within each,
// largest blocks first so the long ones start early.
template<typename D>
void Batch::flush_decomp(Kind kind, std::vector<D>& tasks, std::vector<double>& cost, int64_t lds_max,
                         int64_t (*doubles)(int64_t, int64_t)) {
    if (tasks.empty()) return;
    std::vector<size_t> order(tasks.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    // part 0: global scratch, parts 1..3: LDS classes (largest first),
    // parts 4..6: register-resident SVD with 24 / 16 / 8 rows per lane,
    // part 7: blocked QR
    constexpr int NPART = 8;
    std::vector<D> parts[NPART];
    double units[NPART] = {};
    int64_t need[NPART] = {};
    static const int part_rpt[NPART] = {0, 0, 0, 0, 24, 16, 8, 1};
    for (size_t i : order) {
        const int64_t bytes = doubles(tasks[i].m, tasks[i].n) * 8;
        const int64_t rows = std::max<int64_t>(tasks[i].m, tasks[i].n);
        int p = 0;
        if (kind == SVD && fast_svd_ && bytes > LDS_CLASS[1] && rows <= SVD_REG_MAX_ROWS) {
            p = rows <= 64 ? 6 : (rows <= 128 ? 5 : 4);
        } else if (kind == QR && bytes > lds_max && rows <= QB_MAX_ROWS) {
            p = 7;
        } else if (bytes <= lds_max) {
            p = bytes <= LDS_CLASS[0] ? 3 : (bytes <= LDS_CLASS[1] ? 2 : 1);
        }
        parts[p].push_back(tasks[i]);
        units[p] += cost[i];
        need[p] = std::max(need[p], bytes);
    }
    for (int part = 0; part < NPART; part++) {
        auto& v = parts[part];
        if (v.empty()) continue;
        Cmd c;
        c.kind = kind;
        c.tasks = put(v);
        c.prefix = c.wg = 0;
        c.count = c.nwg = int64_t(v.size());
        c.units = units[part];
        c.lds = (part == 0 || part >= 4) ? 0 : (need[part] + 255) / 256 * 256;  // fixed inside the launcher
        c.rpt = part_rpt[part];
        cmds_.push_back(c);
    }
    pend_units_[kind] = 0;
    tasks.clear();
    cost.clear();
}

void Batch::flush_qr() { flush_decomp(QR, pqr_, pqr_cost_, QR_LDS_MAX, qr_doubles); }

void Batch::svd(const std::vector<SvdTask>& ts, void* const* bufs, int64_t es, double* scratch) {
    double* cur = scratch;
    for (auto& t : ts) {
        SvdDev s{};
        s.in = static_cast<const char*>(bufs[t.in_buf]) + t.in * es;
        s.u = static_cast<char*>(bufs[t.u_buf]) + t.u * es;
        s.s = reinterpret_cast<double*>(static_cast<char*>(bufs[t.s_buf]) + t.s * 8);
        s.v = static_cast<char*>(bufs[t.v_buf]) + t.v * es;
        s.work = cur;
        s.m = t.m;
        s.n = t.n;
        s.max_sweeps = svd_max_sweeps();
        const int64_t mx = std::max(t.m, t.n), kx = std::min(t.m, t.n);
        cur += svd_doubles(t.m, t.n);
        const double fl = 14.0 * double(mx) * double(kx) * double(kx) + 8.0 * double(kx) * double(kx) * double(kx);
        pend_units_[SVD] += fl;
        psvd_.push_back(s);
        psvd_cost_.push_back(fl);
    }
}

void Batch::flush_svd() { flush_decomp(SVD, psvd_, psvd_cost_, SVD_LDS_MAX, svd_doubles); }

void Batch::run(hipStream_t st) {
    flush_diag();
    flush_perm();
    flush_fill();
    flush_gemm();
    flush_qr();
    flush_svd();
    if (cmds_.empty()) return;
    Staging& stg = staging_ ? *staging_ : Staging::get();
    char* base = stg.upload(blob_, st);
    struct Release {
        Staging& s;
        hipStream_t st;
        ~Release() {
            try {
                s.release(st);
            } catch (...) {
            }
        }
    } release{stg, st};
    for (auto& c : cmds_) {
        hipEvent_t ea = nullptr, eb = nullptr;
        if (timer_) {
            ea = timer_->get();
            eb = timer_->get();
            check_hip(hipEventRecord(ea, st), "timer record");
        }
        switch (c.kind) {
            case PERM:
                if (c.conv == 1)
                    launch_permute_from_f64(reinterpret_cast<const PermDev*>(base + c.tasks),
                                            reinterpret_cast<const int64_t*>(base + c.prefix),
                                            reinterpret_cast<const int32_t*>(base + c.wg), c.nwg, 1, st);
                else if (c.conv == 2)
                    launch_permute_to_f64(reinterpret_cast<const PermDev*>(base + c.tasks),
                                          reinterpret_cast<const int64_t*>(base + c.prefix),
                                          reinterpret_cast<const int32_t*>(base + c.wg), c.nwg, st);
                else
                    launch_permute(reinterpret_cast<const PermDev*>(base + c.tasks),
                                   reinterpret_cast<const int64_t*>(base + c.prefix),
                                   reinterpret_cast<const int32_t*>(base + c.wg), c.nwg, dtype_, st);
                break;
            case DIAG:
                launch_permute_from_f64(reinterpret_cast<const PermDev*>(base + c.tasks),
                                        reinterpret_cast<const int64_t*>(base + c.prefix),
                                        reinterpret_cast<const int32_t*>(base + c.wg), c.nwg, dtype_, st);
                break;
            case FILL:
                launch_fill(reinterpret_cast<const FillDev*>(base + c.tasks),
                            reinterpret_cast<const int64_t*>(base + c.prefix),
                            reinterpret_cast<const int32_t*>(base + c.wg), c.nwg, c.dt >= 0 ? c.dt : dtype_, st);
                break;
            case GEMM:
                launch_gemm(reinterpret_cast<const GemmDev*>(base + c.tasks),
                            reinterpret_cast<const TileDev*>(base + c.prefix), c.nwg, c.dt >= 0 ? c.dt : dtype_,
                            c.rpt, st);
                break;
            case QR:
                launch_qr(reinterpret_cast<const QrDev*>(base + c.tasks), c.count, c.lds, c.rpt, dtype_, st);
                break;
            case SVD:
                launch_svd(reinterpret_cast<const SvdDev*>(base + c.tasks), c.count, c.lds, c.rpt, dtype_, st);
                break;
        }
        if (timer_) {
            check_hip(hipEventRecord(eb, st), "timer record");
            int k = int(c.kind);
            if (c.kind == QR && c.rpt == 1) k = KernelTimer::QR_BLOCKED;
            if (c.kind == SVD && c.rpt == 8) k = KernelTimer::SVD_REG8;
            if (c.kind == SVD && c.rpt == 16) k = KernelTimer::SVD_REG16;
            if (c.kind == SVD && c.rpt == 24) k = KernelTimer::SVD_REG24;
            timer_->add(k, ea, eb, c.units);
        }
    }
    reset();
}

void Batch::reset() {
    blob_.clear();
    cmds_.clear();
    pperm_.clear();
    pdiag_.clear();
    pfill_.clear();
    pgemm_.clear();
    for (auto& pt : ptile_) pt.clear();
    for (auto& u : pgemm_units_) u = 0;
    pqr_.clear();
    psvd_.clear();
    pqr_cost_.clear();
    psvd_cost_.clear();
    perm_elems_ = 0;
    for (auto& u : pend_units_) u = 0;
}

}  // namespace tns
