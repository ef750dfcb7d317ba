// Batched amplitude engine (see engine.hpp) and its C ABI.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>

#include "../../include/tnsgpu.h"
#include "handles.hpp"

namespace tns {

namespace {
double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// ------------------------------------------------------------------ pool
Pool::Pool(int n) {
    for (int i = 1; i < n; i++) workers_.emplace_back([this, i] { work(i); });
}

Pool::~Pool() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
}

void Pool::work(int id) {
    int seen = 0;
    while (true) {
        const std::function<void(int64_t, int64_t)>* fn;
        int64_t n;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
            if (stop_) return;
            seen = generation_;
            fn = fn_;
            n = n_;
        }
        const int T = size();
        const int64_t b = n * id / T, e = n * (id + 1) / T;
        if (b < e) (*fn)(b, e);
        {
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) done_cv_.notify_all();
        }
    }
}

void Pool::parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& fn) {
    if (n <= 0) return;
    // waking the workers costs more than planning a few items
    if (workers_.empty() || n < 16) {
        fn(0, n);
        return;
    }
    std::exception_ptr err;
    std::mutex err_mu;
    std::function<void(int64_t, int64_t)> guarded = [&](int64_t b, int64_t e) {
        try {
            fn(b, e);
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
        }
    };
    {
        std::lock_guard<std::mutex> g(mu_);
        fn_ = &guarded;
        n_ = n;
        pending_ = int(workers_.size());
        generation_++;
    }
    cv_.notify_all();
    const int T = size();
    guarded(0, n / T);
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return pending_ == 0; });
    }
    if (err) std::rethrow_exception(err);
}

// ------------------------------------------------------------------ arena
void DevArena::ensure(size_t keep, hipStream_t st) {
    if (used <= cap) return;
    // 1.25x what is used, at least 1.5x the old capacity: the arenas reach
    // their working size in a few steps without doubling past it
    size_t ncap = std::max(used + (used >> 2), cap + (cap >> 1));
    ncap = (ncap + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
    char* nb = nullptr;
    check_hip(hipMalloc(reinterpret_cast<void**>(&nb), ncap), "engine arena alloc");
    if (base) {
        if (keep) check_hip(hipMemcpyAsync(nb, base, keep, hipMemcpyDeviceToDevice, st), "arena grow copy");
        check_hip(hipStreamSynchronize(st), "arena grow sync");
        check_hip(hipFree(base), "arena free");
    }
    base = nb;
    cap = ncap;
}

void DevArena::release() {
    if (base) (void)hipFree(base);
    base = nullptr;
    cap = used = 0;
}

// ------------------------------------------------------------------ engine
enum NameIdx { N_P0, N_T, N_U, N_D, N_L, N_R, N_L1, N_L2, N_L3, N_R1, N_R2, N_P, N_Q, N_FIRST_SITE };

Engine::Engine(int kind, int L1, int L2, const std::vector<Shape>& sites, const std::vector<const void*>& site_data,
               int dtype, const std::vector<int32_t>& names, int threads)
    : kind_(kind), L1_(L1), L2_(L2), dtype_(dtype), es_(dtype == 0 ? 8 : 4), sites_(sites), site_data_(site_data),
      names_(names), pool_(std::max(1, threads)) {
    if (int(sites.size()) != L1 * L2 || int(site_data.size()) != L1 * L2) throw Error("engine: wrong number of sites");
    if (int(names.size()) != N_FIRST_SITE + L1 * L2) throw Error("engine: wrong number of name ids");
    const int p0 = sites_[0].index_of(names_[N_P0]);
    if (p0 < 0) throw Error("engine: site tensors need a P0 edge");
    nphys_ = int(sites_[0].edges[p0].total());
}

Engine::~Engine() {
    for (auto& a : arenas_) a.release();
}

void Engine::trim() {
    // the working arenas regrow on the next evaluation; the projected sites stay
    for (int i = 0; i < NARENA; i++)
        if (i != A_PERSIST) arenas_[i].release();
}

void Engine::grow_all(hipStream_t st) {
    for (int i = 0; i < NARENA; i++) arenas_[i].ensure(keep_[i], st);
}

void Engine::run_batch(Batch& b, hipStream_t st) {
    const double t0 = now();
    stats_.perm_elems += double(b.perm_elements());
    if (profile_) b.set_timer(&timer_);
    b.set_staging(&staging_);
    b.run(st);
    stats_.exec_s += now() - t0;
}

void Engine::rename(std::vector<Slot*>& t, const std::vector<std::pair<int32_t, int32_t>>& m) {
    for (auto* s : t) {
        for (auto& [o, n] : m) {
            const int i = s->sh.index_of(o);
            if (i >= 0) s->sh.names[i] = n;
        }
    }
}

static std::vector<std::pair<int32_t, int32_t>> present_pairs(const Shape& a, const Shape& b,
                                                              const std::vector<std::pair<int32_t, int32_t>>& pairs) {
    std::vector<std::pair<int32_t, int32_t>> out;
    for (auto& pr : pairs)
        if (a.index_of(pr.first) >= 0 && b.index_of(pr.second) >= 0) out.push_back(pr);
    return out;
}

void Engine::contract(std::vector<Slot*>& a, std::vector<Slot*>& b,
                      const std::vector<std::pair<int32_t, int32_t>>& pairs, std::vector<Slot*>& out, int oa,
                      hipStream_t st) {
    const int64_t B = int64_t(a.size());
    const auto pr = present_pairs(a[0]->sh, b[0]->sh, pairs);
    std::vector<ContractPlan> plans(B);
    double t0 = now();
    pool_.parallel_for(B, [&](int64_t s, int64_t e) {
        for (int64_t i = s; i < e; i++) plan_contract(a[i]->sh, b[i]->sh, pr, plans[i]);
    });
    stats_.plan_s += now() - t0;
    // allocation
    arenas_[A_TEMP].reset();
    keep_[A_TEMP] = 0;
    keep_[oa] = arenas_[oa].used;
    // reduced precision: the merged operands and the product are float32
    // (converted by the permutes), everything stored stays float64
    const bool lowp = gemm_prec_ != 0;
    const int64_t eg = lowp ? 4 : es_;
    std::vector<int64_t> ta(B), tb(B), tp(B), to(B);
    for (int64_t i = 0; i < B; i++) {
        ta[i] = int64_t(arenas_[A_TEMP].alloc(size_t(plans[i].ma.size * eg)));
        tb[i] = int64_t(arenas_[A_TEMP].alloc(size_t(plans[i].mb.size * eg)));
        tp[i] = int64_t(arenas_[A_TEMP].alloc(size_t(plans[i].prod.size * eg)));
        to[i] = int64_t(arenas_[oa].alloc(size_t(plans[i].result.size * es_)));
        stats_.gemm_flops += double(plans[i].flops);
    }
    grow_all(st);
    t0 = now();
    if (tune_pending_ && !lowp) {
        // operands first, then every variant on each class's GEMMs (3 timed
        // repetitions after one warm-up); the GEMMs only write the product
        Batch prep(dtype_);
        char* Tp = arenas_[A_TEMP].base;
        for (int64_t i = 0; i < B; i++) {
            void* bufs[6] = {ptr(*a[i]), ptr(*b[i]), Tp + ta[i], Tp + tb[i], Tp + tp[i], nullptr};
            prep.perm(plans[i].perm_a, bufs, es_);
            prep.perm(plans[i].perm_b, bufs, es_);
        }
        prep.flush_perm();
        for (int64_t i = 0; i < B; i++) {
            void* bufs[6] = {nullptr, nullptr, Tp + ta[i], Tp + tb[i], Tp + tp[i], nullptr};
            prep.fill(plans[i].fills, bufs, es_);
        }
        prep.flush_fill();
        run_batch(prep, st);
        for (int c = 0; c < 3; c++) {
                for (int v = 0; v < GEMM_VARIANTS; v++) {
                const int table[3] = {v, v, v};
                KernelTimer tm;
                double ms = 0;
                bool any = false;
                for (int rep = 0; rep < 4; rep++) {
                    Batch g(dtype_);
                    g.set_gemm_table(table);
                    g.set_staging(&staging_);
                    if (rep > 0) g.set_timer(&tm);
                    for (int64_t i = 0; i < B; i++) {
                        void* bufs[6] = {nullptr, nullptr, Tp + ta[i], Tp + tb[i], Tp + tp[i], nullptr};
                        std::vector<GemmTask> sel;
                        for (auto& t : plans[i].gemms)
                            if (Batch::gemm_class(t.m, t.n) == c) sel.push_back(t);
                        g.gemm(sel, bufs, es_);
                    }
                    g.flush_gemm();
                    any = !g.empty();
                    if (!any) break;
                    g.run(st);
                }
                if (!any) break;
                check_hip(hipStreamSynchronize(st), "tuning sync");
                tm.collect();
                ms = tm.acc[KernelTimer::GEMM].ms / 3.0;
                tune_ms_[3 * c + v] += ms;
            }
        }
    }
    Batch bt(dtype_);
    static const int bf16_table[3] = {0, 0, 0};
    bt.set_gemm_table(gemm_prec_ >= 2 ? bf16_table : gemm_table_);
    char* T = arenas_[A_TEMP].base;
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {ptr(*a[i]), ptr(*b[i]), T + ta[i], T + tb[i], T + tp[i], arenas_[oa].base + to[i]};
        if (lowp) {
            bt.perm_cvt(plans[i].perm_a, bufs, es_, eg);
            bt.perm_cvt(plans[i].perm_b, bufs, es_, eg);
        } else {
            bt.perm(plans[i].perm_a, bufs, es_);
            bt.perm(plans[i].perm_b, bufs, es_);
        }
    }
    if (lowp)
        bt.flush_perm_cvt(1);
    else
        bt.flush_perm();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {nullptr, nullptr, T + ta[i], T + tb[i], T + tp[i], nullptr};
        bt.fill(plans[i].fills, bufs, eg);
    }
    if (lowp)
        bt.flush_fill_as(1);
    else
        bt.flush_fill();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {nullptr, nullptr, T + ta[i], T + tb[i], T + tp[i], nullptr};
        bt.gemm(plans[i].gemms, bufs, eg);
    }
    if (lowp)
        bt.flush_gemm_as(gemm_prec_);
    else
        bt.flush_gemm();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {nullptr, nullptr, T + ta[i], T + tb[i], T + tp[i], arenas_[oa].base + to[i]};
        if (lowp)
            bt.perm_cvt(plans[i].perm_out, bufs, eg, es_);
        else
            bt.perm(plans[i].perm_out, bufs, es_);
    }
    if (lowp)
        bt.flush_perm_cvt(2);
    else
        bt.flush_perm();
    stats_.bind_s += now() - t0;
    stats_.launches += 4;
    run_batch(bt, st);
    for (int64_t i = 0; i < B; i++) {
        out[i]->sh = std::move(plans[i].result);
        out[i]->arena = oa;
        out[i]->off = to[i];
    }
}

void Engine::qr(std::vector<Slot*>& t, const std::vector<int32_t>& free_names, int32_t cq, int32_t cr,
                std::vector<Slot*>& q, std::vector<Slot*>& r, int oa, hipStream_t st) {
    const int64_t B = int64_t(t.size());
    std::vector<int32_t> fr;
    for (auto n : free_names)
        if (t[0]->sh.index_of(n) >= 0) fr.push_back(n);
    std::vector<QrPlan> plans(B);
    double t0 = now();
    pool_.parallel_for(B, [&](int64_t s, int64_t e) {
        for (int64_t i = s; i < e; i++) plan_qr(t[i]->sh, 'r', fr, cq, cr, plans[i]);
    });
    stats_.plan_s += now() - t0;
    arenas_[A_TEMP].reset();
    keep_[A_TEMP] = 0;
    keep_[oa] = arenas_[oa].used;
    std::vector<int64_t> tm(B), t1(B), t2(B), ts(B), oq(B), orr(B);
    for (int64_t i = 0; i < B; i++) {
        const QrPlan& p = plans[i];
        tm[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.merged.size * es_)));
        t1[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.t1.size * es_)));
        t2[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.t2.size * es_)));
        int64_t sc = 0;
        for (auto& task : p.qrs) sc += qr_doubles(task.m, task.n);
        ts[i] = int64_t(arenas_[A_TEMP].alloc(size_t(sc * 8)));
        oq[i] = int64_t(arenas_[oa].alloc(size_t(p.q.size * es_)));
        orr[i] = int64_t(arenas_[oa].alloc(size_t(p.r.size * es_)));
        stats_.qr_flops += double(p.flops);
    }
    grow_all(st);
    t0 = now();
    Batch bt(dtype_);
    char* T = arenas_[A_TEMP].base;
    char* O = arenas_[oa].base;
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {ptr(*t[i]), T + tm[i], T + t1[i], T + t2[i], O + oq[i], O + orr[i]};
        bt.perm(plans[i].perm_in, bufs, es_);
    }
    bt.flush_perm();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {ptr(*t[i]), T + tm[i], T + t1[i], T + t2[i], O + oq[i], O + orr[i]};
        bt.qr(plans[i].qrs, bufs, es_, reinterpret_cast<double*>(T + ts[i]));
    }
    bt.flush_qr();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[6] = {ptr(*t[i]), T + tm[i], T + t1[i], T + t2[i], O + oq[i], O + orr[i]};
        bt.perm(plans[i].perm_1, bufs, es_);
        bt.perm(plans[i].perm_2, bufs, es_);
    }
    bt.flush_perm();
    stats_.bind_s += now() - t0;
    stats_.launches += 3;
    run_batch(bt, st);
    for (int64_t i = 0; i < B; i++) {
        q[i]->sh = std::move(plans[i].q);
        q[i]->arena = oa;
        q[i]->off = oq[i];
        r[i]->sh = std::move(plans[i].r);
        r[i]->arena = oa;
        r[i]->off = orr[i];
    }
}

void Engine::svd(std::vector<Slot*>& t, const std::vector<int32_t>& free_u, int32_t cu, int32_t cv, int32_t su,
                 int32_t sv, std::vector<Slot*>& u, std::vector<Slot*>& s, std::vector<Slot*>& v, int oa,
                 hipStream_t st) {
    const int64_t B = int64_t(t.size());
    std::vector<SvdPlan> plans(B);
    double t0 = now();
    pool_.parallel_for(B, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) plan_svd_factor(t[i]->sh, free_u, cu, cv, su, sv, plans[i]);
    });
    stats_.plan_s += now() - t0;
    arenas_[A_TEMP].reset();
    arenas_[A_SVALS].reset();
    keep_[A_TEMP] = keep_[A_SVALS] = 0;
    keep_[oa] = arenas_[oa].used;
    std::vector<int64_t> tm(B), t1(B), t2(B), ts(B), sv_off(B);
    for (int64_t i = 0; i < B; i++) {
        const SvdPlan& p = plans[i];
        tm[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.merged.size * es_)));
        t1[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.t1.size * es_)));
        t2[i] = int64_t(arenas_[A_TEMP].alloc(size_t(p.t2.size * es_)));
        int64_t sc = 0;
        for (auto& task : p.svds) sc += svd_doubles(task.m, task.n);
        ts[i] = int64_t(arenas_[A_TEMP].alloc(size_t(sc * 8)));
        sv_off[i] = int64_t(arenas_[A_SVALS].used);
        arenas_[A_SVALS].used += size_t(p.s_total * 8);  // contiguous, no alignment gaps
        if (arenas_[A_SVALS].used > arenas_[A_SVALS].peak) arenas_[A_SVALS].peak = arenas_[A_SVALS].used;
        stats_.svd_flops += double(p.flops);
    }
    const size_t sv_bytes = arenas_[A_SVALS].used;
    grow_all(st);
    t0 = now();
    Batch bt(dtype_);
    bt.set_fast_svd(true);
    char* T = arenas_[A_TEMP].base;
    char* S = arenas_[A_SVALS].base;
    for (int64_t i = 0; i < B; i++) {
        void* bufs[5] = {ptr(*t[i]), T + tm[i], T + t1[i], T + t2[i], S + sv_off[i]};
        bt.perm(plans[i].perm_in, bufs, es_);
    }
    bt.flush_perm();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[5] = {ptr(*t[i]), T + tm[i], T + t1[i], T + t2[i], S + sv_off[i]};
        bt.svd(plans[i].svds, bufs, es_, reinterpret_cast<double*>(T + ts[i]));
    }
    bt.flush_svd();
    stats_.bind_s += now() - t0;
    stats_.launches += 2;
    run_batch(bt, st);
    svals_host_.resize(sv_bytes / 8 + 1);
    const double t1s = now();
    if (sv_bytes) check_hip(hipMemcpyAsync(svals_host_.data(), S, sv_bytes, hipMemcpyDeviceToHost, st), "svals copy");
    check_hip(hipStreamSynchronize(st), "svals sync");
    stats_.sync_s += now() - t1s;
    if (profile_) timer_.collect();
    t0 = now();
    pool_.parallel_for(B, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++)
            plan_svd_truncate(plans[i], svals_host_.data() + sv_off[i] / 8, remain_cut_, relative_cut_);
    });
    stats_.plan_s += now() - t0;
    std::vector<int64_t> ou(B), os(B), ov(B);
    for (int64_t i = 0; i < B; i++) {
        ou[i] = int64_t(arenas_[oa].alloc(size_t(plans[i].u.size * es_)));
        os[i] = int64_t(arenas_[oa].alloc(size_t(plans[i].s.size * es_)));
        ov[i] = int64_t(arenas_[oa].alloc(size_t(plans[i].v.size * es_)));
    }
    grow_all(st);
    t0 = now();
    Batch b2(dtype_);
    char* O = arenas_[oa].base;
    for (int64_t i = 0; i < B; i++) {
        void* bufs[8] = {nullptr, T + tm[i], T + t1[i], T + t2[i], S + sv_off[i], O + ou[i], O + os[i], O + ov[i]};
        b2.fill(plans[i].fills_s, bufs, es_);
    }
    b2.flush_fill();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[8] = {nullptr, T + tm[i], T + t1[i], T + t2[i], S + sv_off[i], O + ou[i], O + os[i], O + ov[i]};
        b2.perm(plans[i].perm_u, bufs, es_);
        b2.perm(plans[i].perm_v, bufs, es_);
    }
    b2.flush_perm();
    for (int64_t i = 0; i < B; i++) {
        void* bufs[8] = {nullptr, T + tm[i], T + t1[i], T + t2[i], S + sv_off[i], O + ou[i], O + os[i], O + ov[i]};
        b2.diag(plans[i].diag_s, bufs, es_);
    }
    b2.flush_diag();
    stats_.bind_s += now() - t0;
    stats_.launches += 3;
    run_batch(b2, st);
    for (int64_t i = 0; i < B; i++) {
        u[i]->sh = std::move(plans[i].u);
        u[i]->arena = oa;
        u[i]->off = ou[i];
        s[i]->sh = std::move(plans[i].s);
        s[i]->arena = oa;
        s[i]->off = os[i];
        v[i]->sh = std::move(plans[i].v);
        v[i]->arena = oa;
        v[i]->off = ov[i];
    }
}

void Engine::project_sites(hipStream_t st) {
    // shrinker tensors P [(q,1)] x Q conj(phys) with one unit entry
    // (sampling_lattice/lattice.py:291-339), projected once per site and
    // physical index: a projected site depends on nothing else.
    const Kind& k = kind_by_id(kind_);
    const int nsite = L1_ * L2_;
    std::vector<Shape> shr(size_t(nsite) * nphys_);
    std::vector<std::vector<double>> shr_data(shr.size());
    std::vector<Slot> raw(nsite), sh_slots(shr.size());
    arenas_[A_PERSIST].reset();
    for (int si = 0; si < nsite; si++) {
        const Shape& site = sites_[si];
        const Edge& phys = site.edges[site.index_of(names_[N_P0])];
        for (int p = 0; p < nphys_; p++) {
            int64_t off = p;
            int seg = 0;
            while (off >= phys.segs[seg].d) off -= phys.segs[seg++].d;
            const Sym q = phys.segs[seg].q;
            Shape& s = shr[si * nphys_ + p];
            s.kind_id = kind_;
            s.names = {names_[N_FIRST_SITE + si], names_[N_Q]};
            Edge pe;
            pe.segs = {{q, 1}};
            pe.arrow = false;
            s.edges = {pe, edge_conjugate(k, phys)};
            s.finalize();
            std::vector<double> d(size_t(s.size), 0.0);
            // element (P=(q,0), Q=(-q, off)): block (0, seg) at row 0, column off
            int64_t pos[2] = {0, seg};
            d[size_t(s.offset(pos) + off)] = 1.0;
            shr_data[si * nphys_ + p] = std::move(d);
        }
    }
    // upload shrinkers, keep `one`
    std::vector<int64_t> soff(shr.size());
    for (size_t i = 0; i < shr.size(); i++) soff[i] = int64_t(arenas_[A_PERSIST].alloc(size_t(shr[i].size * es_)));
    const int64_t one_off = int64_t(arenas_[A_PERSIST].alloc(size_t(es_)));
    keep_[A_PERSIST] = 0;
    grow_all(st);
    {
        std::vector<char> host(arenas_[A_PERSIST].used, 0);
        for (size_t i = 0; i < shr.size(); i++) {
            for (size_t j = 0; j < shr_data[i].size(); j++) {
                if (dtype_ == 0)
                    reinterpret_cast<double*>(host.data() + soff[i])[j] = shr_data[i][j];
                else
                    reinterpret_cast<float*>(host.data() + soff[i])[j] = float(shr_data[i][j]);
            }
        }
        if (dtype_ == 0)
            *reinterpret_cast<double*>(host.data() + one_off) = 1.0;
        else
            *reinterpret_cast<float*>(host.data() + one_off) = 1.0f;
        check_hip(hipMemcpyAsync(arenas_[A_PERSIST].base, host.data(), host.size(), hipMemcpyHostToDevice, st),
                  "shrinker upload");
        check_hip(hipStreamSynchronize(st), "shrinker sync");
    }
    one_.sh = Shape();
    one_.sh.kind_id = kind_;
    one_.sh.finalize();
    one_.arena = A_PERSIST;
    one_.off = one_off;
    // contract every site with each of its shrinkers in one batch; raw sites
    // are caller-owned buffers, addressed through a pseudo arena slot
    std::vector<Slot> a(shr.size()), b(shr.size()), o(shr.size());
    std::vector<Slot*> pa, pb, po;
    for (size_t i = 0; i < shr.size(); i++) {
        a[i].sh = sites_[i / nphys_];
        b[i].sh = shr[i];
        b[i].arena = A_PERSIST;
        b[i].off = soff[i];
        pa.push_back(&a[i]);
        pb.push_back(&b[i]);
        po.push_back(&o[i]);
    }
    // raw site data lives outside the arenas: temporarily map it by copying
    // into the persistent arena (done once per engine)
    std::vector<int64_t> roff(nsite);
    for (int si = 0; si < nsite; si++) roff[si] = int64_t(arenas_[A_PERSIST].alloc(size_t(sites_[si].size * es_)));
    keep_[A_PERSIST] = arenas_[A_PERSIST].used;
    grow_all(st);
    for (int si = 0; si < nsite; si++) {
        if (sites_[si].size)
            check_hip(hipMemcpyAsync(arenas_[A_PERSIST].base + roff[si], site_data_[si], size_t(sites_[si].size * es_),
                                     hipMemcpyDeviceToDevice, st),
                      "site copy");
    }
    for (size_t i = 0; i < shr.size(); i++) {
        a[i].arena = A_PERSIST;
        a[i].off = roff[i / nphys_];
    }
    // the projection selects elements: it runs in the storage type whatever
    // the GEMM precision, so projected sites equal the raw site values
    const std::vector<std::pair<int32_t, int32_t>> pr{{names_[N_P0], names_[N_Q]}};
    const int prec = gemm_prec_;
    gemm_prec_ = 0;
    contract(pa, pb, pr, po, A_PERSIST, st);
    gemm_prec_ = prec;
    keep_[A_PERSIST] = arenas_[A_PERSIST].used;
    proj_.assign(o.begin(), o.end());
}

void Engine::evaluate(int64_t nconf, const int32_t* occupations, int64_t remain_cut, double relative_cut, double* psi,
                      hipStream_t st) {
    const double t0 = now();
    remain_cut_ = remain_cut;
    relative_cut_ = relative_cut;
    if (proj_.empty()) project_sites(st);
    const int64_t per = int64_t(L1_) * L2_;
    for (int64_t c0 = 0; c0 < nconf; c0 += chunk_) {
        const int64_t nc = std::min(chunk_, nconf - c0);
        evaluate_chunk(nc, occupations + c0 * per, psi + c0, st);
    }
    stats_.total_s += now() - t0;
    stats_.configs += double(nconf);
    size_t peak = 0;
    for (auto& a : arenas_) peak += a.cap;
    stats_.peak_bytes = std::max(stats_.peak_bytes, double(peak));
}

void Engine::evaluate_chunk(int64_t B, const int32_t* occ, double* psi, hipStream_t st) {
    const int64_t per = int64_t(L1_) * L2_;
    for (int64_t i = 0; i < B * per; i++)
        if (occ[i] < 0 || occ[i] >= nphys_) throw Error("occupation index out of range of the physical edge");
    auto proj = [&](int64_t b, int l1, int l2) -> const Slot& {
        return proj_[size_t((l1 * L2_ + l2) * nphys_ + occ[b * per + l1 * L2_ + l2])];
    };
    const int32_t P0 = names_[N_P0], U = names_[N_U], D = names_[N_D], L = names_[N_L], R = names_[N_R];
    const int32_t nL1 = names_[N_L1], nL2 = names_[N_L2], nL3 = names_[N_L3], nR1 = names_[N_R1], nR2 = names_[N_R2];
    (void)P0;
    std::vector<std::vector<Slot>> mps(B, std::vector<Slot>(L2_, one_)), dl(B, std::vector<Slot>(L2_));
    std::vector<Slot> ta(B), tb(B), tc(B), td(B), te(B);
    auto ptrs = [&](std::vector<Slot>& v) {
        std::vector<Slot*> p(v.size());
        for (size_t i = 0; i < v.size(); i++) p[i] = &v[i];
        return p;
    };
    auto col = [&](std::vector<std::vector<Slot>>& m, int i) {
        std::vector<Slot*> p(B);
        for (int64_t b = 0; b < B; b++) p[b] = &m[b][i];
        return p;
    };
    int in_arena = A_ROW0;
    arenas_[A_ROW0].reset();
    arenas_[A_ROW1].reset();
    keep_[A_ROW0] = keep_[A_ROW1] = 0;
    for (int l1 = L1_ - 1; l1 >= 1; l1--) {
        const int oa = in_arena == A_ROW0 ? A_ROW1 : A_ROW0;
        arenas_[oa].reset();
        keep_[oa] = 0;
        // double_line[i] = contract(line_1[i], line_2[i], {(U, D)})
        for (int i = 0; i < L2_; i++) {
            for (int64_t b = 0; b < B; b++) {
                ta[b] = mps[b][i];
                tb[b] = proj(b, l1, i);
            }
            auto pa = ptrs(ta), pb = ptrs(tb);
            rename(pa, {{L, nL1}, {R, nR1}});
            rename(pb, {{L, nL2}, {R, nR2}});
            auto po = col(dl, i);
            contract(pa, pb, {{U, D}}, po, oa, st);
        }
        // left-to-right QR sweep
        for (int i = 0; i + 1 < L2_; i++) {
            auto pt = col(dl, i);
            auto pq = ptrs(ta), pr = ptrs(tb);
            qr(pt, {nR1, nR2}, R, L, pq, pr, oa, st);
            for (int64_t b = 0; b < B; b++) dl[b][i] = ta[b];
            auto pn = col(dl, i + 1);
            auto pc = ptrs(tc);
            contract(pn, pr, {{nL1, nR1}, {nL2, nR2}}, pc, oa, st);
            for (int64_t b = 0; b < B; b++) dl[b][i + 1] = tc[b];
        }
        // right-to-left truncated SVD sweep
        for (int i = L2_ - 1; i >= 1; i--) {
            auto pt = col(dl, i);
            auto pu = ptrs(ta), ps = ptrs(tb), pv = ptrs(tc);
            svd(pt, {L}, R, L, L, R, pu, ps, pv, oa, st);
            for (int64_t b = 0; b < B; b++) dl[b][i] = tc[b];
            auto pp = col(dl, i - 1);
            auto pd = ptrs(td);
            contract(pp, pu, {{R, L}}, pd, oa, st);
            auto pe = ptrs(te);
            contract(pd, ps, {{R, L}}, pe, oa, st);
            for (int64_t b = 0; b < B; b++) dl[b][i - 1] = te[b];
        }
        for (int64_t b = 0; b < B; b++) mps[b] = dl[b];
        in_arena = oa;
    }
    // final row: _inline_right_to_left along row 0 with up_to_down = ones
    const int oa = in_arena == A_ROW0 ? A_ROW1 : A_ROW0;
    if (L1_ > 1) {
        arenas_[oa].reset();
        keep_[oa] = 0;
    }
    std::vector<Slot> inl(B, one_);
    for (int l2 = L2_ - 1; l2 >= 0; l2--) {
        auto pi = ptrs(inl);
        auto pm = col(mps, l2);
        auto pa = ptrs(ta);
        contract(pi, pm, {{nL3, R}}, pa, oa, st);
        rename(pa, {{L, nL3}});
        for (int64_t b = 0; b < B; b++) tb[b] = proj(b, 0, l2);
        auto pb = ptrs(tb);
        rename(pb, {{L, nL2}});
        auto pc = ptrs(tc);
        contract(pa, pb, {{nL2, R}, {U, D}}, pc, oa, st);
        for (int64_t b = 0; b < B; b++) td[b] = one_;
        auto pd = ptrs(td);
        auto pe = ptrs(te);
        contract(pc, pd, {{nL1, R}, {U, D}}, pe, oa, st);
        for (int64_t b = 0; b < B; b++) inl[b] = te[b];
    }
    // gather psi (one element per configuration, zero if no block survives)
    arenas_[A_TEMP].reset();
    keep_[A_TEMP] = 0;
    const int64_t goff = int64_t(arenas_[A_TEMP].alloc(size_t(B * es_)));
    grow_all(st);
    Batch bt(dtype_);
    char* G = arenas_[A_TEMP].base + goff;
    bt.fill_raw(G, B);
    bt.flush_fill();
    for (int64_t b = 0; b < B; b++) {
        if (inl[b].sh.size != 1) continue;
        PermTask t;
        t.rank = 1;
        t.n = 1;
        t.dims[0] = 1;
        t.ss[0] = 1;
        t.ds[0] = 1;
        t.src = 0;
        t.dst = b;
        void* bufs[2] = {ptr(inl[b]), G};
        t.src_buf = 0;
        t.dst_buf = 1;
        bt.perm_one(t, bufs, es_);
    }
    bt.flush_perm();
    run_batch(bt, st);
    std::vector<char> host(size_t(B * es_));
    check_hip(hipMemcpyAsync(host.data(), G, host.size(), hipMemcpyDeviceToHost, st), "psi copy");
    check_hip(hipStreamSynchronize(st), "psi sync");
    if (profile_) timer_.collect();
    for (int64_t b = 0; b < B; b++)
        psi[b] = dtype_ == 0 ? reinterpret_cast<double*>(host.data())[b] : double(reinterpret_cast<float*>(host.data())[b]);
}

}  // namespace tns

// ------------------------------------------------------------------ C ABI
// A handle holds one or more lanes: independent engines over the same PEPS,
// each with its own host thread, HIP stream, arenas and task staging.  A call
// splits the configurations over the lanes so that the decomposition tail
// and host planning of one lane overlap the kernels of the others.
struct tns_engine {
    int kind, L1, L2, dtype, threads;
    std::vector<tns::Shape> sites;
    std::vector<const void*> data;
    std::vector<int32_t> names;
    int64_t chunk = 1024;
    int precision = 0;
    std::vector<int> gemm_table;  // tuned tile variants, empty until autotuned
    bool profile = false;
    std::vector<std::unique_ptr<tns::Engine>> lanes;
    std::vector<hipStream_t> streams;  // streams[0] unused (caller's stream)
    ~tns_engine() {
        lanes.clear();
        for (size_t i = 1; i < streams.size(); i++)
            if (streams[i]) (void)hipStreamDestroy(streams[i]);
    }
    void add_lanes(int n) {
        const int per = std::max(1, threads / std::max(1, n));
        while (int(lanes.size()) < n) {
            lanes.push_back(std::make_unique<tns::Engine>(kind, L1, L2, sites, data, dtype, names, per));
            lanes.back()->set_chunk(chunk);
            lanes.back()->set_precision(precision);
            if (gemm_table.size() == 3) lanes.back()->set_gemm_table(gemm_table.data());
            lanes.back()->set_profile(profile);
            hipStream_t s = nullptr;
            if (lanes.size() > 1) tns::check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "lane stream");
            streams.push_back(s);
        }
    }
};

extern "C" {

int tns_engine_create(int kind, int L1, int L2, const tns_shape* const* sites, const void* const* site_data,
                      int dtype, const int32_t* name_ids, int n_names, int threads, tns_engine** out) {
    TNS_TRY({
        if (dtype != 0 && dtype != 1) throw tns::Error("unsupported dtype");
        auto h = std::make_unique<tns_engine>();
        h->kind = kind;
        h->L1 = L1;
        h->L2 = L2;
        h->dtype = dtype;
        h->threads = std::max(1, threads);
        for (int i = 0; i < L1 * L2; i++) {
            if (sites[i]->s.kind_id != kind) throw tns::Error("site symmetry differs from engine kind");
            h->sites.push_back(sites[i]->s);
            h->data.push_back(site_data[i]);
        }
        h->names.assign(name_ids, name_ids + n_names);
        h->add_lanes(1);
        *out = h.release();
    })
}

int tns_engine_set_chunk(tns_engine* e, int64_t chunk) {
    TNS_TRY({
        e->chunk = chunk;
        for (auto& l : e->lanes) l->set_chunk(chunk);
    })
}

int tns_engine_set_precision(tns_engine* e, int precision) {
    TNS_TRY({
        if (precision < 0 || precision > 3) throw tns::Error("precision must be 0 (fp64), 1 (fp32), 2 (bf16x3) or 3 (bf16x6)");
        if (precision != 0 && e->dtype != 0) throw tns::Error("reduced-precision GEMMs need a float64 engine");
        e->precision = precision;
        for (auto& l : e->lanes) l->set_precision(precision);
    })
}

int tns_engine_autotune(tns_engine* e, int64_t nconf, const int32_t* occupations, int64_t remain_cut,
                        double relative_cut, int32_t* table, double* ms, void* stream) {
    TNS_TRY({
        if (nconf <= 0) throw tns::Error("autotune needs configurations");
        std::vector<double> psi(static_cast<size_t>(nconf), 0.0);
        auto& l0 = *e->lanes[0];
        l0.request_gemm_tuning();
        l0.evaluate(nconf, occupations, remain_cut, relative_cut, psi.data(), static_cast<hipStream_t>(stream));
        l0.finish_gemm_tuning();
        for (auto& l : e->lanes) l->set_gemm_table(l0.gemm_table());
        e->gemm_table.assign(l0.gemm_table(), l0.gemm_table() + 3);
        if (table)
            for (int c = 0; c < 3; c++) table[c] = l0.gemm_table()[c];
        if (ms)
            for (int i = 0; i < 9; i++) ms[i] = l0.gemm_tuning_ms()[i];
    })
}

int tns_engine_set_lanes(tns_engine* e, int lanes) {
    TNS_TRY({
        if (lanes < 1 || lanes > 32) throw tns::Error("lanes must be in [1, 32]");
        if (lanes < int(e->lanes.size())) {
            e->lanes.resize(size_t(lanes));
            for (size_t i = size_t(lanes); i < e->streams.size(); i++)
                if (e->streams[i]) (void)hipStreamDestroy(e->streams[i]);
            e->streams.resize(size_t(lanes));
        }
        e->add_lanes(lanes);
    })
}

int tns_engine_evaluate(tns_engine* e, int64_t nconf, const int32_t* occupations, int64_t remain_cut,
                        double relative_cut, double* psi, void* stream) {
    TNS_TRY({
        hipStream_t caller = static_cast<hipStream_t>(stream);
        const int64_t K = std::min<int64_t>(int64_t(e->lanes.size()), std::max<int64_t>(1, nconf));
        if (K <= 1) {
            e->lanes[0]->evaluate(nconf, occupations, remain_cut, relative_cut, psi, caller);
        } else {
            // lanes run on their own streams: order them after the caller's work
            tns::check_hip(hipStreamSynchronize(caller), "caller stream sync");
            const int64_t per = int64_t(e->L1) * e->L2;
            std::vector<std::exception_ptr> errs;
            errs.resize(size_t(K));
            auto run = [&](int64_t k) {
                const int64_t b = nconf * k / K, en = nconf * (k + 1) / K;
                try {
                    e->lanes[k]->evaluate(en - b, occupations + b * per, remain_cut, relative_cut, psi + b,
                                          k == 0 ? caller : e->streams[k]);
                } catch (...) {
                    errs[k] = std::current_exception();
                }
            };
            std::vector<std::thread> ts;
            for (int64_t k = 1; k < K; k++) ts.emplace_back(run, k);
            run(0);
            for (auto& t : ts) t.join();
            for (auto& ep : errs)
                if (ep) std::rethrow_exception(ep);
        }
    })
}

int tns_engine_stats(const tns_engine* e, double* values, int n) {
    TNS_TRY({
        double v[12] = {0};
        for (auto& l : e->lanes) {
            const auto& s = l->stats();
            const double x[12] = {s.plan_s,     s.exec_s,   s.total_s,    s.gemm_flops, s.qr_flops, s.svd_flops,
                                  s.perm_elems, s.launches, s.peak_bytes, s.configs,    s.bind_s,   s.sync_s};
            for (int i = 0; i < 12; i++) v[i] += x[i];
        }
        // wall-clock style entries: the lanes run concurrently
        double tmax = 0;
        for (auto& l : e->lanes) tmax = std::max(tmax, l->stats().total_s);
        v[2] = tmax;
        for (int i = 0; i < n && i < 12; i++) values[i] = v[i];
    })
}

int tns_engine_reset_stats(tns_engine* e) {
    TNS_TRY({
        for (auto& l : e->lanes) l->reset_stats();
    })
}

int tns_engine_trim(tns_engine* e) {
    TNS_TRY({
        for (auto& l : e->lanes) l->trim();
    })
}

int tns_engine_refresh_sites(tns_engine* e) {
    TNS_TRY({
        for (auto& l : e->lanes) l->refresh_sites();
    })
}

int tns_engine_set_profile(tns_engine* e, int on) {
    TNS_TRY({
        e->profile = on != 0;
        for (auto& l : e->lanes) l->set_profile(on != 0);
    })
}

int tns_engine_kernel_stats(const tns_engine* e, double* values, int n) {
    TNS_TRY({
        std::vector<double> acc(size_t(3 * tns::KernelTimer::NKIND), 0.0);
        for (auto& l : e->lanes) {
            const auto& t = l->timer();
            for (int k = 0; k < tns::KernelTimer::NKIND; k++) {
                acc[3 * k] += t.acc[k].launches;
                acc[3 * k + 1] += t.acc[k].ms;
                acc[3 * k + 2] += t.acc[k].units;
            }
        }
        for (int i = 0; i < n && i < int(acc.size()); i++) values[i] = acc[size_t(i)];
    })
}

int tns_engine_free(tns_engine* e) { TNS_TRY({ delete e; }) }

}  // extern "C"
