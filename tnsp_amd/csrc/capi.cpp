// C ABI of libtnsgpu.so (declared in include/tnsgpu.h): shape handles, plans
// built by the host planners in core.cpp, and their execution on a HIP stream.
//
// Planner introspection (tns_plan_stage_export) record layouts, all int64:
//   perm  : src_buf dst_buf src dst rank neg n dims[8] ss[8] ds[8]   (31)
//   gemm  : a_buf b_buf c_buf a b c m n k lda ldb ldc ta tb alpha    (15)
//   fill  : buf off n                                                (3)
//   qr    : in_buf f1_buf f2_buf in f1 f2 m n lq                     (9)
//   svd   : in_buf u_buf s_buf v_buf in u s v m n                    (10)
// Stages: edge op / conjugate: 0 perm.  contract: 0 perm_a, 1 perm_b,
// 2 gemm, 3 fill, 4 perm_out.  svd: 0 perm_in, 1 svd, 2 perm_u, 3 perm_v,
// 4 fill_s, 5 diag_s.  qr: 0 perm_in, 1 qr, 2 perm_1, 3 perm_2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tnsgpu.h"
#include "core.hpp"
#include "engine.hpp"
#include "handles.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace tns;

struct tns_plan {
    enum Type { EDGEOP, CONTRACT, SVD, QR, CONJ } type;
    Shape eo_out;
    std::vector<PermTask> eo_tasks;
    ContractPlan c;
    SvdPlan s;
    QrPlan q;
    bool factored = false;
    std::vector<double> svals_host;
};

namespace {
thread_local std::string g_error;
}  // namespace

int tns::set_error(const char* what) {
    g_error = what;
    return 1;
}

namespace {

Shape make_shape(int kind, int rank, const int32_t* names, const int32_t* nseg, const int32_t* syms,
                 const int64_t* dims, const uint8_t* arrows) {
    Shape s;
    s.kind_id = kind;
    kind_by_id(kind);
    int64_t seg = 0;
    for (int i = 0; i < rank; i++) {
        s.names.push_back(names[i]);
        Edge e;
        e.arrow = arrows ? arrows[i] != 0 : false;
        for (int j = 0; j < nseg[i]; j++, seg++) {
            Segment g;
            g.q.v[0] = syms[2 * seg];
            g.q.v[1] = syms[2 * seg + 1];
            if (dims[seg] < 0) throw Error("negative segment dimension");
            g.d = dims[seg];
            e.segs.push_back(g);
        }
        s.edges.push_back(std::move(e));
    }
    for (int i = 0; i < rank; i++)
        for (int j = i + 1; j < rank; j++)
            if (s.names[i] == s.names[j]) throw Error("duplicated names in tensor");
    s.finalize();
    return s;
}

tns_shape* new_handle(const Shape& s) {
    auto* h = new tns_shape;
    h->s = s;
    return h;
}

std::vector<int32_t> ids(int n, const int32_t* p) { return std::vector<int32_t>(p, p + n); }

void export_perm(const std::vector<PermTask>& ts, int64_t* r) {
    for (auto& t : ts) {
        int64_t* o = r;
        o[0] = t.src_buf;
        o[1] = t.dst_buf;
        o[2] = t.src;
        o[3] = t.dst;
        o[4] = t.rank;
        o[5] = t.neg;
        o[6] = t.n;
        for (int i = 0; i < MAXR; i++) {
            o[7 + i] = t.dims[i];
            o[15 + i] = t.ss[i];
            o[23 + i] = t.ds[i];
        }
        r += 31;
    }
}

// element counts of every logical buffer of a plan (see the plan structs)
std::vector<int64_t> buffer_sizes(const tns_plan* p) {
    switch (p->type) {
        case tns_plan::EDGEOP:
        case tns_plan::CONJ:
            return {-1, p->eo_out.size};
        case tns_plan::CONTRACT:
            return {-1, -1, p->c.ma.size, p->c.mb.size, p->c.prod.size, p->c.result.size};
        case tns_plan::SVD:
            return {-1,
                    p->s.merged.size,
                    p->s.t1.size,
                    p->s.t2.size,
                    p->s.s_total,
                    p->factored ? p->s.u.size : -1,
                    p->factored ? p->s.s.size : -1,
                    p->factored ? p->s.v.size : -1};
        case tns_plan::QR:
            return {-1, p->q.merged.size, p->q.t1.size, p->q.t2.size, p->q.q.size, p->q.r.size};
    }
    return {};
}

// workspace layout: element buffers then double scratch
struct Layout {
    std::vector<int64_t> off;  // byte offset per buffer id (-1 = not in workspace)
    int64_t scratch = 0;       // byte offset of double scratch
    int64_t total = 0;
};
int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

int64_t qr_scratch(const QrTask& t) { return qr_doubles(t.m, t.n); }
int64_t svd_scratch(const SvdTask& t) { return svd_doubles(t.m, t.n); }

Layout layout_of(const tns_plan* p, int dtype) {
    const int64_t es = dtype == 0 ? 8 : 4;
    Layout L;
    auto sizes = buffer_sizes(p);
    L.off.assign(sizes.size(), -1);
    auto put = [&](int id, int64_t bytes) {
        L.off[id] = L.total;
        L.total = align256(L.total + bytes);
    };
    switch (p->type) {
        case tns_plan::CONTRACT:
            put(2, sizes[2] * es);
            put(3, sizes[3] * es);
            put(4, sizes[4] * es);
            break;
        case tns_plan::SVD: {
            put(1, sizes[1] * es);
            put(2, sizes[2] * es);
            put(3, sizes[3] * es);
            put(4, sizes[4] * 8);
            int64_t sc = 0;
            for (auto& t : p->s.svds) sc += svd_scratch(t);
            L.scratch = L.total;
            L.total = align256(L.total + sc * 8);
            break;
        }
        case tns_plan::QR: {
            put(1, sizes[1] * es);
            put(2, sizes[2] * es);
            put(3, sizes[3] * es);
            int64_t sc = 0;
            for (auto& t : p->q.qrs) sc += qr_scratch(t);
            L.scratch = L.total;
            L.total = align256(L.total + sc * 8);
            break;
        }
        default:
            break;
    }
    return L;
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

const char* tns_last_error(void) { return g_error.c_str(); }
const char* tns_version(void) { return "tnsgpu 0.1 (gfx950)"; }

int tns_set_svd_max_sweeps(int sweeps) { TNS_TRY({ tns::set_svd_max_sweeps(sweeps); }) }

int tns_device_count(int* count) {
    TNS_TRY({
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        if (e != hipSuccess) c = 0;
        *count = c;
    })
}

int tns_shape_create(int kind, int rank, const int32_t* names, const int32_t* nseg, const int32_t* syms,
                     const int64_t* dims, const uint8_t* arrows, tns_shape** out) {
    TNS_TRY({ *out = new_handle(make_shape(kind, rank, names, nseg, syms, dims, arrows)); })
}

int tns_shape_free(tns_shape* s) { TNS_TRY({ delete s; }) }

int tns_shape_counts(const tns_shape* h, int32_t* kind, int32_t* rank, int32_t* total_segments, int64_t* nblocks,
                     int64_t* size) {
    TNS_TRY({
        const Shape& s = h->s;
        if (kind) *kind = s.kind_id;
        if (rank) *rank = s.rank();
        int32_t ts = 0;
        for (auto& e : s.edges) ts += int32_t(e.segs.size());
        if (total_segments) *total_segments = ts;
        if (nblocks) *nblocks = int64_t(s.blocks.size());
        if (size) *size = s.size;
    })
}

int tns_shape_export(const tns_shape* h, int32_t* names, int32_t* nseg, int32_t* syms, int64_t* dims,
                     uint8_t* arrows) {
    TNS_TRY({
        const Shape& s = h->s;
        int64_t seg = 0;
        for (int i = 0; i < s.rank(); i++) {
            names[i] = s.names[i];
            nseg[i] = int32_t(s.edges[i].segs.size());
            arrows[i] = s.edges[i].arrow;
            for (auto& g : s.edges[i].segs) {
                syms[2 * seg] = g.q.v[0];
                syms[2 * seg + 1] = g.q.v[1];
                dims[seg] = g.d;
                seg++;
            }
        }
    })
}

int tns_shape_blocks(const tns_shape* h, int64_t* positions, int64_t* offsets) {
    TNS_TRY({
        const Shape& s = h->s;
        for (size_t b = 0; b < s.blocks.size(); b++) {
            s.positions_of(s.blocks[b], positions + b * s.rank());
            offsets[b] = s.block_off[s.blocks[b]];
        }
    })
}

int tns_shape_rename(const tns_shape* h, int n, const int32_t* old_names, const int32_t* new_names, tns_shape** out) {
    TNS_TRY({
        Shape s = h->s;
        for (int i = 0; i < n; i++) {
            int f = h->s.index_of(old_names[i]);
            if (f < 0) throw Error("Name missing in edge_rename");
            s.names[f] = new_names[i];
        }
        for (int i = 0; i < s.rank(); i++)
            for (int j = i + 1; j < s.rank(); j++)
                if (s.names[i] == s.names[j]) throw Error("duplicated names after edge_rename");
        *out = new_handle(s);
    })
}

int tns_plan_edge_operator(const tns_shape* src, int n_split, const int32_t* split_old, const int32_t* split_count,
                           const int32_t* split_names, const tns_shape* split_edges, int n_rev, const int32_t* rev,
                           int n_merge, const int32_t* merge_new, const int32_t* merge_count, const int32_t* merge_old,
                           int n_new, const int32_t* new_names, int apply_parity, int n_ex_split,
                           const int32_t* ex_split, int n_ex_rb, const int32_t* ex_rb, int n_ex_ra,
                           const int32_t* ex_ra, int n_ex_merge, const int32_t* ex_merge, tns_plan** out) {
    TNS_TRY({
        EdgeOpArgs a;
        int sub = 0;
        for (int i = 0; i < n_split; i++) {
            SplitSpec sp;
            sp.old_name = split_old[i];
            for (int j = 0; j < split_count[i]; j++, sub++) {
                if (!split_edges || sub >= split_edges->s.rank()) throw Error("split edge list too short");
                Edge e = split_edges->s.edges[sub];
                e.arrow = false;
                sp.parts.push_back({split_names[sub], e});
            }
            a.split.push_back(std::move(sp));
        }
        a.reversed = ids(n_rev, rev);
        int mo = 0;
        for (int i = 0; i < n_merge; i++) {
            MergeSpec ms;
            ms.new_name = merge_new[i];
            for (int j = 0; j < merge_count[i]; j++) ms.olds.push_back(merge_old[mo++]);
            a.merge.push_back(std::move(ms));
        }
        a.new_names = ids(n_new, new_names);
        a.apply_parity = apply_parity != 0;
        a.ex_split = ids(n_ex_split, ex_split);
        a.ex_rev_before = ids(n_ex_rb, ex_rb);
        a.ex_rev_after = ids(n_ex_ra, ex_ra);
        a.ex_merge = ids(n_ex_merge, ex_merge);
        auto p = std::make_unique<tns_plan>();
        p->type = tns_plan::EDGEOP;
        plan_edge_operator(src->s, a, p->eo_out, p->eo_tasks, 0, 1);
        *out = p.release();
    })
}

int tns_plan_contract(const tns_shape* a, const tns_shape* b, int npairs, const int32_t* pairs, tns_plan** out) {
    TNS_TRY({
        std::vector<std::pair<int32_t, int32_t>> pr;
        for (int i = 0; i < npairs; i++) pr.push_back({pairs[2 * i], pairs[2 * i + 1]});
        auto p = std::make_unique<tns_plan>();
        p->type = tns_plan::CONTRACT;
        plan_contract(a->s, b->s, pr, p->c);
        *out = p.release();
    })
}

int tns_plan_svd(const tns_shape* t, int nfree, const int32_t* free_u, int32_t cu, int32_t cv, int32_t su,
                 int32_t sv, tns_plan** out) {
    TNS_TRY({
        auto p = std::make_unique<tns_plan>();
        p->type = tns_plan::SVD;
        plan_svd_factor(t->s, ids(nfree, free_u), cu, cv, su, sv, p->s);
        *out = p.release();
    })
}

int tns_plan_qr(const tns_shape* t, char direction, int nfree, const int32_t* free_names, int32_t cq, int32_t cr,
                tns_plan** out) {
    TNS_TRY({
        auto p = std::make_unique<tns_plan>();
        p->type = tns_plan::QR;
        plan_qr(t->s, direction, ids(nfree, free_names), cq, cr, p->q);
        *out = p.release();
    })
}

int tns_plan_conjugate(const tns_shape* t, int trivial_metric, tns_plan** out) {
    TNS_TRY({
        auto p = std::make_unique<tns_plan>();
        p->type = tns_plan::CONJ;
        plan_conjugate(t->s, trivial_metric != 0, p->eo_out, p->eo_tasks);
        *out = p.release();
    })
}

int tns_plan_free(tns_plan* p) { TNS_TRY({ delete p; }) }

int tns_plan_num_outputs(const tns_plan* p, int* n) {
    TNS_TRY({
        switch (p->type) {
            case tns_plan::SVD:
                *n = 3;
                break;
            case tns_plan::QR:
                *n = 2;
                break;
            default:
                *n = 1;
        }
    })
}

int tns_plan_output_shape(const tns_plan* p, int which, tns_shape** out) {
    TNS_TRY({
        const Shape* s = nullptr;
        switch (p->type) {
            case tns_plan::EDGEOP:
            case tns_plan::CONJ:
                if (which == 0) s = &p->eo_out;
                break;
            case tns_plan::CONTRACT:
                if (which == 0) s = &p->c.result;
                break;
            case tns_plan::SVD:
                if (!p->factored) throw Error("svd outputs are known only after tns_svd_factor");
                s = which == 0 ? &p->s.u : which == 1 ? &p->s.s : which == 2 ? &p->s.v : nullptr;
                break;
            case tns_plan::QR:
                s = which == 0 ? &p->q.q : which == 1 ? &p->q.r : nullptr;
                break;
        }
        if (!s) throw Error("no such plan output");
        *out = new_handle(*s);
    })
}

int tns_plan_workspace(const tns_plan* p, int dtype, int64_t* bytes) {
    TNS_TRY({ *bytes = layout_of(p, dtype).total; })
}

int tns_plan_flops(const tns_plan* p, int64_t* flops) {
    TNS_TRY({
        switch (p->type) {
            case tns_plan::CONTRACT:
                *flops = p->c.flops;
                break;
            case tns_plan::SVD:
                *flops = p->s.flops;
                break;
            case tns_plan::QR:
                *flops = p->q.flops;
                break;
            default:
                *flops = 0;
        }
    })
}

int tns_plan_run(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* workspace,
                 void* stream) {
    TNS_TRY({ tns::run_plan(p, dtype, inputs, outputs, workspace, as_stream(stream), nullptr); })
}

}  // extern "C"

int64_t tns::plan_workspace_bytes(const tns_plan* p, int dtype) { return layout_of(p, dtype).total; }

bool tns::plan_is_svd(const tns_plan* p) { return p->type == tns_plan::SVD; }

void tns::run_plan(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* workspace,
                   hipStream_t stream, Staging* staging) {
    {
        if (dtype != 0 && dtype != 1) throw Error("unsupported dtype");
        const int64_t es = dtype == 0 ? 8 : 4;
        Layout L = layout_of(p, dtype);
        char* ws = static_cast<char*>(workspace);
        if (L.total > 0 && !ws) throw Error("workspace required");
        Batch bt(dtype);
        if (staging) bt.set_staging(staging);
        switch (p->type) {
            case tns_plan::EDGEOP:
            case tns_plan::CONJ: {
                void* bufs[2] = {const_cast<void*>(inputs[0]), outputs[0]};
                int64_t covered = 0;
                for (auto& t : p->eo_tasks) covered += t.n;
                if (covered < p->eo_out.size) bt.fill({FillTask{1, 0, p->eo_out.size}}, bufs, es);
                bt.flush_fill();
                bt.perm(p->eo_tasks, bufs, es);
                bt.flush_perm();
                break;
            }
            case tns_plan::CONTRACT: {
                void* bufs[6] = {const_cast<void*>(inputs[0]), const_cast<void*>(inputs[1]), ws + L.off[2],
                                 ws + L.off[3], ws + L.off[4], outputs[0]};
                bt.perm(p->c.perm_a, bufs, es);
                bt.perm(p->c.perm_b, bufs, es);
                bt.flush_perm();
                bt.fill(p->c.fills, bufs, es);
                bt.flush_fill();
                bt.gemm(p->c.gemms, bufs, es);
                bt.flush_gemm();
                bt.perm(p->c.perm_out, bufs, es);
                bt.flush_perm();
                break;
            }
            case tns_plan::QR: {
                void* bufs[6] = {const_cast<void*>(inputs[0]), ws + L.off[1], ws + L.off[2], ws + L.off[3], outputs[0],
                                 outputs[1]};
                bt.perm(p->q.perm_in, bufs, es);
                bt.flush_perm();
                double* sc = reinterpret_cast<double*>(ws + L.scratch);
                bt.qr(p->q.qrs, bufs, es, sc);
                bt.flush_qr();
                bt.perm(p->q.perm_1, bufs, es);
                bt.perm(p->q.perm_2, bufs, es);
                bt.flush_perm();
                break;
            }
            case tns_plan::SVD: {
                if (!p->factored) throw Error("call tns_svd_factor before tns_plan_run for svd");
                void* bufs[8] = {nullptr, ws + L.off[1], ws + L.off[2], ws + L.off[3], ws + L.off[4],
                                 outputs[0], outputs[1], outputs[2]};
                bt.fill(p->s.fills_s, bufs, es);
                bt.flush_fill();
                // singular values live in a double buffer; copy with a cast
                bt.perm(p->s.perm_u, bufs, es);
                bt.perm(p->s.perm_v, bufs, es);
                bt.flush_perm();
                bt.svals_to_diag(p->s.diag_s, bufs, es);
                break;
            }
        }
        bt.run(stream);
    }
}

extern "C" {

// ------------------------------------------------------------- batches
// The same ops for many tensors at once: planning spread over host threads,
// execution as one launch per kernel kind for all plans together.
namespace {
// One persistent pool serves the item loops (a batch op plans tens to
// hundreds of items in tens of microseconds each: spawning threads per call
// would cost more than the work).  A caller that finds the pool busy on
// another thread plans its items serially.
std::mutex pool_mu;

Pool& shared_pool(int threads) {
    static std::unique_ptr<Pool> pool;
    if (!pool || pool->size() != threads) pool = std::make_unique<Pool>(threads);
    return *pool;
}

void parallel_items(int n, int threads, const std::function<void(int)>& fn) {
    const int T = std::max(1, std::min(threads, 64));
    std::unique_lock<std::mutex> lk(pool_mu, std::try_to_lock);
    if (!lk.owns_lock() || T == 1 || n < 4) {
        for (int i = 0; i < n; i++) fn(i);
        return;
    }
    shared_pool(T).parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; i++) fn(int(i));
    });
}

// Host-mapped pinned buffer of this host thread (kernels write into it
// directly); kept until the thread ends.
struct MappedHost {
    void* p = nullptr;
    size_t cap = 0;
    ~MappedHost() {
        if (p) (void)hipHostFree(p);
    }
};

double* mapped_host_buffer(size_t bytes) {
    thread_local MappedHost buf;
    if (bytes > buf.cap) {
        if (buf.p) check_hip(hipHostFree(buf.p), "mapped host free");
        buf.p = nullptr;
        buf.cap = 0;
        const size_t cap = std::max<size_t>(bytes * 2, size_t(1) << 20);
        check_hip(hipHostMalloc(&buf.p, cap, hipHostMallocMapped | hipHostMallocCoherent), "mapped host alloc");
        buf.cap = cap;
    }
    return static_cast<double*>(buf.p);
}

void check_same_type(int n, tns_plan* const* plans) {
    for (int i = 1; i < n; i++)
        if (plans[i]->type != plans[0]->type) throw Error("a plan batch must hold one op type");
}
}  // namespace

int tns_plan_contract_many(int n, const tns_shape* const* a, const tns_shape* const* b, int npairs,
                           const int32_t* pairs, int threads, tns_plan** out) {
    TNS_TRY({
        std::vector<std::pair<int32_t, int32_t>> pr;
        for (int i = 0; i < npairs; i++) pr.push_back({pairs[2 * i], pairs[2 * i + 1]});
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            parallel_items(n, threads, [&](int i) {
                auto p = std::make_unique<tns_plan>();
                p->type = tns_plan::CONTRACT;
                plan_contract(a[i]->s, b[i]->s, pr, p->c);
                out[i] = p.release();
            });
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_plan_qr_many(int n, const tns_shape* const* t, char direction, int nfree, const int32_t* free_names,
                     int32_t cq, int32_t cr, int threads, tns_plan** out) {
    TNS_TRY({
        const auto fr = ids(nfree, free_names);
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            parallel_items(n, threads, [&](int i) {
                auto p = std::make_unique<tns_plan>();
                p->type = tns_plan::QR;
                plan_qr(t[i]->s, direction, fr, cq, cr, p->q);
                out[i] = p.release();
            });
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_plan_svd_many(int n, const tns_shape* const* t, int nfree, const int32_t* free_u, int32_t cu, int32_t cv,
                      int32_t su, int32_t sv, int threads, tns_plan** out) {
    TNS_TRY({
        const auto fr = ids(nfree, free_u);
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            parallel_items(n, threads, [&](int i) {
                auto p = std::make_unique<tns_plan>();
                p->type = tns_plan::SVD;
                plan_svd_factor(t[i]->s, fr, cu, cv, su, sv, p->s);
                out[i] = p.release();
            });
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_plan_conjugate_many(int n, const tns_shape* const* t, int trivial_metric, int threads, tns_plan** out) {
    TNS_TRY({
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            parallel_items(n, threads, [&](int i) {
                auto p = std::make_unique<tns_plan>();
                p->type = tns_plan::CONJ;
                plan_conjugate(t[i]->s, trivial_metric != 0, p->eo_out, p->eo_tasks);
                out[i] = p.release();
            });
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_shapes_free(int n, tns_shape* const* shapes) {
    TNS_TRY({
        for (int i = 0; i < n; i++) delete shapes[i];
    })
}

int tns_plans_free(int n, tns_plan* const* plans) {
    TNS_TRY({
        for (int i = 0; i < n; i++) delete plans[i];
    })
}

int tns_shape_rename_many(int n, const tns_shape* const* shapes, int nren, const int32_t* old_names,
                          const int32_t* new_names, tns_shape** out) {
    TNS_TRY({
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            for (int i = 0; i < n; i++) {
                Shape s = shapes[i]->s;
                for (int k = 0; k < nren; k++) {
                    const int f = s.index_of(old_names[k]);
                    if (f < 0) throw Error("Name missing in edge_rename");
                    s.names[f] = new_names[k];
                }
                for (int a = 0; a < s.rank(); a++)
                    for (int b = a + 1; b < s.rank(); b++)
                        if (s.names[a] == s.names[b]) throw Error("duplicated names after edge_rename");
                out[i] = new_handle(s);
            }
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_plans_output_shapes(int n, tns_plan* const* plans, int which, tns_shape** out, int64_t* sizes) {
    TNS_TRY({
        for (int i = 0; i < n; i++) out[i] = nullptr;
        try {
            for (int i = 0; i < n; i++) {
                if (tns_plan_output_shape(plans[i], which, &out[i]) != 0) throw Error(tns_last_error());
                sizes[i] = out[i]->s.size;
            }
        } catch (...) {
            for (int i = 0; i < n; i++) delete out[i];
            throw;
        }
    })
}

int tns_plans_workspace(int n, tns_plan* const* plans, int dtype, int64_t* bytes) {
    TNS_TRY({
        for (int i = 0; i < n; i++) bytes[i] = layout_of(plans[i], dtype).total;
    })
}

int tns_plans_run(int n, tns_plan* const* plans, int dtype, const void* const* inputs, void* const* outputs,
                  void* const* workspaces, void* stream) {
    TNS_TRY({
        if (n <= 0) return 0;
        if (dtype != 0 && dtype != 1) throw Error("unsupported dtype");
        check_same_type(n, plans);
        const int64_t es = dtype == 0 ? 8 : 4;
        const auto type = plans[0]->type;
        std::vector<Layout> L(static_cast<size_t>(n));
        for (int i = 0; i < n; i++) {
            L[size_t(i)] = layout_of(plans[i], dtype);
            if (L[size_t(i)].total > 0 && !workspaces[i]) throw Error("workspace required");
        }
        auto ws = [&](int i) { return static_cast<char*>(workspaces[i]); };
        Batch bt(dtype);
        if (type == tns_plan::EDGEOP || type == tns_plan::CONJ) {
            // permute-only plans: zero what no block copy covers, then copy
            for (int i = 0; i < n; i++) {
                void* b[2] = {const_cast<void*>(inputs[i]), outputs[i]};
                int64_t covered = 0;
                for (auto& t : plans[i]->eo_tasks) covered += t.n;
                if (covered < plans[i]->eo_out.size) bt.fill({FillTask{1, 0, plans[i]->eo_out.size}}, b, es);
            }
            bt.flush_fill();
            for (int i = 0; i < n; i++) {
                void* b[2] = {const_cast<void*>(inputs[i]), outputs[i]};
                bt.perm(plans[i]->eo_tasks, b, es);
            }
            bt.flush_perm();
        } else if (type == tns_plan::CONTRACT) {
            auto bufs = [&](int i, void** b) {
                b[0] = const_cast<void*>(inputs[2 * i]);
                b[1] = const_cast<void*>(inputs[2 * i + 1]);
                b[2] = ws(i) + L[size_t(i)].off[2];
                b[3] = ws(i) + L[size_t(i)].off[3];
                b[4] = ws(i) + L[size_t(i)].off[4];
                b[5] = outputs[i];
            };
            void* b[6];
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.perm(plans[i]->c.perm_a, b, es);
                bt.perm(plans[i]->c.perm_b, b, es);
            }
            bt.flush_perm();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.fill(plans[i]->c.fills, b, es);
            }
            bt.flush_fill();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.gemm(plans[i]->c.gemms, b, es);
            }
            bt.flush_gemm();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.perm(plans[i]->c.perm_out, b, es);
            }
            bt.flush_perm();
        } else if (type == tns_plan::QR) {
            auto bufs = [&](int i, void** b) {
                b[0] = const_cast<void*>(inputs[i]);
                b[1] = ws(i) + L[size_t(i)].off[1];
                b[2] = ws(i) + L[size_t(i)].off[2];
                b[3] = ws(i) + L[size_t(i)].off[3];
                b[4] = outputs[2 * i];
                b[5] = outputs[2 * i + 1];
            };
            void* b[6];
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.perm(plans[i]->q.perm_in, b, es);
            }
            bt.flush_perm();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.qr(plans[i]->q.qrs, b, es, reinterpret_cast<double*>(ws(i) + L[size_t(i)].scratch));
            }
            bt.flush_qr();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.perm(plans[i]->q.perm_1, b, es);
                bt.perm(plans[i]->q.perm_2, b, es);
            }
            bt.flush_perm();
        } else if (type == tns_plan::SVD) {
            auto bufs = [&](int i, void** b) {
                b[0] = nullptr;
                b[1] = ws(i) + L[size_t(i)].off[1];
                b[2] = ws(i) + L[size_t(i)].off[2];
                b[3] = ws(i) + L[size_t(i)].off[3];
                b[4] = ws(i) + L[size_t(i)].off[4];
                b[5] = outputs[3 * i];
                b[6] = outputs[3 * i + 1];
                b[7] = outputs[3 * i + 2];
            };
            void* b[8];
            for (int i = 0; i < n; i++)
                if (!plans[i]->factored) throw Error("call tns_svd_factor_many before tns_plans_run for svd");
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.fill(plans[i]->s.fills_s, b, es);
            }
            bt.flush_fill();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.perm(plans[i]->s.perm_u, b, es);
                bt.perm(plans[i]->s.perm_v, b, es);
            }
            bt.flush_perm();
            for (int i = 0; i < n; i++) {
                bufs(i, b);
                bt.diag(plans[i]->s.diag_s, b, es);
            }
            bt.flush_diag();
        } else {
            throw Error("tns_plans_run: unknown plan type");
        }
        bt.run(as_stream(stream));
    })
}

int tns_svd_factor_many(int n, tns_plan* const* plans, int dtype, const void* const* inputs,
                        void* const* workspaces, int64_t remain_cut, double relative_cut, int threads, void* stream) {
    TNS_TRY({
        if (n <= 0) return 0;
        for (int i = 0; i < n; i++)
            if (plans[i]->type != tns_plan::SVD) throw Error("not an svd plan");
        const int64_t es = dtype == 0 ? 8 : 4;
        std::vector<Layout> L(static_cast<size_t>(n));
        Batch bt(dtype);
        bt.set_fast_svd(true);
        for (int i = 0; i < n; i++) {
            L[size_t(i)] = layout_of(plans[i], dtype);
            char* ws = static_cast<char*>(workspaces[i]);
            void* b[5] = {const_cast<void*>(inputs[i]), ws + L[size_t(i)].off[1], ws + L[size_t(i)].off[2],
                          ws + L[size_t(i)].off[3], ws + L[size_t(i)].off[4]};
            bt.perm(plans[i]->s.perm_in, b, es);
        }
        bt.flush_perm();
        for (int i = 0; i < n; i++) {
            char* ws = static_cast<char*>(workspaces[i]);
            void* b[5] = {const_cast<void*>(inputs[i]), ws + L[size_t(i)].off[1], ws + L[size_t(i)].off[2],
                          ws + L[size_t(i)].off[3], ws + L[size_t(i)].off[4]};
            bt.svd(plans[i]->s.svds, b, es, reinterpret_cast<double*>(ws + L[size_t(i)].scratch));
        }
        bt.flush_svd();
        hipStream_t st = as_stream(stream);
        bt.run(st);
        // every item's singular values in one launch, written straight into
        // host-mapped pinned memory (one per host thread, grown on demand)
        int64_t total = 0;
        std::vector<int64_t> soff(static_cast<size_t>(n));
        for (int i = 0; i < n; i++) {
            soff[size_t(i)] = total;
            total += plans[i]->s.s_total;
        }
        double* host = mapped_host_buffer(size_t(n) * sizeof(GatherDev) + size_t(total) * 8);
        auto* tasks = reinterpret_cast<GatherDev*>(host);
        double* vals = host + size_t(n) * sizeof(GatherDev) / 8;
        for (int i = 0; i < n; i++)
            tasks[i] = GatherDev{reinterpret_cast<const double*>(static_cast<char*>(workspaces[i]) + L[size_t(i)].off[4]),
                                 plans[i]->s.s_total, soff[size_t(i)]};
        GatherDev* dtasks = nullptr;
        double* dvals = nullptr;
        check_hip(hipHostGetDevicePointer(reinterpret_cast<void**>(&dtasks), tasks, 0), "mapped pointer");
        dvals = reinterpret_cast<double*>(dtasks) + size_t(n) * sizeof(GatherDev) / 8;
        launch_gather_f64(dtasks, n, dvals, st);
        check_hip(hipStreamSynchronize(st), "svd synchronize");
        parallel_items(n, threads, [&](int i) {
            auto* p = plans[i];
            p->svals_host.assign(vals + soff[size_t(i)], vals + soff[size_t(i)] + p->s.s_total);
            plan_svd_truncate(p->s, p->svals_host.data(), remain_cut, relative_cut);
            p->factored = true;
        });
    })
}

int tns_svd_factor(tns_plan* p, int dtype, const void* input, void* workspace, int64_t remain_cut,
                   double relative_cut, void* stream) {
    TNS_TRY({
        if (p->type != tns_plan::SVD) throw Error("not an svd plan");
        const int64_t es = dtype == 0 ? 8 : 4;
        Layout L = layout_of(p, dtype);
        char* ws = static_cast<char*>(workspace);
        void* bufs[5] = {const_cast<void*>(input), ws + L.off[1], ws + L.off[2], ws + L.off[3], ws + L.off[4]};
        Batch bt(dtype);
        bt.set_fast_svd(true);
        bt.perm(p->s.perm_in, bufs, es);
        bt.flush_perm();
        bt.svd(p->s.svds, bufs, es, reinterpret_cast<double*>(ws + L.scratch));
        bt.flush_svd();
        hipStream_t st = as_stream(stream);
        bt.run(st);
        p->svals_host.assign(size_t(p->s.s_total), 0.0);
        if (p->s.s_total) {
            check_hip(hipMemcpyAsync(p->svals_host.data(), ws + L.off[4], size_t(p->s.s_total) * 8,
                                     hipMemcpyDeviceToHost, st),
                      "singular values copy");
        }
        check_hip(hipStreamSynchronize(st), "svd synchronize");
        plan_svd_truncate(p->s, p->svals_host.data(), remain_cut, relative_cut);
        p->factored = true;
    })
}

// CPU-side truncation with externally computed singular values (used by the
// planner tests; the GPU path calls the same rule inside tns_svd_factor)
int tns_svd_truncate(tns_plan* p, const double* svals, int64_t remain_cut, double relative_cut) {
    TNS_TRY({
        if (p->type != tns_plan::SVD) throw Error("not an svd plan");
        plan_svd_truncate(p->s, svals, remain_cut, relative_cut);
        p->factored = true;
    })
}

int tns_plan_buffers(const tns_plan* p, int64_t* sizes, int* n) {
    TNS_TRY({
        auto s = buffer_sizes(p);
        *n = int(s.size());
        if (sizes)
            for (size_t i = 0; i < s.size(); i++) sizes[i] = s[i];
    })
}

int tns_svd_sectors(const tns_plan* p, int64_t* k, int64_t* off, int32_t* syms, int* n) {
    TNS_TRY({
        if (p->type != tns_plan::SVD) throw Error("not an svd plan");
        *n = int(p->s.s_sym.size());
        for (size_t i = 0; i < p->s.s_sym.size(); i++) {
            if (k) k[i] = p->s.s_k[i];
            if (off) off[i] = p->s.s_off[i];
            if (syms) {
                syms[2 * i] = p->s.s_sym[i].v[0];
                syms[2 * i + 1] = p->s.s_sym[i].v[1];
            }
        }
    })
}

int tns_plan_stage_count(const tns_plan* p, int stage, int64_t* n) {
    TNS_TRY({
        int64_t c = -1;
        switch (p->type) {
            case tns_plan::EDGEOP:
            case tns_plan::CONJ:
                if (stage == 0) c = int64_t(p->eo_tasks.size());
                break;
            case tns_plan::CONTRACT: {
                const int64_t cs[5] = {int64_t(p->c.perm_a.size()), int64_t(p->c.perm_b.size()),
                                       int64_t(p->c.gemms.size()), int64_t(p->c.fills.size()),
                                       int64_t(p->c.perm_out.size())};
                if (stage >= 0 && stage < 5) c = cs[stage];
                break;
            }
            case tns_plan::SVD: {
                const int64_t cs[6] = {int64_t(p->s.perm_in.size()), int64_t(p->s.svds.size()),
                                       int64_t(p->s.perm_u.size()), int64_t(p->s.perm_v.size()),
                                       int64_t(p->s.fills_s.size()), int64_t(p->s.diag_s.size())};
                if (stage >= 0 && stage < 6) c = cs[stage];
                break;
            }
            case tns_plan::QR: {
                const int64_t cs[4] = {int64_t(p->q.perm_in.size()), int64_t(p->q.qrs.size()),
                                       int64_t(p->q.perm_1.size()), int64_t(p->q.perm_2.size())};
                if (stage >= 0 && stage < 4) c = cs[stage];
                break;
            }
        }
        if (c < 0) throw Error("no such stage");
        *n = c;
    })
}

int tns_plan_stage_export(const tns_plan* p, int stage, int64_t* r) {
    TNS_TRY({
        auto gemms = [&](const std::vector<GemmTask>& ts) {
            for (auto& t : ts) {
                const int64_t v[15] = {t.a_buf, t.b_buf, t.c_buf, t.a,   t.b,  t.c,  t.m,
                                       t.n,     t.k,     t.lda,   t.ldb, t.ldc, t.ta, t.tb, int64_t(t.alpha)};
                std::memcpy(r, v, sizeof(v));
                r += 15;
            }
        };
        auto fills = [&](const std::vector<FillTask>& ts) {
            for (auto& t : ts) {
                r[0] = t.buf;
                r[1] = t.off;
                r[2] = t.n;
                r += 3;
            }
        };
        auto qrs = [&](const std::vector<QrTask>& ts) {
            for (auto& t : ts) {
                const int64_t v[9] = {t.in_buf, t.f1_buf, t.f2_buf, t.in, t.f1, t.f2, t.m, t.n, t.lq};
                std::memcpy(r, v, sizeof(v));
                r += 9;
            }
        };
        auto svds = [&](const std::vector<SvdTask>& ts) {
            for (auto& t : ts) {
                const int64_t v[10] = {t.in_buf, t.u_buf, t.s_buf, t.v_buf, t.in, t.u, t.s, t.v, t.m, t.n};
                std::memcpy(r, v, sizeof(v));
                r += 10;
            }
        };
        switch (p->type) {
            case tns_plan::EDGEOP:
            case tns_plan::CONJ:
                export_perm(p->eo_tasks, r);
                break;
            case tns_plan::CONTRACT:
                if (stage == 0) export_perm(p->c.perm_a, r);
                if (stage == 1) export_perm(p->c.perm_b, r);
                if (stage == 2) gemms(p->c.gemms);
                if (stage == 3) fills(p->c.fills);
                if (stage == 4) export_perm(p->c.perm_out, r);
                break;
            case tns_plan::SVD:
                if (stage == 0) export_perm(p->s.perm_in, r);
                if (stage == 1) svds(p->s.svds);
                if (stage == 2) export_perm(p->s.perm_u, r);
                if (stage == 3) export_perm(p->s.perm_v, r);
                if (stage == 4) fills(p->s.fills_s);
                if (stage == 5) export_perm(p->s.diag_s, r);
                break;
            case tns_plan::QR:
                if (stage == 0) export_perm(p->q.perm_in, r);
                if (stage == 1) qrs(p->q.qrs);
                if (stage == 2) export_perm(p->q.perm_1, r);
                if (stage == 3) export_perm(p->q.perm_2, r);
                break;
        }
    })
}

}  // extern "C"
