// Host-side model of symmetric block-sparse tensors and the operation
// planners (block matching, offsets, fermion signs).  No device code here:
// planners emit task lists that the HIP kernels in kernels.hip execute.
//
// Semantics follow TAT (paths relative to the TNSP repository root):
//   symmetry  TAT/include/TAT/structure/symmetry.hpp:95-239
//   edge      TAT/include/TAT/structure/edge.hpp:33-292
//   layout    TAT/include/TAT/structure/core.hpp:152-191
//   edge op   TAT/include/TAT/implement/edge_operator.hpp:35-692
//   contract  TAT/include/TAT/implement/contract.hpp:307-620
//   svd       TAT/include/TAT/implement/svd.hpp:214-538
//   qr        TAT/include/TAT/implement/qr.hpp:310-508
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tns {

constexpr int MAXC = 2;  // symmetry components
constexpr int MAXR = 8;  // rank of one permute task after squashing

struct Kind {
    int ncomp = 0;
    bool z2[MAXC] = {false, false};
    bool fermi[MAXC] = {false, false};
    bool is_fermi = false;
};

// 0 No, 1 BoseZ2, 2 BoseU1, 3 FermiU1, 4 FermiZ2, 5 FermiU1BoseZ2,
// 6 FermiU1BoseU1, 7 FermiU1FermiU1
const Kind& kind_by_id(int id);

struct Sym {
    int32_t v[MAXC] = {0, 0};
};
inline bool operator==(const Sym& a, const Sym& b) { return a.v[0] == b.v[0] && a.v[1] == b.v[1]; }
inline bool operator!=(const Sym& a, const Sym& b) { return !(a == b); }
inline bool sym_is_zero(const Sym& a) { return a.v[0] == 0 && a.v[1] == 0; }
inline Sym sym_add(const Kind& k, const Sym& a, const Sym& b) {
    Sym r;
    for (int c = 0; c < k.ncomp; c++) r.v[c] = k.z2[c] ? (a.v[c] ^ b.v[c]) : (a.v[c] + b.v[c]);
    return r;
}
inline Sym sym_neg(const Kind& k, const Sym& a) {
    Sym r;
    for (int c = 0; c < k.ncomp; c++) r.v[c] = k.z2[c] ? a.v[c] : -a.v[c];
    return r;
}
inline bool sym_parity(const Kind& k, const Sym& a) {
    bool p = false;
    for (int c = 0; c < k.ncomp; c++)
        if (k.fermi[c]) p ^= (a.v[c] % 2) != 0;
    return p;
}

struct Segment {
    Sym q;
    int64_t d = 0;
};

struct Edge {
    std::vector<Segment> segs;
    bool arrow = false;
    int find(const Sym& q) const {
        for (size_t i = 0; i < segs.size(); i++)
            if (segs[i].q == q) return int(i);
        return -1;
    }
    int64_t total() const {
        int64_t t = 0;
        for (auto& s : segs) t += s.d;
        return t;
    }
};
inline bool same_edge(const Edge& a, const Edge& b) {
    if (a.arrow != b.arrow || a.segs.size() != b.segs.size()) return false;
    for (size_t i = 0; i < a.segs.size(); i++)
        if (a.segs[i].q != b.segs[i].q || a.segs[i].d != b.segs[i].d) return false;
    return true;
}
Edge edge_conjugate(const Kind& k, const Edge& e);

struct Shape {
    int kind_id = 0;
    std::vector<int32_t> names;
    std::vector<Edge> edges;
    // block table over segment-position tuples, row-major (core.hpp:152-191)
    std::vector<int64_t> pos_lead;
    std::vector<int64_t> block_off;   // -1 for forbidden tuples
    std::vector<int64_t> blocks;      // linear tuple indices of allowed blocks, storage order
    int64_t size = 0;

    const Kind& kind() const { return kind_by_id(kind_id); }
    int rank() const { return int(names.size()); }
    int index_of(int32_t name) const {
        for (size_t i = 0; i < names.size(); i++)
            if (names[i] == name) return int(i);
        return -1;
    }
    void finalize();  // build the block table from edges
    int64_t offset(const int64_t* positions) const {
        int64_t lin = 0;
        for (int i = 0; i < rank(); i++) lin += positions[i] * pos_lead[i];
        return block_off.empty() ? -1 : block_off[lin];
    }
    void positions_of(int64_t lin, int64_t* out) const {
        for (int i = 0; i < rank(); i++) {
            out[i] = pos_lead[i] ? (lin / pos_lead[i]) % int64_t(edges[i].segs.size()) : 0;
        }
    }
};

// ---------------------------------------------------------------- tasks
// Buffer ids let one plan describe several device buffers; a bind step turns
// (buffer, offset) pairs into device pointers.
struct PermTask {
    int32_t src_buf = 0, dst_buf = 0;
    int64_t src = 0, dst = 0;
    int32_t rank = 0;
    int32_t neg = 0;
    int64_t n = 0;
    int32_t dims[MAXR] = {};
    int32_t ss[MAXR] = {};
    int32_t ds[MAXR] = {};
};

struct GemmTask {
    int32_t a_buf = 0, b_buf = 0, c_buf = 0;
    int64_t a = 0, b = 0, c = 0;
    int32_t m = 0, n = 0, k = 0;
    int32_t lda = 0, ldb = 0, ldc = 0;
    int32_t ta = 0, tb = 0;  // C = alpha op(A) op(B); op(A)(i,p) = ta ? A[p*lda+i] : A[i*lda+p]
    double alpha = 1.0;
};

struct FillTask {
    int32_t buf = 0;
    int64_t off = 0, n = 0;
};

// M (m x n, row-major, ld n) -> Q (m x k) . R (k x n), k = min(m, n).
// lq: M = L (m x k) . Q (k x n) with orthonormal rows; stored: first factor at
// f1 (m x k), second at f2 (k x n) in both cases.
struct QrTask {
    int32_t in_buf = 0, f1_buf = 0, f2_buf = 0;
    int64_t in = 0, f1 = 0, f2 = 0;
    int32_t m = 0, n = 0, lq = 0;
};

// M (m x n) = U (m x k) diag(s) V (k x n), k = min(m, n), s descending.
struct SvdTask {
    int32_t in_buf = 0, u_buf = 0, s_buf = 0, v_buf = 0;
    int64_t in = 0, u = 0, s = 0, v = 0;
    int32_t m = 0, n = 0;
};

// ------------------------------------------------------------- edge op
struct SplitSpec {
    int32_t old_name;
    std::vector<std::pair<int32_t, Edge>> parts;  // (new name, segments)
};
struct MergeSpec {
    int32_t new_name;
    std::vector<int32_t> olds;
};
struct CutSpec {
    int32_t name;
    std::vector<std::pair<Sym, int64_t>> dims;
};
struct EdgeOpArgs {
    std::vector<SplitSpec> split;
    std::vector<int32_t> reversed;
    std::vector<MergeSpec> merge;
    std::vector<int32_t> new_names;
    bool apply_parity = false;
    std::vector<int32_t> ex_split, ex_rev_before, ex_rev_after, ex_merge;
    std::vector<CutSpec> cut;
};

// Plans edge_operator (split -> reverse -> transpose -> reverse -> merge, with
// the SVD cut first).  Tasks copy from buffer src_buf to dst_buf.
void plan_edge_operator(const Shape& src, const EdgeOpArgs& args, Shape& out, std::vector<PermTask>& tasks,
                        int32_t src_buf, int32_t dst_buf);

// ------------------------------------------------------------ contract
// Buffers: 0 = A, 1 = B, 2 = merged A, 3 = merged B, 4 = product, 5 = result
struct ContractPlan {
    Shape ma, mb, prod, result;
    std::vector<PermTask> perm_a, perm_b, perm_out;
    std::vector<GemmTask> gemms;
    std::vector<FillTask> fills;
    int64_t flops = 0;
};
void plan_contract(const Shape& a, const Shape& b, const std::vector<std::pair<int32_t, int32_t>>& pairs, ContractPlan& p);

// ----------------------------------------------------------------- svd
// Buffers: 0 = input, 1 = merged, 2 = t1 (full), 3 = t2 (full), 4 = S values,
//          5 = U out, 6 = S out, 7 = V out
struct SvdPlan {
    int kind_id = 0;
    bool put_v_right = true;
    Shape merged, t1, t2;
    std::vector<PermTask> perm_in;
    std::vector<SvdTask> svds;
    // per sector of result_s: symmetry key, k, offset in S values buffer
    std::vector<Sym> s_sym;
    std::vector<int64_t> s_k, s_off;
    int64_t s_total = 0;
    // second phase
    int32_t cu = 0, cv = 0, su = 0, sv = 0;
    EdgeOpArgs args_u, args_v;
    std::vector<int64_t> remain;  // kept per sector
    Shape u, s, v;
    std::vector<PermTask> perm_u, perm_v;
    std::vector<FillTask> fills_s;
    std::vector<PermTask> diag_s;  // copies of kept singular values into S (strided)
    int64_t flops = 0;
};
void plan_svd_factor(const Shape& t, const std::vector<int32_t>& free_u, int32_t cu, int32_t cv, int32_t su, int32_t sv,
                     SvdPlan& p);
// remain_cut < 0: no count limit; relative_cut <= 0: none
void plan_svd_truncate(SvdPlan& p, const double* svals, int64_t remain_cut, double relative_cut);

// ------------------------------------------------------------------ qr
// Buffers: 0 = input, 1 = merged, 2 = t1, 3 = t2, 4 = Q out, 5 = R out
struct QrPlan {
    bool use_qr = true;
    Shape merged, t1, t2, q, r;
    std::vector<PermTask> perm_in, perm_1, perm_2;  // perm_1: t1 -> (use_qr ? Q : R)
    std::vector<QrTask> qrs;
    int64_t flops = 0;
};
void plan_qr(const Shape& t, char direction, const std::vector<int32_t>& free_names, int32_t cq, int32_t cr, QrPlan& p);

// ----------------------------------------------------------- conjugate
// conjugate.hpp:32-120: charges negated, arrows flipped, per-block sign.
void plan_conjugate(const Shape& t, bool trivial_metric, Shape& out, std::vector<PermTask>& tasks);

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace tns
