// Host planners for the block-sparse tensor operations.  See core.hpp for the
// reference files each planner follows.
#include "core.hpp"

#include <algorithm>
#include <cmath>

namespace tns {

namespace {
Kind make_kind(std::initializer_list<std::pair<bool, bool>> comps) {
    Kind k;
    for (auto& c : comps) {
        k.z2[k.ncomp] = c.first;
        k.fermi[k.ncomp] = c.second;
        k.is_fermi |= c.second;
        k.ncomp++;
    }
    return k;
}
const Kind KINDS[8] = {
    make_kind({}),
    make_kind({{true, false}}),
    make_kind({{false, false}}),
    make_kind({{false, true}}),
    make_kind({{true, true}}),
    make_kind({{false, true}, {true, false}}),
    make_kind({{false, true}, {false, false}}),
    make_kind({{false, true}, {false, true}}),
};

template<typename T>
bool contains(const std::vector<T>& v, const T& x) {
    return std::find(v.begin(), v.end(), x) != v.end();
}

// Squash a strided copy like mdspan_transform (multidimension_span.hpp:298-383):
// drop dimension-one axes and fuse axes contiguous in both source and
// destination.  dims/src/dst are in destination order.
void squash_into(PermTask& t, int rank, const int64_t* dims, const int64_t* sst, const int64_t* dst) {
    int64_t d[64 + MAXR], s[64 + MAXR], o[64 + MAXR];
    int r = 0;
    for (int i = 0; i < rank; i++) {
        if (dims[i] == 1) continue;
        if (r > 0 && o[r - 1] == dims[i] * dst[i] && s[r - 1] == dims[i] * sst[i]) {
            d[r - 1] *= dims[i];
            o[r - 1] = dst[i];
            s[r - 1] = sst[i];
            continue;
        }
        d[r] = dims[i];
        s[r] = sst[i];
        o[r] = dst[i];
        r++;
        if (r > 64) throw Error("permute rank too large");
    }
    if (r == 0) {
        d[0] = 1;
        s[0] = 1;
        o[0] = 1;
        r = 1;
    }
    if (r > MAXR) throw Error("permute task exceeds supported rank " + std::to_string(MAXR));
    t.rank = r;
    t.n = 1;
    for (int i = 0; i < r; i++) {
        t.dims[i] = int32_t(d[i]);
        t.ss[i] = int32_t(s[i]);
        t.ds[i] = int32_t(o[i]);
        t.n *= d[i];
    }
}

// a non-owning view of an edge's segments with its own arrow flag
struct EdgeView {
    const Segment* s = nullptr;
    int n = 0;
    bool arrow = false;
    EdgeView() = default;
    explicit EdgeView(const Edge& e) : s(e.segs.data()), n(int(e.segs.size())), arrow(e.arrow) {}
    int size() const { return n; }
    const Segment& seg(int64_t i) const { return s[i]; }
    int find(const Sym& q) const {
        for (int i = 0; i < n; i++)
            if (s[i].q == q) return i;
        return -1;
    }
    Edge to_edge() const {
        Edge e;
        e.segs.assign(s, s + n);
        e.arrow = arrow;
        return e;
    }
};

struct GroupTable {
    // over sub-edge position tuples (row-major): target position, offset, odd count
    std::vector<int64_t> lead;
    std::vector<int32_t> pos;
    std::vector<int64_t> off;
    std::vector<int32_t> npar;
};
}  // namespace

const Kind& kind_by_id(int id) {
    if (id < 0 || id > 7) throw Error("unknown symmetry kind");
    return KINDS[id];
}

Edge edge_conjugate(const Kind& k, const Edge& e) {
    Edge r;
    r.segs.reserve(e.segs.size());
    for (auto& s : e.segs) r.segs.push_back({sym_neg(k, s.q), s.d});
    r.arrow = k.is_fermi ? !e.arrow : false;
    return r;
}

void Shape::finalize() {
    const Kind& k = kind();
    if (!k.is_fermi)
        for (auto& e : edges) e.arrow = false;
    const int r = rank();
    pos_lead.assign(r, 0);
    int64_t total = 1;
    for (int i = r - 1; i >= 0; i--) {
        pos_lead[i] = total;
        total *= int64_t(edges[i].segs.size());
    }
    block_off.assign(total, -1);
    blocks.clear();
    size = 0;
    if (total == 0) return;
    // odometer over segment positions with running charge sums
    std::vector<int64_t> idx(r, 0);
    std::vector<Sym> partial(r + 1);
    for (int i = 0; i < r; i++) partial[i + 1] = sym_add(k, partial[i], edges[i].segs[0].q);
    for (int64_t lin = 0; lin < total; lin++) {
        if (sym_is_zero(partial[r])) {
            int64_t sz = 1;
            for (int i = 0; i < r; i++) sz *= edges[i].segs[idx[i]].d;
            block_off[lin] = size;
            blocks.push_back(lin);
            size += sz;
        }
        int i = r - 1;
        while (i >= 0) {
            if (++idx[i] < int64_t(edges[i].segs.size())) break;
            idx[i] = 0;
            i--;
        }
        if (i < 0) break;
        for (int j = i; j < r; j++) partial[j + 1] = sym_add(k, partial[j], edges[j].segs[idx[j]].q);
    }
}

// ------------------------------------------------------------------------
// edge_operator (edge_operator.hpp:35-692)
// ------------------------------------------------------------------------
namespace {
struct EdgeOpScratch {
    std::vector<EdgeView> edges_bs, edges_as, edges_bt, edges_bm;
    std::vector<int32_t> names_as, names_bm;
    std::vector<int> split_flags, merge_flags, d2s, group_start, active, odd_list;
    std::vector<GroupTable> split_tab, merge_tab;
    std::vector<bool> is_split, am_is_merge, group_merged;
    std::vector<char> rev_before, rev_after, mk_split, mk_rb, mk_ra, mk_merge, rev_flag;
    std::vector<int64_t> pos_bm, pos_as, dims_bm, dims_as, pos_am, off_am, pos_bs, off_bs, pos_src;
    std::vector<int64_t> lead_am, lead_bm, lead_bs, lead_as, tdims, tss, tds, aidx;
};
EdgeOpScratch& edge_op_scratch() {
    thread_local EdgeOpScratch s;
    return s;
}
}  // namespace

void plan_edge_operator(const Shape& src, const EdgeOpArgs& args, Shape& out, std::vector<PermTask>& tasks,
                        int32_t src_buf, int32_t dst_buf) {
    const Kind& k = src.kind();
    const bool fermi = k.is_fermi;
    const int rank_bs = src.rank();
    // per-thread working vectors, reused from call to call (the planner runs
    // thousands of these per configuration; their allocations dominated)
    EdgeOpScratch& S = edge_op_scratch();

    // edges are viewed (segments + arrow), not copied, until the result
    // shape is built
    // step 1: cut
    std::vector<Edge> cut_store;
    cut_store.reserve(args.cut.size());
    auto& edges_bs = S.edges_bs;
    edges_bs.assign(rank_bs, EdgeView());
    for (int i = 0; i < rank_bs; i++) edges_bs[i] = EdgeView(src.edges[i]);
    bool has_cut = false;
    for (auto& c : args.cut) {
        int i = src.index_of(c.name);
        if (i < 0) continue;
        has_cut = true;
        Edge ne;
        ne.arrow = edges_bs[i].arrow;
        for (int si = 0; si < edges_bs[i].size(); si++) {
            const Segment& sg = edges_bs[i].seg(si);
            bool found = false;
            for (auto& [q, nd] : c.dims) {
                if (q == sg.q) {
                    found = true;
                    if (nd != 0) ne.segs.push_back({sg.q, std::min(nd, sg.d)});
                    break;
                }
            }
            if (!found) ne.segs.push_back(sg);
        }
        cut_store.push_back(std::move(ne));
        edges_bs[i] = EdgeView(cut_store.back());
    }

    // step 2: split
    auto& names_as = S.names_as;
    auto& split_flags = S.split_flags;
    auto& edges_as = S.edges_as;
    auto& split_tab = S.split_tab;  // an entry is rebuilt before it is read
    auto& is_split = S.is_split;
    names_as.clear();
    split_flags.clear();
    edges_as.clear();
    if (split_tab.size() < size_t(rank_bs)) split_tab.resize(size_t(rank_bs));
    is_split.assign(rank_bs, false);
    for (int i = 0; i < rank_bs; i++) {
        const SplitSpec* sp = nullptr;
        for (auto& s : args.split)
            if (s.old_name == src.names[i]) sp = &s;
        if (!sp) {
            names_as.push_back(src.names[i]);
            split_flags.push_back(i);
            edges_as.push_back(edges_bs[i]);
            continue;
        }
        is_split[i] = true;
        const int first = int(edges_as.size());
        for (auto& [nm, e] : sp->parts) {
            names_as.push_back(nm);
            split_flags.push_back(i);
            EdgeView ne(e);
            ne.arrow = edges_bs[i].arrow;
            edges_as.push_back(ne);
        }
        const int nsub = int(sp->parts.size());
        GroupTable& g = split_tab[i];
        g.lead.assign(nsub, 0);
        int64_t total = 1;
        for (int s = nsub - 1; s >= 0; s--) {
            g.lead[s] = total;
            total *= int64_t(edges_as[first + s].size());
        }
        g.pos.assign(total, -1);
        g.off.assign(total, 0);
        g.npar.assign(total, 0);
        std::vector<int64_t> bank(edges_bs[i].size(), 0);
        std::vector<int64_t> idx(nsub, 0);
        for (int64_t lin = 0; lin < total; lin++) {
            Sym tot;
            int64_t sz = 1;
            int np = 0;
            for (int s = 0; s < nsub; s++) {
                const Segment& seg = edges_as[first + s].seg(idx[s]);
                tot = sym_add(k, tot, seg.q);
                sz *= seg.d;
                np += sym_parity(k, seg.q);
            }
            int p = edges_bs[i].find(tot);
            if (p >= 0) {
                g.pos[lin] = p;
                g.off[lin] = bank[p];
                g.npar[lin] = np;
                bank[p] += sz;
            }
            for (int s = nsub - 1; s >= 0; s--) {
                if (++idx[s] < int64_t(edges_as[first + s].size())) break;
                idx[s] = 0;
            }
        }
    }
    const int rank_t = int(names_as.size());

    // step 3: reverse before transpose
    auto& rev_before = S.rev_before;
    rev_before.assign(rank_t, 0);
    auto& edges_bt = S.edges_bt;
    edges_bt.assign(edges_as.begin(), edges_as.end());
    if (fermi) {
        for (int i = 0; i < rank_t; i++) {
            if (contains(args.reversed, names_as[i])) {
                rev_before[i] = 1;
                edges_bt[i].arrow = !edges_bt[i].arrow;
            }
        }
    }

    // step 4: names down to up
    const auto& names_am = args.new_names;
    const int rank_am = int(names_am.size());
    auto& names_bm = S.names_bm;
    auto& merge_flags = S.merge_flags;
    auto& am_is_merge = S.am_is_merge;
    names_bm.clear();
    merge_flags.clear();
    am_is_merge.assign(rank_am, false);
    for (int j = 0; j < rank_am; j++) {
        const MergeSpec* ms = nullptr;
        for (auto& m : args.merge)
            if (m.new_name == names_am[j]) ms = &m;
        if (ms) {
            am_is_merge[j] = true;
            for (auto o : ms->olds) {
                names_bm.push_back(o);
                merge_flags.push_back(j);
            }
        } else {
            names_bm.push_back(names_am[j]);
            merge_flags.push_back(j);
        }
    }
    if (int(names_bm.size()) != rank_t) throw Error("Tensor to transpose with Different Rank");
    auto& d2s = S.d2s;
    d2s.assign(rank_t, 0);
    for (int i = 0; i < rank_t; i++) {
        int f = -1;
        for (int s = 0; s < rank_t; s++)
            if (names_as[s] == names_bm[i]) f = s;
        if (f < 0) throw Error("Tensor to transpose with incompatible name list");
        d2s[i] = f;
    }

    // step 5: reverse before merge, merge tables, result edges
    auto& edges_bm = S.edges_bm;
    edges_bm.assign(rank_t, EdgeView());
    for (int i = 0; i < rank_t; i++) edges_bm[i] = edges_bt[d2s[i]];
    auto& rev_after = S.rev_after;
    rev_after.assign(rank_t, 0);
    const bool any_merge = !args.merge.empty();
    out = Shape();
    out.kind_id = src.kind_id;
    out.names = names_am;
    out.edges.resize(rank_am);
    auto& merge_tab = S.merge_tab;  // an entry is rebuilt before it is read
    if (merge_tab.size() < size_t(rank_am)) merge_tab.resize(size_t(rank_am));
    auto& group_merged = S.group_merged;
    group_merged.assign(rank_am, false);
    auto& group_start = S.group_start;
    group_start.assign(rank_am + 1, 0);
    {
        int start = 0;
        for (int j = 0; j < rank_am; j++) {
            int end = start;
            while (end < rank_t && merge_flags[end] == j) end++;
            group_start[j] = start;
            bool arrow = false;
            if (any_merge && fermi) {
                arrow = (start == end) ? false : edges_bm[start].arrow;
                for (int i = start; i < end; i++) {
                    if (edges_bm[i].arrow != arrow) {
                        edges_bm[i].arrow = !edges_bm[i].arrow;
                        rev_after[i] = 1;
                    }
                }
            }
            if (any_merge && end != start + 1) {
                group_merged[j] = true;
                const int nsub = end - start;
                GroupTable& g = merge_tab[j];
                g.lead.assign(nsub, 0);
                int64_t total = 1;
                for (int s = nsub - 1; s >= 0; s--) {
                    g.lead[s] = total;
                    total *= int64_t(edges_bm[start + s].size());
                }
                g.pos.assign(total, -1);
                g.off.assign(total, 0);
                g.npar.assign(total, 0);
                Edge merged;
                merged.arrow = arrow;
                std::vector<int64_t> idx(nsub, 0);
                for (int64_t lin = 0; lin < total; lin++) {
                    Sym tot;
                    int64_t sz = 1;
                    int np = 0;
                    for (int s = 0; s < nsub; s++) {
                        const Segment& seg = edges_bm[start + s].seg(idx[s]);
                        tot = sym_add(k, tot, seg.q);
                        sz *= seg.d;
                        np += sym_parity(k, seg.q);
                    }
                    int p = merged.find(tot);
                    if (p < 0) {
                        p = int(merged.segs.size());
                        merged.segs.push_back({tot, 0});
                    }
                    g.pos[lin] = p;
                    g.off[lin] = merged.segs[p].d;
                    g.npar[lin] = np;
                    merged.segs[p].d += sz;
                    for (int s = nsub - 1; s >= 0; s--) {
                        if (++idx[s] < int64_t(edges_bm[start + s].size())) break;
                        idx[s] = 0;
                    }
                }
                out.edges[j] = std::move(merged);
            } else {
                if (end == start) throw Error("internal: empty group without merge");
                out.edges[j] = edges_bm[start].to_edge();
            }
            start = end;
        }
        group_start[rank_am] = start;
    }
    out.finalize();

    // marks: true => apply that half sign
    auto &mk_split = S.mk_split, &mk_rb = S.mk_rb, &mk_ra = S.mk_ra, &mk_merge = S.mk_merge;
    mk_split.assign(rank_bs, 0);
    mk_rb.assign(rank_t, 0);
    mk_ra.assign(rank_t, 0);
    mk_merge.assign(rank_am, 0);
    for (int i = 0; i < rank_bs; i++) mk_split[i] = contains(args.ex_split, src.names[i]) != args.apply_parity;
    for (int i = 0; i < rank_t; i++) mk_rb[i] = contains(args.ex_rev_before, names_as[i]) != args.apply_parity;
    for (int i = 0; i < rank_t; i++) mk_ra[i] = contains(args.ex_rev_after, names_bm[i]) != args.apply_parity;
    for (int j = 0; j < rank_am; j++) mk_merge[j] = contains(args.ex_merge, names_am[j]) != args.apply_parity;
    auto& rev_flag = S.rev_flag;
    rev_flag.assign(rank_t, 0);
    for (int i = 0; i < rank_t; i++)
        rev_flag[i] = char((rev_after[i] && mk_ra[i]) ^ (rev_before[d2s[i]] && mk_rb[d2s[i]]));

    // main loop over before-merge segment tuples with zero total charge
    auto& active = S.active;
    active.clear();
    Sym const_sum;
    bool const_par = false;
    for (int i = 0; i < rank_t; i++) {
        if (edges_bm[i].size() == 0) return;  // no blocks at all
        if (edges_bm[i].size() == 1) {
            const_sum = sym_add(k, const_sum, edges_bm[i].seg(0).q);
            if (fermi && rev_flag[i] && sym_parity(k, edges_bm[i].seg(0).q)) const_par = !const_par;
        } else {
            active.push_back(i);
        }
    }
    auto &pos_bm = S.pos_bm, &pos_as = S.pos_as, &dims_bm = S.dims_bm, &dims_as = S.dims_as;
    auto &pos_am = S.pos_am, &off_am = S.off_am, &pos_bs = S.pos_bs, &off_bs = S.off_bs, &pos_src = S.pos_src;
    auto &lead_am = S.lead_am, &lead_bm = S.lead_bm, &lead_bs = S.lead_bs, &lead_as = S.lead_as;
    auto &tdims = S.tdims, &tss = S.tss, &tds = S.tds;
    for (auto* v : {&pos_bm, &pos_as, &dims_bm, &dims_as, &lead_bm, &lead_as, &tdims, &tss, &tds}) v->assign(rank_t, 0);
    for (auto* v : {&pos_am, &off_am, &lead_am}) v->assign(rank_am, 0);
    for (auto* v : {&pos_bs, &off_bs, &pos_src, &lead_bs}) v->assign(rank_bs, 0);
    auto& odd_list = S.odd_list;
    odd_list.clear();
    odd_list.reserve(rank_t);
    const int na = int(active.size());
    auto& aidx = S.aidx;
    aidx.assign(na, 0);
    // the odometer runs over all active edges but the last; the last edge's
    // segment is the one whose charge closes the total (charges are unique
    // within an edge), so forbidden tuples are never visited
    const int nfree = na > 0 ? na - 1 : 0;
    while (true) {
        Sym tot = const_sum;
        for (int a = 0; a < nfree; a++) tot = sym_add(k, tot, edges_bm[active[a]].seg(aidx[a]).q);
        bool allowed = true;
        if (na > 0) {
            const int pl = edges_bm[active[na - 1]].find(sym_neg(k, tot));
            if (pl < 0) {
                allowed = false;
            } else {
                aidx[na - 1] = pl;
                tot = sym_add(k, tot, edges_bm[active[na - 1]].seg(pl).q);
            }
        }
        if (allowed && sym_is_zero(tot)) {
            for (int a = 0; a < na; a++) pos_bm[active[a]] = aidx[a];
            bool par = const_par;
            if (fermi) {
                for (int a = 0; a < na; a++) {
                    int i = active[a];
                    if (rev_flag[i] && sym_parity(k, edges_bm[i].seg(pos_bm[i]).q)) par = !par;
                }
                odd_list.clear();
                for (int i = 0; i < rank_t; i++)
                    if (sym_parity(k, edges_bm[i].seg(pos_bm[i]).q)) odd_list.push_back(i);
                for (size_t x = 0; x < odd_list.size(); x++)
                    for (size_t y = x + 1; y < odd_list.size(); y++)
                        if (d2s[odd_list[x]] > d2s[odd_list[y]]) par = !par;
            }
            bool empty = false;
            for (int i = 0; i < rank_t; i++) {
                const int64_t p = pos_bm[i];
                pos_as[d2s[i]] = p;
                const int64_t d = edges_bm[i].seg(p).d;
                dims_bm[i] = d;
                dims_as[d2s[i]] = d;
                if (d == 0) empty = true;
            }
            bool skip = false;
            for (int j = 0; j < rank_am; j++) {
                const int b = group_start[j], e = group_start[j + 1];
                if (group_merged[j]) {
                    const GroupTable& g = merge_tab[j];
                    int64_t lin = 0;
                    for (int s = 0; s < e - b; s++) lin += pos_bm[b + s] * g.lead[s];
                    pos_am[j] = g.pos[lin];
                    off_am[j] = g.off[lin];
                    if ((g.npar[lin] & 2) != 0 && mk_merge[j]) par = !par;
                } else {
                    pos_am[j] = pos_bm[b];
                    off_am[j] = 0;
                }
            }
            {
                int i = 0;
                for (int j = 0; j < rank_bs; j++) {
                    const int b = i;
                    while (i < rank_t && split_flags[i] == j) i++;
                    if (is_split[j]) {
                        const GroupTable& g = split_tab[j];
                        int64_t lin = 0;
                        for (int s = 0; s < i - b; s++) lin += pos_as[b + s] * g.lead[s];
                        if (g.pos[lin] < 0) {
                            skip = true;
                            break;
                        }
                        pos_bs[j] = g.pos[lin];
                        off_bs[j] = g.off[lin];
                        if ((g.npar[lin] & 2) != 0 && mk_split[j]) par = !par;
                    } else {
                        pos_bs[j] = pos_as[b];
                        off_bs[j] = 0;
                    }
                }
            }
            if (!skip && !empty) {
                if (has_cut) {
                    for (int j = 0; j < rank_bs; j++) pos_src[j] = src.edges[j].find(edges_bs[j].seg(pos_bs[j]).q);
                } else {
                    for (int j = 0; j < rank_bs; j++) pos_src[j] = pos_bs[j];
                }
                for (int j = rank_am - 1; j >= 0; j--)
                    lead_am[j] = (j == rank_am - 1) ? 1 : lead_am[j + 1] * out.edges[j + 1].segs[pos_am[j + 1]].d;
                for (int i = rank_t - 1; i >= 0; i--)
                    lead_bm[i] = (i != rank_t - 1 && merge_flags[i] == merge_flags[i + 1]) ? lead_bm[i + 1] * dims_bm[i + 1]
                                                                                          : lead_am[merge_flags[i]];
                for (int j = rank_bs - 1; j >= 0; j--)
                    lead_bs[j] = (j == rank_bs - 1) ? 1 : lead_bs[j + 1] * src.edges[j + 1].segs[pos_src[j + 1]].d;
                for (int i = rank_t - 1; i >= 0; i--)
                    lead_as[i] = (i != rank_t - 1 && split_flags[i] == split_flags[i + 1]) ? lead_as[i + 1] * dims_as[i + 1]
                                                                                          : lead_bs[split_flags[i]];
                int64_t src_off = src.offset(pos_src.data());
                int64_t dst_off = out.offset(pos_am.data());
                if (src_off < 0 || dst_off < 0) throw Error("internal: block lookup failed in edge_operator");
                for (int j = 0; j < rank_bs; j++) src_off += off_bs[j] * lead_bs[j];
                for (int j = 0; j < rank_am; j++) dst_off += off_am[j] * lead_am[j];
                for (int i = 0; i < rank_t; i++) {
                    tdims[i] = dims_bm[i];
                    tss[i] = lead_as[d2s[i]];
                    tds[i] = lead_bm[i];
                }
                PermTask t;
                t.src_buf = src_buf;
                t.dst_buf = dst_buf;
                t.src = src_off;
                t.dst = dst_off;
                t.neg = par ? 1 : 0;
                squash_into(t, rank_t, tdims.data(), tss.data(), tds.data());
                tasks.push_back(t);
            }
        }
        // advance the odometer over the free active edges
        int a = nfree - 1;
        while (a >= 0) {
            if (++aidx[a] < int64_t(edges_bm[active[a]].size())) break;
            aidx[a] = 0;
            a--;
        }
        if (a < 0) break;
    }
}

// ------------------------------------------------------------------------
// contract (contract.hpp:307-620)
// ------------------------------------------------------------------------
namespace {
constexpr int32_t NAME_C1 = -101, NAME_C2 = -102, NAME_SVD_U = -103, NAME_SVD_V = -104, NAME_QR_1 = -105,
                  NAME_QR_2 = -106;
}

void plan_contract(const Shape& a, const Shape& b, const std::vector<std::pair<int32_t, int32_t>>& pairs,
                   ContractPlan& p) {
    const Kind& k = a.kind();
    const bool fermi = k.is_fermi;
    if (a.kind_id != b.kind_id) throw Error("contract of tensors with different symmetry");
    auto in12 = [&](int32_t n) {
        for (auto& pr : pairs)
            if (pr.first == n) return true;
        return false;
    };
    auto in21 = [&](int32_t n) {
        for (auto& pr : pairs)
            if (pr.second == n) return true;
        return false;
    };
    for (auto& [x, y] : pairs) {
        int ia = a.index_of(x), ib = b.index_of(y);
        if (ia < 0 || ib < 0) throw Error("Name missing in contract");
        if (!same_edge(edge_conjugate(k, a.edges[ia]), b.edges[ib])) throw Error("Incompatible edge segments in contract");
    }
    for (size_t i = 0; i < pairs.size(); i++)
        for (size_t j = i + 1; j < pairs.size(); j++)
            if (pairs[i].first == pairs[j].first || pairs[i].second == pairs[j].second)
                throw Error("Duplicated name in contract_names");
    for (auto n1 : a.names)
        for (auto n2 : b.names)
            if (n1 == n2 && !in12(n1) && !in21(n2))
                throw Error("Duplicated name in two contracting tensor but not fused or contracted");

    std::vector<int32_t> free1, free2, common1, common2, rev1, rev2, rev_res, common_rev1, names_res;
    SplitSpec split1{NAME_C1, {}}, split2{NAME_C2, {}};
    for (int i = 0; i < a.rank(); i++) {
        const int32_t n = a.names[i];
        if (!in12(n)) {
            free1.push_back(n);
            Edge e = a.edges[i];
            e.arrow = false;
            split1.parts.push_back({n, e});
            names_res.push_back(n);
            if (fermi && a.edges[i].arrow) {
                rev1.push_back(n);
                rev_res.push_back(n);
            }
        } else if (fermi && !a.edges[i].arrow) {
            rev1.push_back(n);
            common_rev1.push_back(n);
        }
    }
    for (int i = 0; i < b.rank(); i++) {
        const int32_t n = b.names[i];
        if (!in21(n)) {
            free2.push_back(n);
            Edge e = b.edges[i];
            e.arrow = false;
            split2.parts.push_back({n, e});
            names_res.push_back(n);
            if (fermi && b.edges[i].arrow) {
                rev2.push_back(n);
                rev_res.push_back(n);
            }
        } else if (fermi && b.edges[i].arrow) {
            rev2.push_back(n);
        }
    }
    auto fit1 = [&]() {
        for (auto n : a.names)
            for (auto& pr : pairs)
                if (pr.first == n) {
                    common1.push_back(pr.first);
                    common2.push_back(pr.second);
                }
    };
    auto fit2 = [&]() {
        for (auto n : b.names)
            for (auto& pr : pairs)
                if (pr.second == n) {
                    common2.push_back(pr.second);
                    common1.push_back(pr.first);
                }
    };
    bool r1, r2;
    const bool larger1 = a.size > b.size;
    if (free1.empty()) {
        r1 = true;
        fit2();
        r2 = common2.empty() || common2.back() == b.names.back();
    } else if (free2.empty()) {
        r2 = true;
        fit1();
        r1 = common1.empty() || common1.back() == a.names.back();
    } else if (larger1) {
        if (free1.back() != a.names.back()) {
            r1 = true;
            fit1();
            r2 = common2.empty() || common2.back() == b.names.back();
        } else if (free2.back() != b.names.back()) {
            r2 = true;
            fit2();
            r1 = common1.empty() || common1.back() == a.names.back();
        } else {
            r1 = r2 = false;
            fit1();
        }
    } else {
        if (free2.back() != b.names.back()) {
            r2 = true;
            fit2();
            r1 = common1.empty() || common1.back() == a.names.back();
        } else if (free1.back() != a.names.back()) {
            r1 = true;
            fit1();
            r2 = common2.empty() || common2.back() == b.names.back();
        } else {
            r1 = r2 = false;
            fit2();
        }
    }

    EdgeOpArgs ea;
    ea.reversed = rev1;
    ea.merge = {{NAME_C1, free1}, {NAME_C2, common1}};
    ea.new_names = r1 ? std::vector<int32_t>{NAME_C1, NAME_C2} : std::vector<int32_t>{NAME_C2, NAME_C1};
    ea.ex_rev_before = common_rev1;
    ea.ex_merge = {NAME_C2};
    p.perm_a.clear();
    plan_edge_operator(a, ea, p.ma, p.perm_a, 0, 2);
    EdgeOpArgs eb;
    eb.reversed = rev2;
    eb.merge = {{NAME_C2, free2}, {NAME_C1, common2}};
    eb.new_names = r2 ? std::vector<int32_t>{NAME_C2, NAME_C1} : std::vector<int32_t>{NAME_C1, NAME_C2};
    p.perm_b.clear();
    plan_edge_operator(b, eb, p.mb, p.perm_b, 1, 3);

    const Edge& e0 = p.ma.edges[r1 ? 0 : 1];
    const Edge& e1 = p.mb.edges[r2 ? 0 : 1];
    const Edge& ec = p.mb.edges[r2 ? 1 : 0];
    p.prod = Shape();
    p.prod.kind_id = a.kind_id;
    p.prod.names = {NAME_C1, NAME_C2};
    p.prod.edges = {e0, e1};
    p.prod.finalize();
    p.gemms.clear();
    p.fills.clear();
    p.flops = 0;
    for (size_t s0 = 0; s0 < e0.segs.size(); s0++) {
        const Sym q = e0.segs[s0].q;
        const int p0 = e0.find(q);
        const int p1 = e1.find(sym_neg(k, q));
        if (p1 < 0) continue;
        const int64_t m = e0.segs[p0].d, n = e1.segs[p1].d;
        int64_t pos_r[2] = {p0, p1};
        const int64_t c_off = p.prod.offset(pos_r);
        const int pc = ec.find(q);
        if (pc < 0) {
            if (m * n) p.fills.push_back({4, c_off, m * n});
            continue;
        }
        const int64_t kk = ec.segs[pc].d;
        if (m && n && kk) {
            int64_t pos1[2] = {r1 ? p0 : pc, r1 ? pc : p0};
            int64_t pos2[2] = {r2 ? p1 : pc, r2 ? pc : p1};
            GemmTask g;
            g.a_buf = 2;
            g.b_buf = 3;
            g.c_buf = 4;
            g.a = p.ma.offset(pos1);
            g.b = p.mb.offset(pos2);
            g.c = c_off;
            g.m = int32_t(m);
            g.n = int32_t(n);
            g.k = int32_t(kk);
            g.ta = r1 ? 0 : 1;
            g.lda = int32_t(r1 ? kk : m);
            g.tb = r2 ? 1 : 0;
            g.ldb = int32_t(r2 ? kk : n);
            g.ldc = int32_t(n);
            g.alpha = (fermi && ((r1 ^ !r2) && sym_parity(k, q))) ? -1.0 : 1.0;
            p.gemms.push_back(g);
            p.flops += 2 * m * n * kk;
        } else if (m && n) {
            p.fills.push_back({4, c_off, m * n});
        }
    }
    EdgeOpArgs es;
    es.split = {split1, split2};
    es.reversed = rev_res;
    es.new_names = names_res;
    p.perm_out.clear();
    plan_edge_operator(p.prod, es, p.result, p.perm_out, 4, 5);
}

// ------------------------------------------------------------------------
// svd (svd.hpp:260-538)
// ------------------------------------------------------------------------
void plan_svd_factor(const Shape& t, const std::vector<int32_t>& free_u, int32_t cu, int32_t cv, int32_t su, int32_t sv,
                     SvdPlan& p) {
    const Kind& k = t.kind();
    const bool fermi = k.is_fermi;
    for (auto n : free_u)
        if (t.index_of(n) < 0) throw Error("Missing name in svd");
    p = SvdPlan();
    p.kind_id = t.kind_id;
    p.cu = cu;
    p.cv = cv;
    p.su = su;
    p.sv = sv;
    p.put_v_right = t.names.empty() || !contains(free_u, t.names.back());
    std::vector<int32_t> list_u, list_v, rev_in, rev_u, rev_v, res_u, res_v;
    SplitSpec sp_u{NAME_SVD_U, {}}, sp_v{NAME_SVD_V, {}};
    if (p.put_v_right)
        res_v.push_back(cv);
    else
        res_u.push_back(cu);
    for (int i = 0; i < t.rank(); i++) {
        const int32_t n = t.names[i];
        Edge e = t.edges[i];
        e.arrow = false;
        if (contains(free_u, n)) {
            list_u.push_back(n);
            res_u.push_back(n);
            sp_u.parts.push_back({n, e});
            if (fermi && t.edges[i].arrow) {
                rev_u.push_back(n);
                rev_in.push_back(n);
            }
        } else {
            list_v.push_back(n);
            res_v.push_back(n);
            sp_v.parts.push_back({n, e});
            if (fermi && t.edges[i].arrow) {
                rev_v.push_back(n);
                rev_in.push_back(n);
            }
        }
    }
    if (p.put_v_right)
        res_u.push_back(cu);
    else
        res_v.push_back(cv);
    EdgeOpArgs em;
    em.reversed = rev_in;
    em.merge = {{NAME_SVD_U, list_u}, {NAME_SVD_V, list_v}};
    em.new_names = p.put_v_right ? std::vector<int32_t>{NAME_SVD_U, NAME_SVD_V} : std::vector<int32_t>{NAME_SVD_V, NAME_SVD_U};
    plan_edge_operator(t, em, p.merged, p.perm_in, 0, 1);
    const Edge& e0 = p.merged.edges[0];
    const Edge& e1 = p.merged.edges[1];
    Edge c1, c2;
    for (auto& s : e0.segs) {
        int f = e1.find(sym_neg(k, s.q));
        if (f >= 0) {
            int64_t kk = std::min(s.d, e1.segs[f].d);
            c1.segs.push_back({sym_neg(k, s.q), kk});
            c2.segs.push_back({s.q, kk});
        }
    }
    p.t1.kind_id = p.t2.kind_id = t.kind_id;
    p.t1.names = p.put_v_right ? std::vector<int32_t>{NAME_SVD_U, cu} : std::vector<int32_t>{NAME_SVD_V, cv};
    p.t1.edges = {e0, c1};
    p.t1.finalize();
    p.t2.names = p.put_v_right ? std::vector<int32_t>{cv, NAME_SVD_V} : std::vector<int32_t>{cu, NAME_SVD_U};
    p.t2.edges = {c2, e1};
    p.t2.finalize();
    for (auto& s : e0.segs) {
        const int p0 = e0.find(s.q);
        const int p1 = e1.find(sym_neg(k, s.q));
        if (p1 < 0) continue;
        const int pc = p.t2.edges[0].find(s.q);
        const int64_t m = e0.segs[p0].d, n = e1.segs[p1].d, kk = p.t2.edges[0].segs[pc].d;
        p.s_sym.push_back(p.put_v_right ? sym_neg(k, s.q) : s.q);
        p.s_k.push_back(kk);
        p.s_off.push_back(p.s_total);
        if (m * n != 0) {
            SvdTask st;
            st.in_buf = 1;
            st.u_buf = 2;
            st.v_buf = 3;
            st.s_buf = 4;
            int64_t pm[2] = {p0, p1}, pu[2] = {p0, pc}, pv[2] = {pc, p1};
            st.in = p.merged.offset(pm);
            st.u = p.t1.offset(pu);
            st.v = p.t2.offset(pv);
            st.s = p.s_total;
            st.m = int32_t(m);
            st.n = int32_t(n);
            p.svds.push_back(st);
            const int64_t mm = std::max(m, n), nn = std::min(m, n);
            p.flops += 14 * mm * nn * nn + 8 * nn * nn * nn;
        }
        p.s_total += kk;
    }
    p.args_u.split = {sp_u};
    p.args_u.reversed = rev_u;
    if (fermi) p.args_u.reversed.push_back(cu);
    p.args_u.new_names = res_u;
    p.args_v.split = {sp_v};
    p.args_v.reversed = rev_v;
    p.args_v.new_names = res_v;
}

void plan_svd_truncate(SvdPlan& p, const double* svals, int64_t remain_cut, double relative_cut) {
    const Kind& k = kind_by_id(p.kind_id);
    const int ns = int(p.s_sym.size());
    int64_t total_dim = 0;
    double smax = 0;
    for (int i = 0; i < ns; i++) {
        total_dim += p.s_k[i];
        for (int64_t j = 0; j < p.s_k[i]; j++) {
            const double v = svals[p.s_off[i] + j];
            // the Jacobi kernels mark a sector whose iteration did not
            // converge by NaN singular values (LAPACK's gesvd reports info > 0
            // there and TAT raises)
            if (std::isnan(v)) throw Error("svd: the Jacobi iteration of a sector did not converge");
            smax = std::max(smax, v);
        }
    }
    const double threshold = relative_cut > 0 ? relative_cut * smax : 0.0;
    const int64_t rc = (remain_cut < 0 || remain_cut > total_dim) ? total_dim : remain_cut;
    p.remain.assign(ns, 0);
    // svd.hpp:454-472: repeatedly keep the largest next value across sectors
    for (int64_t it = 0; it < rc; it++) {
        int best = -1;
        double bv = 0;
        for (int i = 0; i < ns; i++) {
            if (p.remain[i] != p.s_k[i]) {
                double v = svals[p.s_off[i] + p.remain[i]];
                if (v > bv) {
                    bv = v;
                    best = i;
                }
            }
        }
        if (best >= 0 && bv > threshold)
            p.remain[best]++;
        else
            break;
    }
    CutSpec cut_u{p.cu, {}}, cut_v{p.cv, {}};
    for (int i = 0; i < ns; i++) {
        cut_u.dims.push_back({p.s_sym[i], p.remain[i]});
        cut_v.dims.push_back({sym_neg(k, p.s_sym[i]), p.remain[i]});
    }
    p.args_u.cut = {cut_u};
    p.args_v.cut = {cut_v};
    const Shape& tu = p.put_v_right ? p.t1 : p.t2;
    const Shape& tv = p.put_v_right ? p.t2 : p.t1;
    const int32_t ub = p.put_v_right ? 2 : 3, vb = p.put_v_right ? 3 : 2;
    p.perm_u.clear();
    p.perm_v.clear();
    plan_edge_operator(tu, p.args_u, p.u, p.perm_u, ub, 5);
    plan_edge_operator(tv, p.args_v, p.v, p.perm_v, vb, 7);
    // S tensor (svd.hpp:214-254)
    p.s = Shape();
    p.s.kind_id = p.kind_id;
    p.s.names = {p.su, p.sv};
    Edge eu, ev;
    eu.arrow = false;
    ev.arrow = true;
    for (int i = 0; i < ns; i++) {
        if (p.remain[i] == 0) continue;
        eu.segs.push_back({sym_neg(k, p.s_sym[i]), p.remain[i]});
        ev.segs.push_back({p.s_sym[i], p.remain[i]});
    }
    p.s.edges = {eu, ev};
    p.s.finalize();
    p.fills_s.clear();
    p.diag_s.clear();
    if (p.s.size) p.fills_s.push_back({6, 0, p.s.size});
    for (int i = 0; i < ns; i++) {
        const int64_t r = p.remain[i];
        if (r == 0) continue;
        int64_t pos[2] = {eu.find(sym_neg(k, p.s_sym[i])), ev.find(p.s_sym[i])};
        PermTask t;
        t.src_buf = 4;
        t.dst_buf = 6;
        t.src = p.s_off[i];
        t.dst = p.s.offset(pos);
        t.neg = (k.is_fermi && !p.put_v_right && sym_parity(k, p.s_sym[i])) ? 1 : 0;
        t.rank = 1;
        t.n = r;
        t.dims[0] = int32_t(r);
        t.ss[0] = 1;
        t.ds[0] = int32_t(r + 1);
        p.diag_s.push_back(t);
    }
}

// ------------------------------------------------------------------------
// qr (qr.hpp:310-508)
// ------------------------------------------------------------------------
void plan_qr(const Shape& t, char direction, const std::vector<int32_t>& free_names, int32_t cq, int32_t cr, QrPlan& p) {
    const Kind& k = t.kind();
    const bool fermi = k.is_fermi;
    for (auto n : free_names)
        if (t.index_of(n) < 0) throw Error("Missing name in qr");
    bool use_r;
    if (direction == 'r' || direction == 'R')
        use_r = true;
    else if (direction == 'q' || direction == 'Q')
        use_r = false;
    else
        throw Error("Invalid direction in QR");
    p = QrPlan();
    p.use_qr = t.names.empty() || (contains(free_names, t.names.back()) == use_r);
    std::vector<int32_t> l1, l2, rev_in, rev1, rev2, res1, res2;
    SplitSpec sp1{NAME_QR_1, {}}, sp2{NAME_QR_2, {}};
    res2.push_back(p.use_qr ? cr : cq);
    for (int i = 0; i < t.rank(); i++) {
        const int32_t n = t.names[i];
        Edge e = t.edges[i];
        e.arrow = false;
        if ((contains(free_names, n) == use_r) == p.use_qr) {
            l2.push_back(n);
            res2.push_back(n);
            sp2.parts.push_back({n, e});
            if (fermi && t.edges[i].arrow) {
                rev2.push_back(n);
                rev_in.push_back(n);
            }
        } else {
            l1.push_back(n);
            res1.push_back(n);
            sp1.parts.push_back({n, e});
            if (fermi && t.edges[i].arrow) {
                rev1.push_back(n);
                rev_in.push_back(n);
            }
        }
    }
    res1.push_back(p.use_qr ? cq : cr);
    EdgeOpArgs em;
    em.reversed = rev_in;
    em.merge = {{NAME_QR_1, l1}, {NAME_QR_2, l2}};
    em.new_names = {NAME_QR_1, NAME_QR_2};
    plan_edge_operator(t, em, p.merged, p.perm_in, 0, 1);
    const Edge& e0 = p.merged.edges[0];
    const Edge& e1 = p.merged.edges[1];
    Edge c1, c2;
    for (auto& s : e0.segs) {
        int f = e1.find(sym_neg(k, s.q));
        if (f >= 0) {
            int64_t kk = std::min(s.d, e1.segs[f].d);
            c1.segs.push_back({sym_neg(k, s.q), kk});
            c2.segs.push_back({s.q, kk});
        }
    }
    p.t1.kind_id = p.t2.kind_id = t.kind_id;
    p.t1.names = {NAME_QR_1, p.use_qr ? cq : cr};
    p.t1.edges = {e0, c1};
    p.t1.finalize();
    p.t2.names = {p.use_qr ? cr : cq, NAME_QR_2};
    p.t2.edges = {c2, e1};
    p.t2.finalize();
    for (auto& s : e0.segs) {
        const int p0 = e0.find(s.q);
        const int p1 = e1.find(sym_neg(k, s.q));
        if (p1 < 0) continue;
        const int pc = p.t2.edges[0].find(s.q);
        const int64_t m = e0.segs[p0].d, n = e1.segs[p1].d;
        if (m * n == 0) continue;
        QrTask q;
        q.in_buf = 1;
        q.f1_buf = 2;
        q.f2_buf = 3;
        int64_t pm[2] = {p0, p1}, p1p[2] = {p0, pc}, p2p[2] = {pc, p1};
        q.in = p.merged.offset(pm);
        q.f1 = p.t1.offset(p1p);
        q.f2 = p.t2.offset(p2p);
        q.m = int32_t(m);
        q.n = int32_t(n);
        q.lq = p.use_qr ? 0 : 1;
        p.qrs.push_back(q);
        const int64_t mm = std::max(m, n), nn = std::min(m, n);
        p.flops += 4 * mm * nn * nn - 4 * nn * nn * nn / 3;
    }
    if (fermi) (p.use_qr ? rev1 : rev2).push_back(cq);
    EdgeOpArgs a1;
    a1.split = {sp1};
    a1.reversed = rev1;
    a1.new_names = res1;
    EdgeOpArgs a2;
    a2.split = {sp2};
    a2.reversed = rev2;
    a2.new_names = res2;
    if (!p.use_qr) a2.ex_rev_before = {cq};
    // outputs: buffer 4 = Q, 5 = R
    Shape n1, n2;
    plan_edge_operator(p.t1, a1, n1, p.perm_1, 2, p.use_qr ? 4 : 5);
    plan_edge_operator(p.t2, a2, n2, p.perm_2, 3, p.use_qr ? 5 : 4);
    if (p.use_qr) {
        p.q = std::move(n1);
        p.r = std::move(n2);
    } else {
        p.q = std::move(n2);
        p.r = std::move(n1);
    }
}

// ------------------------------------------------------------------------
// conjugate (conjugate.hpp:32-120)
// ------------------------------------------------------------------------
void plan_conjugate(const Shape& t, bool trivial_metric, Shape& out, std::vector<PermTask>& tasks) {
    const Kind& k = t.kind();
    out = Shape();
    out.kind_id = t.kind_id;
    out.names = t.names;
    for (auto& e : t.edges) out.edges.push_back(k.ncomp ? edge_conjugate(k, e) : e);
    out.finalize();
    std::vector<int64_t> pos(t.rank());
    for (auto lin : t.blocks) {
        t.positions_of(lin, pos.data());
        int nodd = 0;
        bool tm = false;
        int64_t sz = 1;
        for (int i = 0; i < t.rank(); i++) {
            const Segment& s = t.edges[i].segs[pos[i]];
            sz *= s.d;
            if (k.is_fermi && sym_parity(k, s.q)) {
                nodd++;
                tm ^= (t.edges[i].arrow && trivial_metric);
            }
        }
        if (sz == 0) continue;
        PermTask pt;
        pt.src_buf = 0;
        pt.dst_buf = 1;
        pt.src = t.block_off[lin];
        pt.dst = out.block_off[lin];
        pt.neg = (tm ^ ((nodd & 2) != 0)) ? 1 : 0;
        pt.rank = 1;
        pt.n = sz;
        pt.dims[0] = int32_t(sz);
        pt.ss[0] = 1;
        pt.ds[0] = 1;
        tasks.push_back(pt);
    }
}

}  // namespace tns
