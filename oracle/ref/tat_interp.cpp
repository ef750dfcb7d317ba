// Small script interpreter over the reference TAT headers (TEST INFRASTRUCTURE).
//
// Built by oracle/ref/Makefile from this file plus the header-only TAT found
// in the reference checkout; the binary lands in oracle/_ref/ and is used only
// by tests (golden vectors) and by bench.py's cpu_baseline leg.  It restates no
// tensor algorithm itself: every tensor operation is TAT's own
// (contract.hpp, edge_operator.hpp, svd.hpp, qr.hpp).  The amplitude command
// drives TAT in the order tetragono does
// (tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498 for the
// boundary-MPS row, :254-312 and :710-720 for the final row / hole(())).
//
// Script grammar (whitespace separated tokens, one command per line):
//   sym <No|BoseZ2|BoseU1|FermiZ2|FermiU1>
//   tensor <id> <rank>          followed by <rank> lines:
//     edge <name> <nseg> <arrow> (<q> <d>)*
//   data <id> <n> v1 .. vn      set storage
//   range <id> <first> <step>
//   contract <out> <a> <b> <npairs> (<n1> <n2>)*
//   transpose <out> <a> <rank> names...
//   rename <out> <a> <n> (<old> <new>)*
//   svd <u> <s> <v> <a> <nfree> names... <cu> <cv> <su> <sv> <cut>
//   qr <q> <r> <a> <dir> <nfree> names... <cq> <cr>
//   merge <out> <a> <newname> <n> names... <apply_parity 0|1>
//   split <out> <a> <oldname> <n> (<name> <nseg> (<q> <d>)*)* <apply_parity>
//   reverse <out> <a> <n> names... <apply_parity 0|1>
//   conjugate <out> <a>
//   print <id>
//   dump <id> <path>            Tensor::dump() bytes to a file
//   load <id> <path>            Tensor::load() of a file's bytes
//   amplitude <out> <L1> <L2> <cut> <prefix>   sites <prefix>_<l1>_<l2> already
//       shrunk (physical legs contracted away); computes psi with tetragono's
//       contraction order and stores it as tensor <out>.
//   time_amplitude <L1> <L2> <cut> <nconf> <prefix>   shrunk sites
//       <prefix>_<c>_<l1>_<l2>; prints "time <seconds> <sum |psi|>"
//   time_raw <L1> <L2> <cut> <prefix> <nconf> occ...   raw sites
//       <prefix>_<l1>_<l2> (with P0); projects each configuration with a
//       one-hot shrinker (sampling_lattice/lattice.py:291-339) and evaluates
//       psi; projection and amplitude are both inside the timed region
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include <TAT/TAT.hpp>

using Name = TAT::DefaultName;

template<typename Sym>
struct Interp {
    using T = TAT::Tensor<double, Sym, Name>;
    using E = TAT::Edge<Sym>;
    std::map<std::string, T> env;

    static Sym make_sym(long long q) {
        if constexpr (Sym::length == 0) {
            (void)q;
            return Sym();
        } else {
            return Sym(static_cast<std::tuple_element_t<0, typename Sym::base_tuple_t>>(q));
        }
    }
    static long long sym_value(const Sym& s) {
        if constexpr (Sym::length == 0) {
            (void)s;
            return 0;
        } else {
            return static_cast<long long>(std::get<0>(static_cast<const typename Sym::base_tuple_t&>(s)));
        }
    }

    E read_segments(std::istream& in, int nseg, bool arrow) {
        std::vector<std::pair<Sym, TAT::Size>> segs;
        for (int i = 0; i < nseg; i++) {
            long long q;
            unsigned long long d;
            in >> q >> d;
            segs.emplace_back(make_sym(q), d);
        }
        return E(std::move(segs), arrow);
    }

    void print(const std::string& id, const T& t) {
        std::printf("tensor %s %d\n", id.c_str(), int(t.rank()));
        for (int i = 0; i < t.rank(); i++) {
            const auto& e = t.edges(i);
            std::printf("edge %s %d %d", t.names(i).c_str(), int(e.segments_size()), int(e.arrow()));
            for (const auto& [q, d] : e.segments()) {
                std::printf(" %lld %llu", sym_value(q), (unsigned long long)d);
            }
            std::printf("\n");
        }
        std::printf("data %zu", t.storage().size());
        for (auto v : t.storage()) {
            std::printf(" %.17g", v);
        }
        std::printf("\n");
    }

    // tetragono/tetragono/utility.py:340-350: drop pairs whose names are absent.
    static T safe_contract(const T& a, const T& b, const std::vector<std::pair<Name, Name>>& pairs) {
        std::unordered_set<std::pair<Name, Name>> real;
        for (const auto& [x, y] : pairs) {
            if (a.find_by_name(x) != a.names().end() && b.find_by_name(y) != b.names().end()) {
                real.insert({x, y});
            }
        }
        return a.contract(b, real);
    }
    // tetragono/tetragono/utility.py:353-358
    static T safe_rename(const T& a, const std::vector<std::pair<Name, Name>>& m) {
        std::unordered_map<Name, Name> real;
        for (const auto& [x, y] : m) {
            if (a.find_by_name(x) != a.names().end()) {
                real[x] = y;
            }
        }
        return a.edge_rename(real);
    }

    // single_layer_auxiliaries.py:463-498 with udlr = "DULR" (bottom-up sweep).
    static std::vector<T> two_line_to_one_line(const std::vector<T>& line_1, const std::vector<T>& line_2, double cut) {
        const Name up = "D", down = "U", left = "L", right = "R";
        const Name left1 = "L1", left2 = "L2", right1 = "R1", right2 = "R2";
        const auto length = line_1.size();
        std::vector<T> dl;
        for (size_t i = 0; i < length; i++) {
            dl.push_back(safe_contract(
                safe_rename(line_1[i], {{left, left1}, {right, right1}}),
                safe_rename(line_2[i], {{left, left2}, {right, right2}}),
                {{down, up}}
            ));
        }
        for (size_t i = 0; i + 1 < length; i++) {
            std::unordered_set<Name> rn;
            for (const auto& n : {right1, right2}) {
                if (dl[i].find_by_name(n) != dl[i].names().end()) {
                    rn.insert(n);
                }
            }
            auto [q, r] = dl[i].qr('r', rn, right, left);
            dl[i] = q;
            dl[i + 1] = safe_contract(dl[i + 1], r, {{left1, right1}, {left2, right2}});
        }
        for (size_t i = length - 1; i >= 1; i--) {
            TAT::Cut c = cut > 0 ? (cut >= 1 ? TAT::Cut(TAT::Size(cut)) : TAT::Cut(cut)) : TAT::Cut();
            auto [u, s, v] = dl[i].svd({left}, right, left, left, right, c);
            dl[i] = v;
            dl[i - 1] = safe_contract(safe_contract(dl[i - 1], u, {{right, left}}), s, {{right, left}});
        }
        return dl;
    }

    // sampling_lattice/lattice.py:291-339: contract P0 with a one-hot shrinker
    static T shrink(const T& site, int l1, int l2, long long occupation) {
        const auto& phys = site.edges("P0");
        auto q = make_sym(occupation);
        T sh({"P", "Q"}, {E({{q, 1}}, false), phys.conjugate()});
        sh.zero_();
        sh.at(std::unordered_map<Name, std::pair<Sym, TAT::Size>>{{"Q", {-q, 0}}, {"P", {q, 0}}}) = 1;
        auto renamed = sh.edge_rename({{"P", "P_" + std::to_string(l1) + "_" + std::to_string(l2) + "_0"}});
        return site.contract(renamed, {{"P0", "Q"}});
    }

    // hole(()) with hint ("H", 0): single_layer_auxiliaries.py:710-720 ->
    // _inline_right_to_left[0, 0] (:284-312) over _down_to_up[1] (:430-442).
    static T amplitude(const std::vector<std::vector<T>>& lattice, int L1, int L2, double cut) {
        T one(1);
        std::vector<T> mps(L2, one);
        for (int l1 = L1 - 1; l1 >= 1; l1--) {
            mps = two_line_to_one_line(mps, lattice[l1], cut);
        }
        // inline right-to-left along row 0, up_to_down[-1] = ones
        T inline_r = one; // _inline_right_to_left[0, L2]
        for (int l2 = L2 - 1; l2 >= 0; l2--) {
            // tailed[0, l2+1]
            T tailed = (l2 + 1 == 0) ? inline_r : safe_rename(safe_contract(inline_r, mps[l2], {{"L3", "R"}}), {{"L", "L3"}});
            T res = safe_contract(tailed, safe_rename(lattice[0][l2], {{"L", "L2"}}), {{"L2", "R"}, {"U", "D"}});
            res = safe_contract(res, safe_rename(one, {{"L", "L1"}}), {{"L1", "R"}, {"U", "D"}});
            inline_r = res;
        }
        return inline_r;
    }

    void run(std::istream& in) {
        std::string cmd;
        while (in >> cmd) {
            if (cmd == "tensor") {
                std::string id;
                int rank;
                in >> id >> rank;
                std::vector<Name> names;
                std::vector<E> edges;
                for (int i = 0; i < rank; i++) {
                    std::string kw, name;
                    int nseg, arrow;
                    in >> kw >> name >> nseg >> arrow;
                    names.push_back(name);
                    edges.push_back(read_segments(in, nseg, arrow));
                }
                env.insert_or_assign(id, T(names, edges).zero_());
            } else if (cmd == "data") {
                std::string id;
                size_t n;
                in >> id >> n;
                auto& t = env.at(id);
                if (n != t.storage().size()) {
                    throw std::runtime_error("data size mismatch for " + id);
                }
                for (size_t i = 0; i < n; i++) {
                    in >> t.storage()[i];
                }
            } else if (cmd == "range") {
                std::string id;
                double first, step;
                in >> id >> first >> step;
                env.at(id).range_(first, step);
            } else if (cmd == "contract") {
                std::string out, a, b;
                int np;
                in >> out >> a >> b >> np;
                std::unordered_set<std::pair<Name, Name>> pairs;
                for (int i = 0; i < np; i++) {
                    std::string x, y;
                    in >> x >> y;
                    pairs.insert({x, y});
                }
                env.insert_or_assign(out, env.at(a).contract(env.at(b), pairs));
            } else if (cmd == "transpose") {
                std::string out, a;
                int r;
                in >> out >> a >> r;
                std::vector<Name> names(r);
                for (auto& n : names) {
                    in >> n;
                }
                env.insert_or_assign(out, env.at(a).transpose(names));
            } else if (cmd == "rename") {
                std::string out, a;
                int n;
                in >> out >> a >> n;
                std::unordered_map<Name, Name> m;
                for (int i = 0; i < n; i++) {
                    std::string x, y;
                    in >> x >> y;
                    m[x] = y;
                }
                env.insert_or_assign(out, env.at(a).edge_rename(m));
            } else if (cmd == "svd") {
                std::string u, s, v, a, cu, cv, su, sv;
                int nf;
                double cut;
                in >> u >> s >> v >> a >> nf;
                std::unordered_set<Name> fr;
                for (int i = 0; i < nf; i++) {
                    std::string x;
                    in >> x;
                    fr.insert(x);
                }
                in >> cu >> cv >> su >> sv >> cut;
                TAT::Cut c = cut > 0 ? (cut >= 1 ? TAT::Cut(TAT::Size(cut)) : TAT::Cut(cut)) : TAT::Cut();
                auto res = env.at(a).svd(fr, cu, cv, su, sv, c);
                env.insert_or_assign(u, res.U);
                env.insert_or_assign(s, res.S);
                env.insert_or_assign(v, res.V);
            } else if (cmd == "qr") {
                std::string q, r, a, dir, cq, cr;
                int nf;
                in >> q >> r >> a >> dir >> nf;
                std::unordered_set<Name> fr;
                for (int i = 0; i < nf; i++) {
                    std::string x;
                    in >> x;
                    fr.insert(x);
                }
                in >> cq >> cr;
                auto res = env.at(a).qr(dir[0], fr, cq, cr);
                env.insert_or_assign(q, res.Q);
                env.insert_or_assign(r, res.R);
            } else if (cmd == "merge") {
                std::string out, a, nn;
                int n, ap;
                in >> out >> a >> nn >> n;
                std::vector<Name> names(n);
                for (auto& x : names) {
                    in >> x;
                }
                in >> ap;
                env.insert_or_assign(out, env.at(a).merge_edge({{nn, names}}, ap != 0));
            } else if (cmd == "split") {
                std::string out, a, old;
                int n, ap;
                in >> out >> a >> old >> n;
                std::vector<std::pair<Name, TAT::edge_segments_t<Sym>>> plan;
                for (int i = 0; i < n; i++) {
                    std::string nm;
                    int nseg;
                    in >> nm >> nseg;
                    plan.push_back({nm, read_segments(in, nseg, false)});
                }
                in >> ap;
                env.insert_or_assign(out, env.at(a).split_edge({{old, plan}}, ap != 0));
            } else if (cmd == "reverse") {
                std::string out, a;
                int n, ap;
                in >> out >> a >> n;
                std::unordered_set<Name> names;
                for (int i = 0; i < n; i++) {
                    std::string x;
                    in >> x;
                    names.insert(x);
                }
                in >> ap;
                env.insert_or_assign(out, env.at(a).reverse_edge(names, ap != 0));
            } else if (cmd == "conjugate") {
                std::string out, a;
                in >> out >> a;
                env.insert_or_assign(out, env.at(a).conjugate());
            } else if (cmd == "dump") {
                std::string id, path;
                in >> id >> path;
                const std::string bytes = env.at(id).dump();
                std::ofstream f(path, std::ios::binary);
                f.write(bytes.data(), std::streamsize(bytes.size()));
            } else if (cmd == "load") {
                std::string id, path;
                in >> id >> path;
                std::ifstream f(path, std::ios::binary);
                const std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
                T t;
                t.load(bytes);
                env.insert_or_assign(id, std::move(t));
            } else if (cmd == "print") {
                std::string id;
                in >> id;
                print(id, env.at(id));
            } else if (cmd == "amplitude") {
                std::string out, prefix;
                int L1, L2;
                double cut;
                in >> out >> L1 >> L2 >> cut >> prefix;
                std::vector<std::vector<T>> lat(L1);
                for (int l1 = 0; l1 < L1; l1++) {
                    for (int l2 = 0; l2 < L2; l2++) {
                        lat[l1].push_back(env.at(prefix + "_" + std::to_string(l1) + "_" + std::to_string(l2)));
                    }
                }
                env.insert_or_assign(out, amplitude(lat, L1, L2, cut));
            } else if (cmd == "time_amplitude") {
                // time_amplitude <L1> <L2> <cut> <nconf> <prefix>: sites
                // <prefix>_<c>_<l1>_<l2> for c < nconf; prints seconds and sum|psi|
                std::string prefix;
                int L1, L2, nconf;
                double cut;
                in >> L1 >> L2 >> cut >> nconf >> prefix;
                std::vector<std::vector<std::vector<T>>> lats(nconf);
                for (int c = 0; c < nconf; c++) {
                    lats[c].resize(L1);
                    for (int l1 = 0; l1 < L1; l1++) {
                        for (int l2 = 0; l2 < L2; l2++) {
                            lats[c][l1].push_back(env.at(prefix + "_" + std::to_string(c) + "_" + std::to_string(l1) + "_" + std::to_string(l2)));
                        }
                    }
                }
                auto t0 = std::chrono::steady_clock::now();
                double acc = 0;
                for (int c = 0; c < nconf; c++) {
                    auto psi = amplitude(lats[c], L1, L2, cut);
                    acc += std::abs(psi.storage().size() ? psi.storage()[0] : 0.0);
                }
                auto t1 = std::chrono::steady_clock::now();
                std::printf("time %.9f %.17g\n", std::chrono::duration<double>(t1 - t0).count(), acc);
            } else if (cmd == "time_raw") {
                // time_raw <L1> <L2> <cut> <prefix> <nconf> then nconf*L1*L2
                // occupations: raw sites <prefix>_<l1>_<l2>; projection and
                // amplitude both timed; prints seconds and sum |psi|
                std::string prefix;
                int L1, L2, nconf;
                double cut;
                in >> L1 >> L2 >> cut >> prefix >> nconf;
                std::vector<int> occ(size_t(nconf) * L1 * L2);
                for (auto& o : occ) in >> o;
                auto t0 = std::chrono::steady_clock::now();
                double acc = 0;
                for (int c = 0; c < nconf; c++) {
                    std::vector<std::vector<T>> lat(L1);
                    for (int l1 = 0; l1 < L1; l1++)
                        for (int l2 = 0; l2 < L2; l2++)
                            lat[l1].push_back(shrink(env.at(prefix + "_" + std::to_string(l1) + "_" + std::to_string(l2)),
                                                     l1, l2, occ[(size_t(c) * L1 + l1) * L2 + l2]));
                    auto psi = amplitude(lat, L1, L2, cut);
                    acc += std::abs(psi.storage().size() ? psi.storage()[0] : 0.0);
                }
                auto t1 = std::chrono::steady_clock::now();
                std::printf("time %.9f %.17g\n", std::chrono::duration<double>(t1 - t0).count(), acc);
            } else if (cmd == "#") {
                std::string rest;
                std::getline(in, rest);
            } else {
                throw std::runtime_error("unknown command " + cmd);
            }
        }
    }
};

int main() {
    std::ios::sync_with_stdio(false);
    std::string cmd, sym;
    std::cin >> cmd >> sym;
    if (cmd != "sym") {
        std::fprintf(stderr, "script must start with: sym <kind>\n");
        return 2;
    }
    try {
        if (sym == "No") {
            Interp<TAT::NoSymmetry>().run(std::cin);
        } else if (sym == "BoseZ2") {
            Interp<TAT::BoseZ2Symmetry>().run(std::cin);
        } else if (sym == "BoseU1") {
            Interp<TAT::BoseU1Symmetry>().run(std::cin);
        } else if (sym == "FermiZ2") {
            Interp<TAT::FermiZ2Symmetry>().run(std::cin);
        } else if (sym == "FermiU1") {
            Interp<TAT::FermiU1Symmetry>().run(std::cin);
        } else {
            std::fprintf(stderr, "unknown symmetry %s\n", sym.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fflush(stdout);
    return 0;
}
