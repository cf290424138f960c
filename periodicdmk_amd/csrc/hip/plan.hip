// Device plan construction and the evaluate() launch sequence.
#include "plan.hpp"

#include "../host/nufft.hpp"
#include "../host/tables.hpp"
#include "../host/tree_build.hpp"
#include "kernels.hpp"
#include "tree_gpu.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace pdmk::gpu {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

void DevBuf::alloc(size_t nbytes) {
    if (p) { check(hipFree(p), "hipFree"); p = nullptr; }
    bytes = nbytes;
    if (nbytes) check(hipMalloc(&p, nbytes), "hipMalloc");
}

void *Arena::take(size_t nbytes) {
    const size_t off = (used + 255) & ~size_t(255);
    if (!dry() && off + nbytes > cap)
        throw std::length_error("device scratch exhausted: " + std::to_string(off + nbytes) + " bytes needed, " +
                                std::to_string(cap) + " available");
    used = off + nbytes;
    return dry() ? nullptr : base + off;
}

namespace {

constexpr size_t kPinBytes = size_t(512) << 20;  // incoming plane-wave scratch cap
constexpr size_t kTwoPassBytes = size_t(512) << 20;  // two-pass transform scratch cap
// child -> parent levels with at least this many parents take one workgroup item per parent
// (no scratch round trip; clustered N = 4e6: 12 digits 441 -> 418 ms, 9 digits 144 -> 136 ms);
// below it one item per child keeps the GPU filled (N = 1e6: 64 parents, FP64 6 digits 3.94 ->
// 4.16 ms with one item per parent)
constexpr int kC2pFuseParents = 256;

// proxy slots from which the charge -> proxy matrix-core kernel (p = 28) takes its wide
// workgroup shape: 64 boxes per CU (the N=1e6 uniform tree has about 5e3 slots, the
// clustered N=4e6 one 1.2e5)
constexpr int kC2pWideBoxes = 16384;

// f(i0, i1) over contiguous chunks of [0, n) on up to 16 host threads (construction work
// whose chunks write disjoint data)
template <class F> void host_parallel(int n, F f) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + nt - 1) / nt;
    for (int w = 0; w < nt; ++w) {
        const int i0 = w * chunk, i1 = std::min(n, i0 + chunk);
        if (i0 < i1) pool.emplace_back([&f, i0, i1] { f(i0, i1); });
    }
    for (auto &th : pool) th.join();
}

// Tuned knobs per plan kind, shared by the plans of the process; optionally mirrored in the
// text file named by PDMK_TUNE_CACHE (one "key near_wpb near_late p2pw_bk pw2p_bk t3_div"
// line per kind, later lines win).
std::mutex tune_mu;
std::map<std::string, Knobs> tune_cache;
bool tune_file_read = false;

const char *tune_file() { return std::getenv("PDMK_TUNE_CACHE"); }

void tune_cache_load() {
    if (tune_file_read) return;
    tune_file_read = true;
    const char *f = tune_file();
    if (!f) return;
    std::ifstream in(f);
    std::string key;
    Knobs k;
    while (in >> key >> k.near_wpb >> k.near_late >> k.p2pw_bk >> k.pw2p_bk >> k.t3_div) tune_cache[key] = k;
}

int log2_bucket(long long v) {
    int b = 0;
    while (v > 1) { v >>= 1; ++b; }
    return b;
}

int nc_for_digits(int d) {
    static const int nc[] = {5, 5, 5, 6, 9, 11, 12, 15, 17, 18, 22, 23, 25};
    return nc[std::min(12, std::max(0, d))];
}


}  // namespace

template <typename Real>
Plan<Real>::Plan(const double *cell, const double *pos, int N, int n_digits_user, int n_per_leaf, int nz,
                 const double *bloch_in) {
    // eps = 10^-max(1, n_digits) (cpp/src/periodic_dmk.cpp:52): below 1 the tree is built
    // for 1 digit; above 12 there are no tables
    if (n_digits_user > 12) throw std::invalid_argument("PeriodicDMK: n_digits must be at most 12");
    n_digits_user = std::max(1, n_digits_user);
    // streams and events are released on any throw below (a destructor does not run for a
    // partly built object; the DevBuf members free themselves)
    struct HandleGuard {
        Plan *pl;
        bool armed = true;
        ~HandleGuard() {
            if (armed) pl->release_handles();
        }
    } guard{this};
    check(hipStreamCreate(&stream), "hipStreamCreate");
    for (auto &e : ev) check(hipEventCreate(&e), "hipEventCreate");
    check(hipStreamCreate(&s_far), "hipStreamCreate");
    check(hipStreamCreate(&s_near), "hipStreamCreate");
    for (hipEvent_t *e : {&e_in, &e_gather, &e_far, &e_near, &e_up})
        check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    using clk = std::chrono::steady_clock;
    const auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    const auto t_start = clk::now();
    host::Geometry g;
    host::Tree t;
    // per-point tree construction on the GPU; a tree deeper than its path keys hold is
    // built with the points on the host (same result)
    tree_on_gpu = build_tree_gpu(cell, pos, N, nz, n_per_leaf, stream, g, t);
    if (!tree_on_gpu) {
        g = host::make_geometry(cell, pos, N, nz);
        t = host::build_tree(g, n_per_leaf);
    }
    build_ms[1] = ms_since(t_start);
    const host::DmkTables T = host::make_tables(n_digits_user, g.rc, t.nlevels, g.umat);
    n = N;
    rc = g.rc;
    nsall = g.nsall();
    nboxes = t.nboxes;
    nlevels = t.nlevels;
    p = T.p;
    n_pw = T.n_pw;
    n_digits = T.n_digits;
    nc = (int)T.residual.size();
    if (nc != nc_for_digits(n_digits)) throw std::logic_error("PeriodicDMK: residual table mismatch");

    // ---------------- sources and targets (sorted order), coordinates relative to their
    // leaf centre: offsets between boxes are formed in FP64 on the host, so FP32 kernels
    // only see small local coordinates
    // Inside each leaf the points are further ordered by a Morton code of their position
    // in the leaf, so that runs of consecutive points are spatially compact (near-field
    // work items and source chunks); box ranges are unchanged by this.
    std::vector<int> src_leaf(nsall, -1), trg_leaf(n, -1), src_ord(nsall), trg_ord(n);
    // 10 bits of q spread to every third bit
    auto spread3 = [](unsigned long long q) {
        q &= 0x3ffull;
        q = (q | (q << 16)) & 0x030000ffull;
        q = (q | (q << 8)) & 0x0300f00full;
        q = (q | (q << 4)) & 0x030c30c3ull;
        q = (q | (q << 2)) & 0x09249249ull;
        return q;
    };
    auto morton_order = [&](const std::vector<double> &xyz, int beg, int cnt, int b, int *ord,
                            std::vector<std::pair<unsigned long long, int>> &key) {
        key.resize((size_t)cnt);
        const double inv = 1.0 / t.boxsize[t.level[b]];
        for (int i = 0; i < cnt; ++i) {
            unsigned long long code = 0;
            for (int d = 0; d < 3; ++d) {
                const double u = (xyz[3 * (size_t)(beg + i) + d] - t.center[3 * (size_t)b + d]) * inv + 0.5;
                const unsigned long long q = (unsigned long long)std::min(1023.0, std::max(0.0, std::floor(u * 1024.0)));
                code |= spread3(q) << d;
            }
            key[i] = {code, beg + i};
        }
        // (the index breaks ties: the order of a stable sort by code)
        std::sort(key.begin(), key.end());
        for (int i = 0; i < cnt; ++i) ord[beg + i] = key[i].second;
    };
    // leaves own disjoint point ranges: split over the host threads
    host_parallel(nboxes, [&](int b0, int b1) {
        std::vector<std::pair<unsigned long long, int>> key;
        for (int b = b0; b < b1; ++b) {
            if (!t.leaf[b]) continue;
            for (int i = 0; i < t.src_count[b]; ++i) src_leaf[t.src_begin[b] + i] = b;
            for (int i = 0; i < t.trg_count[b]; ++i) trg_leaf[t.trg_begin[b] + i] = b;
            morton_order(t.src_sorted, t.src_begin[b], t.src_count[b], b, src_ord.data(), key);
            morton_order(t.trg_sorted, t.trg_begin[b], t.trg_count[b], b, trg_ord.data(), key);
        }
    });
    auto local = [&](const std::vector<double> &xyz, size_t i, int b, int d) {
        return (Real)(xyz[3 * i + d] - t.center[3 * (size_t)b + d]);
    };
    std::vector<P4<Real>> srcv(nsall);
    std::vector<int> su(nsall);
    host_parallel(nsall, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const int b = src_leaf[i], j = src_ord[i];
            srcv[i] = {local(t.src_sorted, j, b, 0), local(t.src_sorted, j, b, 1), local(t.src_sorted, j, b, 2), Real(0)};
            const int o = t.src_order[j];
            su[i] = o < N ? o : g.img_src[o - N];
        }
    });
    if (bloch_in) {
        // quasi-periodic image phases e^{i b.R}, R = image position - original position in the
        // lattice frame (the translation of generate_images, cpp/include/dmk/pbc.hpp:221-229)
        qp = true;
        for (int d = 0; d < 3; ++d) bloch[d] = bloch_in[d];
        std::vector<double2> ph(nsall);
        for (int i = 0; i < nsall; ++i) {
            const int j = src_ord[i], o = t.src_order[j];
            double phi = 0;
            if (o >= N)
                for (int d = 0; d < 3; ++d)
                    phi += bloch[d] * (t.src_sorted[3 * (size_t)j + d] - g.r_src[3 * (size_t)su[i] + d]);
            ph[i] = make_double2(std::cos(phi), std::sin(phi));
        }
        src_phase.upload(ph);
    }
    std::vector<P4<Real>> trgv(n);
    std::vector<int> tu(n);
    std::vector<double> tw0(n, 0.0);
    host_parallel(n, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const int b = trg_leaf[i], j = trg_ord[i];
            trgv[i] = {local(t.trg_sorted, j, b, 0), local(t.trg_sorted, j, b, 1), local(t.trg_sorted, j, b, 2),
                       Real(0)};
            tu[i] = t.trg_order[j];
        }
    });
    for (int b = 0; b < nboxes; ++b)
        if (t.leaf[b] && t.trg_count[b] > 0) {
            const double w = T.w0[t.level[b] + t.pwexp[b]];
            for (int i = 0; i < t.trg_count[b]; ++i) tw0[t.trg_begin[b] + i] = w;
        }
    srcv.resize((size_t)nsall + NF_CHUNK, P4<Real>{Real(0), Real(0), Real(0), Real(0)});  // chunk padding
    srcv.push_back(P4<Real>{Real(1e18), Real(1e18), Real(1e18), Real(0)});  // sentinel: far from every target
    src.upload(srcv);
    srcv.resize(nsall);
    src_user.upload(su);
    trg.upload(trgv);
    trg_user.upload(tu);
    trg_w0.upload(tw0);
    pot_sorted.alloc(sizeof(Real) * (size_t)n);
    pot_near.alloc(sizeof(Real) * (size_t)n);
    q_dev.alloc(sizeof(double) * (size_t)n);
    pot_dev.alloc(sizeof(Real) * (size_t)n);
    far.alloc(sizeof(double) * (size_t)n);

    // ---------------- periodic folding
    // When the box grid is aligned with the lattice, a level-0 box of the image layer holds
    // the images, under one lattice translation, of the points of an inner box.  If the two
    // subtrees are the same (refinement, plane-wave flags, the same points leaf by leaf at the
    // same positions relative to the box centres), every box of the image subtree has its
    // counterpart's proxy charges and outgoing plane waves, which are relative to the box
    // centre: those boxes are not computed, and the downward pass reads the counterpart's.
    // fold[b] = the counterpart of b, or -1.  PDMK_FOLD=0 turns it off.  Quasi-periodic trees
    // never fold: an image's charge carries its translation's phase, so image boxes are not
    // copies of their counterparts.
    std::vector<int> fold(nboxes, -1);
    {
        const char *fe = std::getenv("PDMK_FOLD");
        const bool enabled = !(fe && std::string(fe) == "0") && !qp;
        auto user_of = [&](int j) {  // tree-order source -> user index
            const int o = t.src_order[j];
            return o < N ? o : g.img_src[o - N];
        };
        const double bs0 = t.boxsize[0], tol = 1e-12 * bs0;
        // (user index, position relative to the box centre) of a box's points, sorted by index
        using Pt = std::pair<int, std::array<double, 3>>;
        auto points_of = [&](int box, std::vector<Pt> &v) {
            v.clear();
            for (int i = 0; i < t.src_count[box]; ++i) {
                const int j = t.src_begin[box] + i;
                std::array<double, 3> x;
                for (int d = 0; d < 3; ++d) x[d] = t.src_sorted[3 * (size_t)j + d] - t.center[3 * box + d];
                v.emplace_back(user_of(j), x);
            }
            std::sort(v.begin(), v.end(), [](const Pt &u, const Pt &w) { return u.first < w.first; });
        };
        struct Scratch { std::vector<int> ub, up; std::vector<Pt> pb, pq; };
        std::function<bool(int, int, Scratch &)> same = [&](int b, int q, Scratch &sc) {
            if (t.level[b] != t.level[q] || t.leaf[b] != t.leaf[q] || t.pwexp[b] != t.pwexp[q] ||
                t.src_count[b] != t.src_count[q] || t.trg_count[b] != 0)
                return false;
            if (!t.leaf[b]) {
                for (int k = 0; k < 8; ++k)
                    if (!same(t.first_child[b] + k, t.first_child[q] + k, sc)) return false;
                return true;
            }
            sc.ub.clear();
            sc.up.clear();
            for (int i = 0; i < t.src_count[b]; ++i) sc.ub.push_back(user_of(t.src_begin[b] + i));
            for (int i = 0; i < t.src_count[q]; ++i) sc.up.push_back(user_of(t.src_begin[q] + i));
            std::sort(sc.ub.begin(), sc.ub.end());
            std::sort(sc.up.begin(), sc.up.end());
            return sc.ub == sc.up;
        };
        std::function<int(int, int)> mark = [&](int b, int q) {
            fold[b] = q;
            int m = 1;
            if (!t.leaf[b])
                for (int k = 0; k < 8; ++k) m += mark(t.first_child[b] + k, t.first_child[q] + k);
            return m;
        };
        // b: a level-0 box of images; its counterpart q, or -1
        auto counterpart = [&](int b, Scratch &sc) {
            const int cnt = t.src_count[b], beg = t.src_begin[b];
            if (cnt == 0 || t.src_order[beg] < N) return -1;
            // the translation of the first point; every point of b must be an image under it
            const int u0 = user_of(beg);
            double R[3];
            for (int d = 0; d < 3; ++d) R[d] = t.src_sorted[3 * (size_t)beg + d] - g.r_src[3 * (size_t)u0 + d];
            int qc[3];
            for (int d = 0; d < 3; ++d) qc[d] = (int)std::lround((t.center[3 * b + d] - R[d] - t.xelo[d]) / bs0 - 0.5);
            const int q = t.find(0, qc[0], qc[1], qc[2]);
            if (q < 0 || t.src_count[q] != cnt) return -1;
            for (int d = 0; d < 3; ++d)
                if (!(std::fabs(t.center[3 * q + d] - (t.center[3 * b + d] - R[d])) <= tol)) return -1;
            // q holds originals only, each once; b holds images only, of the same points at the
            // same positions relative to the box centres
            for (int i = 0; i < cnt; ++i)
                if (t.src_order[t.src_begin[q] + i] >= N || t.src_order[beg + i] < N) return -1;
            points_of(q, sc.pq);
            points_of(b, sc.pb);
            for (int i = 0; i < cnt; ++i) {
                if (i > 0 && sc.pq[i].first == sc.pq[i - 1].first) return -1;
                if (sc.pb[i].first != sc.pq[i].first) return -1;
                for (int d = 0; d < 3; ++d)
                    if (!(std::fabs(sc.pb[i].second[d] - sc.pq[i].second[d]) <= tol)) return -1;
            }
            return same(b, q, sc) ? q : -1;
        };
        // level-0 image boxes are independent (disjoint subtrees): split over host threads
        const std::vector<int> roots = enabled ? t.by_level[0] : std::vector<int>();
        const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<int> found(roots.size(), -1);
        std::vector<std::thread> pool;
        for (int w = 0; w < nthr; ++w)
            pool.emplace_back([&, w] {
                Scratch sc;
                for (size_t i = w; i < roots.size(); i += nthr) found[i] = counterpart(roots[i], sc);
            });
        for (auto &th : pool) th.join();
        for (size_t i = 0; i < roots.size(); ++i)
            if (found[i] >= 0) n_folded += mark(roots[i], found[i]);
    }

    // ---------------- proxy slots
    // up_of: the slot a box's proxy and outgoing plane waves are computed in (-1 for folded
    // boxes), numbered level by level (box order within a level) so that the slots one level's
    // shifts read form one contiguous range; up_src: the slot they are read from (the
    // counterpart's, on the same level, for folded boxes)
    std::vector<int> up_of(nboxes, -1), up_src(nboxes, -1), down_of(nboxes, -1), up_box;
    for (int l = 0; l < nlevels; ++l)
        for (int b : t.by_level[l])
            if (t.pwexp[b] && t.src_count[b] > 0 && fold[b] < 0) {
                up_of[b] = n_up++;
                up_box.push_back(b);
            }
    for (int b = 0; b < nboxes; ++b)
        if ((t.pwexp[b] || t.tpeval[b]) && t.src_count[b] + t.trg_count[b] > 0) down_of[b] = n_down++;
    for (int b = 0; b < nboxes; ++b) up_src[b] = fold[b] < 0 ? up_of[b] : up_of[fold[b]];
    const size_t P3 = (size_t)p * p * p;
    const int H = (n_pw + 1) / 2, NM = n_pw * n_pw * H;
    proxy_up.alloc(sizeof(Real) * P3 * std::max(1, n_up));
    proxy_down.alloc(sizeof(Real) * P3 * std::max(1, n_down));

    // ---------------- upward: one charge->proxy group per proxy slot (possibly without sources)
    {
        std::unordered_map<int, const host::Tree::UpGroup *> grp_of;
        for (const auto &gr : t.up_groups) grp_of[gr.box] = &gr;
        std::vector<int> slot(n_up), rng(n_up + 1, 0), rb, rcnt;
        std::vector<Real> geom(4 * (size_t)n_up), roff;
        for (int k = 0; k < n_up; ++k) {
            const int b = up_box[k];
            slot[k] = k;
            for (int d = 0; d < 3; ++d) geom[4 * k + d] = (Real)t.center[3 * b + d];
            geom[4 * k + 3] = (Real)(2.0 / t.boxsize[t.level[b]]);
            rng[k + 1] = 0;
            auto it = grp_of.find(b);
            if (it == grp_of.end()) continue;
            for (int sb : it->second->src_boxes) {
                // split the source range at leaf boundaries: one centre offset per piece
                const int end = t.src_begin[sb] + t.src_count[sb];
                for (int i = t.src_begin[sb]; i < end;) {
                    const int lf = src_leaf[i];
                    const int stop = std::min(end, t.src_begin[lf] + t.src_count[lf]);
                    rb.push_back(i);
                    rcnt.push_back(stop - i);
                    for (int d = 0; d < 3; ++d) roff.push_back((Real)(t.center[3 * lf + d] - t.center[3 * b + d]));
                    roff.push_back(Real(0));
                    rng[k + 1]++;
                    i = stop;
                }
            }
        }
        // ranges were appended in slot order
        for (int k = 0; k < n_up; ++k) rng[k + 1] += rng[k];
        up_slot.upload(slot);
        up_geom.upload(geom);
        up_rng.upload(rng);
        rng_begin.upload(rb);
        rng_count.upload(rcnt);
        rng_off.upload(roff);
        double c2p_src = 0;
        for (size_t i = 0; i < rcnt.size(); ++i) c2p_src += rcnt[i];
        work.transform_flops += 2.0 * (double)P3 * c2p_src;
    }
    auto make_t3 = [&](const std::vector<int> &out, const std::vector<int> &add, const std::vector<int> &ptr,
                       const std::vector<int> &in, const std::vector<int> &code) {
        Tensor3Launch L;
        L.n_items = (int)out.size();
        L.out.upload(out);
        L.add.upload(add);
        L.ptr.upload(ptr);
        L.in_slot.upload(in);
        L.in_code.upload(code);
        work.transform_flops += 6.0 * std::pow((double)p, 4) * (double)in.size();
        return L;
    };
    // child -> parent, per level: with at least kC2pFuseParents parents, one item per parent,
    // its children (in child order) the inputs, added to the proxy the parent already holds
    // (every slot's charge -> proxy group writes it); with fewer, every child is an item of its
    // own (more workgroups) into a scratch slot, and k_sum_children adds each parent's
    // children's slots in child order (PDMK_C2P_FUSE=<parents> moves the threshold)
    int c2p_max = 0, fuse_min = kC2pFuseParents;
    if (const char *e = std::getenv("PDMK_C2P_FUSE")) fuse_min = std::max(1, std::atoi(e));
    for (int l = nlevels - 1; l >= 0; --l) {
        std::vector<int> pslot, cptr{0}, in, code;
        for (int b : t.by_level[l]) {
            if (!t.proxy_from_children[b] || fold[b] >= 0) continue;
            for (int k = 0; k < 8; ++k) {
                const int ch = t.first_child[b] + k;
                if (t.pwexp[ch] && t.src_count[ch] > 0) {
                    in.push_back(up_of[ch]);
                    code.push_back(k);
                }
            }
            if ((int)in.size() == cptr.back()) continue;
            pslot.push_back(up_of[b]);
            cptr.push_back((int)in.size());
        }
        if (pslot.empty()) continue;
        C2pLevel C;
        C.n_parents = (int)pslot.size();
        C.fused = C.n_parents >= fuse_min;
        if (C.fused) {
            C.t3 = make_t3(pslot, std::vector<int>(pslot.size(), 1), cptr, in, code);
        } else {
            std::vector<int> out(in.size()), add(in.size(), 0), ptr(in.size() + 1);
            for (size_t i = 0; i < in.size(); ++i) out[i] = (int)i;
            for (size_t i = 0; i <= in.size(); ++i) ptr[i] = (int)i;
            C.t3 = make_t3(out, add, ptr, in, code);
            C.parent_slot.upload(pslot);
            C.child_ptr.upload(cptr);
            c2p_max = std::max(c2p_max, (int)in.size());
        }
        c2p_levels.push_back(std::move(C));
    }
    {
        // parent <-> child matrices, row-major U(r, c) = M_s[r + p c] (parent -> child) and
        // M_s[c + p r] (child -> parent), M_s = p2c on the minus / plus side
        std::vector<Real> rm(4 * (size_t)p * p);
        for (int tr = 0; tr < 2; ++tr)
            for (int sd = 0; sd < 2; ++sd)
                for (int r = 0; r < p; ++r)
                    for (int c = 0; c < p; ++c)
                        rm[(size_t)(2 * tr + sd) * p * p + (size_t)r * p + c] =
                            (Real)T.p2c[sd][tr ? c + p * r : r + p * c];
        mats_rm.upload(rm);
    }

    // ---------------- plane-wave tables
    {
        std::vector<Cx<Real>> a(T.poly2pw.size()), b(T.pw2poly.size()), s(T.shift0.size());
        for (size_t i = 0; i < a.size(); ++i) a[i] = {(Real)T.poly2pw[i].real(), (Real)T.poly2pw[i].imag()};
        for (size_t i = 0; i < b.size(); ++i) b[i] = {(Real)T.pw2poly[i].real(), (Real)T.pw2poly[i].imag()};
        for (size_t i = 0; i < s.size(); ++i) s[i] = {(Real)T.shift0[i].real(), (Real)T.shift0[i].imag()};
        poly2pw.upload(a);
        pw2poly.upload(b);
        if (sizeof(Real) == 8) {
            // FP64 shifts from the separable parts of the table (k_pw_shift_sep)
            const int half = n_pw / 2, nf = (n_pw - 1) / 2;
            const double sphase = 2 * T.beta / nf;
            std::vector<Cx<double>> eph(n_pw);
            for (int j = 0; j < n_pw; ++j) {
                const std::complex<double> z = std::polar(1.0, (j - half) * sphase);
                eph[j] = {z.real(), z.imag()};
            }
            shift_e.upload(eph);
            shift_r.upload(T.radial0);
        } else {
            shift0.upload(s);
        }
        // folded tables F[j * 64 + k] = the non-zero part of W(k, j) (real for even j,
        // imaginary for odd j), for the two-pass transforms
        if (n_pw <= 64) {
            std::vector<Real> fa((size_t)p * 64, Real(0)), fb((size_t)p * 64, Real(0));
            for (int j = 0; j < p; ++j)
                for (int k = 0; k < n_pw; ++k) {
                    const Cx<Real> wa = a[k + (size_t)n_pw * j], wb = b[k + (size_t)n_pw * j];
                    fa[(size_t)j * 64 + k] = (j % 2 == 0) ? wa.re : wa.im;
                    fb[(size_t)j * 64 + k] = (j % 2 == 0) ? wb.re : wb.im;
                }
            poly2pw_f.upload(fa);
            pw2poly_f.upload(fb);
        }
        // per box: z step from the real proxy (one multiply-add per term), y and x steps on
        // complex values against matrix entries that are real or imaginary (two per term)
        work.transform_flops += (double)n_up * (2.0 * (double)p * p * H * p + 4.0 * (double)p * n_pw * H * p +
                                                4.0 * (double)NM * p);
    }

    // ---------------- downward per level
    std::vector<uint8_t> set_by_parent(nboxes, 0);
    for (int l = 0; l < nlevels; ++l) {
        std::vector<int> down, zero, nptr{0}, nslot, nsten;
        std::vector<int> out, add, ptr{0}, in, code;
        for (int b : t.by_level[l]) {
            const bool has_out = t.trg_count[b] > 0;
            if (!(t.pwexp[b] && t.src_count[b] + t.trg_count[b] > 0 && has_out)) continue;
            down.push_back(down_of[b]);
            zero.push_back(set_by_parent[b] ? 0 : 1);
            nslot.push_back(up_src[b]);
            nsten.push_back(13);
            for (int s = 0; s < 27; ++s) {
                if (s == 13) continue;
                const int nb = t.colleague(b, s);
                if (nb < 0) continue;
                if (t.leaf[b] && t.leaf[nb]) continue;
                if (up_src[nb] < 0) continue;
                nslot.push_back(up_src[nb]);
                nsten.push_back(26 - s);  // stencil of (centre_box - centre_nbr)
            }
            nptr.push_back((int)nslot.size());
            if (!t.tpeval[b] && !t.leaf[b])
                for (int k = 0; k < 8; ++k) {
                    const int ch = t.first_child[b] + k;
                    if (t.src_count[ch] + t.trg_count[ch] == 0) continue;
                    out.push_back(down_of[ch]);
                    add.push_back(0);
                    in.push_back(down_of[b]);
                    code.push_back(k);
                    ptr.push_back((int)in.size());
                    set_by_parent[ch] = 1;
                }
        }
        EvalLevel E;
        E.level = l;
        E.n_items = (int)down.size();
        if (!nslot.empty()) {
            const auto mm = std::minmax_element(nslot.begin(), nslot.end());
            E.pw_lo = *mm.first;
            E.pw_cnt = *mm.second - *mm.first + 1;
        }
        E.down.upload(down);
        E.zero.upload(zero);
        E.nbr_ptr.upload(nptr);
        E.nbr_slot.upload(nslot);
        E.nbr_sten.upload(nsten);
        {
            // the same lists merged per group of kShiftGroup consecutive items, by source slot
            std::vector<int> gptr{0}, gslot;
            std::vector<unsigned long long> gsten;
            const int ni = (int)down.size();
            std::vector<std::pair<int, int>> ent;  // (slot, target t * 32 + stencil)
            for (int g0 = 0; g0 < ni; g0 += kShiftGroup) {
                ent.clear();
                for (int t = 0; t < kShiftGroup && g0 + t < ni; ++t)
                    for (int e = nptr[g0 + t]; e < nptr[g0 + t + 1]; ++e) ent.push_back({nslot[e], t * 32 + nsten[e]});
                std::stable_sort(ent.begin(), ent.end(),
                                 [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first < y.first; });
                const size_t first = gslot.size();
                for (const auto &pe : ent) {
                    const int t = pe.second / 32, st = pe.second % 32;
                    // the last entry of this slot, unless target t already has it (a slot twice in
                    // one target's list: a folded neighbour)
                    size_t k = gslot.size();
                    if (k > first && gslot[k - 1] == pe.first && ((gsten[k - 1] >> (8 * t)) & 255) == 255) {
                        gsten[k - 1] = (gsten[k - 1] & ~(255ull << (8 * t))) | ((unsigned long long)st << (8 * t));
                    } else {
                        gslot.push_back(pe.first);
                        gsten.push_back(~(255ull << (8 * t)) | ((unsigned long long)st << (8 * t)));
                    }
                }
                gptr.push_back((int)gslot.size());
            }
            E.grp_ptr.upload(gptr);
            E.grp_slot.upload(gslot);
            E.grp_sten.upload(gsten);
        }
        work.transform_flops += 8.0 * (double)nslot.size() * NM;
        // per box: x and y steps complex against real-or-imaginary entries (two multiply-adds
        // per term), z step to a real output (one)
        work.transform_flops += (double)E.n_items *
                                (4.0 * (double)p * n_pw * H * n_pw + 4.0 * (double)p * p * H * n_pw + 2.0 * (double)P3 * H);
        eval_levels.push_back(std::move(E));
        p2c_levels.push_back(make_t3(out, add, ptr, in, code));
    }
    {
        std::vector<int> ld, lt;
        std::vector<Real> lg;
        for (int b = 0; b < nboxes; ++b) {
            if (!(t.tpeval[b] && t.trg_count[b] > 0)) continue;
            ld.push_back(down_of[b]);
            for (int d = 0; d < 3; ++d) lg.push_back(Real(0));  // targets are relative to this leaf's centre
            lg.push_back((Real)(2.0 / t.boxsize[t.level[b]]));
            lt.push_back(t.trg_begin[b]);
            lt.push_back(t.trg_count[b]);
            work.transform_flops += 2.0 * (double)P3 * t.trg_count[b];
        }
        n_leaf_eval = (int)ld.size();
        leaf_down.upload(ld);
        leaf_geom.upload(lg);
        leaf_trg.upload(lt);
    }

    // ---------------- near field
    {
        // source chunks of every leaf: bounding boxes of NF_CHUNK consecutive points, taken
        // over the rounded coordinates the kernel reads
        std::vector<int> cbase(nboxes, -1);
        std::vector<P4<Real>> cb;
        auto box_of = [](const P4<Real> *pts, int cnt, P4<Real> &lo, P4<Real> &hi) {
            lo = pts[0];
            hi = pts[0];
            for (int i = 1; i < cnt; ++i) {
                lo.x = std::min(lo.x, pts[i].x); lo.y = std::min(lo.y, pts[i].y); lo.z = std::min(lo.z, pts[i].z);
                hi.x = std::max(hi.x, pts[i].x); hi.y = std::max(hi.y, pts[i].y); hi.z = std::max(hi.z, pts[i].z);
            }
            lo.w = hi.w = Real(0);
        };
        for (int b = 0; b < nboxes; ++b) {
            if (!t.leaf[b] || t.src_count[b] == 0) continue;
            cbase[b] = (int)(cb.size() / 2);
            for (int c = 0; c < t.src_count[b]; c += NF_CHUNK) {
                P4<Real> lo, hi;
                box_of(&srcv[t.src_begin[b] + c], std::min(NF_CHUNK, t.src_count[b] - c), lo, hi);
                cb.push_back(lo);
                cb.push_back(hi);
            }
        }
        // even/odd (r^2) monomial coefficients, one row per distinct bs:
        // sum_i c_i rsc^i (r + cen)^i expanded in powers of r
        const bool split = n_digits < 6;
        std::vector<double> bs_seen;
        std::vector<Real> split_tab;
        auto cf_index = [&](double bs) {
            if (!split) return 0;
            for (size_t k = 0; k < bs_seen.size(); ++k)
                if (bs_seen[k] == bs) return (int)k;
            bs_seen.push_back(bs);
            const double rsc = 2.0 / bs, cen = -bs / 2.0;
            std::vector<double> cd(nc);
            double rp = 1;
            for (int i = 0; i < nc; ++i) { cd[i] = T.residual[i] * rp; rp *= rsc; }
            for (int i = 0; i < nc; ++i)
                for (int j = nc - 1; j > i; --j) cd[j - 1] += cen * cd[j];
            for (int i = 0; i < NF_SPLIT_STRIDE; ++i) split_tab.push_back(i < nc ? (Real)cd[i] : Real(0));
            return (int)bs_seen.size() - 1;
        };
        std::vector<int> tr, ptr{0};
        std::vector<NfEntry> l1v;
        std::vector<P4<Real>> gv;  // (centre of target leaf - centre of source leaf, bs)
        std::vector<int4> itv;
        std::vector<P4<Real>> wb;
        double cand = 0, src_pts = 0;
        for (int b = 0; b < nboxes; ++b) {
            if (!t.leaf[b] || t.trg_count[b] == 0 || t.l1_begin[b + 1] == t.l1_begin[b]) continue;
            const int li = (int)ptr.size() - 1;
            tr.push_back(t.trg_begin[b]);
            tr.push_back(t.trg_count[b]);
            for (int e = t.l1_begin[b]; e < t.l1_begin[b + 1]; ++e) {
                const int sb = t.l1_box[e];
                double bs = t.boxsize[t.level[sb]];
                if (t.pwexp[sb] && sb == b) bs *= 0.5;
                else if (t.level[sb] < t.level[b]) bs = t.boxsize[t.level[b]];
                if (t.src_count[sb] == 0) continue;
                l1v.push_back({t.src_begin[sb], t.src_count[sb], cbase[sb], cf_index(bs)});
                gv.push_back({(Real)(t.center[3 * b] - t.center[3 * sb]), (Real)(t.center[3 * b + 1] - t.center[3 * sb + 1]),
                              (Real)(t.center[3 * b + 2] - t.center[3 * sb + 2]), (Real)bs});
                cand += (double)t.trg_count[b] * t.src_count[sb];
                src_pts += t.src_count[sb];
            }
            ptr.push_back((int)l1v.size());
            for (int w = 0; w < t.trg_count[b]; w += 64) {
                const int cnt = std::min(64, t.trg_count[b] - w);
                itv.push_back(make_int4(t.trg_begin[b] + w, cnt, li, 0));
                P4<Real> lo, hi;
                box_of(&trgv[t.trg_begin[b] + w], cnt, lo, hi);
                wb.push_back(lo);
                wb.push_back(hi);
            }
        }
        n_nf_leaves = (int)ptr.size() - 1;
        n_nf_items = (int)itv.size();
        nf_trg.upload(tr);
        nf_ptr.upload(ptr);
        nf_l1.upload(l1v);
        nf_geom.upload(gv);
        nf_items.upload(itv);
        nf_wbox.upload(wb);
        nf_cbox.upload(cb);
        if (split) coeffs.upload(split_tab);
        else coeffs.upload(std::vector<Real>(T.residual.begin(), T.residual.end()));
        work.nf_pairs_candidate = cand;
        stats.n_list1 = (long long)l1v.size();
        // step lists: counts per (item, entry), prefix on the host, then the source indices;
        // the first pass also counts the pairs inside the support exactly
        std::vector<int> ieb(n_nf_items);
        int n_ie = 0;
        for (int i = 0; i < n_nf_items; ++i) {
            ieb[i] = n_ie;
            n_ie += ptr[itv[i].z + 1] - ptr[itv[i].z];
        }
        nf_ie_base.upload(ieb);
        // FP32 at 6 digits: quad-row step lists (k_near_rows); PDMK_NEAR_ROWS=0 keeps the
        // wavefront-wide lists of k_near_field
        const char *rows_env = std::getenv("PDMK_NEAR_ROWS");
        nf_rows = !split && (nc == 12 || (std::is_same<Real, double>::value && (nc == 18 || nc == 25))) &&
                  !(rows_env && rows_env[0] == '0');
        std::vector<long long> wbase;
        std::vector<int> rsb, rend;
        long long total = 0;
        unsigned long long np = 0;
        const dim3 grid((n_nf_items + 3) / 4);
        if (n_nf_items > 0 && nf_rows) {
            DevBuf cnt, pairs;
            cnt.alloc(sizeof(int) * (size_t)std::max(1, n_ie) * NF_ROWS);
            pairs.alloc(sizeof(unsigned long long));
            check(hipMemsetAsync(pairs.p, 0, sizeof(unsigned long long), stream), "hipMemsetAsync");
            hipLaunchKernelGGL((k_near_rows_build<Real, 0>), grid, dim3(256), 0, stream, src.as<P4<Real>>(),
                               trg.as<P4<Real>>(), nf_items.as<int4>(), nf_wbox.as<P4<Real>>(),
                               nf_ptr.as<int>(), nf_l1.as<NfEntry>(), nf_geom.as<P4<Real>>(),
                               nf_cbox.as<P4<Real>>(), nf_ie_base.as<int>(), cnt.as<int>(),
                               (const long long *)nullptr, (int *)nullptr, pairs.as<unsigned long long>(), n_nf_items);
            check(hipGetLastError(), "k_near_rows_build count");
            std::vector<int> c((size_t)std::max(1, n_ie) * NF_ROWS);
            check(hipMemcpyAsync(c.data(), cnt.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
            check(hipMemcpyAsync(&np, pairs.p, sizeof(np), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            // per item: K = longest row rounded up to 4, block base; per (item, row): slots of the
            // non-empty entries (geometry, cumulative end) and a terminator
            wbase.assign(n_nf_items, 0);
            rsb.assign(n_nf_items, 0);
            int n_slots = 0, k_item_max = 0;
            nf_ne_max = 0;
            for (int i = 0; i < n_nf_items; ++i) {
                const int ne = ptr[itv[i].z + 1] - ptr[itv[i].z];
                nf_ne_max = std::max(nf_ne_max, ne);
                rsb[i] = n_slots;
                n_slots += ne + 1;
            }
            rend.assign((size_t)n_slots * NF_ROWS, 0);
            // items own disjoint slot ranges: rows of all items on the host threads
            host_parallel(n_nf_items, [&](int i0, int i1) {
                for (int i = i0; i < i1; ++i) {
                    const int ne = ptr[itv[i].z + 1] - ptr[itv[i].z];
                    // slot = (end step << 8) | entry ordinal; unused slots hold the terminator
                    std::fill(rend.begin() + (size_t)rsb[i] * NF_ROWS, rend.begin() + (size_t)(rsb[i] + ne + 1) * NF_ROWS,
                              (0x7fffff << 8) | std::min(ne, 255));
                    int kmax = 0;
                    for (int g = 0; g < NF_ROWS; ++g) {
                        int k = 0, sl = 0;
                        for (int e = 0; e < ne; ++e) {
                            const int ce = c[(size_t)(ieb[i] + e) * NF_ROWS + g];
                            if (ce == 0) continue;
                            k += ce;
                            rend[(size_t)(rsb[i] + sl) * NF_ROWS + g] = (std::min(k, 0x7fffff) << 8) | e;
                            ++sl;
                        }
                        kmax = std::max(kmax, k);
                    }
                    itv[i].w = (kmax + 3) & ~3;  // K
                }
            });
            for (int i = 0; i < n_nf_items; ++i) {
                k_item_max = std::max(k_item_max, itv[i].w);
                wbase[i] = total;
                total += (long long)itv[i].w * NF_ROWS;
            }
            n_nf_steps = total / NF_ROWS;  // steps per row summed over the items
            // beyond int indexing, 8-bit entry ordinals, 16-bit step counts (the kernels decode
            // slots into 16-bit end steps), or source byte offsets past 32 bits (the kernels
            // form them as index * size in 32 bits; the many-vector charges take 64 bytes per
            // source): wavefront lists
            const long long k_lim = 0xfff0;
            const unsigned long long src_bytes = std::is_same<Real, float>::value ? 64 : sizeof(P4<Real>);
            if (total + NF_ROW_PAD > 0x7fffffffLL || nf_ne_max > 255 || k_item_max > k_lim ||
                (unsigned long long)(nsall + NF_CHUNK + 1) * src_bytes > 0xffffffffull)
                nf_rows = false;
            // the row lists take at most a quarter of the free device memory (the plane waves
            // and the rest of the plan come after them)
            size_t mem_free = 0, mem_total = 0;
            check(hipMemGetInfo(&mem_free, &mem_total), "hipMemGetInfo");
            if ((double)(total + NF_ROW_PAD) * sizeof(int) > 0.25 * (double)mem_free) nf_rows = false;
        }
        if (n_nf_items > 0 && nf_rows) {
            nf_items.upload(itv);
            nf_rs_base.upload(rsb);
            nf_rslot.upload(rend);
            nf_wave_base.upload(wbase);
            // the padding points at the sentinel source (far from every target)
            const int sentinel = nsall + NF_CHUNK;
            nf_row_steps.alloc(sizeof(int) * (size_t)(total + NF_ROW_PAD));
            const long long n_fill = total + NF_ROW_PAD;
            hipLaunchKernelGGL(k_fill<int>, dim3((unsigned)((n_fill + 255) / 256)), dim3(256), 0, stream,
                               nf_row_steps.as<int>(), sentinel, (int)n_fill);
            check(hipGetLastError(), "k_fill row steps");
            hipLaunchKernelGGL((k_near_rows_build<Real, 1>), grid, dim3(256), 0, stream, src.as<P4<Real>>(),
                               trg.as<P4<Real>>(), nf_items.as<int4>(), nf_wbox.as<P4<Real>>(),
                               nf_ptr.as<int>(), nf_l1.as<NfEntry>(), nf_geom.as<P4<Real>>(),
                               nf_cbox.as<P4<Real>>(), nf_ie_base.as<int>(), (int *)nullptr,
                               nf_wave_base.as<long long>(), nf_row_steps.as<int>(), (unsigned long long *)nullptr,
                               n_nf_items);
            check(hipGetLastError(), "k_near_rows_build write");
            work.nf_pairs_support = (double)np;
        }
        if (n_nf_items > 0 && !nf_rows) {
            DevBuf cnt, pairs;
            cnt.alloc(sizeof(int) * (size_t)std::max(1, n_ie));
            pairs.alloc(sizeof(unsigned long long));
            check(hipMemsetAsync(pairs.p, 0, sizeof(unsigned long long), stream), "hipMemsetAsync");
            const dim3 grid((n_nf_items + 3) / 4);
            hipLaunchKernelGGL((k_near_steps<Real, 0>), grid, dim3(256), 0, stream, src.as<P4<Real>>(), trg.as<P4<Real>>(),
                               nf_items.as<int4>(), nf_wbox.as<P4<Real>>(), nf_ptr.as<int>(), nf_l1.as<NfEntry>(),
                               nf_geom.as<P4<Real>>(), nf_cbox.as<P4<Real>>(), nf_ie_base.as<int>(), cnt.as<int>(),
                               (const long long *)nullptr, (int *)nullptr, pairs.as<unsigned long long>(), n_nf_items);
            check(hipGetLastError(), "k_near_steps count");
            std::vector<int> c(std::max(1, n_ie));
            unsigned long long np = 0;
            check(hipMemcpyAsync(c.data(), cnt.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
            check(hipMemcpyAsync(&np, pairs.p, sizeof(np), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            std::vector<long long> iep((size_t)n_ie + 1, 0);
            for (int k = 0; k < n_ie; ++k) iep[k + 1] = iep[k] + c[k];
            n_nf_steps = iep.back();
            nf_ie_ptr.upload(iep);
            // zero-filled padding: a group of steps may read up to NF_STEP_PAD indices past the
            // end of the last run (point 0, never used)
            nf_steps.alloc(sizeof(int) * (size_t)(n_nf_steps + NF_STEP_PAD));
            check(hipMemsetAsync(nf_steps.p, 0, sizeof(int) * (size_t)(n_nf_steps + NF_STEP_PAD), stream), "hipMemsetAsync");
            hipLaunchKernelGGL((k_near_steps<Real, 1>), grid, dim3(256), 0, stream, src.as<P4<Real>>(), trg.as<P4<Real>>(),
                               nf_items.as<int4>(), nf_wbox.as<P4<Real>>(), nf_ptr.as<int>(), nf_l1.as<NfEntry>(),
                               nf_geom.as<P4<Real>>(), nf_cbox.as<P4<Real>>(), nf_ie_base.as<int>(), (int *)nullptr,
                               nf_ie_ptr.as<long long>(), nf_steps.as<int>(), (unsigned long long *)nullptr, n_nf_items);
            check(hipGetLastError(), "k_near_steps write");
            work.nf_pairs_support = (double)np;
        }
        stats.n_nf_steps = n_nf_steps;
        // near-field algorithmic HBM bytes: every target read and written once, every
        // list entry's sources read once per target leaf
        work.nf_bytes = src_pts * sizeof(P4<Real>) + (double)n * (sizeof(P4<Real>) + 2 * sizeof(Real));
    }

    // ---------------- far field
    {
        ms = T.ms; mt = T.mt; mu = T.mu;
        std::vector<double> th(3 * (size_t)n);
        const double *A = T.ainv;
        for (int u = 0; u < n; ++u) {
            const double x = g.r_src[3 * (size_t)u], y = g.r_src[3 * (size_t)u + 1], z = g.r_src[3 * (size_t)u + 2];
            double f[3] = {A[0] * x + A[1] * y + A[3] * z, A[2] * y + A[4] * z, A[5] * z};
            for (int d = 0; d < 3; ++d) {
                f[d] -= std::floor(f[d]);
                th[3 * (size_t)u + d] = 2.0 * M_PI * f[d];
            }
        }
        n_ff_modes = (int)T.mode_w.size();
        // the NUFFT (kernel width n_digits + 2, error ~10^-(n_digits+1) of the far field, below
        // the reference's eps/4) once the direct sum's N x modes work is large; PDMK_FAR_FIELD
        // = exact | nufft forces either
        far_nufft = (double)n * n_ff_modes > 1e9;
        if (const char *e = std::getenv("PDMK_FAR_FIELD")) {
            if (std::string(e) == "exact") far_nufft = 0;
            else if (std::string(e) == "nufft") far_nufft = 1;
        }
        // far[] is in user order for the direct sum and in the NUFFT's point order otherwise;
        // far_idx[i] = where the far field of sorted target i sits
        std::vector<int> fidx(tu);
        const auto t_far = clk::now();
        if (far_nufft) {
            nf.setup(th, n, ms, mt, mu, T.mode_ijk, T.mode_w, std::max(4, std::min(14, n_digits + 2)));
            for (int i = 0; i < n; ++i) fidx[i] = nf.position[tu[i]];
        } else {
            // at most 512 point blocks: each writes a partial sum per mode that the reduction
            // reads back (2048 blocks at N = 2e5, 6 digits, made 95 MB of partial sums per call)
            ff.setup(th, n, ms, mt, mu, T.mode_ijk, T.mode_w, nullptr, 512);
        }
        far_idx.upload(fidx);
        if (qp) {
            // quasi-periodic far field (cpp/src/dmk/pbc.cpp:249-303): mode box from
            // dkmax = beta/rc + |b|, weights at the shifted wave vectors, b.x per point
            const double twopi = 2 * M_PI, *U = g.umat;
            const double bn = std::sqrt(bloch[0] * bloch[0] + bloch[1] * bloch[1] + bloch[2] * bloch[2]);
            const double dk = T.beta / rc + bn;
            const double n1 = U[0], n2 = std::sqrt(U[3] * U[3] + U[4] * U[4]),
                         n3 = std::sqrt(U[6] * U[6] + U[7] * U[7] + U[8] * U[8]);
            const int qs = std::max(2, 2 * (int)std::ceil(dk * n1 / twopi));
            const int qt = std::max(2, 2 * (int)std::ceil(dk * n2 / twopi));
            const int qu = std::max(2, 2 * (int)std::ceil(dk * n3 / twopi));
            const double *a = T.ainv;  // a11, a21, a22, a31, a32, a33
            const double dfac = 4 * M_PI / (T.psi0 * U[0] * U[4] * U[8]);
            const host::Pswf pf(T.beta);
            std::vector<double> wq((size_t)qs * qt * qu, 0.0), bx(n);
            for (int k = 0; k < qu; ++k)
                for (int j = 0; j < qt; ++j)
                    for (int i = 0; i < qs; ++i) {
                        const int mi = i - qs / 2, mj = j - qt / 2, mk = k - qu / 2;
                        const double k1 = twopi * a[0] * mi + bloch[0], k2 = twopi * (a[1] * mi + a[2] * mj) + bloch[1],
                                     k3 = twopi * (a[3] * mi + a[4] * mj + a[5] * mk) + bloch[2];
                        const double kk = k1 * k1 + k2 * k2 + k3 * k3;
                        if (kk < 1e-30) continue;
                        const double x = std::sqrt(kk) * rc / T.beta;
                        if (x <= 1) wq[i + (size_t)qs * (j + (size_t)qt * k)] = dfac * pf(x) / kk;
                    }
            for (int u = 0; u < n; ++u)
                bx[u] = bloch[0] * g.r_src[3 * (size_t)u] + bloch[1] * g.r_src[3 * (size_t)u + 1] +
                        bloch[2] * g.r_src[3 * (size_t)u + 2];
            qpf.setup(th, bx, n, qs, qt, qu, wq);
            pot_dev2.alloc(2 * sizeof(Real) * (size_t)n);
        }
        build_ms[2] = ms_since(t_far);
        work.ff_mode_points = 2.0 * (double)n * n_ff_modes;
        work.ff_flops = 8.0 * (double)n * n_ff_modes + 4.0 * (double)n * n_ff_modes;
        if (far_nufft) {
            // the NUFFT's own work: spreading and interpolation (a multiply-add per point and
            // grid cell of the W^3 support), the pruned axis DFTs both ways (a complex
            // multiply-add per input and output pair along each axis), the mode weights
            const double W3 = (double)nf.W * nf.W * nf.W;
            const double g0 = nf.g[0], g1 = nf.g[1], g2 = nf.g[2], M0 = nf.M[0], M1 = nf.M[1], M2 = nf.M[2];
            const double dft = 8.0 * (g2 * g1 * g0 * M0 + g2 * g1 * M0 * M1 + g2 * M1 * M0 * M2);
            work.ff_flops = 4.0 * (double)n * W3 + 2.0 * dft + 8.0 * (double)n_ff_modes;
        }
    }

    // ---------------- incoming plane-wave scratch: one level's boxes at a time, in batches
    {
        int items = 1;
        for (auto &E : eval_levels) items = std::max(items, E.n_items);
        pin_batch = std::max(1, std::min(items, (int)(kPinBytes / ((size_t)NM * 2 * sizeof(Real)))));
        if (pin_batch < items) pin_batch = std::max(kShiftGroup, pin_batch / kShiftGroup * kShiftGroup);
        pin.alloc((size_t)pin_batch * NM * 2 * sizeof(Real));
        // two-pass proxy <-> plane-wave transforms (FP64 at p >= 28, whose slices re-read a
        // box's input kTwoPassSlices times or more): the first contraction of a chunk of boxes
        // (PDMK_TWO_PASS=0 turns them off, PDMK_TWO_PASS=<MiB> caps the chunk scratch)
        const int Hh = (n_pw + 1) / 2;
        long long tp_mib = -1;
        if (const char *e = std::getenv("PDMK_TWO_PASS")) tp_mib = std::atoll(e);
        if (sizeof(Real) == 8 && p >= 28 && tp_mib != 0) {
            const size_t z_box = (size_t)Hh * p * p * 2 * sizeof(Real);  // H p^2 complex
            const size_t x_box = (size_t)p * Hh * n_pw * 2 * sizeof(Real);      // p H n complex
            const size_t want = std::max((size_t)std::max(1, n_up) * z_box, (size_t)pin_batch * x_box);
            const size_t cap = tp_mib > 0 ? (size_t)tp_mib << 20 : kTwoPassBytes;
            const size_t bytes = std::min(want, std::max(cap, std::max(z_box, x_box)));
            tp_scratch.alloc(bytes);
            tp_boxes_p2pw = (int)std::max<size_t>(1, bytes / z_box);
            tp_boxes_pw2p = (int)std::max<size_t>(1, bytes / x_box);
        }
    }
    // ---------------- outgoing plane waves and the child -> parent scratch: all up slots when
    // they fit in the free device memory (less 1 GiB), else one level's slot range at a time
    // (PDMK_PW_LEVELS=1 / 0 forces either)
    {
        const size_t pw_slot = 2 * sizeof(Real) * (size_t)NM;
        const size_t c2p_bytes = sizeof(Real) * P3 * (size_t)std::max(1, c2p_max);
        int lvl_cnt = 1;
        for (auto &E : eval_levels) lvl_cnt = std::max(lvl_cnt, E.pw_cnt);
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const size_t all_bytes = pw_slot * (size_t)std::max(1, n_up) + c2p_bytes;
        pw_by_level = all_bytes + (size_t(1) << 30) > free_b;
        if (const char *e = std::getenv("PDMK_PW_LEVELS")) pw_by_level = e[0] == '1';
        if (const char *e = std::getenv("PDMK_T3_MFMA")) t3_mfma = e[0] != '0';
        // charge -> proxy workgroup shape at p = 28 (same sums either way): clustered N=4e6,
        // 124,287 slots, 9-digit evaluate() 130.9 -> 126.2 ms with the wide shape; N=1e6
        // uniform 7.73 -> 7.53 ms with the 16-wave one
        c2p_wide = n_up >= kC2pWideBoxes;
        if (const char *e = std::getenv("PDMK_C2P_WIDE")) c2p_wide = e[0] != '0';
        // grouped shifts at p >= 28 (N=1e6 9 digits: evaluate() 7.57 -> 7.39 ms; clustered
        // N=4e6 132.5 -> 131.1 ms at 9 digits, 412.5 -> 403.1 ms at 12); at p = 18 one box
        // per workgroup stays faster (6 digits: FP32 2.41 against 2.45 ms, FP64 3.87 against
        // 3.95 ms)
        shift_groups = p >= 28;
        if (const char *e = std::getenv("PDMK_SHIFT_GROUPS")) shift_groups = e[0] != '0';
        if (pw_by_level) {
            pw_out.alloc(std::max(pw_slot * (size_t)lvl_cnt, c2p_bytes));
        } else {
            pw_out.alloc(pw_slot * (size_t)std::max(1, n_up));
            c2p_scratch.alloc(c2p_bytes);
        }
    }

    // ---------------- statistics and exact near-field pair count
    stats.n_src = n; stats.nsall = nsall; stats.nboxes = nboxes; stats.n_levels = nlevels; stats.p = p;
    stats.n_pw = n_pw; stats.ms = ms; stats.mt = mt; stats.mu = mu; stats.n_direct_leaves = n_nf_leaves;
    stats.n_far_modes = n_ff_modes; stats.n_digits = n_digits; stats.is_float = sizeof(Real) == 4;
    for (int b = 0; b < nboxes; ++b) stats.n_leaves += t.leaf[b];
    work.nf_flops = work.nf_pairs_support * (2.0 * nc + 12.0);
    work.on_bytes = (double)nsall * (sizeof(int) + sizeof(Real) + sizeof(double)) +
                    (double)n * (2 * sizeof(Real) + 2 * sizeof(double) + sizeof(int) + sizeof(double) + sizeof(Real));

    // ---------------- topology export
    box7.resize(7 * (size_t)nboxes);
    box_centers.resize(3 * (size_t)nboxes);
    for (int b = 0; b < nboxes; ++b) {
        int v[7] = {t.level[b], t.parent[b], t.leaf[b] ? 0 : 8, t.src_count[b], t.trg_count[b], t.pwexp[b],
                    t.l1_begin[b + 1] - t.l1_begin[b]};
        for (int k = 0; k < 7; ++k) box7[7 * (size_t)b + k] = v[k];
        for (int d = 0; d < 3; ++d) box_centers[3 * (size_t)b + d] = t.center[3 * b + d];
    }
    src_perm_host = t.src_order;
    trg_perm_host = t.trg_order;
    check(hipStreamSynchronize(stream), "plan build");
    {
        // launch shapes of an earlier tuned plan of the same kind
        std::lock_guard<std::mutex> lock(tune_mu);
        tune_cache_load();
        auto it = tune_cache.find(tune_key());
        if (it != tune_cache.end()) {
            knobs = it->second;
            knobs_cached = true;
        }
    }
    build_ms[0] = ms_since(t_start);
    guard.armed = false;
}

template <typename Real>
void Plan<Real>::release_handles() {
    for (hipStream_t st : {stream, s_far, s_near})
        if (st) (void)hipStreamSynchronize(st);
    for (auto e : kev) (void)hipEventDestroy(e);
    kev.clear();
    for (auto &e : ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (hipEvent_t *e : {&e_in, &e_gather, &e_far, &e_near, &e_up}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    for (auto &e : many_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    for (hipStream_t *st : {&stream, &s_far, &s_near}) {
        if (*st) (void)hipStreamDestroy(*st);
        *st = nullptr;
    }
}

template <typename Real>
Plan<Real>::~Plan() {
    release_handles();
}

namespace {

// The arrays one pass of the proxy / plane-wave chain works on: the plan's own (one vector,
// charges in src[].w), or nv vectors' copies, vector v at v times the element strides, with
// charges qv[source * qv_ld + v] (the kernels take v from blockIdx.z).
template <class Real> struct ChainBufs {
    Real *up, *c2p, *down, *pot;
    Cx<Real> *pw, *pin;
    const Real *qv = nullptr;
    int qv_ld = 0, nv = 1;
    long long s_up = 0, s_c2p = 0, s_pw = 0, s_pin = 0, s_down = 0, s_pot = 0;
};

template <class Real> ChainBufs<Real> own_chain(Plan<Real> &pl) {
    ChainBufs<Real> b;
    b.up = pl.proxy_up.template as<Real>();
    b.c2p = pl.c2p_buffer();
    b.down = pl.proxy_down.template as<Real>();
    b.pot = pl.pot_sorted.template as<Real>();
    b.pw = pl.pw_out.template as<Cx<Real>>();
    b.pin = pl.pin.template as<Cx<Real>>();
    return b;
}

// c2p / p2c of one level: output planes are split into slices so that a level with few
// boxes still spreads over the whole GPU (about 1024 workgroups or more); nv vectors along z
template <class Real, int P, bool TRANSPOSE>
void launch_tensor3(const Real *in, Real *out, typename Plan<Real>::Tensor3Launch &L, const Real *mats_rm,
                    hipStream_t s, int div, bool mfma, int nv = 1, long long vs_in = 0, long long vs_out = 0) {
    if (L.n_items == 0) return;
    // FP64 at p = 28 on the matrix cores (clustered N = 4e6, 9 digits: child -> parent 22.3 ->
    // 21.1 ms, parent -> child 12.6 -> 10.5 ms per evaluate).  At p = 38 the matrix-core kernel's
    // a-slices of the input re-read each box several times and ran 1.8x slower (168.8 against
    // 94.4 ms); p = 38 stays on the vector ALUs.
    if constexpr (std::is_same_v<Real, double> && P == 28) {
        if (mfma) {
            hipLaunchKernelGGL((k_tensor3_mfma<P, TRANSPOSE>), dim3(L.n_items * T3MfmaCfg<P>::CT, 1, nv), dim3(64 * T3MfmaCfg<P>::NW), 0, s,
                               in, out, L.out.template as<int>(), L.add.template as<int>(), L.ptr.template as<int>(),
                               L.in_slot.template as<int>(), L.in_code.template as<int>(), mats_rm, L.n_items, vs_in,
                               vs_out);
            check(hipGetLastError(), "k_tensor3_mfma");
            return;
        }
    }
    constexpr int KC = T3Cfg<Real, P>::KC;
    int kc = std::max(1, KC / std::max(1, div));
    while (kc > 1 && (long)L.n_items * ((P + kc - 1) / kc) < 1024) kc = (kc + 1) / 2;
    const dim3 grid(L.n_items * ((P + kc - 1) / kc), 1, nv);  // (item, slice) pairs
    hipLaunchKernelGGL((k_tensor3<Real, P, TRANSPOSE>), grid, dim3(T3Cfg<Real, P>::threads), 0, s, in, out,
                       L.out.template as<int>(), L.add.template as<int>(), L.ptr.template as<int>(),
                       L.in_slot.template as<int>(), L.in_code.template as<int>(), mats_rm, L.n_items, kc, vs_in,
                       vs_out);
    check(hipGetLastError(), "k_tensor3");
}

template <class Real, int P, int NPW>
void launch_transforms(Plan<Real> &pl, hipStream_t s, const ChainBufs<Real> &B) {
    using C = Cx<Real>;
    const int NM = NPW * NPW * ((NPW + 1) / 2);
    const int nv = B.nv;
    if (nv < 1 || nv > 65535) throw std::invalid_argument("launch_transforms: vector count");
    if (pl.n_up > 0) {
        // FP64 at P = 28 (9 digits): matrix cores (a padded to 32 costs 1/8).  At P = 38 (a
        // padded to 48, each box's row tiles split over two workgroups) the matrix-core kernel
        // ran slower than the vector ALUs (clustered N = 4e6: 90.5 against 78.7 ms).  FP64 at
        // P = 18 pads too much.  FP32 at P = 18 on the matrix cores is faster alone (395 -> 307 us at
        // N=1e6) but slower inside evaluate() (3.02 -> 3.24 ms): its 8-wave workgroups take
        // wave slots from the near field beside it; it stays on the vector ALUs (and the
        // many-vector chain with it: the same sums in the same order as one vector's).
        if constexpr (std::is_same_v<Real, double> && P >= 28 && P <= 32) {
            if (nv != 1 || B.qv) throw std::logic_error("launch_transforms: one vector on the matrix-core kernels");
            auto c2p = [&](auto wide) {
                constexpr bool W = decltype(wide)::value;
                hipLaunchKernelGGL((k_charge2proxy_mfma<Real, P, W>), dim3(pl.n_up * C2pMfmaCfg<P, W>::RS),
                                   dim3(64 * C2pMfmaCfg<P, W>::NW), 0, s, pl.src.template as<P4<Real>>(),
                                   pl.up_slot.template as<int>(), pl.up_geom.template as<Real>(),
                                   pl.up_rng.template as<int>(), pl.rng_begin.template as<int>(),
                                   pl.rng_count.template as<int>(), pl.rng_off.template as<Real>(), B.up);
                check(hipGetLastError(), "k_charge2proxy_mfma");
            };
            if (pl.c2p_wide)
                c2p(std::true_type{});
            else
                c2p(std::false_type{});
        } else {
            hipLaunchKernelGGL((k_charge2proxy<Real, P>), dim3(pl.n_up, 1, nv), dim3(C2pCfg<P>::threads), 0, s,
                               pl.src.template as<P4<Real>>(), pl.up_slot.template as<int>(), pl.up_geom.template as<Real>(),
                               pl.up_rng.template as<int>(), pl.rng_begin.template as<int>(), pl.rng_count.template as<int>(),
                               pl.rng_off.template as<Real>(), B.up, B.qv, B.qv_ld, B.s_up);
            check(hipGetLastError(), "k_charge2proxy");
        }
    }
    for (auto &C : pl.c2p_levels) {
        if (C.fused) {
            launch_tensor3<Real, P, true>(B.up, B.up, C.t3, pl.mats_rm.template as<Real>(), s, pl.knobs.t3_div,
                                          pl.t3_mfma, nv, B.s_up, B.s_up);
            continue;
        }
        launch_tensor3<Real, P, true>(B.up, B.c2p, C.t3, pl.mats_rm.template as<Real>(), s, pl.knobs.t3_div, pl.t3_mfma,
                                      nv, B.s_up, B.s_c2p);
        hipLaunchKernelGGL((k_sum_children<Real, P>), dim3(C.n_parents, (P * P * P + 255) / 256, nv), dim3(256), 0, s,
                           B.up, B.c2p, C.parent_slot.template as<int>(), C.child_ptr.template as<int>(), C.n_parents,
                           B.s_up, B.s_c2p);
        check(hipGetLastError(), "k_sum_children");
    }
    // Where a box's slices would each re-read its input kTwoPassSlices times or more, the first
    // contraction of whole boxes goes to HBM first (k_p2pw_z, k_pw2p_x; the same sums), in
    // chunks of boxes that fit the plan's two-pass scratch.
    auto proxy2pw_bk = [&](auto bk, int slot0, int count) {
        constexpr int BK = decltype(bk)::value;
        if constexpr (sizeof(Real) == 8 && P >= 28 && P2pwCfg<Real, P, NPW, BK>::slices >= kTwoPassSlices) {
            if (pl.tp_scratch.p && nv == 1) {
                C *t1 = pl.tp_scratch.template as<C>();
                for (int c0 = 0; c0 < count; c0 += pl.tp_boxes_p2pw) {
                    const int m = std::min(pl.tp_boxes_p2pw, count - c0);
                    hipLaunchKernelGGL((k_p2pw_z<Real, P, NPW>), dim3(m * ((P * P + 255) / 256)), dim3(256), 0, s, B.up, t1,
                                       pl.poly2pw.template as<C>(), slot0 + c0, m);
                    check(hipGetLastError(), "k_p2pw_z");
                    hipLaunchKernelGGL((k_p2pw_yx<P, NPW>), dim3(m * (((NPW + 1) / 2 + 3) / 4)), dim3(256), 0, s,
                                       (const Cx<double> *)t1,
                                       pl.poly2pw_f.template as<double>(), (Cx<double> *)(B.pw + (size_t)c0 * NM), m);
                    check(hipGetLastError(), "k_p2pw_yx");
                }
                return;
            }
        }
        hipLaunchKernelGGL((k_proxy2pw<Real, P, NPW, BK>), dim3(count * P2pwCfg<Real, P, NPW, BK>::slices, 1, nv),
                           dim3(256), 0, s, B.up, B.pw, pl.poly2pw.template as<C>(), slot0, count, B.s_up, B.s_pw);
        check(hipGetLastError(), "k_proxy2pw");
    };
    auto proxy2pw = [&](int slot0, int count) {
        switch (pl.knobs.p2pw_bk) {
        case 16: proxy2pw_bk(std::integral_constant<int, 16>{}, slot0, count); break;
        case 32: proxy2pw_bk(std::integral_constant<int, 32>{}, slot0, count); break;
        case 64: proxy2pw_bk(std::integral_constant<int, 64>{}, slot0, count); break;
        default: proxy2pw_bk(std::integral_constant<int, 0>{}, slot0, count); break;
        }
    };
    auto pw2proxy_bk = [&](auto bk, const typename Plan<Real>::EvalLevel &E, int i0, int nb) {
        constexpr int BK = decltype(bk)::value;
        if constexpr (sizeof(Real) == 8 && P >= 28 && Pw2pCfg<Real, P, NPW, BK>::slices >= kTwoPassSlices) {
            if (pl.tp_scratch.p && nv == 1) {
                C *u1 = pl.tp_scratch.template as<C>();
                constexpr int COL = NPW * ((NPW + 1) / 2);
                for (int c0 = 0; c0 < nb; c0 += pl.tp_boxes_pw2p) {
                    const int m = std::min(pl.tp_boxes_pw2p, nb - c0);
                    hipLaunchKernelGGL((k_pw2p_x<Real, P, NPW>), dim3(m * ((COL + 255) / 256)), dim3(256), 0, s,
                                       B.pin + (size_t)c0 * NM, pl.pw2poly.template as<C>(), m, u1);
                    check(hipGetLastError(), "k_pw2p_x");
                    hipLaunchKernelGGL((k_pw2p_yz<P, NPW>), dim3(m * ((P + 3) / 4)), dim3(256), 0, s, (const Cx<double> *)u1,
                                       pl.pw2poly_f.template as<double>(), E.down.template as<int>(),
                                       E.zero.template as<int>(), i0 + c0, m, (double *)B.down);
                    check(hipGetLastError(), "k_pw2p_yz");
                }
                return;
            }
        }
        hipLaunchKernelGGL((k_pw2proxy<Real, P, NPW, BK>), dim3(nb * Pw2pCfg<Real, P, NPW, BK>::slices, 1, nv), dim3(256),
                           0, s, B.pin, pl.pw2poly.template as<C>(), E.down.template as<int>(),
                           E.zero.template as<int>(), i0, nb, B.down, B.s_pin, B.s_down);
        check(hipGetLastError(), "k_pw2proxy");
    };
    auto pw2proxy = [&](const typename Plan<Real>::EvalLevel &E, int i0, int nb) {
        switch (pl.knobs.pw2p_bk) {
        case 16: pw2proxy_bk(std::integral_constant<int, 16>{}, E, i0, nb); break;
        case 32: pw2proxy_bk(std::integral_constant<int, 32>{}, E, i0, nb); break;
        case 64: pw2proxy_bk(std::integral_constant<int, 64>{}, E, i0, nb); break;
        default: pw2proxy_bk(std::integral_constant<int, 0>{}, E, i0, nb); break;
        }
    };
    if (pl.n_up > 0 && !pl.pw_by_level) proxy2pw(0, pl.n_up);
    if (pl.e_upward) check(hipEventRecord(pl.e_upward, s), "hipEventRecord");
    for (size_t l = 0; l < pl.eval_levels.size(); ++l) {
        auto &E = pl.eval_levels[l];
        const Real lscale = (Real)std::ldexp(1.0, E.level);
        if (pl.pw_by_level && E.pw_cnt > 0) proxy2pw(E.pw_lo, E.pw_cnt);
        const int slot0 = pl.pw_by_level ? E.pw_lo : 0;
        for (int i0 = 0; i0 < E.n_items; i0 += pl.pin_batch) {
            const int nb = std::min(pl.pin_batch, E.n_items - i0);
            if (pl.shift_groups) {
                const int ng = (i0 + nb + kShiftGroup - 1) / kShiftGroup - i0 / kShiftGroup;
                hipLaunchKernelGGL(k_pw_shift_grp<Real>, dim3(ng, (NM + 255) / 256, nv), dim3(256), 0, s, B.pw,
                                   pl.shift0.template as<C>(), pl.shift_r.template as<double>(),
                                   pl.shift_e.template as<Cx<double>>(), lscale, E.grp_ptr.template as<int>(),
                                   E.grp_slot.template as<int>(), E.grp_sten.template as<unsigned long long>(), slot0, NM,
                                   i0, nb, NPW, (NPW + 1) / 2, B.pin, B.s_pw, B.s_pin);
            } else if constexpr (sizeof(Real) == 8) {
                hipLaunchKernelGGL(k_pw_shift_sep, dim3(nb, (NM + 255) / 256, nv), dim3(256), 0, s,
                                   (const Cx<double> *)B.pw, pl.shift_r.template as<double>(),
                                   pl.shift_e.template as<Cx<double>>(), (double)lscale, E.nbr_ptr.template as<int>(),
                                   E.nbr_slot.template as<int>(), E.nbr_sten.template as<int>(), slot0, NM, i0, NPW,
                                   (NPW + 1) / 2, (Cx<double> *)B.pin, B.s_pw, B.s_pin);
            } else {
                hipLaunchKernelGGL(k_pw_shift<Real>, dim3(nb, (NM + 255) / 256, nv), dim3(256), 0, s, B.pw,
                                   pl.shift0.template as<C>(), lscale, E.nbr_ptr.template as<int>(),
                                   E.nbr_slot.template as<int>(), E.nbr_sten.template as<int>(), slot0, NM, i0, B.pin,
                                   B.s_pw, B.s_pin);
            }
            check(hipGetLastError(), "k_pw_shift");
            pw2proxy(E, i0, nb);
        }
        auto &L = pl.p2c_levels[l];
        launch_tensor3<Real, P, false>(B.down, B.down, L, pl.mats_rm.template as<Real>(), s, pl.knobs.t3_div,
                                       pl.t3_mfma, nv, B.s_down, B.s_down);
    }
    if (pl.n_leaf_eval > 0) {
        if constexpr (std::is_same_v<Real, double> && P >= 28) {
            hipLaunchKernelGGL((k_eval_targets_mfma<P>), dim3(pl.n_leaf_eval), dim3(64 * EvMfmaCfg<P>::NW), 0, s, B.down,
                               pl.leaf_down.template as<int>(), pl.leaf_geom.template as<Real>(),
                               pl.leaf_trg.template as<int>(), pl.trg.template as<P4<Real>>(), B.pot);
            check(hipGetLastError(), "k_eval_targets_mfma");
        } else {
            hipLaunchKernelGGL((k_eval_targets<Real, P>), dim3(pl.n_leaf_eval, 1, nv), dim3(EV_THREADS), 0, s, B.down,
                               pl.leaf_down.template as<int>(), pl.leaf_geom.template as<Real>(),
                               pl.leaf_trg.template as<int>(), pl.trg.template as<P4<Real>>(), B.pot, B.s_down, B.s_pot);
            check(hipGetLastError(), "k_eval_targets");
        }
    }
}

// wpb = wavefronts (items) per workgroup: 4 when the near field runs beside the whole
// transform chain; 16 when it starts late beside the long FP64 downward pass, where whole
// 16-wave workgroups retiring at once leave room for the chain's large workgroups.
template <class Real, int NC, bool SPLIT>
void launch_near(Plan<Real> &pl, hipStream_t s, int wpb, int i0, int i1) {
    if (i1 <= i0) return;
    hipLaunchKernelGGL((k_near_field<Real, NC, SPLIT>), dim3((i1 - i0 + wpb - 1) / wpb), dim3(64 * wpb), 0, s,
                       pl.src.template as<P4<Real>>(), pl.trg.template as<P4<Real>>(), pl.nf_items.template as<int4>(),
                       pl.nf_ptr.template as<int>(), pl.nf_l1.template as<NfEntry>(), pl.nf_geom.template as<P4<Real>>(),
                       pl.nf_ie_base.template as<int>(), pl.nf_ie_ptr.template as<long long>(), pl.nf_steps.template as<int>(),
                       pl.coeffs.template as<Real>(), pl.pot_near.template as<Real>(), i0, i1);
    check(hipGetLastError(), "k_near_field");
}

// quad-row near field (FP32 plans built with row step lists)
template <class Real>
void launch_near_rows(Plan<Real> &pl, hipStream_t s, int wpb, int i0, int i1) {
    if constexpr (std::is_same<Real, float>::value) {
        if (i1 <= i0) return;
        // (the LDS tables of a workgroup's wavefronts stay within 64 KiB)
        const size_t per_wave = (size_t)(pl.nf_ne_max + 1) * 6 * sizeof(float4);
        while (wpb > 1 && (size_t)wpb * per_wave > 65536) wpb /= 2;
        const size_t lds = (size_t)wpb * per_wave;
        hipLaunchKernelGGL((k_near_rows<12>), dim3((i1 - i0 + wpb - 1) / wpb), dim3(64 * wpb), lds, s,
                           pl.src.template as<P4<float>>(), pl.trg.template as<P4<float>>(),
                           pl.nf_items.template as<int4>(), pl.nf_ptr.template as<int>(),
                           pl.nf_geom.template as<P4<float>>(), pl.nf_rs_base.template as<int>(),
                           pl.nf_rslot.template as<int>(), pl.nf_wave_base.template as<long long>(),
                           pl.nf_row_steps.template as<int>(), pl.coeffs.template as<float>(),
                           pl.pot_near.template as<float>(), i0, i1, pl.nf_ne_max);
        check(hipGetLastError(), "k_near_rows");
    } else {
        if (i1 <= i0) return;
        const size_t per_wave = near_rows64_lds_per_wave(pl.nf_ne_max);
        while (wpb > 1 && (size_t)wpb * per_wave > 65536) wpb /= 2;
        const size_t lds = (size_t)wpb * per_wave;
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3((i1 - i0 + wpb - 1) / wpb), dim3(64 * wpb), lds, s,
                               pl.src.template as<P4<double>>(), pl.trg.template as<P4<double>>(),
                               pl.nf_items.template as<int4>(), pl.nf_ptr.template as<int>(),
                               pl.nf_geom.template as<P4<double>>(), pl.nf_rs_base.template as<int>(),
                               pl.nf_rslot.template as<int>(), pl.nf_wave_base.template as<long long>(),
                               pl.nf_row_steps.template as<int>(), pl.coeffs.template as<double>(),
                               pl.pot_near.template as<double>(), i0, i1, pl.nf_ne_max);
        };
        switch (pl.nc) {
        case 12: go(k_near_rows64<12>); break;
        case 18: go(k_near_rows64<18>); break;
        case 25: go(k_near_rows64<25>); break;
        default: throw std::logic_error("row near field: unsupported coefficient count");
        }
        check(hipGetLastError(), "k_near_rows64");
    }
}

}  // namespace

template <typename Real>
void Plan<Real>::evaluate(const double *d_q, Real *d_pot, hipStream_t s, int stages, double *times_ms) {
    if (!s) s = stream;
    hipEvent_t *K = nullptr;
    if (kslots > 0) K = &kev[(size_t)(kcall++ % kslots) * 2 * kGroups];
    auto mark = [&](int g, int end, hipStream_t st) {
        if (K) check(hipEventRecord(K[2 * g + end], st), "hipEventRecord");
    };
    // evaluate_timed's split (far field, proxy/plane waves, near field) needs the stages one
    // after the other; otherwise the far field and the near field run on their own streams
    // beside the transform chain and join before the final combination
    const bool serial = times_ms != nullptr || single_stream;
    hipStream_t sf = serial ? s : s_far, sn = serial ? s : s_near;
    if (!serial) {
        check(hipEventRecord(e_in, s), "hipEventRecord");
        check(hipStreamWaitEvent(sf, e_in, 0), "hipStreamWaitEvent");
    }
    if (times_ms) check(hipEventRecord(ev[0], s), "hipEventRecord");
    // far field: depends only on the user-order charges
    mark(0, 0, sf);
    if (stages & 8) {
        if (far_nufft) nf.run(d_q, far.as<double>(), sf);
        else ff.run(d_q, far.as<double>(), sf);
    }
    mark(0, 1, sf);
    if (!serial) check(hipEventRecord(e_far, sf), "hipEventRecord");
    if (times_ms) check(hipEventRecord(ev[1], s), "hipEventRecord");
    mark(1, 0, s);
    {
        const int m = std::max(nsall, n);
        hipLaunchKernelGGL(k_gather_charges<Real>, dim3((m + 255) / 256), dim3(256), 0, s, d_q, src_user.as<int>(),
                           src.as<P4<Real>>(), nsall, pot_sorted.as<Real>(), pot_near.as<Real>(), n,
                           qp_comp >= 0 ? src_phase.as<double2>() : (const double2 *)nullptr, qp_comp < 0 ? 0 : qp_comp);
        check(hipGetLastError(), "k_gather_charges");
    }
    mark(1, 1, s);
    if (!serial) {
        check(hipEventRecord(e_gather, s), "hipEventRecord");
        check(hipStreamWaitEvent(sn, e_gather, 0), "hipStreamWaitEvent");
    }
    const bool near_late = !serial && (knobs.near_late >= 0 ? knobs.near_late == 1 : p >= 28);
    const int wpb = knobs.near_wpb > 0 ? knobs.near_wpb : (near_late ? 16 : 4);
    auto near_field = [&](int i0, int i1) {
        mark(3, 0, sn);
        if (stages & 2) {
            switch (n_digits) {
            case 1: case 2: launch_near<Real, 5, true>(*this, sn, wpb, i0, i1); break;
            case 3: launch_near<Real, 6, true>(*this, sn, wpb, i0, i1); break;
            case 4: launch_near<Real, 9, true>(*this, sn, wpb, i0, i1); break;
            case 5: launch_near<Real, 11, true>(*this, sn, wpb, i0, i1); break;
            case 6:
                if (nf_rows) launch_near_rows<Real>(*this, sn, wpb, i0, i1);
                else launch_near<Real, 12, false>(*this, sn, wpb, i0, i1);
                break;
            case 7: launch_near<Real, 15, false>(*this, sn, wpb, i0, i1); break;
            case 8: launch_near<Real, 17, false>(*this, sn, wpb, i0, i1); break;
            case 9:
                if (nf_rows) launch_near_rows<Real>(*this, sn, wpb, i0, i1);
                else launch_near<Real, 18, false>(*this, sn, wpb, i0, i1);
                break;
            case 10: launch_near<Real, 22, false>(*this, sn, wpb, i0, i1); break;
            case 11: launch_near<Real, 23, false>(*this, sn, wpb, i0, i1); break;
            case 12:
                if (nf_rows) launch_near_rows<Real>(*this, sn, wpb, i0, i1);
                else launch_near<Real, 25, false>(*this, sn, wpb, i0, i1);
                break;
            default: throw std::logic_error("unsupported digit count");
            }
        }
        mark(3, 1, sn);
    };
    // Start of the near field: at once when its cost is comparable to the transform chain's
    // (p < 28); at p >= 28 the chain is several times longer and its first kernels (proxy
    // formation on the matrix cores, proxy -> plane wave) fill the GPU, so the near field
    // starts after them and overlaps the rest.  (Measured at N=1e6: a later start, or the near
    // field cut into two launches, is slower at 6 digits; the late start is faster at 9; a
    // start after the whole chain was slower at 6 and 9 digits, FP32 and FP64.)
    e_upward = near_late ? e_up : nullptr;
    if (!serial && !near_late) near_field(0, n_nf_items);
    mark(2, 0, s);
    if (stages & 1) {
        switch (p) {
        case 9: launch_transforms<Real, 9, 13>(*this, s, own_chain(*this)); break;
        case 18: launch_transforms<Real, 18, 27>(*this, s, own_chain(*this)); break;
        case 28: launch_transforms<Real, 28, 39>(*this, s, own_chain(*this)); break;
        case 38: launch_transforms<Real, 38, 55>(*this, s, own_chain(*this)); break;
        default: throw std::logic_error("unsupported proxy order");
        }
    }
    mark(2, 1, s);
    if (times_ms) check(hipEventRecord(ev[2], s), "hipEventRecord");
    if (serial) near_field(0, n_nf_items);
    if (near_late) {
        check(hipStreamWaitEvent(sn, e_up, 0), "hipStreamWaitEvent");
        near_field(0, n_nf_items);
    }
    if (!serial) {
        check(hipEventRecord(e_near, sn), "hipEventRecord");
        check(hipStreamWaitEvent(s, e_near, 0), "hipStreamWaitEvent");
        check(hipStreamWaitEvent(s, e_far, 0), "hipStreamWaitEvent");
    }
    mark(4, 0, s);
    hipLaunchKernelGGL(k_finalize<Real>, dim3((n + 255) / 256), dim3(256), 0, s, pot_sorted.as<Real>(), pot_near.as<Real>(),
                       trg_w0.as<double>(), trg_user.as<int>(), d_q, far.as<double>(), far_idx.as<int>(), d_pot, n,
                       (stages & 4) ? 1 : 0, (stages & 8) ? 1 : 0, qp_comp >= 0 ? 2 : 1);
    check(hipGetLastError(), "k_finalize");
    mark(4, 1, s);
    if (times_ms) {
        check(hipEventRecord(ev[3], s), "hipEventRecord");
        check(hipEventSynchronize(ev[3]), "hipEventSynchronize");
        float a = 0, b = 0, c = 0;
        check(hipEventElapsedTime(&a, ev[0], ev[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&b, ev[1], ev[2]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&c, ev[2], ev[3]), "hipEventElapsedTime");
        times_ms[0] = a;  // nufft_ms
        times_ms[1] = b;  // dmk_ms
        times_ms[2] = c;  // direct_ms
    }
}

template <typename Real>
bool Plan<Real>::many_ensure() {
    if (many_vc_chain != 0) return many_vc_chain > 0;
    constexpr int VC = kManyVec;
    // per-vector bytes of the chain copies and of the exact far field's partial sums
    const size_t chain_v = proxy_up.bytes + (pw_by_level ? 0 : c2p_scratch.bytes) + pw_out.bytes + pin.bytes +
                           proxy_down.bytes;
    const size_t sp_v = far_nufft ? 0 : sizeof(double2) * ff.sp_elems();
    // at most 256 MiB of chain copies and 128 MiB of partial sums per plan: a large tree
    // already fills the GPU with one vector per launch
    many_vc_chain = (int)std::max<size_t>(1, std::min<size_t>(VC, (size_t(256) << 20) / std::max<size_t>(1, chain_v)));
    many_vc_far = far_nufft ? 0 : (int)std::max<size_t>(1, std::min<size_t>(VC, (size_t(128) << 20) / std::max<size_t>(1, sp_v)));
    const size_t pass = 2 * ((size_t)nsall * VC * sizeof(float) + (size_t)n * VC * (sizeof(float) + sizeof(double)));
    const size_t total = pass + chain_v * many_vc_chain + (size_t)n * VC * sizeof(Real) +
                         (far_nufft ? 0 : sp_v * many_vc_far + sizeof(double2) * ff.s_elems() * VC);
    size_t free_b = 0, total_b = 0;
    check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (pw_by_level || total > free_b / 4) {
        many_vc_chain = many_vc_far = -1;
        return false;
    }
    for (int b = 0; b < 2; ++b) {
        // charges up to the sentinel source of the row lists' padding (zero past nsall)
        many_qv[b].alloc((size_t)(nsall + NF_CHUNK + 1) * VC * sizeof(float));
        check(hipMemset(many_qv[b].p, 0, many_qv[b].bytes), "hipMemset");
        many_near[b].alloc((size_t)n * VC * sizeof(float));
        many_far[b].alloc((size_t)n * VC * sizeof(double));
    }
    many_up.alloc(proxy_up.bytes * many_vc_chain);
    if (!pw_by_level) many_c2p.alloc(c2p_scratch.bytes * many_vc_chain);
    many_pw.alloc(pw_out.bytes * many_vc_chain);
    many_pin.alloc(pin.bytes * many_vc_chain);
    many_down.alloc(proxy_down.bytes * many_vc_chain);
    many_pot.alloc((size_t)n * VC * sizeof(Real));
    if (!far_nufft) {
        many_sp.alloc(sp_v * many_vc_far);
        many_s.alloc(sizeof(double2) * ff.s_elems() * VC);
        check(hipMemset(many_s.p, 0, many_s.bytes), "hipMemset");  // padding of the weighted sums stays zero
    }
    for (auto &e : many_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    return true;
}

template <typename Real>
void Plan<Real>::evaluate_many(int n_vec, const double *d_Q, Real *d_P, hipStream_t s) {
    if (n_vec < 0) throw std::invalid_argument("evaluate_many: negative vector count");
    if (!s) s = stream;
    if (qp) throw std::logic_error("evaluate_many: quasi-periodic tree");
    constexpr int VC = kManyVec;
    const bool batched = std::is_same<Real, float>::value && nf_rows && n_digits == 6 && n_nf_items > 0 && p == 18;
    if (!batched || n_vec < 2 || (size_t)n * VC > (size_t)INT_MAX || !many_ensure()) {
        for (int v = 0; v < n_vec; ++v) evaluate(d_Q + (size_t)v * n, d_P + (size_t)v * n, s);
        return;
    }
    if constexpr (std::is_same<Real, float>::value) {
        const size_t per_wave = (size_t)(nf_ne_max + 1) * 6 * sizeof(float4);
        int wpb = 4;
        while (wpb > 1 && (size_t)wpb * per_wave > 65536) wpb /= 2;
        // the chain copies: vector c of a chain launch at c times these element strides
        ChainBufs<float> cb;
        cb.up = many_up.as<float>();
        cb.c2p = many_c2p.as<float>();
        cb.down = many_down.as<float>();
        cb.pw = many_pw.as<Cx<float>>();
        cb.pin = many_pin.as<Cx<float>>();
        cb.qv_ld = VC;
        cb.s_up = (long long)(proxy_up.bytes / sizeof(float));
        cb.s_c2p = (long long)(c2p_scratch.bytes / sizeof(float));
        cb.s_pw = (long long)(pw_out.bytes / sizeof(Cx<float>));
        cb.s_pin = (long long)(pin.bytes / sizeof(Cx<float>));
        cb.s_down = (long long)(proxy_down.bytes / sizeof(float));
        cb.s_pot = n;
        hipEvent_t *e_gat = many_ev, *e_nr = many_ev + 2, *e_fr = many_ev + 4, *e_fin = many_ev + 6;
        check(hipEventRecord(e_in, s), "hipEventRecord");
        check(hipStreamWaitEvent(s_near, e_in, 0), "hipStreamWaitEvent");
        check(hipStreamWaitEvent(s_far, e_in, 0), "hipStreamWaitEvent");
        e_upward = nullptr;
        for (int v0 = 0, pass = 0; v0 < n_vec; v0 += VC, ++pass) {
            const int vc = std::min(VC, n_vec - v0), b = pass & 1;
            if (pass >= 2) {
                // slot b's buffers are free once pass - 2 has finished on s
                check(hipStreamWaitEvent(s_near, e_fin[b], 0), "hipStreamWaitEvent");
                check(hipStreamWaitEvent(s_far, e_fin[b], 0), "hipStreamWaitEvent");
            }
            // near field of the pass's vectors on s_near
            hipLaunchKernelGGL(k_gather_many<VC>, dim3((nsall + 255) / 256), dim3(256), 0, s_near, d_Q, n_vec, v0, n,
                               src_user.as<int>(), nsall, many_qv[b].as<float>());
            check(hipGetLastError(), "k_gather_many");
            check(hipEventRecord(e_gat[b], s_near), "hipEventRecord");
            hipLaunchKernelGGL((k_near_rows_many<12, VC>), dim3((n_nf_items + wpb - 1) / wpb), dim3(64 * wpb),
                               (size_t)wpb * per_wave, s_near, src.as<P4<float>>(), many_qv[b].as<float>(),
                               trg.as<P4<float>>(), nf_items.as<int4>(), nf_ptr.as<int>(), nf_geom.as<P4<float>>(),
                               nf_rs_base.as<int>(), nf_rslot.as<int>(), nf_wave_base.as<long long>(),
                               nf_row_steps.as<int>(), coeffs.as<float>(), many_near[b].as<float>(), n, vc, 0,
                               n_nf_items, nf_ne_max);
            check(hipGetLastError(), "k_near_rows_many");
            check(hipEventRecord(e_nr[b], s_near), "hipEventRecord");
            // far fields on s_far
            const double *Q0 = d_Q + (size_t)v0 * n;
            double *far0 = many_far[b].as<double>();
            if (far_nufft) {
                for (int c = 0; c < vc; ++c) nf.run(Q0 + (size_t)c * n, far0 + (size_t)c * n, s_far);
            } else {
                ff.run_many(Q0, n, vc, far0, n, many_sp.as<double2>(), many_vc_far, many_s.as<double2>(), s_far);
            }
            check(hipEventRecord(e_fr[b], s_far), "hipEventRecord");
            // the chain on s, many_vc_chain vectors per launch, charges from the gathered copy
            check(hipStreamWaitEvent(s, e_gat[b], 0), "hipStreamWaitEvent");
            hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)(((size_t)n * vc + 255) / 256)), dim3(256), 0, s,
                               many_pot.as<float>(), 0.0f, n * vc);
            check(hipGetLastError(), "k_fill");
            for (int c0 = 0; c0 < vc; c0 += many_vc_chain) {
                ChainBufs<float> B = cb;
                B.nv = std::min(many_vc_chain, vc - c0);
                B.qv = many_qv[b].as<float>() + c0;
                B.pot = many_pot.as<float>() + (size_t)c0 * n;
                launch_transforms<float, 18, 27>(*this, s, B);
            }
            // potentials of the pass once its near and far fields are in
            check(hipStreamWaitEvent(s, e_nr[b], 0), "hipStreamWaitEvent");
            check(hipStreamWaitEvent(s, e_fr[b], 0), "hipStreamWaitEvent");
            hipLaunchKernelGGL(k_finalize<float>, dim3((n + 255) / 256, 1, vc), dim3(256), 0, s, many_pot.as<float>(),
                               many_near[b].as<float>(), trg_w0.as<double>(), trg_user.as<int>(), Q0, far0,
                               far_idx.as<int>(), d_P + (size_t)v0 * n, n, 1, 1, 1, (long long)n, (long long)n,
                               (long long)n, (long long)n, (long long)n);
            check(hipGetLastError(), "k_finalize");
            check(hipEventRecord(e_fin[b], s), "hipEventRecord");
        }
    }
}

namespace {

// Charges of the tuning runs: a fixed neutral-ish pattern (the timings do not depend on them).
__global__ void k_tune_charges(double *q, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = (double)((unsigned)(i * 2654435761u) >> 8) * (1.0 / 16777216.0) - 0.5;
}

}  // namespace

void tune_cache_clear() {
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_cache.clear();
    tune_file_read = true;  // the file is not read again
}

template <typename Real>
std::string Plan<Real>::tune_key() const {
    int dev = 0;
    hipDeviceProp_t prop{};
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    return std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + ":p" + std::to_string(p) +
           ":r" + std::to_string(sizeof(Real)) + ":n" + std::to_string(log2_bucket(n)) + ":b" +
           std::to_string(log2_bucket((long long)n_up + 1)) + ":w" + std::to_string(log2_bucket((long long)n_nf_items + 1));
}

template <typename Real>
void Plan<Real>::autotune(int repeats, double *ms_default, double *ms_tuned) {
    repeats = std::max(1, repeats);
    hipLaunchKernelGGL(k_tune_charges, dim3((n + 255) / 256), dim3(256), 0, stream, q_dev.as<double>(), n);
    check(hipGetLastError(), "k_tune_charges");
    hipEvent_t t0, t1;
    check(hipEventCreate(&t0), "hipEventCreate");
    check(hipEventCreate(&t1), "hipEventCreate");
    auto time_ms = [&]() {
        evaluate(q_dev.as<double>(), pot_dev.as<Real>(), stream);  // warm-up
        check(hipEventRecord(t0, stream), "hipEventRecord");
        for (int r = 0; r < repeats; ++r) evaluate(q_dev.as<double>(), pot_dev.as<Real>(), stream);
        check(hipEventRecord(t1, stream), "hipEventRecord");
        check(hipEventSynchronize(t1), "hipEventSynchronize");
        float ms = 0;
        check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
        return (double)ms / repeats;
    };
    const Knobs saved = knobs;
    knobs = Knobs{};
    const double base = time_ms();
    double best = base;
    // the rows near field keeps its LDS tables under 64 KiB per workgroup
    const size_t rows_lds_wave = nf_rows ? (size_t)(nf_ne_max + 1) * 6 * 16 : 0;
    auto try_values = [&](int Knobs::*field, std::initializer_list<int> values) {
        int best_v = knobs.*field;
        for (int v : values) {
            if (v == best_v) continue;
            if (field == &Knobs::near_wpb && rows_lds_wave * (size_t)v > 65536) continue;
            knobs.*field = v;
            const double t = time_ms();
            if (t < 0.98 * best) {
                best = t;
                best_v = v;
            }
        }
        knobs.*field = best_v;
    };
    try {
        try_values(&Knobs::near_wpb, {4, 8, 16});
        try_values(&Knobs::near_late, {0, 1});
        try_values(&Knobs::p2pw_bk, {16, 32, 64});
        try_values(&Knobs::pw2p_bk, {16, 32, 64});
        try_values(&Knobs::t3_div, {2});
    } catch (...) {
        knobs = saved;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        throw;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (ms_default) *ms_default = base;
    if (ms_tuned) *ms_tuned = best;
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_cache_load();
    const std::string key = tune_key();
    tune_cache[key] = knobs;
    if (const char *f = tune_file()) {
        std::ofstream out(f, std::ios::app);
        out << key << " " << knobs.near_wpb << " " << knobs.near_late << " " << knobs.p2pw_bk << " " << knobs.pw2p_bk
            << " " << knobs.t3_div << "\n";
    }
}

template <typename Real>
void Plan<Real>::evaluate_host(const double *q, Real *pot, double *times_ms) {
    check(hipMemcpyAsync(q_dev.p, q, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    evaluate(q_dev.as<double>(), pot_dev.as<Real>(), stream, 15, times_ms);
    check(hipMemcpyAsync(pot, pot_dev.p, sizeof(Real) * (size_t)n, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

template <typename Real>
void Plan<Real>::enable_kernel_timing(int slots) {
    for (auto e : kev) (void)hipEventDestroy(e);
    kev.clear();
    kslots = std::max(0, slots);
    kcall = 0;
    kev.resize((size_t)kslots * 2 * kGroups);
    for (auto &e : kev) check(hipEventCreate(&e), "hipEventCreate");
}

template <typename Real>
int Plan<Real>::kernel_times(double *avg_ms, int *n_calls) {
    const int nc = (int)std::min<long>(kcall, kslots);
    for (int g = 0; g < kGroups; ++g) avg_ms[g] = 0;
    for (int c = 0; c < nc; ++c)
        for (int g = 0; g < kGroups; ++g) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, kev[(size_t)c * 2 * kGroups + 2 * g], kev[(size_t)c * 2 * kGroups + 2 * g + 1]),
                  "hipEventElapsedTime");
            avg_ms[g] += ms / nc;
        }
    if (n_calls) *n_calls = nc;
    return nc;
}

void QpFarField::setup(const std::vector<double> &theta_host, const std::vector<double> &bx_host, int n_points,
                       int ms_, int mt_, int mu_, const std::vector<double> &w_full) {
    n = n_points;
    ms = ms_;
    mt = mt_;
    mu = mu_;
    M = ms * mt * mu;
    theta.upload(theta_host);
    bx.upload(bx_host);
    w.upload(w_full);
    // points per LDS tile: up to 16 within 48 KB of tables; point columns: enough workgroups
    // to fill the GPU, partials M complex each
    tp = std::max(1, std::min(16, (int)((48 << 10) / (sizeof(double2) * (size_t)(ms + mt + mu + 2)))));
    const int cblk = (((ms + QP_CHUNK - 1) / QP_CHUNK) * mt * mu + 255) / 256;
    nblk = std::max(1, std::min((n + tp - 1) / tp, std::max(1, 4096 / cblk)));
    Sp.alloc(sizeof(double2) * (size_t)nblk * M);
    S.alloc(sizeof(double2) * (size_t)M);
}

template <class R>
void QpFarField::run(const double *d_q, R *d_out, hipStream_t s) const {
    if (n == 0) return;
    const size_t lds = sizeof(double2) * (size_t)tp * (ms + mt + mu + 2);
    const int nch = ((ms + QP_CHUNK - 1) / QP_CHUNK) * mt * mu;
    hipLaunchKernelGGL(k_qp_structure, dim3((nch + 255) / 256, nblk), dim3(256), lds, s, theta.as<double>(), bx.as<double>(),
                       d_q, n, ms, mt, mu, tp, Sp.as<double2>());
    check(hipGetLastError(), "k_qp_structure");
    hipLaunchKernelGGL(k_qp_reduce, dim3((M + 255) / 256), dim3(256), 0, s, Sp.as<double2>(), nblk, w.as<double>(), M,
                       S.as<double2>());
    check(hipGetLastError(), "k_qp_reduce");
    hipLaunchKernelGGL(k_qp_eval<R>, dim3((n + 255) / 256), dim3(256), 0, s, theta.as<double>(), bx.as<double>(), n, ms, mt,
                       mu, S.as<double2>(), d_out);
    check(hipGetLastError(), "k_qp_eval");
}
template void QpFarField::run<float>(const double *, float *, hipStream_t) const;
template void QpFarField::run<double>(const double *, double *, hipStream_t) const;

template <typename Real>
void Plan<Real>::evaluate_complex(const double *d_q, Real *d_out, hipStream_t s, int stages) {
    if (!qp) throw std::logic_error("evaluate_complex: bloch not set at construction");
    if (!s) s = stream;
    // real part: proxy / plane waves, near field and self correction on Re(charges x phase);
    // imaginary part: the same without the self correction (the originals' charges are real)
    qp_comp = 0;
    try {
        evaluate(d_q, d_out, s, stages & 7, nullptr);
        qp_comp = 1;
        evaluate(d_q, d_out + 1, s, stages & 3, nullptr);
    } catch (...) {
        qp_comp = -1;
        throw;
    }
    qp_comp = -1;
    if (stages & 8) qpf.run(d_q, d_out, s);
}

template <typename Real>
void Plan<Real>::evaluate_complex_host(const double *q, Real *out) {
    check(hipMemcpyAsync(q_dev.p, q, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    evaluate_complex(q_dev.as<double>(), pot_dev2.as<Real>(), stream);
    check(hipMemcpyAsync(out, pot_dev2.p, 2 * sizeof(Real) * (size_t)n, hipMemcpyDeviceToHost, stream),
          "hipMemcpyAsync D2H");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void FourierSum::setup(const std::vector<double> &theta_host, int n_points, int ms_, int mt_, int mu_,
                       const std::vector<int> &M, const std::vector<double> &mode_w, Arena *arena, int max_blocks) {
    n = n_points;
    ms = ms_;
    mt = mt_;
    mu = mu_;
    n_modes = (int)mode_w.size();
    // an array in its owned buffer, or in a slice of the arena
    auto put = [&](DevBuf &own, const auto &v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (arena) return static_cast<const T *>(arena->put(v));
        own.upload(v);
        return own.template as<const T>();
    };
    theta_d = put(theta, theta_host);
    w_d = put(w, mode_w);
    // lines of constant (j, k) and consecutive i, cut into chunks of FF_CHUNK modes
    std::vector<int> cijk, cm0, clen, ljk, li, lptr;
    for (int m = 0; m < n_modes;) {
        int e = m + 1;
        while (e < n_modes && M[3 * e + 1] == M[3 * m + 1] && M[3 * e + 2] == M[3 * m + 2] && M[3 * e] == M[3 * (e - 1)] + 1)
            ++e;
        ljk.push_back(M[3 * m + 1]);
        ljk.push_back(M[3 * m + 2]);
        li.push_back(M[3 * m]);
        li.push_back(M[3 * (e - 1)]);
        lptr.push_back(m);
        for (int c = m; c < e; c += FF_CHUNK) {
            cijk.push_back(M[3 * c]);
            cijk.push_back(M[3 * c + 1]);
            cijk.push_back(M[3 * c + 2]);
            cm0.push_back(c);
            clen.push_back(std::min(FF_CHUNK, e - c));
        }
        m = e;
    }
    n_lines = (int)lptr.size();
    n_chunks = (int)cm0.size();
    chunk_ijk_d = put(chunk_ijk, cijk);
    chunk_m0_d = put(chunk_m0, cm0);
    chunk_len_d = put(chunk_len, clen);
    // weighted structure factors per line, padded at the high-i end to a multiple of
    // FF_QUAD modes (zeros), so the evaluation reads them FF_QUAD at a time
    std::vector<int4> lv;
    std::vector<int> lbase, pix(n_modes);
    int npad = 0;
    for (size_t L = 0; L < lptr.size(); ++L) {
        const int len = li[2 * L + 1] - li[2 * L] + 1, nq = (len + FF_QUAD - 1) / FF_QUAD;
        lv.push_back(make_int4(ljk[2 * L], ljk[2 * L + 1], li[2 * L], nq));
        lbase.push_back(npad);
        for (int r = 0; r < len; ++r) pix[lptr[L] + r] = npad + r;
        npad += nq * FF_QUAD;
    }
    lines_d = put(lines, lv);
    line_base_d = put(line_base, lbase);
    pidx_d = put(pidx, pix);
    blocks = std::max(1, std::min(std::min(2048, max_blocks), (n + FF_TP - 1) / FF_TP));
    const size_t sp_n = (size_t)blocks * std::max(1, n_modes), s_n = (size_t)std::max(1, npad);
    s_len = s_n;
    if (arena) {
        Sp_d = arena->take_n<double2>(sp_n);
        S_d = arena->take_n<double2>(s_n);
    } else {
        Sp.alloc(sizeof(double2) * sp_n);
        S.alloc(sizeof(double2) * s_n);
        Sp_d = Sp.as<double2>();
        S_d = S.as<double2>();
    }
    if (S_d) check(hipMemset(S_d, 0, sizeof(double2) * s_n), "hipMemset");  // padding stays zero
}

void FourierSum::run(const double *d_q, double *d_out, hipStream_t s) const {
    run_at(d_q, theta_d, n, d_out, s);
}

void FourierSum::run_at(const double *d_q, const double *d_theta_eval, int n_eval, double *d_out, hipStream_t s) const {
    if (n_eval <= 0) return;
    if (n == 0 || n_modes == 0) {
        hipLaunchKernelGGL(k_fill<double>, dim3((n_eval + 255) / 256), dim3(256), 0, s, d_out, 0.0, n_eval);
        check(hipGetLastError(), "k_fill far");
        return;
    }
    const size_t lds = sizeof(double2) * FF_TP * (ms + mt + mu + 1) + sizeof(double) * FF_TP;
    dim3 grid(blocks, (n_chunks + 255) / 256);
    hipLaunchKernelGGL(k_ff_structure<1>, grid, dim3(256), lds, s, theta_d, d_q, n, ms, mt, mu, chunk_ijk_d, chunk_m0_d,
                       chunk_len_d, n_chunks, n_modes, Sp_d, 0ll, 0ll, 1);
    check(hipGetLastError(), "k_ff_structure");
    hipLaunchKernelGGL(k_ff_reduce, dim3((n_modes + 31) / 32), dim3(256), 0, s, Sp_d, blocks, w_d, pidx_d, n_modes, S_d);
    check(hipGetLastError(), "k_ff_reduce");
    hipLaunchKernelGGL(k_ff_eval<1>, dim3((n_eval + 255) / 256), dim3(256), 0, s, d_theta_eval, n_eval, lines_d,
                       line_base_d, n_lines, S_d, d_out, 0ll, 0ll, 1);
    check(hipGetLastError(), "k_ff_eval");
}

void FourierSum::run_many(const double *d_Q, long long ldq, int nv, double *d_out, long long ldo, double2 *Sp_many,
                          int vc, double2 *S_many, hipStream_t s) const {
    if (nv <= 0 || n <= 0) return;
    if (n_modes == 0) {
        for (int v = 0; v < nv; ++v) run(d_Q + v * ldq, d_out + v * ldo, s);
        return;
    }
    if (vc < 1 || vc > 65535 || nv > 65535) throw std::invalid_argument("FourierSum::run_many: vector count");
    constexpr int VS = FF_MANY_S, VE = FF_MANY_E;
    const size_t lds = sizeof(double2) * FF_TP * (ms + mt + mu + 1) + sizeof(double) * FF_TP * VS;
    const long long spv = (long long)sp_elems(), sv = (long long)s_len;
    // the sums of run() per vector, VS (VE) vectors per workgroup of the structure factors
    // (evaluation) along z
    for (int v0 = 0; v0 < nv; v0 += vc) {
        const int c = std::min(vc, nv - v0);
        hipLaunchKernelGGL(k_ff_structure<VS>, dim3(blocks, (n_chunks + 255) / 256, (c + VS - 1) / VS), dim3(256), lds,
                           s, theta_d, d_Q + v0 * ldq, n, ms, mt, mu, chunk_ijk_d, chunk_m0_d, chunk_len_d, n_chunks,
                           n_modes, Sp_many, (long long)ldq, spv, c);
        check(hipGetLastError(), "k_ff_structure");
        hipLaunchKernelGGL(k_ff_reduce, dim3((n_modes + 31) / 32, 1, c), dim3(256), 0, s, Sp_many, blocks, w_d, pidx_d,
                           n_modes, S_many + v0 * sv, spv, sv);
        check(hipGetLastError(), "k_ff_reduce");
    }
    hipLaunchKernelGGL(k_ff_eval<VE>, dim3((n + 255) / 256, 1, (nv + VE - 1) / VE), dim3(256), 0, s, theta_d, n,
                       lines_d, line_base_d, n_lines, S_many, d_out, sv, (long long)ldo, nv);
    check(hipGetLastError(), "k_ff_eval");
}

void NufftFarField::setup(const std::vector<double> &theta_host, int n_points, int ms_, int mt_, int mu_,
                          const std::vector<int> &mode_ijk, const std::vector<double> &mode_w, int width) {
    const int Mm[3] = {ms_, mt_, mu_};
    host::NufftPlanHost P = host::make_nufft_grid(Mm, width);
    n = n_points;
    n_modes = (int)mode_w.size();
    W = P.W;
    T = P.T;
    lo = P.lo;
    E = P.E;
    SY = P.SY;
    SZ = P.SZ;
    for (int d = 0; d < 3; ++d) {
        g[d] = P.n[d];
        M[d] = P.M[d];
        nt[d] = P.ntile[d];
    }
    // per point on the GPU: keys (tile, base), plan order = stable sort by key, kernel-value
    // rows and packed bases in plan order, points per tile
    {
        NufftGridArgs a{};
        for (int d = 0; d < 3; ++d) {
            a.n[d] = P.n[d];
            a.ntile[d] = P.ntile[d];
            a.h[d] = P.h[d];
        }
        a.T = P.T;
        a.lo = P.lo;
        a.B = P.B;
        a.W = P.W;
        a.E = P.E;
        a.KS = P.KS;
        a.beta = P.beta;
        const int ntiles = P.ntile[0] * P.ntile[1] * P.ntile[2], B3 = P.B * P.B * P.B;
        DevBuf th, key, bad, cnt;
        th.upload(theta_host);
        key.alloc(sizeof(int) * std::max(1, n));
        bad.upload(std::vector<int>{0});
        cnt.upload(std::vector<int>(ntiles, 0));
        order.alloc(sizeof(int) * std::max(1, n));
        kval.alloc(sizeof(double) * (size_t)P.KS * std::max(1, n));
        pbase.alloc(sizeof(int) * std::max(1, n));
        const unsigned nb = (unsigned)((n + 255) / 256);
        hipStream_t s0 = nullptr;
        hipLaunchKernelGGL(k_nufft_keys, dim3(nb), dim3(256), 0, s0, th.as<double>(), n, a, key.as<int>(), bad.as<int>());
        check(hipGetLastError(), "k_nufft_keys");
        int bits = 1;
        while ((1L << bits) < (long)ntiles * B3) ++bits;
        stable_order(key.as<int>(), n, bits, order.as<int>(), s0);
        hipLaunchKernelGGL(k_nufft_rows, dim3(nb), dim3(256), 0, s0, th.as<double>(), order.as<int>(), n, a,
                           kval.as<double>(), pbase.as<int>());
        hipLaunchKernelGGL(k_nufft_tile_hist, dim3(nb), dim3(256), 0, s0, key.as<int>(), n, B3, cnt.as<int>());
        check(hipGetLastError(), "k_nufft_rows");
        int flag = 0;
        std::vector<int> tile_count(ntiles), ord(n);
        check(hipMemcpy(&flag, bad.p, sizeof(int), hipMemcpyDeviceToHost), "D2H");
        if (flag) throw std::logic_error("NUFFT base outside the tile subgrid");
        check(hipMemcpy(tile_count.data(), cnt.p, sizeof(int) * ntiles, hipMemcpyDeviceToHost), "D2H");
        check(hipMemcpy(ord.data(), order.p, sizeof(int) * n, hipMemcpyDeviceToHost), "D2H");
        position.assign(n, 0);
        for (int i = 0; i < n; ++i) position[ord[i]] = i;
        host::nufft_units(P, tile_count);
    }
    unit_ptr.upload(P.unit_ptr);
    unit_tile.upload(P.unit_tile);
    unit_slot.upload(P.unit_slot);
    extra_tile.upload(P.extra_tile);
    extra_ptr.upload(P.extra_ptr);
    n_units = (int)P.unit_tile.size();
    n_tiles = nt[0] * nt[1] * nt[2];
    n_extra_tiles = (int)P.extra_tile.size();
    for (int d = 0; d < 3; ++d) {
        gp[d].upload(P.gptr[d]);
        gt[d].upload(P.gtile[d]);
        gl[d].upload(P.gloc[d]);
        std::vector<double2> f(P.fwd[d].size()), b(P.inv[d].size());
        for (size_t i = 0; i < f.size(); ++i) f[i] = make_double2(P.fwd[d][i].real(), P.fwd[d][i].imag());
        for (size_t i = 0; i < b.size(); ++i) b[i] = make_double2(P.inv[d][i].real(), P.inv[d][i].imag());
        fwd[d].upload(f);
        inv[d].upload(b);
    }
    // listed modes: position in the mode box and weight x deconvolution^2
    std::vector<int> ps(n_modes);
    std::vector<double> c(n_modes);
    for (int i = 0; i < n_modes; ++i) {
        const int k0 = mode_ijk[3 * i] + M[0] / 2, k1 = mode_ijk[3 * i + 1] + M[1] / 2, k2 = mode_ijk[3 * i + 2] + M[2] / 2;
        ps[i] = (k2 * M[1] + k1) * M[0] + k0;
        const double dd = P.dcv[0][k0] * P.dcv[1][k1] * P.dcv[2][k2];
        c[i] = mode_w[i] * dd * dd;
    }
    pos.upload(ps);
    cw.upload(c);
    const size_t E3 = (size_t)E * E * E, G3 = (size_t)g[0] * g[1] * g[2], M3 = (size_t)M[0] * M[1] * M[2];
    const int n_extra = P.extra_ptr.empty() ? 0 : P.extra_ptr.back();
    qs.alloc(sizeof(double) * std::max(1, n));
    // tiles without points keep the zeros written here: the spreading writes only unit slots
    scratch.alloc(sizeof(double) * E3 * ((size_t)n_tiles + n_extra));
    check(hipMemset(scratch.p, 0, scratch.bytes), "hipMemset");
    grid.alloc(sizeof(double) * G3);
    grid2.alloc(sizeof(double) * G3);
    t1.alloc(sizeof(double2) * std::max((size_t)g[2] * g[1] * M[0], (size_t)1));
    t2.alloc(sizeof(double2) * std::max((size_t)g[2] * M[1] * M[0], (size_t)1));
    bbox.alloc(sizeof(double2) * M3);
    cbox.alloc(sizeof(double2) * M3);
    check(hipMemset(cbox.p, 0, cbox.bytes), "hipMemset");  // unlisted modes stay zero
    check(hipDeviceSynchronize(), "NUFFT setup");
}

namespace {

template <int W, int E>
void nufft_spread_interp(const NufftFarField &F, double *d_out, hipStream_t s, bool spread) {
    if (F.E != E) throw std::logic_error("NUFFT subgrid extent mismatch");
    const size_t G = (size_t)E * F.SZ;
    if (spread) {
        hipLaunchKernelGGL((k_nufft_spread<W, E>), dim3(F.n_units), dim3(64), sizeof(double) * G, s, F.unit_ptr.as<int>(),
                           F.unit_slot.as<int>(), F.kval.as<double>(), F.qs.as<double>(), F.pbase.as<int>(), F.SY, F.SZ,
                           F.scratch.as<double>());
        check(hipGetLastError(), "k_nufft_spread");
    } else {
        const size_t lds = sizeof(double) * (G + (size_t)NU_IB * NU_PS) + sizeof(int) * 3 * E;
        hipLaunchKernelGGL((k_nufft_interp<W, E>), dim3(F.n_units), dim3(64), lds, s, F.grid2.as<double>(),
                           F.unit_ptr.as<int>(), F.unit_tile.as<int>(), F.kval.as<double>(), F.pbase.as<int>(), F.g[0], F.g[1],
                           F.g[2], F.nt[0], F.nt[1], F.T, F.lo, F.SY, F.SZ, d_out);
        check(hipGetLastError(), "k_nufft_interp");
    }
}

// Subgrid extent E = T - 1 + W (+1 for odd W) for tile edge T = 2 or 4.
template <int W>
void nufft_width(const NufftFarField &F, double *d_out, hipStream_t s, bool spread) {
    constexpr int EW = W + (W & 1);
    if (F.T == 2) nufft_spread_interp<W, EW + 1>(F, d_out, s, spread);
    else if (F.T == 4) nufft_spread_interp<W, EW + 3>(F, d_out, s, spread);
    else throw std::logic_error("unsupported NUFFT tile edge");
}

void nufft_dispatch(const NufftFarField &F, double *d_out, hipStream_t s, bool spread) {
    switch (F.W) {
    case 4: nufft_width<4>(F, d_out, s, spread); break;
    case 5: nufft_width<5>(F, d_out, s, spread); break;
    case 6: nufft_width<6>(F, d_out, s, spread); break;
    case 7: nufft_width<7>(F, d_out, s, spread); break;
    case 8: nufft_width<8>(F, d_out, s, spread); break;
    case 9: nufft_width<9>(F, d_out, s, spread); break;
    case 10: nufft_width<10>(F, d_out, s, spread); break;
    case 11: nufft_width<11>(F, d_out, s, spread); break;
    case 12: nufft_width<12>(F, d_out, s, spread); break;
    case 13: nufft_width<13>(F, d_out, s, spread); break;
    case 14: nufft_width<14>(F, d_out, s, spread); break;
    default: throw std::logic_error("unsupported NUFFT kernel width");
    }
}

template <bool RIN, bool ROUT>
void dft_axis(const void *in, void *out, const DevBuf &mat, int A, int L, int K, int B, hipStream_t s) {
    const long lines = (long)A * B;
    hipLaunchKernelGGL((k_dft_axis<RIN, ROUT>), dim3((unsigned)((lines + 255) / 256), (K + DFT_KC - 1) / DFT_KC),
                       dim3(256), 0, s, in, out, mat.as<double2>(), A, L, K, B);
    check(hipGetLastError(), "k_dft_axis");
}

}  // namespace

void NufftFarField::run(const double *d_q, double *d_out, hipStream_t s) const {
    if (n == 0) return;
    if (n_modes == 0) {
        hipLaunchKernelGGL(k_fill<double>, dim3((n + 255) / 256), dim3(256), 0, s, d_out, 0.0, n);
        check(hipGetLastError(), "k_fill far");
        return;
    }
    hipLaunchKernelGGL(k_nufft_permute, dim3((n + 255) / 256), dim3(256), 0, s, d_q, order.as<int>(), n, qs.as<double>());
    check(hipGetLastError(), "k_nufft_permute");
    nufft_dispatch(*this, d_out, s, true);
    const int G3 = g[0] * g[1] * g[2], E3 = E * E * E;
    if (n_extra_tiles > 0) {
        hipLaunchKernelGGL(k_nufft_extra_reduce, dim3((E3 + 255) / 256, n_extra_tiles), dim3(256), 0, s,
                           scratch.as<double>(), extra_tile.as<int>(), extra_ptr.as<int>(), n_tiles, E3);
        check(hipGetLastError(), "k_nufft_extra_reduce");
    }
    hipLaunchKernelGGL(k_nufft_gather, dim3((G3 + 63) / 64), dim3(64), 0, s, scratch.as<double>(), gp[0].as<int>(),
                       gt[0].as<int>(), gl[0].as<int>(), gp[1].as<int>(), gt[1].as<int>(), gl[1].as<int>(), gp[2].as<int>(),
                       gt[2].as<int>(), gl[2].as<int>(), g[0], g[1], g[2], nt[0], nt[1], E, grid.as<double>());
    check(hipGetLastError(), "k_nufft_gather");
    // grid [g2][g1][g0] -> mode box [k2][k1][k0]
    dft_axis<true, false>(grid.p, t1.p, fwd[0], g[2] * g[1], g[0], M[0], 1, s);
    dft_axis<false, false>(t1.p, t2.p, fwd[1], g[2], g[1], M[1], M[0], s);
    dft_axis<false, false>(t2.p, bbox.p, fwd[2], 1, g[2], M[2], M[1] * M[0], s);
    hipLaunchKernelGGL(k_nufft_coef, dim3((n_modes + 255) / 256), dim3(256), 0, s, bbox.as<double2>(), pos.as<int>(),
                       cw.as<double>(), n_modes, cbox.as<double2>());
    check(hipGetLastError(), "k_nufft_coef");
    // mode box -> grid (real part)
    dft_axis<false, false>(cbox.p, t2.p, inv[2], 1, M[2], g[2], M[1] * M[0], s);
    dft_axis<false, false>(t2.p, t1.p, inv[1], g[2], M[1], g[1], M[0], s);
    dft_axis<false, true>(t1.p, grid2.p, inv[0], g[2] * g[1], M[0], g[0], 1, s);
    nufft_dispatch(*this, d_out, s, false);
}

template struct Plan<float>;
template struct Plan<double>;

}  // namespace pdmk::gpu
