// Host-side tree construction; see tree_build.hpp for the reference behaviour followed.
#include "tree_build.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace pdmk::host {

Geometry make_frame(const double *cell, const double *pos, int n, int nz) {
    if (n < 1) throw std::invalid_argument("PeriodicDMK: need at least one position");
    if (nz <= 0) throw std::invalid_argument("PeriodicDMK: nz must be positive");
    Geometry g;
    for (int i = 0; i < 9; ++i) g.cell[i] = cell[i];

    // Columns ordered by decreasing length (exchange order as pbc.hpp:35-40), then
    // Gram-Schmidt: umat = Q^T * A_sorted is upper triangular with positive diagonal.
    double len[3];
    int ord[3] = {0, 1, 2};
    for (int j = 0; j < 3; ++j)
        len[j] = std::sqrt(cell[3 * j] * cell[3 * j] + cell[3 * j + 1] * cell[3 * j + 1] + cell[3 * j + 2] * cell[3 * j + 2]);
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (len[j] > len[i]) { std::swap(len[i], len[j]); std::swap(ord[i], ord[j]); }
    double a[3][3];  // a[j] = sorted column j
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) a[j][i] = cell[i + 3 * ord[j]];
    double q[3][3], r[3][3] = {{0}};
    for (int j = 0; j < 3; ++j) {
        double v[3] = {a[j][0], a[j][1], a[j][2]};
        for (int k = 0; k < j; ++k) {
            r[k][j] = q[k][0] * a[j][0] + q[k][1] * a[j][1] + q[k][2] * a[j][2];
            for (int i = 0; i < 3; ++i) v[i] -= r[k][j] * q[k][i];
        }
        r[j][j] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(r[j][j] > 0)) throw std::invalid_argument("PeriodicDMK: singular cell");
        for (int i = 0; i < 3; ++i) q[j][i] = v[i] / r[j][j];
    }
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            g.umat[i + 3 * j] = (i <= j) ? r[i][j] : 0.0;
            g.qmat[i + 3 * j] = q[j][i];
        }
    const double *U = g.umat;
    g.rc = std::min(U[0], std::min(U[4], U[8])) / nz;

    // grid covering the cell in the rotated frame (pbc.hpp:104-117)
    for (int i = 2; i >= 0; --i) {
        double lo = 0, hi = U[i + 3 * i];
        for (int j = i + 1; j < 3; ++j) { lo += std::min(0.0, U[i + 3 * j]); hi += std::max(0.0, U[i + 3 * j]); }
        g.mgrid[i] = std::max(1, (int)std::ceil((hi - lo) / g.rc));
        g.xgrid0[i] = lo;
    }

    // rotate into the umat frame: r = Q^T x
    g.n_src = n;
    g.r_src.resize(3 * (size_t)n);
    for (int p = 0; p < n; ++p)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += g.qmat[k + 3 * j] * pos[3 * (size_t)p + k];
            g.r_src[3 * (size_t)p + j] = s;
        }

    return g;
}

std::vector<std::array<double, 3>> image_shifts(const Geometry &g, double elo[3], double ehi[3]) {
    const double *U = g.umat;
    double clo[3], chi[3];
    for (int i = 0; i < 3; ++i) {
        elo[i] = g.xgrid0[i] - g.rc;
        ehi[i] = g.xgrid0[i] + (g.mgrid[i] + 1) * g.rc;
        clo[i] = g.xgrid0[i];
        double w = U[i + 3 * i];
        for (int j = i + 1; j < 3; ++j) w += std::abs(U[i + 3 * j]);
        chi[i] = g.xgrid0[i] + w;
    }
    std::vector<std::array<double, 3>> out;
    const double rc = g.rc;
    const int n3 = (int)((g.mgrid[2] + 1) * rc / U[8]) + 1;
    for (int k3 = -n3; k3 <= n3; ++k3) {
        const double t3 = U[8] * k3;
        if (clo[2] + t3 > ehi[2] || chi[2] + t3 < elo[2]) continue;
        const int n2 = (int)(((g.mgrid[1] + 1) * rc + std::abs(U[7] * k3)) / U[4]) + 1;
        for (int k2 = -n2; k2 <= n2; ++k2) {
            const double t2 = U[4] * k2 + U[7] * k3;
            if (clo[1] + t2 > ehi[1] || chi[1] + t2 < elo[1]) continue;
            const double o12 = U[3] * k2 + U[6] * k3;
            const int n1 = (int)(((g.mgrid[0] + 1) * rc + std::abs(o12)) / U[0]) + 1;
            for (int k1 = -n1; k1 <= n1; ++k1) {
                if (!k1 && !k2 && !k3) continue;
                const double t1 = U[0] * k1 + o12;
                if (clo[0] + t1 > ehi[0] || chi[0] + t1 < elo[0]) continue;
                out.push_back({t1, t2, t3});
            }
        }
    }
    return out;
}

Geometry make_geometry(const double *cell, const double *pos, int n, int nz) {
    Geometry g = make_frame(cell, pos, n, nz);
    // periodic images inside the extended grid [xgrid0 - rc, xgrid0 + (mgrid+1) rc]
    // (image_layout.hpp:14-88: same shell order, so images carry the same indices)
    double elo[3], ehi[3];
    for (const auto &t : image_shifts(g, elo, ehi))
        for (int p = 0; p < n; ++p) {
            const double x = g.r_src[3 * (size_t)p] + t[0], y = g.r_src[3 * (size_t)p + 1] + t[1],
                         z = g.r_src[3 * (size_t)p + 2] + t[2];
            if (x < elo[0] || x > ehi[0] || y < elo[1] || y > ehi[1] || z < elo[2] || z > ehi[2]) continue;
            g.img_pos.push_back(x);
            g.img_pos.push_back(y);
            g.img_pos.push_back(z);
            g.img_src.push_back(p);
        }
    return g;
}

int Tree::find(int l, int x, int y, int z) const {
    auto it = index.find(Key{l, x, y, z});
    return it == index.end() ? -1 : it->second;
}

int Tree::colleague(int b, int s) const {
    const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
    return find(level[b], coord[b][0] + dx, coord[b][1] + dy, coord[b][2] + dz);
}

namespace {

// Do boxes A (level la, coords a) and B touch or overlap?
bool touching(int la, const std::array<int, 3> &a, int lb, const std::array<int, 3> &b) {
    const int L = std::max(la, lb);
    const int sa = L - la, sb = L - lb;
    for (int d = 0; d < 3; ++d) {
        const long alo = (long)a[d] << sa, ahi = ((long)a[d] + 1) << sa;
        const long blo = (long)b[d] << sb, bhi = ((long)b[d] + 1) << sb;
        if (alo > bhi || blo > ahi) return false;
    }
    return true;
}

struct Builder {
    Tree &t;

    int new_box(int lvl, int par, std::array<int, 3> c) {
        const int id = t.nboxes++;
        t.level.push_back(lvl);
        t.parent.push_back(par);
        t.first_child.push_back(-1);
        t.coord.push_back(c);
        const double bs = t.rc / (double)(1L << lvl);
        for (int d = 0; d < 3; ++d) t.center.push_back(t.xelo[d] + (c[d] + 0.5) * bs);
        t.src_begin.push_back(0);
        t.src_count.push_back(0);
        t.trg_begin.push_back(0);
        t.trg_count.push_back(0);
        t.index.emplace(Tree::Key{lvl, c[0], c[1], c[2]}, id);
        if ((int)t.by_level.size() <= lvl) t.by_level.resize(lvl + 1);
        t.by_level[lvl].push_back(id);
        return id;
    }

    // creates the 8 children of b (ids first_child[b] + k, bit d of k: upper half along d)
    void make_children(int b) {
        const int lvl = t.level[b];
        const auto c = t.coord[b];
        t.first_child[b] = t.nboxes;
        for (int k = 0; k < 8; ++k)
            new_box(lvl + 1, b, {2 * c[0] + (k & 1), 2 * c[1] + ((k >> 1) & 1), 2 * c[2] + ((k >> 2) & 1)});
    }
};

// The points held on the host: level-0 counting sort, stable in-place distribution into
// children (bit d of the child index: x_d >= centre_d).
struct HostPoints final : PointBackend {
    const Geometry &g;
    std::vector<double> spos, tpos;
    std::vector<int> sorder, torder;
    explicit HostPoints(const Geometry &geo) : g(geo) {}

    static void grid_sort(const Tree &t, const std::vector<double> &pos, int n, std::vector<int> &order,
                          std::vector<int> &count) {
        const double inv = 1.0 / t.rc;
        const int nb0 = t.mle[0] * t.mle[1] * t.mle[2];
        std::vector<int> box(n);
        count.assign(nb0, 0);
        for (int i = 0; i < n; ++i) {
            int ii[3];
            for (int d = 0; d < 3; ++d) {
                int v = (int)((pos[3 * (size_t)i + d] - t.xelo[d]) * inv);
                ii[d] = std::min(std::max(v, 0), t.mle[d] - 1);
            }
            box[i] = (ii[2] * t.mle[1] + ii[1]) * t.mle[0] + ii[0];
            count[box[i]]++;
        }
        std::vector<int> fill(nb0, 0);
        for (int b = 1; b < nb0; ++b) fill[b] = fill[b - 1] + count[b - 1];
        order.resize(n);
        for (int i = 0; i < n; ++i) order[fill[box[i]]++] = i;
    }

    void level0(const Tree &t, std::vector<int> &sc, std::vector<int> &tc) override {
        const int nsall = g.nsall(), n = g.n_src;
        std::vector<double> all(3 * (size_t)nsall);
        std::copy(g.r_src.begin(), g.r_src.end(), all.begin());
        std::copy(g.img_pos.begin(), g.img_pos.end(), all.begin() + 3 * (size_t)n);
        grid_sort(t, all, nsall, sorder, sc);
        grid_sort(t, g.r_src, n, torder, tc);
        spos.resize(3 * (size_t)nsall);
        tpos.resize(3 * (size_t)n);
        for (int i = 0; i < nsall; ++i)
            for (int d = 0; d < 3; ++d) spos[3 * (size_t)i + d] = all[3 * (size_t)sorder[i] + d];
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) tpos[3 * (size_t)i + d] = g.r_src[3 * (size_t)torder[i] + d];
    }

    static void split_range(const double *cen, int begin, int count, std::vector<double> &pos, std::vector<int> &order,
                            int child_begin[8], int child_count[8]) {
        std::vector<uint8_t> code(count);
        int cnt[8] = {0};
        for (int i = 0; i < count; ++i) {
            const double *x = &pos[3 * (size_t)(begin + i)];
            const int c = (x[0] >= cen[0] ? 1 : 0) | (x[1] >= cen[1] ? 2 : 0) | (x[2] >= cen[2] ? 4 : 0);
            code[i] = (uint8_t)c;
            cnt[c]++;
        }
        int off[8];
        off[0] = begin;
        for (int c = 1; c < 8; ++c) off[c] = off[c - 1] + cnt[c - 1];
        for (int c = 0; c < 8; ++c) { child_begin[c] = off[c]; child_count[c] = cnt[c]; }
        std::vector<double> p(pos.begin() + 3 * (size_t)begin, pos.begin() + 3 * (size_t)(begin + count));
        std::vector<int> o(order.begin() + begin, order.begin() + begin + count);
        for (int i = 0; i < count; ++i) {
            const int dst = off[code[i]]++;
            pos[3 * (size_t)dst] = p[3 * (size_t)i];
            pos[3 * (size_t)dst + 1] = p[3 * (size_t)i + 1];
            pos[3 * (size_t)dst + 2] = p[3 * (size_t)i + 2];
            order[dst] = o[i];
        }
    }

    void distribute(Tree &t, const std::vector<int> &boxes) override {
        for (int b : boxes) {
            const int first = t.first_child[b];
            const double cen[3] = {t.center[3 * b], t.center[3 * b + 1], t.center[3 * b + 2]};
            int cb[8], cc[8];
            if (t.src_count[b] > 0) {
                split_range(cen, t.src_begin[b], t.src_count[b], spos, sorder, cb, cc);
                for (int k = 0; k < 8; ++k) { t.src_begin[first + k] = cb[k]; t.src_count[first + k] = cc[k]; }
            }
            if (t.trg_count[b] > 0) {
                split_range(cen, t.trg_begin[b], t.trg_count[b], tpos, torder, cb, cc);
                for (int k = 0; k < 8; ++k) { t.trg_begin[first + k] = cb[k]; t.trg_count[first + k] = cc[k]; }
            }
        }
    }

    void finish(Tree &t) override {
        t.src_order = std::move(sorder);
        t.trg_order = std::move(torder);
        t.src_sorted = std::move(spos);
        t.trg_sorted = std::move(tpos);
    }
};

}  // namespace

Tree build_tree(const Geometry &g, int n_per_leaf) {
    HostPoints pts(g);
    return build_tree_with(g, n_per_leaf, pts);
}

Tree build_tree_with(const Geometry &g, int n_per_leaf, PointBackend &pts) {
    if (n_per_leaf < 1) throw std::invalid_argument("PeriodicDMK: n_per_leaf must be positive");
    Tree t;
    t.rc = g.rc;
    for (int d = 0; d < 3; ++d) { t.mle[d] = g.mgrid[d] + 2; t.xelo[d] = g.xgrid0[d] - g.rc; }

    Builder B{t};
    for (int z = 0; z < t.mle[2]; ++z)
        for (int y = 0; y < t.mle[1]; ++y)
            for (int x = 0; x < t.mle[0]; ++x) B.new_box(0, -1, {x, y, z});
    {
        std::vector<int> sc, tc;
        pts.level0(t, sc, tc);
        int so = 0, to = 0;
        for (int id = 0; id < t.nboxes; ++id) {
            t.src_begin[id] = so; t.src_count[id] = sc[id]; so += sc[id];
            t.trg_begin[id] = to; t.trg_count[id] = tc[id]; to += tc[id];
        }
    }

    // refinement while a leaf holds more than ceil(1.5 * n_per_leaf) sources; the boxes of
    // one level are split together, then their points are distributed
    const int limit = (int)std::ceil(n_per_leaf * 1.5);
    // integer box coordinates at level L reach mle * 2^L; children past 2^30 would overflow
    // them (more than `limit` coincident points split without end)
    const long long mle_max = std::max({t.mle[0], t.mle[1], t.mle[2]});
    int max_level = 0;
    for (int lvl = 0; lvl < 50; ++lvl) {
        if ((int)t.by_level.size() <= lvl) break;
        std::vector<int> split;
        for (int b : t.by_level[lvl])
            if (t.first_child[b] < 0 && t.src_count[b] > limit) split.push_back(b);
        if (split.empty()) break;
        if ((mle_max << (lvl + 1)) > (1LL << 30))
            throw std::invalid_argument("PeriodicDMK: more than ceil(1.5 * n_per_leaf) coincident points; the tree "
                                        "would exceed its depth limit");
        for (int b : split) B.make_children(b);
        pts.distribute(t, split);
        max_level = lvl + 1;
    }

    // 2:1 balance: a non-empty leaf at level L touching a non-leaf box of level L+1 is split.
    // Whether a leaf of level L violates depends only on boxes of levels L+1 and L+2, which
    // the splits of level L do not change, so each level's splits are decided together.
    for (int iter = 0; iter < 50; ++iter) {
        bool changed = false;
        for (int L = 0; L <= max_level; ++L) {
            if ((int)t.by_level.size() <= L) break;
            std::vector<int> split;
            for (int b : t.by_level[L]) {
                if (t.first_child[b] >= 0 || t.src_count[b] <= 0) continue;
                bool violate = false;
                for (int s = 0; s < 27 && !violate; ++s) {
                    const int cc = t.colleague(b, s);
                    if (cc < 0 || cc == b || t.first_child[cc] < 0) continue;
                    for (int k = 0; k < 8; ++k) {
                        const int gc = t.first_child[cc] + k;
                        if (t.first_child[gc] < 0) continue;
                        if (touching(L, t.coord[b], L + 1, t.coord[gc])) { violate = true; break; }
                    }
                }
                if (violate) split.push_back(b);
            }
            if (split.empty()) continue;
            for (int b : split) B.make_children(b);
            pts.distribute(t, split);
            changed = true;
            max_level = std::max(max_level, L + 1);
        }
        if (!changed) break;
    }

    t.nlevels = max_level + 1;
    t.by_level.resize(t.nlevels);
    t.boxsize.resize(t.nlevels + 1);
    t.boxsize[0] = t.rc;
    for (int l = 1; l <= t.nlevels; ++l) t.boxsize[l] = t.boxsize[l - 1] / 2;
    const int nb = t.nboxes;
    t.leaf.assign(nb, 0);
    for (int b = 0; b < nb; ++b) t.leaf[b] = t.first_child[b] < 0;
    // subtree counts (non-leaf begin stays at the start of its range)
    for (int b = 0; b < nb; ++b)
        if (!t.leaf[b]) { t.src_count[b] = 0; t.trg_count[b] = 0; }
    for (int l = t.nlevels - 1; l >= 1; --l)
        for (int b : t.by_level[l]) {
            t.src_count[t.parent[b]] += t.src_count[b];
            t.trg_count[t.parent[b]] += t.trg_count[b];
        }
    pts.finish(t);

    // plane-wave flags (pbc_tree.cpp:776-815) and leaf evaluation flags (817-831)
    t.pwexp.assign(nb, 0);
    t.tpeval.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        if (t.level[b] == 0 || !t.leaf[b]) t.pwexp[b] = 1;
        else
            for (int s = 0; s < 27; ++s) {
                const int c = t.colleague(b, s);
                if (c >= 0 && !t.leaf[c]) { t.pwexp[b] = 1; break; }
            }
        if (t.src_count[b] == 0 && t.trg_count[b] == 0) t.pwexp[b] = 0;
        if (t.leaf[b] && t.src_count[b] + t.trg_count[b] > 0) t.tpeval[b] = 1;
    }

    // list1 (pbc_tree.cpp:850-907): leaf colleagues, touching leaf children of
    // non-leaf colleagues, touching leaf colleagues of the parent
    t.l1_begin.assign(nb + 1, 0);
    t.l1_box.clear();
    std::vector<std::vector<int>> l1(nb);
    for (int l = 0; l < t.nlevels; ++l)
        for (int box : t.by_level[l]) {
            if (!t.leaf[box]) continue;
            auto &L1 = l1[box];
            for (int s = 0; s < 27; ++s) {
                const int nbx = t.colleague(box, s);
                if (nbx < 0) continue;
                if (t.leaf[nbx] && t.src_count[nbx] > 0) L1.push_back(nbx);
                else if (!t.leaf[nbx] && l + 1 < t.nlevels)
                    for (int k = 0; k < 8; ++k) {
                        const int kb = t.first_child[nbx] + k;
                        if (t.src_count[kb] == 0 || !t.leaf[kb]) continue;
                        if (touching(l, t.coord[box], l + 1, t.coord[kb])) L1.push_back(kb);
                    }
            }
            if (l == 0) continue;
            const int dad = t.parent[box];
            for (int s = 0; s < 27; ++s) {
                const int nbx = t.colleague(dad, s);
                if (nbx < 0 || !t.leaf[nbx] || t.src_count[nbx] == 0) continue;
                if (touching(l, t.coord[box], l - 1, t.coord[nbx])) L1.push_back(nbx);
            }
        }
    for (int b = 0; b < nb; ++b) {
        t.l1_begin[b + 1] = t.l1_begin[b] + (int)l1[b].size();
        t.l1_box.insert(t.l1_box.end(), l1[b].begin(), l1[b].end());
    }

    // upward work (pbc_tree.cpp:909-943)
    t.proxy_from_children.assign(nb, 0);
    for (int l = t.nlevels - 1; l >= 0; --l)
        for (int b : t.by_level[l]) {
            if (!t.pwexp[b] || t.src_count[b] == 0) continue;
            bool child_proxy = false;
            if (!t.leaf[b])
                for (int k = 0; k < 8; ++k) {
                    const int ch = t.first_child[b] + k;
                    if (t.pwexp[ch] && t.src_count[ch] > 0) { child_proxy = true; break; }
                }
            Tree::UpGroup grp{b, l, {}};
            if (child_proxy) {
                t.proxy_from_children[b] = 1;
                for (int k = 0; k < 8; ++k) {
                    const int ch = t.first_child[b] + k;
                    if (t.src_count[ch] > 0 && !t.pwexp[ch]) grp.src_boxes.push_back(ch);
                }
            } else {
                grp.src_boxes.push_back(b);
            }
            if (!grp.src_boxes.empty()) t.up_groups.push_back(std::move(grp));
        }
    return t;
}

CubicCover cubic_cover(const double *cell, int nz) {
    if (nz <= 0) throw std::invalid_argument("cubic cover nz must be positive");
    const double *a[3] = {cell, cell + 3, cell + 6};  // lattice vectors = columns
    auto dot3 = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[1][0] * (a[0][1] * a[2][2] - a[0][2] * a[2][1]) +
                       a[2][0] * (a[0][1] * a[1][2] - a[0][2] * a[1][1]);
    if (std::fabs(det) < 1e-30) throw std::invalid_argument("cubic cover: singular cell");
    // longest and second-longest vectors; on equal norms the lower index ranks first
    int ord[3] = {0, 1, 2};
    double nrm[3];
    for (int j = 0; j < 3; ++j) nrm[j] = std::sqrt(dot3(a[j], a[j]));
    std::stable_sort(ord, ord + 3, [&](int x, int y) { return nrm[x] > nrm[y]; });
    CubicCover c;
    const double *b1 = a[ord[0]], *b2 = a[ord[1]];
    if (!(nrm[ord[0]] > 0)) throw std::invalid_argument("cubic cover: degenerate lattice (zero-norm vector)");
    for (int d = 0; d < 3; ++d) c.e1[d] = b1[d] / nrm[ord[0]];
    const double p21 = dot3(b2, c.e1);
    double v2[3];
    for (int d = 0; d < 3; ++d) v2[d] = b2[d] - p21 * c.e1[d];
    const double nv2 = std::sqrt(dot3(v2, v2));
    if (!(nv2 > 0)) throw std::invalid_argument("cubic cover: degenerate lattice (parallel vectors)");
    for (int d = 0; d < 3; ++d) c.e2[d] = v2[d] / nv2;
    c.e3[0] = c.e1[1] * c.e2[2] - c.e1[2] * c.e2[1];
    c.e3[1] = c.e1[2] * c.e2[0] - c.e1[0] * c.e2[2];
    c.e3[2] = c.e1[0] * c.e2[1] - c.e1[1] * c.e2[0];
    // widths of the cell's bounding box along the frame axes
    double w[3] = {0, 0, 0};
    const double *e[3] = {c.e1, c.e2, c.e3};
    for (int d = 0; d < 3; ++d)
        for (int j = 0; j < 3; ++j) w[d] += std::fabs(dot3(a[j], e[d]));
    c.s = w[2] / nz;
    c.m = std::max(1, (int)std::ceil(w[0] / c.s));
    c.n = std::max(1, (int)std::ceil(w[1] / c.s));
    c.p = nz;
    c.cost = (double)((c.m + 2) * (c.n + 2) * (c.p + 2)) * c.s * c.s * c.s / std::fabs(det);
    return c;
}

}  // namespace pdmk::host
