// Host-only introspection entry points (pdmk_host.h); they never touch the GPU.
#include <pdmk/pdmk_host.h>

#include "tables.hpp"
#include "tree_build.hpp"

#include <cstdio>
#include <cstring>
#include <exception>

using namespace pdmk::host;

extern "C" {

int pdmk_host_tree(const double *cell, int N, const double *pos, int n_per_leaf, int nz, long long *sizes8, int *box7,
                   double *centers, int *src_perm, int *trg_perm, double *geom16) {
    try {
        const Geometry g = make_geometry(cell, pos, N, nz);
        const Tree t = build_tree(g, n_per_leaf);
        if (sizes8) {
            long long v[8] = {g.n_src, g.nsall(), t.nboxes, t.nlevels, (long long)t.l1_box.size(),
                              (long long)t.up_groups.size(), 0, 0};
            std::memcpy(sizes8, v, sizeof(v));
        }
        for (int b = 0; b < t.nboxes; ++b) {
            if (box7) {
                int v[7] = {t.level[b], t.parent[b], t.leaf[b] ? 0 : 8, t.src_count[b], t.trg_count[b], t.pwexp[b],
                            t.l1_begin[b + 1] - t.l1_begin[b]};
                std::memcpy(box7 + 7 * (size_t)b, v, sizeof(v));
            }
            if (centers)
                for (int d = 0; d < 3; ++d) centers[3 * (size_t)b + d] = t.center[3 * b + d];
        }
        if (src_perm) std::memcpy(src_perm, t.src_order.data(), sizeof(int) * t.src_order.size());
        if (trg_perm) std::memcpy(trg_perm, t.trg_order.data(), sizeof(int) * t.trg_order.size());
        if (geom16) {
            for (int i = 0; i < 9; ++i) geom16[i] = g.umat[i];
            geom16[9] = g.rc;
            for (int d = 0; d < 3; ++d) { geom16[10 + d] = g.mgrid[d]; geom16[13 + d] = g.xgrid0[d]; }
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_host_tree: %s\n", e.what());
        return -1;
    }
}

int pdmk_host_tables(int n_digits, double rc, int n_levels, const double *umat, double *s8, double *w0, double *radial0,
                     double *poly2pw, double *p2c, double *far) {
    try {
        const DmkTables T = make_tables(n_digits, rc, n_levels, umat);
        if (s8) {
            double v[8] = {T.beta, T.psi0, T.c0, T.g0d2, (double)T.n_pw, (double)T.p, (double)T.n_modes,
                           (double)T.mode_w.size()};
            std::memcpy(s8, v, sizeof(v));
        }
        if (w0) std::memcpy(w0, T.w0.data(), sizeof(double) * T.w0.size());
        if (radial0) std::memcpy(radial0, T.radial0.data(), sizeof(double) * T.radial0.size());
        if (poly2pw)
            for (size_t i = 0; i < T.poly2pw.size(); ++i) {
                poly2pw[2 * i] = T.poly2pw[i].real();
                poly2pw[2 * i + 1] = T.poly2pw[i].imag();
            }
        if (p2c) {
            std::memcpy(p2c, T.p2c[0].data(), sizeof(double) * T.p2c[0].size());
            std::memcpy(p2c + T.p2c[0].size(), T.p2c[1].data(), sizeof(double) * T.p2c[1].size());
        }
        if (far)
            for (size_t m = 0; m < T.mode_w.size(); ++m) {
                far[4 * m] = T.mode_ijk[3 * m];
                far[4 * m + 1] = T.mode_ijk[3 * m + 1];
                far[4 * m + 2] = T.mode_ijk[3 * m + 2];
                far[4 * m + 3] = T.mode_w[m];
            }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "pdmk_host_tables: %s\n", e.what());
        return -1;
    }
}

double pdmk_host_pswf(double c, double x) { return Pswf(c)(x); }

}  // extern "C"
