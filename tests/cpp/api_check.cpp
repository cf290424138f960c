// A C++ caller of the library's class API (include/pdmk/periodic_dmk.hpp), written the way
// a caller of the reference's PeriodicDMKTreeT / AutoPeriodicDMKTree is
// (cpp/include/pdmk/periodic_dmk.hpp:17-144; cpp/tests/test_periodic_dmk.cpp builds a tree
// from a Mat3 and std::vector<Vec3> and calls evaluate).  One C entry lets the tests drive it
// through ctypes in their own process.
#include <pdmk/periodic_dmk.hpp>

#include <cstdio>
#include <stdexcept>
#include <vector>

namespace {

template <typename Real>
void run_tree(const pdmk::Mat3 &cell, const std::vector<pdmk::Vec3> &p, const std::vector<double> &q, int n_digits,
              int npl, int nz, double *pot_tree, double *pot_timed, double *timing3, int *flags) {
    pdmk::PeriodicDMKTreeT<Real, 3> t(cell, p, n_digits, npl, nz);
    flags[0] = t.n_points();
    const std::vector<Real> v = t.evaluate(q);
    for (size_t i = 0; i < v.size(); ++i) pot_tree[i] = v[i];
    pdmk::EvalTiming tm;
    const std::vector<Real> w = t.evaluate_timed(q, tm);
    for (size_t i = 0; i < w.size(); ++i) pot_timed[i] = w[i];
    timing3[0] = tm.nufft_ms;
    timing3[1] = tm.dmk_ms;
    timing3[2] = tm.direct_ms;
    try {  // a charge vector of the wrong length
        t.evaluate(std::vector<double>(q.begin(), q.end() - 1));
    } catch (const std::invalid_argument &) {
        flags[2] = 1;
    }
    try {  // evaluate_complex on a tree built without a Bloch vector
        t.evaluate_complex(q);
    } catch (const std::logic_error &) {
        flags[3] = 1;
    }
}

}  // namespace

extern "C" int api_check(const double *cell_colmajor, int n, const double *pos, const double *charges, int n_digits,
                         int npl, int nz, double *pot_tree, double *pot_timed, double *pot_auto, double *timing3,
                         int *flags5) {
    try {
        pdmk::Mat3 cell;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) cell[j][i] = cell_colmajor[i + 3 * j];
        std::vector<pdmk::Vec3> p(n);
        for (int i = 0; i < n; ++i) p[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        const std::vector<double> q(charges, charges + n);
        for (int k = 0; k < 5; ++k) flags5[k] = 0;
        if (n_digits <= 6)
            run_tree<float>(cell, p, q, n_digits, npl, nz, pot_tree, pot_timed, timing3, flags5);
        else
            run_tree<double>(cell, p, q, n_digits, npl, nz, pot_tree, pot_timed, timing3, flags5);
        pdmk::AutoPeriodicDMKTree a = pdmk::make_periodic_dmk_tree(cell, p, n_digits, npl, nz);
        flags5[1] = a.is_single_precision() ? 1 : 0;
        const auto r = a.evaluate(q);
        if (const auto *f = std::get_if<std::vector<float>>(&r))
            for (int i = 0; i < n; ++i) pot_auto[i] = (*f)[i];
        else
            for (int i = 0; i < n; ++i) pot_auto[i] = std::get<std::vector<double>>(r)[i];
        pdmk::PeriodicDMKTreeT<double, 3> qp(cell, p, n_digits, npl, nz, std::array<double, 3>{0.1, 0.2, 0.3});
        try {  // evaluate on a quasi-periodic tree
            qp.evaluate(q);
        } catch (const std::logic_error &) {
            flags5[4] = 1;
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "api_check: %s\n", e.what());
        return -1;
    }
}
