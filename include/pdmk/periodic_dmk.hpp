/// C++ API of the MI355X PeriodicDMK library, mirroring
/// cpp/include/pdmk/periodic_dmk.hpp:17-144 of the reference: build once on the
/// host, evaluate many times on the GPU.  Only DIM = 3 is provided.  With a Bloch
/// vector the tree is quasi-periodic and only evaluate_complex() applies.
#pragma once

#include <pdmk/eval_timing.hpp>
#include <pdmk/types.hpp>

#include <array>
#include <complex>
#include <memory>
#include <optional>
#include <variant>
#include <vector>

namespace pdmk {

namespace gpu { template <typename Real> struct Plan; }

template <typename Real = double, int DIM = 3>
class PeriodicDMKTreeT {
    static_assert(DIM == 3, "the MI355X build provides DIM = 3");

public:
    using scalar_type = Real;

    /// Throws std::invalid_argument for bad arguments (N < 1, n_digits outside
    /// 1..12, nz <= 0, singular cell) and std::runtime_error for HIP failures.
    PeriodicDMKTreeT(const MatND<DIM>& cell_matrix,
                     const std::vector<VecND<DIM>>& positions,
                     int n_digits = 6,
                     int n_per_leaf = 300,
                     int nz = 4,
                     std::optional<std::array<double, DIM>> bloch = std::nullopt);
    ~PeriodicDMKTreeT();

    PeriodicDMKTreeT(const PeriodicDMKTreeT&) = delete;
    PeriodicDMKTreeT& operator=(const PeriodicDMKTreeT&) = delete;
    PeriodicDMKTreeT(PeriodicDMKTreeT&&) = delete;
    PeriodicDMKTreeT& operator=(PeriodicDMKTreeT&&) = delete;

    /// Periodic Coulomb potentials at the stored positions (user order).
    /// Throws std::invalid_argument if charges.size() != n_points().
    std::vector<Real> evaluate(const std::vector<double>& charges);

    /// Same as evaluate(), filling the three-way time breakdown.
    std::vector<Real> evaluate_timed(const std::vector<double>& charges, EvalTiming& timing);

    /// Quasi-periodic potentials (bloch given at construction): sum over images R of
    /// q e^{i bloch.R} / r, (Re, Im) per point.  Throws std::logic_error without bloch,
    /// std::invalid_argument for a size mismatch (cpp/src/periodic_dmk.cpp:251-277).
    /// evaluate() and evaluate_timed() throw std::logic_error when bloch is set.
    std::vector<std::complex<Real>> evaluate_complex(const std::vector<double>& charges);

    int n_points() const { return n_src_; }
    double cube_size() const { return rc_; }

    /// Device-resident evaluate (see pdmk_gpu.h).
    void evaluate_device(const double* d_charges, Real* d_pot, void* stream);

    gpu::Plan<Real>& plan() { return *p_; }

private:
    std::unique_ptr<gpu::Plan<Real>> p_;
    int n_src_;
    double rc_;
};

using PeriodicDMKTree = PeriodicDMKTreeT<double, 3>;

using AnyPDMKTree = std::variant<std::unique_ptr<PeriodicDMKTreeT<float, 3>>,
                                 std::unique_ptr<PeriodicDMKTreeT<double, 3>>>;

/// Picks FP32 kernels when n_digits <= 6, FP64 otherwise
/// (cpp/src/periodic_dmk.cpp:319-336).
class AutoPeriodicDMKTree {
public:
    AutoPeriodicDMKTree(const Mat3& cell, const std::vector<Vec3>& positions, int n_digits,
                        int n_per_leaf = 300, int nz = 4);
    AutoPeriodicDMKTree(AutoPeriodicDMKTree&&) = default;
    AutoPeriodicDMKTree& operator=(AutoPeriodicDMKTree&&) = default;
    bool is_single_precision() const { return single_precision_; }
    std::variant<std::vector<float>, std::vector<double>> evaluate(const std::vector<double>& charges);

private:
    AnyPDMKTree impl_;
    bool single_precision_ = false;
};

AutoPeriodicDMKTree make_periodic_dmk_tree(const Mat3& cell, const std::vector<Vec3>& positions, int n_digits,
                                           int n_per_leaf = 300, int nz = 4);

extern template class PeriodicDMKTreeT<double, 3>;
extern template class PeriodicDMKTreeT<float, 3>;

}  // namespace pdmk
