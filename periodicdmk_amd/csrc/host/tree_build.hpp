// Host-side construction of the periodic DMK tree (built once per tree, FP64).
//
// Produces flat arrays that the device plan uploads.  Box topology, particle
// order and interaction lists follow the reference PBCTree so that the GPU path
// evaluates the same expansion as the reference CPU path:
//   * lattice QR with longest-first column order ..... cpp/include/dmk/pbc.hpp:23-101
//   * level-0 grid and image cover .................... cpp/include/dmk/pbc.hpp:104-117, 265-339,
//                                                       cpp/include/pdmk/image_layout.hpp:14-88
//   * bin sort, refinement with ceil(1.5*npl) ......... cpp/src/dmk/pbc_tree.cpp:247-428
//   * 2:1 balance ..................................... cpp/src/dmk/pbc_tree.cpp:438-598
//   * plane-wave flags, list1, work lists .............. cpp/src/dmk/pbc_tree.cpp:776-959
// Boxes are identified by integer lattice coordinates at their level, so
// adjacency tests are exact integer comparisons (the reference's 1.05 tolerances
// select exactly the touching boxes).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace pdmk::host {

struct Geometry {
    double cell[9];          // column-major input lattice
    double umat[9], qmat[9]; // upper-triangular frame and rotation (column-major)
    double rc = 0;
    int mgrid[3] = {0, 0, 0};
    double xgrid0[3] = {0, 0, 0};
    int n_src = 0;
    std::vector<double> r_src;    // rotated positions, 3*n_src
    std::vector<int> img_src;     // originating source of each image
    std::vector<double> img_pos;  // 3*n_img
    int nsall() const { return n_src + (int)img_src.size(); }
};

// Lattice frame, cutoff cube size rc = min diag(umat)/nz, rotated positions and images.
Geometry make_geometry(const double *cell_colmajor, const double *pos, int n, int nz);

// make_geometry without the images.
Geometry make_frame(const double *cell_colmajor, const double *pos, int n, int nz);

// Image translations in make_geometry's enumeration order (shells k3, k2, k1); an image
// r + t of a rotated position r is kept iff elo <= r + t <= ehi componentwise, and the kept
// images are numbered translation by translation, points in index order within each.
std::vector<std::array<double, 3>> image_shifts(const Geometry &g, double elo[3], double ehi[3]);

// Cover of the cell's bounding box by (m+2) x (n+2) x (p+2) cubes of side s in the frame
// (e1, e2, e3): e1 along the longest lattice vector, e2 in the plane of the two longest,
// p = nz layers across the e3 width (optimize_cubic_cover,
// cpp/src/cubic_cover_optimizer.cpp:36-89).  Throws std::invalid_argument for nz <= 0 or a
// degenerate cell.
struct CubicCover {
    int m = 0, n = 0, p = 0;
    double e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0}, e3[3] = {0, 0, 0};
    double s = 0, cost = 0;
};
CubicCover cubic_cover(const double *cell_colmajor, int nz);

struct Tree {
    int nboxes = 0, nlevels = 0;
    int mle[3] = {0, 0, 0};
    double xelo[3] = {0, 0, 0};
    double rc = 0;
    std::vector<double> boxsize;                 // nlevels + 1
    std::vector<int> level, parent, first_child; // first_child = -1 for leaves (children are consecutive ids)
    std::vector<std::array<int, 3>> coord;       // integer lattice coordinates at the box's level
    std::vector<double> center;                  // 3 * nboxes
    std::vector<int> src_begin, src_count, trg_begin, trg_count;
    std::vector<std::vector<int>> by_level;      // box ids per level, ascending
    std::vector<int> src_order;  // sorted source -> combined index (originals then images)
    std::vector<int> trg_order;  // sorted target -> user index
    std::vector<double> src_sorted, trg_sorted;  // 3*nsall, 3*n_src

    std::vector<uint8_t> leaf, pwexp, tpeval;
    std::vector<int> l1_begin, l1_box;           // list1 CSR over boxes (nboxes + 1)

    // upward work: proxies formed directly from charges, grouped by destination box
    struct UpGroup { int box; int level; std::vector<int> src_boxes; };
    std::vector<UpGroup> up_groups;
    std::vector<uint8_t> proxy_from_children;

    int colleague(int b, int s) const;  // stencil s = (dx+1) + 3(dy+1) + 9(dz+1), -1 if absent
    int find(int lvl, int x, int y, int z) const;

    // (level, x, y, z) -> box id
    struct Key {
        int l, x, y, z;
        bool operator==(const Key &o) const { return l == o.l && x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            uint64_t h = (uint64_t)(uint32_t)k.l * 0x9E3779B97F4A7C15ull;
            h ^= (uint64_t)(uint32_t)k.x + 0x7F4A7C15ull + (h << 6) + (h >> 2);
            h ^= (uint64_t)(uint32_t)k.y * 0xBF58476D1CE4E5B9ull + (h << 6) + (h >> 2);
            h ^= (uint64_t)(uint32_t)k.z * 0x94D049BB133111EBull + (h << 6) + (h >> 2);
            return (size_t)h;
        }
    };
    std::unordered_map<Key, int, KeyHash> index;
};

// Where the points live while the tree is built.  build_tree_with asks the backend for the
// level-0 counts (boxes in grid order, points stable by index within a box), then, after
// each batch of splits (all boxes of one level), for the begin/count of the new children
// (stable distribution, child bit d set when x_d >= centre_d of the parent), and at the end
// for the sorted order and positions.
struct PointBackend {
    virtual ~PointBackend() = default;
    virtual void level0(const Tree &t, std::vector<int> &src_count, std::vector<int> &trg_count) = 0;
    virtual void distribute(Tree &t, const std::vector<int> &split_boxes) = 0;
    virtual void finish(Tree &t) = 0;  // src_order, trg_order, src_sorted, trg_sorted
};

Tree build_tree(const Geometry &g, int n_per_leaf);  // points on the host
Tree build_tree_with(const Geometry &g, int n_per_leaf, PointBackend &points);

}  // namespace pdmk::host
