// gpw_replay -- replays a CP2K "#Grid collocate task v1" file through the
// drop-in C++ signature (include/libgpw_amd.hpp) on the GPU, `cycles` times,
// and compares with the grid values the file holds.
//
// Counterpart of the reference's replay (src/lib/grid_base_ref_replay.cc:18-273,
// src/lib/grid_base_ref_miniapp.cc:15-33), written for this library:
//   * every dumped grid entry is compared (the reference's compare loop only
//     visits the x = 0 plane, src/lib/grid_base_ref_replay.cc:264), and points
//     outside the dumped set are reported separately;
//   * general (non-orthorhombic) tasks carry no lb_cube/ub_cube and are passed
//     through without dereferencing them (the reference dereferences NULL at
//     src/lib/grid_base_ref_replay.cc:226);
//   * errors return a non-zero exit status instead of asserting.
// Usage: gpw_replay <task-file> [cycles]   (exit 0: max diff <= 1e-12 * scale)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "libgpw_amd.hpp"

namespace {

// minimal row-major array with the element access the drop-in header uses
template <int N>
struct RowMajor {
    std::vector<double> data;
    int dims[N];
    double &operator()(int i) { return data[i]; }
    double &operator()(int i, int j) { return data[(size_t)i * dims[1] + j]; }
    const double &operator()(int i, int j) const { return data[(size_t)i * dims[1] + j]; }
    double &operator()(int i, int j, int k) { return data[((size_t)i * dims[1] + j) * dims[2] + k]; }
    const double &operator()(int i, int j, int k) const {
        return data[((size_t)i * dims[1] + j) * dims[2] + k];
    }
};

struct Reader {
    FILE *fp;
    int line = 0;
    char buf[512];
    bool next() {
        line++;
        return fgets(buf, sizeof buf, fp) != nullptr;
    }
    [[noreturn]] void die(const char *what) {
        fprintf(stderr, "task file line %d: %s\n", line, what);
        exit(2);
    }
    template <class... T>
    void get(const char *key, const char *fmt, T *...out) {
        if (!next()) die("unexpected end of file");
        char k[64];
        std::string f = std::string("%63s ") + fmt;
        if (sscanf(buf, f.c_str(), k, out...) != (int)sizeof...(T) + 1 || strcmp(k, key) != 0) die(key);
    }
};

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        printf("Usage: gpw_replay <task-file> [cycles]\n");
        return 1;
    }
    const int cycles = argc > 2 ? atoi(argv[2]) : 1;
    FILE *fp = fopen(argv[1], "r");
    if (!fp) {
        perror(argv[1]);
        return 2;
    }
    Reader r{fp};
    if (!r.next() || strcmp(r.buf, "#Grid collocate task v1\n") != 0) r.die("header");
    int tau, ortho, la_max, la_min, lb_max, lb_min, lmax, nspheres, nnz, maxco, o1, o2, ngnz;
    double zeta, zetb, rscale, radius, dh[3][3], dh_inv[3][3], ra[3], rab[3];
    int npts[3], ngrid[3], lb_grid[3], per[3], lb_cube[3] = {0, 0, 0}, ub_cube[3] = {0, 0, 0};
    r.get("compute_tau", "%d", &tau);
    r.get("use_ortho", "%d", &ortho);
    r.get("la_max", "%d", &la_max);
    r.get("la_min", "%d", &la_min);
    r.get("lb_max", "%d", &lb_max);
    r.get("lb_min", "%d", &lb_min);
    r.get("zeta", "%lf", &zeta);
    r.get("zetb", "%lf", &zetb);
    r.get("rscale", "%lf", &rscale);
    for (int i = 0; i < 3; i++) {
        int j;
        r.get("dh", "%d %lf %lf %lf", &j, &dh[i][0], &dh[i][1], &dh[i][2]);
    }
    for (int i = 0; i < 3; i++) {
        int j;
        r.get("dh_inv", "%d %lf %lf %lf", &j, &dh_inv[i][0], &dh_inv[i][1], &dh_inv[i][2]);
    }
    r.get("ra", "%lf %lf %lf", &ra[0], &ra[1], &ra[2]);
    r.get("rab", "%lf %lf %lf", &rab[0], &rab[1], &rab[2]);
    r.get("npts", "%d %d %d", &npts[0], &npts[1], &npts[2]);
    r.get("ngrid", "%d %d %d", &ngrid[0], &ngrid[1], &ngrid[2]);
    r.get("lb_grid", "%d %d %d", &lb_grid[0], &lb_grid[1], &lb_grid[2]);
    r.get("periodic", "%d %d %d", &per[0], &per[1], &per[2]);
    r.get("lmax", "%d", &lmax);
    r.get("radius", "%lf", &radius);
    if (ortho) {
        r.get("lb_cube", "%d %d %d", &lb_cube[0], &lb_cube[1], &lb_cube[2]);
        r.get("ub_cube", "%d %d %d", &ub_cube[0], &ub_cube[1], &ub_cube[2]);
    }
    r.get("nspheres", "%d", &nspheres);
    r.get("nspheres_nonzero", "%d", &nnz);
    std::vector<int> sphere_bounds(std::max(nspheres, 0), 0);
    for (int n = 0; n < nnz; n++) {
        int i, v;
        r.get("sphere_bounds", "%d %d", &i, &v);
        if (i < 0 || i >= nspheres) r.die("sphere_bounds index");
        sphere_bounds[i] = v;
    }
    r.get("maxco", "%d", &maxco);
    r.get("o1", "%d", &o1);
    r.get("o2", "%d", &o2);
    RowMajor<2> pab{std::vector<double>((size_t)maxco * maxco), {maxco, maxco}};
    for (int i = 0; i < maxco; i++)
        for (int j = 0; j < maxco; j++) {
            int a, b;
            r.get("pab", "%d %d %lf", &a, &b, &pab(i, j));
        }
    r.get("ngrid_nonzero", "%d", &ngnz);
    const size_t ng = (size_t)ngrid[0] * ngrid[1] * ngrid[2];
    RowMajor<3> ref{std::vector<double>(ng, 0.0), {ngrid[2], ngrid[1], ngrid[0]}};
    std::vector<char> dumped(ng, 0);
    for (int n = 0; n < ngnz; n++) {
        int i, j, k;
        double v;
        r.get("grid", "%d %d %d %lf", &i, &j, &k, &v);
        if (i < 0 || i >= ngrid[2] || j < 0 || j >= ngrid[1] || k < 0 || k >= ngrid[0]) r.die("grid index");
        ref(i, j, k) = v;
        dumped[((size_t)i * ngrid[1] + j) * ngrid[0] + k] = 1;
    }
    fclose(fp);
    const bool periodic[3] = {per[0] != 0, per[1] != 0, per[2] != 0};
    RowMajor<3> grid{std::vector<double>(ng, 0.0), {ngrid[2], ngrid[1], ngrid[0]}};
    printf("task %s: l=(%d..%d, %d..%d) tau=%d ortho=%d grid %d %d %d, %d cycles\n", argv[1], la_min,
           la_max, lb_min, lb_max, tau, ortho, ngrid[0], ngrid[1], ngrid[2], cycles);
    const auto t0 = std::chrono::steady_clock::now();
    try {
        for (int c = 0; c < cycles; c++)
            grid_collocate_pgf_product_rspace(tau != 0, ortho != 0, la_max, la_min, lb_max, lb_min,
                                              zeta, zetb, rscale, dh, dh_inv, ra, rab, npts, ngrid,
                                              lb_grid, periodic, lmax, radius, lb_cube, ub_cube,
                                              nspheres, sphere_bounds, maxco, o1, o2, pab, grid);
    } catch (const gpw::error &e) {
        fprintf(stderr, "collocate failed: status %d: %s\n", e.code, e.what());
        return 3;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double max_diff = 0.0, scale = 0.0;
    for (size_t q = 0; q < ng; q++) {
        if (!dumped[q]) continue;
        max_diff = std::max(max_diff, std::fabs(grid.data[q] - cycles * ref.data[q]));
        scale = std::max(scale, std::fabs(cycles * ref.data[q]));
    }
    printf("compared %d dumped values: max diff %.3e (scale %.3e), %.3f ms per cycle\n", ngnz, max_diff,
           scale, 1e3 * secs / std::max(cycles, 1));
    return max_diff <= 1e-12 * std::max(scale, 1e-300) ? 0 : 4;
}
