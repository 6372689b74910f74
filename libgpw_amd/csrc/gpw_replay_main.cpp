// gpw_replay -- replays CP2K "#Grid collocate task v1" files on the GPU and
// compares with the grid values the files hold.
//
// Counterpart of the reference's replay (src/lib/grid_base_ref_replay.cc:18-273,
// src/lib/grid_base_ref_miniapp.cc:15-33), written for this library:
//   * every grid entry of the file is compared (the reference's compare loop
//     only visits the x = 0 plane, src/lib/grid_base_ref_replay.cc:264); with
//     --complete the file lists every non-zero point (the .full.task goldens),
//     so every other point must come out zero;
//   * general (non-orthorhombic) tasks carry no lb_cube/ub_cube and are passed
//     through without dereferencing them (the reference dereferences NULL at
//     src/lib/grid_base_ref_replay.cc:226);
//   * errors return a non-zero exit status instead of asserting.
//
// Usage:
//   gpw_replay <task-file> [cycles] [--complete]
//       one task through the drop-in C++ signature (include/libgpw_amd.hpp),
//       `cycles` calls accumulating into one host grid; compares with
//       cycles x the file's values.
//   gpw_replay --batch <cycles> [--complete] <task-file>...
//       every file once as a pair of a device task list: files on the same
//       grid (npts, ngrid, window, periodicity, steps) with the same
//       compute_tau form one list, collocated `cycles` times into one device
//       grid; compares with cycles x the sum of the files' values. A file
//       whose geometry the task list would not reproduce exactly (the cube
//       rule of an orthorhombic list differs from its lb_cube / ub_cube /
//       sphere_bounds, or use_ortho disagrees with the cell) goes through the
//       drop-in instead, into the same grid.
// Exit status 0: max diff <= 1e-12 x scale for every compared grid.
#include <hip/hip_runtime_api.h>

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

struct Task {
    std::string path;
    int tau, ortho, la_max, la_min, lb_max, lb_min, lmax, nspheres, maxco, o1, o2;
    double zeta, zetb, rscale, radius, dh[3][3], dh_inv[3][3], ra[3], rab[3];
    int npts[3], ngrid[3], lb_grid[3], per[3], lb_cube[3] = {0, 0, 0}, ub_cube[3] = {0, 0, 0};
    std::vector<int> sphere_bounds;
    RowMajor<2> pab;
    std::vector<double> ref;     // file values, (z, y, x)
    std::vector<char> listed;    // points the file lists
    int nlisted = 0;
};

struct Reader {
    FILE *fp;
    const char *path;
    int line = 0;
    char buf[512];
    bool next() {
        line++;
        return fgets(buf, sizeof buf, fp) != nullptr;
    }
    [[noreturn]] void die(const char *what) {
        fprintf(stderr, "%s line %d: %s\n", path, line, what);
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

void read_task(const char *path, Task &t) {
    FILE *fp = fopen(path, "r");
    if (!fp) {
        perror(path);
        exit(2);
    }
    Reader r{fp, path};
    t.path = path;
    if (!r.next() || strcmp(r.buf, "#Grid collocate task v1\n") != 0) r.die("header");
    r.get("compute_tau", "%d", &t.tau);
    r.get("use_ortho", "%d", &t.ortho);
    r.get("la_max", "%d", &t.la_max);
    r.get("la_min", "%d", &t.la_min);
    r.get("lb_max", "%d", &t.lb_max);
    r.get("lb_min", "%d", &t.lb_min);
    r.get("zeta", "%lf", &t.zeta);
    r.get("zetb", "%lf", &t.zetb);
    r.get("rscale", "%lf", &t.rscale);
    for (int i = 0; i < 3; i++) {
        int j;
        r.get("dh", "%d %lf %lf %lf", &j, &t.dh[i][0], &t.dh[i][1], &t.dh[i][2]);
    }
    for (int i = 0; i < 3; i++) {
        int j;
        r.get("dh_inv", "%d %lf %lf %lf", &j, &t.dh_inv[i][0], &t.dh_inv[i][1], &t.dh_inv[i][2]);
    }
    r.get("ra", "%lf %lf %lf", &t.ra[0], &t.ra[1], &t.ra[2]);
    r.get("rab", "%lf %lf %lf", &t.rab[0], &t.rab[1], &t.rab[2]);
    r.get("npts", "%d %d %d", &t.npts[0], &t.npts[1], &t.npts[2]);
    r.get("ngrid", "%d %d %d", &t.ngrid[0], &t.ngrid[1], &t.ngrid[2]);
    r.get("lb_grid", "%d %d %d", &t.lb_grid[0], &t.lb_grid[1], &t.lb_grid[2]);
    r.get("periodic", "%d %d %d", &t.per[0], &t.per[1], &t.per[2]);
    r.get("lmax", "%d", &t.lmax);
    r.get("radius", "%lf", &t.radius);
    if (t.ortho) {
        r.get("lb_cube", "%d %d %d", &t.lb_cube[0], &t.lb_cube[1], &t.lb_cube[2]);
        r.get("ub_cube", "%d %d %d", &t.ub_cube[0], &t.ub_cube[1], &t.ub_cube[2]);
    }
    int nnz;
    r.get("nspheres", "%d", &t.nspheres);
    r.get("nspheres_nonzero", "%d", &nnz);
    t.sphere_bounds.assign(std::max(t.nspheres, 0), 0);
    for (int n = 0; n < nnz; n++) {
        int i, v;
        r.get("sphere_bounds", "%d %d", &i, &v);
        if (i < 0 || i >= t.nspheres) r.die("sphere_bounds index");
        t.sphere_bounds[i] = v;
    }
    r.get("maxco", "%d", &t.maxco);
    r.get("o1", "%d", &t.o1);
    r.get("o2", "%d", &t.o2);
    t.pab = RowMajor<2>{std::vector<double>((size_t)t.maxco * t.maxco), {t.maxco, t.maxco}};
    for (int i = 0; i < t.maxco; i++)
        for (int j = 0; j < t.maxco; j++) {
            int a, b;
            r.get("pab", "%d %d %lf", &a, &b, &t.pab(i, j));
        }
    r.get("ngrid_nonzero", "%d", &t.nlisted);
    const size_t ng = (size_t)t.ngrid[0] * t.ngrid[1] * t.ngrid[2];
    t.ref.assign(ng, 0.0);
    t.listed.assign(ng, 0);
    for (int n = 0; n < t.nlisted; n++) {
        int i, j, k;
        double v;
        r.get("grid", "%d %d %d %lf", &i, &j, &k, &v);
        if (i < 0 || i >= t.ngrid[2] || j < 0 || j >= t.ngrid[1] || k < 0 || k >= t.ngrid[0])
            r.die("grid index");
        const size_t q = ((size_t)i * t.ngrid[1] + j) * t.ngrid[0] + k;
        t.ref[q] = v;
        t.listed[q] = 1;
    }
    fclose(fp);
}

// grid += one call of the drop-in per cycle
bool dropin(const Task &t, int cycles, std::vector<double> &grid) {
    const bool periodic[3] = {t.per[0] != 0, t.per[1] != 0, t.per[2] != 0};
    RowMajor<3> g{std::vector<double>(), {t.ngrid[2], t.ngrid[1], t.ngrid[0]}};
    g.data.swap(grid);
    bool ok = true;
    try {
        for (int c = 0; c < cycles; c++)
            grid_collocate_pgf_product_rspace(t.tau != 0, t.ortho != 0, t.la_max, t.la_min, t.lb_max,
                                              t.lb_min, t.zeta, t.zetb, t.rscale, t.dh, t.dh_inv, t.ra,
                                              t.rab, t.npts, t.ngrid, t.lb_grid, periodic, t.lmax,
                                              t.radius, t.lb_cube, t.ub_cube, t.nspheres,
                                              t.sphere_bounds, t.maxco, t.o1, t.o2, t.pab, g);
    } catch (const gpw::error &e) {
        fprintf(stderr, "%s: collocate failed: status %d: %s\n", t.path.c_str(), e.code, e.what());
        ok = false;
    }
    g.data.swap(grid);
    return ok;
}

// compares grid with cycles x ref on the listed points (complete: every point)
bool compare(const char *what, const std::vector<double> &grid, const std::vector<double> &ref,
             const std::vector<char> &listed, int cycles, bool complete, double ms_per_cycle) {
    double max_diff = 0.0, scale = 0.0, outside = 0.0;
    long long n = 0;
    for (size_t q = 0; q < grid.size(); q++) {
        if (listed[q]) {
            n++;
            max_diff = std::max(max_diff, std::fabs(grid[q] - cycles * ref[q]));
            scale = std::max(scale, std::fabs(cycles * ref[q]));
        } else {
            outside = std::max(outside, std::fabs(grid[q]));
        }
    }
    if (complete) max_diff = std::max(max_diff, outside);
    printf("%s: compared %lld listed values%s: max diff %.3e (scale %.3e), %.3f ms per cycle\n", what, n,
           complete ? " and every other point" : "", max_diff, scale, ms_per_cycle);
    return max_diff <= 1e-12 * std::max(scale, 1e-300);
}

bool same_grid(const Task &a, const Task &b) {
    if (a.tau != b.tau) return false;
    for (int d = 0; d < 3; d++) {
        if (a.npts[d] != b.npts[d] || a.ngrid[d] != b.ngrid[d] || a.lb_grid[d] != b.lb_grid[d] ||
            a.per[d] != b.per[d])
            return false;
        for (int e = 0; e < 3; e++)
            if (a.dh[d][e] != b.dh[d][e] || a.dh_inv[d][e] != b.dh_inv[d][e]) return false;
    }
    return true;
}

gpw_grid_desc grid_of(const Task &t) {
    gpw_grid_desc g{};
    for (int d = 0; d < 3; d++) {
        g.npts[d] = t.npts[d];
        g.ngrid[d] = t.ngrid[d];
        g.lb_grid[d] = t.lb_grid[d];
        g.periodic[d] = t.per[d] != 0;
        for (int e = 0; e < 3; e++) {
            g.dh[d][e] = t.dh[d][e];
            g.dh_inv[d][e] = t.dh_inv[d][e];
        }
    }
    return g;
}

gpw_pair_desc pair_of(const Task &t, long long block_offset) {
    gpw_pair_desc p{};
    p.la_max = t.la_max;
    p.la_min = t.la_min;
    p.lb_max = t.lb_max;
    p.lb_min = t.lb_min;
    p.zeta = t.zeta;
    p.zetb = t.zetb;
    p.rscale = t.rscale;
    for (int d = 0; d < 3; d++) {
        p.ra[d] = t.ra[d];
        p.rab[d] = t.rab[d];
    }
    p.radius = t.radius;
    p.block_offset = block_offset;
    return p;
}

// whether a task list on this grid reproduces the task's own point set
bool batchable(const Task &t, const gpw_grid_desc &g) {
    bool ortho_cell = true;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (i != j && g.dh[i][j] != 0.0) ortho_cell = false;
    if ((t.ortho != 0) != ortho_cell) return false;
    if (!t.ortho) return true;
    const gpw_pair_desc p = pair_of(t, 0);
    int cc[3], lb[3], ub[3], n = 0;
    double roff[3];
    std::vector<int> sb(t.sphere_bounds.size() + 1);
    if (gpw_pair_geometry(&g, &p, t.tau, cc, roff, lb, ub, sb.data(), (int)sb.size(), &n) != GPW_OK)
        return false;
    // the file's array may be longer than the entries the cube uses (CP2K
    // dumps its whole buffer); the rest must be unused (zero)
    if (n > t.nspheres) return false;
    for (int d = 0; d < 3; d++)
        if (lb[d] != t.lb_cube[d] || ub[d] != t.ub_cube[d]) return false;
    for (int i = 0; i < t.nspheres; i++)
        if ((i < n ? sb[i] : 0) != t.sphere_bounds[i]) return false;
    return true;
}

#define HIP_OK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                     \
            return false;                                                               \
        }                                                                               \
    } while (0)

// one group of tasks on one grid: task list of the batchable ones, drop-in
// for the rest, all into one grid
bool run_group(const std::vector<const Task *> &group, int cycles, bool complete) {
    const Task &t0 = *group[0];
    const gpw_grid_desc g = grid_of(t0);
    const size_t ng = (size_t)t0.ngrid[0] * t0.ngrid[1] * t0.ngrid[2];
    std::vector<gpw_pair_desc> pairs;
    std::vector<double> pab;
    std::vector<const Task *> rest;
    for (const Task *t : group) {
        if (!batchable(*t, g)) {
            rest.push_back(t);
            continue;
        }
        // packed block ncoset(lb_max) x ncoset(la_max) of the file's pab
        const int na = (t->la_max + 1) * (t->la_max + 2) * (t->la_max + 3) / 6;
        const int nb = (t->lb_max + 1) * (t->lb_max + 2) * (t->lb_max + 3) / 6;
        pairs.push_back(pair_of(*t, (long long)pab.size()));
        for (int j = 0; j < nb; j++)
            for (int i = 0; i < na; i++) pab.push_back(t->pab(t->o2 + j, t->o1 + i));
    }
    std::vector<double> grid(ng, 0.0);
    const auto t_start = std::chrono::steady_clock::now();
    if (!pairs.empty()) {
        gpw_tasklist *tl = nullptr;
        if (gpw_tasklist_create(&g, (int)pairs.size(), pairs.data(), t0.tau, nullptr, 1, &tl) != GPW_OK) {
            fprintf(stderr, "task list: %s\n", gpw_last_error());
            return false;
        }
        double *d_pab = nullptr, *d_grid = nullptr;
        bool ok = hipMalloc((void **)&d_pab, pab.size() * sizeof(double)) == hipSuccess &&
                  hipMalloc((void **)&d_grid, ng * sizeof(double)) == hipSuccess &&
                  hipMemcpy(d_pab, pab.data(), pab.size() * sizeof(double), hipMemcpyHostToDevice) ==
                      hipSuccess &&
                  hipMemset(d_grid, 0, ng * sizeof(double)) == hipSuccess;
        for (int c = 0; ok && c < cycles; c++)
            if (gpw_tasklist_collocate(tl, d_pab, d_grid, nullptr) != GPW_OK) {
                fprintf(stderr, "collocate: %s\n", gpw_last_error());
                ok = false;
            }
        ok = ok && hipMemcpy(grid.data(), d_grid, ng * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
        if (d_pab) (void)hipFree(d_pab);
        if (d_grid) (void)hipFree(d_grid);
        gpw_tasklist_destroy(tl);
        if (!ok) return false;
    }
    for (const Task *t : rest)
        if (!dropin(*t, cycles, grid)) return false;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::vector<double> ref(ng, 0.0);
    std::vector<char> listed(ng, 0);
    for (const Task *t : group)
        for (size_t q = 0; q < ng; q++) {
            ref[q] += t->ref[q];
            listed[q] |= t->listed[q];
        }
    char what[256];
    snprintf(what, sizeof what, "batch of %zu tasks on %s grid %d %d %d (%zu in the task list, %zu drop-in)",
             group.size(), t0.ortho ? "an orthorhombic" : "a general", t0.ngrid[0], t0.ngrid[1],
             t0.ngrid[2], pairs.size(), rest.size());
    return compare(what, grid, ref, listed, cycles, complete, 1e3 * secs / std::max(cycles, 1));
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<const char *> files;
    bool batch = false, complete = false;
    int cycles = 1;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--batch")) {
            batch = true;
            if (i + 1 >= argc) break;
            cycles = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--complete")) {
            complete = true;
        } else if (!batch && !files.empty()) {
            cycles = atoi(argv[i]);
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.empty() || cycles < 1) {
        printf("Usage: gpw_replay <task-file> [cycles] [--complete]\n"
               "       gpw_replay --batch <cycles> [--complete] <task-file>...\n");
        return 1;
    }
    std::vector<Task> tasks(files.size());
    for (size_t i = 0; i < files.size(); i++) read_task(files[i], tasks[i]);
    if (!batch) {
        const Task &t = tasks[0];
        printf("task %s: l=(%d..%d, %d..%d) tau=%d ortho=%d grid %d %d %d, %d cycles\n", t.path.c_str(),
               t.la_min, t.la_max, t.lb_min, t.lb_max, t.tau, t.ortho, t.ngrid[0], t.ngrid[1], t.ngrid[2],
               cycles);
        std::vector<double> grid(t.ref.size(), 0.0);
        const auto t0 = std::chrono::steady_clock::now();
        if (!dropin(t, cycles, grid)) return 3;
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return compare(t.path.c_str(), grid, t.ref, t.listed, cycles, complete,
                       1e3 * secs / std::max(cycles, 1)) ? 0 : 4;
    }
    // group the files by grid, in order of first appearance
    std::vector<std::vector<const Task *>> groups;
    for (const Task &t : tasks) {
        bool placed = false;
        for (auto &gr : groups)
            if (same_grid(*gr[0], t)) {
                gr.push_back(&t);
                placed = true;
                break;
            }
        if (!placed) groups.push_back({&t});
    }
    bool ok = true;
    for (const auto &gr : groups) ok = run_group(gr, cycles, complete) && ok;
    return ok ? 0 : 4;
}
