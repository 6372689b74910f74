// gpw_kernels.hip -- HIP kernels for gfx950 (MI355X): collocate / integrate of
// primitive Gaussian pairs on an orthorhombic periodic (or windowed) grid.
//
// Math follows the reference's CP2K path (src/lib/grid_base_ref_c.cc):
//   value at cube offset (i, j, k), displacement X = i*dx - roff_x etc.:
//     sum_{lzp,lyp,lxp} coef[lzp][lyp][lxp] Z^lzp Y^lyp X^lxp exp(-zetp (X^2+Y^2+Z^2))
//   contracted in the order z, y, x (grid_base_ref_c.cc:431-494), restricted to
//   the sphere encoded by sphere_bounds (:413-470), mapped periodically (:318-351).
// Integrate is the exact transpose of that contraction.
//
// Work decomposition (MI355X-first, not a translation of the reference loops):
//   * tile kernels: one workgroup owns a TX x TY x TZ tile of the grid, each
//     thread a 4-row x 8-point register block of it; the workgroup sweeps every
//     (sub-box, tile) work item of its tile, staging per-pair 1D tables and
//     z-stage coefficients in LDS, Q pairs per barrier. Work items are
//     bucketed by polynomial length into classes, one kernel instantiation
//     each. Each grid point is written by exactly one workgroup (or one split
//     buffer): no atomics, results are deterministic.
//   * plain kernels: one launch per sub-box of one pair (the baseline form).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "gpw_device.hpp"
#include "gpw_host.hpp"

namespace gpw {

// cab -> cxyz of one pair per workgroup, factorised over the Cartesian axes
// (the three-step contraction of grid_base_ref_c.cc:234-315, in two passes):
//   Ty[az][bz][(ly,lx)] = sum_{ay,by} sum_{ax,bx} eff[b][a] alpha_x(bx,ax,lx) alpha_y(by,ay,ly)
//   coef[(lz,ly,lx)]    = prefactor sum_{az,bz} Ty[az][bz][(ly,lx)] alpha_z(bz,az,lz)
// with (ax, bx) over the shells of the (tau-shifted) block; Ty holds only
// ly + lx <= (lam - az) + (lbm - bz). Dynamic LDS: alpha, eff, Ty (+ cxyz).
// General cells (lt.general()): coef_ijk = T coef_xyz, T the multinomial
// expansion of grid_base_ref_c.cc:647-698 for each lp (tables, or the two
// passes of lattice_stepped).
// Workgroups loop over pairs (pair0 + blockIdx.x, stride gridDim.x); eff, Ty
// and cxyz live in LDS after alpha, or, when gws is given (blocks too large
// for LDS), in the workgroup's slice gws + blockIdx.x * ws_stride.
__global__ __launch_bounds__(TR_NT) void k_coef(const PairDev *__restrict__ pairs,
                                             const int *__restrict__ order, int pair0,
                                             int npairs, int tau, const double *__restrict__ pab,
                                             double *__restrict__ coef,
                                             LatticeTab lt,
                                             double *__restrict__ gws, long long ws_stride) {
    extern __shared__ double sm_dyn[];
    __shared__ int dec2[NC2MAX];
    // stepped cells: the transform's tables first in the dynamic LDS
    StepTabs &stt = *reinterpret_cast<StepTabs *>(sm_dyn);
    double *sm_coef = sm_dyn + (lt.stepped ? STEP_TABS_DOUBLES : 0);
    if (lt.stepped) fill_step_tabs(lt, stt, threadIdx.x, blockDim.x);
    for (int ip = pair0 + blockIdx.x; ip < pair0 + npairs; ip += gridDim.x) {
    __syncthreads();  // LDS / workspace reuse across pairs
    const int pi = order ? order[ip] : ip;
    const PairDev p = pairs[pi];
    const int lam = p.la_max + tau, lbm = p.lb_max + tau;
    const int lam_lo = max(p.la_min - tau, 0), lbm_lo = max(p.lb_min - tau, 0);
    const int lp = lam + lbm;
    AlphaView A(sm_coef, lam, lbm);
    const int nalpha = alpha_doubles(lam, lbm);
    const int nla = ncoset(lam), nlb = ncoset(lbm);
    const int n2 = ncoef2(lp), n3 = ncoset(lp);
    double *eff = gws ? gws + blockIdx.x * ws_stride : sm_coef + nalpha;
    double *Ty = eff + nla * nlb;
    double *cxyz = Ty + (lam + 1) * (lbm + 1) * n2;
    fill_alpha(&pairs[pi], A, threadIdx.x, blockDim.x);
    fill_decode2(dec2, lp, threadIdx.x, blockDim.x);
    build_eff(p, tau, pab, eff, nla, nlb, nla, threadIdx.x, blockDim.x);  // ends with a barrier
    const int nzz = (lam + 1) * (lbm + 1);
    for (int w = threadIdx.x; w < nzz * n2; w += blockDim.x) {
        const int zz = w / n2, az = zz / (lbm + 1), bz = zz % (lbm + 1);
        const int dc = dec2[w % n2], ly = dc & 255, lx = dc >> 8;
        if (ly + lx > (lam - az) + (lbm - bz)) continue;
        double acc = 0.0;
        for (int ay = 0; ay <= lam - az; ay++) {
            const int ax_lo = max(0, lam_lo - ay - az), ax_hi = lam - ay - az;
            for (int by = 0; by <= lbm - bz; by++) {
                if (ly > ay + by) continue;
                const double cy = A.at(1, by, ay, ly);
                const int bx_lo = max(0, lbm_lo - by - bz), bx_hi = lbm - by - bz;
                for (int ax = ax_lo; ax <= ax_hi; ax++) {
                    const double *e = eff + coset(ax, ay, az);
                    for (int bx = max(bx_lo, lx - ax); bx <= bx_hi; bx++)
                        acc += e[coset(bx, by, bz) * nla] * A.at(0, bx, ax, lx) * cy;
                }
            }
        }
        Ty[w] = acc;
    }
    __syncthreads();
    double *dst = lt.general() ? cxyz : coef + p.coef_off;
    for (int c = threadIdx.x; c < n3; c += blockDim.x) {
        int lz, ly, lx;
        decode_c3(lp, c, lz, ly, lx);
        const int c2 = c2idx(lp, ly, lx);
        double acc = 0.0;
        for (int az = 0; az <= lam; az++)
            for (int bz = max(0, lz - az); bz <= lbm; bz++) {
                if (ly + lx > (lam - az) + (lbm - bz)) break;
                acc += Ty[(az * (lbm + 1) + bz) * n2 + c2] * A.at(2, bz, az, lz);
            }
        // stepped cells: scaled by c2^lz for the transform
        dst[c] = p.prefactor * acc * (lt.stepped ? stt.pw[5][lz] : 1.0);
    }
    if (!lt.general()) continue;
    __syncthreads();
    // Ty (no longer read) is the stepped transform's scratch
    if (lt.stepped)
        lattice_stepped(stt, lp, 0, cxyz, Ty, coef + p.coef_off, threadIdx.x, blockDim.x);
    else
        lattice_transform(lt, lp, 0, cxyz, coef + p.coef_off, threadIdx.x, blockDim.x);
    }
}

// ---------------------------------------------------------------------------
// cab -> cxyz of one pair per wave (tau = 0), in slices of fixed (az, bz), so
// that the per-pair LDS is the alpha table, the pair's block and one slice,
// and many pairs are resident per CU (one 64-lane workgroup each; workgroup
// barriers of a single wave are cheap). Global memory is touched once per
// pair (block in, coefficients out); the slice loops read only LDS. The
// factorisation is the three-step contraction of grid_base_ref_c.cc:234-315,
// one axis per step.

constexpr int TW = 64;

// shells la_min..la_max of a block: first coset index and count
__device__ __forceinline__ void shell_range(int lmin, int lmax, int &lo, int &n) {
    lo = ncoset(lmin - 1);
    n = ncoset(lmax) - lo;
}

// LDS doubles of k_coef_wave for a block of up to (la, lb) whose shells hold
// at most nab (a, b) entries: alpha, the block, the slice S1[by][ay][lx] and
// (general cells) the xyz coefficients before the lattice transform
__host__ __device__ constexpr int coefw_doubles(int la, int lb, int nab) {
    return alpha_doubles(la, lb) + nab + (la + 1) * (lb + 1) * (la + lb + 1) + ncoset(la + lb);
}
// the part before the xyz coefficients (alpha, the smallest block of single
// shells, the slice) holds ncoset(la + lb) doubles for all la, lb <= LMAX:
// the scratch of the stepped lattice transform
constexpr bool coefw_scratch_fits() {
    for (int la = 0; la <= LMAX; la++)
        for (int lb = 0; lb <= LMAX; lb++)
            if (alpha_doubles(la, lb) + (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) +
                    (la + 1) * (lb + 1) * (la + lb + 1) < ncoset(la + lb))
                return false;
    return true;
}
static_assert(coefw_scratch_fits(), "k_coef_wave scratch of the stepped lattice transform");

// cab -> cxyz of one pair per wave: for each slice (az, bz)
//   S1[by][ay][lx] = sum_{ax, bx} pab[b][a] alpha_x(bx, ax, lx)
//   t(ly, lx)      = sum_{ay, by} alpha_y(by, ay, ly) S1[by][ay][lx]
//   c(lz, ly, lx) += t(ly, lx) alpha_z(bz, az, lz)
// with (ax, bx) over the shells la_min..la_max, lb_min..lb_max; lane s owns
// the columns (ly, lx) = c2 = lane + 64 s of c in registers. LPM bounds
// lp + 1 (register arrays).
template <int LPM>
__global__ __launch_bounds__(TW) void k_coef_wave(const PairDev *__restrict__ pairs,
                                                  const int *__restrict__ order, int pair0, int npairs,
                                                  const double *__restrict__ pab, double *__restrict__ coef,
                                                  LatticeTab lt) {
    constexpr int NCOL = (ncoef2(LPM - 1) + TW - 1) / TW;
    extern __shared__ double sm_dyn[];
    // stepped cells: the transform's tables first in the dynamic LDS
    StepTabs &stt = *reinterpret_cast<StepTabs *>(sm_dyn);
    double *smw = sm_dyn + (lt.stepped ? STEP_TABS_DOUBLES : 0);
    const int lane = threadIdx.x;
    if (lt.stepped) fill_step_tabs(lt, stt, lane, TW);
    for (int ip = pair0 + blockIdx.x; ip < pair0 + npairs; ip += gridDim.x) {
        __syncthreads();  // LDS reuse across pairs
        const int pi = order ? order[ip] : ip;
        const PairDev p = pairs[pi];
        const int la = p.la_max, lb = p.lb_max, lp = la + lb;
        int na_lo, nA, nb_lo, nB;
        shell_range(p.la_min, la, na_lo, nA);
        shell_range(p.lb_min, lb, nb_lo, nB);
        AlphaView A(smw, la, lb);
        double *P = smw + alpha_doubles(la, lb);  // P[jb][ia] = pab[nb_lo + jb][na_lo + ia]
        double *S1 = P + nA * nB;
        double *cx = S1 + (la + 1) * (lb + 1) * (lp + 1);
        fill_alpha(&pairs[pi], A, lane, TW);
        for (int w = lane; w < nA * nB; w += TW)
            P[w] = pab[p.block_off + (long long)(nb_lo + w / nA) * p.ld + na_lo + w % nA];
        const int n2 = ncoef2(lp);
        int ly[NCOL], lx[NCOL];
        double c[NCOL][LPM];
#pragma unroll
        for (int s = 0; s < NCOL; s++) {
            ly[s] = lx[s] = 0;
            if (lane + TW * s < n2) decode_c2(lp, lane + TW * s, ly[s], lx[s]);
#pragma unroll
            for (int k = 0; k < LPM; k++) c[s][k] = 0.0;
        }
        __syncthreads();
        const double *A0 = A.table(0), *A1 = A.table(1);
        for (int az = 0; az <= la; az++) {
            for (int bz = 0; bz <= lb; bz++) {
                const int ma = la - az, mb = lb - bz, X = ma + mb + 1;
                const int nyy = (ma + 1) * (mb + 1);
                // w -> (l, ay, by) by float reciprocals: exact quotients for w < 2^12
                const float rX = 1.0f / (float)X, rA = 1.0f / (float)(ma + 1);
                for (int w = lane; w < nyy * X; w += TW) {
                    const int yy = (int)(((float)w + 0.5f) * rX), l = w - yy * X;
                    const int by = (int)(((float)yy + 0.5f) * rA), ay = yy - by * (ma + 1);
                    const int ax_lo = max(0, p.la_min - ay - az), ax_hi = ma - ay;
                    const int bx_lo = max(0, p.lb_min - by - bz), bx_hi = mb - by;
                    double s = 0.0;
                    for (int ax = ax_lo; ax <= ax_hi; ax++) {
                        const double *col = P + coset(ax, ay, az) - na_lo;
                        const int bx0 = max(bx_lo, l - ax);
                        if (bx0 > bx_hi) continue;
                        // along bx: the block row and the alpha_x row step
                        // by closed-form differences (coset(bx + 1, by, bz) -
                        // coset(bx, by, bz) = (lb + 1)(lb + 2)/2, lb = bx + by + bz)
                        int lbx = bx0 + by + bz;
                        int jb = (coset(bx0, by, bz) - nb_lo) * nA;
                        int ao = A.rowoff(bx0, ax) + l;
                        for (int bx = bx0; bx <= bx_hi; bx++) {
                            s += col[jb] * A0[ao];
                            jb += ((lbx + 1) * (lbx + 2) / 2) * nA;
                            lbx++;
                            ao += A.rowstep(bx, ax);
                        }
                    }
                    S1[w] = s;
                }
                __syncthreads();
                const double *Az = A.table(2) + A.rowoff(bz, az);
#pragma unroll
                for (int s = 0; s < NCOL; s++) {
                    if (lane + TW * s >= n2 || ly[s] + lx[s] > ma + mb) continue;
                    double t = 0.0;
                    for (int ay = 0; ay <= ma; ay++) {
                        const int by0 = max(0, ly[s] - ay);
                        int ao = A.rowoff(by0, ay) + ly[s];
                        int si = (by0 * (ma + 1) + ay) * X + lx[s];
                        for (int by = by0; by <= mb; by++) {
                            t += A1[ao] * S1[si];
                            ao += A.rowstep(by, ay);
                            si += (ma + 1) * X;
                        }
                    }
#pragma unroll
                    for (int lz = 0; lz < LPM; lz++)
                        if (lz <= az + bz) c[s][lz] += t * Az[lz];
                }
                __syncthreads();
            }
        }
        double *dst = lt.general() ? cx : coef + p.coef_off;
#pragma unroll
        for (int s = 0; s < NCOL; s++) {
            if (lane + TW * s >= n2) continue;
#pragma unroll
            for (int lz = 0; lz < LPM; lz++)
                if (lz <= lp - ly[s] - lx[s])
                    dst[cidx(lp, lz, ly[s], lx[s])] =
                        p.prefactor * c[s][lz] * (lt.stepped ? stt.pw[5][lz] : 1.0);  // see k_coef
        }
        if (lt.general()) {
            __syncthreads();
            // stepped: the alpha table, block and slice (no longer read)
            // are the scratch; they hold ncoset(lp) doubles for every
            // la, lb <= LMAX (coefw_scratch_fits)
            if (lt.stepped)
                lattice_stepped(stt, lp, 0, cx, smw, coef + p.coef_off, lane, TW);
            else
                lattice_transform(lt, lp, 0, cx, coef + p.coef_off, lane, TW);
        }
    }
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int splits, long long ng,
                                  double *__restrict__ grid) {
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < ng;
         g += (long long)gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int s = 0; s < splits; s++) a += partial[s * ng + g];
        grid[g] += a;
    }
}

// ---------------------------------------------------------------------------
// moments per pair from the partial buffer, then hab / forces

__global__ void k_fill(double *__restrict__ p, long long n, double v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Mbuf[pair*mstride + c] = sum over the pair's entries of partial[ent_off[e] + c];
// entries are numbered pair-major, so the pair's partials are one contiguous run.
// One workgroup of SUM_NT threads per pair.
constexpr int SUM_NT = 64;
__global__ __launch_bounds__(SUM_NT) void k_sum_entries(const PairDev *__restrict__ pairs, int tau, int lp_add,
                              const int *__restrict__ pe_start, const long long *__restrict__ ent_off,
                              const double *__restrict__ partial, int mstride,
                              const unsigned char *__restrict__ skip, double *__restrict__ Mbuf) {
    const int p = blockIdx.x;
    if (skip && skip[p]) return;  // moments written by k_moments_small
    const PairDev P = pairs[p];
    const int lp = P.lp + lp_add;
    const int n3 = ncoset(lp);
    const int b = pe_start[p], e = pe_start[p + 1];
    const double *src = b < e ? partial + ent_off[b] : partial;
    const int ne = e - b;
    const int t = threadIdx.x;
    // short moment vectors: G = 64 / n3 groups of lanes take every G-th item,
    // then the group sums are added in group order (fixed order: the result
    // does not depend on scheduling)
    const int G = n3 <= SUM_NT / 2 ? SUM_NT / n3 : 1;
    if (G > 1) {
        __shared__ double red[SUM_NT];
        const int g = t / n3, c = t % n3;
        if (g < G) {
            double a0 = 0.0, a1 = 0.0;
            int k = g;
            for (; k + G < ne; k += 2 * G) {
                a0 += src[(long long)k * n3 + c];
                a1 += src[(long long)(k + G) * n3 + c];
            }
            if (k < ne) a0 += src[(long long)k * n3 + c];
            red[t] = a0 + a1;
        }
        __syncthreads();
        if (t < n3) {
            double a = 0.0;
            for (int q = 0; q < G; q++) a += red[q * n3 + t];
            Mbuf[(long long)p * mstride + t] = a;
        }
        return;
    }
    for (int c = t; c < n3; c += blockDim.x) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = 0;
        for (; k + 4 <= ne; k += 4) {
            a0 += src[(long long)k * n3 + c];
            a1 += src[(long long)(k + 1) * n3 + c];
            a2 += src[(long long)(k + 2) * n3 + c];
            a3 += src[(long long)(k + 3) * n3 + c];
        }
        for (; k < ne; k++) a0 += src[(long long)k * n3 + c];
        Mbuf[(long long)p * mstride + c] = (a0 + a1) + (a2 + a3);
    }
}

// heff = prefactor * alpha^T M over the extended block, then hab (rho copy /
// tau adjoint) and forces (d phi/dA = 2 zeta phi_{a+1} - a phi_{a-1}).
// Workgroups loop over pairs (pair0 + blockIdx.x, stride gridDim.x). heff,
// eff and U live in LDS, or, when gws is given (blocks too large for LDS), in
// the workgroup's slice gws + blockIdx.x * ws_stride of a global workspace.
// Storage kept small so that most blocks fit in LDS:
//   * U[az][bz] holds only its triangle ly + lx <= (lamx - az) + (lbmx - bz),
//     the slices packed one after the other (offsets uoff);
//   * with forces and without tau heff holds only the cross that the hab and
//     force stages read (hab_cross_entries): the block's own b shells with
//     every a shell of the extended block, then the b shells outside the
//     block with the block's own a shells; otherwise the dense extended block;
//   * without tau the forces weight heff with pab itself (no eff copy).
// Work unit: a segment of pairs that share one hab block (segs: CSR over
// positions of `order`, segments seg0 .. seg0 + nseg); one workgroup walks a
// segment's pairs in order, so the += into the shared block is race-free and
// deterministic. segs NULL: one pair per segment, positions seg0 .. seg0 + nseg.
// Registers capped for eight waves per SIMD (more pairs in flight; measured
// faster than the uncapped 75 VGPRs). LK: the cell's lattice transform (0
// none: orthorhombic; 1 tables; 2 stepped), one instance each so that the
// orthorhombic kernel carries no transform code.
template <int LK>
__global__ __launch_bounds__(TR_NT, 8) void k_hab(const PairDev *__restrict__ pairs,
                                            const int *__restrict__ order,
                                            const int *__restrict__ segs, int seg0, int nseg,
                                            int tau, int with_forces, const double *__restrict__ Mbuf,
                                            int mstride, const double *__restrict__ pab,
                                            double *__restrict__ hab, double *__restrict__ forces,
                                            LatticeTab lt,
                                            double *__restrict__ gws, long long ws_stride) {
    extern __shared__ double sm_dyn[];
    __shared__ int dec[NCOSET_LMAX];
    __shared__ int dec2[NC2MAX];
    __shared__ int uoff[(LMAX + 1) * (LMAX + 1)];
    // stepped cells: the transform's tables first in the dynamic LDS
    StepTabs &stt = *reinterpret_cast<StepTabs *>(sm_dyn);
    double *sm_hab = sm_dyn + ((LK == 2) ? STEP_TABS_DOUBLES : 0);
    fill_decode(dec, threadIdx.x, blockDim.x);  // first read after the barrier below
    if ((LK == 2)) fill_step_tabs(lt, stt, threadIdx.x, blockDim.x);
    const int fx = with_forces ? 1 : 0;
    const bool cross = with_forces && !tau;
    for (int sg = seg0 + blockIdx.x; sg < seg0 + nseg; sg += gridDim.x) {
    const int ip_end = segs ? segs[sg + 1] : sg + 1;
    for (int ip = segs ? segs[sg] : sg; ip < ip_end; ip++) {
    __syncthreads();  // LDS / workspace reuse across pairs
    const int pi = order ? order[ip] : ip;
    const PairDev p = pairs[pi];
    const int lam = p.la_max + tau, lbm = p.lb_max + tau;          // local
    const int lamx = lam + fx, lbmx = lbm + fx;                    // extended
    const int lamx_lo = max(max(p.la_min - tau, 0) - fx, 0);
    const int lbmx_lo = max(max(p.lb_min - tau, 0) - fx, 0);
    const int lpx = lamx + lbmx;
    AlphaView A(sm_hab, lamx, lbmx);
    const int nalpha = alpha_doubles(lamx, lbmx);
    const int nlax = ncoset(lamx), nlbx = ncoset(lbmx);
    const int n3x = ncoset(lpx);
    const int n2x = ncoef2(lpx);
    const int nzz = (lamx + 1) * (lbmx + 1);
    double *M = sm_hab + nalpha;                    // ncoset(lpx)
    double *Mijk = M + n3x;                         // general cells: lattice moments
    double *red = Mijk + n3x;                       // 6 x (waves per workgroup)
    // only the shells lamx_lo..lamx, lbmx_lo..lbmx: the hab and force stages
    // read nothing else
    const int na_lo = ncoset(lamx_lo - 1), nb_lo = ncoset(lbmx_lo - 1);
    const int nAx = nlax - na_lo, nBx = nlbx - nb_lo;
    const int na_mid = ncoset(p.la_min - 1), nAmid = ncoset(p.la_max) - na_mid;
    const int nb_mid = ncoset(p.lb_min - 1), nBmid = ncoset(p.lb_max) - nb_mid;
    const int n1 = cross ? nBmid * nAx : nAx * nBx;
    const int nout = cross ? n1 + (nBx - nBmid) * nAmid : n1;
    // heff (nout cross entries, or nlbx x nlax), eff (tau only), U (packed slices)
    double *heff = gws ? gws + blockIdx.x * ws_stride : red + 6 * (TR_NT / 64);
    double *eff = heff + (cross ? nout : nlax * nlbx);
    double *U = eff + (tau ? ncoset(lam) * ncoset(lbm) : 0);
    fill_alpha(&pairs[pi], A, threadIdx.x, blockDim.x);
    fill_decode2(dec2, lpx, threadIdx.x, blockDim.x);
    // offset of slice (az, bz): the triangles of the slices before it, in
    // closed form (sums of ncoset over a range are differences of c4)
    for (int zz = threadIdx.x; zz < nzz; zz += blockDim.x) {
        const int az = zz / (lbmx + 1), bz = zz % (lbmx + 1), ma = lamx - az;
        uoff[zz] = (c4(lpx) - c4(lpx - az)) - (c4(lamx - 1) - c4(lamx - 1 - az)) +
                   (ncoset(ma + lbmx) - ncoset(ma + lbmx - bz));
    }
    if (!(LK != 0)) {
        for (int c = threadIdx.x; c < n3x; c += blockDim.x) M[c] = Mbuf[(long long)pi * mstride + c];
    } else {
        // M_xyz = T^T M_ijk (adjoint of the lattice transform); stepped
        // cells: scaled by a0^i for the transform
        for (int c = threadIdx.x; c < n3x; c += blockDim.x) {
            double v = Mbuf[(long long)pi * mstride + c];
            if ((LK == 2)) {
                int iz, iy, ix;
                decode_c3(lpx, c, iz, iy, ix);
                v *= stt.pw[0][ix];
            }
            Mijk[c] = v;
        }
        __syncthreads();
        if ((LK == 2)) {
            lattice_stepped(stt, lpx, 1, Mijk, M, M, threadIdx.x, blockDim.x);
        } else {
            lattice_transform(lt, lpx, 1, Mijk, M, threadIdx.x, blockDim.x);
        }
    }
    __syncthreads();
    // adjoint of the factorised cab -> cxyz transform (see k_coef):
    //   U[az][bz][(ly,lx)] = sum_lz M[(lz,ly,lx)] alpha_z(bz,az,lz)
    //   heff[b][a] = prefactor sum_{ly,lx} U[az][bz][(ly,lx)] alpha_y(by,ay,ly) alpha_x(bx,ax,lx)
    const float rn2 = __builtin_amdgcn_rcpf((float)n2x), rnb = __builtin_amdgcn_rcpf((float)(lbmx + 1));
    for (int w = threadIdx.x; w < nzz * n2x; w += blockDim.x) {
        // quotients by float reciprocals, corrected to the exact ones
        int zz = (int)((float)w * rn2);
        zz -= zz * n2x > w;
        zz += (zz + 1) * n2x <= w;
        int az = (int)((float)zz * rnb);
        az -= az * (lbmx + 1) > zz;
        az += (az + 1) * (lbmx + 1) <= zz;
        const int bz = zz - az * (lbmx + 1);
        const int dc = dec2[w - zz * n2x], ly = dc & 255, lx = dc >> 8;
        const int nl = (lamx - az) + (lbmx - bz);
        if (ly + lx > nl) continue;
        // two accumulators: shorter dependency chains; the moment index
        // steps by cidx(lpx, lz + 1, ly, lx) - cidx(lpx, lz, ly, lx) =
        // (m + 1)(m + 2)/2 - ly, m = lpx - lz
        double acc = 0.0;
        const double *Az = &A.at(2, bz, az, 0);
        double acc2 = 0.0;
        int lz = 0, at = cidx(lpx, 0, ly, lx);
        for (; lz + 1 <= az + bz; lz += 2) {
            const int m = lpx - lz;
            const int at1 = at + (m + 1) * (m + 2) / 2 - ly;
            acc += M[at] * Az[lz];
            acc2 += M[at1] * Az[lz + 1];
            at = at1 + m * (m + 1) / 2 - ly;
        }
        if (lz <= az + bz) acc += M[at] * Az[lz];
        acc += acc2;
        U[uoff[zz] + c2idx(nl, ly, lx)] = acc;
    }
    __syncthreads();
    // heff entry of (jco, ico): cross storage or dense
    auto hix = [&](int jco, int ico) -> int {
        if (!cross) return jco * nlax + ico;
        if (jco >= nb_mid && jco < nb_mid + nBmid) return (jco - nb_mid) * nAx + (ico - na_lo);
        const int r = jco < nb_mid ? jco - nb_lo : (jco - nb_mid - nBmid) + (nb_mid - nb_lo);
        return n1 + r * nAmid + (ico - na_mid);
    };
    // Forces without tau read a cross of the extended block: the block's own
    // b shells with every a shell, and its own a shells with every b shell
    // (not the corners a +- 1 with b +- 1), so only the cross is built.
    const double *A0 = A.table(0), *A1 = A.table(1);
    const float rAx = __builtin_amdgcn_rcpf((float)nAx), rAmid = __builtin_amdgcn_rcpf((float)nAmid);
    for (int w = threadIdx.x; w < nout; w += blockDim.x) {
        int ico, jco;
        if (w < n1) {
            int q, r;
            fdivmod(w, nAx, rAx, q, r);
            ico = na_lo + r;
            jco = (cross ? nb_mid : nb_lo) + q;
        } else {
            const int nlow = nb_mid - nb_lo;
            int r, rem;
            fdivmod(w - n1, nAmid, rAmid, r, rem);
            ico = na_mid + rem;
            jco = r < nlow ? nb_lo + r : nb_mid + nBmid + (r - nlow);
        }
        const int da = dec[ico], db = dec[jco];
        const int ax = da & 15, ay = (da >> 4) & 15, az = da >> 8;
        const int bx = db & 15, by = (db >> 4) & 15, bz = db >> 8;
        const int nl = (lamx - az) + (lbmx - bz);
        const double *Uz = U + uoff[az * (lbmx + 1) + bz];
        const double *Ax = A0 + A.rowoff(bx, ax), *Ay = A1 + A.rowoff(by, ay);
        double acc = 0.0;
        const double *Ur = Uz;  // row ly of the slice's (ly, lx) triangle over nl
        for (int ly = 0; ly <= ay + by; ly++, Ur += nl + 2 - ly) {
            double sx = 0.0;
            double sx2 = 0.0;
            int lx = 0;
            for (; lx + 1 <= ax + bx; lx += 2) {
                sx += Ur[lx] * Ax[lx];
                sx2 += Ur[lx + 1] * Ax[lx + 1];
            }
            if (lx <= ax + bx) sx += Ur[lx] * Ax[lx];
            sx += sx2;
            acc += sx * Ay[ly];
        }
        heff[cross ? w : jco * nlax + ico] = p.prefactor * acc;
    }
    __syncthreads();
    // hab: loop over the original block entries
    const int na0 = ncoset(p.la_max), na_lo0 = ncoset(p.la_min - 1);
    const int nb0 = ncoset(p.lb_max), nb_lo0 = ncoset(p.lb_min - 1);
    const int nA = na0 - na_lo0, nB = nb0 - nb_lo0;
    const float rA = __builtin_amdgcn_rcpf((float)nA);
    for (int w = threadIdx.x; w < nA * nB; w += blockDim.x) {
        int q, r;
        fdivmod(w, nA, rA, q, r);
        const int ico = na_lo0 + r, jco = nb_lo0 + q;
        double acc;
        if (!tau) {
            acc = heff[hix(jco, ico)];
        } else {
            const int da = dec[ico], db = dec[jco];
            const int a[3] = {da & 15, (da >> 4) & 15, da >> 8};
            const int b[3] = {db & 15, (db >> 4) & 15, db >> 8};
            acc = 0.0;
            for (int d = 0; d < 3; d++) {
                int am[3] = {a[0], a[1], a[2]}, ap[3] = {a[0], a[1], a[2]};
                int bm[3] = {b[0], b[1], b[2]}, bp[3] = {b[0], b[1], b[2]};
                am[d] = a[d] > 0 ? a[d] - 1 : 0;
                bm[d] = b[d] > 0 ? b[d] - 1 : 0;
                ap[d] = a[d] + 1;
                bp[d] = b[d] + 1;
                const int iam = coset(am[0], am[1], am[2]), iap = coset(ap[0], ap[1], ap[2]);
                const int jbm = coset(bm[0], bm[1], bm[2]), jbp = coset(bp[0], bp[1], bp[2]);
                acc += 0.5 * a[d] * b[d] * heff[jbm * nlax + iam];
                acc += 0.5 * (-2.0 * a[d] * p.zetb) * heff[jbp * nlax + iam];
                acc += 0.5 * (-2.0 * p.zeta * b[d]) * heff[jbm * nlax + iap];
                acc += 0.5 * (4.0 * p.zeta * p.zetb) * heff[jbp * nlax + iap];
            }
        }
        hab[p.block_off + (long long)jco * p.ld + ico] += acc;
    }
    if (!with_forces) continue;
    const int nla = ncoset(lam), nlb = ncoset(lbm);
    if (tau) build_eff(p, tau, pab, eff, nla, nlb, nla, threadIdx.x, blockDim.x);
    double f[6] = {0, 0, 0, 0, 0, 0};
    const int la_lo = ncoset(max(p.la_min - tau, 0) - 1), lb_lo = ncoset(max(p.lb_min - tau, 0) - 1);
    // the entries of shells la_lo.., lb_lo.. only
    const int nAf = nla - la_lo, nBf = nlb - lb_lo;
    const float rAf = __builtin_amdgcn_rcpf((float)nAf);
    for (int w = threadIdx.x; w < nAf * nBf; w += blockDim.x) {
        int jco, ico;
        fdivmod(w, nAf, rAf, jco, ico);
        ico += la_lo;
        jco += lb_lo;
        const double e = tau ? eff[jco * nla + ico] : pab[p.block_off + (long long)jco * p.ld + ico];
        if (e == 0.0) continue;
        const int da = dec[ico], db = dec[jco];
        const int a[3] = {da & 15, (da >> 4) & 15, da >> 8};
        const int b[3] = {db & 15, (db >> 4) & 15, db >> 8};
        for (int d = 0; d < 3; d++) {
            int ap[3] = {a[0], a[1], a[2]}, am[3] = {a[0], a[1], a[2]};
            int bp[3] = {b[0], b[1], b[2]}, bm[3] = {b[0], b[1], b[2]};
            ap[d]++;
            bp[d]++;
            double ga = 2.0 * p.zeta * heff[hix(jco, coset(ap[0], ap[1], ap[2]))];
            double gb = 2.0 * p.zetb * heff[hix(coset(bp[0], bp[1], bp[2]), ico)];
            if (a[d] > 0) {
                am[d]--;
                ga -= a[d] * heff[hix(jco, coset(am[0], am[1], am[2]))];
            }
            if (b[d] > 0) {
                bm[d]--;
                gb -= b[d] * heff[hix(coset(bm[0], bm[1], bm[2]), ico)];
            }
            f[d] += e * ga;
            f[3 + d] += e * gb;
        }
    }
    // sums over the wave (butterfly), then over the waves
    const int nw = (int)(blockDim.x / 64);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = f[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[k * (TR_NT / 64) + (threadIdx.x >> 6)] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = 0.0;
        for (int i = 0; i < nw; i++) s += red[threadIdx.x * (TR_NT / 64) + i];
        forces[(long long)pi * 6 + threadIdx.x] += s;
    }
    }
    }
}

// ---------------------------------------------------------------------------
// Integrate of small spheres: one workgroup per pair whose cube is a single
// sub-box of at most SMALL_NX x SMALL_NY x SMALL_NZ points. The tile kernels
// sweep whole 32 x 16 x 16 tiles per (sub-box, tile) item, which for a sphere
// a few points wide is mostly overhead; here the threads take the rows of the
// sub-box instead. x stage along each row in registers, the Gaussian by the
// recursion e(i+1) = e(i) rho(i), rho(i+1) = rho(i) kappa started at the
// row's first point inside the sphere (no underflow); then the y and z stages
// through LDS, in passes of whole z-planes of at most NROW rows. The point set
// of a row is the tile kernels' (orthorhombic: the sphere profile; general:
// the reference's quadratic bounds). Writes the pair's moments Mbuf[pair]
// (k_sum_entries skips these pairs).

constexpr int SMALL_PMAX = 16;  // z-planes per pass (LDS of t2)

template <int LM>
struct SmallSmem {
    double rows[NROW][LM];                    // x-stage result per row of the pass
    double Yt[SMALL_NY][LM], Zt[SMALL_NZ][LM];  // powers of the row displacements
    double t2[SMALL_PMAX][LM * (LM + 1) / 2]; // per plane of the pass: (ly, lx) sums over y
    double M[LM * (LM + 1) * (LM + 2) / 6];
};

template <int LM>
__global__ __launch_bounds__(NROW) void k_moments_small(TileGrid tg, const SubDev *__restrict__ subs,
                                                        const int *__restrict__ list,
                                                        const int *__restrict__ prof,
                                                        const double *__restrict__ grid, int lp_add,
                                                        int mstride, double *__restrict__ Mbuf) {
    __shared__ SmallSmem<LM> sm;
    const SubDev s = subs[list[blockIdx.x]];
    const bool gen = tg.general != 0;
    const int lp = s.lp + lp_add, L = lp + 1;  // L <= LM (host classification)
    const int ny = s.ub[1] - s.lb[1] + 1, nz = s.ub[2] - s.lb[2] + 1;
    const int t = threadIdx.x;
    const int n2 = ncoef2(lp), n3 = ncoset(lp);
    for (int w = t; w < ny + nz; w += NROW) {
        const int d = w < ny ? 1 : 2, idx = w < ny ? w : w - ny;
        const int o = s.lb[d] + idx;
        const double r = gen ? __dsub_rn((double)o, s.roff[d]) : o * tg.dr[d] - s.roff[d];
        double *dst = d == 1 ? sm.Yt[idx] : sm.Zt[idx];
        double v = 1.0;
        for (int l = 0; l < L; l++) {
            dst[l] = v;
            v *= r;
        }
    }
    for (int c = t; c < n3; c += NROW) sm.M[c] = 0.0;
    // x recursion: exponent step kap2 between consecutive ratios
    double qa;
    if (gen) {
        double a = 0.0;
        for (int x = 0; x < 3; x++) a += tg.dh[0][x] * tg.dh[0][x];
        qa = -s.zetp * a;
    } else {
        qa = -s.zetp * tg.dr[0] * tg.dr[0];
    }
    const double kap = exp(2.0 * qa);
    const int P = max(1, min(min(nz, SMALL_PMAX), NROW / ny));  // planes per pass
    for (int k0 = 0; k0 < nz; k0 += P) {
        const int np = min(P, nz - k0);
        double tx[LM];
#pragma unroll
        for (int l = 0; l < LM; l++) tx[l] = 0.0;
        if (t < np * ny) {
            const int y = t % ny, zp = t / ny;
            const int j = s.lb[1] + y, k = s.lb[2] + k0 + zp;
            int ilo = s.lb[0], ihi = s.ub[0];
            double ex0 = 0.0, d0 = 0.0;
            if (gen) {
                const double dj = __dsub_rn((double)j, s.roff[1]), dk = __dsub_rn((double)k, s.roff[2]);
                int lo, hi;
                double qb_, qc_;
                general_row_dev(tg, s.roff[0], dj, dk, s.R2, s.zetp, lo, hi, qb_, qc_);
                ilo = max(ilo, lo);
                ihi = min(ihi, hi);
                // exponent qa u^2 + qb u + qc of the displacement u = i - gp0
                double b = 0.0, c = 0.0;
                for (int x = 0; x < 3; x++) {
                    const double w = tg.dh[1][x] * dj + tg.dh[2][x] * dk;
                    b += w * tg.dh[0][x];
                    c += w * w;
                }
                const double qb = -2.0 * s.zetp * b, qc = -s.zetp * c;
                const double u0 = (double)ilo - s.roff[0];
                ex0 = (qa * u0 + qb) * u0 + qc;
                d0 = qa * (2.0 * u0 + 1.0) + qb;
            } else {
                const int T = row_threshold(prof, s.prof_off, s.prof_stride, fold_off(k), fold_off(j));
                if (T < 0) {
                    ihi = ilo - 1;
                } else {
                    ilo = max(ilo, -T);
                    ihi = min(ihi, T + 1);
                }
                const double rx = ilo * tg.dr[0] - s.roff[0];
                const double ry = j * tg.dr[1] - s.roff[1], rz = k * tg.dr[2] - s.roff[2];
                ex0 = -s.zetp * (rx * rx + ry * ry + rz * rz);
                d0 = -s.zetp * (2.0 * rx * tg.dr[0] + tg.dr[0] * tg.dr[0]);
            }
            if (ilo <= ihi) {
                const int gz = wrap_once(s.base[2] + (k - s.lb[2]), s.period[2]);
                const int gy = wrap_once(s.base[1] + (j - s.lb[1]), s.period[1]);
                const double *row = grid + ((long long)gz * tg.ny + gy) * tg.nx;
                double e = exp(ex0), rho = exp(d0);
                int gx = wrap_once(s.base[0] + (ilo - s.lb[0]), s.period[0]);  // then stepped, wrapped
                const int P0 = s.period[0];
                int i = ilo;
                // four points per step: independent power chains, loads issued
                // together; the sums keep the point order of the one-point loop
                for (; i + 3 <= ihi; i += 4) {
                    const int g1 = gx + 1 == P0 ? 0 : gx + 1;
                    const int g2 = g1 + 1 == P0 ? 0 : g1 + 1;
                    const int g3 = g2 + 1 == P0 ? 0 : g2 + 1;
                    double p[4] = {row[gx], row[g1], row[g2], row[g3]};
                    gx = g3 + 1 == P0 ? 0 : g3 + 1;
                    double u[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        u[q] = gen ? (double)(i + q) - s.roff[0] : (i + q) * tg.dr[0] - s.roff[0];
                        p[q] *= e;
                        e *= rho;
                        rho *= kap;
                    }
#pragma unroll
                    for (int l = 0; l < LM; l++) {
                        if (l < L) {
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                tx[l] += p[q];
                                p[q] *= u[q];
                            }
                        }
                    }
                }
                for (; i <= ihi; i++) {
                    const double u = gen ? (double)i - s.roff[0] : i * tg.dr[0] - s.roff[0];
                    double p = row[gx] * e;
                    gx = gx + 1 == P0 ? 0 : gx + 1;
#pragma unroll
                    for (int l = 0; l < LM; l++) {
                        if (l < L) {
                            tx[l] += p;
                            p *= u;
                        }
                    }
                    e *= rho;
                    rho *= kap;
                }
            }
#pragma unroll
            for (int l = 0; l < LM; l++)
                if (l < L) sm.rows[t][l] = tx[l];
        }
        __syncthreads();
        // y and z stages with four partial sums (shorter dependency chains)
        for (int w = t; w < np * n2; w += NROW) {
            const int zp = w / n2;
            int ly, lx;
            decode_c2(lp, w % n2, ly, lx);
            const double *rw = &sm.rows[zp * ny][lx];
            double a[4] = {0.0, 0.0, 0.0, 0.0};
            int y = 0;
            for (; y + 3 < ny; y += 4)
#pragma unroll
                for (int q = 0; q < 4; q++) a[q] += sm.Yt[y + q][ly] * rw[(y + q) * LM];
            for (; y < ny; y++) a[0] += sm.Yt[y][ly] * rw[y * LM];
            sm.t2[zp][w % n2] = (a[0] + a[1]) + (a[2] + a[3]);
        }
        __syncthreads();
        for (int c = t; c < n3; c += NROW) {
            int lz, ly, lx;
            decode_c3(lp, c, lz, ly, lx);
            const int c2 = c2idx(lp, ly, lx);
            double a[4] = {0.0, 0.0, 0.0, 0.0};
            int zp = 0;
            for (; zp + 3 < np; zp += 4)
#pragma unroll
                for (int q = 0; q < 4; q++) a[q] += sm.Zt[k0 + zp + q][lz] * sm.t2[zp + q][c2];
            for (; zp < np; zp++) a[0] += sm.Zt[k0 + zp][lz] * sm.t2[zp][c2];
            sm.M[c] += (a[0] + a[1]) + (a[2] + a[3]);
        }
        __syncthreads();
    }
    for (int c = t; c < n3; c += NROW) Mbuf[(long long)s.pair * mstride + c] = sm.M[c];
}

// ---------------------------------------------------------------------------
// Collocate of small spheres, in two passes so that every grid point is still
// summed in one fixed order (no atomics):
//  1. k_collocate_small: one workgroup per pair whose cube is a single
//     sub-box of at most CSMALL_N points per axis; its values on the whole
//     box (zero outside the sphere) go to the pair's slice of a scratch
//     buffer, box-major (z, y, x). Same stages as k_moments_small in the
//     collocate direction: z stage cxy = sum_lz coef Z^lz in LDS, then one
//     thread per row: cx = sum_ly cxy Y^ly, and the points of the row's
//     sphere range by Horner in the displacement times the Gaussian
//     recursion started at the range's first point.
//  2. k_gather_small: one workgroup per tile adds the boxes of the tile's
//     small pairs (its tile list, in item order) to the tile's grid points.
// Tile kernels leave these pairs out of their lists.

template <int LM>
struct CSmallSmem {
    double coef[LM * (LM + 1) * (LM + 2) / 6];
    double Yt[CSMALL_N][LM], Zt[CSMALL_N][LM];
    double cxy[CSMALL_N][LM * (LM + 1) / 2];
};

template <int LM>
__global__ __launch_bounds__(NROW) void k_collocate_small(TileGrid tg, const SubDev *__restrict__ subs,
                                                          const int *__restrict__ list,
                                                          const double *__restrict__ coef,
                                                          const int *__restrict__ prof,
                                                          const long long *__restrict__ box_off,
                                                          double *__restrict__ sval) {
    __shared__ CSmallSmem<LM> sm;
    const int si = list[blockIdx.x];
    const SubDev s = subs[si];
    const bool gen = tg.general != 0;
    const int lp = s.lp, L = lp + 1;  // L <= LM (host classification)
    const int nx = s.ub[0] - s.lb[0] + 1, ny = s.ub[1] - s.lb[1] + 1, nz = s.ub[2] - s.lb[2] + 1;
    const int t = threadIdx.x;
    const int n2 = ncoef2(lp), n3 = ncoset(lp);
    for (int c = t; c < n3; c += NROW) sm.coef[c] = coef[s.coef_off + c];
    for (int w = t; w < ny + nz; w += NROW) {
        const int d = w < ny ? 1 : 2, idx = w < ny ? w : w - ny;
        const int o = s.lb[d] + idx;
        const double r = gen ? __dsub_rn((double)o, s.roff[d]) : o * tg.dr[d] - s.roff[d];
        double *dst = d == 1 ? sm.Yt[idx] : sm.Zt[idx];
        double v = 1.0;
        for (int l = 0; l < L; l++) {
            dst[l] = v;
            v *= r;
        }
    }
    __syncthreads();
    // z stage: cxy[z][(ly, lx)] = sum_lz coef[(lz, ly, lx)] Z[z]^lz
    for (int w = t; w < nz * n2; w += NROW) {
        const int z = w / n2, c2 = w % n2;
        int ly, lx;
        decode_c2(lp, c2, ly, lx);
        double a = 0.0;
        for (int lz = 0; lz <= lp - ly - lx; lz++) a += sm.coef[cidx(lp, lz, ly, lx)] * sm.Zt[z][lz];
        sm.cxy[z][c2] = a;
    }
    __syncthreads();
    double qa;
    if (gen) {
        double a = 0.0;
        for (int x = 0; x < 3; x++) a += tg.dh[0][x] * tg.dh[0][x];
        qa = -s.zetp * a;
    } else {
        qa = -s.zetp * tg.dr[0] * tg.dr[0];
    }
    const double kap = exp(2.0 * qa);
    double *box = sval + box_off[si];
    for (int row = t; row < ny * nz; row += NROW) {
        const int y = row % ny, z = row / ny;
        const int j = s.lb[1] + y, k = s.lb[2] + z;
        // y stage of the row
        double cx[LM];
#pragma unroll
        for (int l = 0; l < LM; l++) cx[l] = 0.0;
        for (int ly = 0; ly <= lp; ly++) {
            const double yv = sm.Yt[y][ly];
            const double *c = sm.cxy[z] + c2idx(lp, ly, 0);
#pragma unroll
            for (int lx = 0; lx < LM; lx++)
                if (lx <= lp - ly) cx[lx] += c[lx] * yv;
        }
        // the row's points inside the sphere (the tile kernels' set)
        int ilo = s.lb[0], ihi = s.ub[0];
        double ex0 = 0.0, d0 = 0.0;
        if (gen) {
            const double dj = __dsub_rn((double)j, s.roff[1]), dk = __dsub_rn((double)k, s.roff[2]);
            int lo, hi;
            double qb_, qc_;
            general_row_dev(tg, s.roff[0], dj, dk, s.R2, s.zetp, lo, hi, qb_, qc_);
            ilo = max(ilo, lo);
            ihi = min(ihi, hi);
            double b = 0.0, c = 0.0;
            for (int x = 0; x < 3; x++) {
                const double w = tg.dh[1][x] * dj + tg.dh[2][x] * dk;
                b += w * tg.dh[0][x];
                c += w * w;
            }
            const double qb = -2.0 * s.zetp * b, qc = -s.zetp * c;
            const double u0 = (double)ilo - s.roff[0];
            ex0 = (qa * u0 + qb) * u0 + qc;
            d0 = qa * (2.0 * u0 + 1.0) + qb;
        } else {
            const int T = row_threshold(prof, s.prof_off, s.prof_stride, fold_off(k), fold_off(j));
            if (T < 0) {
                ihi = ilo - 1;
            } else {
                ilo = max(ilo, -T);
                ihi = min(ihi, T + 1);
            }
            const double rx = ilo * tg.dr[0] - s.roff[0];
            const double ry = j * tg.dr[1] - s.roff[1], rz = k * tg.dr[2] - s.roff[2];
            ex0 = -s.zetp * (rx * rx + ry * ry + rz * rz);
            d0 = -s.zetp * (2.0 * rx * tg.dr[0] + tg.dr[0] * tg.dr[0]);
        }
        double *out = box + ((long long)z * ny + y) * nx - s.lb[0];
        double e = ilo <= ihi ? exp(ex0) : 0.0, rho = ilo <= ihi ? exp(d0) : 0.0;
        for (int i = s.lb[0]; i < min(ilo, s.ub[0] + 1); i++) out[i] = 0.0;
        int i = ilo;
        // four points per step: independent Horner chains
        for (; i + 3 <= ihi; i += 4) {
            double u[4], p[4], ee[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                u[q] = gen ? (double)(i + q) - s.roff[0] : (i + q) * tg.dr[0] - s.roff[0];
                p[q] = cx[LM - 1];
                ee[q] = e;
                e *= rho;
                rho *= kap;
            }
#pragma unroll
            for (int l = LM - 2; l >= 0; l--)
#pragma unroll
                for (int q = 0; q < 4; q++) p[q] = p[q] * u[q] + cx[l];
#pragma unroll
            for (int q = 0; q < 4; q++) out[i + q] = p[q] * ee[q];
        }
        for (; i <= ihi; i++) {
            const double u = gen ? (double)i - s.roff[0] : i * tg.dr[0] - s.roff[0];
            double p = cx[LM - 1];
#pragma unroll
            for (int l = LM - 2; l >= 0; l--) p = p * u + cx[l];
            out[i] = p * e;
            e *= rho;
            rho *= kap;
        }
        for (i = max(i, s.lb[0]); i <= s.ub[0]; i++) out[i] = 0.0;
    }
}

// one workgroup per tile, one thread per tile row (y, z) holding its TX
// points in registers: the boxes listed for the tile (recs: base, lb, ub of
// the box; ent_off its offset in the scratch buffer) are added in list order,
// each thread reading the part of the box row that lies on its row; the
// records pass through LDS in chunks
constexpr int GQ = 64;

__global__ __launch_bounds__(NROW) void k_gather_small(TileGrid tg, const EntryRec *__restrict__ recs,
                                                       const int *__restrict__ tile_start,
                                                       const double *__restrict__ sval,
                                                       double *__restrict__ grid) {
    __shared__ EntryRec rec[GQ];
    const int tile = blockIdx.x;
    const int beg = tile_start[tile], end = tile_start[tile + 1];
    if (beg == end) return;
    const int tix = tile % tg.ntx, tiy = (tile / tg.ntx) % tg.nty, tiz = tile / (tg.ntx * tg.nty);
    const int t = threadIdx.x;
    const int gx0 = tix * TX, gy = tiy * TY + (t % TY), gz = tiz * TZ + t / TY;
    const bool live = gy < tg.ny && gz < tg.nz;
    // grid index g -> box offset (g - base) mod period; g and base lie in
    // [0, n), so the difference is in (-period, period)
    auto boff = [](int g, int base, int period) {
        const int d = g - base;
        return d < 0 ? d + period : d;
    };
    double acc[TX];
#pragma unroll
    for (int j = 0; j < TX; j++) acc[j] = 0.0;
    const int *rec_words = reinterpret_cast<const int *>(recs);
    for (int c0 = beg; c0 < end; c0 += GQ) {
        const int ng = min(GQ, end - c0);
        __syncthreads();
        for (int w = t; w < ng * RECW; w += NROW)
            reinterpret_cast<int *>(rec)[w] = rec_words[(long long)c0 * RECW + w];
        __syncthreads();
        if (!live) continue;
        for (int g = 0; g < ng; g++) {
            const EntryRec &R = rec[g];
            const int nx = R.ub[0] - R.lb[0] + 1, ny = R.ub[1] - R.lb[1] + 1, nz = R.ub[2] - R.lb[2] + 1;
            const int kz = boff(gz, R.base[2], tg.period[2]);
            const int jy = boff(gy, R.base[1], tg.period[1]);
            if (kz >= nz || jy >= ny) continue;
            const double *row = sval + R.ent_off + ((long long)kz * ny + jy) * nx;
            int ix = boff(gx0, R.base[0], tg.period[0]);
#pragma unroll
            for (int j = 0; j < TX; j++, ix++) {
                if (ix == tg.period[0]) ix = 0;
                if (ix < nx && gx0 + j < tg.nx) acc[j] += row[ix];
            }
        }
    }
    if (!live) return;
    double *out = grid + ((long long)gz * tg.ny + gy) * tg.nx + gx0;
#pragma unroll
    for (int j = 0; j < TX; j++)
        if (gx0 + j < tg.nx) out[j] += acc[j];
}

// ---------------------------------------------------------------------------
// plain kernels: one launch per sub-box. Workgroup = 8 rows x 32 lanes over
// one z-plane of the sub-box; blockIdx.x -> z offset, blockIdx.y -> row chunk.

constexpr int PR = 8, PL = 32;

__global__ __launch_bounds__(PR *PL) void k_collocate_plain(TileGrid tg, const SubDev *__restrict__ subs,
                                                            int si, const double *__restrict__ coef,
                                                            const int *__restrict__ prof,
                                                            double *__restrict__ grid) {
    __shared__ double cxy[ncoef2(LPMAX)];
    __shared__ double zt[LPMAX + 1];
    const SubDev s = subs[si];
    const bool gen = tg.general != 0;
    const int lp = s.lp;
    const int k = s.lb[2] + blockIdx.x;
    const int t = threadIdx.x, r = t / PL, lane = t % PL;
    const double rz = gen ? (double)k - s.roff[2] : k * tg.dr[2] - s.roff[2];
    if (t == 0) {
        double v = gen ? 1.0 : exp(-s.zetp * rz * rz);
        for (int l = 0; l <= lp; l++) {
            zt[l] = v;
            v *= rz;
        }
    }
    __syncthreads();
    for (int c2 = t; c2 < ncoef2(lp); c2 += blockDim.x) {
        int lyp, lxp;
        decode_c2(lp, c2, lyp, lxp);
        double a = 0.0;
        for (int lzp = 0; lzp <= lp - lyp - lxp; lzp++) a += coef[s.coef_off + cidx(lp, lzp, lyp, lxp)] * zt[lzp];
        cxy[c2] = a;
    }
    __syncthreads();
    const int j = s.lb[1] + blockIdx.y * PR + r;
    if (j > s.ub[1]) return;
    int ilo = s.lb[0], ihi = s.ub[0], T = BIG;
    double qa = 0.0, qb = 0.0, qc = 0.0;
    const double ry = gen ? (double)j - s.roff[1] : j * tg.dr[1] - s.roff[1];
    if (gen) {
        int lo, hi;
        general_row_dev(tg, s.roff[0], ry, rz, s.R2, s.zetp, lo, hi, qb, qc);
        ilo = max(ilo, lo);
        ihi = min(ihi, hi);
        for (int x = 0; x < 3; x++) qa = __dadd_rn(qa, __dmul_rn(tg.dh[0][x], tg.dh[0][x]));
        qa = qa * -s.zetp;
    } else {
        T = row_threshold(prof, s.prof_off, s.prof_stride, fold_off(k), fold_off(j));
        if (T < 0) return;
    }
    if (ilo > ihi) return;
    double cx[LPMAX + 1];
    for (int l = 0; l <= lp; l++) cx[l] = 0.0;
    double yp = gen ? 1.0 : exp(-s.zetp * ry * ry);
    for (int lyp = 0; lyp <= lp; lyp++) {
        for (int lxp = 0; lxp <= lp - lyp; lxp++) cx[lxp] += cxy[c2idx(lp, lyp, lxp)] * yp;
        yp *= ry;
    }
    const int gz = imod(s.base[2] + (k - s.lb[2]), s.period[2]);
    const int gy = imod(s.base[1] + (j - s.lb[1]), s.period[1]);
    double *row = grid + ((long long)gz * tg.ny + gy) * tg.nx;
    for (int i = ilo + lane; i <= ihi; i += PL) {
        if (!gen && fold_off(i) > T) continue;
        const double rx = gen ? (double)i - s.roff[0] : i * tg.dr[0] - s.roff[0];
        double xp = gen ? 1.0 : exp(-s.zetp * rx * rx), v = 0.0;
        for (int l = 0; l <= lp; l++) {
            v += cx[l] * xp;
            xp *= rx;
        }
        if (gen) v *= exp((qa * i + qb) * i + qc);
        row[imod(s.base[0] + (i - s.lb[0]), s.period[0])] += v;
    }
}

// moments of one sub-box, atomically accumulated into Mbuf[pair]
__global__ __launch_bounds__(PR *PL) void k_moments_plain(TileGrid tg, const SubDev *__restrict__ subs,
                                                          int si, int lp_add,
                                                          const int *__restrict__ prof,
                                                          const double *__restrict__ grid,
                                                          int mstride, double *__restrict__ Mbuf) {
    __shared__ double tyl[PR][ncoef2(LPMAX)];
    const SubDev s = subs[si];
    const bool gen = tg.general != 0;
    const int lp = s.lp + lp_add;
    const int k = s.lb[2] + blockIdx.x;
    const int t = threadIdx.x, r = t / PL, lane = t % PL;
    const int j = s.lb[1] + blockIdx.y * PR + r;
    double tx[LPMAX + 1];
    for (int l = 0; l <= lp; l++) tx[l] = 0.0;
    const double rz = gen ? (double)k - s.roff[2] : k * tg.dr[2] - s.roff[2];
    const double ry = gen ? (double)j - s.roff[1] : j * tg.dr[1] - s.roff[1];
    bool rowok = j <= s.ub[1];
    int ilo = s.lb[0], ihi = s.ub[0], T = BIG;
    double qa = 0.0, qb = 0.0, qc = 0.0;
    if (rowok) {
        if (gen) {
            int lo, hi;
            general_row_dev(tg, s.roff[0], ry, rz, s.R2, s.zetp, lo, hi, qb, qc);
            ilo = max(ilo, lo);
            ihi = min(ihi, hi);
            for (int x = 0; x < 3; x++) qa = __dadd_rn(qa, __dmul_rn(tg.dh[0][x], tg.dh[0][x]));
            qa = qa * -s.zetp;
        } else {
            T = row_threshold(prof, s.prof_off, s.prof_stride, fold_off(k), fold_off(j));
            rowok = T >= 0;
        }
    }
    if (rowok) {
        const int gz = imod(s.base[2] + (k - s.lb[2]), s.period[2]);
        const int gy = imod(s.base[1] + (j - s.lb[1]), s.period[1]);
        const double *row = grid + ((long long)gz * tg.ny + gy) * tg.nx;
        for (int i = ilo + lane; i <= ihi; i += PL) {
            if (!gen && fold_off(i) > T) continue;
            const double rx = gen ? (double)i - s.roff[0] : i * tg.dr[0] - s.roff[0];
            double xp = (gen ? exp((qa * i + qb) * i + qc) : exp(-s.zetp * rx * rx)) *
                        row[imod(s.base[0] + (i - s.lb[0]), s.period[0])];
            for (int l = 0; l <= lp; l++) {
                tx[l] += xp;
                xp *= rx;
            }
        }
    }
    // reduce tx over the 32 lanes of the row
    for (int l = 0; l <= lp; l++) {
        double v = tx[l];
        for (int o = PL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, PL);
        tx[l] = v;
    }
    if (lane == 0) {
        double yp = rowok ? (gen ? 1.0 : exp(-s.zetp * ry * ry)) : 0.0;
        for (int lyp = 0; lyp <= lp; lyp++) {
            for (int lxp = 0; lxp <= lp - lyp; lxp++) tyl[r][c2idx(lp, lyp, lxp)] = tx[lxp] * yp;
            yp *= ry;
        }
    }
    __syncthreads();
    const double ez = gen ? 1.0 : exp(-s.zetp * rz * rz);
    for (int c = t; c < ncoset(lp); c += blockDim.x) {
        int lzp, lyp, lxp;
        decode_c3(lp, c, lzp, lyp, lxp);
        double zp = ez;
        for (int q = 0; q < lzp; q++) zp *= rz;
        const int c2 = c2idx(lp, lyp, lxp);
        double a = 0.0;
        for (int rr = 0; rr < PR; rr++) a += tyl[rr][c2];
        atomicAdd(&Mbuf[(long long)s.pair * mstride + c], a * zp);
    }
}

// ---------------------------------------------------------------------------
// libgpw tensor-path contractions (reference src/lib/collocate.cc:438-623)

// one thread per (z, y) row: xyz[a] = sum_{g,b} co[a][g][b] Pz[g][z] Py[b][y];
// V[z][y][x] = sum_a xyz[a] Px[a][x]
__global__ void k_collocate_core(int n0, int n1, int n2, int L, const double *__restrict__ co,
                                 const double *__restrict__ P, int ldp, double *__restrict__ V) {
    const long long rowid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (rowid >= (long long)n0 * n1) return;
    const int z = rowid / n1, y = rowid % n1;
    const double *Pz = P, *Py = P + (long long)L * ldp, *Px = P + 2LL * L * ldp;
    double xyz[LPMAX + 1];
    for (int a = 0; a < L; a++) {
        double s = 0.0;
        for (int g = 0; g < L; g++) {
            const double pz = Pz[(long long)g * ldp + z];
            for (int b = 0; b < L; b++) s += co[((long long)a * L + g) * L + b] * pz * Py[(long long)b * ldp + y];
        }
        xyz[a] = s;
    }
    double *row = V + rowid * n2;
    for (int x = 0; x < n2; x++) {
        double v = 0.0;
        for (int a = 0; a < L; a++) v += xyz[a] * Px[(long long)a * ldp + x];
        row[x] = v;
    }
}

// one workgroup per z plane: Velems[g][b][a] += sum_y sum_x ...
__global__ __launch_bounds__(256) void k_integrate_core(int n0, int n1, int n2, int L,
                                                        const double *__restrict__ P, int ldp,
                                                        const double *__restrict__ V,
                                                        double *__restrict__ Velems) {
    extern __shared__ double sm_ic[];  // [L][L] partial (b, a)
    const int z = blockIdx.x;
    const double *Pz = P, *Py = P + (long long)L * ldp, *Px = P + 2LL * L * ldp;
    for (int w = threadIdx.x; w < L * L; w += blockDim.x) sm_ic[w] = 0.0;
    __syncthreads();
    for (int y = threadIdx.x; y < n1; y += blockDim.x) {
        const double *row = V + ((long long)z * n1 + y) * n2;
        double t[LPMAX + 1];
        for (int a = 0; a < L; a++) t[a] = 0.0;
        for (int x = 0; x < n2; x++) {
            const double v = row[x];
            for (int a = 0; a < L; a++) t[a] += v * Px[(long long)a * ldp + x];
        }
        for (int b = 0; b < L; b++) {
            const double py = Py[(long long)b * ldp + y];
            for (int a = 0; a < L; a++) atomicAdd(&sm_ic[b * L + a], t[a] * py);
        }
    }
    __syncthreads();
    for (int w = threadIdx.x; w < L * L * L; w += blockDim.x) {
        const int g = w / (L * L), ba = w % (L * L);
        atomicAdd(&Velems[w], sm_ic[ba] * Pz[(long long)g * ldp + z]);
    }
}

// ---------------------------------------------------------------------------
// device planner of the tile work lists (see plan_device in gpw_plan.cpp)

// tiles of axis d a sub-box touches, ascending: r0..r1, then q0..q1 (empty
// when q0 > q1) -- a run that wraps around a periodic axis touches the
// tiles from 0 and those up to the last one (the host's axis_tiles)
__device__ __forceinline__ void plan_axis_ranges(const SubDev &s, int d, int n, int T, int ntile, int &r0,
                                                 int &r1, int &q0, int &q1) {
    const int len = s.ub[d] - s.lb[d] + 1;
    const int start = s.base[d];
    q0 = 1;
    q1 = 0;
    if (s.period[d] != BIG) {
        const int stop = start + len - 1;
        if (stop < n) {
            r0 = start / T;
            r1 = stop / T;
        } else {
            const int w1 = (stop - n) / T, s0 = start / T;
            if (w1 + 1 >= s0) {
                r0 = 0;
                r1 = ntile - 1;
            } else {
                r0 = 0;
                r1 = w1;
                q0 = s0;
                q1 = ntile - 1;
            }
        }
    } else {
        r0 = start / T;
        r1 = (start + len - 1) / T;
    }
}

__device__ __forceinline__ int plan_axis_count(const SubDev &s, int d, const TileGrid &tg) {
    int r0, r1, q0, q1;
    const int n = d == 0 ? tg.nx : d == 1 ? tg.ny : tg.nz;
    const int T = d == 0 ? TX : d == 1 ? TY : TZ;
    const int nt = d == 0 ? tg.ntx : d == 1 ? tg.nty : tg.ntz;
    plan_axis_ranges(s, d, n, T, nt, r0, r1, q0, q1);
    return (r1 - r0 + 1) + (q1 >= q0 ? q1 - q0 + 1 : 0);
}

__device__ __forceinline__ int plan_axis_tile(const SubDev &s, int d, const TileGrid &tg, int k) {
    int r0, r1, q0, q1;
    const int n = d == 0 ? tg.nx : d == 1 ? tg.ny : tg.nz;
    const int T = d == 0 ? TX : d == 1 ? TY : TZ;
    const int nt = d == 0 ? tg.ntx : d == 1 ? tg.nty : tg.ntz;
    plan_axis_ranges(s, d, n, T, nt, r0, r1, q0, q1);
    const int n1 = r1 - r0 + 1;
    return k < n1 ? r0 + k : q0 + (k - n1);
}

// x offsets lo..hi of row (j, k) of sub-box s inside its sphere, as the tile
// kernels decide point by point: the general quadratic bounds, or the
// orthorhombic profile threshold (fold offsets <= T <=> -T <= i <= T + 1)
__device__ __forceinline__ bool plan_row(const SubDev &s, const TileGrid &tg, const int *__restrict__ prof, int j,
                                         int k, int &lo, int &hi) {
    if (tg.general) {
        double qb, qc;
        general_row_dev(tg, s.roff[0], (double)j - s.roff[1], (double)k - s.roff[2], s.R2, s.zetp, lo, hi, qb, qc);
        return lo <= hi;
    }
    const int T = row_threshold(prof, s.prof_off, s.prof_stride, fold_off(k), fold_off(j));
    lo = -T;
    hi = T + 1;
    return T >= 0;
}

// Work items of sub-box s in the tiles of candidate row (iy, iz) (the sub-box's
// iy-th y tile and iz-th z tile): bit ix - ix0 is set when the sphere has a
// point in the tile of the ix-th x tile, ix0 <= ix < ix0 + 64 (the test of
// item_has_points). The wave's lanes take the tile rows, each row its x
// offset range against every candidate x tile; OR over the wave.
__device__ unsigned long long plan_row_mask(const SubDev &s, const TileGrid &tg, const int *__restrict__ prof,
                                            int iy, int iz, int ix0, int nxs, int lane) {
    const int ty = plan_axis_tile(s, 1, tg, iy), tz = plan_axis_tile(s, 2, tg, iz);
    const int ix1 = min(nxs, ix0 + 64);
    unsigned long long m = 0;
    for (int r = lane; r < TY * TZ; r += 64) {
        const int gz = tz * TZ + r / TY, gy = ty * TY + r % TY;
        if (gz >= tg.nz || gy >= tg.ny) continue;
        const int k = s.lb[2] + imod_hd(gz - s.base[2], s.period[2]);
        const int j = s.lb[1] + imod_hd(gy - s.base[1], s.period[1]);
        int lo, hi;
        if (k > s.ub[2] || j > s.ub[1] || !plan_row(s, tg, prof, j, k, lo, hi)) continue;
        for (int ix = ix0; ix < ix1; ix++) {
            const int tx = plan_axis_tile(s, 0, tg, ix);
            int xa[2], xb[2];
            tile_offset_runs(s, 0, tx * TX, min((tx + 1) * TX, tg.nx) - 1, xa, xb);
            if (max(lo, xa[0]) <= min(hi, xb[0]) || max(lo, xa[1]) <= min(hi, xb[1])) m |= 1ULL << (ix - ix0);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o);
    return m;
}

// One wave per sub-box: its number of (sub-box, tile) items -- the tiles
// that hold a point of its sphere -- and its sphere points, rows and
// z-planes added to the per-lp counters work[3 lp + 0..2] (the algorithmic
// work of the stats; integer sums, so order-independent).
__global__ __launch_bounds__(256) void k_plan_count(const SubDev *__restrict__ subs, int nsub,
                                                    const int *__restrict__ prof, TileGrid tg,
                                                    int *__restrict__ counts,
                                                    unsigned long long *__restrict__ work) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long si = (long long)blockIdx.x * 4 + wave;
    if (si >= nsub) return;
    const SubDev s = subs[si];
    {
        const int nxs = plan_axis_count(s, 0, tg), nys = plan_axis_count(s, 1, tg), nzs = plan_axis_count(s, 2, tg);
        int cnt = 0;
        for (int iz = 0; iz < nzs; iz++)
            for (int iy = 0; iy < nys; iy++)
                for (int ix0 = 0; ix0 < nxs; ix0 += 64) cnt += __popcll(plan_row_mask(s, tg, prof, iy, iz, ix0, nxs, lane));
        if (lane == 0) counts[si] = cnt;
    }
    unsigned long long pts = 0, rows = 0, planes = 0;
    for (int k = s.lb[2]; k <= s.ub[2]; k++) {
        int rows_k = 0;
        for (int j = s.lb[1] + lane; j <= s.ub[1]; j += 64) {
            int lo, hi;
            if (tg.general) {
                double qb, qc;
                general_row_dev(tg, s.roff[0], (double)j - s.roff[1], (double)k - s.roff[2], s.R2, s.zetp,
                                lo, hi, qb, qc);
                lo = max(lo, s.lb[0]);
                hi = min(hi, s.ub[0]);
            } else {
                const int T = s.prof_off < 0 ? BIG : prof[s.prof_off + fold_off(k) * s.prof_stride + fold_off(j)];
                if (T < 0) continue;
                lo = max(s.lb[0], -T);
                hi = min(s.ub[0], T + 1);
            }
            if (hi >= lo) {
                pts += (unsigned long long)(hi - lo + 1);
                rows_k++;
            }
        }
        rows += (unsigned long long)rows_k;
        if (__any(rows_k > 0)) planes++;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pts += __shfl_xor(pts, o);
        rows += __shfl_xor(rows, o);
    }
    if (lane == 0) {
        atomicAdd(&work[3 * s.lp + 0], pts);
        atomicAdd(&work[3 * s.lp + 1], rows);
        atomicAdd(&work[3 * s.lp + 2], planes);
    }
}

// One wave per sub-box: its items e = ent_start[sub] + k (k-th non-empty
// candidate in candidate order; sub-boxes are pair-major), each with its
// sub-box, partial-moment offsets and sort key per tile-list set, class *
// ntiles + tile, or NCLS * ntiles when the set leaves it out (pairs of the
// small-sphere kernels; set 3 holds only the small-sphere collocate pairs);
// hist counts the items per set and key. small_mask: [3][npairs] (integrate
// without / with forces, collocate).
__global__ __launch_bounds__(256) void k_plan_entries(const SubDev *__restrict__ subs, int nsub,
                               const long long *__restrict__ ent_start, const long long *__restrict__ base0,
                               const long long *__restrict__ base2, const int *__restrict__ prof, TileGrid tg,
                               int ntiles, const unsigned char *__restrict__ small_mask, int npairs, int ne,
                               int *__restrict__ ent_sub, long long *__restrict__ off0,
                               long long *__restrict__ off2, unsigned *__restrict__ keys,
                               int *__restrict__ vals, int *__restrict__ hist) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int si = blockIdx.x * 4 + wave;
    if (si >= nsub) return;
    const SubDev sd = subs[si];
    const int nxs = plan_axis_count(sd, 0, tg), nys = plan_axis_count(sd, 1, tg), nzs = plan_axis_count(sd, 2, tg);
    const long long e0 = ent_start[si];
    const unsigned sent = (unsigned)(NCLS * ntiles);
    const bool skip1 = small_mask[npairs + sd.pair] != 0, skip2 = small_mask[sd.pair] != 0;
    const bool csmall = small_mask[2 * (long long)npairs + sd.pair] != 0;
    int done = 0;
    for (int iz = 0; iz < nzs; iz++)
    for (int iy = 0; iy < nys; iy++)
    for (int ix0 = 0; ix0 < nxs; ix0 += 64) {
        const unsigned long long m = plan_row_mask(sd, tg, prof, iy, iz, ix0, nxs, lane);
        // lane b takes candidate x tile ix0 + b
        const bool pass = (m >> lane) & 1ULL;
        if (pass) {
            const int tile = (plan_axis_tile(sd, 2, tg, iz) * tg.nty + plan_axis_tile(sd, 1, tg, iy)) * tg.ntx +
                             plan_axis_tile(sd, 0, tg, ix0 + lane);
            const int k = done + __popcll(m & ((1ULL << lane) - 1ULL));
            const long long e = e0 + k;
            ent_sub[e] = si;
            off0[e] = base0[si] + (long long)k * ncoset(sd.lp);
            off2[e] = base2[si] + (long long)k * ncoset(sd.lp + 2);
            vals[e] = (int)e;
            const unsigned kc = (unsigned)(class_of(sd.lp) * ntiles + tile);
            const unsigned k0 = csmall ? sent : kc;
            const unsigned k1 = skip1 ? sent : (unsigned)(class_of(sd.lp + 2) * ntiles + tile);
            const unsigned k2 = skip2 ? sent : kc;
            const unsigned k3 = csmall ? (unsigned)tile : sent;
            const unsigned kk[NSET] = {k0, k1, k2, k3};
#pragma unroll
            for (int a = 0; a < NSET; a++) {
                keys[(long long)a * ne + e] = kk[a];
                if (kk[a] != sent) atomicAdd(&hist[(long long)a * NCLS * ntiles + kk[a]], 1);
            }
        }
        done += __popcll(m);
    }
}

// records of one set in sorted order: position p holds item sorted_e[p]
// (sub_off: ent_off is the sub-box's offset sub_off[sub] instead of off[item])
__global__ void k_plan_records(const SubDev *__restrict__ subs, const int *__restrict__ ent_sub,
                               const long long *__restrict__ off, const long long *__restrict__ sub_off,
                               const int *__restrict__ sorted_e, int n, EntryRec *__restrict__ recs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int e = sorted_e[p];
    const SubDev sd = subs[ent_sub[e]];
    EntryRec r;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        r.roff[d] = sd.roff[d];
        r.base[d] = sd.base[d];
        r.lb[d] = sd.lb[d];
        r.ub[d] = sd.ub[d];
    }
    r.zetp = sd.zetp;
    r.R2 = sd.R2;
    r.ent_off = sub_off ? sub_off[ent_sub[e]] : off[e];
    r.lp = sd.lp;
    r.coef_off = sd.coef_off;
    r.prof_off = sd.prof_off;
    r.prof_stride = sd.prof_stride;
    r.pad = 0;
    recs[p] = r;
}

// ---------------------------------------------------------------------------
// libgpw folded-table path (reference collocate.cc:32-57, 276-331, 626-745)

__device__ __forceinline__ int pmod_dev(long long a, int m) {
    const long long r = a % m;
    return (int)(r < 0 ? r + m : r);
}

// Periodic fold of the 1D tables: block (a, d) folds power a of direction d
// (a == L: the support mask of direction d). Entry i of an unfolded table is
// grid coordinate p0 + i, p0 = center - (length - 1) / 2; folded[g] gathers
// the entries i = g - p0 (mod period) in increasing i, so the sums are
// deterministic.
__global__ void k_fold_tables(int L, int3 length, int3 center, int3 period,
                              const double *__restrict__ unfolded, int lu, int ldu,
                              double *__restrict__ folded, int ldp, int *__restrict__ mask, int ldm) {
    const int a = blockIdx.x, d = blockIdx.y;
    const int len = d == 0 ? length.x : d == 1 ? length.y : length.z;
    const int per = d == 0 ? period.x : d == 1 ? period.y : period.z;
    const long long p0 = (long long)(d == 0 ? center.x : d == 1 ? center.y : center.z) - (len - 1) / 2;
    if (a == L) {
        if (!mask) return;
        for (int g = threadIdx.x; g < ldm; g += blockDim.x) {
            int m = 0;
            if (g < per) m = len >= per ? 1 : pmod_dev((long long)g - p0, per) < len;
            mask[(long long)d * ldm + g] = m;
        }
        return;
    }
    const double *u = unfolded + ((long long)d * lu + a) * ldu;
    double *f = folded + ((long long)d * L + a) * ldp;
    for (int g = threadIdx.x; g < ldp; g += blockDim.x) {
        double acc = 0.0;
        if (g < per)
            for (int i = pmod_dev((long long)g - p0, per); i < len; i += per) acc += u[i];
        f[g] = acc;
    }
}

// collocate<T>: one thread per (z, y) row of the local window whose z and y
// masks are set; xyz[a] = sum_{g,b} co[a][g][b] Pz[g][c0+z] Py[b][c1+y], then
// density[z][y][x] += sum_a xyz[a] Px[a][c2+x] on the x points whose mask is set
__global__ void k_collocate_folded(int L, const double *__restrict__ co, const double *__restrict__ P,
                                   int ldp, const int *__restrict__ mask, int ldm, int3 gs, int3 lc,
                                   double *__restrict__ density) {
    const long long rowid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (rowid >= (long long)gs.x * gs.y) return;
    const int z = rowid / gs.y, y = rowid % gs.y;
    if (!mask[lc.x + z] || !mask[ldm + lc.y + y]) return;
    const double *Pz = P, *Py = P + (long long)L * ldp, *Px = P + 2LL * L * ldp;
    const int *mx = mask + 2LL * ldm + lc.z;
    double xyz[LPMAX + 1];
    for (int a = 0; a < L; a++) {
        double s = 0.0;
        for (int g = 0; g < L; g++) {
            const double pz = Pz[(long long)g * ldp + lc.x + z];
            for (int b = 0; b < L; b++) s += co[((long long)a * L + g) * L + b] * pz * Py[(long long)b * ldp + lc.y + y];
        }
        xyz[a] = s;
    }
    double *row = density + rowid * gs.z;
    for (int x = 0; x < gs.z; x++) {
        if (!mx[x]) continue;
        double v = 0.0;
        for (int a = 0; a < L; a++) v += xyz[a] * Px[(long long)a * ldp + lc.z + x];
        row[x] += v;
    }
}

// integrate<T>: one workgroup per z plane of the window (mask set):
// co[g][b][a] += Pz[g][c0+z] sum_y Py[b][c1+y] sum_x V[z][y][x] Px[a][c2+x]
// over the set masks (co cleared beforehand)
__global__ __launch_bounds__(256) void k_integrate_folded(int L, const double *__restrict__ P, int ldp,
                                                          const int *__restrict__ mask, int ldm, int3 gs,
                                                          int3 lc, const double *__restrict__ V,
                                                          double *__restrict__ co) {
    extern __shared__ double sm_if[];  // [L][L] partial (b, a)
    const int z = blockIdx.x;
    if (!mask[lc.x + z]) return;
    const double *Pz = P, *Py = P + (long long)L * ldp, *Px = P + 2LL * L * ldp;
    const int *my = mask + ldm + lc.y, *mx = mask + 2LL * ldm + lc.z;
    for (int w = threadIdx.x; w < L * L; w += blockDim.x) sm_if[w] = 0.0;
    __syncthreads();
    for (int y = threadIdx.x; y < gs.y; y += blockDim.x) {
        if (!my[y]) continue;
        const double *row = V + ((long long)z * gs.y + y) * gs.z;
        double t[LPMAX + 1];
        for (int a = 0; a < L; a++) t[a] = 0.0;
        for (int x = 0; x < gs.z; x++) {
            if (!mx[x]) continue;
            const double v = row[x];
            for (int a = 0; a < L; a++) t[a] += v * Px[(long long)a * ldp + lc.z + x];
        }
        for (int b = 0; b < L; b++) {
            const double py = Py[(long long)b * ldp + lc.y + y];
            for (int a = 0; a < L; a++) atomicAdd(&sm_if[b * L + a], t[a] * py);
        }
    }
    __syncthreads();
    for (int w = threadIdx.x; w < L * L * L; w += blockDim.x) {
        const int g = w / (L * L), ba = w % (L * L);
        atomicAdd(&co[w], sm_if[ba] * Pz[(long long)g * ldp + lc.x + z]);
    }
}

// ---------------------------------------------------------------------------
// launchers (host side)

hipError_t launch_collocate_tiles(int cls, const TileGrid &tg, const EntryRec *recs,
                                  const int *tile_start, const double *coef, int ntiles, int splits,
                                  double *out, int out_mode, hipStream_t st) {
    return tg.general ? launch_collocate_tiles_t<true>(cls, tg, recs, tile_start, coef, ntiles, splits,
                                                       out, out_mode, st)
                      : launch_collocate_tiles_t<false>(cls, tg, recs, tile_start, coef, ntiles, splits,
                                                        out, out_mode, st);
}

hipError_t launch_integrate_tiles(int cls, const TileGrid &tg, const EntryRec *recs,
                                  const int *tile_start, const double *grid, int lp_add,
                                  int ntiles, int splits, double *partial, hipStream_t st) {
    return tg.general ? launch_integrate_tiles_t<true>(cls, tg, recs, tile_start, grid, lp_add, ntiles,
                                                       splits, partial, st)
                      : launch_integrate_tiles_t<false>(cls, tg, recs, tile_start, grid, lp_add, ntiles,
                                                        splits, partial, st);
}

hipError_t launch_reduce_partials(const double *partial, int splits, long long ng, double *grid,
                                  hipStream_t st) {
    const int blocks = (int)std::min<long long>((ng + 255) / 256, 8192);
    hipLaunchKernelGGL(k_reduce_partials, dim3(blocks), dim3(256), 0, st, partial, splits, ng, grid);
    return hipGetLastError();
}

// workgroup size of k_coef / k_hab: one wave for small blocks, four waves
// (more latency hiding per LDS allocation) from lp = 4 on
static int transform_threads(int lp) { return lp >= 4 ? TR_NT : 64; }

// k_coef: per-pair blocks (eff, Ty, cxyz) and the LDS with / without them
size_t coef_ws_doubles(int lam, int lbm) {
    const int lp = lam + lbm;
    return (size_t)ncoset(lam) * ncoset(lbm) + (size_t)(lam + 1) * (lbm + 1) * ncoef2(lp) +
           (size_t)ncoset(lp);
}

size_t coef_smem_bytes(int lam, int lbm, bool lds_blocks) {
    size_t n = (size_t)alpha_doubles(lam, lbm);
    if (lds_blocks) n += coef_ws_doubles(lam, lbm);
    return sizeof(double) * n;
}

bool coef_needs_workspace(int lam, int lbm) { return coef_smem_bytes(lam, lbm, true) > COEF_LDS_MAX; }

int block_shell_entries(int la_min, int la_max, int lb_min, int lb_max) {
    return (ncoset(la_max) - ncoset(la_min - 1)) * (ncoset(lb_max) - ncoset(lb_min - 1));
}

// pairs order[pair0 .. pair0 + npairs) (or pair0 .. when order is NULL), whose
// la_max / lb_max are at most la_max / lb_max (LDS sizing)
hipError_t launch_coef(const PairDev *pairs, const int *order, int pair0, int npairs, int tau,
                       int lam_max, int lbm_max, int nab, const double *pab, double *coef,
                       const LatticeTab &lt, double *ws,
                       long long ws_stride, int ws_blocks, hipStream_t st) {
    if (npairs <= 0) return hipSuccess;
    const size_t xs = lt.stepped ? sizeof(StepTabs) : 0;  // stepped transform's tables
    if (!tau && sizeof(double) * coefw_doubles(lam_max, lbm_max, nab) + xs <= WAVE_LDS_MAX) {
        // one pair per wave
        const size_t smem = sizeof(double) * coefw_doubles(lam_max, lbm_max, nab) + xs;
        switch (class_of(lam_max + lbm_max)) {
        case 0: hipLaunchKernelGGL(k_coef_wave<4>, dim3(npairs), dim3(TW), smem, st, pairs, order, pair0, npairs, pab, coef, lt); break;
        case 1: hipLaunchKernelGGL(k_coef_wave<8>, dim3(npairs), dim3(TW), smem, st, pairs, order, pair0, npairs, pab, coef, lt); break;
        case 2: hipLaunchKernelGGL(k_coef_wave<12>, dim3(npairs), dim3(TW), smem, st, pairs, order, pair0, npairs, pab, coef, lt); break;
        default: hipLaunchKernelGGL(k_coef_wave<17>, dim3(npairs), dim3(TW), smem, st, pairs, order, pair0, npairs, pab, coef, lt); break;
        }
        return hipGetLastError();
    }
    const int lam = lam_max + tau, lbm = lbm_max + tau;
    const bool big = coef_needs_workspace(lam, lbm);
    if (big && (!ws || ws_stride < (long long)coef_ws_doubles(lam, lbm) || ws_blocks < 1))
        return hipErrorInvalidValue;
    const size_t smem = coef_smem_bytes(lam, lbm, !big) + xs;
    const int nblk = big ? std::min(npairs, ws_blocks) : npairs;
    hipLaunchKernelGGL(k_coef, dim3(nblk), dim3(transform_threads(lam + lbm)), smem, st, pairs, order, pair0, npairs, tau, pab,
                       coef, lt, big ? ws : nullptr, big ? ws_stride : 0LL);
    return hipGetLastError();
}

// k_hab per-pair blocks (heff, eff, packed U) in doubles, for an extended
// block (lamx, lbmx) of a local block (lam, lbm); nhx >= 0: heff holds only
// the cross of nhx entries (forces without tau), else the dense block
size_t hab_ws_doubles(int lamx, int lbmx, int lam, int lbm, int tau, int nhx) {
    size_t u = 0;
    for (int az = 0; az <= lamx; az++)
        for (int bz = 0; bz <= lbmx; bz++) u += (size_t)ncoef2((lamx - az) + (lbmx - bz));
    const size_t heff = nhx >= 0 ? (size_t)nhx : (size_t)ncoset(lamx) * ncoset(lbmx);
    const size_t eff = tau ? (size_t)ncoset(lam) * ncoset(lbm) : 0;
    return heff + eff + u;
}

// LDS of k_hab: alpha, M, Mijk, the force reduction and (LDS mode) the blocks
size_t hab_smem_bytes(int lamx, int lbmx, int lam, int lbm, int tau, int nhx, bool lds_blocks) {
    const int lpx = lamx + lbmx;
    size_t n = (size_t)alpha_doubles(lamx, lbmx) + 2 * (size_t)ncoset(lpx) + 6 * (TR_NT / 64);
    if (lds_blocks) n += hab_ws_doubles(lamx, lbmx, lam, lbm, tau, nhx);
    return sizeof(double) * n;
}

bool hab_needs_workspace(int lamx, int lbmx, int lam, int lbm, int tau, int nhx) {
    return hab_smem_bytes(lamx, lbmx, lam, lbm, tau, nhx, true) > HAB_LDS_MAX;
}

// workspace doubles per workgroup k_hab needs for blocks up to (la_max,
// lb_max) with or without forces (nhx: the most cross entries, see
// hab_cross_entries); 0 when every case fits in LDS
size_t hab_ws_need(int la_max, int lb_max, int tau, int nhx) {
    const int lam = la_max + tau, lbm = lb_max + tau;
    size_t need = 0;
    for (int fx = 0; fx <= 1; fx++) {
        const int h = (fx && !tau) ? nhx : -1;
        if (hab_needs_workspace(lam + fx, lbm + fx, lam, lbm, tau, h))
            need = std::max(need, hab_ws_doubles(lam + fx, lbm + fx, lam, lbm, tau, h));
    }
    return need;
}

// nhx: the most heff cross entries (hab_cross_entries) among the launched
// pairs, used with forces and without tau
hipError_t launch_hab(const PairDev *pairs, const int *order, const int *segs, int seg0, int nseg,
                      int tau, int with_forces, int la_max, int lb_max, int nhx, const double *Mbuf,
                      int mstride, const double *pab, double *hab, double *forces,
                      const LatticeTab &lt, double *ws,
                      long long ws_stride, int ws_blocks, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    const int npairs = nseg;
    const int fx = with_forces ? 1 : 0;
    const int lamx = la_max + tau + fx, lbmx = lb_max + tau + fx;
    const int h = (with_forces && !tau) ? nhx : -1;
    const int lam = la_max + tau, lbm = lb_max + tau;
    const bool big = hab_needs_workspace(lamx, lbmx, lam, lbm, tau, h);
    if (big && (!ws || ws_stride < (long long)hab_ws_doubles(lamx, lbmx, lam, lbm, tau, h) || ws_blocks < 1))
        return hipErrorInvalidValue;
    const size_t smem = hab_smem_bytes(lamx, lbmx, lam, lbm, tau, h, !big) +
                        (lt.stepped ? sizeof(StepTabs) : 0);  // stepped transform's tables
    const int nblk = big ? std::min(npairs, ws_blocks) : npairs;
    auto kern = lt.stepped ? k_hab<2> : lt.blk ? k_hab<1> : k_hab<0>;
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(transform_threads(lamx + lbmx)), smem, st, pairs, order, segs, seg0, nseg, tau,
                       with_forces, Mbuf, mstride, pab, hab, forces, lt,
                       big ? ws : nullptr, big ? ws_stride : 0LL);
    return hipGetLastError();
}

hipError_t launch_sum_entries(const PairDev *pairs, int npairs, int tau, int lp_add,
                              const int *pe_start, const long long *ent_off,
                              const double *partial, int mstride, const unsigned char *skip,
                              double *Mbuf, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_entries, dim3(npairs), dim3(SUM_NT), 0, st, pairs, tau, lp_add, pe_start,
                       ent_off, partial, mstride, skip, Mbuf);
    return hipGetLastError();
}

// class c: L = lp + lp_add + 1 in (4c, 4c + 4], the last class (c = 3) up to 17
hipError_t launch_moments_small(int cls, const TileGrid &tg, const SubDev *subs, const int *list,
                                int n, const int *prof, const double *grid, int lp_add, int mstride,
                                double *Mbuf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    switch (cls) {
    case 0: hipLaunchKernelGGL(k_moments_small<4>, dim3(n), dim3(NROW), 0, st, tg, subs, list, prof, grid, lp_add, mstride, Mbuf); break;
    case 1: hipLaunchKernelGGL(k_moments_small<8>, dim3(n), dim3(NROW), 0, st, tg, subs, list, prof, grid, lp_add, mstride, Mbuf); break;
    case 2: hipLaunchKernelGGL(k_moments_small<12>, dim3(n), dim3(NROW), 0, st, tg, subs, list, prof, grid, lp_add, mstride, Mbuf); break;
    case 3: hipLaunchKernelGGL(k_moments_small<17>, dim3(n), dim3(NROW), 0, st, tg, subs, list, prof, grid, lp_add, mstride, Mbuf); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill(double *p, long long n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_collocate_plain(const TileGrid &tg, const SubDev *subs, int si, int nz_box,
                                  int ny_box, const double *coef, const int *prof, double *grid,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_collocate_plain, dim3(nz_box, (ny_box + PR - 1) / PR), dim3(PR * PL), 0, st,
                       tg, subs, si, coef, prof, grid);
    return hipGetLastError();
}

hipError_t launch_moments_plain(const TileGrid &tg, const SubDev *subs, int si, int nz_box,
                                int ny_box, int lp_add, const int *prof, const double *grid,
                                int mstride, double *Mbuf, hipStream_t st) {
    hipLaunchKernelGGL(k_moments_plain, dim3(nz_box, (ny_box + PR - 1) / PR), dim3(PR * PL), 0, st,
                       tg, subs, si, lp_add, prof, grid, mstride, Mbuf);
    return hipGetLastError();
}

hipError_t launch_collocate_core(const int *length, int L, const double *co, const double *P,
                                 int ldp, double *V, hipStream_t st) {
    const long long rows = (long long)length[0] * length[1];
    const int blocks = (int)((rows + 127) / 128);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_collocate_core, dim3(blocks), dim3(128), 0, st, length[0], length[1],
                       length[2], L, co, P, ldp, V);
    return hipGetLastError();
}

hipError_t launch_integrate_core(const int *length, int L, const double *P, int ldp,
                                 const double *V, double *Velems, hipStream_t st) {
    hipError_t e = launch_fill(Velems, (long long)L * L * L, 0.0, st);
    if (e != hipSuccess || length[0] == 0) return e;
    hipLaunchKernelGGL(k_integrate_core, dim3(length[0]), dim3(256), sizeof(double) * L * L, st,
                       length[0], length[1], length[2], L, P, ldp, V, Velems);
    return hipGetLastError();
}

hipError_t launch_plan_count(const SubDev *subs, int nsub, const int *prof, const TileGrid &tg, int *counts,
                             unsigned long long *work, hipStream_t st) {
    if (nsub <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_count, dim3((nsub + 3) / 4), dim3(256), 0, st, subs, nsub, prof, tg, counts, work);
    return hipGetLastError();
}

hipError_t launch_plan_entries(const SubDev *subs, int nsub, const long long *ent_start, const long long *base0,
                               const long long *base2, const int *prof, const TileGrid &tg, int ntiles,
                               const unsigned char *small_mask, int npairs, int ne, int *ent_sub,
                               long long *off0, long long *off2, unsigned *keys, int *vals, int *hist,
                               hipStream_t st) {
    if (ne <= 0 || nsub <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_entries, dim3((nsub + 3) / 4), dim3(256), 0, st, subs, nsub, ent_start, base0,
                       base2, prof, tg, ntiles, small_mask, npairs, ne, ent_sub, off0, off2, keys, vals, hist);
    return hipGetLastError();
}

hipError_t launch_plan_records(const SubDev *subs, const int *ent_sub, const long long *off,
                               const long long *sub_off, const int *sorted_e, int n, EntryRec *recs,
                               hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plan_records, dim3((n + 255) / 256), dim3(256), 0, st, subs, ent_sub, off, sub_off,
                       sorted_e, n, recs);
    return hipGetLastError();
}

// class c: L = lp + 1 in (4c, 4c + 4], the last class (c = 3) up to 17
hipError_t launch_collocate_small(int cls, const TileGrid &tg, const SubDev *subs, const int *list, int n,
                                  const double *coef, const int *prof, const long long *box_off, double *sval,
                                  hipStream_t st) {
    if (n <= 0) return hipSuccess;
    switch (cls) {
    case 0: hipLaunchKernelGGL(k_collocate_small<4>, dim3(n), dim3(NROW), 0, st, tg, subs, list, coef, prof, box_off, sval); break;
    case 1: hipLaunchKernelGGL(k_collocate_small<8>, dim3(n), dim3(NROW), 0, st, tg, subs, list, coef, prof, box_off, sval); break;
    case 2: hipLaunchKernelGGL(k_collocate_small<12>, dim3(n), dim3(NROW), 0, st, tg, subs, list, coef, prof, box_off, sval); break;
    case 3: hipLaunchKernelGGL(k_collocate_small<17>, dim3(n), dim3(NROW), 0, st, tg, subs, list, coef, prof, box_off, sval); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gather_small(const TileGrid &tg, const EntryRec *recs, const int *tile_start, int ntiles,
                               const double *sval, double *grid, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_small, dim3(ntiles), dim3(NROW), 0, st, tg, recs, tile_start, sval, grid);
    return hipGetLastError();
}

// stable sort of (key, value) pairs on the low end_bit bits of the keys
// (temp NULL: *temp_bytes returns the scratch size)
hipError_t sort_pairs_u32(void *temp, size_t *temp_bytes, const unsigned *keys_in, unsigned *keys_out,
                          const int *vals_in, int *vals_out, int n, int end_bit, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0,
                                              end_bit, st);
}

hipError_t launch_fold_tables(int L, const int *length, const int *center, const int *period,
                              const double *unfolded, int lu, int ldu, double *folded, int ldp,
                              int *mask, int ldm, hipStream_t st) {
    hipLaunchKernelGGL(k_fold_tables, dim3(L + 1, 3), dim3(256), 0, st, L,
                       make_int3(length[0], length[1], length[2]),
                       make_int3(center[0], center[1], center[2]),
                       make_int3(period[0], period[1], period[2]), unfolded, lu, ldu, folded, ldp,
                       mask, ldm);
    return hipGetLastError();
}

hipError_t launch_collocate_folded(int L, const double *co, const double *P, int ldp, const int *mask,
                                   int ldm, const int *gs, const int *lc, double *density,
                                   hipStream_t st) {
    const long long rows = (long long)gs[0] * gs[1];
    const int blocks = (int)((rows + 127) / 128);
    if (blocks == 0 || gs[2] == 0) return hipSuccess;
    hipLaunchKernelGGL(k_collocate_folded, dim3(blocks), dim3(128), 0, st, L, co, P, ldp, mask, ldm,
                       make_int3(gs[0], gs[1], gs[2]), make_int3(lc[0], lc[1], lc[2]), density);
    return hipGetLastError();
}

hipError_t launch_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                                   const int *gs, const int *lc, const double *V, double *co,
                                   hipStream_t st) {
    hipError_t e = launch_fill(co, (long long)L * L * L, 0.0, st);
    if (e != hipSuccess || gs[0] == 0 || gs[1] == 0 || gs[2] == 0) return e;
    hipLaunchKernelGGL(k_integrate_folded, dim3(gs[0]), dim3(256), sizeof(double) * L * L, st, L, P,
                       ldp, mask, ldm, make_int3(gs[0], gs[1], gs[2]), make_int3(lc[0], lc[1], lc[2]),
                       V, co);
    return hipGetLastError();
}

}  // namespace gpw
