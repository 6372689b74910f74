// gpw_device.hpp -- device helpers shared by the HIP translation units:
// index helpers, the binomial table and alpha (cab -> cxyz coefficients),
// coset decoding, the effective (tau-shifted) block, the lattice transform of
// general cells, the register-block layout of the tile kernels and the
// exact row bounds of general cells.
#pragma once

#include <hip/hip_runtime.h>

#include "gpw_internal.hpp"

namespace gpw {

__device__ __forceinline__ int imod(int a, int m) {
    int r = a % m;
    return r < 0 ? r + m : r;
}
// a mod m for 0 <= a < 2m: the grid index of a sub-box offset, base + (i - lb),
// where the base lies in [0, m) and a sub-box spans at most one period
// (non-periodic axes: m = BIG and a below the grid size)
__device__ __forceinline__ int wrap_once(int a, int m) { return a >= m ? a - m : a; }

// mirror index of an offset: the reference pairs offsets ig <= 0 with 1 - ig
// (grid_base_ref_c.cc:418-474); ii' = -ig for both members of the pair.
__device__ __forceinline__ int fold_off(int i) { return i <= 0 ? -i : i - 1; }

// row threshold of the sphere: offsets along x with fold_off(i) <= T are inside
__device__ __forceinline__ int row_threshold(const int *__restrict__ prof, int prof_off,
                                             int stride, int kk, int jj) {
    if (kk < 0 || jj < 0) return -1;
    if (prof_off < 0) return BIG;
    return prof[prof_off + kk * stride + jj];
}

// ---------------------------------------------------------------------------
// coefficient transform: pab block -> coef_xyz (cab -> cxyz)
// (grid_base_ref_c.cc:65-315 as the direct triple product)

// alpha_d(lb, la, k) for k <= la + lb (the only non-zero ones), packed: per
// axis d the rows (lb, la), lb outer, each of length la + lb + 1
__host__ __device__ constexpr int alpha_doubles(int lam, int lbm) {
    return 3 * ((lam + 1) * (lbm + 1) * (lam + lbm + 2)) / 2;
}
struct AlphaView {
    double *a;
    int lam, lbm;
    __device__ AlphaView(double *a_, int lam_, int lbm_) : a(a_), lam(lam_), lbm(lbm_) {}
    // offset of row (lb, la) within a table: sum over lb' < lb, la' <= lam of
    // (la' + lb' + 1), then over la' < la of (la' + lb + 1)
    __device__ int rowoff(int lb, int la) const {
        return lb * ((lam + 1) * lam / 2) + (lam + 1) * (lb * (lb + 1) / 2) + la * (la - 1) / 2 +
               la * (lb + 1);
    }
    // rowoff(lb + 1, la) - rowoff(lb, la)
    __device__ int rowstep(int lb, int la) const { return (lam + 1) * lam / 2 + (lam + 1) * (lb + 1) + la; }
    __device__ double *table(int d) const { return a + d * (alpha_doubles(lam, lbm) / 3); }
    __device__ double &at(int d, int lb, int la, int k) const { return table(d)[rowoff(lb, la) + k]; }
};

// binomial coefficients C(n, k), n < 20 (Pascal's triangle at compile time)
struct BinomTab {
    int c[20][20];
};
__host__ __device__ constexpr BinomTab make_binom() {
    BinomTab t{};
    for (int n = 0; n < 20; n++) {
        t.c[n][0] = 1;
        for (int k = 1; k <= n; k++) t.c[n][k] = t.c[n - 1][k - 1] + (k < n ? t.c[n - 1][k] : 0);
    }
    return t;
}
static __constant__ BinomTab kBinom = make_binom();

// alpha_d(lb, la, k): coefficient of (x - xp)^k in (x - xa)^la (x - xb)^lb,
// from the binomial expansions about xp (grid_base_ref_c.cc:195-230, with
// exact integer binomials).
// pp points to the pair in global memory: its per-axis fields are read with a
// run-time axis index, which on a register copy would force it to scratch
__device__ inline void fill_alpha(const PairDev *__restrict__ pp, AlphaView A, int tid, int nthr) {
    const int n = 3 * (A.lbm + 1) * (A.lam + 1);
    for (int w = tid; w < n; w += nthr) {
        const int la = w % (A.lam + 1);
        const int lb = (w / (A.lam + 1)) % (A.lbm + 1);
        const int d = w / ((A.lam + 1) * (A.lbm + 1));
        double *row = A.table(d) + A.rowoff(lb, la);
        for (int k = 0; k <= la + lb; k++) row[k] = 0.0;
        const double pa = pp->rp[d] - pp->ra[d];
        const double pb = pp->rp[d] - pp->rb[d];
        double powa = 1.0;
        // binomials by the exact integer recurrence C(n, i + 1) = C(n, i)
        // (n - i) / (i + 1) in float (values < 2^24; the rounding of the
        // approximate reciprocal is far below 1/2)
        float ca = 1.0f;
        for (int i = 0; i <= la; i++) {
            // C(la, la - i) = C(la, i): the term (x - xp)^(la - i) (xp - xa)^i
            double powb = 1.0;
            float cb = 1.0f;
            for (int j = 0; j <= lb; j++) {
                row[(la - i) + (lb - j)] += (double)ca * (double)cb * powa * powb;
                powb *= pb;
                cb = rintf(cb * (float)(lb - j) * __builtin_amdgcn_rcpf((float)(j + 1)));
            }
            powa *= pa;
            ca = rintf(ca * (float)(la - i) * __builtin_amdgcn_rcpf((float)(i + 1)));
        }
    }
}

// q = w / d, r = w % d for 0 <= w < 2^16, 1 <= d < 2^12, with rd = an
// approximate 1 / d: the float quotient is off by at most one, then corrected
__device__ __forceinline__ void fdivmod(int w, int d, float rd, int &q, int &r) {
    q = (int)((float)w * rd);
    q -= q * d > w;
    q += (q + 1) * d <= w;
    r = w - q * d;
}

__device__ __forceinline__ void decode_coset(int l, int c, int &lx, int &ly, int &lz) {
    // inverse of coset() within shell l: c = ((l-lx)(l-lx+1))/2 + lz
    int m = 0;
    while (((m + 1) * (m + 2)) / 2 <= c) m++;
    lx = l - m;
    lz = c - (m * (m + 1)) / 2;
    ly = l - lx - lz;
}

__device__ __forceinline__ void decode_c3(int lp, int c, int &lzp, int &lyp, int &lxp) {
    lzp = 0;
    while (c >= ncoef2(lp - lzp)) {
        c -= ncoef2(lp - lzp);
        lzp++;
    }
    const int m = lp - lzp;
    lyp = 0;
    while (c > m - lyp) {
        c -= m - lyp + 1;
        lyp++;
    }
    lxp = c;
}

__device__ __forceinline__ void decode_c2(int lp, int c, int &lyp, int &lxp) {
    lyp = 0;
    while (c > lp - lyp) {
        c -= lp - lyp + 1;
        lyp++;
    }
    lxp = c;
}

// Cartesian exponents of every coset index c < NCOSET_LMAX, packed as
// lx | ly << 4 | lz << 8 (a table in LDS instead of the shell search and
// decode per lookup)
__device__ inline void fill_decode(int *tab, int tid, int nthr) {
    for (int c = tid; c < NCOSET_LMAX; c += nthr) {
        int l = 0;
        while (ncoset(l) <= c) l++;
        int x, y, z;
        decode_coset(l, c - ncoset(l - 1), x, y, z);
        tab[c] = x | (y << 4) | (z << 8);
    }
}

// (lyp, lxp) of every packed index c < C2(lp), as lyp | lxp << 8
constexpr int NC2MAX = ncoef2(LPMAX + 2);
__device__ inline void fill_decode2(int *tab, int lp, int tid, int nthr) {
    for (int c = tid; c < ncoef2(lp); c += nthr) {
        int y, x;
        decode_c2(lp, c, y, x);
        tab[c] = y | (x << 8);
    }
}

// effective block of the local (tau-shifted) ranges, eff[jco*ldE + ico]
__device__ inline void build_eff(const PairDev &p, int tau, const double *__restrict__ pab, double *eff,
                          int ldE, int nlb, int nla, int tid, int nthr) {
    for (int w = tid; w < nla * nlb; w += nthr) eff[(w / nla) * ldE + (w % nla)] = 0.0;
    __syncthreads();
    // one thread per original (a, b) entry; tau scatters to neighbours, which
    // can collide, so the tau branch is serialised over b inside one thread.
    const int na0 = ncoset(p.la_max), na_lo = ncoset(p.la_min - 1);
    const int nb0 = ncoset(p.lb_max), nb_lo = ncoset(p.lb_min - 1);
    if (!tau) {
        for (int w = tid; w < (na0 - na_lo) * (nb0 - nb_lo); w += nthr) {
            const int ico = na_lo + w % (na0 - na_lo);
            const int jco = nb_lo + w / (na0 - na_lo);
            eff[jco * ldE + ico] = pab[p.block_off + (long long)jco * p.ld + ico];
        }
        __syncthreads();
        return;
    }
    if (tid == 0) {
        const double za = p.zeta, zb = p.zetb;
        for (int jco = nb_lo; jco < nb0; jco++) {
            int lb = 0;
            while (ncoset(lb) <= jco) lb++;
            int bx, by, bz;
            decode_coset(lb, jco - ncoset(lb - 1), bx, by, bz);
            const int b[3] = {bx, by, bz};
            for (int ico = na_lo; ico < na0; ico++) {
                int la = 0;
                while (ncoset(la) <= ico) la++;
                int ax, ay, az;
                decode_coset(la, ico - ncoset(la - 1), ax, ay, az);
                const int a[3] = {ax, ay, az};
                const double v = pab[p.block_off + (long long)jco * p.ld + ico];
                for (int d = 0; d < 3; d++) {
                    int am[3] = {a[0], a[1], a[2]}, ap[3] = {a[0], a[1], a[2]};
                    int bm[3] = {b[0], b[1], b[2]}, bp[3] = {b[0], b[1], b[2]};
                    am[d] = a[d] > 0 ? a[d] - 1 : 0;
                    bm[d] = b[d] > 0 ? b[d] - 1 : 0;
                    ap[d] = a[d] + 1;
                    bp[d] = b[d] + 1;
                    const int iam = coset(am[0], am[1], am[2]), iap = coset(ap[0], ap[1], ap[2]);
                    const int jbm = coset(bm[0], bm[1], bm[2]), jbp = coset(bp[0], bp[1], bp[2]);
                    eff[jbm * ldE + iam] += 0.5 * a[d] * b[d] * v;
                    eff[jbp * ldE + iam] += 0.5 * (-2.0 * a[d] * zb) * v;
                    eff[jbm * ldE + iap] += 0.5 * (-2.0 * za * b[d]) * v;
                    eff[jbp * ldE + iap] += 0.5 * (4.0 * za * zb) * v;
                }
            }
        }
    }
    __syncthreads();
}

// General cells: out = T in (dir 0) or T^T in (dir 1) for the monomials of
// total degree <= lp, from the CSR rows of (lp, dir).
__device__ inline void lattice_transform(const LatticeTab &lt, int lp, int dir, const double *in,
                                  double *out, int tid, int nthr) {
    const LatBlk b = lt.blk[2 * lp + dir];
    const int *col = lt.col + b.base;
    const double *val = lt.val + b.base;
    for (int r = tid; r < b.n; r += nthr) {
        double a = 0.0;
        // in row order: the fill adds exact zeros at the end
#pragma unroll 4
        for (int k = 0; k < b.width; k++) a += val[k * b.n + r] * in[col[k * b.n + r]];
        out[r] = a;
    }
}

// Powers pw[m][n] = h[m]^n (n <= LPMAX) of the six steps of a stepped cell,
// in LDS.
struct StepTabs {
    double pw[6][LPMAX + 1];
};
constexpr int STEP_TABS_DOUBLES = sizeof(StepTabs) / sizeof(double);
__device__ inline void fill_step_tabs(const LatticeTab &lt, StepTabs &st, int tid, int nthr) {
    for (int m = tid; m < 6; m += nthr) {
        double v = 1.0;
        for (int n = 0; n <= LPMAX; n++) {
            st.pw[m][n] = v;
            v *= lt.h[m];
        }
    }
}

// One pass of lattice_stepped (OP 0, 1, 2: A, B, C; 3, 4, 5: A', B', C'),
// src -> dst. Output (s2, s1, s0) sums src along one line of the index
// triangle, e = 0 .. the slot that bounds it; the source index moves along
// the line by the closed-form steps of cidx (cidx(lp, z, y, x) =
// ncoset(lp) - ncoset(m) + y (m + 1) - y (y - 1) / 2 + x, m = lp - z).
template <int OP>
__device__ inline void stepped_pass(const StepTabs &st, int lp, const double *src, double *dst,
                                    int tid, int nthr) {
    constexpr int SLOT = (OP == 0 || OP == 1) ? 0 : (OP == 2 || OP == 3) ? 1 : 2;
    constexpr int SM = (OP == 0 || OP == 3) ? 1 : (OP == 1 || OP == 4) ? 2 : 4;  // the pass's step
    const double step = st.pw[SM][1];
    const int n = ncoset(lp);
    for (int w = tid; w < n; w += nthr) {
        int s0, s1, s2;
        if (SLOT == 0)
            decode_c3(lp, w, s0, s1, s2);
        else if (SLOT == 1)
            decode_c3(lp, w, s1, s0, s2);
        else
            decode_c3(lp, w, s2, s1, s0);
        const int ne = step == 0.0 ? 0 : SLOT == 0 ? s0 : SLOT == 1 ? s1 : s2;
        const int out = cidx(lp, s2, s1, s0);
        int at = out, z = s2, y = s1;  // source index and its z-like, y-like slots
        double acc = 0.0, pe = 1.0;    // pe = step^e
        // binomial of the term, C(s, e) (A, B, C) or C(s + e, e) (A', B',
        // C'), by the exact integer recurrence in float (values < 2^24; the
        // rounding of the approximate reciprocal is far below 1/2)
        const int sb = OP == 2 || OP == 5 ? s1 : s0;
        float bin = 1.0f;
        for (int e = 0; e <= ne; e++) {
            const int m = lp - z;
            const double c = (double)bin;
            int d;
            if (OP == 0) {  // (s2, s1 + e, s0 - e)
                d = m - y;
                y++;
            } else if (OP == 1) {  // (s2 + e, s1, s0 - e)
                d = (m + 1) * (m + 2) / 2 - y - 1;
                z++;
            } else if (OP == 2) {  // (s2 + e, s1 - e, s0)
                d = (m + 1) * m / 2;
                z++;
                y--;
            } else if (OP == 3) {  // (s2, s1 - e, s0 + e)
                d = y - m - 1;
                y--;
            } else if (OP == 4) {  // (s2 - e, s1, s0 + e)
                d = y + 1 - (m + 2) * (m + 3) / 2;
                z--;
            } else {  // (s2 - e, s1 + e, s0)
                d = -((m + 2) * (m + 1) / 2);
                z--;
                y++;
            }
            acc += c * pe * src[at];
            pe *= step;
            at += d;
            const int num = OP <= 2 ? sb - e : sb + e + 1;
            bin = rintf(bin * (float)num * __builtin_amdgcn_rcpf((float)(e + 1)));
        }
        if (OP == 1 || OP == 5) acc *= st.pw[3][s1];
        if (OP == 2) acc *= st.pw[5][s2];
        if (OP == 3) acc *= st.pw[0][s0];
        dst[out] = acc;
    }
}

// Lattice transform of a stepped cell (LatticeTab::stepped) in three passes
// of one Cartesian substitution each: with u = a0 i + b0 j, x = u + c0 k,
// y = b1 j + c1 k, z = c2 k and triples (z-like, y-like, x-like) powers
// packed by cidx,
// dir 1, moments of i^p j^q k^r -> of x^lx y^ly z^lz:
//   A: m1(r, q, p)     = sum_{e <= p} C(p, e) a0^(p-e) b0^e in(r, q + e, p - e)
//   B: m2(r, q, l)     = sum_{e <= l} C(l, e) c0^e m1(r + e, q, l - e)
//   C: out(n, m, l)    = c2^n sum_{e <= m} C(m, e) b1^(m-e) c1^e m2(n + e, m - e, l)
// dir 0, coefficients of x^lx y^ly z^lz -> of i^p j^q k^r (the transposes,
// in reverse order):
//   C': g(t, q, l)     = b1^q sum_{e <= t} C(q + e, e) c1^e c2^(t-e) in(t - e, q + e, l)
//   B': h(s, q, m)     = sum_{e <= s} C(m + e, e) c0^e g(s - e, q, m + e)
//   A': out(r, n, m)   = a0^m sum_{e <= n} C(m + e, e) b0^e h(r, n - e, m + e)
// (the multinomial expansion of grid_base_ref_c.cc:647-698, the tables' T,
// factorised one substitution at a time; terms of a zero step are skipped).
// The powers that vary along a sum travel with the data: the input comes
// scaled, dir 1 by a0^p (its x-like power), dir 0 by c2^lz (its z-like
// power), and pass B scales m2 by b1^q, so that each term is a binomial,
// one running power of the pass's step and one LDS read.
// Buffers: a -> b -> a -> out, a barrier between passes; a is overwritten,
// out may be b. Each pass visits its outputs ordered by the slot that
// bounds its sum, so the lanes of a wave run loops of nearly equal length.
__device__ inline void lattice_stepped(const StepTabs &st, int lp, int dir, double *a, double *b,
                                       double *out, int tid, int nthr) {
    if (dir == 1) {
        stepped_pass<0>(st, lp, a, b, tid, nthr);
        __syncthreads();
        stepped_pass<1>(st, lp, b, a, tid, nthr);
        __syncthreads();
        stepped_pass<2>(st, lp, a, out, tid, nthr);
    } else {
        stepped_pass<5>(st, lp, a, b, tid, nthr);
        __syncthreads();
        stepped_pass<4>(st, lp, b, a, tid, nthr);
        __syncthreads();
        stepped_pass<3>(st, lp, a, out, tid, nthr);
    }
}

// ---------------------------------------------------------------------------
// tile kernels
//
// Workgroup = one TX x TY x TZ tile (32 x 16 x 16 points), 256 threads. Thread
// t owns a register block of RB = 4 consecutive rows (same z) x XB = 8 points:
//   xg = t & 3 (x block), yb = (t >> 2) & 3 (row block), z = t >> 4.
// The 16 threads of one z plane are one DPP row (16 consecutive lanes).
// Every LDS table value read in the x stage feeds RB FMAs.
//
// A workgroup walks its work items in chunks of Q pairs. Table phase: wave q
// builds everything of pair q of the chunk (1D tables, x codes and block
// classes, the z-stage, the sphere threshold of every tile row) using only
// LDS and in-wave shuffles. The item records of the chunk after next and the
// coefficients of the next chunk are loaded into registers at the start of a
// chunk and stored to LDS rings at its end, so global latency stays off the
// critical path.

constexpr int XB = 8, RB = 4;
constexpr int NXG = TX / XB;  // x blocks per tile
constexpr int RECW = (int)(sizeof(EntryRec) / 4);
static_assert(TX + TY + TZ == 64, "table phase maps one wave per pair");
static_assert(NXG * (TY / RB) * TZ == NROW, "one register block per thread");
static_assert(NXG * (TY / RB) == 16, "one z plane per DPP row");
static_assert(TX == 4 * XB && TY == 16 && TZ == 16, "lane layout of the table phase");

// Index-set predicates. They must round exactly as the host planner (built
// with -ffp-contract=off) does, so every product is kept apart from the sum
// that consumes it: the functions below are written with plain operators
// under "#pragma clang fp contract(off)" (the __dmul_rn / __dadd_rn helpers
// of the HIP headers are plain operators compiled in the header's scope and
// are fused into FMAs after inlining, which moves points that lie exactly on
// a sphere or row bound in or out).

// General cells: i-range of lattice row (j, k) inside the sphere, with the
// operation order of grid_base_ref_c.cc:790-811 and no contraction, so the
// index set equals the host's and the oracle's; qb, qc: the row's quadratic
// scaled by -zetp (exponent of point i: (qa*i + qb)*i + qc).
__device__ __forceinline__ void general_row_dev(const TileGrid &tg, double gp0, double dj, double dk,
                                                double R2, double zetp, int &lo, int &hi,
                                                double &qb, double &qc) {
#pragma clang fp contract(off)
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        const double pointk = tg.dh[2][x] * dk;
        const double pointj = pointk + tg.dh[1][x] * dj;
        const double v = pointj - gp0 * tg.dh[0][x];
        a = a + tg.dh[0][x] * tg.dh[0][x];
        b = b + (2.0 * v) * tg.dh[0][x];
        c = c + v * v;
    }
    double d = b * b - (4.0 * a) * (c - R2);
    if (d < 0.0) {
        lo = 1;
        hi = 0;
        qb = qc = 0.0;
        return;
    }
    d = __dsqrt_rn(d);
    const double a2 = 2.0 * a;
    lo = (int)ceil((-b - d) / a2);
    hi = (int)floor((-b + d) / a2);
    qb = b * -zetp;
    qc = c * -zetp;
}

}  // namespace gpw
