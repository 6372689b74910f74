// gpw_tiles.hpp -- shared parts of the collocate and integrate tile kernels:
// per-class LDS layout, the table phase (1D tables, x codes, sphere and row
// thresholds of a tile for one pair), the per-block classification and the
// dispatch over the polynomial length.
#pragma once

#include "gpw_device.hpp"

namespace gpw {

template <int LM>
struct TileCfg {
    static constexpr int Q = LM <= 4 ? 8 : LM <= 8 ? 4 : 2;  // pairs per chunk
    static constexpr int C2M = LM * (LM + 1) / 2;       // (lyp, lxp) terms
    static constexpr int C3M = LM * (LM + 1) * (LM + 2) / 6;
    static constexpr int GS = XB * LM + 4;              // doubles per x block (+32 B spreads banks)
    // y-table row stride, odd: the 4 rows 4 * yb + r a wave reads at once
    // (yb = 0..3) fall into distinct banks
    static constexpr int YS = LM | 1;
    static constexpr int XS = NXG * GS;
    static constexpr int NPF = (Q * C3M + NROW - 1) / NROW;  // coef prefetch registers
    __device__ static int xat(int x) { return (x / XB) * GS + (x % XB) * LM; }
};


// per-row data of the general path (unused storage on orthorhombic grids).
// Row (y, z) is the lattice row at displacements dj = j - gp1, dk = k - gp2;
// lo..hi its points inside the sphere. With u = i - gp0 and
// w = dj h1 + dk h2, the exponent -zetp |u h0 + w|^2 of point i is
// qa u^2 + qb u + qc, qb = c01 dj + c02 dk, qc = (c11 dj + c12 dk) dj + c22 dk^2.
// kap = exp(2 qa) is the step ratio of the exponential recursion along a row.
template <int Q, bool GEN>
struct GenRows {
    int lo[Q][NROW], hi[Q][NROW];
    double dj[Q][TY], dk[Q][TZ];
    double qa[Q], kap[Q], gp0[Q], c01[Q], c02[Q], c11[Q], c12[Q], c22[Q];
    int ball[Q][NXG];  // 1: every point of the x block is in the sub-box
    // c0: when the valid points j of the x block have the lattice indices
    // c0 + j (no periodic wrap between them), else BIG; vmask: valid points
    int c0[Q][NXG], vmask[Q][NXG];
};
template <int Q>
struct GenRows<Q, false> {};

// Table phase of pair slot q, run by one whole wave (lane = t % 64).
// Lanes 0..31: x entries, 32..47: y entries, 48..63: z entries.
// WITH_CXY: z lanes also contract coef with their Z powers (collocate).
// GEN (general cells): tables are pure powers of the lattice displacement
// i - gp; the Gaussian is applied per point from the row quadratic.
template <int L, int LM, bool WITH_CXY, bool GEN, int Q>
__device__ __forceinline__ void tile_phase1(const EntryRec &R, int q, int lane, int x0, int y0,
                                            int z0, const TileGrid &tg,
                                            const double *__restrict__ coef, double *X,
                                            double (*Y)[TileCfg<LM>::YS], double (*Z)[LM],
                                            double (*cxy)[TileCfg<LM>::C2M], int *xi, int *T,
                                            int *bmax, int *bmin, GenRows<Q, GEN> &gr) {
    constexpr int lp = L - 1;
    const int d = lane < TX ? 0 : lane < TX + TY ? 1 : 2;
    const int pos = d == 0 ? lane : d == 1 ? lane - TX : lane - TX - TY;
    const int g = (d == 0 ? x0 : d == 1 ? y0 : z0) + pos;
    const int nmax = d == 0 ? tg.nx : d == 1 ? tg.ny : tg.nz;
    const int period = d == 0 ? tg.period[0] : d == 1 ? tg.period[1] : tg.period[2];
    const double dr = d == 0 ? tg.dr[0] : d == 1 ? tg.dr[1] : tg.dr[2];
    int i = BIG;
    bool ok = g < nmax;
    if (ok) {
        i = R.lb[d] + imod(g - R.base[d], period);
        ok = i <= R.ub[d];
    }
    double tv[L];
    {
        const double r = GEN ? (double)i - R.roff[d] : i * dr - R.roff[d];
        double v = ok ? (GEN ? 1.0 : exp(-R.zetp * r * r)) : 0.0;
#pragma unroll
        for (int k = 0; k < L; k++) {
            tv[k] = v;
            v *= r;
        }
    }
    int code;
    if constexpr (GEN)
        code = ok ? i : (d == 0 ? BIG : -BIG);
    else
        code = ok ? fold_off(i) : (d == 0 ? BIG : -1);
    if (d == 0) {
        double *dst = X + TileCfg<LM>::xat(pos);
#pragma unroll
        for (int k = 0; k < L; k++) dst[k] = tv[k];
        xi[pos] = code;
    } else if (d == 1) {
#pragma unroll
        for (int k = 0; k < L; k++) Y[pos][k] = tv[k];
    } else {
        if (Z) {
#pragma unroll
            for (int k = 0; k < L; k++) Z[pos][k] = tv[k];
        }
        if constexpr (WITH_CXY) {
            // z-stage: cxy[z][(lyp,lxp)] = sum_lzp coef[(lzp,lyp,lxp)] Z^lzp
#pragma unroll
            for (int lyp = 0; lyp <= lp; lyp++)
#pragma unroll
                for (int lxp = 0; lxp <= lp - lyp; lxp++) {
                    double a = 0.0;
#pragma unroll
                    for (int lzp = 0; lzp <= lp - lyp - lxp; lzp++)
                        a += coef[cidx(lp, lzp, lyp, lxp)] * tv[lzp];
                    cxy[pos][c2idx(lp, lyp, lxp)] = ok ? a : 0.0;
                }
        }
    }
    if constexpr (!GEN) {
        // x blocks: max code (BIG if any point invalid) and min code over 8 lanes
        int mx = code, mn = code;
#pragma unroll
        for (int o = 1; o < XB; o <<= 1) {
            mx = max(mx, __shfl_xor(mx, o, XB));
            mn = min(mn, __shfl_xor(mn, o, XB));
        }
        if (d == 0 && (lane % XB) == 0) {
            bmax[lane / XB] = mx;
            bmin[lane / XB] = mn;
        }
        // thresholds of the 256 tile rows (row = z * TY + y), 4 per lane
#pragma unroll
        for (int rr = 0; rr < NROW / 64; rr++) {
            const int row = lane + 64 * rr;
            const int kk = __shfl(code, TX + TY + row / TY);
            const int jj = __shfl(code, TX + row % TY);
            T[row] = (kk < 0 || jj < 0) ? -1
                     : R.prof_off < 0   ? BIG
                                        : tg.prof[R.prof_off + kk * R.prof_stride + jj];
        }
    } else {
        // x blocks: min / max lattice index of the valid points, all valid?
        // and whether the valid points share one offset i - j
        int mn = ok ? i : BIG, mx = ok ? i : -BIG, all = ok ? 1 : 0;
        int omn = ok ? i - (lane & (XB - 1)) : BIG, omx = ok ? omn : -BIG;
        int vm = ok ? 1 << (lane & (XB - 1)) : 0;
#pragma unroll
        for (int o = 1; o < XB; o <<= 1) {
            mx = max(mx, __shfl_xor(mx, o, XB));
            mn = min(mn, __shfl_xor(mn, o, XB));
            all = min(all, __shfl_xor(all, o, XB));
            omx = max(omx, __shfl_xor(omx, o, XB));
            omn = min(omn, __shfl_xor(omn, o, XB));
            vm |= __shfl_xor(vm, o, XB);
        }
        if (d == 0 && (lane % XB) == 0) {
            bmax[lane / XB] = mx;
            bmin[lane / XB] = mn;
            gr.ball[q][lane / XB] = all;
            gr.c0[q][lane / XB] = omn == omx ? omn : BIG;
            gr.vmask[q][lane / XB] = vm;
        }
#pragma unroll
        for (int rr = 0; rr < NROW / 64; rr++) {
            const int row = lane + 64 * rr;
            const int kidx = __shfl(code, TX + TY + row / TY);
            const int jidx = __shfl(code, TX + row % TY);
            int lo = 1, hi = 0;
            double qb, qc;
            if (kidx != -BIG && jidx != -BIG)
                general_row_dev(tg, R.roff[0], __dsub_rn((double)jidx, R.roff[1]),
                                __dsub_rn((double)kidx, R.roff[2]), R.R2, R.zetp, lo, hi, qb, qc);
            gr.lo[q][row] = lo;
            gr.hi[q][row] = hi;
        }
        // displacements of the tile's y and z lattice rows (0 when invalid:
        // such rows are empty)
        if (d == 1) gr.dj[q][pos] = ok ? __dsub_rn((double)i, R.roff[1]) : 0.0;
        if (d == 2) gr.dk[q][pos] = ok ? __dsub_rn((double)i, R.roff[2]) : 0.0;
        if (lane == 0) {
            double a = 0.0;
#pragma unroll
            for (int x = 0; x < 3; x++) a = __dadd_rn(a, __dmul_rn(tg.dh[0][x], tg.dh[0][x]));
            gr.qa[q] = a * -R.zetp;
            gr.kap[q] = exp(2.0 * (a * -R.zetp));
            gr.gp0[q] = R.roff[0];
            double h01 = 0.0, h02 = 0.0, h11 = 0.0, h12 = 0.0, h22 = 0.0;
#pragma unroll
            for (int x = 0; x < 3; x++) {
                h01 += tg.dh[0][x] * tg.dh[1][x];
                h02 += tg.dh[0][x] * tg.dh[2][x];
                h11 += tg.dh[1][x] * tg.dh[1][x];
                h12 += tg.dh[1][x] * tg.dh[2][x];
                h22 += tg.dh[2][x] * tg.dh[2][x];
            }
            gr.c01[q] = -2.0 * R.zetp * h01;
            gr.c02[q] = -2.0 * R.zetp * h02;
            gr.c11[q] = -R.zetp * h11;
            gr.c12[q] = -2.0 * R.zetp * h12;
            gr.c22[q] = -R.zetp * h22;
        }
        (void)T;
    }
}

// DPP move of a double (two 32-bit halves) within a lane quad; CTRL is the
// dpp_ctrl code (0xB1: quad_perm [1,0,3,2], lane ^ 1; 0x4E: [2,3,0,1],
// lane ^ 2; 0x1B: [3,2,1,0], lane ^ 3). Every lane of the quad must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double a) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(a), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(a), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Phase-3 data of one thread's register block on the general path.
// contig: the valid points j of the block have the lattice indices c0 + j;
// then bit j of bits[r] says whether point j of row r is valid and inside
// the sphere, and u0 = c0 - gp0.
struct GenBlock {
    int bits[RB];
    double qa, kap, gp0, u0;
    int bmn, bmx;
    bool contig;
};

// exponent coefficients qb, qc of row yb * RB + r of plane z (see GenRows)
template <int Q>
__device__ __forceinline__ void gen_row_coef(const GenRows<Q, true> &gr, int q, int z, int y,
                                             double &qb, double &qc) {
    const double dk = gr.dk[q][z], dj = gr.dj[q][y];
    qb = gr.c01[q] * dj + gr.c02[q] * dk;
    qc = (gr.c11[q] * dj + gr.c12[q] * dk) * dj + gr.c22[q] * dk * dk;
}

// Gaussian factors of the block's rows by the recursion along the row:
// e(u+1) = e(u) rho(u), rho(u+1) = rho(u) kap, started at u0 (two exp per
// row instead of one per point). Returns false when the recursion cannot be
// used: the block is not contiguous, or the start value of a row with points
// inside underflows (later factors would be lost). Register r holds row
// yb * RB + (r ^ rx).
template <int Q>
__device__ __forceinline__ bool gen_row_start(const GenRows<Q, true> &gr, int q, int z, int yb,
                                              const GenBlock &gb, double (&e)[RB],
                                              double (&rho)[RB], int rx = 0) {
    bool fast = gb.contig;
#pragma unroll
    for (int r = 0; r < RB; r++) {
        double qb, qc;
        gen_row_coef(gr, q, z, yb * RB + (r ^ rx), qb, qc);
        e[r] = exp((gb.qa * gb.u0 + qb) * gb.u0 + qc);
        rho[r] = exp(gb.qa * (2.0 * gb.u0 + 1.0) + qb);
        fast = fast && (gb.bits[r] == 0 || e[r] >= 1e-290);
    }
    return fast;
}

#define GPW_LP_DISPATCH(LM, CALL)                                                   \
    switch (lp) {                                                                   \
    case 0: CALL(1); break;                                                         \
    case 1: CALL(2); break;                                                         \
    case 2: CALL(3); break;                                                         \
    case 3: CALL(4); break;                                                         \
    default:                                                                        \
        if constexpr (LM > 4) {                                                     \
            switch (lp) {                                                           \
            case 4: CALL(5); break;                                                 \
            case 5: CALL(6); break;                                                 \
            case 6: CALL(7); break;                                                 \
            case 7: CALL(8); break;                                                 \
            default:                                                                \
                if constexpr (LM > 8) {                                             \
                    switch (lp) {                                                   \
                    case 8: CALL(9); break;                                         \
                    case 9: CALL(10); break;                                        \
                    case 10: CALL(11); break;                                       \
                    case 11: CALL(12); break;                                       \
                    default:                                                        \
                        if constexpr (LM > 12) {                                    \
                            switch (lp) {                                           \
                            case 12: CALL(13); break;                               \
                            case 13: CALL(14); break;                               \
                            case 14: CALL(15); break;                               \
                            case 15: CALL(16); break;                               \
                            default: CALL(17); break;                               \
                            }                                                       \
                        }                                                           \
                        break;                                                      \
                    }                                                               \
                }                                                                   \
                break;                                                              \
            }                                                                       \
        }                                                                           \
        break;                                                                      \
    }


// classification of a thread's register block for pair q: returns false when
// no point of the block is inside; sets full when every point is inside.
// Register r of Tr and gb.bits describes row yb * RB + (r ^ rx).
template <bool GEN, class Smem>
__device__ __forceinline__ bool classify_block(const Smem &sm, int q, int xg, int yb, int z,
                                               int (&Tr)[RB], GenBlock &gb, bool &full,
                                               int rx = 0) {
    const int bmn = sm.bmin[q][xg], bmx = sm.bmax[q][xg];
    if constexpr (!GEN) {
        int tmin = BIG, tmax = -1;
#pragma unroll
        for (int r = 0; r < RB; r++) {
            Tr[r] = sm.T[q][z * TY + yb * RB + (r ^ rx)];
            tmin = min(tmin, Tr[r]);
            tmax = max(tmax, Tr[r]);
        }
        if (tmax < 0 || bmn > tmax) return false;
        full = bmx <= tmin;
        return true;
    } else {
        bool any = false;
        full = sm.gr.ball[q][xg] != 0;
        const int c0 = sm.gr.c0[q][xg], vm = sm.gr.vmask[q][xg];
        gb.contig = c0 != BIG;
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int row = z * TY + yb * RB + (r ^ rx);
            const int lo = sm.gr.lo[q][row], hi = sm.gr.hi[q][row];
            any = any || (lo <= hi && lo <= bmx && hi >= bmn);
            full = full && lo <= bmn && bmx <= hi;
            // points j = max(lo - c0, 0) .. min(hi - c0, 7) (contiguous blocks)
            const int a = max(lo - c0, 0), b = min(hi - c0, XB - 1);
            gb.bits[r] = (gb.contig && a <= b) ? ((0xFF >> (XB - 1 - b)) & (0xFF << a)) & vm : 0;
            Tr[r] = 0;
        }
        gb.qa = sm.gr.qa[q];
        gb.kap = sm.gr.kap[q];
        gb.gp0 = sm.gr.gp0[q];
        gb.bmn = bmn;
        gb.bmx = bmx;
        gb.u0 = (double)c0 - gb.gp0;
        return any && bmn <= bmx;
    }
}

}  // namespace gpw
