// gpw_tiles_integ.hpp -- integrate tile kernels k_integrate_tiles<LM, GEN>
// (and two classes in one launch, k_integrate_tiles2) and their launchers:
// partial moments of every (sub-box, tile) work item, the transpose of the
// collocate contraction (libgpw src/lib/collocate.cc:438-541 order).
#pragma once

#include "gpw_tiles.hpp"

namespace gpw {

// Per-wave buffer of the integrate transposition: the x-stage results of the
// wave's 64 rows (4 z-planes x 16 y), row (zl, y) at zl * ZS + y * RS. A store
// writes rows 4 * yb + r (yb, zl = 0..3) of one r at once; RS = 1 (mod 4) and
// ZS = 1 (mod 16) put those 16 rows on 16 distinct 8-byte bank pairs.
template <int LM>
struct RowT1 {
    static constexpr int RS = LM + ((1 - LM) % 4 + 4) % 4;
    static constexpr int ZS = TY * RS + 1;
    static constexpr int SZ = 4 * ZS;
};

template <int LM, bool GEN>
struct IntegSmem {
    using C = TileCfg<LM>;
    GenRows<C::Q, GEN> gr;
    EntryRec rec[2][C::Q];
    double X[C::Q][C::XS];
    double Y[C::Q][TY][C::YS];
    double Z[C::Q][TZ][LM];
    double uz[C::Q][TZ][C::C2M];  // per-plane (lyp, lxp) moments
    double t1[NROW / 64][RowT1<LM>::SZ];
    int xi[C::Q][TX];
    int T[C::Q][NROW];
    int bmax[C::Q][NXG], bmin[C::Q][NXG];
    int lp[C::Q];
};

// x stage of integ_block for the powers L0 <= l < L1 of the x table, summed
// over the lane quad and stored to the wave's transposition rows t1. Register
// row r of lane xg holds grid row yb * RB + (r ^ xg), so the quad sum is a
// reduce-scatter with the same register operands on every lane: three DPP
// moves and adds per power leave row xg's sum in register row 0.
template <int L0, int L1, int LM, bool GEN>
__device__ __forceinline__ void integ_x_part(IntegSmem<LM, GEN> &sm, int q, int xg, int yb, int z,
                                             const int (&Tr)[RB], int mode, const GenBlock &gb,
                                             const double (&v)[RB][XB], double *t1, bool fast,
                                             const double (&e0)[RB], const double (&rho0)[RB]) {
    constexpr int NL = L1 - L0;
    double tx[RB][NL];
#pragma unroll
    for (int r = 0; r < RB; r++)
#pragma unroll
        for (int l = 0; l < NL; l++) tx[r][l] = 0.0;
    if (mode != 0) {  // 1: masked, 2: whole block inside
        const double *Xg = sm.X[q] + xg * TileCfg<LM>::GS;
        const bool wfull = __all(mode == 2);  // wave-uniform among the active lanes
        if constexpr (GEN) {
            double e[RB], rho[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) {
                e[r] = e0[r];
                rho[r] = rho0[r];
            }
            if (fast) {
#pragma unroll
                for (int j = 0; j < XB; j++) {
                    double xv[NL];
#pragma unroll
                    for (int l = 0; l < NL; l++) xv[l] = Xg[j * LM + L0 + l];
#pragma unroll
                    for (int r = 0; r < RB; r++) {
                        const double w = (gb.bits[r] & (1 << j)) ? v[r][j] * e[r] : 0.0;
#pragma unroll
                        for (int l = 0; l < NL; l++) tx[r][l] += w * xv[l];
                        if (j < XB - 1) {
                            e[r] *= rho[r];
                            rho[r] *= gb.kap;
                        }
                    }
                }
            } else {
                // per-point exponentials (a block across a periodic wrap, or
                // far from a steep Gaussian): one point per iteration, so that
                // only RB exponentials are live at a time
#pragma unroll 1
                for (int j = 0; j < XB; j++) {
                    double xv[NL];
#pragma unroll
                    for (int l = 0; l < NL; l++) xv[l] = Xg[j * LM + L0 + l];
                    const int code = sm.xi[q][xg * XB + j];
                    const double u = (double)code - gb.gp0;
#pragma unroll
                    for (int r = 0; r < RB; r++) {
                        double vj = 0.0;
#pragma unroll
                        for (int jj = 0; jj < XB; jj++) vj = jj == j ? v[r][jj] : vj;
                        const int row = z * TY + yb * RB + (r ^ xg);
                        double qb, qc;
                        gen_row_coef(sm.gr, q, z, yb * RB + (r ^ xg), qb, qc);
                        const bool in = code >= sm.gr.lo[q][row] && code <= sm.gr.hi[q][row];
                        const double w = in ? vj * exp((gb.qa * u + qb) * u + qc) : 0.0;
#pragma unroll
                        for (int l = 0; l < NL; l++) tx[r][l] += w * xv[l];
                    }
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < XB; j++) {
                double xv[NL];
#pragma unroll
                for (int l = 0; l < NL; l++) xv[l] = Xg[j * LM + L0 + l];
                const int code = sm.xi[q][xg * XB + j];
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    double w = v[r][j];
                    if (!wfull) w = code <= Tr[r] ? w : 0.0;
#pragma unroll
                    for (int l = 0; l < NL; l++) tx[r][l] += w * xv[l];
                }
            }
        }
    }
    // rows of the quad summed over its four x blocks: lane xg ^ 1 holds row xg
    // in register row 1 and row xg ^ 2 in register row 3, lane xg ^ 2 the
    // pair sum of row xg in register row 2; lane xg stores row xg
    static_assert(RB == 4, "reduce-scatter over a lane quad");
#pragma unroll
    for (int l = 0; l < NL; l++) {
        tx[0][l] += dpp_f64<0xB1>(tx[1][l]);
        tx[2][l] += dpp_f64<0xB1>(tx[3][l]);
        tx[0][l] += dpp_f64<0x4E>(tx[2][l]);
        t1[(yb * RB + xg) * RowT1<LM>::RS + L0 + l] = tx[0][l];
    }
}

// Register block (rows y = 4*yb + r of plane z, points x = 8*xg + j) of pair
// q: x stage in registers, sum of the four x blocks of each row over the lane
// quad (DPP), then a wave-local transposition through LDS for the y stage:
// lane (zl, c) of the wave sums Y[y][lyp] * tx[zl][y][lxp] over the 16 rows
// of its plane for the (lyp, lxp) terms c, c + 16, ... (from L = 6 on, for
// the 2 x 2 blocks of terms c, c + 16, ...).  A wave holds four
// whole z-planes, so the transposition needs no workgroup barrier (LDS
// operations of one wave complete in order).
template <int L, int LM, bool GEN>
__device__ __forceinline__ void integ_block(IntegSmem<LM, GEN> &sm, int q, int xg, int yb, int z,
                                            const int (&Tr)[RB], int mode, const GenBlock &gb,
                                            const double (&v)[RB][XB], int t) {
    constexpr int lp = L - 1;
    constexpr int N2 = L * (L + 1) / 2;
    constexpr int ZS = RowT1<LM>::ZS;
    const int lane = t & 63, wave = t >> 6;
    const int zl = lane >> 4, cq = lane & 15;
    double *uzp = sm.uz[q][wave * 4 + zl];
    if (!__any(mode != 0)) {  // no point of the wave's four planes is inside
        for (int c = cq; c < N2; c += 16) uzp[c] = 0.0;
        return;
    }
    // general cells above L = 8: two passes over half the powers each (half
    // the accumulator registers: no spills at two workgroups per CU)
    constexpr int LA = (GEN && L > 8) ? (L + 1) / 2 : L;
    double *t1 = sm.t1[wave] + zl * ZS;
    // general cells: the Gaussian's row factors once for both halves
    double e0[RB], rho0[RB];
    bool fast = false;
    if constexpr (GEN) {
        if (mode != 0) fast = gen_row_start(sm.gr, q, z, yb, gb, e0, rho0, xg);
    }
    integ_x_part<0, LA>(sm, q, xg, yb, z, Tr, mode, gb, v, t1, fast, e0, rho0);
    if constexpr (LA < L) integ_x_part<LA, L>(sm, q, xg, yb, z, Tr, mode, gb, v, t1, fast, e0, rho0);
    __builtin_amdgcn_wave_barrier();
    const double *Yq = &sm.Y[q][0][0];
    constexpr int YS = TileCfg<LM>::YS;
    constexpr int RS = RowT1<LM>::RS;
    if constexpr (L >= 6) {
        // 2 x 2 blocks (lyp0 .. lyp0 + 1, lxp0 .. lxp0 + 1), lyp0 and lxp0
        // even, lyp0 + lxp0 <= lp: two paired LDS reads per four FMAs. Terms
        // of a block past the triangle are computed from in-bounds LDS words
        // (the next row's first entries) and not stored.
        constexpr int NB = (lp / 2 + 1) * (lp / 2 + 2) / 2;
        for (int b = cq; b < NB; b += 16) {
            int by = 0, bx = b;
            while (bx > (lp - 2 * by) / 2) {
                bx -= (lp - 2 * by) / 2 + 1;
                by++;
            }
            const int ly = 2 * by, lx = 2 * bx;
            double a00 = 0.0, a01 = 0.0, a10 = 0.0, a11 = 0.0;
#pragma unroll 4
            for (int y = 0; y < TY; y++) {
                const double y0 = Yq[y * YS + ly], y1 = Yq[y * YS + ly + 1];
                const double x0 = t1[y * RS + lx], x1 = t1[y * RS + lx + 1];
                a00 += y0 * x0;
                a01 += y0 * x1;
                a10 += y1 * x0;
                a11 += y1 * x1;
            }
            const int c = c2idx(lp, ly, lx), cn = c + lp - ly + 1;
            uzp[c] = a00;
            if (lx + 1 <= lp - ly) uzp[c + 1] = a01;
            if (lx + 1 <= lp - ly - 1) uzp[cn + 1] = a11;
            if (lx <= lp - ly - 1) uzp[cn] = a10;
        }
        __builtin_amdgcn_wave_barrier();
        return;
    }
    for (int c = cq; c < N2; c += 16) {
        int lyp = 0, lxp = c;
        while (lxp > lp - lyp) {
            lxp -= lp - lyp + 1;
            lyp++;
        }
        double a = 0.0;
#pragma unroll
        for (int y = 0; y < TY; y++) a += Yq[y * YS + lyp] * t1[y * RowT1<LM>::RS + lxp];
        uzp[c] = a;
    }
    __builtin_amdgcn_wave_barrier();
}

// partial moments of every (sub-box, tile) item: partial[rec.ent_off + c3];
// workgroup bid = tile * splits + split
template <int LM, bool GEN>
__device__ __forceinline__ void integrate_tile(IntegSmem<LM, GEN> &sm, int bid, const TileGrid &tg,
                                               const EntryRec *__restrict__ recs,
                                               const int *__restrict__ tile_start,
                                               const double *__restrict__ grid, int lp_add,
                                               int splits, double *__restrict__ partial) {
    using C = TileCfg<LM>;
    constexpr int Q = C::Q;
    static_assert(Q * RECW <= NROW, "one record word per thread in the prefetch");
    const int tile = bid / splits, split = bid % splits;
    const int tix = tile % tg.ntx, tiy = (tile / tg.ntx) % tg.nty, tiz = tile / (tg.ntx * tg.nty);
    const int x0 = tix * TX, y0 = tiy * TY, z0 = tiz * TZ;
    const int t = threadIdx.x, xg = t & 3, yb = (t >> 2) & 3, z = t >> 4;
    const int beg = tile_start[tile], end = tile_start[tile + 1];
    const int per = (end - beg + splits - 1) / splits;
    const int mybeg = min(end, beg + split * per), myend = min(end, mybeg + per);
    if (mybeg >= myend) return;
    const int *rec_words = reinterpret_cast<const int *>(recs);

    for (int w = t; w < Q * RECW; w += NROW) {
        const int item = mybeg + w / RECW;
        if (item < myend)
            reinterpret_cast<int *>(&sm.rec[0][0])[w] = rec_words[(long long)item * RECW + w % RECW];
    }

    double v[RB][XB];
    {
        const int gz = z0 + z;
#pragma unroll
        for (int r = 0; r < RB; r++) {  // register row r: grid row r ^ xg
            const int gy = y0 + yb * RB + (r ^ xg);
            const bool rowok = gy < tg.ny && gz < tg.nz;
            const long long row = ((long long)gz * tg.ny + gy) * tg.nx + x0 + xg * XB;
#pragma unroll
            for (int j = 0; j < XB; j++)
                v[r][j] = (rowok && x0 + xg * XB + j < tg.nx) ? grid[row + j] : 0.0;
        }
    }
    __syncthreads();

    int k = 0;
    for (int c = mybeg; c < myend; c += Q, k++) {
        const int nq = min(Q, myend - c);
        const int rs = k & 1;
        int pf_rec = 0;
        if (t < Q * RECW && c + Q + t / RECW < myend)
            pf_rec = rec_words[(long long)(c + Q) * RECW + t];
        {
            const int lane = t % 64;
            for (int q = t / 64; q < nq; q += NROW / 64) {
                const EntryRec &R = sm.rec[rs][q];
                const int lp = R.lp + lp_add;
                if (lane == 0) sm.lp[q] = lp;
#define GPW_P1_CALL(L)                                                                         \
    tile_phase1<L, LM, false, GEN, Q>(R, q, lane, x0, y0, z0, tg, nullptr, sm.X[q], sm.Y[q],    \
                                      sm.Z[q], nullptr, sm.xi[q], sm.T[q], sm.bmax[q],          \
                                      sm.bmin[q], sm.gr)
                GPW_LP_DISPATCH(LM, GPW_P1_CALL)
#undef GPW_P1_CALL
            }
        }
        __syncthreads();
        // commit the prefetched records (a ring slot this chunk does not read)
        if (t < Q * RECW) reinterpret_cast<int *>(&sm.rec[rs ^ 1][0])[t] = pf_rec;
#pragma unroll 1
        for (int q = 0; q < nq; q++) {
            int Tr[RB];
            GenBlock gb;
            bool full = false;
            const int mode =
                classify_block<GEN>(sm, q, xg, yb, z, Tr, gb, full, xg) ? (full ? 2 : 1) : 0;
            const int lp = sm.lp[q];
#define GPW_INTEG_CALL(L) integ_block<L, LM, GEN>(sm, q, xg, yb, z, Tr, mode, gb, v, t)
            GPW_LP_DISPATCH(LM, GPW_INTEG_CALL)
#undef GPW_INTEG_CALL
        }
        __syncthreads();
        // z-stage of all pairs of the chunk at once, thread -> (q, c3):
        // M[(lzp,lyp,lxp)] = sum_z uz[q][z][(lyp,lxp)] Z[z][lzp]
        {
            int n3q[Q];
            int tot = 0;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                n3q[q] = q < nq ? ncoset(sm.lp[q]) : 0;
                tot += n3q[q];
            }
            for (int w = t; w < tot; w += NROW) {
                int q = 0, c3 = w;
#pragma unroll
                for (int qq = 0; qq < Q - 1; qq++)
                    if (q == qq && c3 >= n3q[qq]) {
                        c3 -= n3q[qq];
                        q++;
                    }
                const int lp = sm.lp[q];
                int lzp, lyp, lxp;
                decode_c3(lp, c3, lzp, lyp, lxp);
                const double *uzq = &sm.uz[q][0][c2idx(lp, lyp, lxp)];
                const double *zq = &sm.Z[q][0][lzp];
                double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
                for (int zz = 0; zz < TZ; zz += 4) {
                    a0 += uzq[(zz + 0) * C::C2M] * zq[(zz + 0) * LM];
                    a1 += uzq[(zz + 1) * C::C2M] * zq[(zz + 1) * LM];
                    a2 += uzq[(zz + 2) * C::C2M] * zq[(zz + 2) * LM];
                    a3 += uzq[(zz + 3) * C::C2M] * zq[(zz + 3) * LM];
                }
                partial[sm.rec[rs][q].ent_off + c3] = (a0 + a1) + (a2 + a3);
            }
        }
        __syncthreads();
    }
}

template <int LM, bool GEN>
__global__ __launch_bounds__(NROW, (GEN && LM >= 17) ? 1 : 2) void k_integrate_tiles(
    TileGrid tg, const EntryRec *__restrict__ recs, const int *__restrict__ tile_start,
    const double *__restrict__ grid, int lp_add, int splits, double *__restrict__ partial) {
    __shared__ IntegSmem<LM, GEN> sm;
    integrate_tile<LM, GEN>(sm, blockIdx.x, tg, recs, tile_start, grid, lp_add, splits, partial);
}

// Two classes in one launch (workgroups of class A first, then class B), so
// the tail of one class overlaps the other; LDS is the union of both.
template <int LMA, int LMB, bool GEN>
__global__ __launch_bounds__(NROW, 2) void k_integrate_tiles2(
    TileGrid tg, const EntryRec *__restrict__ recsA, const int *__restrict__ startA, int splitsA,
    const EntryRec *__restrict__ recsB, const int *__restrict__ startB, int splitsB,
    const double *__restrict__ grid, int lp_add, int ntiles, double *__restrict__ partial) {
    __shared__ union {
        IntegSmem<LMA, GEN> a;
        IntegSmem<LMB, GEN> b;
    } sm;
    const int na = ntiles * splitsA;
    if ((int)blockIdx.x < na)
        integrate_tile<LMA, GEN>(sm.a, blockIdx.x, tg, recsA, startA, grid, lp_add, splitsA, partial);
    else
        integrate_tile<LMB, GEN>(sm.b, blockIdx.x - na, tg, recsB, startB, grid, lp_add, splitsB,
                                 partial);
}

// ---------------------------------------------------------------------------
// launcher (host side), instantiated once per cell kind in
// gpw_tiles_integ_ortho.hip / gpw_tiles_integ_general.hip
template <bool GEN>
hipError_t launch_integrate_tiles_t(int cls, const TileGrid &tg, const EntryRec *recs,
                                    const int *tile_start, const double *grid, int lp_add,
                                    int ntiles, int splits, double *partial, hipStream_t st) {
    const dim3 g(ntiles * splits), block(NROW);
    switch (cls) {
    case 0:
        hipLaunchKernelGGL((k_integrate_tiles<4, GEN>), g, block, 0, st, tg, recs, tile_start, grid,
                           lp_add, splits, partial);
        break;
    case 1:
        hipLaunchKernelGGL((k_integrate_tiles<8, GEN>), g, block, 0, st, tg, recs, tile_start, grid,
                           lp_add, splits, partial);
        break;
    case 2:
        hipLaunchKernelGGL((k_integrate_tiles<12, GEN>), g, block, 0, st, tg, recs, tile_start, grid,
                           lp_add, splits, partial);
        break;
    default:
        hipLaunchKernelGGL((k_integrate_tiles<17, GEN>), g, block, 0, st, tg, recs, tile_start, grid,
                           lp_add, splits, partial);
        break;
    }
    return hipGetLastError();
}

}  // namespace gpw
