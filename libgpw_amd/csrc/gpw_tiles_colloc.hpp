// gpw_tiles_colloc.hpp -- collocate tile kernels k_collocate_tiles<LM, GEN>
// and their launcher: grid (or a split buffer) += sum over the (sub-box, tile)
// work items of each tile (z, y, x contraction of the reference's
// src/lib/grid_base_ref_c.cc:403-508 and :593-846).
#pragma once

#include "gpw_tiles.hpp"

namespace gpw {

// Collocate chunk configuration: the general-cell kernels run one workgroup
// per CU, so their class L <= 12 takes four pairs per chunk (with four the
// table phase keeps every wave busy). The class L <= 8 keeps four too: eight
// shorten its own time a little but double its LDS, and less of the other
// work of the step then fits beside it on a CU.
template <int LM, bool GEN>
struct CollocCfg : TileCfg<LM> {
    static constexpr int Q = (GEN && LM == 12) ? 4 : TileCfg<LM>::Q;
    static constexpr int NPF = (Q * TileCfg<LM>::C3M + NROW - 1) / NROW;
};

template <int LM, bool GEN>
struct CollocSmem {
    using C = CollocCfg<LM, GEN>;
    GenRows<C::Q, GEN> gr;
    EntryRec rec[3][C::Q];
    // coefficients: orthorhombic cells at cidx(lp, lzp, lyp, lxp); general
    // cells in the layout of the class's largest degree, cidx(LM - 1, ...),
    // zero above the pair's lp
    double coef[2][C::Q][C::C3M];
    double X[C::Q][C::XS];
    double Y[C::Q][TY][C::YS];
    double cxy[C::Q][TZ][C::C2M];  // at c2idx(lp, ...); general cells c2idx(LM - 1, ...)
    int dec[GEN ? C::C3M : 1];     // general cells: (lzp, lyp, lxp) of each coef entry
    int xi[C::Q][TX];
    int T[C::Q][NROW];
    int bmax[C::Q][NXG], bmin[C::Q][NXG];
    int lp[C::Q];
};

// Register row r of a thread holds grid row 4*yb + (r ^ rx): general cells
// permute the rows by rx = xg, so that the four lanes of a quad share the y
// stage (each computes one row; orthorhombic cells, whose kernels are short
// of registers, compute all four rows per lane and keep rx = 0).
template <bool GEN>
__device__ __forceinline__ int colloc_rx(int xg) { return GEN ? xg : 0; }

// register block of one thread for one pair: rows y = 4*yb + (r ^ rx) of
// plane z in register row r, points x = 8*xg + j. General cells: entered by
// every lane of the wave (the quad exchanges the y stage's rows); lanes whose
// block has no point inside (active false) stop after the y stage.
template <int L, int LM, bool GEN, bool SLOW>
__device__ __forceinline__ bool colloc_block(const CollocSmem<LM, GEN> &sm, int q, int xg, int yb,
                                             int z, const int (&Tr)[RB], bool full,
                                             const GenBlock &gb, double (&acc)[RB][XB],
                                             bool active) {
    constexpr int lp = L - 1;
    (void)full;
    // general cells: the Gaussian's row factors before the y stage, so that
    // their exponentials are not live beside cx; a block the recursion
    // cannot take is left to the per-point pass (SLOW)
    double e[RB], rho[RB];
    bool fast = true;
    const int rx = colloc_rx<GEN>(xg);
    if constexpr (GEN && !SLOW) fast = gen_row_start(sm.gr, q, z, yb, gb, e, rho, rx);
    double cx[RB][L];
    const double *cxy = sm.cxy[q][z];
    if constexpr (GEN && L >= 3) {
        // y stage of the lane's own row 4*yb + xg; lane xg ^ s holds row
        // 4*yb + (s ^ xg) in its register row 0, so three DPP moves per
        // power fill the other register rows (a quarter of the FMAs of each
        // lane computing all four rows)
#pragma unroll
        for (int l = 0; l < L; l++) cx[0][l] = 0.0;
#pragma unroll
        for (int lyp = 0; lyp <= lp; lyp++) {
            const double y = sm.Y[q][yb * RB + xg][lyp];
#pragma unroll
            for (int lxp = 0; lxp <= lp - lyp; lxp++) cx[0][lxp] += cxy[c2idx(LM - 1, lyp, lxp)] * y;
        }
        static_assert(RB == 4, "row exchange within a lane quad");
#pragma unroll
        for (int l = 0; l < L; l++) {
            cx[1][l] = dpp_f64<0xB1>(cx[0][l]);
            cx[2][l] = dpp_f64<0x4E>(cx[0][l]);
            cx[3][l] = dpp_f64<0x1B>(cx[0][l]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < RB; r++)
#pragma unroll
            for (int l = 0; l < L; l++) cx[r][l] = 0.0;
        // y stage, rows innermost: each cxy value is loaded once and used by
        // its RB FMAs at once
#pragma unroll
        for (int lyp = 0; lyp <= lp; lyp++) {
            double y[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) y[r] = sm.Y[q][yb * RB + (r ^ rx)][lyp];
#pragma unroll
            for (int lxp = 0; lxp <= lp - lyp; lxp++) {
                const double c = cxy[GEN ? c2idx(LM - 1, lyp, lxp) : c2idx(lp, lyp, lxp)];
#pragma unroll
                for (int r = 0; r < RB; r++) cx[r][lxp] += c * y[r];
            }
        }
    }
    if (!active) return true;
    if (!fast) return false;
    const double *Xg = sm.X[q] + xg * TileCfg<LM>::GS;
    if constexpr (GEN) {
        if constexpr (!SLOW) {
#pragma unroll
            for (int j = 0; j < XB; j++) {
                double xv[L];
#pragma unroll
                for (int l = 0; l < L; l++) xv[l] = Xg[j * LM + l];
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    double v = 0.0;
#pragma unroll
                    for (int l = 0; l < L; l++) v += cx[r][l] * xv[l];
                    const double m = (gb.bits[r] & (1 << j)) ? e[r] : 0.0;
                    acc[r][j] = __builtin_fma(v, m, acc[r][j]);
                    if (j < XB - 1) {
                        e[r] *= rho[r];
                        rho[r] *= gb.kap;
                    }
                }
            }
        } else {
            // per-point exponentials (a block across a periodic wrap, or far
            // from a steep Gaussian): one point per iteration, so that only
            // RB exponentials are live at a time
#pragma unroll 1
            for (int j = 0; j < XB; j++) {
                double xv[L];
#pragma unroll
                for (int l = 0; l < L; l++) xv[l] = Xg[j * LM + l];
                const int code = sm.xi[q][xg * XB + j];
                const double u = (double)code - gb.gp0;
                double d[RB];
#pragma unroll
                for (int r = 0; r < RB; r++) {
                    double v = 0.0;
#pragma unroll
                    for (int l = 0; l < L; l++) v += cx[r][l] * xv[l];
                    const int row = z * TY + yb * RB + (r ^ rx);
                    double qb, qc;
                    gen_row_coef(sm.gr, q, z, yb * RB + (r ^ rx), qb, qc);
                    const bool in = code >= sm.gr.lo[q][row] && code <= sm.gr.hi[q][row];
                    d[r] = in ? v * exp((gb.qa * u + qb) * u + qc) : 0.0;
                }
#pragma unroll
                for (int jj = 0; jj < XB; jj++)
#pragma unroll
                    for (int r = 0; r < RB; r++) acc[r][jj] += jj == j ? d[r] : 0.0;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < XB; j++) {
            double xv[L];
#pragma unroll
            for (int l = 0; l < L; l++) xv[l] = Xg[j * LM + l];
            const int code = sm.xi[q][xg * XB + j];
#pragma unroll
            for (int r = 0; r < RB; r++) {
                double v = 0.0;
#pragma unroll
                for (int l = 0; l < L; l++) v += cx[r][l] * xv[l];
                // the sphere mask as a 0/1 factor of one FMA (cheaper than a
                // 64-bit select and an add)
                const double m = (full || code <= Tr[r]) ? 1.0 : 0.0;
                acc[r][j] = __builtin_fma(v, m, acc[r][j]);
            }
        }
    }
    return true;
}


// coefficient at entry (lzp, lyp, lxp) = d of the LM - 1 layout for the
// pair of record R (zero above its lp); coefficients are stored per pair in
// the cidx(lp, ...) order
template <int LM>
__device__ __forceinline__ double coef_entry(int d, const EntryRec &R, const double *__restrict__ coef) {
    const int lz = d & 31, ly = (d >> 5) & 31, lx = d >> 10;
    return lz + ly + lx <= R.lp ? coef[R.coef_off + cidx(R.lp, lz, ly, lx)] : 0.0;
}

// first (lyp, lxp) term (c2idx(LM - 1, ...) order) of part P of the z stage:
// four parts of about equal numbers of lzp terms
template <int LM>
__host__ __device__ constexpr int zpart_begin(int P) {
    int total = 0;
    for (int ly = 0; ly < LM; ly++)
        for (int lx = 0; lx < LM - ly; lx++) total += LM - ly - lx;
    int acc = 0, c = 0;
    for (int ly = 0; ly < LM; ly++)
        for (int lx = 0; lx < LM - ly; lx++) {
            if (4 * acc >= P * total) return c;
            acc += LM - ly - lx;
            c++;
        }
    return c;
}

// z stage, part P: cxy[q][z][(lyp, lxp)] = sum_lzp coef[q][(lzp, lyp, lxp)]
// Z_q(z)^lzp for the terms of part P, all pairs of the chunk at once (lane
// -> pair q, z-plane), over the LM - 1 layout: the entries above a pair's lp
// are zero and add exact zeros, so each sum equals the one over its own lp
// in the same order
template <int LM, int P, bool GEN>
__device__ __forceinline__ void colloc_zpart(CollocSmem<LM, GEN> &sm, int rs, int cs, int nq,
                                             int lane, int z0, const TileGrid &tg) {
    constexpr int lpm = LM - 1;
    constexpr int BEG = zpart_begin<LM>(P), END = zpart_begin<LM>(P + 1);
    using C = CollocCfg<LM, GEN>;
    for (int sb = 0; sb < C::Q * TZ; sb += 64) {
        const int slot = sb + lane, q = slot / TZ, pos = slot % TZ;
        if (q >= nq) continue;
        const EntryRec &R = sm.rec[rs][q];
        // the z table entry of the table phase
        const int g = z0 + pos;
        int i = BIG;
        bool ok = g < tg.nz;
        if (ok) {
            i = R.lb[2] + imod(g - R.base[2], tg.period[2]);
            ok = i <= R.ub[2];
        }
        const double r = GEN ? (double)i - R.roff[2] : i * tg.dr[2] - R.roff[2];
        double v = ok ? (GEN ? 1.0 : exp(-R.zetp * r * r)) : 0.0;
        double tv[LM];
#pragma unroll
        for (int l = 0; l < LM; l++) {
            tv[l] = v;
            v *= r;
        }
        const double *cf = sm.coef[cs][q];
        double *out = sm.cxy[q][pos];
#pragma unroll
        for (int lyp = 0; lyp <= lpm; lyp++)
#pragma unroll
            for (int lxp = 0; lxp <= lpm - lyp; lxp++) {
                const int c2 = c2idx(lpm, lyp, lxp);
                if (c2 < BEG || c2 >= END) continue;
                double a = 0.0;
#pragma unroll
                for (int lzp = 0; lzp <= lpm - lyp - lxp; lzp++) a += cf[cidx(lpm, lzp, lyp, lxp)] * tv[lzp];
                out[c2] = ok ? a : 0.0;
            }
    }
}

// out_mode 0: grid += acc (splits == 1); 1: partial[split] = acc; 2: partial[split] += acc
// General-cell collocate kernels run one workgroup per CU: the row data of
// the general path would otherwise spill registers in the inner loops
// (measured faster despite the lower occupancy). The general integrate
// kernels keep two (their x stage runs in two halves of the powers above
// L = 8, which measured faster than one workgroup per CU without spills).
// One workgroup's share of a tile: workgroup bid = tile * splits + split;
// split buffers are out + (split_base + split) * grid points.
template <int LM, bool GEN>
__device__ __forceinline__ void collocate_tile(CollocSmem<LM, GEN> &sm, int bid, const TileGrid &tg,
                                               const EntryRec *__restrict__ recs,
                                               const int *__restrict__ tile_start,
                                               const double *__restrict__ coef, int splits,
                                               double *__restrict__ out, int out_mode, int split_base) {
    using C = CollocCfg<LM, GEN>;
    constexpr int Q = C::Q;
    static_assert(Q * RECW <= NROW, "one record word per thread in the prefetch");
    const int tile = bid / splits, split = bid % splits;
    const int tix = tile % tg.ntx, tiy = (tile / tg.ntx) % tg.nty, tiz = tile / (tg.ntx * tg.nty);
    const int x0 = tix * TX, y0 = tiy * TY, z0 = tiz * TZ;
    const int t = threadIdx.x, xg = t & 3, yb = (t >> 2) & 3, z = t >> 4;
    const int beg = tile_start[tile], end = tile_start[tile + 1];
    const int per = (end - beg + splits - 1) / splits;
    const int mybeg = min(end, beg + split * per), myend = min(end, mybeg + per);
    if (out_mode != 1 && mybeg >= myend) return;
    const int *rec_words = reinterpret_cast<const int *>(recs);

    // prologue: records of chunks 0 and 1, the coefficient layout's decode
    // table, then coefficients of chunk 0
    if constexpr (GEN) {
        for (int c = t; c < C::C3M; c += NROW) {
            int lz, ly, lx;
            decode_c3(LM - 1, c, lz, ly, lx);
            sm.dec[c] = lz | ly << 5 | lx << 10;
        }
    }
    for (int w = t; w < 2 * Q * RECW; w += NROW) {
        const int item = mybeg + w / RECW;
        if (item < myend)
            reinterpret_cast<int *>(&sm.rec[w / (Q * RECW)][0])[w % (Q * RECW)] =
                rec_words[(long long)item * RECW + w % RECW];
    }
    __syncthreads();
    for (int w = t; w < Q * C::C3M; w += NROW) {
        const int q = w / C::C3M, idx = w % C::C3M;
        if constexpr (GEN)
            sm.coef[0][q][idx] = mybeg + q < myend ? coef_entry<LM>(sm.dec[idx], sm.rec[0][q], coef) : 0.0;
        else if (mybeg + q < myend && idx < ncoset(sm.rec[0][q].lp))
            sm.coef[0][q][idx] = coef[sm.rec[0][q].coef_off + idx];
    }
    __syncthreads();

    double acc[RB][XB];
#pragma unroll
    for (int r = 0; r < RB; r++)
#pragma unroll
        for (int j = 0; j < XB; j++) acc[r][j] = 0.0;

    int k = 0;
    for (int c = mybeg; c < myend; c += Q, k++) {
        const int nq = min(Q, myend - c);
        const int rs = k % 3, cs = k & 1;
        // prefetch: records of chunk k+2, coefficients of chunk k+1
        int pf_rec = 0;
        if (t < Q * RECW && c + 2 * Q + t / RECW < myend)
            pf_rec = rec_words[(long long)(c + 2 * Q) * RECW + t];
        double pf_coef[C::NPF];
#pragma unroll
        for (int j = 0; j < C::NPF; j++) {
            pf_coef[j] = 0.0;
            const int w = t + j * NROW;
            if (w < Q * C::C3M) {
                const int q = w / C::C3M, idx = w % C::C3M;
                const EntryRec &Rn = sm.rec[(k + 1) % 3][q];
                if constexpr (GEN) {
                    if (c + Q + q < myend) pf_coef[j] = coef_entry<LM>(sm.dec[idx], Rn, coef);
                } else if (c + Q + q < myend && idx < ncoset(Rn.lp)) {
                    pf_coef[j] = coef[Rn.coef_off + idx];
                }
            }
        }
        // table phase: wave w -> pairs w, w + 4, ...
        {
            const int lane = t % 64;
            for (int q = t / 64; q < nq; q += NROW / 64) {
                const EntryRec &R = sm.rec[rs][q];
                const int lp = R.lp;
                if (lane == 0) sm.lp[q] = lp;
            // orthorhombic cells: the z lanes of the table phase also run the
            // z stage of their pair
#define GPW_P1_CALL(L)                                                                        \
    tile_phase1<L, LM, !GEN, GEN, Q>(R, q, lane, x0, y0, z0, tg, GEN ? nullptr : sm.coef[cs][q], \
                                     sm.X[q], sm.Y[q], nullptr, GEN ? nullptr : sm.cxy[q],      \
                                     sm.xi[q], sm.T[q], sm.bmax[q], sm.bmin[q], sm.gr)
                GPW_LP_DISPATCH(LM, GPW_P1_CALL)
#undef GPW_P1_CALL
            }
            if constexpr (GEN) {
                // general cells: z stage of every pair of the chunk, wave w
                // takes part w of the (lyp, lxp) terms, lane -> (pair, z-plane)
                switch (t / 64) {
                case 0: colloc_zpart<LM, 0, GEN>(sm, rs, cs, nq, lane, z0, tg); break;
                case 1: colloc_zpart<LM, 1, GEN>(sm, rs, cs, nq, lane, z0, tg); break;
                case 2: colloc_zpart<LM, 2, GEN>(sm, rs, cs, nq, lane, z0, tg); break;
                default: colloc_zpart<LM, 3, GEN>(sm, rs, cs, nq, lane, z0, tg); break;
                }
            }
        }
        __syncthreads();
        // commit the prefetched data to the LDS rings (slots this chunk does
        // not read), so that its registers are free in the point phase
        if (t < Q * RECW) reinterpret_cast<int *>(&sm.rec[(k + 2) % 3][0])[t] = pf_rec;
#pragma unroll
        for (int j = 0; j < C::NPF; j++) {
            const int w = t + j * NROW;
            if (w < Q * C::C3M) sm.coef[cs ^ 1][w / C::C3M][w % C::C3M] = pf_coef[j];
        }
        unsigned slow = 0;  // general cells: pairs whose block takes the per-point pass
#pragma unroll 1
        for (int q = 0; q < nq; q++) {
            int Tr[RB];
            GenBlock gb;
            bool full;
            const bool active =
                classify_block<GEN>(sm, q, xg, yb, z, Tr, gb, full, colloc_rx<GEN>(xg));
            // general cells: wave-uniform, so the quads stay whole
            if (GEN ? !__any(active) : !active) continue;
            const int lp = sm.lp[q];
            bool done = true;
#define GPW_COLLOC_CALL(L)                                                                     \
    done = colloc_block<L, LM, GEN, false>(sm, q, xg, yb, z, Tr, full, gb, acc, active)
            GPW_LP_DISPATCH(LM, GPW_COLLOC_CALL)
#undef GPW_COLLOC_CALL
            if (!done) slow |= 1u << q;
        }
        if constexpr (GEN) {
            // the rare blocks with per-point exponentials, in a region of
            // their own (its registers do not add to the pass above)
#pragma unroll 1
            for (int q = 0; q < nq; q++) {
                const bool mine = slow >> q & 1u;
                if (!__any(mine)) continue;
                int Tr[RB];
                GenBlock gb;
                bool full;
                classify_block<GEN>(sm, q, xg, yb, z, Tr, gb, full, colloc_rx<GEN>(xg));
                const int lp = sm.lp[q];
#define GPW_COLLOC_CALL(L) colloc_block<L, LM, GEN, true>(sm, q, xg, yb, z, Tr, full, gb, acc, mine)
                GPW_LP_DISPATCH(LM, GPW_COLLOC_CALL)
#undef GPW_COLLOC_CALL
            }
        }
        __syncthreads();
    }
    const int gz = z0 + z;
    if (gz >= tg.nz) return;
    const long long ng = (long long)tg.nx * tg.ny * tg.nz;
    double *dst = out_mode == 0 ? out : out + (split_base + split) * ng;
#pragma unroll
    for (int r = 0; r < RB; r++) {  // register row r: grid row r ^ rx
        const int gy = y0 + yb * RB + (r ^ colloc_rx<GEN>(xg));
        if (gy >= tg.ny) continue;
        double *row = dst + ((long long)gz * tg.ny + gy) * tg.nx + x0 + xg * XB;
#pragma unroll
        for (int j = 0; j < XB; j++) {
            if (x0 + xg * XB + j < tg.nx) {
                if (out_mode == 1)
                    row[j] = acc[r][j];
                else
                    row[j] += acc[r][j];
            }
        }
    }
}

template <int LM, bool GEN>
__global__ __launch_bounds__(NROW, GEN ? 1 : 2) void k_collocate_tiles(
    TileGrid tg, const EntryRec *__restrict__ recs, const int *__restrict__ tile_start,
    const double *__restrict__ coef, int splits, double *__restrict__ out, int out_mode) {
    __shared__ CollocSmem<LM, GEN> sm;
    collocate_tile<LM, GEN>(sm, blockIdx.x, tg, recs, tile_start, coef, splits, out, out_mode, 0);
}

// Two classes in one launch (workgroups of class A first, then class B),
// each into split buffers of its own (A: 0 .. splitsA - 1, B: splitsA ..
// splitsA + splitsB - 1, written, not added), so the tail of one class
// overlaps the other; LDS is the union of both.
template <int LMA, int LMB, bool GEN>
__global__ __launch_bounds__(NROW, 2) void k_collocate_tiles2(
    TileGrid tg, const EntryRec *__restrict__ recsA, const int *__restrict__ startA, int splitsA,
    const EntryRec *__restrict__ recsB, const int *__restrict__ startB, int splitsB,
    const double *__restrict__ coef, int ntiles, double *__restrict__ partial) {
    __shared__ union {
        CollocSmem<LMA, GEN> a;
        CollocSmem<LMB, GEN> b;
    } sm;
    const int na = ntiles * splitsA;
    if ((int)blockIdx.x < na)
        collocate_tile<LMA, GEN>(sm.a, blockIdx.x, tg, recsA, startA, coef, splitsA, partial, 1, 0);
    else
        collocate_tile<LMB, GEN>(sm.b, blockIdx.x - na, tg, recsB, startB, coef, splitsB, partial, 1,
                                 splitsA);
}

// ---------------------------------------------------------------------------
// launcher (host side), instantiated once per cell kind in
// gpw_tiles_colloc_ortho.hip / gpw_tiles_colloc_general.hip so the two
// kernel families compile side by side.
// class c handles L = lp + 1 in (4c, 4c + 4], the last class up to 17
template <bool GEN>
hipError_t launch_collocate_tiles_t(int cls, const TileGrid &tg, const EntryRec *recs,
                                    const int *tile_start, const double *coef, int ntiles, int splits,
                                    double *out, int out_mode, hipStream_t st) {
    const dim3 grid(ntiles * splits), block(NROW);
    switch (cls) {
    case 0:
        hipLaunchKernelGGL((k_collocate_tiles<4, GEN>), grid, block, 0, st, tg, recs, tile_start, coef,
                           splits, out, out_mode);
        break;
    case 1:
        hipLaunchKernelGGL((k_collocate_tiles<8, GEN>), grid, block, 0, st, tg, recs, tile_start, coef,
                           splits, out, out_mode);
        break;
    case 2:
        hipLaunchKernelGGL((k_collocate_tiles<12, GEN>), grid, block, 0, st, tg, recs, tile_start, coef,
                           splits, out, out_mode);
        break;
    default:
        hipLaunchKernelGGL((k_collocate_tiles<17, GEN>), grid, block, 0, st, tg, recs, tile_start, coef,
                           splits, out, out_mode);
        break;
    }
    return hipGetLastError();
}

}  // namespace gpw
