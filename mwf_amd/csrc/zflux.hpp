// zflux.hpp -- the z pass of the flux (conservative) form of the nonlinear
// term.
//
// For the divergence-free Galerkin field, (u.grad)u = div(u u), and the
// wall-normal derivative moves onto the bases exactly (gen_galerkin.py has
// the derivation).  The z pass then needs only the KK base lines of U (u_k,
// v_k, w_k; no x- or z-derivatives): per point it forms the 6 K0 - 3
// projected products uu_i, uw_i, ww_i, uv_i, vw_i, vv_i (flux_gen.hpp),
// transforms them forward, and writes per spec slot s a line B_s (to be
// differentiated in x by the x kernel) and a line A_s = (ak1 terms) + i gamma
// n (z terms).  NL thus holds 2 KK lines per x-row (B at line s, A at KK + s),
// U only KK.  Compared with the advective z kernel this transforms
// ceil(KK/2) + ceil(NR/2) pairs instead of ceil(3 KK/2) + ceil(KK/2) (M4: 17
// against 23) and keeps KK values per point instead of 2 KK.
//
// Block: T = max(512, NN) threads, XB = T/NN x-rows; thread (xl, r) owns the
// points z = 3r + b, b = 0..2, of row xl, and the output index n = r of that
// row in every line, so a line whose rows straddle two batches is summed in
// the thread's registers.  Sequences of ZP pairs per batch; lane group g of
// G = min(64, NN/8) lanes transforms sequence g of the batch in place with
// wave-local exchanges only (group_fft), so the radix passes need no block
// barrier: a batch has one barrier between its fold and its transforms and
// one between the transforms and the unfold (NN = 1024 and the FP64 one-
// buffer form, whose registers do not hold a lane group's values, keep
// block-wide Stockham passes).  With two buffers (PP) batches
// alternate between them and need no third barrier; one buffer (at or under
// 80 KB, so that two blocks can share a CU) adds a barrier after the unfold.
#pragma once
#include "kernels.hpp"
#include "flux_gen.hpp"

template <int LOGNN> constexpr int zc_threads() { return (1 << LOGNN) > ZT ? (1 << LOGNN) : ZT; }

// Lane-group transforms where their registers fit (NN <= 512, and FP32 or
// the two-buffer form), with seven-entry pass tables; otherwise block-wide
// Stockham passes (ip_fft) with three-entry ones.
template <typename R, int LOGNN, bool PP> constexpr bool zc_wave() { return LOGNN <= 9 && (PP || sizeof(R) == 4); }
template <typename R, int LOGNN, bool PP> constexpr int zc_pw() { return zc_wave<R, LOGNN, PP>() ? 7 : 3; }
// LDS layout of the sequences: padded (lslot shift 4) for the FP32 lane-group
// transforms, XOR-swizzled otherwise (padding with shift 3 made FP64 k_zc 2 %
// slower at M4 1024^2)
template <typename R, int LOGNN, bool PP> constexpr int zc_ls() { return zc_wave<R, LOGNN, PP>() && sizeof(R) == 4 ? 4 : 0; }

// Sequence-granular forward batches: with one x-row per block, lane-group
// transforms and two buffers, a forward batch takes the next T/G (8)
// sequences in the order (pair q, phase b) -> 3 q + b, so that every lane
// group transforms; a pair whose three phases straddle two batches keeps its
// partial sums in registers.  Pair-granular batches of ZP pairs (3 ZP = 6
// sequences) leave two of the eight groups idle.  FP64 keeps batches of 6
// (its transforms are bound by LDS bandwidth, 16-byte values, and eight at
// once made FP64 k_zc 2 % slower at M4 1024^2 despite one batch in nine
// fewer; FP32 1 % faster), except that a tail of up to 2 sequences joins the
// last batch.
template <typename R, int LOGNN, bool PP> constexpr bool zc_sg()
{
    return PP && zc_wave<R, LOGNN, PP>() && zc_threads<LOGNN>() == (1 << LOGNN);
}
// forward sequences per batch (the last batch may take up to 8)
template <typename R> constexpr int zc_spb() { return sizeof(R) == 4 ? 8 : 6; }
// forward batches of NSF sequences, SPB a batch, a tail of at most 8 - SPB
// sequences joining the last full batch
template <int NSF, int SPB> constexpr int zc_nbf()
{
    const int nb = (NSF + SPB - 1) / SPB;
    return nb > 1 && NSF - (nb - 1) * SPB <= 8 - SPB ? nb - 1 : nb;
}
// sequences per buffer
template <typename R, int LOGNN, int ZP, bool PP> constexpr int zc_bseq()
{
    constexpr int NN = 1 << LOGNN, T = zc_threads<LOGNN>(), SEQ = T / NN * ZP * 3;
    constexpr int SPB = zc_sg<R, LOGNN, PP>() ? T / group_lanes<NN>() : SEQ;
    return SEQ > SPB ? SEQ : SPB;
}

// transform buffer(s), then the pass twiddle table of the NN-point
// transforms and the first NN entries of exp(2 pi i j/Q) (twq2 extends them)
template <typename R, int LOGNN, int ZP, bool PP> constexpr size_t zc_lds_bytes()
{
    constexpr int NN = 1 << LOGNN;
    constexpr int LPZ = lspan<NN, zc_ls<R, LOGNN, PP>()>();
    return sizeof(cplx_t<R>) * ((PP ? 2 : 1) * (size_t)zc_bseq<R, LOGNN, ZP, PP>() * LPZ +
                                ptw_size<NN, 1, zc_pw<R, LOGNN, PP>()>() + (size_t)NN);
}

template <typename R, int K0, int LOGNN, int ZP, int WAVES, bool PP>
__global__ __launch_bounds__(zc_threads<LOGNN>(), WAVES) void k_zc(Batch bt, const cplx_t<R>* __restrict__ U,
                                                                   cplx_t<R>* __restrict__ NL,
                                                                   const cplx_t<R>* __restrict__ twNN,
                                                                   const cplx_t<R>* __restrict__ twQ)
{
    using CT = cplx_t<R>;
    using FX = Flux<K0>;
    const Geo g = bt.geo[blockIdx.z];
    U += blockIdx.z * bt.u_stride;
    NL += blockIdx.z * bt.nl_stride;
    constexpr int KK = 3 * K0 - 1;
    constexpr int NN = 1 << LOGNN;
    constexpr int T = zc_threads<LOGNN>();
    constexpr int XB = T / NN;
    constexpr int H = NN / 2;
    constexpr int NPI = (KK + 1) / 2;                 // inverse pairs: base lines 2p, 2p + 1
    constexpr int NBI = (NPI + ZP - 1) / ZP;
    constexpr int NR = FX::NR;
    constexpr int NPO = (NR + 1) / 2;                 // forward pairs: rows order(2p), order(2p + 1)
    constexpr int NBO = (NPO + ZP - 1) / ZP;
    constexpr int SEQ = XB * ZP * 3;
    constexpr int FT = XB * ZP * H / T;               // fold tasks per thread
    static_assert(FT * T == XB * ZP * H, "fold tasks must tile the block");
    extern __shared__ unsigned char smem_raw[];
    constexpr int LS = zc_ls<R, LOGNN, PP>(), LPZ = lspan<NN, LS>(); // sequence layout and span
    constexpr bool SG = zc_sg<R, LOGNN, PP>();
    constexpr int BSEQ = zc_bseq<R, LOGNN, ZP, PP>(); // sequences per buffer
    CT* bufA = reinterpret_cast<CT*>(smem_raw);       // [BSEQ][LPZ]: the sequences of a batch
    CT* bufB = bufA + (PP ? BSEQ * LPZ : 0);          // [BSEQ][LPZ]: those of the next batch (PP)
    constexpr int PW = zc_pw<R, LOGNN, PP>();
    CT* twNl = bufB + BSEQ * LPZ;                     // [ptw_size<NN, 1, PW>()] pass twiddles
    // buffer of batch k (inverse batches first, then forward ones)
    auto buf_of = [&](int k) { return (PP && (k & 1)) ? bufB : bufA; };
    // transforms of the first S sequences of b, in place, natural order in
    // and out; ends with a block barrier.  Lane-group transforms where their
    // registers fit (NN <= 512, and FP32 or the two-wave form); otherwise
    // Stockham passes over the whole block (ip_fft, a barrier per half-pass).
    constexpr bool WAVE = zc_wave<R, LOGNN, PP>();
    auto transform = [&](auto sign, int S, CT* b, int t) {
        if constexpr (!WAVE) {
            ip_fft<CT, LOGNN, decltype(sign)::value, T, SEQ, true>(b, S, twNl, t);
            return;
        }
        constexpr int G = group_lanes<NN>(), PL = NN / G;
        const int gi = t / G, gl = t % G;
        if (gi < S) {
            CT* sq = b + gi * LPZ;
            CT v[PL];
            const int pg = lslot<LS>(gl);
            static_for<0, PL>([&](auto qi) { v[qi] = lsw<decltype(qi)::value * G, LS>(sq, pg); });
            group_fft<CT, LOGNN, decltype(sign)::value, true, PW, 1, 0, LS>(v, sq, gl, twNl);
            static_for<0, PL>([&](auto qi) { lsw<decltype(qi)::value * G, LS>(sq, pg) = v[qi]; });
        }
        __syncthreads();
    };
    static_assert(!WAVE || BSEQ * group_lanes<NN>() <= T, "one lane group per sequence");
    CT* twQl = twNl + ptw_size<NN, 1, PW>();          // [NN]: w^k, k < NN

    const int tid = threadIdx.x;
    const int xl = tid >> LOGNN, r = tid & (NN - 1);
    const int x0 = xcd_tile(blockIdx.x, gridDim.x) * XB;
    const int x = x0 + xl;
    const bool valid = x < g.P;
    const bool wide = g.NTs * (int)sizeof(CT) >= 64;
    // with one x-row per block the sequences of the valid pairs of a batch
    // come first, so a part-padding batch transforms only those
    auto nseq = [](int nv) { return XB == 1 ? 3 * nv : SEQ; };

    // fold task f of this thread: tau = tid + T f -> (row fx, pair fp, index fr);
    // its four values of the next batch are loaded while this one transforms
    CT pf[FT][4];
    auto prefetch = [&](int batch) {
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int tau = tid + T * f;
            const int fx = tau / (ZP * H), fp = (tau / H) % ZP, fr = tau % H;
            const int xg = x0 + fx;
            const int pr = batch * ZP + fp;
            const int r2 = fr == 0 ? H : NN - fr;
            const CT z{(R)0, (R)0};
            pf[f][0] = pf[f][1] = pf[f][2] = pf[f][3] = z;
            if (xg < g.P && pr < NPI) {
                // U[x][line][n] with KK lines of NN (uidx, with the flux form's sizes)
                const CT* u0 = U + ((size_t)(xg * KK + 2 * pr) << LOGNN);
                pf[f][0] = u0[fr];
                pf[f][1] = u0[r2];
                if (2 * pr + 1 < KK) {
                    pf[f][2] = u0[NN + fr];
                    pf[f][3] = u0[NN + r2];
                }
            }
        }
    };

    R vb[3][KK];
    prefetch(0);
    ptw_fill<CT, LOGNN, 1, PW>(twNl, twNN, tid, T);
    for (int i = tid; i < NN; i += T) twQl[i] = twQ[i];
    __syncthreads();

    // In FP64, thread-derived indices are recomputed per batch from an opaque
    // copy of the thread index, so that the compiler does not hoist the LDS
    // addresses of every batch's passes into registers next to the point
    // values (hoisted, FP64 k_zc is 2 % slower at M4 1024^2).  The FP32
    // kernels have the registers to hold them (at most 119 VGPRs, no spills):
    // hoisted, they issue 9 % (NN = 512) to 22 % (NN = 1024) fewer VALU
    // instructions.
    auto opaque_tid = [] {
        int t = threadIdx.x;
        if constexpr (sizeof(R) == 8) asm volatile("" : "+v"(t));
        return t;
    };

    // ---- inverse z transforms of the base lines (c2r, +1) ----
    static_for<0, NBI>([&](auto bi) {
        constexpr int BN = decltype(bi)::value;
        constexpr int NV = NPI - BN * ZP < ZP ? NPI - BN * ZP : ZP; // valid pairs
        const int tid = opaque_tid();
        const int xl = tid >> LOGNN, r = tid & (NN - 1);
        CT* buf = buf_of(BN);
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int tau = tid + T * f;
            const int fx = tau / (ZP * H), fp = (tau / H) % ZP, fr = tau % H;
            if (XB == 1 && fp >= NV) continue;
            const int fm = fr == 0 ? H : NN - fr;
            CT* seq = buf + ((fx * ZP + fp) * 3) * LPZ;
            const CT one{(R)1, (R)0};
            zfold_bw<R, LS>(seq, 0, fr, NN, pf[f][0], pf[f][1], pf[f][2], pf[f][3], one, one);
            zfold_bw<R, LS>(seq + LPZ, 1, fr, NN, pf[f][0], pf[f][1], pf[f][2], pf[f][3], twQl[fr], twQl[fm]);
            zfold_bw<R, LS>(seq + 2 * LPZ, 2, fr, NN, pf[f][0], pf[f][1], pf[f][2], pf[f][3], twq2<R>(twQl, NN, 2 * fr),
                        twq2<R>(twQl, NN, 2 * fm));
        }
        if constexpr (BN + 1 < NBI) prefetch(BN + 1);
        __syncthreads();
        transform(std::integral_constant<int, +1>{}, nseq(NV), buf, tid);
        static_for<0, NV>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int s1 = 2 * (BN * ZP + p);
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const CT zv = buf[((xl * ZP + p) * 3 + b) * LPZ + lslot<LS>(r)];
                vb[b][s1] = zv.x;
                if constexpr (s1 + 1 < KK) vb[b][s1 + 1] = zv.y;
            }
        });
        // with one buffer the next batch overwrites this one's results
        if constexpr (!PP) __syncthreads();
    });

    // ---- products, forward z transforms (r2c, -1), line assembly ----
    const R scale = (R)g.scale;
    const R gn = (R)(g.gamma * r);       // i gamma n at the thread's output n = r
    const R beta = (R)g.beta;
    CT acc[2 * KK];
    // line assembly of the two rows q1, q1 + 1 of a forward pair from their
    // transforms X: each row's terms go into the accumulators of its lines,
    // and a line is stored after its last row
    auto assemble = [&](auto q1c, const CT* X, CT* nlr, size_t nls) {
        constexpr int q1 = decltype(q1c)::value;
        static_for<0, 2>([&](auto hi) {
            constexpr int q = q1 + decltype(hi)::value;
            constexpr int row = FX::order(q);
            if constexpr (row >= 0) {
                const CT xs = cscale(X[decltype(hi)::value], scale);
                static_for<0, FX::nuse(row)>([&](auto ui) {
                    constexpr int u = decltype(ui)::value;
                    constexpr int L = FX::use_line(row, u), kind = FX::use_kind(row, u);
                    constexpr int sg = FX::use_sign(row, u);
                    CT term = xs;
                    if constexpr (kind == 1) term = cscale(xs, (R)sg * beta);
                    else if constexpr (kind == 2) term = CT{-(xs.y * gn), xs.x * gn};
                    if constexpr (FX::first(L) == q) acc[L] = term;
                    else acc[L] = cadd(acc[L], term);
                    if constexpr (FX::last(L) == q) {
                        if (valid) st_piece(nlr + L * nls, acc[L], wide);
                    }
                });
            }
        });
    };
    if constexpr (SG) {
        constexpr int SPB = zc_spb<R>();             // sequences per forward batch
        static_assert(T / group_lanes<NN>() >= 8, "a batch of up to 8 sequences");
        // forward sequences: pair q, phase b at 3 q + b.  With an odd number
        // of rows the last pair holds one row, whose three phases take two
        // sequences (phases 0 and 1 as the real and imaginary parts of the
        // first, phase 2 in the second): M4's 21 rows fill 4 batches of 8
        // instead of 5
        constexpr bool LONE = NR % 2 == 1;
        constexpr int QL = NPO - 1;                  // the last pair
        constexpr int NSF = 3 * NPO - (LONE ? 1 : 0);
        constexpr int NBF = zc_nbf<NSF, SPB>();
        CT X[2];                                      // transforms of the current pair's two rows
        static_for<0, NBF>([&](auto ri) {
            constexpr int RD = decltype(ri)::value;
            constexpr int J0 = RD * SPB, J1 = RD == NBF - 1 ? NSF : J0 + SPB;
            const int tid = opaque_tid();
            const int r = tid & (NN - 1); // one x-row per block
            const int rm = (NN - r) & (NN - 1);
            CT* buf = buf_of(NBI + RD);
            CT* nlr = NL + ((((size_t)(r >> g.lgNS) * (2 * KK) * g.P) + x0) << g.lgNS) + (r & (g.NTs - 1));
            const size_t nls = (size_t)g.P << g.lgNS;
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int sl = 0; sl < KK; sl++) asm volatile("" : "+v"(vb[b][sl]));
            static_for<J0, J1>([&](auto ji) {
                constexpr int j = decltype(ji)::value, q = j / 3, b = j % 3;
                constexpr int r1 = FX::order(2 * q), r2 = FX::order(2 * q + 1);
                R v1 = (R)0, v2 = (R)0;
                if constexpr (LONE && q == QL) { // b = 0: phases 0 and 1; b = 1: phase 2
                    v1 = FX::template row<r1>(vb[2 * b]);
                    if constexpr (b == 0) v2 = FX::template row<r1>(vb[1]);
                } else {
                    if constexpr (r1 >= 0) v1 = FX::template row<r1>(vb[b]);
                    if constexpr (r2 >= 0) v2 = FX::template row<r2>(vb[b]);
                }
                buf[(j - J0) * LPZ + lslot<LS>(r)] = CT{v1, v2};
            });
            __syncthreads();
            transform(std::integral_constant<int, -1>{}, J1 - J0, buf, tid);
            static_for<J0, J1>([&](auto ji) {
                constexpr int j = decltype(ji)::value, q = j / 3, b = j % 3;
                const CT zn = buf[(j - J0) * LPZ + lslot<LS>(r)];
                const CT zm = buf[(j - J0) * LPZ + lslot<LS>(rm)];
                const CT F1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                const CT F2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                if constexpr (LONE && q == QL) {
                    if constexpr (b == 0) { // phase 0 (F1) and phase 1 (F2) of the one row
                        X[0] = cadd(F1, cmul(F2, cconj(twq2<R>(twQl, NN, r))));
                        X[1] = CT{(R)0, (R)0};
                    } else { // phase 2
                        X[0] = cadd(X[0], cmul(F1, cconj(twq2<R>(twQl, NN, 2 * r))));
                        assemble(std::integral_constant<int, 2 * q>{}, X, nlr, nls);
                    }
                } else {
                    if constexpr (b == 0) {
                        X[0] = F1;
                        X[1] = F2;
                    } else {
                        const CT w = cconj(twq2<R>(twQl, NN, r * b));
                        X[0] = cadd(X[0], cmul(F1, w));
                        X[1] = cadd(X[1], cmul(F2, w));
                    }
                    if constexpr (b == 2) assemble(std::integral_constant<int, 2 * q>{}, X, nlr, nls);
                }
            });
        });
        return;
    }
    static_for<0, NBO>([&](auto ri) {
        constexpr int RD = decltype(ri)::value;
        constexpr int NV = NPO - RD * ZP < ZP ? NPO - RD * ZP : ZP; // valid output pairs
        const int tid = opaque_tid();
        const int xl = tid >> LOGNN, r = tid & (NN - 1);
        CT* buf = buf_of(NBI + RD);
        // this thread's element of NL line L is nlr[L * nls] (nidx with the
        // flux form's 2 KK lines): one base per batch instead of the full
        // index per store
        CT* nlr = NL + ((((size_t)(r >> g.lgNS) * (2 * KK) * g.P) + (x0 + xl)) << g.lgNS) + (r & (g.NTs - 1));
        const size_t nls = (size_t)g.P << g.lgNS;
        // the base values pass through an empty asm per batch: products shared
        // by rows of different batches are formed again in each batch instead
        // of being computed once and held in registers across all of them
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int sl = 0; sl < KK; sl++) asm volatile("" : "+v"(vb[b][sl]));
        static_for<0, NV>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int q1 = 2 * (RD * ZP + p);
            constexpr int r1 = FX::order(q1), r2 = FX::order(q1 + 1);
#pragma unroll
            for (int b = 0; b < 3; b++) {
                R v1 = (R)0, v2 = (R)0;
                if constexpr (r1 >= 0) v1 = FX::template row<r1>(vb[b]);
                if constexpr (r2 >= 0) v2 = FX::template row<r2>(vb[b]);
                buf[((xl * ZP + p) * 3 + b) * LPZ + lslot<LS>(r)] = CT{v1, v2};
            }
        });
        __syncthreads();
        transform(std::integral_constant<int, -1>{}, nseq(NV), buf, tid);
        static_for<0, NV>([&](auto pi) {
            constexpr int p = decltype(pi)::value;
            constexpr int q1 = 2 * (RD * ZP + p);
            const int rm = (NN - r) & (NN - 1);
            CT X[2] = {CT{(R)0, (R)0}, CT{(R)0, (R)0}};
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const CT zn = buf[((xl * ZP + p) * 3 + b) * LPZ + lslot<LS>(r)];
                const CT zm = buf[((xl * ZP + p) * 3 + b) * LPZ + lslot<LS>(rm)];
                const CT F1 = CT{(R)0.5 * (zn.x + zm.x), (R)0.5 * (zn.y - zm.y)};
                const CT F2 = CT{(R)0.5 * (zn.y + zm.y), (R)0.5 * (zm.x - zn.x)};
                if (b == 0) {
                    X[0] = F1;
                    X[1] = F2;
                } else {
                    const CT w = cconj(twq2<R>(twQl, NN, r * b));
                    X[0] = cadd(X[0], cmul(F1, w));
                    X[1] = cadd(X[1], cmul(F2, w));
                }
            }
            assemble(std::integral_constant<int, q1>{}, X, nlr, nls);
        });
        if constexpr (!PP) __syncthreads();
    });
}
