// DP-refine chains (align.h:54-145) on gfx950: one wavefront per chain.
//
// A chain is one gap-open branch of dp_refine_chunk: `iters` rounds of
//   score matrix (fused, never stored) -> Needleman-Wunsch fill -> traceback
//   -> Kabsch on the aligned rows -> detailed search -> keep the best.
//
// NW fill: lane l computes rows l, l+64, l+128, ... one after the other; the
// wavefront sweeps them as a systolic pipeline in which every cell's
// up/diagonal neighbours arrive from lane l-1 by a cross-lane shift and its
// left neighbour is the lane's own previous result.  Lane 0 takes the
// previous round's last row from an LDS hand-over buffer.  Trace codes
// (0 diag, 1 up, 2 left) are packed 16 steps per 32-bit word into LDS, or
// into a per-wavefront global slot when the LDS copy would cost resident
// wavefronts, and walked back from there.
//
// Work distribution: a persistent grid pulls chain indices from an atomic
// queue, so chains of very different cost balance across the CUs.
#include "ia_kernels.h"
#include "wave.h"

#include <math.h>
#include <type_traits>

namespace ia {

namespace {

// d0 multipliers of the detailed search (seeds.h:12)
IA_D double d0_mult(int g) {
    switch (g) {
        case 0: return 2.0;
        case 1: return 1.5;
        case 2: return 1.0;
        case 3: return 0.75;
        case 4: return 0.5;
        default: return 0.25;
    }
}

struct ChainLds {
    unsigned *trace;  // trace words (when the trace fits in LDS)
    double *colv;     // the chain's compact native coordinates, x y z per column
    char *cols;                  // compact native secondary structure
    double *rowHG;    // row hand-over buffer: (H, G) of one row's nlen cells
    unsigned *al;     // aligned pairs, reverse order: pred | native << 16 (alias rowHG)
    int *misc;        // [0] aligned-pair count, [2..3] the operation count (ops_word)
    P8 *gq;           // this wavefront's global slot: the aligned pairs as 64-byte records
};

// Aligned pair q of the traceback list (indices below 2^16).
IA_D int al_p(const ChainLds &L, int q) { return (int)(L.al[q] & 0xffffu); }
IA_D int al_n(const ChainLds &L, int q) { return (int)(L.al[q] >> 16); }
IA_D unsigned al_pack(int p, int n) { return (unsigned)p | ((unsigned)n << 16); }

// The wavefront's count of algorithmic FP64 operations outside the NW cells
// (Arena::chain_work): Kabsch 31 n + 3400 per solve, TM-score 29 n,
// selection pass 26 n, n the points of the set.  It lives in the
// wavefront's LDS, so no register holds it across the fill.
IA_D unsigned long long *ops_word(const ChainLds &L) { return (unsigned long long *)(L.misc + 2); }

// Add v to the count: once per wavefront (v wave-uniform), or once per
// lite-search group (the lanes with on set).  Only the FP64 kernels count;
// the fast method's kernels compile without it (their registers are the
// tighter ones).
template <bool FAST>
IA_D void add_ops(const ChainLds &L, unsigned v, bool on) {
    if constexpr (!FAST)
        if (on) atomicAdd(ops_word(L), (unsigned long long)v);
}
template <bool FAST>
IA_D void add_ops(const ChainLds &L, unsigned v) { add_ops<FAST>(L, v, (threadIdx.x & 63) == 0); }

// Gather the iteration's aligned pairs into the wavefront's slot in list
// order (record q = compact pred al_p(q), native al_n(q)), followed by 32 zero
// records, so the passes over them read linear addresses; the loads past na
// of the lite search's batches land on the zeros.
IA_D void gather_aligned(const Arena &A, int off, const ChainLds &L, int na) {
    const int lane = threadIdx.x & 63;
    for (int q = lane; q < na + 32; q += 64) {
        P8 r;
        if (q < na) {
            const unsigned e = L.al[q];
            r.p = A.pc4[off + (int)(e & 0xffffu)];
            r.q = A.qc4[off + (int)(e >> 16)];
        } else {
            r.p = P4{0.0, 0.0, 0.0, 0.0};
            r.q = P4{0.0, 0.0, 0.0, 0.0};
        }
        L.gq[q] = r;
    }
    // the same wavefront reads the records back (one-wavefront workgroup)
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Shift up by one lane (DPP wave_shr:1, a VALU move): lane l receives lane
// l-1's v, lane 0 receives its own `old`.
IA_D int shr1_old(int v, int old) { return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false); }
IA_D double shr1_old(double v, double old) {
    const int lo = shr1_old(__double2loint(v), __double2loint(old));
    const int hi = shr1_old(__double2hiint(v), __double2hiint(old));
    return __hiloint2double(hi, lo);
}
IA_D float shr1_old(float v, float old) {
    return __int_as_float(shr1_old(__float_as_int(v), __float_as_int(old)));
}

// 1.0 or 0.0 from a condition by one 32-bit select (the high word; the low
// word of both is zero).
template <class V>
IA_D V cut_weight(bool c) {
    if constexpr (sizeof(V) == 8) return __hiloint2double(c ? 0x3FF00000 : 0, 0);
    else return __int_as_float(c ? 0x3F800000 : 0);
}
IA_D double fma_v(double a, double b, double c) { return fma(a, b, c); }
IA_D float fma_v(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// max of two finite values as one v_max instruction.  fmax() would first
// quiet possible signalling NaNs of operands that arrive through lane moves
// (an extra v_max per operand); the fill never holds a NaN.
IA_D double vmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
IA_D float vmax(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Needleman-Wunsch fill (nw.h:10-38) over the compact plen x nlen matrix,
// score computed on the fly.
//
// Schedule: lane l computes rows l, l+64, l+128, ... one after another, a
// row taking P = max(nlen, 65) steps, so row a = 64m + l holds cell (a, b)
// at step s(a, b) = mP + l + b (fill_row_start(a, P) + b).  Every cell's up
// and diagonal operands were computed one step earlier by lane l-1 and
// arrive by a lane shift; lane 0 reads the previous round's last row (lane
// 63's) from an LDS buffer, which round 0 finds zeroed (the H = 0 boundary
// with no gap).  The left operand is the lane's own previous value.  A
// round's first 64 steps (phase A) are those where lane q = step - mP moves
// on to its next row; the rest (phase B, two steps per loop iteration) need
// no per-lane row switch.
//
// Trace codes (0 diag, 1 up, 2 left) are packed by step: lane l shifts each
// code into its word, and every 16 steps the wavefront stores its 64 words
// side by side (word (s >> 4) * 64 + l), one coalesced 256-byte store; the
// traceback finds cell (a, b) from its step s(a, b) alone.
// H = max(diag, up, left) is exact; ties only decide the trace code;
// G = H + gap_open for diagonally entered cells, else H, is what down and
// right neighbours add.
IA_HD int fill_period(int nlen) { return nlen > 65 ? nlen : 65; }
enum StepKind { ST_A, ST_AP, ST_G, ST_GP, ST_B, ST_F };
IA_HD int fill_row_start(int a, int P) { return (a >> 6) * P + (a & 63); }
// Steps of a plen x nlen fill, and its trace words: word (s >> 4) * 64 + l
// holds lane l's codes of steps (s & ~15) .. (s | 15).
IA_HD int fill_steps(int plen, int nlen) {
    const int M = (plen + 63) >> 6;
    return (M - 1) * fill_period(nlen) + (plen - 1 - 64 * (M - 1)) + nlen;
}
IA_HD long long trace_words(int plen, int nlen) {
    return (long long)((fill_steps(plen, nlen) + 15) >> 4) * 64;
}

//
// SM_MAT (test entry only): the score of cell (a, b) is smat[a * nlen + b]
// instead of a function of the coordinates, and every cell's H is also
// written to Hout[(a + 1) * (nlen + 1) + b + 1].
template <int SMODE, class TR, bool CL, class V>
IA_D void nw_fill(const Arena &A, const PairInfo &pi, const double *R, const double *t,
                  double gap_d, double cut8_d, double d0sq_s, ChainLds &L, TR *trace,
                  const double *smat = nullptr, double *Hout = nullptr) {
    constexpr bool kXYZ = SMODE == SM_TM || SMODE == SM_IA4 || SMODE == SM_TORCH;  // coordinates scored
    constexpr bool kTorch = SMODE == SM_TORCH;  // full-index rows and columns
    constexpr bool kSS = SMODE == SM_SS || SMODE == SM_IA4;   // SS codes scored
    const int lane = threadIdx.x & 63;
    const int plen = pi.plen, nlen = pi.nlen, off = pi.off;
    constexpr bool F32 = sizeof(V) == 4;
    const int P = fill_period(nlen);
    const int M = (plen + 63) >> 6;
    const int lastl = plen - 1 - 64 * (M - 1);
    // column operands: the wavefront's LDS copy (FP64, or FP32 for the fast
    // method) or, for chains too long to stage, the arena
    const V *colv = (const V *)L.colv;
    auto colx = [&](int b) -> V { if constexpr (F32) return CL ? colv[3 * b] : A.qfx[off + b]; else return CL ? colv[3 * b] : kTorch ? A.qx[off + b] : A.qcx[off + b]; };
    auto coly = [&](int b) -> V { if constexpr (F32) return CL ? colv[3 * b + 1] : A.qfy[off + b]; else return CL ? colv[3 * b + 1] : kTorch ? A.qy[off + b] : A.qcy[off + b]; };
    auto colz = [&](int b) -> V { if constexpr (F32) return CL ? colv[3 * b + 2] : A.qfz[off + b]; else return CL ? colv[3 * b + 2] : kTorch ? A.qz[off + b] : A.qcz[off + b]; };
    auto cols = [&](int b) { return CL ? L.cols[b] : A.ssn[off + A.nmap[off + b]]; };
    const V gap = (V)gap_d, cut8 = (V)cut8_d;
    const double inv_d0sq_d = 1.0 / d0sq_s;
    const V inv_d0sq = (V)inv_d0sq_d;
    typedef typename std::conditional<F32, float2, double2>::type V2;
    V2 *rowHG = (V2 *)L.rowHG;
    for (int b = lane; b < nlen; b += 64) rowHG[b] = V2{(V)0, (V)0};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();

    // transformed coordinates / SS code of the lane's next row
    V nmx = 0, nmy = 0, nmz = 0;
    char nsp = 0;
    // SM_TORCH: the moved row's |p|^2 and its validity (dp.py:18, 27)
    double np2 = 0.0, p2 = 0.0;
    bool nvr = false, vr = false;
    auto next_row = [&](int r) {
        const int rc = min(r, plen - 1);
        if (kXYZ) {
            double ex, ey, ez;
            if constexpr (kTorch)
                xform(R, t, A.px[off + rc], A.py[off + rc], A.pz[off + rc], ex, ey, ez);
            else
                xform(R, t, A.pcx[off + rc], A.pcy[off + rc], A.pcz[off + rc], ex, ey, ez);
            nmx = (V)ex; nmy = (V)ey; nmz = (V)ez;
            if constexpr (kTorch) {
                np2 = ex * ex + ey * ey + ez * ez;
                nvr = A.v[off + rc] != 0;
            }
        }
        if (kSS) nsp = A.ssp[off + A.pmap[off + rc]];
    };
    next_row(lane);
    V mx = 0, my = 0, mz = 0;
    char sp = 0;
    // current row and column; lane l's column counter reaches P at step l,
    // where phase A starts its first row (a = l, b = 0)
    int a = lane - 64, b = P - lane;
    V H = 0, G = 0, D = 0;
    unsigned tw = 0;
    // column operands of the current step.  FP64: loaded per lane one step
    // ahead.  FP32: lane l's column is the one lane l-1 had a step earlier,
    // so they move down the lanes at the start of each step and only lane
    // 0's (c0*) is loaded, one step ahead, at a wave-uniform address.
    V cx = 0, cy = 0, cz = 0, c0x = 0, c0y = 0, c0z = 0;
    char cs = 0, c0s = 0;
    if constexpr (F32) {
        if (kXYZ) { c0x = colx(0); c0y = coly(0); c0z = colz(0); }
        if (kSS) c0s = cols(0);
    } else {  // step 0: lane 0 is on column 0, the others are idle
        if (kXYZ) { cx = colx(0); cy = coly(0); cz = colz(0); }
        if (kSS) cs = cols(0);
    }
    V rH = 0, rG = 0;
    int st = 0;
    // FP64 TM chains with the columns in LDS, steps between a round's first
    // 64 and its last: the next column's operands are read at a running byte
    // offset from colv (clamped to the last column) instead of re-deriving
    // the index from q and the lane (0.6% of the chain kernel's time).
    constexpr bool OPT_C = !F32 && CL && SMODE == SM_TM;
    int cbyte = 0;
    const int cbyte_max = 3 * (int)sizeof(V) * (nlen - 1);

    int rnd = 0;  // the round (M rounds of 64 rows)

    // store the word of steps (st & ~15) .. st, the last nsteps codes
    auto flush = [&](int nsteps) {
        trace[(long long)(st >> 4) * 64 + lane] = nsteps == 16 ? tw : tw >> (32 - 2 * nsteps);
    };

    // Step kinds.  ST_A: phase A (lane q starts its next row; general
    // next-column rule).  ST_G: a round's last step (general rule).  ST_B:
    // phase B, lane l's next column is q + 1 - l, clamped to the last column.
    // ST_F: phase B of a round that is not the last, with nlen >= 65 (P =
    // nlen): every lane's row is below plen and its column q - l lies in
    // [1, nlen - 2], so every cell is active, no index needs a clamp, and
    // the column counter b is set once after the loop instead of per step.
    //
    // ST_AP, ST_GP: ST_A and ST_G for FP64 TM chains with the columns in LDS
    // and nlen >= 65 (P = nlen): every lane's next column comes from the
    // running byte offset cbyte = 24 b (b, the lane's column, in [0, P]), the
    // LDS copy holding column 0 again at index P, so a lane on its row's last
    // column reads its next row's first; the row switch resets cbyte.
    auto step = [&](auto kind, int q, int q_next) {
        constexpr int K = decltype(kind)::value;
        constexpr bool PA = K == ST_A || K == ST_AP;
        constexpr bool GC = K == ST_A || K == ST_G;
        constexpr bool PER = K == ST_AP || K == ST_GP;
        constexpr bool PB = K == ST_B;
        constexpr bool PF = K == ST_F;
        const V hu = shr1_old(H, rH);
        const V gu = shr1_old(G, rG);
        if constexpr (F32) {
            if (kXYZ) {
                cx = shr1_old(cx, c0x);
                cy = shr1_old(cy, c0y);
                cz = shr1_old(cz, c0z);
            }
            if (kSS) cs = (char)shr1_old((int)cs, (int)c0s);
        }
        if constexpr (PA) {
            // lane q starts its next row: its diagonal and left operands are
            // the H = 0 boundary.  One lane switches per step, so this is a
            // branch (moves under a one-lane exec mask), not selects in all
            // 64 lanes; the empty volatile asm keeps it from being if-converted.
            if (lane == q) {
                asm volatile("");
                if constexpr (PER) {
                    cbyte = 0;  // a and b are not used on this path
                } else {
                    a += 64; b = 0;
                }
                mx = nmx; my = nmy; mz = nmz; sp = nsp;
                if constexpr (kTorch) { p2 = np2; vr = nvr; }
                D = (V)0; G = (V)0;
            }
        }
        const V Dc = D, Gl = G;
        // per-lane activity, needed only by the matrix test entry: inactive
        // cells (rows >= plen, columns >= nlen, round 0's lanes before their
        // first row) never feed an active one and their codes are never read
        const bool act = PF || SMODE != SM_MAT || ((unsigned)a < (unsigned)plen && b < nlen);
        const int bc = PF ? q - lane : b;  // this step's column
        // next step's column operands (prefetch)
        const int rq = PF ? q_next : min(q_next, nlen - 1);  // lane 0's next column
        V nx = 0, ny = 0, nz = 0;
        char ns = 0;
        if constexpr (F32) {
            if (kXYZ) { c0x = colx(rq); c0y = coly(rq); c0z = colz(rq); }
            if (kSS) c0s = cols(rq);
        } else if constexpr (OPT_C && (PB || PF || PER)) {
            cbyte += 3 * (int)sizeof(V);
            const V *cp = (const V *)((const char *)colv + (PB ? min(cbyte, cbyte_max) : cbyte));
            nx = cp[0]; ny = cp[1]; nz = cp[2];
        } else {
            int bn;
            if constexpr (GC) bn = b + 1 >= P ? 0 : min(b + 1, nlen - 1);
            else if constexpr (PF) bn = q + 1 - lane;
            else bn = min(q + 1 - lane, nlen - 1);
            if (kXYZ) { nx = colx(bn); ny = coly(bn); nz = colz(bn); }
            if (kSS) ns = cols(bn);
        }
        const V2 rn = rowHG[rq];
        V s, cut_w = 0;
        if constexpr (SMODE == SM_MAT) {
            s = act ? (V)smat[(long long)a * nlen + bc] : (V)0;
        } else if (SMODE == SM_SS) {
            s = (sp == cs) ? (V)1 : (V)0;
        } else if (SMODE == SM_TORCH) {
            // d2 = clamp(p2 + n2 - 2 p.q, 0) in the engine's order; the score
            // is kept where d2 <= score_d8^2 and row and column are valid
            const double n2 = cx * cx + cy * cy + cz * cz;
            const double cr = mx * cx + my * cy + mz * cz;
            const double d2 = vmax((double)(p2 + n2 - 2.0 * cr), 0.0);
            s = (V)(1.0 / (1.0 + d2 / d0sq_s));
            const bool keep = d2 <= cut8 && vr && A.v[off + min(bc, nlen - 1)] != 0;
            cut_w = cut_weight<V>(keep);
        } else {
            const V dx = mx - cx, dy = my - cy, dz = mz - cz;
            V d2, tm;
            if constexpr (F32) {
                d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                tm = __builtin_amdgcn_rcpf(__builtin_fmaf(d2, inv_d0sq, 1.0f));
            } else {
                d2 = dx * dx + dy * dy + dz * dz;
                const double y = 1.0 + div_by(d2, d0sq_s, inv_d0sq_d);
                // TM mode keeps only cells with d2 <= cut8, whose y is far
                // below 2^1022; IA4 scores every cell, so it divides
                tm = SMODE == SM_TM ? recip_in_range(y) : 1.0 / y;
            }
            if (SMODE == SM_TM) {
                s = tm;
            } else {
                s = tm;
                if (sp == cs) s += (V)0.5;
            }
            if constexpr (SMODE == SM_TM) cut_w = cut_weight<V>(d2 <= cut8);
        }
        // TM mode: diag = Dc + (d2 <= cut8 ? tm : 0) as fma(w, tm, Dc) with
        // w = 1 or 0 (one rounding; Dc >= +0.0, so w = 0 leaves Dc exactly)
        V diag;
        if constexpr (SMODE == SM_TM || SMODE == SM_TORCH) diag = fma_v(cut_w, s, Dc);
        else diag = Dc + s;
        const V mgl = vmax(gu, Gl);
        const V hm = vmax(diag, mgl);
        const bool cD = diag >= mgl;
        // ties: diag, then left, then up (nw.h:29-33); the PyTorch engine's
        // first maximum of (diag, up, left) takes up before left (dp.py:51)
        const unsigned tn = cD ? 0u : kTorch ? (gu >= Gl ? 1u : 2u) : (Gl >= gu ? 2u : 1u);
        // G = diag + gap when entered diagonally (then hm = diag), else mgl
        // (= hm > 0): fma(1, gap, hm) = diag + gap, fma(0, gap, hm) = hm
        const V gn = fma_v(cut_weight<V>(cD), gap, hm);
        // to the next round's lane 0.  Lane 63's cell is wave-uniform: before
        // it switches rows (q < 63) it is on row 64 rnd - 1 at column
        // P + q - 63, from then on on row 64 rnd + 63 at column q - 63.
        if constexpr (PF) {
            if (lane == 63) rowHG[q - 63] = V2{hm, gn};
        } else {
            const int r63 = q < 63 ? 64 * rnd - 1 : 64 * rnd + 63;
            const int c63 = q < 63 ? P + q - 63 : q - 63;
            if (lane == 63 && r63 >= 0 && r63 < plen && c63 < nlen) rowHG[c63] = V2{hm, gn};
        }
        if constexpr (SMODE == SM_MAT) {
            if (act) Hout[(long long)(a + 1) * (nlen + 1) + bc + 1] = (double)hm;
        }
        tw = __builtin_amdgcn_alignbit(tn, tw, 2);
        D = hu; H = hm; G = gn;
        if constexpr (!F32) { cx = nx; cy = ny; cz = nz; cs = ns; }
        rH = rn.x; rG = rn.y;
        if ((st & 15) == 15) flush(16);
        st++;
        if constexpr (!PF && !PER) b++;
    };

    // periodic column offsets: lane l starts on column P - l (round 0's
    // lanes before their first row count up to it)
    const bool per = OPT_C && nlen >= 65;
    if constexpr (OPT_C) cbyte = 3 * (int)sizeof(V) * (P - lane);
    using KA = std::integral_constant<int, ST_A>;
    using KAP = std::integral_constant<int, ST_AP>;
    using KGP = std::integral_constant<int, ST_GP>;
    using KG = std::integral_constant<int, ST_G>;
    using KB = std::integral_constant<int, ST_B>;
    using KF = std::integral_constant<int, ST_F>;
    for (int m = 0; m < M; m++) {
        rnd = m;
        const bool last = m == M - 1;
        const int qend = last ? lastl + nlen : P;
        const int qa = min(64, qend);
        if (per) {
            int q = 0;
            for (; q + 1 < qa; q += 2) {
                step(KAP(), q, q + 1 < qend ? q + 1 : 0);
                step(KAP(), q + 1, q + 2 < qend ? q + 2 : 0);
            }
            if (q < qa) step(KAP(), q, q + 1 < qend ? q + 1 : 0);
        } else {
            int q = 0;
            for (; q + 1 < qa; q += 2) {
                step(KA(), q, q + 1 < qend ? q + 1 : 0);
                step(KA(), q + 1, q + 2 < qend ? q + 2 : 0);
            }
            if (q < qa) step(KA(), q, q + 1 < qend ? q + 1 : 0);
        }
        // phase B, two steps per iteration (lets loop-carried values
        // alternate registers instead of being copied back every step).
        // Phase B starts at q = 64 (a shorter last round has none), where
        // lane l's next column is q + 1 - l.
        const int qb = last ? qend : P - 1;
        if (!last) next_row(64 * (m + 1) + lane);
        if constexpr (OPT_C) cbyte = 3 * (int)sizeof(V) * (qa - lane);
        int q = qa;
        if (!last && nlen >= 65) {
            for (; q + 1 < qb; q += 2) {
                step(KF(), q, q + 1);
                step(KF(), q + 1, q + 2);
            }
            if (q < qb) step(KF(), q, q + 1);
            b = qb - lane;  // the column of step qb, the next one
        } else {
            for (; q + 1 < qb; q += 2) {
                step(KB(), q, q + 1);
                step(KB(), q + 1, q + 2);
            }
            if (q < qb) step(KB(), q, q + 1);
        }
        if (!last) {
            if (per) step(KGP(), P - 1, 0);
            else step(KG(), P - 1, 0);
        }
    }
    if (st & 15) {  // the last, partial word
        st--;
        flush((st & 15) + 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Trace word of cell (a, b) and the code's bit position in it.
IA_D int trace_word_index(int a, int b, int P, int &shift) {
    const int s = fill_row_start(a, P) + b;
    shift = (s & 15) * 2;
    return (s >> 4) * 64 + (a & 63);
}

// Needleman-Wunsch fill + traceback; returns the aligned count (pairs in
// L.al, reverse order).  nw.h:10-68 semantics.
template <int SMODE, class TR, bool CL, bool FAST>
IA_D int nw_align(const Arena &A, const PairInfo &pi, const double *R, const double *t,
                  double gap, double cut8, double d0sq_s, ChainLds &L, TR *trace,
                  const double *smat = nullptr, double *Hout = nullptr) {
    const int lane = threadIdx.x & 63;
    const int plen = pi.plen, nlen = pi.nlen;
    const int P = fill_period(nlen);
    typedef typename std::conditional<FAST, float, double>::type V;
    nw_fill<SMODE, TR, CL, V>(A, pi, R, t, gap, cut8, d0sq_s, L, trace, smat, Hout);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // traceback (nw.h:42-68)
    if constexpr (!std::is_same<TR, unsigned>::value) {
        // trace in LDS: lane 0 walks it
        if (lane == 0) {
            int i = plen, j = nlen, k = 0;
            while (i > 0 && j > 0) {
                int sh;
                const unsigned w = trace[trace_word_index(i - 1, j - 1, P, sh)];
                const unsigned d = (w >> sh) & 3u;
                if (d == 0) {
                    L.al[k] = al_pack(i - 1, j - 1);
                    k++;
                    i--;
                    j--;
                } else if (d == 1) {
                    i--;
                } else {
                    j--;
                }
            }
            L.misc[0] = k;
        }
    } else {
        // trace in global memory: a window of four step-words (lane q holds
        // lane q's words of steps 16 (w0 - k) .. 16 (w0 - k) + 15, k < 4) is
        // loaded with four coalesced loads; the wavefront walks it with
        // wave-uniform cell (a, b) and step s, reading each cell's word from
        // lane a & 63 by readlane, and reloads when the path's step leaves
        // the window.  Along the path the step only decreases: a move that
        // leaves row a for a - 1 takes 1 step, or P - 63 from lane 0 (the
        // previous round's lane 63), and a move that leaves column b takes
        // 1 more.  Aligned pair k goes to lane k & 63 of a register; each 64
        // of them are stored with one write.
        int a = plen - 1, b = nlen - 1, k = 0;
        int s = fill_row_start(a, P) + b;
        unsigned acc = 0;
        while (a >= 0 && b >= 0) {
            const int w0 = s >> 4;
            unsigned wv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) wv[q] = w0 - q >= 0 ? trace[(long long)(w0 - q) * 64 + lane] : 0u;
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int wk = (int)wv[kk];
                const int lo = (w0 - kk) << 4;  // the word's first step
                while (a >= 0 && b >= 0 && s >= lo) {
                    const unsigned w = (unsigned)__builtin_amdgcn_readlane(wk, a & 63);
                    const unsigned d = (w >> ((s & 15) * 2)) & 3u;
                    // a move to row a - 1 from lane 0 takes P - 64 more steps
                    const int wrap = (a & 63) == 0 ? P - 64 : 0;
                    if (d == 0) {
                        acc = lane == (k & 63) ? al_pack(a, b) : acc;
                        k++;
                        if ((k & 63) == 0) L.al[k - 64 + lane] = acc;
                        s -= 2 + wrap;
                        a--;
                        b--;
                    } else if (d == 1) {
                        s -= 1 + wrap;
                        a--;
                    } else {
                        s -= 1;
                        b--;
                    }
                }
            }
        }
        if (lane < (k & 63)) L.al[(k & ~63) + lane] = acc;
        if (lane == 0) L.misc[0] = k;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    return L.misc[0];
}

// Lite detailed search (align.h:15-50 with use_frag = false): the whole
// aligned set as the only seed, cutoffs d0s x {2,1.5,1,.75,.5,.25}, 8
// refinement rounds, TM-score with the d8 cut.  Eight lanes per cutoff; lane
// s of a group takes aligned pairs s, s+8, s+16, ... in that order.
//
// The passes over the aligned pairs read them from the wavefront's slot of
// gathered 64-byte records (gather_aligned), two pairs per lane before using
// either, and the sums take every pair through an FMA with a 0/1 weight w:
// fma(w, v, sum) rounds once, so w = 1 adds v as written and w = 0 adds a
// signed zero.  The sums start at +0.0 and so never hold -0.0, which makes
// adding a zero the same as skipping the pair, bit for bit.  The six
// cutoffs' rotations are solved with V split by rows over each quad
// (horn_rotation_quad).
constexpr int kLiteBatch = 2;

template <bool FAST>
IA_D void lite_search(const Arena &A, const PairInfo &pi, const ChainLds &L, int na,
                      const double *nxR, const double *nxt, double sd8, Hit &det) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 3, s = lane & 7;
    const bool gact = g < 6;
    const double cut = pi.d0s * d0_mult(g);
    const double lim = cut * cut;
    double R[9], t[3];
#pragma unroll
    for (int q = 0; q < 9; q++) R[q] = nxR[q];
    t[0] = nxt[0]; t[1] = nxt[1]; t[2] = nxt[2];
    // the aligned pairs of one batch: lane s takes pairs base + s + 8k
    struct Batch {
        double x[kLiteBatch], y[kLiteBatch], z[kLiteBatch];
        double a[kLiteBatch], b[kLiteBatch], c[kLiteBatch];
        bool in[kLiteBatch];
    };
    // the gathered records: q < na + 32 stays inside the slot (zeros past na)
    const P8 *gl = L.gq + s;
    auto load = [&](int base, bool on, Batch &B) {
#pragma unroll
        for (int k = 0; k < kLiteBatch; k++) {
            const int q = base + s + 8 * k;
            B.in[k] = on && q < na;
            const P8 r = gl[base + 8 * k];
            B.x[k] = r.p.x; B.y[k] = r.p.y; B.z[k] = r.p.z;
            B.a[k] = r.q.x; B.b[k] = r.q.y; B.c[k] = r.q.z;
        }
    };
    bool done = !gact;
    int last = -1;
    // the selection (this lane's bits, aligned sets up to 512 pairs) that
    // the group's current (R, t) was solved from
    unsigned long long rset = 0;
    bool rset_ok = false;
    const bool lead = gact && s == 0;  // counts the group's operations
    for (int it = 0; it < 8; it++) {
        if (__all(done)) break;
        add_ops<FAST>(L, 26u * na, lead && !done);
        double c[6] = {0, 0, 0, 0, 0, 0};
        int cnt = 0;
        if (FAST) {  // one pass: selection + raw second moments
            double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int base = 0; base < na; base += 8 * kLiteBatch) {
                Batch B;
                load(base, !done, B);
#pragma unroll
                for (int k = 0; k < kLiteBatch; k++) {
                    const bool sel = B.in[k] &&
                        dist2_moved_fma(R, t, B.x[k], B.y[k], B.z[k], B.a[k], B.b[k], B.c[k]) <= lim;
                    const double w = sel ? 1.0 : 0.0;
                    const double x = w * B.x[k], y = w * B.y[k], z = w * B.z[k];
                    cnt += sel ? 1 : 0;
                    c[0] += x; c[1] += y; c[2] += z;
                    c[3] = fma(w, B.a[k], c[3]); c[4] = fma(w, B.b[k], c[4]); c[5] = fma(w, B.c[k], c[5]);
                    h[0] = fma(x, B.a[k], h[0]); h[1] = fma(x, B.b[k], h[1]); h[2] = fma(x, B.c[k], h[2]);
                    h[3] = fma(y, B.a[k], h[3]); h[4] = fma(y, B.b[k], h[4]); h[5] = fma(y, B.c[k], h[5]);
                    h[6] = fma(z, B.a[k], h[6]); h[7] = fma(z, B.b[k], h[7]); h[8] = fma(z, B.c[k], h[8]);
                }
            }
            cnt = group8_isum(cnt);
#pragma unroll
            for (int d = 0; d < 6; d++) c[d] = group8_sum(c[d]);
#pragma unroll
            for (int k = 0; k < 9; k++) h[k] = group8_sum(h[k]);
            if (!done && cnt >= 3) {
                add_ops<FAST>(L, 31u * cnt + 3400u, lead);
                const double fc = (double)cnt;
                const double cp[3] = {c[0] / fc, c[1] / fc, c[2] / fc};
                const double cq[3] = {c[3] / fc, c[4] / fc, c[5] / fc};
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) h[3 * i + j] -= fc * cp[i] * cq[j];
                solve_rotation_quad<FAST>(h, R);
                finish_t(cp, cq, R, t);
            }
            if (!done) {
                if (cnt == last) done = true;
                last = cnt;
            }
            continue;
        }
        // lane s's pair base + s + 8k is its pair (base >> 3) + k: with at most
        // 64 per lane, the first pass leaves its selection in a bit mask
        const bool keep_sel = na <= 8 * 64;
        unsigned long long selmask = 0;
        for (int base = 0; base < na; base += 8 * kLiteBatch) {
            Batch B;
            load(base, !done, B);
#pragma unroll
            for (int k = 0; k < kLiteBatch; k++) {
                const bool sel = B.in[k] &&
                    dist2_moved(R, t, B.x[k], B.y[k], B.z[k], B.a[k], B.b[k], B.c[k]) <= lim;
                if (keep_sel) selmask |= (unsigned long long)(sel ? 1 : 0) << ((base >> 3) + k);
                cnt += sel ? 1 : 0;
                // fma(1, v, c) = c + v and fma(0, v, c) = c exactly
                const double w = sel ? 1.0 : 0.0;
                c[0] = fma(w, B.x[k], c[0]); c[1] = fma(w, B.y[k], c[1]); c[2] = fma(w, B.z[k], c[2]);
                c[3] = fma(w, B.a[k], c[3]); c[4] = fma(w, B.b[k], c[4]); c[5] = fma(w, B.c[k], c[5]);
            }
        }
        cnt = group8_isum(cnt);
#pragma unroll
        for (int d = 0; d < 6; d++) c[d] = group8_sum(c[d]);
        // a group whose selection is the set its (R, t) was solved from
        // would solve the same sums in the same order again: its (R, t)
        // stands as is (it then stops, its count having repeated)
        const bool same = group8_isum(keep_sel && rset_ok && selmask == rset ? 0 : 1) == 0;
        const bool solve = !done && cnt >= 3 && !same;
        if (__any(solve)) {
            double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            double cp[3], cq[3];
            const double fc = (double)cnt;
            cp[0] = c[0] / fc; cp[1] = c[1] / fc; cp[2] = c[2] / fc;
            cq[0] = c[3] / fc; cq[1] = c[4] / fc; cq[2] = c[5] / fc;
            for (int base = 0; base < na; base += 8 * kLiteBatch) {
                Batch B;
                load(base, solve, B);
#pragma unroll
                for (int k = 0; k < kLiteBatch; k++) {
                    bool sel;
                    if (keep_sel)
                        sel = (selmask >> ((base >> 3) + k)) & 1ull;
                    else
                        sel = B.in[k] &&
                              dist2_moved(R, t, B.x[k], B.y[k], B.z[k], B.a[k], B.b[k], B.c[k]) <= lim;
                    // h += (d e) w: the product rounded as written, then
                    // added (w = 1) or skipped (w = 0) exactly
                    const double w = sel ? 1.0 : 0.0;
                    const double d0 = B.x[k] - cp[0], d1 = B.y[k] - cp[1], d2 = B.z[k] - cp[2];
                    const double e0 = B.a[k] - cq[0], e1 = B.b[k] - cq[1], e2 = B.c[k] - cq[2];
                    h[0] = fma(d0 * e0, w, h[0]); h[1] = fma(d0 * e1, w, h[1]); h[2] = fma(d0 * e2, w, h[2]);
                    h[3] = fma(d1 * e0, w, h[3]); h[4] = fma(d1 * e1, w, h[4]); h[5] = fma(d1 * e2, w, h[5]);
                    h[6] = fma(d2 * e0, w, h[6]); h[7] = fma(d2 * e1, w, h[7]); h[8] = fma(d2 * e2, w, h[8]);
                }
            }
#pragma unroll
            for (int k = 0; k < 9; k++) h[k] = group8_sum(h[k]);
            if (solve) {
                add_ops<FAST>(L, 31u * cnt + 3400u, lead);
                solve_rotation_quad<FAST>(h, R);
                finish_t(cp, cq, R, t);
                rset = selmask;
                rset_ok = keep_sel;
            }
        }
        if (!done) {
            if (cnt == last) done = true;
            last = cnt;
        }
    }
    // score with the d8 cut over the aligned set; a kept term has
    // 1 + d2 / d0^2 <= 1 + sd8^2 / d0^2, far inside the exact reciprocal's range
    const double d0sq = fmax(pi.d0 * pi.d0, 1e-12), c8 = sd8 * sd8;
    const double inv_d0sq = 1.0 / d0sq;
    double acc = 0.0;
    for (int base = 0; base < na; base += 8 * kLiteBatch) {
        Batch B;
        load(base, gact, B);
#pragma unroll
        for (int k = 0; k < kLiteBatch; k++) {
            const double d2 = dist2_moved(R, t, B.x[k], B.y[k], B.z[k], B.a[k], B.b[k], B.c[k]);
            const bool sel = B.in[k] && d2 <= c8;
            const double tm = recip_in_range(1.0 + div_by(d2, d0sq, inv_d0sq));
            acc = fma(sel ? 1.0 : 0.0, tm, acc);
        }
    }
    acc = group8_sum(acc);
    const double score = acc / fmax(pi.Lnorm, 1.0);
    add_ops<FAST>(L, 29u * na, lead);
    // first maximum over the cutoffs in order (search.h:375)
    double bs = kNeg;
    int bg = -1;
    for (int q = 0; q < 6; q++) {
        const double sq = __shfl(score, q * 8, 64);
        if (sq > bs) { bs = sq; bg = q; }
    }
    const int src = bg * 8;
#pragma unroll
    for (int q = 0; q < 9; q++) det.R[q] = __shfl(R[q], src, 64);
#pragma unroll
    for (int q = 0; q < 3; q++) det.t[q] = __shfl(t[q], src, 64);
    det.score = bs;
}

IA_D int find_group(const ChainGroup *groups, int n_groups, long long task) {
    int lo = 0, hi = n_groups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (groups[mid].chain_off <= task) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int SMODE, class TR, bool CL, bool FAST>
IA_D void run_chain(const Arena &A, const ChainGroup &G, const PairInfo &pi, int c,
                    const Hit *xf, Hit *out, ListOut *lout, int *lists_p, int *lists_n,
                    int *list_na, int list_cap, ChainLds &L, TR *trace) {
    const int lane = threadIdx.x & 63;
    const int off = pi.off;
    // SM_TORCH chains fill all N x N residue pairs (full indices)
    PairInfo pf = pi;
    if (SMODE == SM_TORCH) pf.plen = pf.nlen = pi.N;
    for (int b = lane; CL && b < pf.nlen; b += 64) {
        if (SMODE == SM_TORCH) {
            L.colv[3 * b] = A.qx[off + b];
            L.colv[3 * b + 1] = A.qy[off + b];
            L.colv[3 * b + 2] = A.qz[off + b];
        } else if (SMODE != SM_SS) {
            if (FAST) {
                float *cf = (float *)L.colv;
                cf[3 * b] = A.qfx[off + b];
                cf[3 * b + 1] = A.qfy[off + b];
                cf[3 * b + 2] = A.qfz[off + b];
            } else {
                L.colv[3 * b] = A.qcx[off + b];
                L.colv[3 * b + 1] = A.qcy[off + b];
                L.colv[3 * b + 2] = A.qcz[off + b];
            }
        }
        if (SMODE == SM_SS || SMODE == SM_IA4) L.cols[b] = A.ssn[off + A.nmap[off + b]];
    }
    // column 0 again at index nlen: the fill's running column offset reads
    // index P = nlen where a lane's row ends and its next row starts
    if (CL && SMODE == SM_TM && !FAST && lane == 0) {
        L.colv[3 * pi.nlen] = A.qcx[off];
        L.colv[3 * pi.nlen + 1] = A.qcy[off];
        L.colv[3 * pi.nlen + 2] = A.qcz[off];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const bool tm = SMODE == SM_TM;
    const double gap = tm ? ((c & 1) ? 0.0 : -0.6) : G.gap_ss;
    const int item = tm ? (c >> 1) : c;
    const int g = tm ? (c & 1) : 0;
    // CI_XFORM with a selection: the chain of selected item i runs item
    // isel[sel_off + i] in place (start and result slots of that item)
    const int xitem = (G.init == CI_XFORM && G.sel) ? A.isel[G.sel_off + item] : item;
    Hit *dst = out + G.out_off + (long long)xitem * G.out_stride + g;
    double R[9], t[3];
    bool skip = false;
    if (G.init == CI_XFORM) {
        const Hit &h = xf[G.xf_off + (long long)xitem * G.xf_stride + (G.xf_per_chain ? g : 0)];
#pragma unroll
        for (int q = 0; q < 9; q++) R[q] = h.R[q];
        t[0] = h.t[0]; t[1] = h.t[1]; t[2] = h.t[2];
    } else if (G.init == CI_FRAG) {
        const int fi = G.sel ? A.isel[G.sel_off + item] : item;
        int L_ = 0;
        while (L_ + 1 < G.n_lens && G.items_prefix[L_ + 1] <= fi) L_++;
        const int r = fi - G.items_prefix[L_];
        const int flen = G.lens[L_];
        const int ps = (r / G.n_ns[L_]) * G.jp, ns = (r % G.n_ns[L_]) * G.jn;
        wave_kabsch<FAST>(flen,
                    [&](int q, double &x, double &y, double &z, double &a, double &b, double &cc) {
                        x = A.pcx[off + ps + q]; y = A.pcy[off + ps + q]; z = A.pcz[off + ps + q];
                        a = A.qcx[off + ns + q]; b = A.qcy[off + ns + q]; cc = A.qcz[off + ns + q];
                    },
                    R, t);
        add_ops<FAST>(L, 31u * flen + 3400u);
    } else {
        const int k = G.k0 + item * G.kstep;
        const int N = pi.N;
        int cnt = 0;
        for (int i = lane; i < N; i += 64) {
            const int j = i - k;
            cnt += (j >= 0 && j < N && A.pv[off + i] && A.nv[off + j]) ? 1 : 0;
        }
        for (int m = 32; m > 0; m >>= 1) cnt += __shfl_xor(cnt, m, 64);
        if (k == 0 || cnt < 3) {
            skip = true;
        } else {
            // compact the pairing in index order through a prefix count
            double c6[6] = {0, 0, 0, 0, 0, 0};
            for (int i = lane; i < N; i += 64) {
                const int j = i - k;
                if (j >= 0 && j < N && A.pv[off + i] && A.nv[off + j]) {
                    c6[0] += A.px[off + i]; c6[1] += A.py[off + i]; c6[2] += A.pz[off + i];
                    c6[3] += A.qx[off + j]; c6[4] += A.qy[off + j]; c6[5] += A.qz[off + j];
                }
            }
#pragma unroll
            for (int d = 0; d < 6; d++) c6[d] = wave_sum(c6[d]);
            const double fc = (double)cnt;
            const double cp[3] = {c6[0] / fc, c6[1] / fc, c6[2] / fc};
            const double cq[3] = {c6[3] / fc, c6[4] / fc, c6[5] / fc};
            double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = lane; i < N; i += 64) {
                const int j = i - k;
                if (j >= 0 && j < N && A.pv[off + i] && A.nv[off + j]) {
                    const double d0 = A.px[off + i] - cp[0], d1 = A.py[off + i] - cp[1], d2 = A.pz[off + i] - cp[2];
                    const double e0 = A.qx[off + j] - cq[0], e1 = A.qy[off + j] - cq[1], e2 = A.qz[off + j] - cq[2];
                    h[0] += d0 * e0; h[1] += d0 * e1; h[2] += d0 * e2;
                    h[3] += d1 * e0; h[4] += d1 * e1; h[5] += d1 * e2;
                    h[6] += d2 * e0; h[7] += d2 * e1; h[8] += d2 * e2;
                }
            }
#pragma unroll
            for (int q = 0; q < 9; q++) h[q] = wave_sum(h[q]);
            solve_rotation_quad<FAST>(h, R);
            finish_t(cp, cq, R, t);
            add_ops<FAST>(L, 31u * cnt + 3400u);
        }
    }
    Hit best;
    set_identity(best.R, best.t);
    best.score = 0.0;  // align.h:61-64
    if (skip) {
        best.score = -INFINITY;
        if (lane == 0) *dst = best;
        return;
    }
    const double d0dp = fmax(pi.d0, 1.0);
    double d0sq_s;
    if (SMODE == SM_TM) d0sq_s = fmax(d0dp * d0dp, 1e-12);
    else if (SMODE == SM_TORCH) d0sq_s = fmax(pi.d0 * pi.d0, 1e-12);  // d0, not max(d0, 1)
    else {
        const double d01 = fmax(pi.d0 + 1.5, pi.d0s);  // ss.h:118-119
        d0sq_s = d01 * d01;
    }
    const double cut8 = G.sd8 * G.sd8;
    const double d0sq = fmax(pi.d0 * pi.d0, 1e-12);
    for (int it = 0; it < G.iters; it++) {
        const int na = nw_align<SMODE, TR, CL, FAST>(A, pf, R, t, gap, cut8, d0sq_s, L, trace);
        if (SMODE == SM_TORCH) {
            // dp.py:189-205: the list keeps every diagonal step; entries
            // whose residues are not both valid are masked out of the
            // Kabsch, the iteration score and the detailed search
            const int slot = G.list_slot0 + c;
            auto entry = [&](int q, double &x, double &y, double &z, double &a, double &b,
                             double &cc) -> bool {
                const int ap = al_p(L, q), an = al_n(L, q);
                x = A.px[off + ap]; y = A.py[off + ap]; z = A.pz[off + ap];
                a = A.qx[off + an]; b = A.qy[off + an]; cc = A.qz[off + an];
                return A.v[off + ap] != 0 && A.v[off + an] != 0;
            };
            int nv = 0;
            for (int q = lane; q < na; q += 64) nv += (A.v[off + al_p(L, q)] != 0 && A.v[off + al_n(L, q)] != 0) ? 1 : 0;
            for (int m = 32; m > 0; m >>= 1) nv += __shfl_xor(nv, m, 64);
            double nR[9], nt[3];
            double sc = 0.0;
            if (nv >= 3) {
                wave_kabsch_masked<FAST>(na, entry, nR, nt);
                double acc = 0.0;
                for (int q = lane; q < na; q += 64) {
                    double x, y, z, a, b, cc;
                    const bool in = entry(q, x, y, z, a, b, cc);
                    acc += in ? tm_term(dist2_moved(nR, nt, x, y, z, a, b, cc), d0sq) : 0.0;
                }
                sc = wave_sum(acc) / fmax(pi.Lnorm, 1.0);
                add_ops<FAST>(L, 60u * nv + 3400u);  // Kabsch + TM-score over the valid entries
            } else {
                set_identity(nR, nt);
            }
            for (int q = lane; q < na; q += 64) {
                const int rq = na - 1 - q;
                lists_p[(long long)slot * list_cap + q] = al_p(L, rq);
                lists_n[(long long)slot * list_cap + q] = al_n(L, rq);
            }
            if (lane == 0) {
                list_na[slot] = na;
                list_na[G.nv_slot0 + c] = nv;
                ListOut o;
#pragma unroll
                for (int q = 0; q < 9; q++) o.R[q] = nR[q];
                o.t[0] = nt[0]; o.t[1] = nt[1]; o.t[2] = nt[2];
                o.it_score = sc;
                o.na = na;
                o.nvalid = nv;
                lout[slot] = o;
            }
            return;
        }
        if (G.mode == CM_LIST) {
            const int slot = G.list_slot0 + c;
            double nR[9], nt[3];
            double sc = 0.0;
            if (na >= 3) {
                wave_kabsch<FAST>(na,
                            [&](int q, double &x, double &y, double &z, double &a, double &b, double &cc) {
                                const int ap = al_p(L, q), an = al_n(L, q);
                                x = A.pcx[off + ap]; y = A.pcy[off + ap]; z = A.pcz[off + ap];
                                a = A.qcx[off + an]; b = A.qcy[off + an]; cc = A.qcz[off + an];
                            },
                            nR, nt);
                double acc = 0.0;
                for (int q = lane; q < na; q += 64) {
                    const int ap = al_p(L, q), an = al_n(L, q);
                    acc += tm_term(dist2_moved(nR, nt, A.pcx[off + ap], A.pcy[off + ap], A.pcz[off + ap],
                                               A.qcx[off + an], A.qcy[off + an], A.qcz[off + an]), d0sq);
                }
                sc = wave_sum(acc) / fmax(pi.Lnorm, 1.0);
                add_ops<FAST>(L, 60u * na + 3400u);  // Kabsch + TM-score
            } else {
                set_identity(nR, nt);
            }
            // aligned pairs in increasing order, full indices
            for (int q = lane; q < na; q += 64) {
                const int rq = na - 1 - q;
                lists_p[(long long)slot * list_cap + q] = A.pmap[off + al_p(L, rq)];
                lists_n[(long long)slot * list_cap + q] = A.nmap[off + al_n(L, rq)];
            }
            if (lane == 0) {
                list_na[slot] = na;
                ListOut o;
#pragma unroll
                for (int q = 0; q < 9; q++) o.R[q] = nR[q];
                o.t[0] = nt[0]; o.t[1] = nt[1]; o.t[2] = nt[2];
                o.it_score = sc;
                o.na = na;
                o.nvalid = na;
                lout[slot] = o;
            }
            return;
        }
        if (na < 3) continue;  // align.h:106
        gather_aligned(A, off, L, na);
        double nR[9], nt[3];
        wave_kabsch<FAST>(na,
                    [&](int q, double &x, double &y, double &z, double &a, double &b, double &cc) {
                        const P8 r = L.gq[q];
                        x = r.p.x; y = r.p.y; z = r.p.z;
                        a = r.q.x; b = r.q.y; cc = r.q.z;
                    },
                    nR, nt);
        double acc = 0.0;
        for (int q = lane; q < na; q += 64) {
            const P8 r = L.gq[q];
            acc += tm_term(dist2_moved(nR, nt, r.p.x, r.p.y, r.p.z, r.q.x, r.q.y, r.q.z), d0sq);
        }
        const double it_score = wave_sum(acc) / fmax(pi.Lnorm, 1.0);
        add_ops<FAST>(L, 60u * na + 3400u);  // Kabsch + TM-score
        Hit det;
        lite_search<FAST>(A, pi, L, na, nR, nt, G.sd8, det);
        Hit step;
        if (det.score > it_score) {
            step = det;
        } else {
#pragma unroll
            for (int q = 0; q < 9; q++) step.R[q] = nR[q];
            step.t[0] = nt[0]; step.t[1] = nt[1]; step.t[2] = nt[2];
            step.score = it_score;
        }
#pragma unroll
        for (int q = 0; q < 9; q++) R[q] = step.R[q];
        t[0] = step.t[0]; t[1] = step.t[1]; t[2] = step.t[2];
        if (step.score > best.score) best = step;
    }
    if (lane == 0) *dst = best;
}

// TORCH: the instantiation that runs the PyTorch engine's DP chains
// (SM_TORCH) and nothing else; the others never compile that score mode.
template <class TR, bool CL, bool FAST, bool TORCH>
IA_D void chain_loop(const Arena &A, const ChainGroup *groups, int n_groups, long long n_chains,
                     const Hit *xf, Hit *out, ListOut *lout, int *lists_p, int *lists_n,
                     int *list_na, int list_cap, int *queue, ChainLds &L, TR *trace) {
    const int lane = threadIdx.x & 63;
    while (true) {
        long long task = 0;
        if (lane == 0) task = atomicAdd(queue, 1);
        // lane 0's value in scalar registers: everything derived from the
        // chain index (group, pair, sizes, loop bounds) is then wave-uniform
        task = (long long)__builtin_amdgcn_readfirstlane((int)task);
        if (task >= n_chains) break;
        const int gi = find_group(groups, n_groups, task);
        const ChainGroup &G = groups[gi];
        const int c = (int)(task - G.chain_off);
        const PairInfo &pi = A.info[G.pair];
        if constexpr (TORCH)
            run_chain<SM_TORCH, TR, CL, FAST>(A, G, pi, c, xf, out, lout, lists_p, lists_n, list_na, list_cap, L, trace);
        else if (G.smat == SM_TM)
            run_chain<SM_TM, TR, CL, FAST>(A, G, pi, c, xf, out, lout, lists_p, lists_n, list_na, list_cap, L, trace);
        else if (G.smat == SM_SS)
            run_chain<SM_SS, TR, CL, FAST>(A, G, pi, c, xf, out, lout, lists_p, lists_n, list_na, list_cap, L, trace);
        else
            run_chain<SM_IA4, TR, CL, FAST>(A, G, pi, c, xf, out, lout, lists_p, lists_n, list_na, list_cap, L, trace);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    // one atomic per wavefront: its chains' operations outside the NW cells
    if constexpr (!FAST)
        if (lane == 0 && *ops_word(L) > 0) atomicAdd(&A.chain_work[0], *ops_word(L));
}

// LDS carve-out of one wavefront (shared with chain_lds_bytes).  The
// traceback's aligned-pair lists reuse the row hand-over buffer, which is
// dead once the fill has finished (and is cleared by the next fill).
struct LdsPlan {
    size_t rowHG, misc, al, colv, cols, trace, total;
    IA_HD LdsPlan(int max_plen, int max_nlen, bool trace_in_lds, bool cols_in_lds) {
        (void)max_plen;
        const size_t n = (size_t)max_nlen, nc = cols_in_lds ? n : 0;
        size_t o = 0;
        rowHG = o; o += 16 * n;
        al = rowHG;  // 4 min(plen, nlen) <= 16 nlen bytes
        colv = o; o += nc ? 24 * (nc + 1) : 0;  // + a copy of column 0 at index nlen
        misc = o; o += 16;
        cols = o; o += (nc + 15) & ~(size_t)15;
        trace = o;
        if (trace_in_lds) o += (size_t)trace_words(max_plen, max_nlen) * 4;
        total = o;
    }
};

// W wavefronts per SIMD: W = 3 caps the kernel at 168 VGPRs (three resident
// wavefronts hide each other's latency; the spills this costs sit outside
// the fill loop), W = 2 at 256.  Which one runs is a launch-tuning choice.
// TORCH (FP64 only) runs the PyTorch engine's DP chains.
template <bool FAST, int W, bool TORCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W))) void chain_kernel(Arena A, const ChainGroup *groups,
                                                   int n_groups, long long n_chains,
                                                   const Hit *xf, Hit *out, ListOut *lout,
                                                   int *lists_p, int *lists_n, int *list_na,
                                                   int list_cap, int max_plen, int max_nlen,
                                                   int *queue, unsigned *trace_g,
                                                   long long slot_words, int cols_in_lds,
                                                   P8 *gq, long long gq_stride) {
    extern __shared__ __align__(16) unsigned char lds[];
    const LdsPlan P(max_plen, max_nlen, trace_g == nullptr, cols_in_lds != 0);
    ChainLds L;
    L.rowHG = (double *)(lds + P.rowHG);
    L.colv = (double *)(lds + P.colv);
    L.misc = (int *)(lds + P.misc);
    L.cols = (char *)(lds + P.cols);
    L.al = (unsigned *)(lds + P.al);
    L.trace = (unsigned *)(lds + P.trace);
    L.gq = gq + (long long)blockIdx.x * gq_stride;
    if (!FAST && threadIdx.x == 0) *ops_word(L) = 0;
    if (trace_g) {
        unsigned *tg = trace_g + (long long)blockIdx.x * slot_words;
        if (cols_in_lds)
            chain_loop<unsigned, true, FAST, TORCH>(A, groups, n_groups, n_chains, xf, out, lout, lists_p,
                                       lists_n, list_na, list_cap, queue, L, tg);
        else
            chain_loop<unsigned, false, FAST, TORCH>(A, groups, n_groups, n_chains, xf, out, lout, lists_p,
                                        lists_n, list_na, list_cap, queue, L, tg);
    } else {
        typedef __attribute__((address_space(3))) unsigned lds_u32;
        if (cols_in_lds)
            chain_loop<lds_u32, true, FAST, TORCH>(A, groups, n_groups, n_chains, xf, out, lout, lists_p,
                                      lists_n, list_na, list_cap, queue, L, (lds_u32 *)L.trace);
        else
            chain_loop<lds_u32, false, FAST, TORCH>(A, groups, n_groups, n_chains, xf, out, lout, lists_p,
                                       lists_n, list_na, list_cap, queue, L, (lds_u32 *)L.trace);
    }
}

// NW fill + traceback of one supplied score matrix per wavefront (the
// device NW of the chains, checked bit for bit against the reference's
// nw_dp / nw_traceback on the same matrix).
template <bool FAST, bool GT>
__global__ __launch_bounds__(64) void nw_matrix_kernel(const NwProblem *probs, const double *smat,
                                                       double *H, signed char *T, int *ap,
                                                       int *an, int *na_out, int max_plen,
                                                       int max_nlen, unsigned *trace_g) {
    extern __shared__ __align__(16) unsigned char lds[];
    const LdsPlan P(max_plen, max_nlen, !GT, false);
    ChainLds L;
    L.rowHG = (double *)(lds + P.rowHG);
    L.colv = (double *)(lds + P.colv);
    L.misc = (int *)(lds + P.misc);
    L.cols = (char *)(lds + P.cols);
    L.al = (unsigned *)(lds + P.al);
    L.trace = (unsigned *)(lds + P.trace);
    L.gq = nullptr;
    const int lane = threadIdx.x & 63;
    const NwProblem pr = probs[blockIdx.x];
    PairInfo pi{};
    pi.plen = pr.plen;
    pi.nlen = pr.nlen;
    const Arena A{};
    double R[9], t[3];
    set_identity(R, t);
    typedef __attribute__((address_space(3))) unsigned lds_u32;
    typedef typename std::conditional<GT, unsigned, lds_u32>::type TR;
    TR *trace = GT ? (TR *)(trace_g + (long long)blockIdx.x * trace_words(max_plen, max_nlen))
                   : (TR *)L.trace;
    const int na = nw_align<SM_MAT, TR, false, FAST>(A, pi, R, t, pr.gap, 0.0, 1.0, L, trace,
                                                     smat + pr.s_off, H + pr.h_off);
    const int Pd = fill_period(pr.nlen);
    const long long cells = (long long)pr.plen * pr.nlen;
    for (long long c = lane; c < cells; c += 64) {
        const int a = (int)(c / pr.nlen), b = (int)(c % pr.nlen);
        int sh;
        const unsigned w = trace[trace_word_index(a, b, Pd, sh)];
        T[pr.h_off + (long long)(a + 1) * (pr.nlen + 1) + b + 1] = (signed char)((w >> sh) & 3u);
    }
    for (int q = lane; q < na; q += 64) {
        ap[pr.a_off + q] = al_p(L, na - 1 - q);
        an[pr.a_off + q] = al_n(L, na - 1 - q);
    }
    if (lane == 0) na_out[blockIdx.x] = na;
}

}  // namespace

size_t nw_matrix_lds_bytes(int max_plen, int max_nlen, bool trace_in_lds) {
    return LdsPlan(max_plen, max_nlen, trace_in_lds, false).total;
}

long long nw_matrix_trace_words(int max_plen, int max_nlen) {
    return trace_words(max_plen, max_nlen);
}

void launch_nw_matrix(hipStream_t s, const NwProblem *probs, int n, const double *smat, double *H,
                      signed char *T, int *ap, int *an, int *na, int max_plen, int max_nlen,
                      unsigned *trace_g, bool fast) {
    if (n <= 0) return;
    const size_t lds = nw_matrix_lds_bytes(max_plen, max_nlen, trace_g == nullptr);
#define IA_LAUNCH_NWM(F, G)                                                                       \
    hipLaunchKernelGGL((nw_matrix_kernel<F, G>), dim3(n), dim3(64), lds, s, probs, smat, H, T, ap, \
                       an, na, max_plen, max_nlen, trace_g)
    if (trace_g) {
        if (fast) IA_LAUNCH_NWM(true, true);
        else IA_LAUNCH_NWM(false, true);
    } else {
        if (fast) IA_LAUNCH_NWM(true, false);
        else IA_LAUNCH_NWM(false, false);
    }
#undef IA_LAUNCH_NWM
}

int chain_blocks_per_cu(size_t lds, bool fast, int waves, bool torch) {
    int nb = 0;
    hipError_t e;
#define IA_OCC(F, W, T) hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, chain_kernel<F, W, T>, 64, lds)
    if (waves == 2)
        e = torch ? IA_OCC(false, 2, true) : fast ? IA_OCC(true, 2, false) : IA_OCC(false, 2, false);
    else
        e = torch ? IA_OCC(false, 3, true) : fast ? IA_OCC(true, 3, false) : IA_OCC(false, 3, false);
#undef IA_OCC
    return e == hipSuccess ? nb : 1;
}

long long chain_trace_words(int max_plen, int max_nlen) { return trace_words(max_plen, max_nlen); }

int chain_max_nlen() {
    constexpr size_t kLdsMax = 160 * 1024;
    // the smallest plan (trace and columns outside LDS) is 16 bytes per
    // compact native residue plus 16
    int n = (int)((kLdsMax - 16) / 16);
    while (n > 1 && LdsPlan(1, n, false, false).total > kLdsMax) n--;
    return n;
}

size_t chain_lds_bytes(int max_plen, int max_nlen, bool trace_in_lds, bool cols_in_lds) {
    return LdsPlan(max_plen, max_nlen, trace_in_lds, cols_in_lds).total;
}

void launch_chains(hipStream_t s, const Arena &A, const ChainGroup *groups, int n_groups,
                   long long n_chains, const Hit *xf, Hit *out, ListOut *lout, int *lists_p,
                   int *lists_n, int *list_na, int list_cap, int max_plen, int max_nlen,
                   int *queue, unsigned *trace_scratch, long long trace_slot_words,
                   int n_slots, bool cols_in_lds, bool fast, int waves, P8 *gq,
                   long long gq_stride, bool torch) {
    if (n_chains <= 0) return;
    const size_t lds = chain_lds_bytes(max_plen, max_nlen, trace_scratch == nullptr, cols_in_lds);
    const int grid = n_slots;
#define IA_LAUNCH_CHAIN(F, W, T)                                                                 \
    hipLaunchKernelGGL((chain_kernel<F, W, T>), dim3(grid), dim3(64), lds, s, A, groups, n_groups, \
                       n_chains, xf, out, lout, lists_p, lists_n, list_na, list_cap, max_plen,   \
                       max_nlen, queue, trace_scratch, trace_slot_words, (int)cols_in_lds, gq,  \
                       gq_stride)
    if (waves == 2) {
        if (torch) IA_LAUNCH_CHAIN(false, 2, true);
        else if (fast) IA_LAUNCH_CHAIN(true, 2, false);
        else IA_LAUNCH_CHAIN(false, 2, false);
    } else {
        if (torch) IA_LAUNCH_CHAIN(false, 3, true);
        else if (fast) IA_LAUNCH_CHAIN(true, 3, false);
        else IA_LAUNCH_CHAIN(false, 3, false);
    }
#undef IA_LAUNCH_CHAIN
}

}  // namespace ia
